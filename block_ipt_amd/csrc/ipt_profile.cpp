// Kernel timing instrumentation: HIP events on the stream the main kernels
// run on -- either the dispatches' own start and end timestamps
// (hipExtLaunchKernelGGL, no extra packets) or markers recorded around a
// span of launches -- summed on request.
#include <cstring>
#include <mutex>

#include "ipt_internal.h"
#include "../../include/ipt_hip_ext.h"

namespace ipt {

namespace {
const char *kNames[PROF_COUNT] = {"gemm_f64", "lowp_gemm", "block_update", "spmm", "spmm_f32", "spmm_tiles"};
std::mutex g_mu;
unsigned g_mask = 0;  // bit i: family i is timed
struct Timed {
    hipEvent_t start;
    hipEvent_t stop;
    long long cols;
};
std::vector<Timed> g_ev[PROF_COUNT];
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event()
{
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
        return nullptr;
    }
    return e;
}
}  // namespace

bool profile_on() { return g_mask != 0; }

ProfileToken profile_begin(Context *ctx, int id)
{
    ProfileToken t;
    t.id = -1;
    if (id < 0 || id >= PROF_COUNT || !((g_mask >> id) & 1u)) {
        return t;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    t.start = take_event();
    t.stop = take_event();
    if (t.start == nullptr || t.stop == nullptr) {
        return t;
    }
    (void)hipEventRecord(t.start, ctx->stream);
    t.id = id;
    return t;
}

ProfileToken profile_begin_launches(int id)
{
    ProfileToken t;
    t.id = -1;
    t.markers = false;
    if (id < 0 || id >= PROF_COUNT || !((g_mask >> id) & 1u)) {
        return t;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    t.start = take_event();
    t.stop = take_event();
    if (t.start == nullptr || t.stop == nullptr) {
        return t;
    }
    t.id = id;
    return t;
}

void profile_end(Context *ctx, const ProfileToken &t, long long cols)
{
    if (t.id < 0) {
        return;
    }
    if (t.markers) {
        (void)hipEventRecord(t.stop, ctx->stream);
    }
    std::lock_guard<std::mutex> lock(g_mu);
    g_ev[t.id].push_back(Timed{t.start, t.stop, cols});
}

}  // namespace ipt

using namespace ipt;

extern "C" void ipt_hip_profile_enable(int on) { g_mask = on != 0 ? (1u << PROF_COUNT) - 1u : 0u; }

extern "C" int ipt_hip_profile_enable_only(const char *kernels)
{
    if (kernels == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    unsigned mask = 0;
    const char *p = kernels;
    while (*p != '\0') {
        const char *e = strchr(p, ',');
        const size_t len = e != nullptr ? (size_t)(e - p) : strlen(p);
        int id = -1;
        for (int i = 0; i < PROF_COUNT; ++i) {
            if (strlen(kNames[i]) == len && !strncmp(p, kNames[i], len)) {
                id = i;
            }
        }
        if (id < 0) {
            return IPT_CUDA_INVALID_ARGUMENT;
        }
        mask |= 1u << id;
        p += len;
        if (*p == ',') {
            ++p;
        }
    }
    g_mask = mask;
    return IPT_CUDA_SUCCESS;
}

extern "C" void ipt_hip_profile_reset(void)
{
    std::lock_guard<std::mutex> lock(g_mu);
    for (int i = 0; i < PROF_COUNT; ++i) {
        for (auto &p : g_ev[i]) {
            (void)hipEventSynchronize(p.stop);
            g_pool.push_back(p.start);
            g_pool.push_back(p.stop);
        }
        g_ev[i].clear();
    }
}

extern "C" int ipt_hip_profile_get(const char *kernel, double *total_ms, long long *launches)
{
    if (kernel == nullptr || total_ms == nullptr || launches == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    int id = -1;
    for (int i = 0; i < PROF_COUNT; ++i) {
        if (!strcmp(kernel, kNames[i])) {
            id = i;
        }
    }
    if (id < 0) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    double sum = 0.0;
    for (auto &p : g_ev[id]) {
        float ms = 0.f;
        if (hipEventSynchronize(p.stop) != hipSuccess || hipEventElapsedTime(&ms, p.start, p.stop) != hipSuccess) {
            return IPT_CUDA_CUDA_ERROR;
        }
        sum += ms;
    }
    *total_ms = sum;
    *launches = (long long)g_ev[id].size();
    return IPT_CUDA_SUCCESS;
}

extern "C" int ipt_hip_profile_get_columns(const char *kernel, long long *columns)
{
    if (kernel == nullptr || columns == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    for (int i = 0; i < PROF_COUNT; ++i) {
        if (!strcmp(kernel, kNames[i])) {
            std::lock_guard<std::mutex> lock(g_mu);
            long long c = 0;
            for (auto &p : g_ev[i]) {
                c += p.cols;
            }
            *columns = c;
            return IPT_CUDA_SUCCESS;
        }
    }
    return IPT_CUDA_INVALID_ARGUMENT;
}
