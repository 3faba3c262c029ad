// Library context (stream, pinned scalars, scratch), option parsing, status
// strings and result release.

#include <climits>
#include <cstring>
#include <map>
#include <chrono>
#include <mutex>
#include <unordered_map>

#include <algorithm>

#include "ipt_internal.h"

namespace ipt {

// ------------------------------------------------------------------ allocation cache

namespace {
std::mutex g_pool_mutex;
thread_local hipStream_t t_pool_stream = nullptr;
struct PoolBlock {
    hipStream_t stream;
    size_t bytes;
};
std::multimap<std::pair<hipStream_t, size_t>, void *> g_pool_free;
std::unordered_map<void *, PoolBlock> g_pool_live;
size_t g_pool_cached = 0;
constexpr size_t kPoolCap = (size_t)32 << 30;  // cached bytes kept at most

size_t pool_round(size_t b)
{
    const size_t g = b <= ((size_t)1 << 20) ? 256 : ((size_t)2 << 20);
    return (b + g - 1) / g * g;
}

void pool_trim_all()
{
    std::vector<void *> drop;
    {
        std::lock_guard<std::mutex> lock(g_pool_mutex);
        for (auto &e : g_pool_free) {
            drop.push_back(e.second);
        }
        g_pool_free.clear();
        g_pool_cached = 0;
    }
    if (!drop.empty()) {
        (void)hipDeviceSynchronize();
        for (void *q : drop) {
            (void)hipFree(q);
        }
    }
}
}  // namespace

void pool_set_stream(hipStream_t s)
{
    t_pool_stream = s;
}

namespace {
bool g_trace_on = false;
std::chrono::steady_clock::time_point g_trace_t0;
}  // namespace

void trace_begin()
{
    g_trace_on = env_flag("IPT_HIP_HOST_TRACE") != 0;
    g_trace_t0 = std::chrono::steady_clock::now();
}

bool trace_on() { return g_trace_on; }

void trace_mark(const char *what)
{
    if (g_trace_on) {
        const double us = 1e6 * std::chrono::duration<double>(std::chrono::steady_clock::now() - g_trace_t0).count();
        fprintf(stderr, "IPT host trace %9.1f us  %s\n", us, what);
    }
}

int pool_alloc(void **p, size_t bytes)
{
    const size_t rb = pool_round(bytes);
    {
        std::lock_guard<std::mutex> lock(g_pool_mutex);
        auto it = g_pool_free.find({t_pool_stream, rb});
        if (it != g_pool_free.end()) {
            *p = it->second;
            g_pool_free.erase(it);
            g_pool_cached -= rb;
            g_pool_live[*p] = {t_pool_stream, rb};
            return IPT_CUDA_SUCCESS;
        }
    }
    void *q = nullptr;
    if (g_trace_on) {
        char msg[96];
        snprintf(msg, sizeof(msg), "hipMalloc of %zu bytes (none cached)", rb);
        trace_mark(msg);
    }
    if (hipMalloc(&q, rb) != hipSuccess) {
        (void)hipGetLastError();
        pool_trim_all();  // cached blocks of other sizes may be what is missing
        if (hipMalloc(&q, rb) != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            return IPT_CUDA_ALLOCATION_FAILED;
        }
    }
    std::lock_guard<std::mutex> lock(g_pool_mutex);
    g_pool_live[q] = {t_pool_stream, rb};
    *p = q;
    return IPT_CUDA_SUCCESS;
}

void pool_free(void *p)
{
    if (p == nullptr) {
        return;
    }
    {
        std::lock_guard<std::mutex> lock(g_pool_mutex);
        auto it = g_pool_live.find(p);
        if (it != g_pool_live.end()) {
            const PoolBlock b = it->second;
            g_pool_live.erase(it);
            if (g_pool_cached + b.bytes <= kPoolCap) {
                g_pool_free.insert({{b.stream, b.bytes}, p});
                g_pool_cached += b.bytes;
                return;
            }
        }
    }
    if (g_trace_on) {
        trace_mark("hipFree of a block not cached");
    }
    (void)hipFree(p);  // not cached (foreign, detached, or over the cap)
}

void pool_detach(void *p)
{
    if (p == nullptr) {
        return;
    }
    std::lock_guard<std::mutex> lock(g_pool_mutex);
    g_pool_live.erase(p);
}

void pool_release_stream(hipStream_t s)
{
    std::vector<void *> drop;
    {
        std::lock_guard<std::mutex> lock(g_pool_mutex);
        for (auto it = g_pool_free.begin(); it != g_pool_free.end();) {
            if (it->first.first == s) {
                drop.push_back(it->second);
                g_pool_cached -= it->first.second;
                it = g_pool_free.erase(it);
            } else {
                ++it;
            }
        }
    }
    for (void *q : drop) {
        (void)hipFree(q);
    }
}

namespace {
std::mutex g_ctx_mutex;
Context *g_default_ctx[64] = {nullptr};

int context_init(Context *ctx, hipStream_t stream)
{
    IPT_HIPCHK(hipGetDevice(&ctx->device));
    if (stream != nullptr) {
        ctx->stream = stream;
        ctx->own_stream = false;
    } else {
        // default priority only
        IPT_HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->own_stream = true;
    }
    IPT_HIPCHK(hipHostMalloc((void **)&ctx->host_scalars, 4096 * sizeof(double), hipHostMallocDefault));
    IPT_HIPCHK(hipHostMalloc((void **)&ctx->host_flags, 4096 * sizeof(int), hipHostMallocDefault));
    IPT_HIPCHK(hipHostMalloc((void **)&ctx->host_status, 4096 * sizeof(double), hipHostMallocDefault));
    return IPT_CUDA_SUCCESS;
}

void context_destroy(Context *ctx)
{
    if (ctx == nullptr) {
        return;
    }
    if (ctx->stream != nullptr) {
        (void)hipStreamSynchronize(ctx->stream);
        pool_release_stream(ctx->stream);
    }
    for (int i = 0; i < 8; ++i) {
        if (ctx->slots[i] != nullptr) {
            (void)hipFree(ctx->slots[i]);
        }
    }
    if (ctx->host_scalars != nullptr) {
        (void)hipHostFree(ctx->host_scalars);
    }
    if (ctx->host_status != nullptr) {
        (void)hipHostFree(ctx->host_status);
    }
    if (ctx->host_flags != nullptr) {
        (void)hipHostFree(ctx->host_flags);
    }
    if (ctx->host_pinned != nullptr) {
        (void)hipHostFree(ctx->host_pinned);
    }
    if (ctx->own_stream && ctx->stream != nullptr) {
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}
}  // namespace

Context *context_from_handle(void *handle, int *status)
{
    *status = IPT_CUDA_SUCCESS;
    if (handle != nullptr) {
        Context *c = static_cast<Context *>(handle);
        pool_set_stream(c->stream);
        return c;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        *status = IPT_CUDA_CUDA_ERROR;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    if (g_default_ctx[dev] == nullptr) {
        Context *ctx = new Context();
        int st = context_init(ctx, nullptr);
        if (st != IPT_CUDA_SUCCESS) {
            context_destroy(ctx);
            *status = st;
            return nullptr;
        }
        g_default_ctx[dev] = ctx;
    }
    pool_set_stream(g_default_ctx[dev]->stream);
    return g_default_ctx[dev];
}

int context_host_pinned(Context *ctx, size_t bytes, void **out)
{
    if (bytes > ctx->host_pinned_bytes) {
        if (ctx->host_pinned != nullptr) {
            IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
            IPT_HIPCHK(hipHostFree(ctx->host_pinned));
            ctx->host_pinned = nullptr;
            ctx->host_pinned_bytes = 0;
        }
        const size_t want = std::max(bytes, (size_t)1 << 20);
        IPT_HIPCHK(hipHostMalloc(&ctx->host_pinned, want, hipHostMallocDefault));
        ctx->host_pinned_bytes = want;
    }
    *out = ctx->host_pinned;
    return IPT_CUDA_SUCCESS;
}

int context_scratch(Context *ctx, int slot, size_t bytes, void **out)
{
    if (slot < 0 || slot >= 8) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    if (bytes == 0) {
        bytes = 256;
    }
    if (ctx->slot_bytes[slot] < bytes) {
        if (ctx->slots[slot] != nullptr) {
            IPT_HIPCHK(hipStreamSynchronize(ctx->stream));
            (void)hipFree(ctx->slots[slot]);
            ctx->slots[slot] = nullptr;
            ctx->slot_bytes[slot] = 0;
        }
        size_t grow = bytes + bytes / 4;
        if (hipMalloc(&ctx->slots[slot], grow) != hipSuccess) {
            (void)hipGetLastError();
            ctx->slots[slot] = nullptr;
            return IPT_CUDA_ALLOCATION_FAILED;
        }
        ctx->slot_bytes[slot] = grow;
    }
    *out = ctx->slots[slot];
    return IPT_CUDA_SUCCESS;
}

// Environment parsing, IPT/src/ipt_cuda.cu:18-73: empty, unparsable,
// negative or non-finite values fall back to the default.
double env_double(const char *name, double default_value)
{
    const char *raw = getenv(name);
    if (raw == nullptr || raw[0] == '\0') {
        return default_value;
    }
    char *end = nullptr;
    double v = strtod(raw, &end);
    if (end == raw || v < 0.0 || !std::isfinite(v)) {
        return default_value;
    }
    return v;
}

int env_int(const char *name, int default_value)
{
    const char *raw = getenv(name);
    if (raw == nullptr || raw[0] == '\0') {
        return default_value;
    }
    char *end = nullptr;
    long v = strtol(raw, &end, 10);
    if (end == raw || v < 0 || v > INT_MAX) {
        return default_value;
    }
    return (int)v;
}

int env_flag(const char *name)
{
    const char *raw = getenv(name);
    if (raw == nullptr || raw[0] == '\0') {
        return 0;
    }
    if (!strcmp(raw, "0") || !strcmp(raw, "false") || !strcmp(raw, "FALSE") || !strcmp(raw, "no") ||
        !strcmp(raw, "NO")) {
        return 0;
    }
    return 1;
}

int env_flag_default(const char *name, int default_value)
{
    const char *raw = getenv(name);
    if (raw == nullptr || raw[0] == '\0') {
        return default_value;
    }
    return env_flag(name);
}

}  // namespace ipt

using namespace ipt;

extern "C" const char *ipt_cuda_status_string(int status)
{
    switch (status) {
    case IPT_CUDA_SUCCESS: return "success";
    case IPT_CUDA_INVALID_ARGUMENT: return "invalid argument";
    case IPT_CUDA_ALLOCATION_FAILED: return "allocation failed";
    case IPT_CUDA_CUDA_ERROR: return "HIP runtime error";
    case IPT_CUDA_CUBLAS_ERROR: return "device BLAS error";
    case IPT_CUDA_CUSPARSE_ERROR: return "sparse kernel error";
    case IPT_CUDA_CUSOLVER_ERROR: return "dense solver error";
    default: return "unknown error";
    }
}

// Reset every field to the reference's "empty" values, IPT/src/ipt_cuda.cu:398-475.
extern "C" void ipt_cuda_free_result(IPTCudaResult *r)
{
    if (r == nullptr) {
        return;
    }
    free(r->vectors);
    free(r->values);
    free(r->davidson_history);
    free(r->davidson_selection_history);
    free(r->davidson_block_history);
    memset(r, 0, sizeof(*r));
    r->fixed_point_residual = NAN;
    r->basis_orthogonality_frobenius_error = NAN;
    r->basis_orthogonality_max_abs_error = NAN;
    r->ritz_vectors_orthogonality_frobenius_error = NAN;
    r->ritz_vectors_orthogonality_max_abs_error = NAN;
    r->davidson_target_index = -1;
    r->davidson_residual_before = NAN;
    r->davidson_residual_after = NAN;
    r->davidson_denom_clip = NAN;
    r->adaptive_coupling_tau = NAN;
    r->adaptive_target_block_start = -1;
    r->adaptive_target_block_end = -1;
    r->best_so_far_step = -1;
    r->best_so_far_max_residual = NAN;
    r->best_so_far_max_residual_index = -1;
    r->pair28_best_residual = NAN;
    r->pair29_best_residual = NAN;
    r->gap28_29 = NAN;
    r->relative_gap28_29 = NAN;
    r->cluster_residual_fro = NAN;
    r->cluster_residual_max = NAN;
    snprintf(r->pair28_lock_state, sizeof(r->pair28_lock_state), "unknown");
    snprintf(r->pair29_lock_state, sizeof(r->pair29_lock_state), "unknown");
}

extern "C" void ipt_cuda_free_device_result(double *d_vectors, double *d_values)
{
    if (d_vectors != nullptr) {
        (void)hipFree(d_vectors);
    }
    if (d_values != nullptr) {
        (void)hipFree(d_values);
    }
}

extern "C" int ipt_hip_create_handle(void **handle_out, void *hip_stream)
{
    if (handle_out == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    *handle_out = nullptr;
    Context *ctx = new Context();
    int st = context_init(ctx, (hipStream_t)hip_stream);
    if (st != IPT_CUDA_SUCCESS) {
        context_destroy(ctx);
        return st;
    }
    *handle_out = ctx;
    return IPT_CUDA_SUCCESS;
}

extern "C" int ipt_hip_destroy_handle(void *handle)
{
    if (handle == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    context_destroy(static_cast<Context *>(handle));
    return IPT_CUDA_SUCCESS;
}
