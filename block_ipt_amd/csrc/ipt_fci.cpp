// Determinant-space (full CI) Hamiltonian, FCIDUMP integral files and the
// binary CSC matrix cache: the inputs of the reference's real-chemistry
// sparse benchmarks (IPT/test/sparse/CH4/benchmark_ch4_sto6g_fci_15876_ipt.cu).
//
// The reference builds each column by applying the second-quantised
// operators term by term (one-body links, same-spin two-body links,
// opposite-spin products of one-body links) and summing what lands on the
// same row (:743-903).  Here each nonzero is formed once, from the
// Slater-Condon rules for determinants that differ by zero, one or two
// spin-orbitals; the phases follow the same operator-ordering convention
// (annihilators right to left, creators after them), so the matrix is the
// same up to the order of the floating-point sums.  Columns are built in
// parallel (OpenMP) and concatenated.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ipt_internal.h"
#include "ipt_prep.h"

#include "../../include/ipt_hip_fci.h"

namespace ipt {
namespace {

constexpr int kMaxOrbitals = 24;  // string address tables of 2^norb entries

// Occupation strings of nel electrons in norb orbitals, ordered by their
// occupied-orbital lists in lexicographic order (the enumeration order of
// the reference's generate_strings, benchmark_ch4_..._ipt.cu:637-661, which
// fixes the determinant numbering and so the order of equal diagonals).
std::vector<uint32_t> occupation_strings(int norb, int nel)
{
    std::vector<uint32_t> out;
    if (nel == 0) {
        out.push_back(0u);
        return out;
    }
    std::vector<int> idx((size_t)nel);
    for (int i = 0; i < nel; ++i) {
        idx[(size_t)i] = i;
    }
    for (;;) {
        uint32_t bits = 0;
        for (int i : idx) {
            bits |= 1u << i;
        }
        out.push_back(bits);
        int i = nel - 1;
        while (i >= 0 && idx[(size_t)i] == norb - nel + i) {
            --i;
        }
        if (i < 0) {
            break;
        }
        ++idx[(size_t)i];
        for (int j = i + 1; j < nel; ++j) {
            idx[(size_t)j] = idx[(size_t)j - 1] + 1;
        }
    }
    return out;
}

long long binomial(int n, int k)
{
    if (k < 0 || k > n) {
        return 0;
    }
    long long r = 1;
    for (int i = 1; i <= k; ++i) {
        r = r * (n - k + i) / i;
    }
    return r;
}

inline int popc(uint32_t x) { return __builtin_popcount(x); }

// (-1)^(occupied orbitals of s below bit) -- the phase of moving an operator
// on orbital `bit` past the electrons in front of it
inline int order_sign(uint32_t s, int bit) { return (popc(s & ((1u << bit) - 1u)) & 1) ? -1 : 1; }

// phase of a^+_a a_i |s> (i occupied, a empty or a == i)
inline int single_phase(uint32_t s, int i, int a)
{
    const int s1 = order_sign(s, i);
    const uint32_t t = s & ~(1u << i);
    return s1 * order_sign(t, a);
}

// phase of a^+_a a^+_b a_j a_i |s> (i, j occupied, distinct; a, b empty
// after the annihilations, distinct)
inline int double_phase(uint32_t s, int i, int j, int a, int b)
{
    int sg = order_sign(s, i);
    s &= ~(1u << i);
    sg *= order_sign(s, j);
    s &= ~(1u << j);
    sg *= order_sign(s, b);
    s |= 1u << b;
    sg *= order_sign(s, a);
    return sg;
}

struct Integrals {
    int norb;
    const double *h1;
    const double *eri;
    double ecore;
    double h(int p, int q) const { return h1[(size_t)p * norb + q]; }
    double g(int p, int q, int r, int s) const
    {
        return eri[(((size_t)p * norb + q) * norb + r) * norb + s];
    }
};

void occupied(uint32_t s, int norb, std::vector<int> *occ)
{
    occ->clear();
    for (int o = 0; o < norb; ++o) {
        if (s >> o & 1u) {
            occ->push_back(o);
        }
    }
}

// <D|H|D> for the determinant of strings (sa, sb).  The terms are added in
// the order of the reference's determinant_diagonal
// (benchmark_ch4_..._ipt.cu:694-730: ecore, h_ii orbital by orbital with the
// alpha before the beta electron, 0.5 [(ii|jj) - (ij|ji)] over ordered
// same-spin pairs, (ii|jj) over alpha-beta pairs): the value is the sort key
// of the row order, and determinants of equal energy (spin-swapped ones)
// must tie exactly, as they do there, for the stable sort to keep them in
// determinant order.
double diagonal_element(const Integrals &in, uint32_t sa, uint32_t sb, const std::vector<int> &oa,
                        const std::vector<int> &ob)
{
    double v = in.ecore;
    for (int o = 0; o < in.norb; ++o) {
        if (sa >> o & 1u) {
            v += in.h(o, o);
        }
        if (sb >> o & 1u) {
            v += in.h(o, o);
        }
    }
    for (const std::vector<int> *occ : {&oa, &ob}) {
        for (int i : *occ) {
            for (int j : *occ) {
                v += 0.5 * (in.g(i, i, j, j) - in.g(i, j, j, i));
            }
        }
    }
    for (int i : oa) {
        for (int j : ob) {
            v += in.g(i, i, j, j);
        }
    }
    return v;
}

// <D'|H|D> for D' = a^+_a a_i D in spin sigma, without the phase: occ_same
// and occ_other are the occupations of D in sigma and the other spin
double single_element(const Integrals &in, int i, int a, const std::vector<int> &occ_same,
                      const std::vector<int> &occ_other)
{
    double v = in.h(a, i);
    for (int k : occ_same) {
        if (k != i) {
            v += in.g(a, i, k, k) - in.g(a, k, k, i);
        }
    }
    for (int k : occ_other) {
        v += in.g(a, i, k, k);
    }
    return v;
}

struct Entry {
    int row;
    double v;
};

}  // namespace

int fci_build(int norb, int neleca, int nelecb, const double *h1, const double *eri, double ecore, int sort_diag,
              IPTFciCsc *out)
{
    memset(out, 0, sizeof(*out));
    if (norb <= 0 || norb > kMaxOrbitals || neleca < 0 || nelecb < 0 || neleca > norb || nelecb > norb ||
        h1 == nullptr || eri == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    const long long na = binomial(norb, neleca), nb = binomial(norb, nelecb);
    if (na * nb > 0x7fffffffLL) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    const int n = (int)(na * nb);
    const Integrals in{norb, h1, eri, ecore};
    const std::vector<uint32_t> sa = occupation_strings(norb, neleca), sb = occupation_strings(norb, nelecb);
    std::vector<int> addr_a((size_t)1 << norb, -1), addr_b((size_t)1 << norb, -1);
    for (size_t i = 0; i < sa.size(); ++i) {
        addr_a[sa[i]] = (int)i;
    }
    for (size_t i = 0; i < sb.size(); ++i) {
        addr_b[sb[i]] = (int)i;
    }
    // diagonal and the sorted order
    std::vector<double> diag((size_t)n);
#pragma omp parallel for schedule(dynamic, 64)
    for (int ia = 0; ia < (int)na; ++ia) {
        std::vector<int> oa, ob;
        occupied(sa[(size_t)ia], norb, &oa);
        for (int ib = 0; ib < (int)nb; ++ib) {
            occupied(sb[(size_t)ib], norb, &ob);
            diag[(size_t)ia * nb + ib] = diagonal_element(in, sa[(size_t)ia], sb[(size_t)ib], oa, ob);
        }
    }
    std::vector<int> perm((size_t)n), inv((size_t)n);
    for (int i = 0; i < n; ++i) {
        perm[(size_t)i] = i;
    }
    if (sort_diag) {
        std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return diag[(size_t)x] < diag[(size_t)y]; });
    }
    for (int s = 0; s < n; ++s) {
        inv[(size_t)perm[(size_t)s]] = s;
    }
    // columns in parallel
    std::vector<std::vector<Entry>> cols((size_t)n);
    bool bad_alloc = false;
#pragma omp parallel
    {
        std::vector<int> oa, ob, va, vb;
        std::vector<Entry> e;
#pragma omp for schedule(dynamic, 32)
        for (int c = 0; c < n; ++c) {
            const int det = perm[(size_t)c];
            const int ia = det / (int)nb, ib = det % (int)nb;
            const uint32_t a0 = sa[(size_t)ia], b0 = sb[(size_t)ib];
            occupied(a0, norb, &oa);
            occupied(b0, norb, &ob);
            occupied(~a0 & ((1u << norb) - 1u), norb, &va);
            occupied(~b0 & ((1u << norb) - 1u), norb, &vb);
            e.clear();
            e.push_back({c, diag[(size_t)det]});
            // single excitations in one spin
            for (int spin = 0; spin < 2; ++spin) {
                const uint32_t s0 = spin == 0 ? a0 : b0;
                const std::vector<int> &occ = spin == 0 ? oa : ob, &vir = spin == 0 ? va : vb,
                                       &other = spin == 0 ? ob : oa;
                for (int i : occ) {
                    for (int a : vir) {
                        const uint32_t s1 = (s0 & ~(1u << i)) | (1u << a);
                        const int tdet = spin == 0 ? addr_a[s1] * (int)nb + ib : ia * (int)nb + addr_b[s1];
                        const double v = single_phase(s0, i, a) * single_element(in, i, a, occ, other);
                        e.push_back({inv[(size_t)tdet], v});
                    }
                }
                // double excitations in the same spin: i < j -> a < b
                for (size_t x = 0; x < occ.size(); ++x) {
                    for (size_t y = x + 1; y < occ.size(); ++y) {
                        const int i = occ[x], j = occ[y];
                        for (size_t p = 0; p < vir.size(); ++p) {
                            for (size_t q = p + 1; q < vir.size(); ++q) {
                                const int a = vir[p], b = vir[q];
                                const uint32_t s1 = (s0 & ~(1u << i) & ~(1u << j)) | (1u << a) | (1u << b);
                                const int tdet = spin == 0 ? addr_a[s1] * (int)nb + ib : ia * (int)nb + addr_b[s1];
                                const double v =
                                    double_phase(s0, i, j, a, b) * (in.g(a, i, b, j) - in.g(a, j, b, i));
                                e.push_back({inv[(size_t)tdet], v});
                            }
                        }
                    }
                }
            }
            // one excitation in each spin
            for (int i : oa) {
                for (int a : va) {
                    const int pa = single_phase(a0, i, a);
                    const int ta = addr_a[(a0 & ~(1u << i)) | (1u << a)];
                    for (int j : ob) {
                        for (int b : vb) {
                            const int pb = single_phase(b0, j, b);
                            const int tb = addr_b[(b0 & ~(1u << j)) | (1u << b)];
                            e.push_back({inv[(size_t)ta * nb + tb], pa * pb * in.g(a, i, b, j)});
                        }
                    }
                }
            }
            std::sort(e.begin(), e.end(), [](const Entry &x, const Entry &y) { return x.row < y.row; });
            std::vector<Entry> &col = cols[(size_t)c];
            try {
                for (size_t q = 0; q < e.size();) {
                    const int row = e[q].row;
                    double v = 0.0;
                    for (; q < e.size() && e[q].row == row; ++q) {
                        v += e[q].v;
                    }
                    if (v != 0.0) {
                        col.push_back({row, v});
                    }
                }
            } catch (...) {
#pragma omp critical
                bad_alloc = true;
            }
        }
    }
    if (bad_alloc) {
        return IPT_CUDA_ALLOCATION_FAILED;
    }
    long long nnz = 0;
    for (const auto &col : cols) {
        nnz += (long long)col.size();
    }
    if (nnz > 0x7fffffffLL) {
        fprintf(stderr, "IPT: %lld nonzeros exceed the 32-bit CSC index range\n", nnz);
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    out->col_ptr = (int *)malloc(sizeof(int) * ((size_t)n + 1));
    out->row_ind = (int *)malloc(sizeof(int) * (size_t)std::max<long long>(nnz, 1));
    out->values = (double *)malloc(sizeof(double) * (size_t)std::max<long long>(nnz, 1));
    out->diagonal = (double *)malloc(sizeof(double) * (size_t)n);
    out->perm = (int *)malloc(sizeof(int) * (size_t)n);
    if (!out->col_ptr || !out->row_ind || !out->values || !out->diagonal || !out->perm) {
        ipt_hip_fci_free_csc(out);
        return IPT_CUDA_ALLOCATION_FAILED;
    }
    std::vector<double> rowsum((size_t)n, 0.0);
    long long p = 0;
    for (int c = 0; c < n; ++c) {
        out->col_ptr[c] = (int)p;
        out->diagonal[c] = 0.0;
        bool seen = false;
        for (const Entry &x : cols[(size_t)c]) {
            out->row_ind[p] = x.row;
            out->values[p] = x.v;
            rowsum[(size_t)x.row] += std::fabs(x.v);
            if (x.row == c && !seen) {
                out->diagonal[c] = x.v;
                seen = true;
            }
            ++p;
        }
        std::vector<Entry>().swap(cols[(size_t)c]);
    }
    out->col_ptr[n] = (int)p;
    memcpy(out->perm, perm.data(), sizeof(int) * (size_t)n);
    out->n = n;
    out->nnz = (int)nnz;
    out->matrix_inf_norm = rowsum.empty() ? 0.0 : *std::max_element(rowsum.begin(), rowsum.end());
    out->norb = norb;
    out->nelec = neleca + nelecb;
    out->neleca = neleca;
    out->nelecb = nelecb;
    out->na = (int)na;
    out->nb = (int)nb;
    return IPT_CUDA_SUCCESS;
}

// ------------------------------------------------------------------ FCIDUMP

namespace {

// "&FCI NORB= 9,NELEC=10,MS2=0, ORBSYM=... ISYM=1, &END": the value of key
double namelist_value(const std::string &head, const char *key, bool *found)
{
    std::string up(head);
    for (char &ch : up) {
        ch = (char)toupper((unsigned char)ch);
    }
    const std::string k(key);
    size_t pos = 0;
    while ((pos = up.find(k, pos)) != std::string::npos) {
        // whole word: not preceded by a letter
        const bool start = pos == 0 || !isalpha((unsigned char)up[pos - 1]);
        size_t q = pos + k.size();
        while (q < up.size() && up[q] == ' ') {
            ++q;
        }
        if (start && q < up.size() && up[q] == '=') {
            *found = true;
            return atof(up.c_str() + q + 1);
        }
        pos += k.size();
    }
    *found = false;
    return 0.0;
}

// reads the header up to &END or '/', returns the file positioned after it
FILE *fcidump_open(const char *path, std::string *head)
{
    FILE *f = path ? fopen(path, "r") : nullptr;
    if (f == nullptr) {
        return nullptr;
    }
    head->clear();
    char line[4096];
    bool done = false;
    while (!done && fgets(line, sizeof(line), f)) {
        std::string s(line);
        std::string up(s);
        for (char &ch : up) {
            ch = (char)toupper((unsigned char)ch);
        }
        const size_t e1 = up.find("&END"), e2 = up.find('/');
        if (e1 != std::string::npos || e2 != std::string::npos) {
            s = s.substr(0, std::min(e1, e2));
            done = true;
        }
        *head += s;
        *head += ' ';
    }
    if (!done) {
        fclose(f);
        return nullptr;
    }
    return f;
}

}  // namespace

}  // namespace ipt

using namespace ipt;

extern "C" int ipt_hip_fci_build_csc(int norb, int neleca, int nelecb, const double *h1, const double *eri,
                                     double ecore, int sort_by_diagonal, IPTFciCsc *out)
{
    if (out == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    return fci_build(norb, neleca, nelecb, h1, eri, ecore, sort_by_diagonal, out);
}

extern "C" void ipt_hip_fci_free_csc(IPTFciCsc *m)
{
    if (m == nullptr) {
        return;
    }
    free(m->col_ptr);
    free(m->row_ind);
    free(m->values);
    free(m->diagonal);
    free(m->perm);
    memset(m, 0, sizeof(*m));
}

extern "C" int ipt_hip_fcidump_read_header(const char *path, int *norb, int *nelec, int *ms2)
{
    std::string head;
    FILE *f = fcidump_open(path, &head);
    if (f == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    fclose(f);
    bool fn = false, fe = false, fm = false;
    const double vn = namelist_value(head, "NORB", &fn);
    const double ve = namelist_value(head, "NELEC", &fe);
    const double vm = namelist_value(head, "MS2", &fm);
    if (!fn || !fe || vn < 1 || vn > kMaxOrbitals || ve < 0) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    if (norb) *norb = (int)vn;
    if (nelec) *nelec = (int)ve;
    if (ms2) *ms2 = fm ? (int)vm : 0;
    return IPT_CUDA_SUCCESS;
}

extern "C" int ipt_hip_fcidump_read(const char *path, int norb, double *h1, double *eri, double *ecore)
{
    int nh = 0;
    if (ipt_hip_fcidump_read_header(path, &nh, nullptr, nullptr) != IPT_CUDA_SUCCESS || nh != norb ||
        h1 == nullptr || eri == nullptr || ecore == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    std::string head;
    FILE *f = fcidump_open(path, &head);
    if (f == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    const size_t n2 = (size_t)norb * norb;
    std::fill(h1, h1 + n2, 0.0);
    std::fill(eri, eri + n2 * n2, 0.0);
    *ecore = 0.0;
    auto g = [&](int p, int q, int r, int s) -> double & { return eri[(((size_t)p * norb + q) * norb + r) * norb + s]; };
    char line[512];
    int st = IPT_CUDA_SUCCESS;
    while (fgets(line, sizeof(line), f)) {
        double v = 0.0;
        int i = 0, j = 0, k = 0, l = 0;
        const int got = sscanf(line, "%lf %d %d %d %d", &v, &i, &j, &k, &l);
        if (got <= 0) {
            continue;  // blank line
        }
        if (got != 5 || i < 0 || j < 0 || k < 0 || l < 0 || i > norb || j > norb || k > norb || l > norb) {
            st = IPT_CUDA_INVALID_ARGUMENT;
            break;
        }
        if (i > 0 && j > 0 && k > 0 && l > 0) {
            const int p = i - 1, q = j - 1, r = k - 1, s = l - 1;
            g(p, q, r, s) = g(q, p, r, s) = g(p, q, s, r) = g(q, p, s, r) = v;
            g(r, s, p, q) = g(s, r, p, q) = g(r, s, q, p) = g(s, r, q, p) = v;
        } else if (i > 0 && j > 0 && k == 0 && l == 0) {
            h1[(size_t)(i - 1) * norb + (j - 1)] = v;
            h1[(size_t)(j - 1) * norb + (i - 1)] = v;
        } else if (i == 0 && j == 0 && k == 0 && l == 0) {
            *ecore = v;
        }
        // i > 0, j = k = l = 0: orbital energy, not needed
    }
    fclose(f);
    return st;
}

extern "C" int ipt_hip_fcidump_write(const char *path, int norb, int nelec, int ms2, const double *h1,
                                     const double *eri, double ecore)
{
    if (path == nullptr || norb <= 0 || norb > kMaxOrbitals || h1 == nullptr || eri == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    FILE *f = fopen(path, "w");
    if (f == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    fprintf(f, " &FCI NORB=%d,NELEC=%d,MS2=%d,\n  ORBSYM=", norb, nelec, ms2);
    for (int i = 0; i < norb; ++i) {
        fprintf(f, "1,");
    }
    fprintf(f, "\n  ISYM=1,\n &END\n");
    auto g = [&](int p, int q, int r, int s) { return eri[(((size_t)p * norb + q) * norb + r) * norb + s]; };
    // one representative per 8-fold class: p >= q, r >= s, pq >= rs
    for (int p = 0; p < norb; ++p) {
        for (int q = 0; q <= p; ++q) {
            for (int r = 0; r <= p; ++r) {
                for (int s = 0; s <= r; ++s) {
                    if (p * (p + 1) / 2 + q < r * (r + 1) / 2 + s) {
                        continue;
                    }
                    const double v = g(p, q, r, s);
                    if (v != 0.0) {
                        fprintf(f, "%.17e %d %d %d %d\n", v, p + 1, q + 1, r + 1, s + 1);
                    }
                }
            }
        }
    }
    for (int p = 0; p < norb; ++p) {
        for (int q = 0; q <= p; ++q) {
            const double v = h1[(size_t)p * norb + q];
            if (v != 0.0) {
                fprintf(f, "%.17e %d %d 0 0\n", v, p + 1, q + 1);
            }
        }
    }
    fprintf(f, "%.17e 0 0 0 0\n", ecore);
    return fclose(f) == 0 ? IPT_CUDA_SUCCESS : IPT_CUDA_INVALID_ARGUMENT;
}

// ------------------------------------------------------------------ CSC cache

namespace {

struct CacheHeader {
    char magic[8];
    uint32_t version;
    int32_t n;
    int32_t nnz;
    double matrix_inf_norm;
    int32_t norb;
    int32_t nelec;
    int32_t neleca;
    int32_t nelecb;
    int32_t na;
    int32_t nb;
};
static_assert(sizeof(CacheHeader) == 56, "the cache header is 56 bytes");
constexpr char kCacheMagic[8] = {'I', 'P', 'T', 'C', 'H', '4', 'C', '1'};

}  // namespace

extern "C" int ipt_hip_csc_cache_write(const char *path, const IPTFciCsc *m)
{
    if (path == nullptr || path[0] == '\0' || m == nullptr || m->n <= 0 || m->nnz < 0 || m->col_ptr == nullptr ||
        (m->nnz > 0 && (m->row_ind == nullptr || m->values == nullptr)) || m->diagonal == nullptr ||
        !csc_valid(m->col_ptr, m->row_ind, m->n, m->nnz)) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    CacheHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, kCacheMagic, sizeof(h.magic));
    h.version = 1;
    h.n = m->n;
    h.nnz = m->nnz;
    h.matrix_inf_norm = m->matrix_inf_norm;
    h.norb = m->norb;
    h.nelec = m->nelec;
    h.neleca = m->neleca;
    h.nelecb = m->nelecb;
    h.na = m->na;
    h.nb = m->nb;
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (f == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    bool ok = fwrite(&h, sizeof(h), 1, f) == 1 &&
              fwrite(m->col_ptr, sizeof(int), (size_t)m->n + 1, f) == (size_t)m->n + 1 &&
              fwrite(m->row_ind, sizeof(int), (size_t)m->nnz, f) == (size_t)m->nnz &&
              fwrite(m->values, sizeof(double), (size_t)m->nnz, f) == (size_t)m->nnz &&
              fwrite(m->diagonal, sizeof(double), (size_t)m->n, f) == (size_t)m->n && fflush(f) == 0;
    ok = (fclose(f) == 0) && ok;
    if (!ok || rename(tmp.c_str(), path) != 0) {
        remove(tmp.c_str());
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    return IPT_CUDA_SUCCESS;
}

extern "C" int ipt_hip_csc_cache_read(const char *path, IPTFciCsc *out)
{
    if (out == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    memset(out, 0, sizeof(*out));
    FILE *f = path ? fopen(path, "rb") : nullptr;
    if (f == nullptr) {
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    CacheHeader h;
    if (fread(&h, sizeof(h), 1, f) != 1 || memcmp(h.magic, kCacheMagic, sizeof(h.magic)) != 0 || h.version != 1 ||
        h.n <= 0 || h.nnz < 0 || !std::isfinite(h.matrix_inf_norm) || h.matrix_inf_norm < 0.0) {
        fclose(f);
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    out->col_ptr = (int *)malloc(sizeof(int) * ((size_t)h.n + 1));
    out->row_ind = (int *)malloc(sizeof(int) * (size_t)std::max(h.nnz, 1));
    out->values = (double *)malloc(sizeof(double) * (size_t)std::max(h.nnz, 1));
    out->diagonal = (double *)malloc(sizeof(double) * (size_t)h.n);
    bool ok = out->col_ptr && out->row_ind && out->values && out->diagonal;
    ok = ok && fread(out->col_ptr, sizeof(int), (size_t)h.n + 1, f) == (size_t)h.n + 1 &&
         fread(out->row_ind, sizeof(int), (size_t)h.nnz, f) == (size_t)h.nnz &&
         fread(out->values, sizeof(double), (size_t)h.nnz, f) == (size_t)h.nnz &&
         fread(out->diagonal, sizeof(double), (size_t)h.n, f) == (size_t)h.n;
    fclose(f);
    if (!ok || !csc_valid(out->col_ptr, out->row_ind, h.n, h.nnz)) {
        ipt_hip_fci_free_csc(out);
        return IPT_CUDA_INVALID_ARGUMENT;
    }
    out->n = h.n;
    out->nnz = h.nnz;
    out->matrix_inf_norm = h.matrix_inf_norm;
    out->norb = h.norb;
    out->nelec = h.nelec;
    out->neleca = h.neleca;
    out->nelecb = h.nelecb;
    out->na = h.na;
    out->nb = h.nb;
    return IPT_CUDA_SUCCESS;
}
