// Host-side preparation: CSC validation, diagonal sort / permutation,
// transposition and target-cluster detection.
#include "ipt_prep.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>

#include "ipt_internal.h"

namespace ipt {

DegeneracyOptions degeneracy_options_from_env(double degeneracy_threshold)
{
    // IPT/src/ipt_preparation.cu:126-146: negative / non-finite values fall
    // back to defaults; integer options must be positive.
    DegeneracyOptions o;
    o.absolute_gap = degeneracy_threshold > 0.0 ? degeneracy_threshold : 0.0;
    o.eps_multiplier = env_double("IPT_EPS_MULTIPLIER", 1000.0);
    o.rel_gap_tol = env_double("IPT_REL_GAP_TOL", 1.0e-12);
    o.ratio_tol = env_double("IPT_COUPLING_GAP_RATIO_TOL", 0.2);
    o.eta = env_double("IPT_COUPLING_ETA", 0.02);
    int mb = env_int("IPT_MAX_LIFT_BLOCK_SIZE", 64);
    o.max_block = mb > 0 ? mb : 64;
    return o;
}

bool csc_valid(const int *col_ptr, const int *row_ind, int n, int nnz)
{
    if (col_ptr == nullptr || row_ind == nullptr || n <= 0 || nnz < 0 || col_ptr[0] != 0 ||
        col_ptr[n] != nnz) {
        return false;
    }
    for (int c = 0; c < n; ++c) {
        if (col_ptr[c] > col_ptr[c + 1]) {
            return false;
        }
    }
    for (int p = 0; p < nnz; ++p) {
        if (row_ind[p] < 0 || row_ind[p] >= n) {
            return false;
        }
    }
    return true;
}

void csc_diagonal_host(const int *col_ptr, const int *row_ind, const double *values, int n,
                       std::vector<double> *diag)
{
    diag->assign((size_t)n, 0.0);
    for (int c = 0; c < n; ++c) {
        for (int p = col_ptr[c]; p < col_ptr[c + 1]; ++p) {
            if (row_ind[p] == c) {
                (*diag)[(size_t)c] = values[p];
                break;
            }
        }
    }
}

void stable_sort_permutation(const std::vector<double> &diag, std::vector<int> *perm,
                             std::vector<int> *inverse)
{
    const int n = (int)diag.size();
    perm->resize((size_t)n);
    inverse->resize((size_t)n);
    std::iota(perm->begin(), perm->end(), 0);
    // already sorted (the common case): identity, no sort
    if (!std::is_sorted(diag.begin(), diag.end())) {
        std::stable_sort(perm->begin(), perm->end(),
                         [&](int a, int b) { return diag[(size_t)a] < diag[(size_t)b]; });
    }
    for (int s = 0; s < n; ++s) {
        (*inverse)[(size_t)(*perm)[(size_t)s]] = s;
    }
}

void permute_csc(const int *col_ptr, const int *row_ind, const double *values, int n,
                 const std::vector<int> &inverse, HostCsc *out)
{
    // counting pass per new column, then sort each column by row and merge
    // duplicates
    std::vector<int> cnt((size_t)n + 1, 0);
    for (int oc = 0; oc < n; ++oc) {
        cnt[(size_t)inverse[(size_t)oc] + 1] += col_ptr[oc + 1] - col_ptr[oc];
    }
    for (int c = 0; c < n; ++c) {
        cnt[(size_t)c + 1] += cnt[(size_t)c];
    }
    std::vector<std::pair<int, double>> ent((size_t)cnt[(size_t)n]);
    std::vector<int> cur(cnt.begin(), cnt.end() - 1);
    for (int oc = 0; oc < n; ++oc) {
        const int nc = inverse[(size_t)oc];
        for (int p = col_ptr[oc]; p < col_ptr[oc + 1]; ++p) {
            ent[(size_t)cur[(size_t)nc]++] = std::make_pair(inverse[(size_t)row_ind[p]], values[p]);
        }
    }
    out->n = n;
    out->col_ptr.assign((size_t)n + 1, 0);
    out->row_ind.clear();
    out->values.clear();
    out->row_ind.reserve(ent.size());
    out->values.reserve(ent.size());
    for (int c = 0; c < n; ++c) {
        auto b = ent.begin() + cnt[(size_t)c], e = ent.begin() + cnt[(size_t)c + 1];
        std::stable_sort(b, e, [](const std::pair<int, double> &x, const std::pair<int, double> &y) {
            return x.first < y.first;
        });
        out->col_ptr[(size_t)c] = (int)out->row_ind.size();
        for (auto it = b; it != e; ++it) {
            const int row = it->first;
            double v = it->second;
            while (it + 1 != e && (it + 1)->first == row) {
                ++it;
                v += it->second;
            }
            if (v != 0.0) {
                out->row_ind.push_back(row);
                out->values.push_back(v);
            }
        }
    }
    out->col_ptr[(size_t)n] = (int)out->row_ind.size();
}

void csc_to_csr(const int *col_ptr, const int *row_ind, const double *values, int n,
                std::vector<int> *row_ptr, std::vector<int> *col_ind, std::vector<double> *vals)
{
    const int nnz = col_ptr[n];
    row_ptr->assign((size_t)n + 1, 0);
    col_ind->resize((size_t)nnz);
    vals->resize((size_t)nnz);
    for (int p = 0; p < nnz; ++p) {
        (*row_ptr)[(size_t)row_ind[p] + 1]++;
    }
    for (int r = 0; r < n; ++r) {
        (*row_ptr)[(size_t)r + 1] += (*row_ptr)[(size_t)r];
    }
    std::vector<int> cur(row_ptr->begin(), row_ptr->end() - 1);
    // columns visited in increasing order -> each row's entries sorted by column
    for (int c = 0; c < n; ++c) {
        for (int p = col_ptr[c]; p < col_ptr[c + 1]; ++p) {
            const int r = row_ind[p];
            const int q = cur[(size_t)r]++;
            (*col_ind)[(size_t)q] = c;
            (*vals)[(size_t)q] = values[p];
        }
    }
}

static double csc_lookup(const HostCsc &a, int row, int col)
{
    const int b = a.col_ptr[(size_t)col], e = a.col_ptr[(size_t)col + 1];
    // rows are sorted after permute_csc; fall back to a scan otherwise
    auto first = a.row_ind.begin() + b, last = a.row_ind.begin() + e;
    if (std::is_sorted(first, last)) {
        auto it = std::lower_bound(first, last, row);
        if (it != last && *it == row) {
            return a.values[(size_t)(it - a.row_ind.begin())];
        }
        return 0.0;
    }
    for (int p = b; p < e; ++p) {
        if (a.row_ind[(size_t)p] == row) {
            return a.values[(size_t)p];
        }
    }
    return 0.0;
}

bool csc_is_symmetric(const HostCsc &a)
{
    for (int c = 0; c < a.n; ++c) {
        for (int p = a.col_ptr[(size_t)c]; p < a.col_ptr[(size_t)c + 1]; ++p) {
            const int r = a.row_ind[(size_t)p];
            const double v = a.values[(size_t)p];
            const double t = csc_lookup(a, c, r);
            const double scale = std::max(1.0, std::max(std::fabs(v), std::fabs(t)));
            if (std::fabs(v - t) > 1.0e-12 * scale) {
                return false;
            }
        }
    }
    return true;
}

double csc_inf_norm(const int *col_ptr, const int *row_ind, const double *values, int n)
{
    std::vector<double> rs((size_t)n, 0.0);
    for (int c = 0; c < n; ++c) {
        for (int p = col_ptr[c]; p < col_ptr[c + 1]; ++p) {
            rs[(size_t)row_ind[p]] += std::fabs(values[p]);
        }
    }
    double m = 0.0;
    for (double x : rs) {
        m = std::max(m, x);
    }
    return m;
}

double CscCoupling::diag(int i) const
{
    for (int p = a->col_ptr[(size_t)i]; p < a->col_ptr[(size_t)i + 1]; ++p) {
        if (a->row_ind[(size_t)p] == i) {
            return a->values[(size_t)p];
        }
    }
    return 0.0;
}

double CscCoupling::value(int row, int col) const { return csc_lookup(*a, row, col); }

void CscCoupling::offdiag_strengths(std::vector<double> *s) const
{
    const int nn = a->n;
    std::vector<double> colm((size_t)nn, 0.0), rowm((size_t)nn, 0.0);
    for (int c = 0; c < nn; ++c) {
        for (int p = a->col_ptr[(size_t)c]; p < a->col_ptr[(size_t)c + 1]; ++p) {
            const int r = a->row_ind[(size_t)p];
            if (r == c) {
                continue;
            }
            const double v = std::fabs(a->values[(size_t)p]);
            colm[(size_t)c] = std::max(colm[(size_t)c], v);
            rowm[(size_t)r] = std::max(rowm[(size_t)r], v);
        }
    }
    s->resize((size_t)nn);
    for (int i = 0; i < nn; ++i) {
        (*s)[(size_t)i] = std::max(colm[(size_t)i], rowm[(size_t)i]);
    }
}

void DenseCoupling::offdiag_strengths(std::vector<double> *s) const
{
    std::vector<double> colm((size_t)nn, 0.0), rowm((size_t)nn, 0.0);
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nn; ++c) {
        double m = 0.0;
        const double *hc = h + (long)c * ld;
        for (int r = 0; r < nn; ++r) {
            if (r != c) {
                m = std::max(m, std::fabs(hc[r]));
            }
        }
        colm[(size_t)c] = m;
    }
    // row maxima: for symmetric input equal to column maxima; compute anyway
#pragma omp parallel for schedule(static)
    for (int r = 0; r < nn; ++r) {
        double m = 0.0;
        for (int c = 0; c < nn; ++c) {
            if (r != c) {
                m = std::max(m, std::fabs(h[(long)r + (long)c * ld]));
            }
        }
        rowm[(size_t)r] = m;
    }
    s->resize((size_t)nn);
    for (int i = 0; i < nn; ++i) {
        (*s)[(size_t)i] = std::max(colm[(size_t)i], rowm[(size_t)i]);
    }
}

namespace {

struct UnionFind {
    std::vector<int> parent, rank;
    explicit UnionFind(int n) : parent((size_t)n), rank((size_t)n, 0)
    {
        std::iota(parent.begin(), parent.end(), 0);
    }
    int find(int x)
    {
        while (parent[(size_t)x] != x) {
            parent[(size_t)x] = parent[(size_t)parent[(size_t)x]];
            x = parent[(size_t)x];
        }
        return x;
    }
    void unite(int a, int b)
    {
        a = find(a);
        b = find(b);
        if (a == b) {
            return;
        }
        if (rank[(size_t)a] < rank[(size_t)b]) {
            std::swap(a, b);
        }
        parent[(size_t)b] = a;
        if (rank[(size_t)a] == rank[(size_t)b]) {
            ++rank[(size_t)a];
        }
    }
};

double gap_floor(double a, double b, double tau_num, const DegeneracyOptions &o)
{
    const double scale = std::max(1.0, std::max(std::fabs(a), std::fabs(b)));
    return std::max(std::max(o.absolute_gap, tau_num), o.rel_gap_tol * scale);
}

}  // namespace

std::vector<std::pair<int, int>> target_clusters(const CouplingSource &m, const std::vector<double> &diag,
                                                 int k, const DegeneracyOptions &o, bool *too_large)
{
    const int n = m.n();
    const double eps = DBL_EPSILON;
    double maxabs = 0.0;
    for (double x : diag) {
        maxabs = std::max(maxabs, std::fabs(x));
    }
    const double tau_num = o.eps_multiplier * eps * std::max(1.0, maxabs);
    const double ratio_tol = std::max(o.ratio_tol, eps);
    std::vector<double> off;
    m.offdiag_strengths(&off);

    std::vector<int> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (diag[(size_t)a] != diag[(size_t)b]) {
            return diag[(size_t)a] < diag[(size_t)b];
        }
        return a < b;
    });
    std::vector<double> sd((size_t)n);
    for (int i = 0; i < n; ++i) {
        sd[(size_t)i] = diag[(size_t)order[(size_t)i]];
    }

    UnionFind uf(n);
    for (int col = 0; col < k; ++col) {
        const double dc = diag[(size_t)col];
        const double tau_c = std::max(gap_floor(dc, dc, tau_num, o), off[(size_t)col] / ratio_tol);
        const int lo = (int)(std::lower_bound(sd.begin(), sd.end(), dc - tau_c) - sd.begin());
        const int hi = (int)(std::upper_bound(sd.begin(), sd.end(), dc + tau_c) - sd.begin());
        for (int pos = lo; pos < hi; ++pos) {
            const int row = order[(size_t)pos];
            if (row == col) {
                continue;
            }
            const double dr = diag[(size_t)row];
            double coupling = std::max(std::fabs(m.value(row, col)), std::fabs(m.value(col, row)));
            if (!(coupling > 0.0)) {
                coupling = o.eta * std::min(off[(size_t)row], off[(size_t)col]);
            }
            const double gap = std::fabs(dr - dc);
            const double fl = gap_floor(dr, dc, tau_num, o);
            const double ratio = coupling > 0.0 ? coupling / std::max(gap, fl) : 0.0;
            if (gap <= fl || (coupling > 0.0 && ratio >= ratio_tol)) {
                uf.unite(row, col);
            }
        }
    }

    std::vector<int> cmin((size_t)n, n), cmax((size_t)n, -1);
    std::vector<char> has((size_t)n, 0);
    for (int i = 0; i < n; ++i) {
        const int r = uf.find(i);
        cmin[(size_t)r] = std::min(cmin[(size_t)r], i);
        cmax[(size_t)r] = std::max(cmax[(size_t)r], i);
        if (i < k) {
            has[(size_t)r] = 1;
        }
    }
    std::vector<std::pair<int, int>> ranges;
    for (int r = 0; r < n; ++r) {
        if (has[(size_t)r] && cmin[(size_t)r] < cmax[(size_t)r]) {
            ranges.emplace_back(cmin[(size_t)r], cmax[(size_t)r]);
        }
    }
    std::sort(ranges.begin(), ranges.end());
    std::vector<std::pair<int, int>> merged;
    for (const auto &r : ranges) {
        if (!merged.empty() && r.first <= merged.back().second + 1) {
            merged.back().second = std::max(merged.back().second, r.second);
        } else {
            merged.push_back(r);
        }
    }
    *too_large = false;
    std::vector<char> covered((size_t)n, 0);
    for (const auto &r : merged) {
        if (r.second - r.first + 1 > o.max_block) {
            *too_large = true;
        }
        for (int i = r.first; i <= r.second; ++i) {
            covered[(size_t)i] = 1;
        }
    }
    for (int t = 0; t < k; ++t) {
        if (!covered[(size_t)t]) {
            merged.emplace_back(t, t);
        }
    }
    std::sort(merged.begin(), merged.end());
    return merged;
}

}  // namespace ipt
