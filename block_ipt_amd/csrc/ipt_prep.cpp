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
    int mr = env_int("IPT_MAX_LIFT_ROUNDS", 3);
    o.max_lift_rounds = mr > 0 ? mr : 3;
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

namespace {

// Pairs (row, col) of the cluster rule for target column col: rows whose
// diagonal lies within the target window (IPT/src/ipt_preparation.cu:426-441)
// and that the pairwise test joins (443-478).
template <typename F>
void for_joined_pairs(const CouplingSource &m, const std::vector<double> &diag, int k, const DegeneracyOptions &o,
                      F &&visit)
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
    // a sorted diagonal (the block path sorts it first) is its own order
    if (!std::is_sorted(diag.begin(), diag.end())) {
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            if (diag[(size_t)a] != diag[(size_t)b]) {
                return diag[(size_t)a] < diag[(size_t)b];
            }
            return a < b;
        });
    }
    std::vector<double> sd((size_t)n);
    for (int i = 0; i < n; ++i) {
        sd[(size_t)i] = diag[(size_t)order[(size_t)i]];
    }
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
                visit(row, col);
            }
        }
    }
}

}  // namespace

int target_gap_violations(const CouplingSource &m, const std::vector<double> &diag, int k,
                          const DegeneracyOptions &o)
{
    int bad = 0;
    for_joined_pairs(m, diag, k, o, [&](int, int) { ++bad; });
    return bad;
}

std::vector<std::pair<int, int>> target_clusters(const CouplingSource &m, const std::vector<double> &diag,
                                                 int k, const DegeneracyOptions &o, bool *too_large,
                                                 bool add_targets)
{
    const int n = m.n();
    UnionFind uf(n);
    for_joined_pairs(m, diag, k, o, [&](int row, int col) { uf.unite(row, col); });

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
    if (add_targets) {
        for (int t = 0; t < k; ++t) {
            if (!covered[(size_t)t]) {
                merged.emplace_back(t, t);
            }
        }
        std::sort(merged.begin(), merged.end());
    }
    return merged;
}

void sym_block_eigenvectors(const std::vector<double> &a, int m, std::vector<double> *v)
{
    // Classical Jacobi: rotate away the largest off-diagonal entry until
    // all are <= 1e-14 (or 64 m^2 rotations); the same rotation formulas,
    // stop rule and ordering as the reference, so that the lifted matrix
    // and the vectors' normalisation agree with it.
    std::vector<double> w = a;
    v->assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) {
        (*v)[(size_t)i * m + i] = 1.0;
    }
    auto at = [m](std::vector<double> &x, int r, int c) -> double & { return x[(size_t)r + (size_t)c * m]; };
    const int max_rot = 64 * m * m;
    for (int it = 0; it < max_rot; ++it) {
        int p = 0, q = 1;
        double big = 0.0;
        for (int c = 1; c < m; ++c) {
            for (int r = 0; r < c; ++r) {
                const double x = std::fabs(at(w, r, c));
                if (x > big) {
                    big = x;
                    p = r;
                    q = c;
                }
            }
        }
        if (big <= 1.0e-14) {
            break;
        }
        const double app = at(w, p, p), aqq = at(w, q, q), apq = at(w, p, q);
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t);
        const double sn = t * c;
        for (int r = 0; r < m; ++r) {
            if (r == p || r == q) {
                continue;
            }
            const double arp = at(w, r, p), arq = at(w, r, q);
            const double np = c * arp - sn * arq, nq = sn * arp + c * arq;
            at(w, r, p) = np;
            at(w, p, r) = np;
            at(w, r, q) = nq;
            at(w, q, r) = nq;
        }
        at(w, p, p) = c * c * app - 2.0 * sn * c * apq + sn * sn * aqq;
        at(w, q, q) = sn * sn * app + 2.0 * sn * c * apq + c * c * aqq;
        at(w, p, q) = 0.0;
        at(w, q, p) = 0.0;
        for (int r = 0; r < m; ++r) {
            const double vp = at(*v, r, p), vq = at(*v, r, q);
            at(*v, r, p) = c * vp - sn * vq;
            at(*v, r, q) = sn * vp + c * vq;
        }
    }
    // order by the Rayleigh quotients v^T a v (stable)
    std::vector<double> lam((size_t)m, 0.0);
    for (int j = 0; j < m; ++j) {
        double x = 0.0;
        for (int r = 0; r < m; ++r) {
            for (int u = 0; u < m; ++u) {
                x += (*v)[(size_t)r + (size_t)j * m] * a[(size_t)r + (size_t)u * m] * (*v)[(size_t)u + (size_t)j * m];
            }
        }
        lam[(size_t)j] = x;
    }
    std::vector<int> order((size_t)m);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return lam[(size_t)x] < lam[(size_t)y]; });
    std::vector<double> sorted((size_t)m * m);
    for (int j = 0; j < m; ++j) {
        std::copy(v->begin() + (size_t)order[(size_t)j] * m, v->begin() + (size_t)(order[(size_t)j] + 1) * m,
                  sorted.begin() + (size_t)j * m);
    }
    v->swap(sorted);
}

namespace {

// value at (row, col), duplicates summed (IPT/src/ipt_preparation.cu:830-844)
double csc_value_sum(const HostCsc &a, int row, int col)
{
    double v = 0.0;
    for (int p = a.col_ptr[(size_t)col]; p < a.col_ptr[(size_t)col + 1]; ++p) {
        if (a.row_ind[(size_t)p] == row) {
            v += a.values[(size_t)p];
        }
    }
    return v;
}

void csc_diagonal_sum(const HostCsc &a, std::vector<double> *d)
{
    d->assign((size_t)a.n, 0.0);
    for (int i = 0; i < a.n; ++i) {
        (*d)[(size_t)i] = csc_value_sum(a, i, i);
    }
}

// A <- S^-1 A S for S = I except on the block B = [first, last], where
// S|B = R and S^-1|B = L (here L = R^T: R is orthogonal).  Only entries with
// a row or a column in B change:
//     rows outside B, columns in B:   A(o, B)  <- A(o, B) R
//     rows in B, columns outside B:   A(B, o)  <- L A(B, o)
//     the block itself:               A(B, B)  <- L A(B, B) R
// The B columns and the entries of B rows are gathered into dense m-vectors
// (duplicate entries summed), transformed by the small dense products
// above, and every column is rebuilt once, rows ascending, duplicates
// summed and exact zeros dropped.  The reference applies the same
// similarity entry by entry (IPT/src/ipt_preparation.cu:1109-1220); sums are
// formed in another order, so results agree to rounding.
void local_similarity(HostCsc *a, int first, int last, const std::vector<double> &right,
                      const std::vector<double> &left)
{
    const int n = a->n, m = last - first + 1;
    auto inb = [&](int i) { return i >= first && i <= last; };
    // dense block A(B, B), column-major m x m
    std::vector<double> blk((size_t)m * m, 0.0);
    // A(o, B) by outside row: slot per row, m values each
    std::vector<int> row_slot((size_t)n, -1), slot_row;
    std::vector<double> rowv;
    for (int j = 0; j < m; ++j) {
        const int col = first + j;
        for (int p = a->col_ptr[(size_t)col]; p < a->col_ptr[(size_t)col + 1]; ++p) {
            const int row = a->row_ind[(size_t)p];
            if (inb(row)) {
                blk[(size_t)(row - first) + (size_t)j * m] += a->values[(size_t)p];
                continue;
            }
            if (row_slot[(size_t)row] < 0) {
                row_slot[(size_t)row] = (int)slot_row.size();
                slot_row.push_back(row);
                rowv.resize(rowv.size() + (size_t)m, 0.0);
            }
            rowv[(size_t)row_slot[(size_t)row] * m + j] += a->values[(size_t)p];
        }
    }
    // A(B, o) by outside column: slot per column, m values each
    std::vector<int> col_slot((size_t)n, -1);
    std::vector<double> colv;
    int ncs = 0;
    for (int col = 0; col < n; ++col) {
        if (inb(col)) {
            continue;
        }
        for (int p = a->col_ptr[(size_t)col]; p < a->col_ptr[(size_t)col + 1]; ++p) {
            const int row = a->row_ind[(size_t)p];
            if (!inb(row)) {
                continue;
            }
            if (col_slot[(size_t)col] < 0) {
                col_slot[(size_t)col] = ncs++;
                colv.resize(colv.size() + (size_t)m, 0.0);
            }
            colv[(size_t)col_slot[(size_t)col] * m + (row - first)] += a->values[(size_t)p];
        }
    }
    // the three dense products (left and right are column-major m x m)
    auto mat = [m](const std::vector<double> &x, int i, int j) { return x[(size_t)i + (size_t)j * m]; };
    std::vector<double> tmp((size_t)m * m, 0.0), nblk((size_t)m * m, 0.0);
    for (int j = 0; j < m; ++j) {  // tmp = A(B,B) R
        for (int l = 0; l < m; ++l) {
            const double r = mat(right, l, j);
            for (int i = 0; i < m; ++i) {
                tmp[(size_t)i + (size_t)j * m] += blk[(size_t)i + (size_t)l * m] * r;
            }
        }
    }
    for (int j = 0; j < m; ++j) {  // nblk = L tmp
        for (int l = 0; l < m; ++l) {
            const double t = tmp[(size_t)l + (size_t)j * m];
            for (int i = 0; i < m; ++i) {
                nblk[(size_t)i + (size_t)j * m] += mat(left, i, l) * t;
            }
        }
    }
    std::vector<double> nrow(rowv.size(), 0.0), ncol(colv.size(), 0.0);
    for (size_t s = 0; s < slot_row.size(); ++s) {  // row vector times R
        const double *x = rowv.data() + s * m;
        for (int j = 0; j < m; ++j) {
            double acc = 0.0;
            for (int l = 0; l < m; ++l) {
                acc += x[l] * mat(right, l, j);
            }
            nrow[s * m + (size_t)j] = acc;
        }
    }
    for (int s = 0; s < ncs; ++s) {  // L times column vector
        const double *u = colv.data() + (size_t)s * m;
        for (int i = 0; i < m; ++i) {
            double acc = 0.0;
            for (int l = 0; l < m; ++l) {
                acc += mat(left, i, l) * u[l];
            }
            ncol[(size_t)s * m + (size_t)i] = acc;
        }
    }
    // rebuild every column once
    HostCsc out;
    out.n = n;
    out.col_ptr.assign((size_t)n + 1, 0);
    std::vector<std::pair<int, double>> e;
    for (int col = 0; col < n; ++col) {
        e.clear();
        if (inb(col)) {
            const int j = col - first;
            for (size_t s = 0; s < slot_row.size(); ++s) {
                e.emplace_back(slot_row[s], nrow[s * m + (size_t)j]);
            }
            for (int i = 0; i < m; ++i) {
                e.emplace_back(first + i, nblk[(size_t)i + (size_t)j * m]);
            }
        } else {
            for (int p = a->col_ptr[(size_t)col]; p < a->col_ptr[(size_t)col + 1]; ++p) {
                if (!inb(a->row_ind[(size_t)p])) {
                    e.emplace_back(a->row_ind[(size_t)p], a->values[(size_t)p]);
                }
            }
            if (col_slot[(size_t)col] >= 0) {
                const double *u = ncol.data() + (size_t)col_slot[(size_t)col] * m;
                for (int i = 0; i < m; ++i) {
                    e.emplace_back(first + i, u[i]);
                }
            }
        }
        std::sort(e.begin(), e.end(), [](const std::pair<int, double> &x, const std::pair<int, double> &y) {
            return x.first < y.first;
        });
        out.col_ptr[(size_t)col] = (int)out.row_ind.size();
        for (size_t q = 0; q < e.size();) {
            const int row = e[q].first;
            double v = 0.0;
            for (; q < e.size() && e[q].first == row; ++q) {
                v += e[q].second;
            }
            if (v != 0.0) {
                out.row_ind.push_back(row);
                out.values.push_back(v);
            }
        }
    }
    out.col_ptr[(size_t)n] = (int)out.row_ind.size();
    *a = std::move(out);
}

}  // namespace

int prepare_lifted_csc(const int *col_ptr, const int *row_ind, const double *values, int n, int k, int nnz,
                       double degeneracy_threshold, LiftedCsc *out)
{
    out->perm.clear();
    out->rot_first.clear();
    out->rot_m.clear();
    out->rot_q.clear();
    std::vector<double> dg;
    csc_diagonal_host(col_ptr, row_ind, values, n, &dg);
    std::vector<int> perm, inv;
    stable_sort_permutation(dg, &perm, &inv);
    bool permuted = false;
    for (int i = 0; i < n; ++i) {
        if (perm[(size_t)i] != i) {
            permuted = true;
            break;
        }
    }
    if (permuted) {
        permute_csc(col_ptr, row_ind, values, n, inv, &out->a);
        out->perm = perm;
    } else {
        out->a.n = n;
        out->a.col_ptr.assign(col_ptr, col_ptr + (size_t)n + 1);
        out->a.row_ind.assign(row_ind, row_ind + (size_t)nnz);
        out->a.values.assign(values, values + (size_t)nnz);
    }
    const DegeneracyOptions o = degeneracy_options_from_env(degeneracy_threshold);
    const bool hermitian = csc_is_symmetric(out->a);
    std::vector<double> diag;
    bool clean = false;
    for (int round = 0; round < o.max_lift_rounds; ++round) {
        csc_diagonal_sum(out->a, &diag);
        const CscCoupling cm(&out->a);
        bool too_large = false;
        const auto subspaces = target_clusters(cm, diag, k, o, &too_large, false);
        if (too_large) {
            return IPT_CUDA_INVALID_ARGUMENT;
        }
        if (subspaces.empty()) {
            clean = target_gap_violations(cm, diag, k, o) == 0;
            break;
        }
        if (!hermitian) {
            // the reference lifts a non-symmetric block with LAPACK dgeev
            // (and fails where that is absent); not built here
            return IPT_CUDA_INVALID_ARGUMENT;
        }
        for (const auto &sp : subspaces) {
            const int first = sp.first, m = sp.second - sp.first + 1;
            std::vector<double> block((size_t)m * m, 0.0);
            for (int c = 0; c < m; ++c) {
                const int gc = first + c;
                for (int p = out->a.col_ptr[(size_t)gc]; p < out->a.col_ptr[(size_t)gc + 1]; ++p) {
                    const int gr = out->a.row_ind[(size_t)p];
                    if (gr >= first && gr < first + m) {
                        block[(size_t)(gr - first) + (size_t)c * m] += out->a.values[(size_t)p];
                    }
                }
            }
            std::vector<double> q, qt((size_t)m * m);
            sym_block_eigenvectors(block, m, &q);
            for (int c = 0; c < m; ++c) {
                for (int r = 0; r < m; ++r) {
                    qt[(size_t)r + (size_t)c * m] = q[(size_t)c + (size_t)r * m];
                }
            }
            local_similarity(&out->a, first, sp.second, q, qt);
            out->rot_first.push_back(first);
            out->rot_m.push_back(m);
            out->rot_q.insert(out->rot_q.end(), q.begin(), q.end());
        }
        csc_diagonal_sum(out->a, &diag);
        const CscCoupling cm2(&out->a);
        if (target_gap_violations(cm2, diag, k, o) == 0) {
            clean = true;
            break;
        }
    }
    if (!clean) {
        csc_diagonal_sum(out->a, &diag);
        const CscCoupling cm(&out->a);
        if (target_gap_violations(cm, diag, k, o) > 0) {
            return IPT_CUDA_INVALID_ARGUMENT;
        }
    }
    return IPT_CUDA_SUCCESS;
}

void lifted_to_original(const LiftedCsc &p, double *vectors, int k)
{
    const int n = p.a.n;
    size_t off = p.rot_q.size();
    std::vector<double> tmp;
    for (int r = (int)p.rot_first.size() - 1; r >= 0; --r) {
        const int first = p.rot_first[(size_t)r], m = p.rot_m[(size_t)r];
        off -= (size_t)m * m;
        const double *q = p.rot_q.data() + off;
        tmp.assign((size_t)m, 0.0);
        for (int c = 0; c < k; ++c) {
            double *x = vectors + (size_t)c * n + first;
            for (int i = 0; i < m; ++i) {
                double sum = 0.0;
                for (int u = 0; u < m; ++u) {
                    sum += q[(size_t)i + (size_t)u * m] * x[u];
                }
                tmp[(size_t)i] = sum;
            }
            std::copy(tmp.begin(), tmp.end(), x);
        }
    }
    if (!p.perm.empty()) {
        tmp.assign((size_t)n, 0.0);
        for (int c = 0; c < k; ++c) {
            double *x = vectors + (size_t)c * n;
            for (int s = 0; s < n; ++s) {
                tmp[(size_t)p.perm[(size_t)s]] = x[s];
            }
            std::copy(tmp.begin(), tmp.end(), x);
        }
    }
}

BlockTargets block_targets_from_env(int n, int k)
{
    BlockTargets t;
    t.adaptive = env_flag("IPT_BLOCK_CLUSTER_ADAPTIVE") != 0;
    t.tau = env_double("IPT_BLOCK_CLUSTER_COUPLING_TAU", 0.1);
    t.oversample = env_int("IPT_BLOCK_CLUSTER_OVERSAMPLE", 0);
    if (t.adaptive || t.oversample < 0) {
        t.oversample = 0;
    }
    t.target_k = k;
    if (t.oversample > 0) {
        t.target_k = t.oversample > n - k ? n : k + t.oversample;
    }
    return t;
}

bool adaptive_expand(const HostCsc &a, const std::vector<double> &d, int k, int max_block, double tau,
                     std::vector<std::pair<int, int>> *ranges, std::vector<int> *added)
{
    const double eps_gap = 1.0e-14;
    const int n = (int)d.size();
    bool limit_hit = false;
    added->clear();
    for (;;) {
        int id = -1;
        for (size_t i = 0; i < ranges->size(); ++i) {
            if ((*ranges)[i].first <= k - 1 && k - 1 <= (*ranges)[i].second) {
                id = (int)i;
                break;
            }
        }
        if (id < 0) {
            break;
        }
        const int first = (*ranges)[(size_t)id].first, last = (*ranges)[(size_t)id].second;
        int best = -1;
        double best_ratio = 0.0;
        for (int col = first; col <= last; ++col) {
            for (int p = a.col_ptr[(size_t)col]; p < a.col_ptr[(size_t)col + 1]; ++p) {
                const int row = a.row_ind[(size_t)p];
                if (row >= first && row <= last) {
                    continue;
                }
                const double ratio =
                    std::fabs(a.values[(size_t)p]) / std::max(std::fabs(d[(size_t)col] - d[(size_t)row]), eps_gap);
                if (std::max(last, row) - std::min(first, row) + 1 > max_block) {
                    if (ratio > tau) {
                        limit_hit = true;
                    }
                    continue;
                }
                if (ratio > best_ratio) {
                    best_ratio = ratio;
                    best = row;
                }
            }
        }
        if (best < 0 || best_ratio <= tau) {
            break;
        }
        int lo = std::min(first, best), hi = std::max(last, best);
        // rows whose diagonal equals a boundary row's to 1000 eps join too
        auto same = [&](int i, int j) {
            const double scale = std::max(1.0, std::max(std::fabs(d[(size_t)i]), std::fabs(d[(size_t)j])));
            return std::fabs(d[(size_t)i] - d[(size_t)j]) <= std::max(eps_gap, 1000.0 * DBL_EPSILON * scale);
        };
        for (bool grew = true; grew;) {
            grew = false;
            if (lo > 0 && same(lo, lo - 1)) {
                --lo;
                grew = true;
            }
            if (hi + 1 < n && same(hi + 1, hi)) {
                ++hi;
                grew = true;
            }
        }
        if (hi - lo + 1 > max_block) {
            limit_hit = true;
            break;
        }
        for (int i = lo; i <= hi; ++i) {
            if (i < first || i > last) {
                added->push_back(i);
            }
        }
        (*ranges)[(size_t)id] = {lo, hi};
        std::sort(ranges->begin(), ranges->end());
        std::vector<std::pair<int, int>> merged;
        for (const auto &r : *ranges) {
            if (merged.empty() || r.first > merged.back().second) {
                merged.push_back(r);
            } else {
                merged.back().second = std::max(merged.back().second, r.second);
            }
        }
        ranges->swap(merged);
    }
    std::sort(added->begin(), added->end());
    added->erase(std::unique(added->begin(), added->end()), added->end());
    return limit_hit;
}

}  // namespace ipt
