// Small dense symmetric-definite generalized eigenproblem on the host:
//     P c = theta G c,   G symmetric positive definite, P symmetric,
// for the Rayleigh-Ritz projections of the block path (a few dozen to a
// few hundred columns).  Same method as LAPACK dsygv / the reference's
// cusolverDnDsygvd call (IPT/src/ipt_cuda.cu:3223-3226): Cholesky G = L L^T,
// standard problem A = L^-1 P L^-T, A = Z diag(theta) Z^T, C = L^-T Z, so
// that C^T G C = I.  The standard problem is solved by Householder
// tridiagonalisation followed by the implicit QL iteration with shifts
// (the EISPACK tred2/tql2 method), eigenvalues returned ascending.
//
// Header-only so a CPU test can compile it directly.
#pragma once

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace ipt {

namespace small_eig_detail {

// In place: G (m x m, column-major) -> lower Cholesky factor L (upper
// triangle zeroed).  False if G is not numerically positive definite.
// Left-looking, every inner loop runs down a contiguous column.
inline bool cholesky_lower(double *G, int m)
{
    for (int j = 0; j < m; ++j) {
        double *cj = G + (size_t)j * m;
        for (int k = 0; k < j; ++k) {
            const double *ck = G + (size_t)k * m;
            const double ljk = ck[j];
            for (int i = j; i < m; ++i) {
                cj[i] -= ljk * ck[i];
            }
        }
        const double djj = cj[j];
        if (!(djj > 0.0) || !std::isfinite(djj)) {
            return false;
        }
        const double ljj = std::sqrt(djj);
        const double inv = 1.0 / ljj;
        cj[j] = ljj;
        for (int i = j + 1; i < m; ++i) {
            cj[i] *= inv;
        }
        for (int i = 0; i < j; ++i) {
            cj[i] = 0.0;
        }
    }
    return true;
}

// X <- L^-1 X for the m x m column-major X (column-oriented forward
// substitution: contiguous axpys down the columns of L).
inline void forward_substitute(const double *L, double *X, int m)
{
    for (int c = 0; c < m; ++c) {
        double *x = X + (size_t)c * m;
        for (int k = 0; k < m; ++k) {
            const double *lk = L + (size_t)k * m;
            const double xk = x[k] / lk[k];
            x[k] = xk;
            for (int i = k + 1; i < m; ++i) {
                x[i] -= xk * lk[i];
            }
        }
    }
}

// A <- L^-1 A L^-T for symmetric A: with X = L^-1 A, the result is
// L^-1 X^T (A = A^T), i.e. two forward substitutions around a transpose.
inline void reduce_to_standard(const double *L, double *A, int m)
{
    forward_substitute(L, A, m);
    for (int j = 0; j < m; ++j) {
        for (int i = j + 1; i < m; ++i) {
            std::swap(A[i + (size_t)j * m], A[j + (size_t)i * m]);
        }
    }
    forward_substitute(L, A, m);
    // symmetrise the rounding
    for (int j = 0; j < m; ++j) {
        for (int i = j + 1; i < m; ++i) {
            const double v = 0.5 * (A[i + (size_t)j * m] + A[j + (size_t)i * m]);
            A[i + (size_t)j * m] = v;
            A[j + (size_t)i * m] = v;
        }
    }
}

// Householder tridiagonalisation A = Q T Q^T.  On return d = diag(T),
// e[i] = T(i+1, i) (e[m-1] = 0) and Q (m x m, column-major) overwrites A.
inline void tridiagonalize(double *A, int m, double *d, double *e)
{
    std::vector<double> Q((size_t)m * m, 0.0), v((size_t)m), p((size_t)m), w((size_t)m);
    for (int i = 0; i < m; ++i) {
        Q[i + (size_t)i * m] = 1.0;
    }
    for (int k = 0; k + 2 < m; ++k) {
        // x = A(k+1:m, k)
        double sigma = 0.0;
        for (int i = k + 2; i < m; ++i) {
            sigma += A[i + (size_t)k * m] * A[i + (size_t)k * m];
        }
        const double alpha = A[(k + 1) + (size_t)k * m];
        if (sigma == 0.0) {
            continue;
        }
        const double nrm = std::sqrt(alpha * alpha + sigma);
        const double beta = alpha <= 0.0 ? nrm : -nrm;  // new subdiagonal
        const double tau = (beta - alpha) / beta;
        const double scal = 1.0 / (alpha - beta);
        std::fill(v.begin(), v.end(), 0.0);
        v[(size_t)k + 1] = 1.0;
        for (int i = k + 2; i < m; ++i) {
            v[(size_t)i] = A[i + (size_t)k * m] * scal;
        }
        // trailing block B = A(k+1:m, k+1:m): B <- H B H, H = I - tau v v^T
        // p = tau B v, w = p - (tau/2)(p^T v) v, B <- B - v w^T - w v^T
        for (int i = k + 1; i < m; ++i) {
            p[(size_t)i] = 0.0;
        }
        for (int j = k + 1; j < m; ++j) {
            const double vj = tau * v[(size_t)j];
            const double *aj = A + (size_t)j * m;
            for (int i = k + 1; i < m; ++i) {
                p[(size_t)i] += aj[i] * vj;
            }
        }
        double pv = 0.0;
        for (int i = k + 1; i < m; ++i) {
            pv += p[(size_t)i] * v[(size_t)i];
        }
        for (int i = k + 1; i < m; ++i) {
            w[(size_t)i] = p[(size_t)i] - 0.5 * tau * pv * v[(size_t)i];
        }
        for (int j = k + 1; j < m; ++j) {
            for (int i = k + 1; i < m; ++i) {
                A[i + (size_t)j * m] -= v[(size_t)i] * w[(size_t)j] + w[(size_t)i] * v[(size_t)j];
            }
        }
        A[(k + 1) + (size_t)k * m] = beta;
        A[k + (size_t)(k + 1) * m] = beta;
        for (int i = k + 2; i < m; ++i) {
            A[i + (size_t)k * m] = 0.0;
            A[k + (size_t)i * m] = 0.0;
        }
        // Q <- Q H = Q - tau (Q v) v^T
        std::fill(p.begin(), p.end(), 0.0);
        for (int j = k + 1; j < m; ++j) {
            const double vj = tau * v[(size_t)j];
            const double *qj = Q.data() + (size_t)j * m;
            for (int r = 0; r < m; ++r) {
                p[(size_t)r] += qj[r] * vj;
            }
        }
        for (int j = k + 1; j < m; ++j) {
            const double vj = v[(size_t)j];
            double *qj = Q.data() + (size_t)j * m;
            for (int r = 0; r < m; ++r) {
                qj[r] -= p[(size_t)r] * vj;
            }
        }
    }
    for (int i = 0; i < m; ++i) {
        d[i] = A[i + (size_t)i * m];
        e[i] = (i + 1 < m) ? A[(i + 1) + (size_t)i * m] : 0.0;
    }
    std::copy(Q.begin(), Q.end(), A);
}

// Implicit QL with shifts on the symmetric tridiagonal (d, e), applying the
// rotations to the columns of Z (m x m, column-major).  False if an
// eigenvalue needs more than 60 sweeps.
inline bool tridiagonal_ql(double *d, double *e, double *Z, int m)
{
    const double eps = 2.220446049250313e-16;
    for (int l = 0; l < m; ++l) {
        int iter = 0;
        for (;;) {
            int mm = l;
            for (; mm + 1 < m; ++mm) {
                const double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
                if (std::fabs(e[mm]) <= eps * dd) {
                    break;
                }
            }
            if (mm == l) {
                break;
            }
            if (++iter > 60) {
                return false;
            }
            // shift from the leading 2 x 2 block
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::sqrt(g * g + 1.0);
            g = d[mm] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = mm - 1;
            bool underflow = false;
            for (; i >= l; --i) {
                const double f = s * e[i];
                const double b = c * e[i];
                r = std::sqrt(f * f + g * g);  // magnitudes bounded by ||A||: no overflow
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[mm] = 0.0;
                    underflow = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                double *zi = Z + (size_t)i * m;
                double *zi1 = Z + (size_t)(i + 1) * m;
                for (int k = 0; k < m; ++k) {
                    const double t = zi1[k];
                    zi1[k] = s * zi[k] + c * t;
                    zi[k] = c * zi[k] - s * t;
                }
            }
            if (underflow) {
                continue;
            }
            d[l] -= p;
            e[l] = g;
            e[mm] = 0.0;
        }
    }
    return true;
}

}  // namespace small_eig_detail

// theta[m] ascending, C (m x m, column-major) with C^T G C = I.  G and P
// are column-major m x m and are not modified.  Returns 0, or 1 if G is not
// positive definite, 2 if the QL iteration did not converge.
inline int sym_generalized_eig(int m, const double *G, const double *P, double *theta, double *C)
{
    using namespace small_eig_detail;
    std::vector<double> L(G, G + (size_t)m * m), A(P, P + (size_t)m * m), d((size_t)m), e((size_t)m);
    if (!cholesky_lower(L.data(), m)) {
        return 1;
    }
    reduce_to_standard(L.data(), A.data(), m);
    tridiagonalize(A.data(), m, d.data(), e.data());  // A now holds Q
    if (!tridiagonal_ql(d.data(), e.data(), A.data(), m)) {
        return 2;
    }
    std::vector<int> order((size_t)m);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return d[(size_t)a] < d[(size_t)b]; });
    // C = L^-T Z (back substitution with L^T per column), sorted; the dot
    // product of row i of L^T with c runs down column i of L (contiguous)
    for (int j = 0; j < m; ++j) {
        const int src = order[(size_t)j];
        theta[j] = d[(size_t)src];
        double *c = C + (size_t)j * m;
        const double *z = A.data() + (size_t)src * m;
        for (int i = m - 1; i >= 0; --i) {
            const double *li = L.data() + (size_t)i * m;
            double s = z[i];
            for (int k = i + 1; k < m; ++k) {
                s -= li[k] * c[k];
            }
            c[i] = s / li[i];
        }
    }
    return 0;
}

// As sym_generalized_eig, but a Gram matrix G that is not numerically
// positive definite (a nearly dependent basis) is handled by canonical
// orthogonalisation: G = W diag(g) W^T, the directions with
// g <= 1e-13 max(g) are dropped, B = W_kept diag(g_kept)^-1/2, and
// (B^T P B) z = theta z gives C = B Z.  The first `need` columns of C and
// entries of theta are defined (the rest are zero / +inf).  Returns 0, 1 if
// fewer than `need` directions survive, 2 if an eigenvalue iteration did
// not converge.
inline int sym_generalized_eig_robust(int m, const double *G, const double *P, double *theta, double *C, int need)
{
    using namespace small_eig_detail;
    const int st = sym_generalized_eig(m, G, P, theta, C);
    if (st != 1) {
        return st;
    }
    std::vector<double> W(G, G + (size_t)m * m), g((size_t)m), e((size_t)m);
    tridiagonalize(W.data(), m, g.data(), e.data());
    if (!tridiagonal_ql(g.data(), e.data(), W.data(), m)) {
        return 2;
    }
    double gmax = 0.0;
    for (double x : g) {
        gmax = std::max(gmax, x);
    }
    std::vector<int> keep;
    for (int i = 0; i < m; ++i) {
        if (g[(size_t)i] > 1e-13 * gmax) {
            keep.push_back(i);
        }
    }
    const int r = (int)keep.size();
    if (r < need || !(gmax > 0.0)) {
        return 1;
    }
    // B (m x r) = W_kept / sqrt(g_kept)
    std::vector<double> B((size_t)m * r), PB((size_t)m * r), A((size_t)r * r), d((size_t)r), ee((size_t)r);
    for (int j = 0; j < r; ++j) {
        const double sc = 1.0 / std::sqrt(g[(size_t)keep[(size_t)j]]);
        for (int i = 0; i < m; ++i) {
            B[i + (size_t)j * m] = W[i + (size_t)keep[(size_t)j] * m] * sc;
        }
    }
    for (int j = 0; j < r; ++j) {
        for (int i = 0; i < m; ++i) {
            double acc = 0.0;
            for (int l = 0; l < m; ++l) {
                acc += P[i + (size_t)l * m] * B[l + (size_t)j * m];
            }
            PB[i + (size_t)j * m] = acc;
        }
    }
    for (int j = 0; j < r; ++j) {
        for (int i = 0; i < r; ++i) {
            double acc = 0.0;
            for (int l = 0; l < m; ++l) {
                acc += B[l + (size_t)i * m] * PB[l + (size_t)j * m];
            }
            A[i + (size_t)j * r] = acc;
        }
    }
    for (int j = 0; j < r; ++j) {
        for (int i = j + 1; i < r; ++i) {
            const double v = 0.5 * (A[i + (size_t)j * r] + A[j + (size_t)i * r]);
            A[i + (size_t)j * r] = v;
            A[j + (size_t)i * r] = v;
        }
    }
    tridiagonalize(A.data(), r, d.data(), ee.data());
    if (!tridiagonal_ql(d.data(), ee.data(), A.data(), r)) {
        return 2;
    }
    std::vector<int> order((size_t)r);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return d[(size_t)a] < d[(size_t)b]; });
    std::fill(C, C + (size_t)m * m, 0.0);
    for (int j = 0; j < m; ++j) {
        theta[j] = INFINITY;
    }
    for (int j = 0; j < r; ++j) {
        const int src = order[(size_t)j];
        theta[j] = d[(size_t)src];
        for (int i = 0; i < m; ++i) {
            double acc = 0.0;
            for (int l = 0; l < r; ++l) {
                acc += B[i + (size_t)l * m] * A[l + (size_t)src * r];
            }
            C[i + (size_t)j * m] = acc;
        }
    }
    return 0;
}

}  // namespace ipt
