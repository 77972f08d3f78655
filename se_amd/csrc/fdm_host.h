// fdm_host.h -- host side of the fast-diagonalisation (FDM) preconditioner:
// a dense symmetric eigensolver and the 1-D generalised eigenproblems of the
// separable operator.
//
// The mixed-precision solvers see the operator in its separable form on the
// unknown grid (rows J, columns I):
//     A = Sy (x) Gx + Gy (x) Sx + My (x) Mx
// with Gx, Gy the assembled 1-D band operators and Sx, Sy, Mx, My diagonal
// line factors.  With the generalised eigenvectors Gx Vx = Sx Vx Lx,
// Vx^T Sx Vx = I (and the same for y),
//     (Vy (x) Vx)^T A (Vy (x) Vx) = I (x) Lx + Ly (x) I + (Vy^T My Vy) (x) (Vx^T Mx Vx)
// and Mx, My are multiples of Sx, Sy up to round-off on the reference's
// uniform meshes, so the last term is diagonal to that accuracy and
//     A^-1 ~= (Vy (x) Vx) D^-1 (Vy (x) Vx)^T,  D(J, I) = Lx[I] + Ly[J] + ex[I] ey[J]
// with ex = diag(Vx^T Mx Vx), ey = diag(Vy^T My Vy) (R. E. Lynch, J. R. Rice,
// D. H. Thomas, "Direct solution of partial difference equations by tensor
// product methods", Numer. Math. 6, 1964; E. M. Ronquist's use of it for
// spectral elements).  Applying it is four dense transforms: GEMMs.
//
// Pure host C++ (no HIP types), so that it also builds in host-only checks.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <numeric>
#include <vector>

namespace se_fdm {

// Eigen-decomposition of the symmetric n x n matrix a (row-major, destroyed).
// On return w[k] are the eigenvalues in ascending order and z[k * n + i] is
// component i of the unit eigenvector k (eigenvectors stored as rows).
// Householder reduction to tridiagonal form (Golub & Van Loan, Matrix
// Computations, Alg. 8.3.1), explicit accumulation of the reflections, then
// the implicit symmetric QR iteration with Wilkinson shifts (Alg. 8.3.3)
// applied to the rows of the accumulated basis.  False if the iteration
// does not converge.
inline bool sym_eig(int n, std::vector<double>& a, std::vector<double>& w, std::vector<double>& z) {
    w.assign(n, 0.0);
    z.assign((size_t)n * n, 0.0);
    if (n == 0) return true;
    if (n == 1) {
        w[0] = a[0];
        z[0] = 1.0;
        return true;
    }
    std::vector<double> d(n), e(n, 0.0);
    // reflections H_k = I - beta_k v_k v_k^T acting on rows/columns k+1 .. n-1;
    // v_k (with v_k[k+1] = 1) is kept in row k of a, beta_k in beta
    std::vector<double> beta(n, 0.0), v(n), p(n), q(n);
    for (int k = 0; k + 2 < n; k++) {
        const int m = n - k - 1;  // length of the column below the diagonal
        double* col = &v[0];
        for (int i = 0; i < m; i++) col[i] = a[(size_t)(k + 1 + i) * n + k];
        double sig = 0.0;
        for (int i = 1; i < m; i++) sig += col[i] * col[i];
        const double x0 = col[0];
        double b = 0.0;
        if (sig == 0.0) {
            // already tridiagonal in this column
            e[k] = x0;
            for (int i = 0; i < m; i++) a[(size_t)k * n + k + 1 + i] = 0.0;
            beta[k] = 0.0;
            d[k] = a[(size_t)k * n + k];
            continue;
        }
        const double mu = std::sqrt(x0 * x0 + sig);
        const double v0 = x0 <= 0.0 ? x0 - mu : -sig / (x0 + mu);
        b = 2.0 * v0 * v0 / (sig + v0 * v0);
        col[0] = 1.0;
        for (int i = 1; i < m; i++) col[i] /= v0;
        e[k] = mu;  // H x = mu e1 (Parlett's v0 avoids cancellation for x0 > 0)
        d[k] = a[(size_t)k * n + k];
        // trailing block T = a[k+1.., k+1..]: T := H T H
        // p = b T v, w = p - (b/2)(p.v) v, T -= v w^T + w v^T
        for (int i = 0; i < m; i++) {
            const double* row = &a[(size_t)(k + 1 + i) * n + k + 1];
            double s = 0.0;
            for (int j = 0; j < m; j++) s += row[j] * col[j];
            p[i] = b * s;
        }
        double pv = 0.0;
        for (int i = 0; i < m; i++) pv += p[i] * col[i];
        const double c = 0.5 * b * pv;
        for (int i = 0; i < m; i++) q[i] = p[i] - c * col[i];
        for (int i = 0; i < m; i++) {
            double* row = &a[(size_t)(k + 1 + i) * n + k + 1];
            const double vi = col[i], qi = q[i];
            for (int j = 0; j < m; j++) row[j] -= vi * q[j] + qi * col[j];
        }
        for (int i = 0; i < m; i++) a[(size_t)k * n + k + 1 + i] = col[i];
        beta[k] = b;
    }
    d[n - 2] = a[(size_t)(n - 2) * n + (n - 2)];
    d[n - 1] = a[(size_t)(n - 1) * n + (n - 1)];
    e[n - 2] = a[(size_t)(n - 1) * n + (n - 2)];
    // accumulate Q = H_0 H_1 ... H_{n-3}, stored transposed (row k of z is
    // column k of Q), backwards: Q := H_k Q acts on rows k+1.. of Q, i.e. on
    // columns k+1.. of Q^T
    for (int k = 0; k < n; k++) z[(size_t)k * n + k] = 1.0;
    for (int k = n - 3; k >= 0; k--) {
        const double b = beta[k];
        if (b == 0.0) continue;
        const int m = n - k - 1;
        const double* vk = &a[(size_t)k * n + k + 1];
        for (int r = k + 1; r < n; r++) {  // rows of Q^T = columns of Q (only r > k are touched)
            double* zr = &z[(size_t)r * n + k + 1];
            double s = 0.0;
            for (int i = 0; i < m; i++) s += zr[i] * vk[i];
            s *= b;
            for (int i = 0; i < m; i++) zr[i] -= s * vk[i];
        }
    }
    // implicit symmetric QR on (d, e[0..n-2]), rotations applied to rows of z
    long steps = 0;
    for (int hi = n - 1; hi > 0;) {
        if (++steps > 60L * n) return false;  // no convergence (does not happen for finite input)
        // deflate negligible off-diagonals
        int lo = hi;
        while (lo > 0) {
            const double t = std::fabs(d[lo - 1]) + std::fabs(d[lo]);
            if (std::fabs(e[lo - 1]) <= 1e-16 * t || e[lo - 1] == 0.0) {
                e[lo - 1] = 0.0;
                break;
            }
            lo--;
        }
        if (lo == hi) {
            hi--;
            continue;
        }
        // Wilkinson shift from the trailing 2 x 2 block
        const double dd = 0.5 * (d[hi - 1] - d[hi]);
        const double ee = e[hi - 1];
        const double den = dd + (dd >= 0.0 ? 1.0 : -1.0) * std::hypot(dd, ee);
        const double shift = den == 0.0 ? d[hi] : d[hi] - ee * ee / den;
        double x = d[lo] - shift, y = e[lo];
        for (int k = lo; k < hi; k++) {
            // Givens rotation zeroing y against x
            double c, s;
            if (y == 0.0) {
                c = 1.0;
                s = 0.0;
            } else if (std::fabs(y) > std::fabs(x)) {
                const double t = -x / y;
                s = 1.0 / std::sqrt(1.0 + t * t);
                c = s * t;
            } else {
                const double t = -y / x;
                c = 1.0 / std::sqrt(1.0 + t * t);
                s = c * t;
            }
            if (k > lo) e[k - 1] = c * e[k - 1] - s * y;
            const double a11 = d[k], a22 = d[k + 1], a12 = e[k];
            // G^T [a11 a12; a12 a22] G with G = [c s; -s c]
            d[k] = c * c * a11 - 2.0 * c * s * a12 + s * s * a22;
            d[k + 1] = s * s * a11 + 2.0 * c * s * a12 + c * c * a22;
            e[k] = c * s * (a11 - a22) + (c * c - s * s) * a12;
            if (k + 1 < hi) {
                x = e[k];
                y = -s * e[k + 1];
                e[k + 1] = c * e[k + 1];
            }
            double* zk = &z[(size_t)k * n];
            double* zk1 = &z[(size_t)(k + 1) * n];
            for (int i = 0; i < n; i++) {
                const double u0 = zk[i], u1 = zk1[i];
                zk[i] = c * u0 - s * u1;
                zk1[i] = s * u0 + c * u1;
            }
        }
    }
    // ascending order
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int i, int j) { return d[i] < d[j]; });
    std::vector<double> zz((size_t)n * n);
    for (int k = 0; k < n; k++) {
        w[k] = d[idx[k]];
        std::copy(&z[(size_t)idx[k] * n], &z[(size_t)idx[k] * n] + n, &zz[(size_t)k * n]);
    }
    z.swap(zz);
    return true;
}

// One direction of the FDM: the generalised eigenproblem G v = lambda S v of a
// band operator G (2N+1 coefficients per line, periodic lines wrap) with the
// positive diagonal S, and e = diag(V^T M V) for the diagonal M.
// V[i * nU + k] = component i of generalised eigenvector k (V^T S V = I).
struct Fdm1D {
    int nU = 0;
    std::vector<double> V, lambda, e;
};

inline bool fdm_1d(int nU, int N, int periodic, const std::vector<double>& g, const std::vector<double>& s,
                   const std::vector<double>& m, Fdm1D& out) {
    const int B = 2 * N + 1;
    out.nU = nU;
    for (int i = 0; i < nU; i++)
        if (!(s[i] > 0.0)) return false;
    std::vector<double> C((size_t)nU * nU, 0.0), isq(nU);
    for (int i = 0; i < nU; i++) isq[i] = 1.0 / std::sqrt(s[i]);
    for (int i = 0; i < nU; i++)
        for (int t = 0; t < B; t++) {
            int j = i + t - N;
            if (periodic) j = ((j % nU) + nU) % nU;
            if (j < 0 || j >= nU) continue;
            C[(size_t)i * nU + j] += isq[i] * g[(size_t)i * B + t] * isq[j];
        }
    // symmetrise away the round-off of the assembly
    for (int i = 0; i < nU; i++)
        for (int j = i + 1; j < nU; j++) {
            const double h = 0.5 * (C[(size_t)i * nU + j] + C[(size_t)j * nU + i]);
            C[(size_t)i * nU + j] = C[(size_t)j * nU + i] = h;
        }
    std::vector<double> z;
    if (!sym_eig(nU, C, out.lambda, z)) return false;
    out.V.assign((size_t)nU * nU, 0.0);
    out.e.assign(nU, 0.0);
    for (int k = 0; k < nU; k++)
        for (int i = 0; i < nU; i++) {
            const double v = isq[i] * z[(size_t)k * nU + i];
            out.V[(size_t)i * nU + k] = v;
            out.e[k] += v * v * m[i];
        }
    return true;
}

}  // namespace se_fdm
