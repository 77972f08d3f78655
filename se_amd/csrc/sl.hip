// sl.hip -- semi-Lagrangian advection (departure-point search + GLL
// interpolation), velocity from the stream function, CFL time step.
//
// One thread per field node.  The coordinate arithmetic that decides the
// departure element (RK2 midpoint, wrap/clamp, GetEl's modf and its 1e-12
// edge rule, the local-coordinate map) follows the reference's operation order
// with FP contraction disabled.  P_N(x_j) is tabulated on the host with the
// reference's evaluation; P_N'(x) is evaluated on the device with the same
// monomial sum, its powers bit-identical to the C library's pow
// (glibc_pow.h), so interpolated values and departure elements agree with
// the reference's bit for bit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "glibc_pow.h"
#include "se_common.h"

using namespace se;

namespace {

struct SLGeom {
    int N, nEls0, nEls1, nv0, nv1, per0, per1;
    int cubic;  // the advected field by bicubic Lagrange interpolation (Advector::Interpolate)
    double min0, min1, max0, max1, dx0, dx1;
    const double* xk;   // abscissae
    const double* D;    // derivative matrix
    const double* pnx;  // P_N(x_j)
    const double* pn;   // P_N monomial coefficients
    const double* X;    // field column coordinates
    const double* Y;    // field row coordinates
    const double* J00;
    const double* J10;
    const double* J11;
    const double* J01;
};

#define SE_NUMAPPROX(var, val) ((var) >= (val)-1.0e-12 && (var) <= (val) + 1.0e-12)

// Mesh::GetEl, src/base/Mesh.cpp:106-137
__device__ __forceinline__ int get_el(const SLGeom& g, double cx, double cy) {
#pragma clang fp contract(off)
    int t0, t1;
    double integer, frac;
    if (SE_NUMAPPROX(cx - g.max0, 0.0)) {
        t0 = g.nEls0 - 1;
    } else {
        double out = (cx - g.min0) / g.dx0;
        frac = modf(out, &integer);
        t0 = (int)integer;
        if (t0 > 0 && SE_NUMAPPROX(frac, 0.0)) t0--;
    }
    if (SE_NUMAPPROX(cy - g.max1, 0.0)) {
        t1 = g.nEls1 - 1;
    } else {
        double out = (cy - g.min1) / g.dx1;
        frac = modf(out, &integer);
        t1 = (int)integer;
        if (t1 > 0 && SE_NUMAPPROX(frac, 0.0)) t1--;
    }
    // A coordinate outside the domain (displacement beyond one period, or a
    // non-finite velocity) has no element in the reference either; clamp so
    // the gather stays inside the field arrays.
    t0 = t0 < 0 ? 0 : (t0 >= g.nEls0 ? g.nEls0 - 1 : t0);
    t1 = t1 < 0 ? 0 : (t1 >= g.nEls1 ? g.nEls1 - 1 : t1);
    return t1 * g.nEls0 + t0;
}

// Advector::PeriodicUpdate, src/base/Advector.cpp:122-134
__device__ __forceinline__ void wrap(const SLGeom& g, double& cx, double& cy) {
#pragma clang fp contract(off)
    if (cx < g.min0) cx = g.per0 ? g.max0 - g.min0 + cx : g.min0;
    if (cx > g.max0) cx = g.per0 ? g.min0 - g.max0 + cx : g.max0;
    if (cy < g.min1) cy = g.per1 ? g.max1 - g.min1 + cy : g.min1;
    if (cy > g.max1) cy = g.per1 ? g.min1 - g.max1 + cy : g.max1;
}

// The kernels take the order as a template argument NN (the launches
// instantiate N = 2..8) or, with NN = 0, from the geometry (any N up to 12):
// the same operations either way, with fixed trip counts and arrays of N + 1
// when NN is known.
template <int NN> struct Ord {
    static constexpr int MAX = NN ? NN : 12;  // array bound
};
template <int NN>
__device__ __forceinline__ int order(const SLGeom& g) {
    return NN ? NN : g.N;
}

// Jacobi::EvalDeriv (Jacobi.cpp:58-66) of P_N; its powers of x share one
// log (pow_int_at: each power the same bits as pow(x, i - 1))
template <int NN>
__device__ __forceinline__ double pn_deriv(const SLGeom& g, double x) {
#pragma clang fp contract(off)
    const se_pow::PowBase b = se_pow::pow_base(x);
    double dp = 0.0;
    for (int i = order<NN>(g); i >= 1; i -= 2) dp += g.pn[i] * i * se_pow::pow_int_at(b, i - 1);
    return dp;
}

// 1-D GLL cardinal functions at local coordinate x (Legendre::IsRoot and the
// per-dimension factor of ShapeFuncsAtCoord, Legendre.cpp:145-169,291-305)
template <int NN>
__device__ __forceinline__ void cardinal(const SLGeom& g, double x, double (&L)[Ord<NN>::MAX + 1]) {
#pragma clang fp contract(off)
    const int N = order<NN>(g);
    int root = -1;
    for (int k = 0; k <= N; k++)
        if (fabs(x - g.xk[k]) < 1.0e-8) {
            root = k;
            break;
        }
    if (root >= 0) {
        for (int j = 0; j <= N; j++) L[j] = (j == root) ? 1.0 : 0.0;
        return;
    }
    const double dpx = pn_deriv<NN>(g, x);
    for (int j = 0; j <= N; j++)
        L[j] = -((1.0 - x * x) / (N * (N + 1) * g.pnx[j] * (x - g.xk[j]))) * dpx;
}

// Field::InterpGlobal (Field.cpp:57-76) through Mesh::GlobalToLocal
// (Mesh.cpp:163-177); vals node-major with ND dofs
template <int NN, int ND>
__device__ __forceinline__ void interp_global(const SLGeom& g, const double* __restrict__ vals, int nd, double cx,
                                              double cy, double (&out)[ND]) {
#pragma clang fp contract(off)
    const int N = order<NN>(g);
    int el = get_el(g, cx, cy);
    int ex = el % g.nEls0, ey = el / g.nEls0;
    double lx = (cx - g.min0 - ex * g.dx0) * (2.0 / g.dx0) - 1.0;
    double ly = (cy - g.min1 - ey * g.dx1) * (2.0 / g.dx1) - 1.0;
    if (lx < -1.0) lx = -1;
    if (lx > +1.0) lx = +1;
    if (ly < -1.0) ly = -1;
    if (ly > +1.0) ly = +1;
    double Lx[Ord<NN>::MAX + 1], Ly[Ord<NN>::MAX + 1];
    cardinal<NN>(g, lx, Lx);
    cardinal<NN>(g, ly, Ly);
    for (int d = 0; d < ND; d++) out[d] = 0.0;
    const int n0 = N * ey * g.nv0 + N * ex;
#pragma unroll 1
    for (int jj = 0; jj <= N; jj++)
        for (int ii = 0; ii <= N; ii++) {
            double Ni = Lx[ii] * Ly[jj];
            const double* v = &vals[(size_t)(n0 + jj * g.nv0 + ii) * nd];
            for (int d = 0; d < ND; d++)
                if (d < nd) out[d] += Ni * v[d];
        }
}

// interp_global of two node-major fields at one point (the weights computed
// once): each output is summed exactly as interp_global sums it
template <int NN, int NA, int NB>
__device__ __forceinline__ void interp_global2(const SLGeom& g, const double* __restrict__ va, int na,
                                               const double* __restrict__ vb, int nb, double cx, double cy,
                                               double (&oa)[NA], double (&ob)[NB]) {
#pragma clang fp contract(off)
    const int N = order<NN>(g);
    int el = get_el(g, cx, cy);
    int ex = el % g.nEls0, ey = el / g.nEls0;
    double lx = (cx - g.min0 - ex * g.dx0) * (2.0 / g.dx0) - 1.0;
    double ly = (cy - g.min1 - ey * g.dx1) * (2.0 / g.dx1) - 1.0;
    if (lx < -1.0) lx = -1;
    if (lx > +1.0) lx = +1;
    if (ly < -1.0) ly = -1;
    if (ly > +1.0) ly = +1;
    double Lx[Ord<NN>::MAX + 1], Ly[Ord<NN>::MAX + 1];
    cardinal<NN>(g, lx, Lx);
    cardinal<NN>(g, ly, Ly);
    for (int d = 0; d < NA; d++) oa[d] = 0.0;
    for (int d = 0; d < NB; d++) ob[d] = 0.0;
    const int n0 = N * ey * g.nv0 + N * ex;
#pragma unroll 1
    for (int jj = 0; jj <= N; jj++)
        for (int ii = 0; ii <= N; ii++) {
            double Ni = Lx[ii] * Ly[jj];
            const size_t node = (size_t)(n0 + jj * g.nv0 + ii);
            for (int d = 0; d < NA; d++)
                if (d < na) oa[d] += Ni * va[node * na + d];
            for (int d = 0; d < NB; d++)
                if (d < nb) ob[d] += Ni * vb[node * nb + d];
        }
}

// Advector::InterpLagrange, Advector.cpp:248-275
template <int ND>
__device__ __forceinline__ void interp_lagrange(double x, const double* c, double (&vals)[4][ND], double (&psi)[ND]) {
#pragma clang fp contract(off)
    for (int d = 0; d < ND; d++) psi[d] = 0.0;
    for (int node = 0; node < 4; node++) {
        double factor = 1.0;
        for (int o = 0; o < 4; o++)
            if (o != node) factor *= (x - c[o]) / (c[node] - c[o]);
        for (int d = 0; d < ND; d++) psi[d] += vals[node][d] * factor;
    }
}

// Advector::Interpolate (bicubic on the 4x4 node stencil, used when N < 3),
// Advector.cpp:278-337
template <int ND>
__device__ __forceinline__ void interp_bicubic(const SLGeom& g, const double* __restrict__ vals, int nd, double cx,
                                               double cy, double (&out)[ND]) {
#pragma clang fp contract(off)
    const int N = g.N;
    int el = get_el(g, cx, cy);
    int ex = el % g.nEls0, ey = el / g.nEls0;
    int node0 = N * ey * g.nv0 + N * ex;
    int x0 = node0 % g.nv0;
    int y0 = (node0 / g.nv0) % g.nv1;
    double ddx = g.dx0 / N, ddy = g.dx1 / N;
    if (cx <= g.X[x0 + 1]) x0--;
    if (cy <= g.Y[ey * N + 3 / (N + 1)]) y0--;  // row of entry 3 of the element's node list
    if (cx <= g.min0 + ddx) x0++;
    else if (cx >= g.max0 - ddx) x0--;
    if (cy <= g.min1 + ddy) y0++;
    else if (cy >= g.max1 - ddy) y0--;
    x0 = x0 < 0 ? 0 : (x0 > g.nv0 - 4 ? g.nv0 - 4 : x0);
    y0 = y0 < 0 ? 0 : (y0 > g.nv1 - 4 ? g.nv1 - 4 : y0);
    double px[4], py[4], pts[4][ND], rows[4][ND];
    for (int i = 0; i < 4; i++) px[i] = g.X[x0 + i];
    for (int j = 0; j < 4; j++) py[j] = g.Y[y0 + j];
    for (int j = 0; j < 4; j++) {
        for (int i = 0; i < 4; i++)
            for (int d = 0; d < ND; d++) pts[i][d] = d < nd ? vals[(size_t)(x0 + i + (y0 + j) * g.nv0) * nd + d] : 0.0;
        double r[ND];
        interp_lagrange<ND>(cx, px, pts, r);
        for (int d = 0; d < ND; d++) rows[j][d] = r[d];
    }
    interp_lagrange<ND>(cy, py, rows, out);
}

template <int NN, int ND>
__device__ __forceinline__ void interp_any(const SLGeom& g, const double* __restrict__ vals, int nd, double cx,
                                           double cy, double (&out)[ND]) {
    if (g.cubic) interp_bicubic<ND>(g, vals, nd, cx, cy, out);
    else interp_global<NN, ND>(g, vals, nd, cx, cy, out);
}

// Advector::FirstOrder + IntegrateRK2 (Advector.cpp:136-148,181-201)
template <int NN>
__global__ __launch_bounds__(256) void k_sl_first(SLGeom g, const double* __restrict__ q, int nd,
                                                   const double* __restrict__ vel, double dt, double* __restrict__ qsl,
                                                   int32_t* __restrict__ dep) {
#pragma clang fp contract(off)
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.nv0 * g.nv1) return;
    const double ox = g.X[v % g.nv0], oy = g.Y[v / g.nv0];
    // velocity at the origin: the origin is a GLL node, so InterpGlobal
    // returns the nodal value exactly (IsRoot selects a Kronecker delta)
    double vel0[2];
    interp_global<NN, 2>(g, vel, 2, ox, oy, vel0);
    double hx = ox - 0.5 * dt * vel0[0];
    double hy = oy - 0.5 * dt * vel0[1];
    wrap(g, hx, hy);
    double vh[2];
    interp_global<NN, 2>(g, vel, 2, hx, hy, vh);
    double cx = ox - dt * vh[0];
    double cy = oy - dt * vh[1];
    wrap(g, cx, cy);
    double val[3];
    interp_any<NN, 3>(g, q, nd, cx, cy, val);
    for (int d = 0; d < nd; d++) qsl[(size_t)v * nd + d] = val[d];
    if (dep) dep[v] = get_el(g, cx, cy);
}

// Advector::ExtrapolateVelocities (Advector.cpp:339-349)
__global__ void k_extrap(int n2, const double* __restrict__ vel, const double* __restrict__ velPrev,
                         double* __restrict__ vp1, double* __restrict__ vph, double* __restrict__ vmh) {
#pragma clang fp contract(off)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n2) return;
    double a = vel[i], b = velPrev[i];
    vp1[i] = 2.0 * a - 1.0 * b;
    vph[i] = 1.5 * a - 0.5 * b;
    vmh[i] = 0.5 * a + 0.5 * b;
}

// Advector::SecondOrder (Advector.cpp:203-246).  Four waves per SIMD (at
// most 128 VGPRs, ~160 bytes of spills) ran 7% faster than three (167 VGPRs):
// the six gathers per node are latency-bound.
template <int NN>
__global__ __launch_bounds__(256, 4) void k_sl_second(SLGeom g, const double* __restrict__ q, const double* __restrict__ qp,
                                                    int nd, const double* __restrict__ vel,
                                                    const double* __restrict__ vp1, const double* __restrict__ vph,
                                                    const double* __restrict__ vmh, double dt,
                                                    double* __restrict__ qsl) {
#pragma clang fp contract(off)
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.nv0 * g.nv1) return;
    const double ox = g.X[v % g.nv0], oy = g.Y[v / g.nv0];
    double w[2];
    interp_global<NN, 2>(g, vp1, 2, ox, oy, w);
    double hx = ox - 0.5 * dt * w[0], hy = oy - 0.5 * dt * w[1];
    wrap(g, hx, hy);
    interp_global<NN, 2>(g, vph, 2, hx, hy, w);
    double x1 = ox - dt * w[0], y1 = oy - dt * w[1];
    wrap(g, x1, y1);
    // the velocity and (GLL interpolation) the advected field at (x1, y1)
    // share one set of weights
    double p1[3], p2[3];
    if (g.cubic) {
        interp_global<NN, 2>(g, vel, 2, x1, y1, w);
        interp_bicubic<3>(g, q, nd, x1, y1, p1);
    } else {
        interp_global2<NN, 2, 3>(g, vel, 2, q, nd, x1, y1, w, p1);
    }
    hx = x1 - 0.5 * dt * w[0];
    hy = y1 - 0.5 * dt * w[1];
    wrap(g, hx, hy);
    interp_global<NN, 2>(g, vmh, 2, hx, hy, w);
    double x2 = x1 - dt * w[0], y2 = y1 - dt * w[1];
    wrap(g, x2, y2);
    interp_any<NN, 3>(g, qp, nd, x2, y2, p2);
    for (int d = 0; d < nd; d++) qsl[(size_t)v * nd + d] = 2.0 * p1[d] - 0.5 * p2[d];
}

// Field::PeriodicUpdate (Field.cpp:232-252) applied to the advected field
__global__ void k_periodic_copy(int nv0, int nv1, int per0, int per1, int nd, double* __restrict__ f) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv0 * nv1) return;
    int I = v % nv0, J = v / nv0;
    bool dupI = per0 && I == nv0 - 1, dupJ = per1 && J == nv1 - 1;
    if (!dupI && !dupJ) return;
    int src = (dupJ ? 0 : J) * nv0 + (dupI ? 0 : I);
    for (int d = 0; d < nd; d++) f[(size_t)v * nd + d] = f[(size_t)src * nd + d];
}

// Field::InterpGlobal (mode GLL) or Advector::Interpolate (mode BICUBIC) at
// arbitrary points, one thread per point; xy: npts x 2, out: npts x nd
template <int NN>
__global__ __launch_bounds__(256) void k_interp_points(SLGeom g, const double* __restrict__ vals, int nd, int npts,
                                                       const double* __restrict__ xy, int cubic,
                                                       double* __restrict__ out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npts) return;
    double v[3];
    if (cubic) interp_bicubic<3>(g, vals, nd, xy[2 * (size_t)p], xy[2 * (size_t)p + 1], v);
    else interp_global<NN, 3>(g, vals, nd, xy[2 * (size_t)p], xy[2 * (size_t)p + 1], v);
    for (int d = 0; d < nd; d++) out[(size_t)p * nd + d] = v[d];
}

// QuasiGeostrophicEqn::GenVel via Field::InterpDerivsGlobal and
// Mesh::ShapeFuncDerivsAtCoord (Mesh.cpp:252-285).  At a vertex both local
// coordinates are GLL roots, so the local derivative rows are columns of D
// and the Jacobian is the tabulated one of that GLL point.
template <int NN>
__global__ __launch_bounds__(256) void k_genvel(SLGeom g, const double* __restrict__ psi, double* __restrict__ vel) {
#pragma clang fp contract(off)
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.nv0 * g.nv1) return;
    const int N = order<NN>(g), n1 = N + 1;
    const double ox = g.X[v % g.nv0], oy = g.Y[v / g.nv0];
    int el = get_el(g, ox, oy);
    int ex = el % g.nEls0, ey = el / g.nEls0;
    double lx = (ox - g.min0 - ex * g.dx0) * (2.0 / g.dx0) - 1.0;
    double ly = (oy - g.min1 - ey * g.dx1) * (2.0 / g.dx1) - 1.0;
    if (lx < -1.0) lx = -1;
    if (lx > +1.0) lx = +1;
    if (ly < -1.0) ly = -1;
    if (ly > +1.0) ly = +1;
    int xi = -1, yi = -1;
    for (int k = 0; k <= N; k++)
        if (xi < 0 && fabs(lx - g.xk[k]) < 1.0e-8) xi = k;
    for (int k = 0; k <= N; k++)
        if (yi < 0 && fabs(ly - g.xk[k]) < 1.0e-8) yi = k;
    if (xi < 0 || yi < 0) {  // cannot happen at a vertex
        vel[2 * (size_t)v] = vel[2 * (size_t)v + 1] = __builtin_nan("");
        return;
    }
    double j00 = g.J00[ex * n1 + xi], j01 = g.J01[(ey * n1 + yi) * n1 + xi];
    double j10 = g.J10[(ex * n1 + xi) * n1 + yi], j11 = g.J11[ey * n1 + yi];
    double det = j00 * j11 - j01 * j10;
    double i00 = j11 / det, i11 = j00 / det, i01 = -j01 / det, i10 = -j10 / det;
    double g0 = 0.0, g1 = 0.0;
    const int n0 = N * ey * g.nv0 + N * ex;
    // the element's nodes (jj, ii) in row-major order; only row yi and
    // column xi can carry a non-zero shape-function derivative, so the walk
    // visits those in the same order (the others were skipped anyway)
    auto node = [&](int jj, int ii) {
        double a0 = (jj == yi) ? g.D[xi * n1 + ii] : 0.0;  // GNx * Ny
        double a1 = (ii == xi) ? g.D[yi * n1 + jj] : 0.0;  // Nx * GNy
        if (a0 == 0.0 && a1 == 0.0) return;
        double t0 = 0.0;
        t0 += a0 * i00;
        t0 += a1 * i01;
        double t1 = 0.0;
        t1 += a0 * i10;
        t1 += a1 * i11;
        double pv = psi[n0 + jj * g.nv0 + ii];
        g0 += t0 * pv;
        g1 += t1 * pv;
    };
    for (int jj = 0; jj < yi; jj++) node(jj, xi);
    for (int ii = 0; ii <= N; ii++) node(yi, ii);
    for (int jj = yi + 1; jj <= N; jj++) node(jj, xi);
    vel[2 * (size_t)v] = -1.0 * g1;
    vel[2 * (size_t)v + 1] = +1.0 * g0;
}

__global__ void k_maxvel2(int nV, const double* __restrict__ vel, unsigned long long* __restrict__ out) {
    double m = 0.0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nV; v += gridDim.x * blockDim.x) {
        double a = vel[2 * (size_t)v], b = vel[2 * (size_t)v + 1];
        double s = a * a + b * b;
        if (s > m) m = s;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        double o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    // non-negative doubles order like their bit patterns
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

SLGeom sl_geom(const se_mesh* m) {
    SLGeom g;
    g.N = m->N;
    // Advector's cubicInterp: N < 3 unless the caller chose (Advector.cpp:33,60)
    g.cubic = m->sl_interp == SE_INTERP_BICUBIC || (m->sl_interp == SE_INTERP_DEFAULT && m->N < 3);
    g.nEls0 = m->nEls[0];
    g.nEls1 = m->nEls[1];
    g.nv0 = m->nVerts[0];
    g.nv1 = m->nVerts[1];
    g.per0 = m->periodic[0];
    g.per1 = m->periodic[1];
    g.min0 = m->min[0];
    g.min1 = m->min[1];
    g.max0 = m->max[0];
    g.max1 = m->max[1];
    g.dx0 = m->dx[0];
    g.dx1 = m->dx[1];
    const int n1 = m->N + 1;
    const double* b = m->d_basis.as<double>();
    g.xk = b;
    g.D = b + 2 * n1;
    g.pnx = b + 2 * n1 + n1 * n1;
    g.pn = b + 3 * n1 + n1 * n1;
    g.X = m->d_X.as<double>();
    g.Y = m->d_Y.as<double>();
    g.J00 = m->d_J00.as<double>();
    g.J10 = m->d_J10.as<double>();
    g.J11 = m->d_J11.as<double>();
    g.J01 = m->d_J01.as<double>();
    return g;
}

int blocks_for(int n) { return (n + 255) / 256; }

// KERNEL<N> for the orders 2..8, KERNEL<0> (the order read from the
// geometry) for the others; 256 threads per workgroup
#define SE_SL_LAUNCH(KERNEL, N, GRID, ...)                                          \
    switch (N) {                                                                    \
        case 2: hipLaunchKernelGGL(KERNEL<2>, GRID, dim3(256), 0, __VA_ARGS__); break; \
        case 3: hipLaunchKernelGGL(KERNEL<3>, GRID, dim3(256), 0, __VA_ARGS__); break; \
        case 4: hipLaunchKernelGGL(KERNEL<4>, GRID, dim3(256), 0, __VA_ARGS__); break; \
        case 5: hipLaunchKernelGGL(KERNEL<5>, GRID, dim3(256), 0, __VA_ARGS__); break; \
        case 6: hipLaunchKernelGGL(KERNEL<6>, GRID, dim3(256), 0, __VA_ARGS__); break; \
        case 7: hipLaunchKernelGGL(KERNEL<7>, GRID, dim3(256), 0, __VA_ARGS__); break; \
        case 8: hipLaunchKernelGGL(KERNEL<8>, GRID, dim3(256), 0, __VA_ARGS__); break; \
        default: hipLaunchKernelGGL(KERNEL<0>, GRID, dim3(256), 0, __VA_ARGS__); break; \
    }

struct SLScratch {
    DevBuf a, b, c, d, e, f, g, h;
    ~SLScratch() {
        DevBuf* all[] = {&a, &b, &c, &d, &e, &f, &g, &h};
        for (DevBuf* p : all) p->release();
    }
};

}  // namespace

extern "C" int se_sl_advect_dev(se_mesh* m, const double* d_q, int nd, const double* d_vel, double dt, double* d_qsl,
                                int32_t* d_dep) {
    SE_CHECK(m && d_q && d_vel && d_qsl, SE_ERR_ARG, "se_sl_advect_dev: NULL argument");
    SE_CHECK(nd >= 1 && nd <= 3, SE_ERR_ARG, "se_sl_advect_dev: nDofs must be 1..3");
    SE_CHECK(m->N <= 12, SE_ERR_UNSUPPORTED, "se_sl_advect_dev: order above 12");
    SE_CHECK(m->N >= 3 || (m->nVerts[0] >= 4 && m->nVerts[1] >= 4), SE_ERR_UNSUPPORTED,
             "bicubic advection (N < 3) needs at least 4 nodes per direction");
    SLGeom g = sl_geom(m);
    hipStream_t s = m->ctx->stream;
    SE_SL_LAUNCH(k_sl_first, m->N, dim3(blocks_for(m->nV)), s, g, d_q, nd, d_vel, dt, d_qsl, d_dep);
    if (m->periodic[0] || m->periodic[1])
        hipLaunchKernelGGL(k_periodic_copy, dim3(blocks_for(m->nV)), dim3(256), 0, s, m->nVerts[0], m->nVerts[1],
                           m->periodic[0], m->periodic[1], nd, d_qsl);
    SE_HIP(hipGetLastError());
    return SE_OK;
}

namespace {

int sl_second_check(const se_mesh* m, const void* d_q, const void* d_qp, int nd, const void* d_vel,
                    const void* d_velp, const void* d_qsl, const char* who) {
    SE_CHECK(m && d_q && d_qp && d_vel && d_velp && d_qsl, SE_ERR_ARG, std::string(who) + ": NULL argument");
    SE_CHECK(nd >= 1 && nd <= 3, SE_ERR_ARG, std::string(who) + ": nDofs must be 1..3");
    SE_CHECK(m->N <= 12, SE_ERR_UNSUPPORTED, std::string(who) + ": order above 12");
    SE_CHECK(m->N >= 3 || (m->nVerts[0] >= 4 && m->nVerts[1] >= 4), SE_ERR_UNSUPPORTED,
             "bicubic advection (N < 3) needs at least 4 nodes per direction");
    return SE_OK;
}

// the second-order step with the three extrapolated velocity fields in
// vp1 | vph | vmh (2 nV doubles each)
int sl_second(se_mesh* m, const double* d_q, const double* d_qp, int nd, const double* d_vel, const double* d_velp,
              double dt, double* d_qsl, double* vp1) {
    SLGeom g = sl_geom(m);
    hipStream_t s = m->ctx->stream;
    const size_t n2 = 2 * (size_t)m->nV;
    double* vph = vp1 + n2;
    double* vmh = vph + n2;
    hipLaunchKernelGGL(k_extrap, dim3(blocks_for((int)n2)), dim3(256), 0, s, (int)n2, d_vel, d_velp, vp1, vph, vmh);
    SE_SL_LAUNCH(k_sl_second, m->N, dim3(blocks_for(m->nV)), s, g, d_q, d_qp, nd, d_vel, vp1, vph, vmh, dt, d_qsl);
    if (m->periodic[0] || m->periodic[1])
        hipLaunchKernelGGL(k_periodic_copy, dim3(blocks_for(m->nV)), dim3(256), 0, s, m->nVerts[0], m->nVerts[1],
                           m->periodic[0], m->periodic[1], nd, d_qsl);
    SE_HIP(hipGetLastError());
    return SE_OK;
}

// Scratch of the allocation-free second-order step: its three extrapolated
// velocity fields, 3 x 2 x SE_SL_NOALLOC_MAX_NODES doubles (24 MiB).  One
// copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the device
// symbol): it passes the per-device address that hipGetSymbolAddress gives.
static __device__ __align__(256) double scratch[3 * 2 * (size_t)SE_SL_NOALLOC_MAX_NODES];

std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr;  // device -> address of that device's scratch

// this device's scratch address, looked up once per device with that device
// current (the caller's current device is restored afterwards)
int scratch_of(int device, double** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it == g_scratch_ptr.end()) {
        double* ptr = nullptr;
        int prev = 0;
        SE_HIP(hipGetDevice(&prev));
        SE_HIP(hipSetDevice(device));
        hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        SE_HIP(hipSetDevice(prev));
        SE_CHECK(e == hipSuccess, SE_ERR_HIP,
                 std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        SE_CHECK(ptr != nullptr, SE_ERR_HIP, "no device address for the SL scratch");
        it = g_scratch_ptr.emplace(device, ptr).first;
    }
    *out = it->second;
    return SE_OK;
}

}  // namespace

extern "C" int se_sl_advect2_dev_noalloc(se_mesh* m, const double* d_q, const double* d_qp, int nd,
                                         const double* d_vel, const double* d_velp, double dt, double* d_qsl) {
    int rc = sl_second_check(m, d_q, d_qp, nd, d_vel, d_velp, d_qsl, "se_sl_advect2_dev_noalloc");
    if (rc) return rc;
    SE_CHECK((size_t)m->nV <= (size_t)SE_SL_NOALLOC_MAX_NODES, SE_ERR_ARG,
             "se_sl_advect2_dev_noalloc: the mesh has " + std::to_string(m->nV) +
                 " field nodes, more than the fixed scratch holds (SE_SL_NOALLOC_MAX_NODES = " +
                 std::to_string(SE_SL_NOALLOC_MAX_NODES) + "); use se_sl_advect2_dev");
    double* sc = nullptr;
    rc = scratch_of(m->ctx->device, &sc);
    if (rc) return rc;
    return sl_second(m, d_q, d_qp, nd, d_vel, d_velp, dt, d_qsl, sc);
}

extern "C" int se_sl_advect2_dev(se_mesh* m, const double* d_q, const double* d_qp, int nd, const double* d_vel,
                                 const double* d_velp, double dt, double* d_qsl) {
    int rc = sl_second_check(m, d_q, d_qp, nd, d_vel, d_velp, d_qsl, "se_sl_advect2_dev");
    if (rc) return rc;
    size_t n2 = 2 * (size_t)m->nV;
    rc = m->d_sl.alloc(3 * n2 * sizeof(double));
    if (rc) return rc;
    return sl_second(m, d_q, d_qp, nd, d_vel, d_velp, dt, d_qsl, m->d_sl.as<double>());
}

extern "C" int se_genvel_dev(se_mesh* m, const double* d_psi, double* d_vel) {
    SE_CHECK(m && d_psi && d_vel, SE_ERR_ARG, "se_genvel_dev: NULL argument");
    SLGeom g = sl_geom(m);
    SE_SL_LAUNCH(k_genvel, m->N, dim3(blocks_for(m->nV)), m->ctx->stream, g, d_psi, d_vel);
    SE_HIP(hipGetLastError());
    return SE_OK;
}

extern "C" int se_calc_timestep_dev(se_mesh* m, const double* d_vel, double courant, double* dt) {
    SE_CHECK(m && d_vel && dt, SE_ERR_ARG, "se_calc_timestep_dev: NULL argument");
    int rc = m->d_red.alloc(sizeof(unsigned long long));
    if (rc) return rc;
    hipStream_t s = m->ctx->stream;
    SE_HIP(hipMemsetAsync(m->d_red.ptr, 0, sizeof(unsigned long long), s));
    int nb = std::min(blocks_for(m->nV), 1024);
    hipLaunchKernelGGL(k_maxvel2, dim3(nb), dim3(256), 0, s, m->nV, d_vel, m->d_red.as<unsigned long long>());
    unsigned long long bits = 0;
    SE_HIP(hipMemcpyAsync(&bits, m->d_red.ptr, sizeof(bits), hipMemcpyDeviceToHost, s));
    SE_HIP(hipStreamSynchronize(s));
    double maxVelSq;
    std::memcpy(&maxVelSq, &bits, sizeof(double));
    // Utils.cpp:23-40
    double minSep = (m->dx[0] < m->dx[1]) ? m->dx[0] : m->dx[1];
    *dt = courant * minSep / std::sqrt(maxVelSq);
    return SE_OK;
}

// ---- host-pointer wrappers ---------------------------------------------------
namespace {
int up(hipStream_t s, DevBuf& b, const void* h, size_t bytes) {
    int rc = b.alloc(std::max<size_t>(bytes, 8));
    if (rc) return rc;
    SE_HIP(hipMemcpyAsync(b.ptr, h, bytes, hipMemcpyHostToDevice, s));
    return SE_OK;
}
int down(hipStream_t s, void* h, const DevBuf& b, size_t bytes) {
    SE_HIP(hipMemcpyAsync(h, b.ptr, bytes, hipMemcpyDeviceToHost, s));
    SE_HIP(hipStreamSynchronize(s));
    return SE_OK;
}
}  // namespace

extern "C" int se_sl_advect(se_mesh* m, const double* q, int nd, const double* vel, double dt, double* qsl,
                            int32_t* dep) {
    SE_CHECK(m && q && vel && qsl, SE_ERR_ARG, "se_sl_advect: NULL argument");
    SE_CHECK(nd >= 1 && nd <= 3, SE_ERR_ARG, "se_sl_advect: nDofs must be 1..3");
    SE_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    SLScratch b;
    size_t nV = m->nV;
    int rc = up(s, b.a, q, nV * nd * sizeof(double));
    if (!rc) rc = up(s, b.b, vel, nV * 2 * sizeof(double));
    if (!rc) rc = b.c.alloc(nV * nd * sizeof(double));
    if (!rc) rc = b.d.alloc(nV * sizeof(int32_t));
    if (!rc) rc = se_sl_advect_dev(m, b.a.as<double>(), nd, b.b.as<double>(), dt, b.c.as<double>(),
                                   dep ? b.d.as<int32_t>() : nullptr);
    if (!rc) rc = down(s, qsl, b.c, nV * nd * sizeof(double));
    if (!rc && dep) rc = down(s, dep, b.d, nV * sizeof(int32_t));
    return rc;
}

extern "C" int se_sl_advect2(se_mesh* m, const double* q, const double* qp, int nd, const double* vel,
                             const double* velp, double dt, double* qsl) {
    SE_CHECK(m && q && qp && vel && velp && qsl, SE_ERR_ARG, "se_sl_advect2: NULL argument");
    SE_CHECK(nd >= 1 && nd <= 3, SE_ERR_ARG, "se_sl_advect2: nDofs must be 1..3");
    SE_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    SLScratch b;
    size_t nV = m->nV;
    int rc = up(s, b.a, q, nV * nd * sizeof(double));
    if (!rc) rc = up(s, b.b, qp, nV * nd * sizeof(double));
    if (!rc) rc = up(s, b.c, vel, nV * 2 * sizeof(double));
    if (!rc) rc = up(s, b.d, velp, nV * 2 * sizeof(double));
    if (!rc) rc = b.e.alloc(nV * nd * sizeof(double));
    if (!rc) rc = se_sl_advect2_dev(m, b.a.as<double>(), b.b.as<double>(), nd, b.c.as<double>(), b.d.as<double>(), dt,
                                    b.e.as<double>());
    if (!rc) rc = down(s, qsl, b.e, nV * nd * sizeof(double));
    return rc;
}

extern "C" int se_genvel(se_mesh* m, const double* psi, double* vel) {
    SE_CHECK(m && psi && vel, SE_ERR_ARG, "se_genvel: NULL argument");
    SE_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    SLScratch b;
    size_t nV = m->nV;
    int rc = up(s, b.a, psi, nV * sizeof(double));
    if (!rc) rc = b.b.alloc(nV * 2 * sizeof(double));
    if (!rc) rc = se_genvel_dev(m, b.a.as<double>(), b.b.as<double>());
    if (!rc) rc = down(s, vel, b.b, nV * 2 * sizeof(double));
    return rc;
}

extern "C" int se_calc_timestep(se_mesh* m, const double* vel, double courant, double* dt) {
    SE_CHECK(m && vel && dt, SE_ERR_ARG, "se_calc_timestep: NULL argument");
    SE_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    SLScratch b;
    int rc = up(s, b.a, vel, (size_t)m->nV * 2 * sizeof(double));
    if (!rc) rc = se_calc_timestep_dev(m, b.a.as<double>(), courant, dt);
    return rc;
}

extern "C" int se_mesh_set_sl_interp(se_mesh* m, int mode) {
    SE_CHECK(m, SE_ERR_ARG, "se_mesh_set_sl_interp: NULL argument");
    SE_CHECK(mode == SE_INTERP_DEFAULT || mode == SE_INTERP_GLL || mode == SE_INTERP_BICUBIC, SE_ERR_ARG,
             "se_mesh_set_sl_interp: mode must be SE_INTERP_DEFAULT, SE_INTERP_GLL or SE_INTERP_BICUBIC");
    const bool cubic = mode == SE_INTERP_BICUBIC || (mode == SE_INTERP_DEFAULT && m->N < 3);
    SE_CHECK(!cubic || (m->nVerts[0] >= 4 && m->nVerts[1] >= 4), SE_ERR_UNSUPPORTED,
             "bicubic interpolation needs at least 4 nodes per direction");
    m->sl_interp = mode;
    return SE_OK;
}

extern "C" int se_interp_points_dev(se_mesh* m, const double* d_vals, int nd, int npts, const double* d_xy, int mode,
                                    double* d_out) {
    SE_CHECK(m && (npts == 0 || (d_vals && d_xy && d_out)), SE_ERR_ARG, "se_interp_points_dev: NULL argument");
    SE_CHECK(nd >= 1 && nd <= 3 && npts >= 0, SE_ERR_ARG, "se_interp_points_dev: nDofs 1..3, npts >= 0");
    SE_CHECK(mode == SE_INTERP_DEFAULT || mode == SE_INTERP_GLL || mode == SE_INTERP_BICUBIC, SE_ERR_ARG,
             "se_interp_points_dev: bad mode");
    SE_CHECK(m->N <= 12, SE_ERR_UNSUPPORTED, "se_interp_points_dev: order above 12");
    const int cubic = mode == SE_INTERP_BICUBIC || (mode == SE_INTERP_DEFAULT && m->N < 3);
    SE_CHECK(!cubic || (m->nVerts[0] >= 4 && m->nVerts[1] >= 4), SE_ERR_UNSUPPORTED,
             "bicubic interpolation needs at least 4 nodes per direction");
    if (npts == 0) return SE_OK;
    SLGeom g = sl_geom(m);
    SE_SL_LAUNCH(k_interp_points, m->N, dim3(blocks_for(npts)), m->ctx->stream, g, d_vals, nd, npts, d_xy, cubic,
                 d_out);
    SE_HIP(hipGetLastError());
    return SE_OK;
}

extern "C" int se_interp_points(se_mesh* m, const double* vals, int nd, int npts, const double* xy, int mode,
                                double* out) {
    SE_CHECK(m && (npts == 0 || (vals && xy && out)), SE_ERR_ARG, "se_interp_points: NULL argument");
    SE_CHECK(nd >= 1 && nd <= 3 && npts >= 0, SE_ERR_ARG, "se_interp_points: nDofs 1..3, npts >= 0");
    if (npts == 0) return SE_OK;
    SE_HIP(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    SLScratch b;
    const size_t nv = (size_t)m->nV * nd;
    int rc = up(s, b.a, vals, nv * sizeof(double));
    if (!rc) rc = up(s, b.b, xy, 2 * (size_t)npts * sizeof(double));
    if (!rc) rc = b.c.alloc((size_t)npts * nd * sizeof(double));
    if (!rc) rc = se_interp_points_dev(m, b.a.as<double>(), nd, npts, b.b.as<double>(), mode, b.c.as<double>());
    if (!rc) rc = down(s, out, b.c, (size_t)npts * nd * sizeof(double));
    return rc;
}
