// mesh.hip -- context, GLL basis, mesh coordinates, global numbering and the
// exact per-element Jacobian tables used by the operator kernels.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <mutex>

#include "se_common.h"

namespace se {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int DevBuf::alloc(size_t b) {
    if (b <= bytes && ptr) return SE_OK;
    release();
    if (b == 0) return SE_OK;
    hipError_t e = hipMalloc(&ptr, b);
    if (e != hipSuccess) {
        ptr = nullptr;
        return fail(SE_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    bytes = b;
    return SE_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

// ---------------------------------------------------------------------------
// GLL basis on the host.  Same algorithm and operation order as
// Legendre::Legendre (src/base/Legendre.cpp:15-95): Legendre monomial
// coefficients from the three-term recurrence (Jacobi.cpp:21-36), interior
// abscissae as roots of P_N' by Newton from a 1e-3 upward scan
// (Element.cpp:54-77), mirrored and sorted, weights 2/(N(N+1)P_N(x)^2),
// derivative matrix D[i][j] = l_j'(x_i).  The operation order is kept so that
// the abscissae match the reference to the last bit (they seed the mesh
// coordinates, and hence the departure-element search).
// ---------------------------------------------------------------------------
namespace {

struct Poly {
    std::vector<double> c;
};

Poly recur(int n, const Poly& m1, const Poly& m2) {
    const double a = 0.0;
    Poly p;
    p.c.assign(n + 1, 0.0);
    double c1 = 2 * n * (2 * a + n) * (2 * a + 2 * n - 2);
    double c2 = (2 * a + 2 * n - 1) * (2 * a + 2 * n) * (2 * a + 2 * n - 2);
    double c3 = 2 * (a + n - 1) * (a + n - 1) * (2 * a + 2 * n);
    c2 /= c1;
    c3 /= c1;
    for (size_t i = 0; i < m1.c.size(); i++) p.c[i + 1] += c2 * m1.c[i];
    for (size_t i = 0; i < m2.c.size(); i++) p.c[i] -= c3 * m2.c[i];
    return p;
}

// parity-definite evaluations: the highest-degree term and every second one
double peval(const std::vector<double>& c, double x) {
    int n = (int)c.size() - 1;
    double s = 0.0;
    for (int i = n; i >= 0; i -= 2) s += c[i] * std::pow(x, i);
    return s;
}
double pderiv(const std::vector<double>& c, double x) {
    int n = (int)c.size() - 1;
    double s = 0.0;
    for (int i = n; i >= 1; i -= 2) s += c[i] * i * std::pow(x, i - 1);
    return s;
}
double pderiv2(const std::vector<double>& c, double x) {
    int n = (int)c.size() - 1;
    double s = 0.0;
    for (int i = n; i >= 2; i -= 2) s += c[i] * i * (i - 1) * std::pow(x, i - 2);
    return s;
}

double newton(double x, const std::vector<double>& c) {
    for (int it = 0; it < 100000; it++) {
        double prev = x;
        x = prev - pderiv(c, prev) / pderiv2(c, prev);
        if (std::fabs(x - prev) < 1.0e-12) return x;
    }
    return x;
}

}  // namespace

int gll_build(int N, GLL& g) {
    if (N < 2 || N > 16) return fail(SE_ERR_UNSUPPORTED, "GLL order must be in [2,16]");
    std::vector<Poly> P(N + 2);
    P[0].c = {1.0};
    P[1].c = {0.0, 1.0};
    for (int i = 2; i <= N + 1; i++) P[i] = recur(i, P[i - 1], P[i - 2]);
    const std::vector<double>& pN = P[N].c;
    std::vector<double> x(N + 1, 0.0);
    x[0] = -1.0;
    int half = ((N + 1) % 2 == 1) ? N / 2 : (N + 1) / 2;
    for (int i = 1; i < half; i++) {
        double x0 = x[i - 1];
        for (;;) {
            x0 += 1.0e-3;
            x[i] = newton(x0, pN);
            bool seen = false;
            for (int k = 0; k < i; k++)
                if (std::fabs(x[k] - x[i]) < 1.0e-6) seen = true;
            if (!seen && !(x[i] > 0.0 - 1.0e-3)) break;
        }
        x[N - i] = -x[i];
    }
    x[N] = 1.0;
    if ((N + 1) % 2 == 1) x[(N + 1) / 2] = 0.0;
    // ascending order by repeated selection of the smallest remaining entry
    {
        std::vector<double> out(N + 1);
        std::vector<char> used(N + 1, 0);
        int pick = 0;
        for (int i = 0; i <= N; i++) {
            double best = 1.0e+99;
            for (int j = 0; j <= N; j++)
                if (!used[j] && x[j] < best) {
                    best = x[j];
                    pick = j;
                }
            out[i] = best;
            used[pick] = 1;
        }
        x = out;
    }
    g.N = N;
    g.x = x;
    g.w.assign(N + 1, 0.0);
    g.w[0] = 2.0 / ((N + 1) * N);
    for (int i = 1; i < N; i++) g.w[i] = 2.0 / ((N + 1) * N * peval(pN, x[i]) * peval(pN, x[i]));
    g.w[N] = 2.0 / ((N + 1) * N);
    // the cardinal-function polynomial is regenerated from P_{N-1}, P_{N-2}
    Poly card = recur(N, P[N - 1], P[N - 2]);
    g.pn = card.c;
    g.pnx.assign(N + 1, 0.0);
    for (int j = 0; j <= N; j++) g.pnx[j] = peval(g.pn, x[j]);
    g.D.assign((N + 1) * (N + 1), 0.0);
    for (int i = 0; i <= N; i++)
        for (int j = 0; j <= N; j++) {
            double v;
            if (i == j && i == 0) v = -0.25 * N * (N + 1);
            else if (i == j && i == N) v = +0.25 * N * (N + 1);
            else if (i == j) v = 0.0;
            else v = peval(g.pn, x[i]) / (peval(g.pn, x[j]) * (x[i] - x[j]));
            g.D[i * (N + 1) + j] = v;
        }
    return SE_OK;
}

}  // namespace se

using namespace se;

// ---------------------------------------------------------------------------
// device kernels: numbering maps and vertex coordinates
// ---------------------------------------------------------------------------
namespace {

struct MapGeom {
    int nv0, nv1, nU0, per0, per1, bcb, bct, bcl, bcr;
};

// unknown index of one field-grid line coordinate, or -1 on a Dirichlet line.
// periodic: the last line aliases line 0; otherwise Dirichlet lines are cut
// off and the remaining lines are numbered in order (BCs.cpp:91-172).
__device__ __host__ inline int line_unknown(int I, int nv, int per, int bc_lo, int bc_hi) {
    if (per) return (I == nv - 1) ? 0 : I;
    if (bc_lo && I == 0) return -1;
    if (bc_hi && I == nv - 1) return -1;
    return I - bc_lo;
}

__global__ void k_build_maps(MapGeom g, int32_t* __restrict__ f2v, int32_t* __restrict__ v2f) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    int nV = g.nv0 * g.nv1;
    if (v >= nV) return;
    int I = v % g.nv0, J = v / g.nv0;
    int iu = line_unknown(I, g.nv0, g.per0, g.bcl, g.bcr);
    int ju = line_unknown(J, g.nv1, g.per1, g.bcb, g.bct);
    int u = (iu < 0 || ju < 0) ? -1 : ju * g.nU0 + iu;
    f2v[v] = u;
    // the canonical node of an unknown is the one the reference lists in
    // vecToFieldMap: not on a periodic duplicate line
    bool dupI = g.per0 && I == g.nv0 - 1;
    bool dupJ = g.per1 && J == g.nv1 - 1;
    if (u >= 0 && !dupI && !dupJ) v2f[u] = v;
}

__global__ void k_build_verts(int nv0, int nv1, const double* __restrict__ X, const double* __restrict__ Y,
                              double* __restrict__ verts) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv0 * nv1) return;
    verts[2 * (size_t)v] = X[v % nv0];
    verts[2 * (size_t)v + 1] = Y[v / nv0];
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI: context
// ---------------------------------------------------------------------------
extern "C" const char* se_last_error(void) { return g_last_error.c_str(); }
extern "C" int se_version(void) { return 1; }

extern "C" int se_ctx_create(int device, se_ctx** out) {
    SE_CHECK(out != nullptr, SE_ERR_ARG, "se_ctx_create: out is NULL");
    int count = 0;
    SE_HIP(hipGetDeviceCount(&count));
    SE_CHECK(device >= 0 && device < count, SE_ERR_ARG, "se_ctx_create: no such device");
    SE_HIP(hipSetDevice(device));
    se_ctx* c = new se_ctx;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(SE_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cu_count = prop.multiProcessorCount;
    if (c->cu_count <= 0) c->cu_count = 256;
    (void)hipEventCreate(&c->ev0);
    (void)hipEventCreate(&c->ev1);
    *out = c;
    return SE_OK;
}

extern "C" int se_ctx_destroy(se_ctx* c) {
    if (!c) return SE_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->d_stage.release();
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SE_OK;
}

extern "C" int se_ctx_synchronize(se_ctx* c) {
    SE_CHECK(c != nullptr, SE_ERR_ARG, "se_ctx_synchronize: NULL context");
    SE_HIP(hipStreamSynchronize(c->stream));
    return SE_OK;
}

extern "C" void* se_ctx_stream(se_ctx* c) { return c ? (void*)c->stream : nullptr; }

// Plain streaming copy of known size: the yardstick for the HBM byte counters
// (FETCH_SIZE / WRITE_SIZE) and the attainable copy bandwidth.
// Each thread moves four 16-byte words per step (four loads in flight).
static __global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ in, float4* __restrict__ out,
                                                          size_t n) {
    const size_t stride = (size_t)gridDim.x * 1024;
    size_t i = (size_t)blockIdx.x * 1024 + threadIdx.x;
    for (; i + 768 < n; i += stride) {
        const float4 v0 = in[i], v1 = in[i + 256], v2 = in[i + 512], v3 = in[i + 768];
        out[i] = v0;
        out[i + 256] = v1;
        out[i + 512] = v2;
        out[i + 768] = v3;
    }
    for (; i < n; i += 256) out[i] = in[i];  // the last partial step
}

extern "C" int se_ctx_copy_probe(se_ctx* c, size_t bytes, int reps, double* us) {
    SE_CHECK(c && us, SE_ERR_ARG, "se_ctx_copy_probe: NULL argument");
    SE_CHECK(reps > 0 && bytes >= 16 && bytes % 16 == 0, SE_ERR_ARG,
             "se_ctx_copy_probe: reps must be positive and bytes a positive multiple of 16");
    SE_HIP(hipSetDevice(c->device));
    se::DevBuf a, b;
    if (int rc = a.alloc(bytes)) return rc;
    if (int rc = b.alloc(bytes)) {
        a.release();
        return rc;
    }
    const size_t n = bytes / 16;
    const int grid = (int)std::min<size_t>((n + 1023) / 1024, (size_t)c->cu_count * 16);
    hipError_t e = hipMemsetAsync(a.ptr, 0, bytes, c->stream);
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, c->stream, a.as<const float4>(), b.as<float4>(), n);
    if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
    for (int i = 0; i < reps && e == hipSuccess; i++) {
        hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, c->stream, a.as<const float4>(), b.as<float4>(), n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(c->ev1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    (void)hipStreamSynchronize(c->stream);
    a.release();
    b.release();
    if (e != hipSuccess) return se::fail(SE_ERR_HIP, std::string("se_ctx_copy_probe: ") + hipGetErrorString(e));
    *us = 1e3 * (double)ms / reps;
    return SE_OK;
}

// ---------------------------------------------------------------------------
// C ABI: mesh
// ---------------------------------------------------------------------------
extern "C" int se_mesh_create(se_ctx* ctx, const int nEls[2], int N, const double min[2], const double max[2],
                              const int periodic[2], const int dirichlet[4], se_mesh** out) {
    SE_CHECK(ctx && nEls && min && max && periodic && dirichlet && out, SE_ERR_ARG, "se_mesh_create: NULL argument");
    SE_CHECK(nEls[0] >= 1 && nEls[1] >= 1, SE_ERR_ARG, "se_mesh_create: nEls must be >= 1");
    SE_CHECK(max[0] > min[0] && max[1] > min[1], SE_ERR_ARG, "se_mesh_create: empty domain");
    SE_CHECK((long)nEls[0] * N + 1 < 46341 && (long)nEls[1] * N + 1 < 46341, SE_ERR_ARG,
             "se_mesh_create: mesh too large for 32-bit node indices");
    int bcb = dirichlet[0] != 0, bct = dirichlet[1] != 0, bcl = dirichlet[2] != 0, bcr = dirichlet[3] != 0;
    int per0 = periodic[0] != 0, per1 = periodic[1] != 0;
    // BCs.cpp:29-36: periodic and Dirichlet on the same direction is an error
    SE_CHECK(!((bcb || bct) && per1), SE_ERR_UNSUPPORTED, "both periodic and non-periodic bcs in the vertical direction");
    SE_CHECK(!((bcl || bcr) && per0), SE_ERR_UNSUPPORTED, "both periodic and non-periodic bcs in the horizontal direction");
    // BCs.cpp:91-172 assigns no map for y-periodic-only meshes and exits for
    // an x-periodic mesh with a bottom wall and no top wall
    SE_CHECK(!(per1 && !per0), SE_ERR_UNSUPPORTED, "y-periodic-only meshes are not supported by the reference numbering");
    SE_CHECK(!(per0 && !per1 && bcb && !bct), SE_ERR_UNSUPPORTED, "x-periodic mesh with only a bottom wall is rejected by the reference numbering");
    SE_HIP(hipSetDevice(ctx->device));

    se_mesh* m = new se_mesh;
    m->ctx = ctx;
    int rc = gll_build(N, m->gll);
    if (rc != SE_OK) {
        delete m;
        return rc;
    }
    m->N = N;
    m->nl = (N + 1) * (N + 1);
    for (int d = 0; d < 2; d++) {
        m->nEls[d] = nEls[d];
        m->nVerts[d] = nEls[d] * N + 1;
        m->min[d] = min[d];
        m->max[d] = max[d];
        m->periodic[d] = periodic[d] != 0;
        m->dx[d] = (max[d] - min[d]) / nEls[d];
    }
    m->nElsTotal = nEls[0] * nEls[1];
    m->nV = m->nVerts[0] * m->nVerts[1];
    // effective wall flags: a periodic direction marks its last line (BCs.cpp:39-44)
    m->bc[0] = bcb;
    m->bc[1] = bct || per1;
    m->bc[2] = bcl;
    m->bc[3] = bcr || per0;
    int s0 = 0;
    if (m->bc[0]) s0 += m->nVerts[0];
    if (m->bc[1]) s0 += m->nVerts[0];
    if (m->bc[2]) s0 += m->nVerts[1];
    if (m->bc[3]) s0 += m->nVerts[1];
    if (m->bc[0] && m->bc[2]) s0--;
    if (m->bc[0] && m->bc[3]) s0--;
    if (m->bc[1] && m->bc[2]) s0--;
    if (m->bc[1] && m->bc[3]) s0--;
    m->size0 = s0;
    m->nU[0] = per0 ? m->nVerts[0] - 1 : m->nVerts[0] - bcl - bcr;
    m->nU[1] = per1 ? m->nVerts[1] - 1 : m->nVerts[1] - bcb - bct;
    m->off[0] = per0 ? 0 : bcl;
    m->off[1] = per1 ? 0 : bcb;
    m->n = m->nU[0] * m->nU[1];
    if (m->n != m->nV - m->size0) {
        delete m;
        return fail(SE_ERR_UNSUPPORTED, "numbering inconsistent with BCs::size");
    }

    // field-line coordinates, Mesh.cpp:63-72:  min + el*dx + ((xi+1)/2)*dx
    const std::vector<double>& xi = m->gll.x;
    m->X.resize(m->nVerts[0]);
    m->Y.resize(m->nVerts[1]);
    for (int I = 0; I < m->nVerts[0]; I++)
        m->X[I] = m->min[0] + (I / N) * m->dx[0] + ((xi[I % N] + 1.0) / 2.0) * m->dx[0];
    for (int J = 0; J < m->nVerts[1]; J++)
        m->Y[J] = m->min[1] + (J / N) * m->dx[1] + ((xi[J % N] + 1.0) / 2.0) * m->dx[1];

    // Jacobian entries exactly as Mesh::ShapeFuncDerivs / DetJac sum them
    // (Mesh.cpp:183-198,217-226): at quadrature point (qx,qy) of element
    // (ex,ey) the x-derivative row of the local gradient is non-zero only on
    // the element row qy, the y-derivative column only on column qx, so
    //   J00(ex,qx)    = sum_ix D[qx][ix] X[ex N + ix]
    //   J01(ey,qy,qx) = sum_ix D[qx][ix] Y[ey N + qy]
    //   J10(ex,qx,qy) = sum_iy D[qy][iy] X[ex N + qx]
    //   J11(ey,qy)    = sum_iy D[qy][iy] Y[ey N + iy]
    // with terms added in ascending local index from 0.0 (the zero entries of
    // the reference's dense sum do not change these values).
    const int n1 = N + 1;
    const std::vector<double>& D = m->gll.D;
    m->J00.assign((size_t)nEls[0] * n1, 0.0);
    m->J10.assign((size_t)nEls[0] * n1 * n1, 0.0);
    m->J11.assign((size_t)nEls[1] * n1, 0.0);
    m->J01.assign((size_t)nEls[1] * n1 * n1, 0.0);
    for (int ex = 0; ex < nEls[0]; ex++)
        for (int qx = 0; qx < n1; qx++) {
            double s = 0.0;
            for (int i = 0; i < n1; i++) s += D[qx * n1 + i] * m->X[ex * N + i];
            m->J00[(size_t)ex * n1 + qx] = s;
            for (int qy = 0; qy < n1; qy++) {
                double t = 0.0;
                for (int i = 0; i < n1; i++) t += D[qy * n1 + i] * m->X[ex * N + qx];
                m->J10[((size_t)ex * n1 + qx) * n1 + qy] = t;
            }
        }
    for (int ey = 0; ey < nEls[1]; ey++)
        for (int qy = 0; qy < n1; qy++) {
            double s = 0.0;
            for (int i = 0; i < n1; i++) s += D[qy * n1 + i] * m->Y[ey * N + i];
            m->J11[(size_t)ey * n1 + qy] = s;
            for (int qx = 0; qx < n1; qx++) {
                double t = 0.0;
                for (int i = 0; i < n1; i++) t += D[qx * n1 + i] * m->Y[ey * N + qy];
                m->J01[((size_t)ey * n1 + qy) * n1 + qx] = t;
            }
        }

    auto upload = [&](DevBuf& b, const void* src, size_t bytes) -> int {
        int r = b.alloc(bytes);
        if (r != SE_OK) return r;
        hipError_t e = hipMemcpy(b.ptr, src, bytes, hipMemcpyHostToDevice);
        return e == hipSuccess ? SE_OK : fail(SE_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    };
    std::vector<double> basis;
    basis.insert(basis.end(), m->gll.x.begin(), m->gll.x.end());
    basis.insert(basis.end(), m->gll.w.begin(), m->gll.w.end());
    basis.insert(basis.end(), m->gll.D.begin(), m->gll.D.end());
    basis.insert(basis.end(), m->gll.pnx.begin(), m->gll.pnx.end());
    basis.insert(basis.end(), m->gll.pn.begin(), m->gll.pn.end());
    for (int k = 0; k <= m->N; k++)  // D transposed
        for (int q = 0; q <= m->N; q++) basis.push_back(m->gll.D[(size_t)q * (m->N + 1) + k]);
    rc = SE_OK;
    if (!rc) rc = upload(m->d_X, m->X.data(), m->X.size() * sizeof(double));
    if (!rc) rc = upload(m->d_Y, m->Y.data(), m->Y.size() * sizeof(double));
    if (!rc) rc = upload(m->d_J00, m->J00.data(), m->J00.size() * sizeof(double));
    if (!rc) rc = upload(m->d_J10, m->J10.data(), m->J10.size() * sizeof(double));
    if (!rc) rc = upload(m->d_J11, m->J11.data(), m->J11.size() * sizeof(double));
    if (!rc) rc = upload(m->d_J01, m->J01.data(), m->J01.size() * sizeof(double));
    if (!rc) rc = upload(m->d_basis, basis.data(), basis.size() * sizeof(double));
    if (!rc) rc = m->d_f2v.alloc(sizeof(int32_t) * (size_t)m->nV);
    if (!rc) rc = m->d_v2f.alloc(sizeof(int32_t) * (size_t)(m->n > 0 ? m->n : 1));
    if (!rc) rc = m->d_verts.alloc(sizeof(double) * 2 * (size_t)m->nV);
    if (rc) {
        se_mesh_destroy(m);
        return rc;
    }
    MapGeom g{m->nVerts[0], m->nVerts[1], m->nU[0], per0, per1, m->bc[0] && !per1, bct, bcl, bcr};
    int threads = 256, blocks = (m->nV + threads - 1) / threads;
    hipLaunchKernelGGL(k_build_maps, dim3(blocks), dim3(threads), 0, ctx->stream, g, m->d_f2v.as<int32_t>(),
                       m->d_v2f.as<int32_t>());
    hipLaunchKernelGGL(k_build_verts, dim3(blocks), dim3(threads), 0, ctx->stream, m->nVerts[0], m->nVerts[1],
                       m->d_X.as<double>(), m->d_Y.as<double>(), m->d_verts.as<double>());
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        se_mesh_destroy(m);
        return fail(SE_ERR_HIP, std::string("mesh kernels: ") + hipGetErrorString(e));
    }
    *out = m;
    return SE_OK;
}

extern "C" int se_mesh_destroy(se_mesh* m) {
    if (!m) return SE_OK;
    if (m->ctx) (void)hipSetDevice(m->ctx->device);
    m->d_f2v.release();
    m->d_v2f.release();
    m->d_verts.release();
    m->d_X.release();
    m->d_Y.release();
    m->d_J00.release();
    m->d_J10.release();
    m->d_J11.release();
    m->d_J01.release();
    m->d_basis.release();
    m->d_sl.release();
    m->d_red.release();
    delete m;
    return SE_OK;
}

extern "C" int se_mesh_get_info(const se_mesh* m, se_mesh_info* o) {
    SE_CHECK(m && o, SE_ERR_ARG, "se_mesh_get_info: NULL argument");
    o->N = m->N;
    o->nEls[0] = m->nEls[0];
    o->nEls[1] = m->nEls[1];
    o->nVerts[0] = m->nVerts[0];
    o->nVerts[1] = m->nVerts[1];
    o->nV = m->nV;
    o->nElsTotal = m->nElsTotal;
    o->n = m->n;
    o->size0 = m->size0;
    return SE_OK;
}

extern "C" int se_mesh_gll(const se_mesh* m, double* a, double* w, double* D) {
    SE_CHECK(m, SE_ERR_ARG, "se_mesh_gll: NULL mesh");
    if (a) std::memcpy(a, m->gll.x.data(), sizeof(double) * m->gll.x.size());
    if (w) std::memcpy(w, m->gll.w.data(), sizeof(double) * m->gll.w.size());
    if (D) std::memcpy(D, m->gll.D.data(), sizeof(double) * m->gll.D.size());
    return SE_OK;
}

extern "C" int se_mesh_maps(const se_mesh* m, int32_t* f2v, int32_t* v2f, int32_t* elNodes) {
    SE_CHECK(m, SE_ERR_ARG, "se_mesh_maps: NULL mesh");
    SE_HIP(hipSetDevice(m->ctx->device));
    if (f2v) SE_HIP(hipMemcpy(f2v, m->d_f2v.ptr, sizeof(int32_t) * (size_t)m->nV, hipMemcpyDeviceToHost));
    if (v2f && m->n > 0) SE_HIP(hipMemcpy(v2f, m->d_v2f.ptr, sizeof(int32_t) * (size_t)m->n, hipMemcpyDeviceToHost));
    if (elNodes) {
        // Mesh::ElNodes closed form (Mesh.cpp:138-152)
        const int N = m->N;
        for (int el = 0; el < m->nElsTotal; el++) {
            int n0 = N * (el / m->nEls[0]) * m->nVerts[0] + N * (el % m->nEls[0]);
            int k = 0;
            for (int jj = 0; jj <= N; jj++)
                for (int ii = 0; ii <= N; ii++) elNodes[(size_t)el * m->nl + k++] = n0 + jj * m->nVerts[0] + ii;
        }
    }
    return SE_OK;
}

extern "C" int se_mesh_verts(const se_mesh* m, double* verts) {
    SE_CHECK(m && verts, SE_ERR_ARG, "se_mesh_verts: NULL argument");
    SE_HIP(hipSetDevice(m->ctx->device));
    SE_HIP(hipMemcpy(verts, m->d_verts.ptr, sizeof(double) * 2 * (size_t)m->nV, hipMemcpyDeviceToHost));
    return SE_OK;
}

extern "C" const int32_t* se_mesh_dev_field_to_vec(const se_mesh* m) { return m ? m->d_f2v.as<int32_t>() : nullptr; }
extern "C" const double* se_mesh_dev_verts(const se_mesh* m) { return m ? m->d_verts.as<double>() : nullptr; }

// ---------------------------------------------------------------------------
// field <-> unknown vector
// ---------------------------------------------------------------------------
namespace {

// Vector::UpdateVec: every non-Dirichlet, non-duplicate node writes its unknown
__global__ void k_field_to_vec(int n, const int32_t* __restrict__ v2f, const double* __restrict__ field,
                               double* __restrict__ vec) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) vec[i] = field[v2f[i]];
}

// Vector::UpdateField followed by Field::PeriodicUpdate: unknown nodes take
// their unknown's value; a node on a periodic duplicate line then takes the
// value of its source node on line 0 (for a Dirichlet source, the source's
// own field value), which is what the copy of Field.cpp:232-252 produces.
__global__ void k_vec_to_field(int nv0, int nv1, int per0, int per1, const int32_t* __restrict__ f2v,
                               const double* __restrict__ vec, double* __restrict__ field) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv0 * nv1) return;
    int I = v % nv0, J = v / nv0;
    bool dup = (per0 && I == nv0 - 1) || (per1 && J == nv1 - 1);
    if (!dup) {
        int u = f2v[v];
        if (u >= 0) field[v] = vec[u];
        return;
    }
    int Is = (per0 && I == nv0 - 1) ? 0 : I;
    int Js = (per1 && J == nv1 - 1) ? 0 : J;
    int src = Js * nv0 + Is;
    int u = f2v[src];
    field[v] = (u >= 0) ? vec[u] : field[src];
}

}  // namespace

extern "C" int se_field_to_vec_dev(se_mesh* m, const double* d_field, double* d_vec) {
    SE_CHECK(m && d_field && d_vec, SE_ERR_ARG, "se_field_to_vec_dev: NULL argument");
    if (m->n == 0) return SE_OK;
    hipLaunchKernelGGL(k_field_to_vec, dim3((m->n + 255) / 256), dim3(256), 0, m->ctx->stream, m->n,
                       m->d_v2f.as<int32_t>(), d_field, d_vec);
    SE_HIP(hipGetLastError());
    return SE_OK;
}

extern "C" int se_vec_to_field_dev(se_mesh* m, const double* d_vec, double* d_field) {
    SE_CHECK(m && d_field && d_vec, SE_ERR_ARG, "se_vec_to_field_dev: NULL argument");
    hipLaunchKernelGGL(k_vec_to_field, dim3((m->nV + 255) / 256), dim3(256), 0, m->ctx->stream, m->nVerts[0],
                       m->nVerts[1], m->periodic[0], m->periodic[1], m->d_f2v.as<int32_t>(), d_vec, d_field);
    SE_HIP(hipGetLastError());
    return SE_OK;
}
