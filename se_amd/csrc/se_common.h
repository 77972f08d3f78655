// se_common.h -- internal declarations shared by the se_amd HIP library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "se_amd.h"

namespace se {

// ---- error handling ------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define SE_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::se::fail(SE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define SE_CHECK(cond, code, msg)                  \
    do {                                           \
        if (!(cond)) return ::se::fail(code, msg); \
    } while (0)

// ---- device buffer (owned by a context) ------------------------------------
// Device allocation owned by its holder: freed by release() or on
// destruction (the owner makes its device current first); move-only.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes) {
        o.ptr = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            ptr = o.ptr;
            bytes = o.bytes;
            o.ptr = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    int alloc(size_t b);
    void release();
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// ---- GLL basis (host, setup) ----------------------------------------------
struct GLL {
    int N = 0;
    std::vector<double> x;    // abscissae, ascending
    std::vector<double> w;    // weights
    std::vector<double> D;    // D[i*(N+1)+j] = l_j'(x_i)
    std::vector<double> pn;   // monomial coefficients of P_N
    std::vector<double> pnx;  // P_N(x_j) (as evaluated by the parity-definite sum)
};
int gll_build(int N, GLL& out);

}  // namespace se

// ---- opaque handle bodies ----------------------------------------------------
struct se_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    // pinned staging used by host-pointer entry points
    void* h_stage = nullptr;
    size_t h_stage_bytes = 0;
    se::DevBuf d_stage;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // launch shapes chosen by the batched solver's autotuner, per
    // (order, log2 of the work items): later batches of that shape reuse them
    std::map<std::pair<int, int>, int> tuned_grid;
};

struct se_mesh {
    se_ctx* ctx = nullptr;
    int N = 0, nl = 0;
    int nEls[2] = {0, 0}, nVerts[2] = {0, 0};
    int nV = 0, nElsTotal = 0, n = 0, size0 = 0;
    int periodic[2] = {0, 0};
    int bc[4] = {0, 0, 0, 0};      // effective bottom, top, left, right
    int sl_interp = 0;             // SE_INTERP_* of the advected field in the SL steps
    int off[2] = {0, 0};           // first unknown column / row of the field grid
    int nU[2] = {0, 0};            // unknown grid extents
    double min[2] = {0, 0}, max[2] = {0, 0}, dx[2] = {0, 0};
    se::GLL gll;
    std::vector<double> X, Y;      // field column / row coordinates (reference verts)
    // exact reference Jacobian entries, see geometry tables in mesh.hip
    std::vector<double> J00, J10, J11, J01;
    se::DevBuf d_f2v, d_v2f, d_verts, d_X, d_Y;
    se::DevBuf d_J00, d_J10, d_J11, d_J01;
    se::DevBuf d_basis;            // x, w, D, pnx, pn, D^T  (N+1, N+1, (N+1)^2, N+1, N+1, (N+1)^2 doubles)
    se::DevBuf d_sl, d_red;        // SL velocity extrapolations, reduction scratch
};
