"""se_sl_advect2_dev_noalloc, the allocation-free second-order semi-Lagrangian
step (no counterpart in the reference: for callers that cannot allocate
device memory), against se_sl_advect2_dev on the same device inputs: bit for
bit, on two calls in a row (the second on a smaller mesh), with no change in
the device's free memory across the call, and a clear refusal above its
fixed scratch size.  Both calls go through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import se_amd
from se_amd import workloads as W

pytestmark = pytest.mark.gpu


def inputs(m, seed):
    a = m.gll()[0]
    cfg = dict(nEls=m.nEls, N=m.N, lo=m.lo, hi=m.hi)
    X, Y = W.mesh_coords(cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], a)
    rng = W.Rng(seed)
    q = W.gaussian_vortices(X, Y, 6, rng, cfg["lo"], cfg["hi"])
    qp = W.gaussian_vortices(X, Y, 6, rng, cfg["lo"], cfg["hi"])
    yy = np.repeat(Y, len(X))
    xx = np.tile(X, len(Y))
    vel = np.stack([np.sin(2 * np.pi * yy), 0.5 * np.cos(2 * np.pi * xx)], 1).ravel() * 0.2
    velp = 0.9 * vel
    return q, qp, vel, velp


def hip():
    """The HIP runtime libse_amd.so itself uses (loaded with it), for the
    test's own device buffers and the free-memory readings."""
    se_amd.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return h


class DeviceArray:
    """A device copy of a float64 numpy array (hipMalloc / hipMemcpy)."""

    def __init__(self, a):
        self.h = hip()
        self.a = np.ascontiguousarray(a, dtype=np.float64)
        self.p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(self.p), max(self.a.nbytes, 8)) == 0
        assert self.h.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.a)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, self.a.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        self.h.hipFree(self.p)


def free_memory():
    f, t = C.c_size_t(), C.c_size_t()
    assert hip().hipMemGetInfo(C.byref(f), C.byref(t)) == 0
    return f.value


def call(name, m, t, out, dt):
    """The entry point on device arrays t = (q, qPrev, vel, velPrev), into out."""
    rc = getattr(se_amd.lib(), name)(m.h, t[0].p, t[1].p, 1, t[2].p, t[3].p, C.c_double(dt), out.p)
    se_amd.check(rc, name)
    m.ctx.synchronize()


def both(m, args, dt, check_memory=False):
    """(no-alloc result, allocating result); with check_memory, the device's
    free memory is read right before and after the no-alloc call."""
    t = [DeviceArray(a) for a in args]
    out_n = DeviceArray(np.full(m.nV, np.nan))
    out_a = DeviceArray(np.full(m.nV, np.nan))
    m.ctx.synchronize()
    free0 = free_memory()
    call("se_sl_advect2_dev_noalloc", m, t, out_n, dt)
    free1 = free_memory()
    call("se_sl_advect2_dev", m, t, out_a, dt)
    res = out_n.get(), out_a.get()
    for x in t + [out_n, out_a]:
        x.free()
    if check_memory:
        assert free1 == free0, (free0, free1)
    return res


def test_noalloc_second_order_step_equals_allocating_one(gpu_ctx):
    # a channel like the served requests' (x-periodic, walls), then a smaller doubly periodic mesh
    big = se_amd.Mesh(gpu_ctx, (40, 36), 6, (0, 0), (1, 1), (1, 0), (1, 1, 0, 0))
    small = se_amd.Mesh(gpu_ctx, (12, 10), 5, (0, 0), (1, 1), (1, 1), (0, 0, 0, 0))
    assert big.nV <= se_amd.SL_NOALLOC_MAX_NODES and small.nV < big.nV
    got, ref = both(big, inputs(big, 3), 0.01)
    assert np.all(np.isfinite(ref)) and np.array_equal(got, ref)
    # second call, smaller input, on a mesh that has never allocated its own
    # scratch: no device memory changes hands during the call
    got, ref = both(small, inputs(small, 4), 0.02, check_memory=True)
    assert np.all(np.isfinite(ref)) and np.array_equal(got, ref)
    small.close()
    big.close()


def test_noalloc_refuses_meshes_beyond_its_scratch(gpu_ctx):
    m = se_amd.Mesh(gpu_ctx, (128, 128), 6, (0, 0), (1, 1), (1, 1), (0, 0, 0, 0))
    assert m.nV > se_amd.SL_NOALLOC_MAX_NODES
    t = [DeviceArray(np.zeros(k * m.nV)) for k in (1, 1, 2, 2)]
    out = DeviceArray(np.zeros(m.nV))
    with pytest.raises(se_amd.SEError, match="SE_SL_NOALLOC_MAX_NODES"):
        call("se_sl_advect2_dev_noalloc", m, t, out, 0.01)
    for x in t + [out]:
        x.free()
    m.close()
