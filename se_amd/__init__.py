"""se_amd -- MI355X (gfx950) native spectral-element hot path of davelee2804/SE.

Python view of the C ABI declared in include/se_amd.h (libse_amd.so, built
in-tree by se_amd/build.py).  The compute path is the HIP library; this module
only marshals numpy arrays (host entry points) or raw device pointers (the
*_dev entry points, e.g. from torch tensors) and raises on every non-zero
status.  There is no CPU fallback: without the built library every call fails.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libse_amd.so")

DETJ_LITERAL, DETJ_PHYSICAL = 0, 1
PREC_FP64, PREC_FP32_IR, PREC_BF16_IR, PREC_FP64_FP32C, PREC_FP64_BF16C = 0, 1, 2, 3, 4
PC_DEFAULT, PC_JACOBI, PC_FDM = 0, 1, 2
SL_NOALLOC_MAX_NODES = 524288  # SE_SL_NOALLOC_MAX_NODES of include/se_amd.h
INTERP_DEFAULT, INTERP_GLL, INTERP_BICUBIC = 0, 1, 2
SOLVE_CONVERGED, SOLVE_MAXIT, SOLVE_BREAKDOWN, SOLVE_NONFINITE, SOLVE_STAGNATED = 0, 1, 2, 3, 4

_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_F = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_lib = None


class SEError(RuntimeError):
    pass


class SolveOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("maxit", C.c_int), ("precision", C.c_int), ("null_space", C.c_int),
                ("check_every", C.c_int), ("inner_rtol", C.c_double), ("max_outer", C.c_int),
                ("history", C.c_void_p), ("preconditioner", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_int), ("outer_iters", C.c_int),
                ("true_rel_residual", C.c_double), ("seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SolveInfos:
    """Read-only sequence of per-member solve results (dicts with SolveInfo's
    fields), converted on access: a batch call of hundreds of members then
    costs no per-member Python work until the results are read."""

    _names = [k for k, _ in SolveInfo._fields_]

    def __init__(self, arr, n):
        self._arr = arr
        self._n = n

    def __len__(self):
        return self._n

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[j] for j in range(*i.indices(self._n))]
        if i < 0:
            i += self._n
        if not 0 <= i < self._n:
            raise IndexError(i)
        r = self._arr[i]
        return {k: getattr(r, k) for k in self._names}

    def __iter__(self):
        return (self[i] for i in range(self._n))

    def __eq__(self, other):
        return list(self) == list(other)

    def __repr__(self):
        return repr(list(self))


class ServerOpts(C.Structure):
    _fields_ = [("memory_budget", C.c_double), ("max_active", C.c_int), ("precision", C.c_int),
                ("rtol", C.c_double), ("detj_mode", C.c_int), ("preconditioner", C.c_int)]


class QG2Request(C.Structure):
    _fields_ = [("nEls", C.c_int * 2), ("N", C.c_int), ("min", C.c_double * 2), ("max", C.c_double * 2),
                ("periodic", C.c_int * 2), ("bc", C.c_int * 4), ("F1", C.c_double), ("F2", C.c_double),
                ("beta", C.c_double), ("steps", C.c_int), ("dt", C.c_double), ("psi1", C.c_void_p),
                ("omega1", C.c_void_p), ("psi2", C.c_void_p), ("omega2", C.c_void_p), ("E2", C.c_double),
                ("r", C.c_double), ("nu", C.c_double), ("tau_e", C.c_void_p)]


class RequestResult(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int), ("steps_done", C.c_int), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("inner_iters", C.c_int64),
                ("unknowns", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("admitted", C.c_int64), ("completed", C.c_int64),
                ("steps", C.c_int64), ("ticks", C.c_int64), ("device_bytes", C.c_double)]


class MeshInfo(C.Structure):
    _fields_ = [("N", C.c_int), ("nEls", C.c_int * 2), ("nVerts", C.c_int * 2), ("nV", C.c_int),
                ("nElsTotal", C.c_int), ("n", C.c_int), ("size0", C.c_int)]


def lib():
    """Load libse_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SEError("se_amd: %s is not built; run `python -m se_amd.build`" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        pvp = C.POINTER(C.c_void_p)
        sig = {
            "se_last_error": (C.c_char_p, []),
            "se_version": (i, []),
            "se_ctx_create": (i, [i, pvp]),
            "se_ctx_destroy": (i, [vp]),
            "se_ctx_synchronize": (i, [vp]),
            "se_ctx_stream": (vp, [vp]),
            "se_ctx_copy_probe": (i, [vp, C.c_size_t, i, C.POINTER(d)]),
            "se_mesh_create": (i, [vp, _I, i, _D, _D, _I, _I, pvp]),
            "se_mesh_destroy": (i, [vp]),
            "se_mesh_get_info": (i, [vp, C.POINTER(MeshInfo)]),
            "se_mesh_gll": (i, [vp, _D, _D, _D]),
            "se_mesh_maps": (i, [vp, vp, vp, vp]),
            "se_mesh_verts": (i, [vp, _D]),
            "se_mesh_dev_field_to_vec": (vp, [vp]),
            "se_mesh_dev_verts": (vp, [vp]),
            "se_field_to_vec_dev": (i, [vp, vp, vp]),
            "se_vec_to_field_dev": (i, [vp, vp, vp]),
            "se_helm_create": (i, [vp, d, d, i, pvp]),
            "se_helm_destroy": (i, [vp]),
            "se_helm_apply": (i, [vp, _D, _D]),
            "se_helm_apply_dev": (i, [vp, vp, vp]),
            "se_helm_apply_inner": (i, [vp, i, _F, _F, C.POINTER(i)]),
            "se_helm_apply_fdm": (i, [vp, i, _D, _D]),
            "se_helm_fdm_probe": (i, [vp, i, _D, _D, C.POINTER(d), C.POINTER(d)]),
            "se_batch_fdm_probe": (i, [vp, i, _D, _D, C.POINTER(d), C.POINTER(d)]),
            "se_helm_inv_diag": (i, [vp, _D]),
            "se_helm_set_dirichlet": (i, [vp, _D]),
            "se_helm_set_dirichlet_dev": (i, [vp, vp]),
            "se_helm_lift": (i, [vp, _D]),
            "se_helm_rhs": (i, [vp, _D, _D]),
            "se_helm_rhs_dev": (i, [vp, vp, vp]),
            "se_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
            "se_helm_pcg": (i, [vp, _D, _D, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
            "se_helm_pcg_dev": (i, [vp, vp, vp, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
            "se_helm_solve": (i, [vp, _D, _D, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
            "se_helm_solve_dev": (i, [vp, vp, vp, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
            "se_batch_create": (i, [vp, i, vp, pvp]),
            "se_batch_destroy": (i, [vp]),
            "se_batch_size": (i, [vp]),
            "se_batch_launch_shape": (i, [vp, C.POINTER(i), C.POINTER(i)]),
            "se_batch_pcg_dev": (i, [vp, vp, vp, C.POINTER(SolveOpts), vp]),
            "se_batch_pcg": (i, [vp, vp, vp, C.POINTER(SolveOpts), vp]),
            "se_batch_probe": (i, [vp, i, i, i, _I, _D, _D, _D, _D]),
            "se_helm_probe": (i, [vp, i, i, C.POINTER(d), C.POINTER(d), C.POINTER(d), C.POINTER(d)]),
            "se_helm_set_svv": (i, [vp, d, i]),
            "se_qg2_create": (i, [vp, d, d, d, i, pvp]),
            "se_qg2_destroy": (i, [vp]),
            "se_qg2_set_state": (i, [vp, _D, _D, _D, _D]),
            "se_qg2_get_state": (i, [vp, _D, _D, _D, _D]),
            "se_qg2_state_dev": (i, [vp, pvp, pvp, pvp, pvp]),
            "se_qg2_timestep": (i, [vp, d, C.POINTER(d)]),
            "se_qg2_step": (i, [vp, d, C.POINTER(SolveOpts), vp]),
            "se_qg2_set_dissipation": (i, [vp, d, d, d, vp]),
            "se_qg_create": (i, [vp, d, d, i, pvp]),
            "se_qg_destroy": (i, [vp]),
            "se_qg_set_state": (i, [vp, _D, _D]),
            "se_qg_get_state": (i, [vp, _D, _D]),
            "se_qg_timestep": (i, [vp, d, C.POINTER(d)]),
            "se_qg_step": (i, [vp, d, C.POINTER(SolveOpts), vp]),
            "se_qg_set_dissipation": (i, [vp, d, d]),
            "se_qg_set_topography": (i, [vp, vp]),
            "se_server_opts_default": (None, [C.POINTER(ServerOpts)]),
            "se_server_create": (i, [i, C.POINTER(ServerOpts), pvp]),
            "se_server_submit": (i, [vp, C.POINTER(QG2Request), C.POINTER(C.c_int64)]),
            "se_server_wait": (i, [vp, C.c_int64, C.POINTER(RequestResult), vp, vp, vp, vp]),
            "se_server_get_stats": (i, [vp, C.POINTER(ServerStats)]),
            "se_server_destroy": (i, [vp]),
            "se_sl_advect": (i, [vp, _D, i, _D, d, _D, vp]),
            "se_sl_advect_dev": (i, [vp, vp, i, vp, d, vp, vp]),
            "se_sl_advect2": (i, [vp, _D, _D, i, _D, _D, d, _D]),
            "se_sl_advect2_dev": (i, [vp, vp, vp, i, vp, vp, d, vp]),
            "se_sl_advect2_dev_noalloc": (i, [vp, vp, vp, i, vp, vp, d, vp]),
            "se_mesh_set_sl_interp": (i, [vp, i]),
            "se_interp_points": (i, [vp, _D, i, i, _D, i, _D]),
            "se_interp_points_dev": (i, [vp, vp, i, i, vp, i, vp]),
            "se_genvel": (i, [vp, _D, _D]),
            "se_genvel_dev": (i, [vp, vp, vp]),
            "se_calc_timestep": (i, [vp, _D, d, C.POINTER(d)]),
            "se_calc_timestep_dev": (i, [vp, vp, d, C.POINTER(d)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().se_last_error().decode(errors="replace")
        raise SEError("%s failed (%d): %s" % (what, rc, msg))


def _d(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def solve_opts(rtol=1e-10, maxit=100000, precision=PREC_FP64, null_space=False, check_every=32,
               inner_rtol=0.0, max_outer=1000, history=None, preconditioner=PC_DEFAULT):
    o = SolveOpts()
    lib().se_solve_opts_default(C.byref(o))
    o.rtol, o.maxit, o.precision, o.null_space = rtol, maxit, precision, int(bool(null_space))
    o.preconditioner = preconditioner
    o.check_every, o.inner_rtol, o.max_outer = check_every, inner_rtol, max_outer
    o.history = history.ctypes.data if history is not None else None
    return o


def _fdm_probe(fn, h, reps):
    us, fl = np.zeros(4), np.zeros(4)
    ua, ba = C.c_double(), C.c_double()
    check(fn(h, reps, us, fl, C.byref(ua), C.byref(ba)), "fdm probe")
    return dict(us_pass=us.tolist(), flops_pass=fl.tolist(), us_apply=ua.value, bytes_apply=ba.value)


IO_LIB_PATH = os.path.join(os.path.dirname(LIB_PATH), "libse_amd_io.so")
_io = None


def io_lib():
    """Load libse_amd_io.so (field files in the reference's HDF5 layout)."""
    global _io
    if _io is None:
        if not os.path.exists(IO_LIB_PATH):
            raise SEError("se_amd: %s is not built (needs HDF5); run `python -m se_amd.build`" % IO_LIB_PATH)
        L = C.CDLL(IO_LIB_PATH)
        i, pd, pi = C.c_int, _D, C.POINTER(C.c_int)
        sig = {"se_io_last_error": (C.c_char_p, []),
               "se_field_save": (i, [C.c_char_p, C.c_uint, pd, i, i, pi]),
               "se_field_read": (i, [C.c_char_p, C.c_uint, pd, i, i]),
               "se_field_info": (i, [C.c_char_p, C.c_uint, pi, pi, pi, pi]),
               "se_mesh_save": (i, [C.c_char_p, pi, pd, pd, pd]),
               "se_mesh_read": (i, [C.c_char_p, pi, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
               "se_xdmf_header": (i, [C.c_uint]),
               "se_xdmf_footer": (i, [C.c_uint]),
               "se_xdmf_write": (i, [C.c_char_p, pi, C.c_void_p, i, C.c_uint, C.c_double, C.c_double]),
               "se_xdmf_temporal": (i, [i, i])}
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype, f.argtypes = r, a
        _io = L
    return _io


def _io_check(rc, what):
    if rc:
        raise SEError("%s: %s" % (what, io_lib().se_io_last_error().decode()))


def field_save(name, step, vals, nVerts):
    """Field::Save: vals (nV,) or (nV, nDofs) node-major -> <name>.<step:05u>.h5."""
    v = np.ascontiguousarray(vals, dtype=np.float64)
    nV = v.shape[0]
    nd = 1 if v.ndim == 1 else v.shape[1]
    nv = (C.c_int * 2)(*nVerts)
    _io_check(io_lib().se_field_save(name.encode(), step, v.ravel(), nV, nd, nv), "se_field_save")


def field_read(name, step, nV, nDofs=1):
    """Field::Read: the (nV, nDofs) values of <name>.<step:05u>.h5."""
    out = np.zeros(nV * nDofs)
    _io_check(io_lib().se_field_read(name.encode(), step, out, nV, nDofs), "se_field_read")
    return out.reshape(nV, nDofs)


def field_info(name, step):
    """(nVertsTotal, nDofs, dimensions, mesh resolution) of a field file."""
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    res = (C.c_int * 2)()
    _io_check(io_lib().se_field_info(name.encode(), step, C.byref(a), C.byref(b), C.byref(c), res), "se_field_info")
    return a.value, b.value, c.value, (res[0], res[1])


def mesh_save(name, nVerts, lo, hi, verts):
    """Mesh::Save: <name>.h5 with the vertices (nVertsTotal, 2) and the linear-quad connectivity."""
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 2)
    nv = (C.c_int * 2)(*nVerts)
    _io_check(io_lib().se_mesh_save(name.encode(), nv, _d(lo), _d(hi), v.ravel()), "se_mesh_save")


def mesh_read(name):
    """(nVerts, min, max, vertices (nV, 2), connectivity (nQuads, 4)) of a mesh file."""
    nv = (C.c_int * 2)()
    _io_check(io_lib().se_mesh_read(name.encode(), nv, None, None, None, None), "se_mesh_read")
    nV, nq = nv[0] * nv[1], (nv[0] - 1) * (nv[1] - 1)
    lo, hi, verts = np.zeros(2), np.zeros(2), np.zeros((nV, 2))
    conn = np.zeros((nq, 4), dtype=np.int32)
    _io_check(io_lib().se_mesh_read(name.encode(), nv, lo.ctypes.data, hi.ctypes.data, verts.ctypes.data,
                                    conn.ctypes.data), "se_mesh_read")
    return (nv[0], nv[1]), lo, hi, verts, conn


class XdmfField(C.Structure):
    _fields_ = [("name", C.c_char_p), ("nDofs", C.c_int), ("vals", C.c_void_p)]


def xdmf_write(mesh_name, nVerts, fields, step, time, dt):
    """WriteXDMF: fields = [(name, vals (nV,) or (nV, nDofs))]; the field files, time.<step>.ts and
    the step's grid in XDMF.<step>.xmf (working directory)."""
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for _, v in fields]
    names = [n.encode() for n, _ in fields]
    F = (XdmfField * max(1, len(fields)))()
    for k, (nm, a) in enumerate(zip(names, arrs)):
        F[k].name, F[k].nDofs, F[k].vals = nm, 1 if a.ndim == 1 else a.shape[1], a.ctypes.data
    nv = (C.c_int * 2)(*nVerts)
    _io_check(io_lib().se_xdmf_write(mesh_name.encode(), nv, C.cast(F, C.c_void_p), len(fields), step, time, dt),
              "se_xdmf_write")


def xdmf_header(step):
    _io_check(io_lib().se_xdmf_header(step), "se_xdmf_header")


def xdmf_footer(step):
    _io_check(io_lib().se_xdmf_footer(step), "se_xdmf_footer")


def xdmf_temporal(n_steps, dump_every):
    _io_check(io_lib().se_xdmf_temporal(n_steps, dump_every), "se_xdmf_temporal")


class Context:
    """One HIP stream + the device memory of everything created from it."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().se_ctx_create(device, C.byref(h)), "se_ctx_create")
        self.h = h
        self._children = []  # weak references, in creation order

    def _own(self, obj):
        self._children.append(weakref.ref(obj))

    @property
    def stream(self):
        return lib().se_ctx_stream(self.h)

    def synchronize(self):
        check(lib().se_ctx_synchronize(self.h), "se_ctx_synchronize")

    def copy_probe(self, nbytes, reps=10):
        """Average microseconds of one plain device copy of nbytes (the
        yardstick for the HBM byte counters)."""
        us = C.c_double()
        check(lib().se_ctx_copy_probe(self.h, int(nbytes), int(reps), C.byref(us)), "se_ctx_copy_probe")
        return us.value

    def close(self):
        """Destroy the context and, first, everything still alive that was
        created from it (batches, operators, meshes: newest first)."""
        if self.h:
            for ref in reversed(self._children):
                obj = ref()
                if obj is not None:
                    obj.close()
            self._children = []
            lib().se_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mesh:
    """Mesh + BCs of the reference (legendre elements).  bc = (bottom, top, left, right)."""

    def __init__(self, ctx, nEls, N, lo, hi, periodic, bc=(0, 0, 0, 0)):
        self.ctx = ctx
        self.h = None
        h = C.c_void_p()
        check(lib().se_mesh_create(ctx.h, np.array(nEls, dtype=np.int32), N, _d(lo), _d(hi),
                                   np.array(periodic, dtype=np.int32), np.array(bc, dtype=np.int32),
                                   C.byref(h)), "se_mesh_create")
        self.h = h
        info = MeshInfo()
        check(lib().se_mesh_get_info(h, C.byref(info)), "se_mesh_get_info")
        self.N, self.nV, self.n, self.nElsTotal, self.size0 = info.N, info.nV, info.n, info.nElsTotal, info.size0
        self.nEls = (info.nEls[0], info.nEls[1])
        self.nVerts = (info.nVerts[0], info.nVerts[1])
        ctx._own(self)
        self.nl = (self.N + 1) ** 2
        self.lo, self.hi, self.periodic, self.bc = tuple(lo), tuple(hi), tuple(periodic), tuple(bc)

    def gll(self):
        N = self.N
        a, w, D = np.zeros(N + 1), np.zeros(N + 1), np.zeros((N + 1) ** 2)
        check(lib().se_mesh_gll(self.h, a, w, D), "se_mesh_gll")
        return a, w, D.reshape(N + 1, N + 1)

    def maps(self, el_nodes=False):
        f2v = np.zeros(self.nV, dtype=np.int32)
        v2f = np.zeros(max(self.n, 1), dtype=np.int32)
        en = np.zeros(self.nElsTotal * self.nl, dtype=np.int32) if el_nodes else None
        check(lib().se_mesh_maps(self.h, f2v.ctypes.data, v2f.ctypes.data, en.ctypes.data if el_nodes else None),
              "se_mesh_maps")
        if el_nodes:
            return f2v, v2f[: self.n], en.reshape(self.nElsTotal, self.nl)
        return f2v, v2f[: self.n]

    def verts(self):
        v = np.zeros(2 * self.nV)
        check(lib().se_mesh_verts(self.h, v), "se_mesh_verts")
        return v.reshape(self.nV, 2)

    def set_sl_interp(self, mode):
        """Interpolation of the advected field in the SL steps (INTERP_*)."""
        check(lib().se_mesh_set_sl_interp(self.h, mode), "se_mesh_set_sl_interp")

    def interp_points(self, vals, xy, nDofs=1, mode=INTERP_GLL):
        """Field values at points xy (k x 2): GLL (Field::InterpGlobal) or bicubic."""
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        out = np.zeros(len(xy) * nDofs)
        check(lib().se_interp_points(self.h, _d(vals), nDofs, len(xy), xy.ravel(), mode, out), "se_interp_points")
        return out.reshape(len(xy), nDofs)

    def advect(self, q, vel, dt, nDofs=1, want_dep=True):
        qsl = np.zeros(self.nV * nDofs)
        dep = np.zeros(self.nV, dtype=np.int32)
        check(lib().se_sl_advect(self.h, _d(q), nDofs, _d(vel), dt, qsl, dep.ctypes.data if want_dep else None),
              "se_sl_advect")
        return (qsl, dep) if want_dep else qsl

    def advect2(self, q, qprev, vel, velprev, dt, nDofs=1):
        qsl = np.zeros(self.nV * nDofs)
        check(lib().se_sl_advect2(self.h, _d(q), _d(qprev), nDofs, _d(vel), _d(velprev), dt, qsl), "se_sl_advect2")
        return qsl

    def genvel(self, psi):
        vel = np.zeros(2 * self.nV)
        check(lib().se_genvel(self.h, _d(psi), vel), "se_genvel")
        return vel.reshape(self.nV, 2)

    def calc_timestep(self, vel, courant):
        dt = C.c_double()
        check(lib().se_calc_timestep(self.h, _d(vel), courant, C.byref(dt)), "se_calc_timestep")
        return dt.value

    def close(self):
        if self.h:
            lib().se_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Helm:
    """Matrix-free c_mass*M + c_lap*K~ (HelmholtzEqn: c_mass=c, c_lap=nu*dt; PoissonEqn: 0, c)."""

    def __init__(self, mesh, c_mass, c_lap, detj_mode=DETJ_LITERAL):
        self.mesh = mesh
        self.h = None
        h = C.c_void_p()
        check(lib().se_helm_create(mesh.h, c_mass, c_lap, detj_mode, C.byref(h)), "se_helm_create")
        self.h = h
        mesh.ctx._own(self)

    def set_svv(self, dt, cutoff):
        """HelmholtzEqn's spectral vanishing viscosity term (FP64 solves only)."""
        check(lib().se_helm_set_svv(self.h, dt, cutoff), "se_helm_set_svv")

    def apply(self, x):
        y = np.zeros(self.mesh.n)
        check(lib().se_helm_apply(self.h, _d(x), y), "se_helm_apply")
        return y

    def apply_inner(self, x, precision=PREC_FP32_IR):
        """Separable low-precision inner operator: returns (sign*A~ x as float32, sign)."""
        xf = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(self.mesh.n, dtype=np.float32)
        sg = C.c_int(0)
        check(lib().se_helm_apply_inner(self.h, precision, xf, y, C.byref(sg)), "se_helm_apply_inner")
        return y, sg.value

    def apply_fdm(self, r, precision=PREC_FP64):
        """Fast-diagonalisation preconditioner: z ~= (sign*A)^-1 r (separable
        form; the null mode of a singular operator is left out)."""
        z = np.zeros(self.mesh.n)
        check(lib().se_helm_apply_fdm(self.h, precision, _d(r), z), "se_helm_apply_fdm")
        return z

    def fdm_probe(self, reps=20):
        """After an FDM solve: per-pass us and flops of the transforms, us and bytes of the FP64 apply."""
        return _fdm_probe(lib().se_helm_fdm_probe, self.h, reps)

    def probe(self, precision=PREC_FP32_IR, reps=50):
        """After a mixed solve: dict(us_dir, us_upd, bytes_dir, bytes_upd) of the inner kernels."""
        v = [C.c_double() for _ in range(4)]
        check(lib().se_helm_probe(self.h, precision, reps, *[C.byref(x) for x in v]), "se_helm_probe")
        return dict(us_dir=v[0].value, us_upd=v[1].value, bytes_dir=v[2].value, bytes_upd=v[3].value)

    def apply_dev(self, x_ptr, y_ptr):
        check(lib().se_helm_apply_dev(self.h, x_ptr, y_ptr), "se_helm_apply_dev")

    def inv_diag(self):
        d = np.zeros(self.mesh.n)
        check(lib().se_helm_inv_diag(self.h, d), "se_helm_inv_diag")
        return d

    def set_dirichlet(self, sol_field):
        check(lib().se_helm_set_dirichlet(self.h, _d(sol_field)), "se_helm_set_dirichlet")

    def lift(self):
        b = np.zeros(self.mesh.n)
        check(lib().se_helm_lift(self.h, b), "se_helm_lift")
        return b

    def rhs(self, rhs_field):
        f = np.zeros(self.mesh.n)
        check(lib().se_helm_rhs(self.h, _d(rhs_field), f), "se_helm_rhs")
        return f

    def pcg(self, f, **kw):
        hist = None
        if kw.pop("history", False):
            hist = np.zeros(kw.get("maxit", 100000) + 1)
        o = solve_opts(history=hist, **kw)
        x = np.zeros(self.mesh.n)
        info = SolveInfo()
        check(lib().se_helm_pcg(self.h, _d(f), x, C.byref(o), C.byref(info)), "se_helm_pcg")
        out = info.as_dict()
        if hist is not None:
            out["hist"] = hist[: info.iters + 1]
        return x, out

    def pcg_dev(self, f_ptr, x_ptr, **kw):
        o = solve_opts(**kw)
        info = SolveInfo()
        check(lib().se_helm_pcg_dev(self.h, f_ptr, x_ptr, C.byref(o), C.byref(info)), "se_helm_pcg_dev")
        return info.as_dict()

    def solve(self, rhs_field, sol_field, **kw):
        o = solve_opts(**kw)
        sol = _d(sol_field).copy()
        info = SolveInfo()
        check(lib().se_helm_solve(self.h, _d(rhs_field), sol, C.byref(o), C.byref(info)), "se_helm_solve")
        return sol, info.as_dict()

    def solve_dev(self, rhs_ptr, sol_ptr, **kw):
        o = solve_opts(**kw)
        info = SolveInfo()
        check(lib().se_helm_solve_dev(self.h, rhs_ptr, sol_ptr, C.byref(o), C.byref(info)), "se_helm_solve_dev")
        return info.as_dict()

    def close(self):
        if self.h:
            lib().se_helm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent mixed-precision solves of several Helm operators at once."""

    def __init__(self, ctx, helms):
        self.helms = list(helms)
        self.h = None
        arr = (C.c_void_p * max(1, len(self.helms)))(*[h.h.value for h in self.helms])
        h = C.c_void_p()
        check(lib().se_batch_create(ctx.h, len(self.helms), arr, C.byref(h)), "se_batch_create")
        self.h = h
        ctx._own(self)

    def _infos(self):
        return (SolveInfo * max(1, len(self.helms)))()

    def pcg(self, fs, precision=PREC_FP32_IR, **kw):
        """fs: list of host vectors; returns (list of solutions, list of info dicts)."""
        o = solve_opts(precision=precision, **kw)
        fs = [_d(f) for f in fs]
        xs = [np.zeros(h.mesh.n) for h in self.helms]
        fp = (C.c_void_p * max(1, len(fs)))(*[f.ctypes.data for f in fs])
        xp = (C.c_void_p * max(1, len(xs)))(*[x.ctypes.data for x in xs])
        infos = self._infos()
        check(lib().se_batch_pcg(self.h, fp, xp, C.byref(o), infos), "se_batch_pcg")
        return xs, [infos[i].as_dict() for i in range(len(self.helms))]

    def _ptrs(self, slot, ptrs):
        """ctypes array of device pointers, reused while the caller passes the same ones."""
        key = tuple(ptrs)
        cache = self.__dict__.setdefault("_ptr_cache", {})
        hit = cache.get(slot)
        if hit is None or hit[0] != key:
            hit = (key, (C.c_void_p * max(1, len(key)))(*key))
            cache[slot] = hit
        return hit[1]

    def pcg_dev(self, f_ptrs, x_ptrs, precision=PREC_FP32_IR, **kw):
        """Device vectors in and out; returns a SolveInfos sequence (one dict per member)."""
        o = solve_opts(precision=precision, **kw)
        infos = self._infos()
        check(lib().se_batch_pcg_dev(self.h, self._ptrs("f", f_ptrs), self._ptrs("x", x_ptrs), C.byref(o), infos),
              "se_batch_pcg_dev")
        return SolveInfos(infos, len(self.helms))

    def fdm_probe(self, reps=20):
        """After an FDM solve of the batch: as Helm.fdm_probe over all members."""
        return _fdm_probe(lib().se_batch_fdm_probe, self.h, reps)

    def probe(self, precision=PREC_FP32_IR, reps=50, max_groups=8):
        """Per order group: dict(order, us_dir, us_upd, bytes_dir, bytes_upd) (after a solve)."""
        o = np.zeros(max_groups, dtype=np.int32)
        a, b_, c, e = (np.zeros(max_groups) for _ in range(4))
        ng = lib().se_batch_probe(self.h, precision, reps, max_groups, o, a, b_, c, e)
        check(min(ng, 0), "se_batch_probe")
        return [dict(order=int(o[g]), us_dir=a[g], us_upd=b_[g], bytes_dir=c[g], bytes_upd=e[g]) for g in range(ng)]

    def launch_shape(self):
        """(direction, update) workgroups per CU chosen by the autotuner (0: one
        per item; -1: not yet tuned)."""
        d, u = C.c_int(), C.c_int()
        check(lib().se_batch_launch_shape(self.h, C.byref(d), C.byref(u)), "se_batch_launch_shape")
        return d.value, u.value

    def close(self):
        if self.h:
            lib().se_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class QG2:
    """Two-layer QG model on a mesh (TwoLayerQGEqn steps on the device: first
    order after set_state, second order after that)."""

    def __init__(self, mesh, F1, F2, beta, detj_mode=DETJ_LITERAL):
        self.mesh = mesh
        self.h = None
        h = C.c_void_p()
        check(lib().se_qg2_create(mesh.h, F1, F2, beta, detj_mode, C.byref(h)), "se_qg2_create")
        self.h = h
        mesh.ctx._own(self)

    def set_state(self, psi1, om1, psi2, om2):
        check(lib().se_qg2_set_state(self.h, _d(psi1), _d(om1), _d(psi2), _d(om2)), "se_qg2_set_state")

    def set_dissipation(self, E2=0.0, r=0.0, nu=0.0, tau_e=None):
        """Ekman drag of layer 2, thermal relaxation toward tau_e (None: zero), viscosity."""
        te = None if tau_e is None else _d(tau_e)
        check(lib().se_qg2_set_dissipation(self.h, E2, r, nu, None if te is None else te.ctypes.data),
              "se_qg2_set_dissipation")

    def state(self):
        out = [np.zeros(self.mesh.nV) for _ in range(4)]
        check(lib().se_qg2_get_state(self.h, *out), "se_qg2_get_state")
        return out

    def timestep(self, courant=0.5):
        dt = C.c_double()
        check(lib().se_qg2_timestep(self.h, courant, C.byref(dt)), "se_qg2_timestep")
        return dt.value

    def step(self, dt, **kw):
        o = solve_opts(**kw)
        infos = (SolveInfo * 2)()
        check(lib().se_qg2_step(self.h, dt, C.byref(o), infos), "se_qg2_step")
        return [infos[0].as_dict(), infos[1].as_dict()]

    def close(self):
        if self.h:
            lib().se_qg2_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class QG:
    """One-layer QG channel model on a mesh (QuasiGeostrophicEqn steps on the
    device: first order after set_state, second order after that)."""

    def __init__(self, mesh, F, beta, detj_mode=DETJ_LITERAL):
        self.mesh = mesh
        self.h = None
        h = C.c_void_p()
        check(lib().se_qg_create(mesh.h, F, beta, detj_mode, C.byref(h)), "se_qg_create")
        self.h = h
        mesh.ctx._own(self)

    def set_state(self, psi, omega):
        check(lib().se_qg_set_state(self.h, _d(psi), _d(omega)), "se_qg_set_state")

    def set_dissipation(self, nu=0.0, r=0.0):
        """Viscosity and linear drag on the vorticity."""
        check(lib().se_qg_set_dissipation(self.h, nu, r), "se_qg_set_dissipation")

    def set_topography(self, topo):
        """Topography field (nV), or None for none."""
        tp = None if topo is None else _d(topo)
        check(lib().se_qg_set_topography(self.h, None if tp is None else tp.ctypes.data), "se_qg_set_topography")

    def state(self):
        out = [np.zeros(self.mesh.nV) for _ in range(2)]
        check(lib().se_qg_get_state(self.h, *out), "se_qg_get_state")
        return out

    def timestep(self, courant=0.5):
        dt = C.c_double()
        check(lib().se_qg_timestep(self.h, courant, C.byref(dt)), "se_qg_timestep")
        return dt.value

    def step(self, dt, **kw):
        o = solve_opts(**kw)
        info = SolveInfo()
        check(lib().se_qg_step(self.h, dt, C.byref(o), C.byref(info)), "se_qg_step")
        return info.as_dict()

    def close(self):
        if self.h:
            lib().se_qg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request server for two-layer QG step requests (own context and thread)."""

    def __init__(self, device=0, memory_budget=200e9, max_active=0, precision=PREC_FP32_IR, rtol=1e-10,
                 detj_mode=DETJ_LITERAL, preconditioner=PC_DEFAULT):
        o = ServerOpts()
        lib().se_server_opts_default(C.byref(o))
        o.memory_budget, o.max_active, o.precision, o.rtol, o.detj_mode = memory_budget, max_active, precision, \
            rtol, detj_mode
        o.preconditioner = preconditioner
        h = C.c_void_p()
        check(lib().se_server_create(device, C.byref(o), C.byref(h)), "se_server_create")
        self.h = h
        self._keep = {}

    def submit(self, req, psi1, om1, psi2, om2, dt=0.0):
        """req: dict with nEls, N, lo, hi, periodic, bc, F1, F2, beta, steps and
        optionally E2, r, nu, tau_e (dissipation, as QG2.set_dissipation)."""
        r = QG2Request()
        r.nEls[:] = list(req["nEls"])
        r.N = req["N"]
        r.min[:] = list(req["lo"])
        r.max[:] = list(req["hi"])
        r.periodic[:] = list(req["periodic"])
        r.bc[:] = list(req["bc"])
        r.F1, r.F2, r.beta, r.steps, r.dt = req["F1"], req["F2"], req["beta"], req["steps"], dt
        arrs = [_d(a) for a in (psi1, om1, psi2, om2)]
        r.psi1, r.omega1, r.psi2, r.omega2 = [a.ctypes.data for a in arrs]
        r.E2, r.r, r.nu = req.get("E2", 0.0), req.get("r", 0.0), req.get("nu", 0.0)
        te = None if req.get("tau_e") is None else _d(req["tau_e"])
        r.tau_e = None if te is None else te.ctypes.data
        i = C.c_int64()
        check(lib().se_server_submit(self.h, C.byref(r), C.byref(i)), "se_server_submit")
        self._keep[i.value] = len(arrs[0])
        return i.value

    def wait(self, rid, fields=True):
        nV = self._keep.pop(rid)
        res = RequestResult()
        out = [np.zeros(nV) for _ in range(4)] if fields else [None] * 4
        ptrs = [a.ctypes.data if a is not None else None for a in out]
        check(lib().se_server_wait(self.h, rid, C.byref(res), *ptrs), "se_server_wait")
        return res.as_dict(), (out if fields else None)

    def stats(self):
        st = ServerStats()
        check(lib().se_server_get_stats(self.h, C.byref(st)), "se_server_get_stats")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self.h:
            lib().se_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
