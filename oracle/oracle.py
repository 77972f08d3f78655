"""ctypes bindings for the CPU oracle (liboracle.so) and, when it has been built
in this container, the compiled reference element layer (_ref/libseref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product package (se_amd)
never imports this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = np.ctypeslib.ndpointer
_D = _P(dtype=np.float64, flags="C_CONTIGUOUS")
_I = _P(dtype=np.int32, flags="C_CONTIGUOUS")
_L = _P(dtype=np.int64, flags="C_CONTIGUOUS")
_lib = None
_ref = None


def _d(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def set_threads(n):
    """Threads of the oracle's parallel loops (results do not depend on it)."""
    lib().orc_set_threads(int(n))


def get_threads():
    return int(lib().orc_get_threads())


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        sig = {
            "orc_gll": (i, [i, _D, _D, _D, _D]),
            "orc_leg_eval": (d, [_D, i, d]),
            "orc_mesh_create": (vp, [_I, i, _D, _D, _I, _I, C.POINTER(i)]),
            "orc_mesh_destroy": (None, [vp]),
            "orc_mesh_sizes": (None, [vp, _I]),
            "orc_mesh_verts": (None, [vp, _D]),
            "orc_mesh_maps": (None, [vp, _I, _I]),
            "orc_mesh_gll": (None, [vp, _D, _D, _D]),
            "orc_el_nodes": (None, [vp, i, _I]),
            "orc_get_el": (None, [vp, _D, C.POINTER(i)]),
            "orc_global_to_local": (None, [vp, _D, C.POINTER(i), _D]),
            "orc_shape_funcs_at_coord": (None, [vp, _D, _D]),
            "orc_interp_global": (None, [vp, _D, i, _D, _D]),
            "orc_det_jac": (d, [vp, i, i]),
            "orc_shape_func_derivs": (None, [vp, i, i, _D, C.POINTER(d)]),
            "orc_elem_laplacian": (None, [vp, i, d, i, _D]),
            "orc_elem_mass": (None, [vp, i, d, _D]),
            "orc_elem_field_rhs": (None, [vp, i, d, _D, _D]),
            "orc_helm_create": (vp, [vp, d, d, i]),
            "orc_helm_destroy": (None, [vp]),
            "orc_helm_assemble": (None, [vp, _D]),
            "orc_helm_nnz": (C.c_long, [vp]),
            "orc_helm_csr": (None, [vp, _L, _I, _D]),
            "orc_helm_lift": (None, [vp, _D]),
            "orc_helm_inv_diag": (None, [vp, _D]),
            "orc_helm_matvec": (None, [vp, _D, _D]),
            "orc_set_threads": (None, [i]),
            "orc_get_threads": (i, []),
            "orc_helm_rhs": (None, [vp, _D, _D]),
            "orc_helm_pcg": (i, [vp, _D, _D, d, i, i, C.POINTER(i), C.POINTER(d), vp]),
            "orc_helm_solve": (i, [vp, _D, _D, d, i, i, C.POINTER(i), C.POINTER(d)]),
            "orc_update_field": (None, [vp, _D, _D]),
            "orc_update_vec": (None, [vp, _D, _D]),
            "orc_field_periodic_update": (None, [vp, _D, i]),
            "orc_helm_apply_ideal": (None, [vp, d, d, i, _D, _D]),
            "orc_advect_first": (None, [vp, _D, i, _D, d, _D, vp, vp]),
            "orc_advect_first_interp": (None, [vp, _D, i, _D, d, _D, i]),
            "orc_interp_bicubic": (None, [vp, _D, i, _D, _D]),
            "orc_advect_second": (None, [vp, _D, _D, i, _D, _D, d, _D]),
            "orc_sl_points": (None, [vp, i, _D, _D, d, _D]),
            "orc_genvel": (None, [vp, _D, _D]),
            "orc_calc_timestep": (d, [vp, _D, d]),
            "orc_qg_solveq": (None, [vp, d, d, _D, d, d, d, vp, vp, vp, _D]),
            "orc_elem_svv": (None, [vp, i, d, i, _D]),
            "orc_helm_set_svv": (None, [vp, d, i]),
            "orc_qg_step": (i, [vp, d, d, d, d, vp, d, i, _D, _D, vp, vp, d, i, C.POINTER(i)]),
            "orc_qg2_step": (i, [vp, d, d, d, d, d, d, vp, d, i, _D, _D, _D, _D, vp, vp, vp, vp, d, i,
                                 C.POINTER(i)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libseref.so"))


def ref():
    """The compiled reference element layer (only present where it was built)."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libseref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libseref.so is not built")
        L = C.CDLL(path)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        sig = {
            "ref_mesh_create": (vp, [_I, i, _D, _D, _I, _I]),
            "ref_mesh_destroy": (None, [vp]),
            "ref_mesh_sizes": (None, [vp, _I]),
            "ref_mesh_verts": (None, [vp, _D]),
            "ref_mesh_maps": (None, [vp, _I, _I]),
            "ref_mesh_gll": (None, [vp, _D, _D, _D]),
            "ref_el_nodes": (None, [vp, i, _I]),
            "ref_get_el": (None, [vp, _D, C.POINTER(i)]),
            "ref_global_to_local": (None, [vp, _D, C.POINTER(i), _D]),
            "ref_shape_funcs_at_coord": (None, [vp, _D, _D]),
            "ref_interp_global": (None, [vp, _D, i, _D, _D]),
            "ref_det_jac": (d, [vp, i, i]),
            "ref_shape_func_derivs": (None, [vp, i, i, _D, C.POINTER(d)]),
            "ref_elem_laplacian": (None, [vp, i, d, _D]),
            "ref_elem_mass": (None, [vp, i, d, _D]),
            "ref_elem_field_rhs": (None, [vp, i, d, _D, _D]),
            "ref_field_periodic_update": (None, [vp, _D, i]),
            "ref_advect_first": (None, [vp, _D, i, _D, d, _D, vp, vp]),
            "ref_advect_second": (None, [vp, _D, _D, i, _D, _D, d, _D]),
            "ref_genvel": (None, [vp, _D, _D]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _ref = L
    return _ref


def gll(N):
    a = np.zeros(N + 1)
    w = np.zeros(N + 1)
    D = np.zeros((N + 1) * (N + 1))
    pn = np.zeros(N + 1)
    if lib().orc_gll(N, a, w, D, pn) != 0:
        raise ValueError("orc_gll failed for N=%d" % N)
    return a, w, D.reshape(N + 1, N + 1), pn


class _MeshBase:
    """Common helpers over an oracle or reference mesh handle."""

    def _init_sizes(self):
        s = np.zeros(7, dtype=np.int32)
        self._fn("mesh_sizes")(self.h, s)
        self.nVerts = (int(s[0]), int(s[1]))
        self.nV, self.nElsTotal, self.n, self.size0, self.N = map(int, s[2:7])
        self.nl = (self.N + 1) ** 2

    def verts(self):
        v = np.zeros(2 * self.nV)
        self._fn("mesh_verts")(self.h, v)
        return v.reshape(self.nV, 2)

    def maps(self):
        f2v = np.zeros(self.nV, dtype=np.int32)
        v2f = np.zeros(max(self.n, 1), dtype=np.int32)
        self._fn("mesh_maps")(self.h, f2v, v2f)
        return f2v, v2f[: self.n]

    def gll(self):
        N = self.N
        a, w, D = np.zeros(N + 1), np.zeros(N + 1), np.zeros((N + 1) ** 2)
        self._fn("mesh_gll")(self.h, a, w, D)
        return a, w, D.reshape(N + 1, N + 1)

    def el_nodes(self, el):
        out = np.zeros(self.nl, dtype=np.int32)
        self._fn("el_nodes")(self.h, el, out)
        return out

    def get_el(self, xy):
        e = C.c_int()
        self._fn("get_el")(self.h, _d(xy), C.byref(e))
        return e.value

    def global_to_local(self, xy):
        e = C.c_int()
        l = np.zeros(2)
        self._fn("global_to_local")(self.h, _d(xy), C.byref(e), l)
        return e.value, l

    def shape_funcs_at_coord(self, l):
        out = np.zeros(self.nl)
        self._fn("shape_funcs_at_coord")(self.h, _d(l), out)
        return out

    def interp_global(self, vals, nDofs, xy):
        out = np.zeros(3)
        self._fn("interp_global")(self.h, _d(vals), nDofs, _d(xy), out)
        return out[:nDofs]

    def det_jac(self, el, pt):
        return self._fn("det_jac")(self.h, el, pt)

    def shape_func_derivs(self, el, pt):
        g = np.zeros(2 * self.nl)
        dj = C.c_double()
        self._fn("shape_func_derivs")(self.h, el, pt, g, C.byref(dj))
        return g.reshape(2, self.nl), dj.value

    def elem_mass(self, el, c):
        M = np.zeros(self.nl * self.nl)
        self._fn("elem_mass")(self.h, el, c, M)
        return M.reshape(self.nl, self.nl)

    def elem_field_rhs(self, el, c, field):
        r = np.zeros(self.nl)
        self._fn("elem_field_rhs")(self.h, el, c, _d(field), r)
        return r

    def periodic_update(self, field, nDofs=1):
        f = _d(field).copy()
        self._fn("field_periodic_update")(self.h, f, nDofs)
        return f

    def advect_first(self, q, nDofs, vel, dt, want_dep=True):
        qsl = np.zeros(self.nV * nDofs)
        dep = np.zeros(self.nV, dtype=np.int32)
        xy = np.zeros(2 * self.nV)
        self._fn("advect_first")(self.h, _d(q), nDofs, _d(vel), dt, qsl,
                                 dep.ctypes.data if want_dep else None,
                                 xy.ctypes.data if want_dep else None)
        return qsl, dep, xy.reshape(self.nV, 2)

    def advect_second(self, q, qprev, nDofs, vel, velprev, dt):
        qsl = np.zeros(self.nV * nDofs)
        self._fn("advect_second")(self.h, _d(q), _d(qprev), nDofs, _d(vel), _d(velprev), dt, qsl)
        return qsl

    def genvel(self, psi):
        vel = np.zeros(2 * self.nV)
        self._fn("genvel")(self.h, _d(psi), vel)
        return vel.reshape(self.nV, 2)

    def close(self):
        if self.h:
            self._fn("mesh_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OracleMesh(_MeshBase):
    """orc_mesh handle. bc = (bottom, top, left, right)."""

    def __init__(self, nEls, N, lo, hi, periodic, bc=(0, 0, 0, 0)):
        self.L = lib()
        st = C.c_int()
        self.h = self.L.orc_mesh_create(np.array(nEls, dtype=np.int32), N, _d(lo), _d(hi),
                                        np.array(periodic, dtype=np.int32),
                                        np.array(bc, dtype=np.int32), C.byref(st))
        self.status = st.value
        if not self.h:
            raise ValueError("orc_mesh_create failed, status %d" % st.value)
        self.nEls, self.lo, self.hi, self.periodic, self.bc = tuple(nEls), tuple(lo), tuple(hi), tuple(periodic), tuple(bc)
        self._init_sizes()


    def interp_bicubic(self, vals, nDofs, xy):
        """Advector::Interpolate (bicubic on the 4 x 4 node stencil) at one point."""
        out = np.zeros(3)
        self.L.orc_interp_bicubic(self.h, _d(vals), nDofs, _d(xy), out)
        return out[:nDofs]

    def advect_first_interp(self, q, nDofs, vel, dt, interp):
        """First-order advection with interp 0 (reference default), 1 GLL, 2 bicubic."""
        qsl = np.zeros(self.nV * nDofs)
        self.L.orc_advect_first_interp(self.h, _d(q), nDofs, _d(vel), dt, qsl, int(interp))
        return qsl
    def _fn(self, name):
        return getattr(self.L, "orc_" + name)

    def elem_laplacian(self, el, c, detj_mode=0):
        K = np.zeros(self.nl * self.nl)
        self.L.orc_elem_laplacian(self.h, el, c, detj_mode, K)
        return K.reshape(self.nl, self.nl)

    def calc_timestep(self, vel, courant):
        return self.L.orc_calc_timestep(self.h, _d(vel), courant)

    def sl_points(self, order, vel, dt, velprev=None):
        K = 4 if order == 2 else 2
        pts = np.zeros(self.nV * K * 2)
        self.L.orc_sl_points(self.h, order, _d(vel), _d(velprev if velprev is not None else vel), dt, pts)
        return pts.reshape(self.nV, K, 2)

    def node_distance(self, pts):
        """Smallest distance, in local element coordinates, from any of the given
        points to a GLL line that is not a root hit (< 1e-8 counts as on the line,
        where the reference evaluates a Kronecker delta exactly).  inf if none."""
        a, _, _ = self.gll()
        flat = pts.reshape(-1, 2)
        d = np.full(len(flat), np.inf)
        for k, p in enumerate(flat):
            _, l = self.global_to_local(p)
            for c in l:
                dist = np.min(np.abs(c - a))
                if dist >= 1e-8:
                    d[k] = min(d[k], dist)
        return d.reshape(pts.shape[:-1]).min(axis=-1)

    def update_field(self, x, field):
        f = _d(field).copy()
        self.L.orc_update_field(self.h, _d(x), f)
        return f

    def update_vec(self, field):
        x = np.zeros(self.n)
        self.L.orc_update_vec(self.h, _d(field), x)
        return x

    def apply_ideal(self, c_mass, c_lap, detj_mode, x):
        y = np.zeros(self.n)
        self.L.orc_helm_apply_ideal(self.h, c_mass, c_lap, detj_mode, _d(x), y)
        return y

    def elem_svv(self, el, c, mN):
        A = np.zeros(self.nl * self.nl)
        self.L.orc_elem_svv(self.h, el, c, mN, A)
        return A.reshape(self.nl, self.nl)

    def qg_solveq(self, a, dt, qhat, alpha, rF=0.0, Ei=0.0, nu=0.0, tau=None, omega_hat=None, tau_e=None):
        """SolveQ with the dissipation right-hand side; returns alpha with its
        unknowns replaced (Dirichlet nodes keep the given alpha)."""
        out = _d(alpha).copy()
        keep = [None if v is None else _d(v) for v in (tau, omega_hat, tau_e)]
        self.L.orc_qg_solveq(self.h, a, dt, _d(qhat), rF, Ei, nu,
                             *[None if v is None else v.ctypes.data for v in keep], out)
        return out

    def qg_step(self, F, beta, dt, psi, omega, detj_mode=1, rtol=1e-10, maxit=100000, prev=None, nu=0.0, r=0.0,
                topo=None):
        """One QG channel step; prev = (psi, omega) before the previous step
        selects the second-order step; nu, r, topo: viscosity, linear drag,
        topography."""
        psi, omega = _d(psi).copy(), _d(omega).copy()
        pp = [_d(v) for v in prev] if prev is not None else [None, None]
        tp = None if topo is None else _d(topo)
        it = C.c_int()
        st = self.L.orc_qg_step(self.h, F, beta, nu, r, None if tp is None else tp.ctypes.data, dt, detj_mode,
                                psi, omega, *[None if v is None else v.ctypes.data for v in pp], rtol, maxit,
                                C.byref(it))
        return st, it.value, psi, omega

    def qg2_step(self, F1, F2, beta, dt, psi1, omega1, psi2, omega2, detj_mode=1, rtol=1e-10, maxit=100000,
                 prev=None, E2=0.0, r=0.0, nu=0.0, tau_e=None):
        """One two-layer step; prev = (psi1, omega1, psi2, omega2) before the
        previous step selects the second-order step; E2, r, nu, tau_e: Ekman
        drag of layer 2, thermal relaxation toward tau_e, viscosity."""
        a = [_d(v).copy() for v in (psi1, omega1, psi2, omega2)]
        pp = [_d(v) for v in prev] if prev is not None else [None] * 4
        te = None if tau_e is None else _d(tau_e)
        it = C.c_int()
        st = self.L.orc_qg2_step(self.h, F1, F2, beta, E2, r, nu, None if te is None else te.ctypes.data, dt,
                                 detj_mode, a[0], a[1], a[2], a[3],
                                 *[None if v is None else v.ctypes.data for v in pp], rtol, maxit, C.byref(it))
        return st, it.value, a[0], a[1], a[2], a[3]


class RefMesh(_MeshBase):
    """Reference Mesh + BCs objects (compiled reference, golden-vector source)."""

    def __init__(self, nEls, N, lo, hi, periodic, bc=(0, 0, 0, 0)):
        self.L = ref()
        self.nEls, self.lo, self.hi, self.periodic, self.bc = tuple(nEls), tuple(lo), tuple(hi), tuple(periodic), tuple(bc)
        self.h = self.L.ref_mesh_create(np.array(nEls, dtype=np.int32), N, _d(lo), _d(hi),
                                        np.array(periodic, dtype=np.int32), np.array(bc, dtype=np.int32))
        self._init_sizes()

    def _fn(self, name):
        return getattr(self.L, "ref_" + name)

    def elem_laplacian(self, el, c):
        K = np.zeros(self.nl * self.nl)
        self.L.ref_elem_laplacian(self.h, el, c, K)
        return K.reshape(self.nl, self.nl)


class Helm:
    """orc_helm: PETSc-pattern assembled c_mass*M + c_lap*K~ with lifting and PCG."""

    def __init__(self, mesh, c_mass, c_lap, detj_mode=0, sol_field=None, svv=None):
        """svv = (dt, cutoff): HelmholtzEqn's SVV operator."""
        self.m = mesh
        self.L = mesh.L
        self.h = self.L.orc_helm_create(mesh.h, c_mass, c_lap, detj_mode)
        if svv is not None:
            self.L.orc_helm_set_svv(self.h, float(svv[0]), int(svv[1]))
        if sol_field is None:
            sol_field = np.zeros(mesh.nV)
        self.L.orc_helm_assemble(self.h, _d(sol_field))

    def csr(self):
        nnz = self.L.orc_helm_nnz(self.h)
        rp = np.zeros(self.m.n + 1, dtype=np.int64)
        cols = np.zeros(max(nnz, 1), dtype=np.int32)
        vals = np.zeros(max(nnz, 1))
        self.L.orc_helm_csr(self.h, rp, cols, vals)
        return rp, cols[:nnz], vals[:nnz]

    def scipy(self):
        import scipy.sparse as sp
        rp, cols, vals = self.csr()
        return sp.csr_matrix((vals, cols, rp), shape=(self.m.n, self.m.n))

    def lift(self):
        b = np.zeros(self.m.n)
        self.L.orc_helm_lift(self.h, b)
        return b

    def inv_diag(self):
        d = np.zeros(self.m.n)
        self.L.orc_helm_inv_diag(self.h, d)
        return d

    def matvec(self, x):
        y = np.zeros(self.m.n)
        self.L.orc_helm_matvec(self.h, _d(x), y)
        return y

    def rhs(self, rhs_field):
        f = np.zeros(self.m.n)
        self.L.orc_helm_rhs(self.h, _d(rhs_field), f)
        return f

    def pcg(self, f, rtol=1e-10, maxit=100000, null_space=0, hist=False):
        x = np.zeros(self.m.n)
        it = C.c_int()
        rr = C.c_double()
        hh = np.zeros(maxit + 1) if hist else None
        st = self.L.orc_helm_pcg(self.h, _d(f), x, rtol, maxit, null_space, C.byref(it), C.byref(rr),
                                 hh.ctypes.data if hist else None)
        out = dict(status=st, iters=it.value, relres=rr.value)
        if hist:
            out["hist"] = hh[: it.value + 1]
        return x, out

    def solve(self, rhs_field, sol_field, rtol=1e-10, maxit=100000, null_space=0):
        sol = _d(sol_field).copy()
        it = C.c_int()
        rr = C.c_double()
        st = self.L.orc_helm_solve(self.h, _d(rhs_field), sol, rtol, maxit, null_space, C.byref(it), C.byref(rr))
        return sol, dict(status=st, iters=it.value, relres=rr.value)

    def close(self):
        if self.h:
            self.L.orc_helm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
