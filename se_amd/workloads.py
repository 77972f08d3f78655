"""Seeded synthetic workloads for the SE hot path.

All randomness comes from this module's SplitMix64 stream with Box-Muller
normals, so the same seed gives the same bytes on every machine and for every
consumer (HIP path, CPU oracle, bench).  Field values are evaluated at the
reference's vertex coordinates (nV nodes, periodic duplicates included).

Workloads (sizes from the benchmark configuration list):
  poisson_8x8      8x8 elements, p=4, doubly periodic [0,2pi]^2, 16 Fourier modes
  qg_channel       32x32, p=6, [0,1]^2 x-periodic channel, 20 Gaussian vortices
  helmholtz_128    128x128, p=8, doubly periodic, lambda=100, k^-3 spectrum
  ensemble_512     512 Helmholtz instances, uneven size/order/lambda
  qg2_stream       two-layer QG step requests arriving as a Poisson process
"""
import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
TWO_PI = 2.0 * math.pi


class Rng:
    """SplitMix64 stream (Steele, Lea & Flood 2014) with Box-Muller normals."""

    def __init__(self, seed):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def u64(self, n):
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * _GOLDEN
            self.state = self.state + np.uint64(n) * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            return z ^ (z >> np.uint64(31))

    def uniform(self, n):
        """n doubles in [0, 1) from the top 53 bits."""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n):
        """n standard normals; each pair of uniforms gives one Box-Muller pair."""
        m = (n + 1) // 2
        u = self.uniform(2 * m).reshape(m, 2)
        r = np.sqrt(-2.0 * np.log(1.0 - u[:, 0]))
        t = TWO_PI * u[:, 1]
        return np.stack([r * np.cos(t), r * np.sin(t)], axis=1).ravel()[:n]

    def integers(self, n, lo, hi):
        """n integers uniform in [lo, hi)."""
        return lo + np.floor(self.uniform(n) * (hi - lo)).astype(np.int64)


def mesh_coords(nEls, N, lo, hi, abcissa):
    """Vertex coordinates of the reference mesh (Mesh.cpp:63-72), per line."""
    out = []
    for d in range(2):
        dx = (hi[d] - lo[d]) / nEls[d]
        I = np.arange(nEls[d] * N + 1)
        out.append(lo[d] + (I // N) * dx + ((abcissa[I % N] + 1.0) / 2.0) * dx)
    return out


def field_from_lines(X, Y, fn):
    """fn(Xgrid, Ygrid) evaluated on the node grid, flattened node-major."""
    XX, YY = np.meshgrid(X, Y)
    return np.ascontiguousarray(fn(XX, YY).ravel())


# ---------------------------------------------------------------------------
def poisson_8x8(seed=1):
    """Doubly periodic Poisson problem: 16 Fourier modes with |k| <= 4."""
    ks = [(kx, ky) for kx in range(-4, 5) for ky in range(-4, 5) if 0 < kx * kx + ky * ky <= 16]
    rng = Rng(seed)
    ks = list(ks)
    for i in range(16):  # partial Fisher-Yates draw without replacement
        j = i + int(rng.uniform(1)[0] * (len(ks) - i))
        ks[i], ks[j] = ks[j], ks[i]
    modes = ks[:16]
    amp = rng.normal(16)
    phase = TWO_PI * rng.uniform(16)
    cfg = dict(workload="poisson_8x8", nEls=(8, 8), N=4, lo=(0.0, 0.0), hi=(TWO_PI, TWO_PI),
               periodic=(1, 1), bc=(0, 0, 0, 0), c_mass=0.0, c_lap=1.0, null_space=True)

    def rhs(X, Y):
        f = np.zeros_like(X)
        for (kx, ky), a, p in zip(modes, amp, phase):
            f += a * np.cos(kx * X + ky * Y + p)
        return f

    cfg["modes"], cfg["amp"], cfg["phase"] = modes, amp, phase
    cfg["rhs_fn"] = rhs
    return cfg


def spectrum_field(X, Y, L, kmax, rng, slope=-1.5):
    """Real random field with isotropic amplitude |k|^slope on 1 <= |k| <= kmax
    (slope -3/2 gives an energy spectrum E(k) ~ k^-3), evaluated separably."""
    kmax = int(kmax)
    k = np.arange(-kmax, kmax + 1)
    KX, KY = np.meshgrid(k, k, indexing="ij")
    kk = np.sqrt(KX * KX + KY * KY)
    mask = (kk >= 1.0) & (kk <= kmax)
    nm = int(mask.sum())
    z = rng.normal(2 * nm).reshape(nm, 2)
    F = np.zeros(KX.shape, dtype=np.complex128)
    F[mask] = (z[:, 0] + 1j * z[:, 1]) * kk[mask] ** slope
    s = TWO_PI / L
    Ex = np.exp(1j * s[0] * np.outer(X, k))      # (nX, nk)
    Ey = np.exp(1j * s[1] * np.outer(Y, k))      # (nY, nk)
    G = Ex @ F @ Ey.T                              # (nX, nY)
    f = G.real.T                                   # row J, column I
    return np.ascontiguousarray(f / max(np.sqrt(np.mean(f * f)), 1e-300))


def helmholtz_128(nEls=128, N=8, lam=100.0, seed=3, detj_mode=1):
    """Single large periodic Helmholtz solve: (lambda M + K) psi = M f."""
    return dict(workload="helmholtz_128" if nEls == 128 and N == 8 else "helmholtz_%d_p%d" % (nEls, N),
                nEls=(nEls, nEls), N=N, lo=(0.0, 0.0), hi=(TWO_PI, TWO_PI), periodic=(1, 1),
                bc=(0, 0, 0, 0), c_mass=lam, c_lap=1.0, detj_mode=detj_mode, seed=seed,
                kmax=max(1, nEls * N // 4), null_space=False)


def make_rhs(cfg, abcissa):
    X, Y = mesh_coords(cfg["nEls"], cfg["N"], cfg["lo"], cfg["hi"], abcissa)
    if "rhs_fn" in cfg:
        return field_from_lines(X, Y, cfg["rhs_fn"])
    L = (cfg["hi"][0] - cfg["lo"][0], cfg["hi"][1] - cfg["lo"][1])
    return spectrum_field(X, Y, np.array(L), cfg["kmax"], Rng(cfg["seed"])).ravel()


def ensemble_512(count=512, seed0=1000, detj_mode=1):
    """Uneven ensemble: nEls ~ U{8,16,32,64}, p ~ U{4,6,8}, lambda ~ 10^U[0,3]."""
    out = []
    for i in range(count):
        rng = Rng(seed0 + i)
        ne = [8, 16, 32, 64][int(rng.integers(1, 0, 4)[0])]
        N = [4, 6, 8][int(rng.integers(1, 0, 3)[0])]
        lam = 10.0 ** (3.0 * rng.uniform(1)[0])
        c = helmholtz_128(nEls=ne, N=N, lam=lam, seed=seed0 + i, detj_mode=detj_mode)
        c["workload"] = "ensemble_member"
        out.append(c)
    return out


def gaussian_vortices(X, Y, count, rng, lo, hi, xper=True):
    """Sum of Gaussians: centres U(domain), radii U[0.02,0.08]L, amplitudes N(0,1)."""
    L = hi[0] - lo[0]
    cx = lo[0] + (hi[0] - lo[0]) * rng.uniform(count)
    cy = lo[1] + (hi[1] - lo[1]) * rng.uniform(count)
    r = L * (0.02 + 0.06 * rng.uniform(count))
    a = rng.normal(count)
    XX, YY = np.meshgrid(X, Y)
    q = np.zeros_like(XX)
    for i in range(count):
        dxv = XX - cx[i]
        if xper:
            dxv = dxv - L * np.round(dxv / L)
        q += a[i] * np.exp(-(dxv * dxv + (YY - cy[i]) ** 2) / (r[i] * r[i]))
    return np.ascontiguousarray(q.ravel())


def qg_channel(seed=2, nEls=32, N=6):
    """One-layer QG channel: beta = 1, F = 1/Ld^2 with Ld = 0.1, psi = 0 on the walls."""
    return dict(workload="qg_channel", nEls=(nEls, nEls), N=N, lo=(0.0, 0.0), hi=(1.0, 1.0), periodic=(1, 0),
                bc=(1, 1, 0, 0), beta=1.0, F=100.0, courant=0.5, seed=seed, vortices=20)


def qg2_stream(n_requests=200, rate=2000.0, seed=5, N=6):
    """Two-layer QG step requests: Poisson arrivals, mesh m^2 with m log-uniform
    in [16, 96], 1..50 steps, F1 = F2 = 200 (lambda = F1 + F2 = 1/Ld^2, Ld = 0.05)."""
    rng = Rng(seed)
    gaps = -np.log(1.0 - rng.uniform(n_requests)) / rate
    t = np.cumsum(gaps)
    m = np.rint(16.0 * 6.0 ** rng.uniform(n_requests)).astype(int)
    steps = rng.integers(n_requests, 1, 51)
    reqs = []
    for i in range(n_requests):
        reqs.append(dict(id=i, t_arrive=float(t[i]), nEls=(int(m[i]), int(m[i])), N=N, steps=int(steps[i]),
                         seed=seed * 100003 + i, F1=200.0, F2=200.0, beta=1.0, lo=(0.0, 0.0), hi=(1.0, 1.0),
                         periodic=(1, 0), bc=(1, 1, 0, 0)))
    return reqs


def qg2_initial(req, X, Y):
    """Initial two-layer state of a stream request on the vertex lines X, Y:
    psi_1, psi_2 seeded Gaussian vortices times sin(pi y / L_y) (zero on the
    walls), omega_1 = omega_2 = 0."""
    rng = Rng(req["seed"])
    lo, hi = req["lo"], req["hi"]
    wall = np.sin(np.pi * (np.asarray(Y) - lo[1]) / (hi[1] - lo[1]))
    win = np.repeat(wall, len(X))
    p1 = gaussian_vortices(X, Y, 10, rng, lo, hi) * win
    p2 = 0.5 * gaussian_vortices(X, Y, 10, rng, lo, hi) * win
    z = np.zeros_like(p1)
    return p1, z, p2, z.copy()
