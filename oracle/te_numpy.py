"""ORACLE — test infrastructure only; never imported by the etalon_amd package.

A from-scratch CPU restatement, in NumPy, of the reference's 2D TE FDTD path,
used by tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py as the checker of the HIP kernels.  It follows the reference's host
setup and its NumPy backend operation for operation, so that on the same host
its float32/float64 results are the bits the reference produces:

    grid            etalon/grid/yee.py:82-115, etalon/grid/materials.py:12-35
    CPML profiles   etalon/fdtd/cpml.py:43-92
    sampling        etalon/fdtd/dft.py:39-119
    sources         etalon/fdtd/sources.py:26-83
    state           etalon/fdtd/state.py:52-99
    update / DFT    etalon/fdtd/backends/numpy_backend.py:41-67
    time loop       etalon/fdtd/solver.py:83-128

What pins it (tests/test_oracle.py): golden vectors written from the reference's
own functions (tests/golden/), the reference tests' known answers (pulse spectrum,
CPML closed form at ρ = 1, injection weights, sampling rules, DFT of a tone), the
discrete energy invariant of a PEC cavity, exact mirror/transpose symmetry, and
the M0 physics checks.  Geometry beyond the reference's Box (disks, rings,
pixel grids) follows the same staircase rule and is pinned by brute-force
point-in-shape checks only.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# CODATA 2018, etalon/constants.py:13-29
C0_SI = 299_792_458.0
C0 = C0_SI * 1e6          # µm/s
MU0 = 1.25663706212e-6
EPS0 = 1.0 / (MU0 * C0_SI**2)
ETA0 = MU0 * C0_SI
UM = 1e-6
ON_NODE = 1e-9
TOL_UM = 1e-9
HALF = {"Hz": (True, True), "Ex": (True, False), "Ey": (False, True)}


# ------------------------------------------------------------------ grid
@dataclass(frozen=True)
class Grid:
    nx: int
    ny: int
    dx: float
    dy: float
    dt: float
    layers: int
    x_min: float
    y_min: float
    z: float

    def shape(self, comp):
        hx, hy = HALF[comp]
        return (self.nx if hx else self.nx + 1, self.ny if hy else self.ny + 1)

    def coords(self, comp):
        hx, hy = HALF[comp]
        n_x, n_y = self.shape(comp)
        x = self.x_min + (np.arange(n_x) + (0.5 if hx else 0.0)) * self.dx
        y = self.y_min + (np.arange(n_y) + (0.5 if hy else 0.0)) * self.dy
        return np.asarray(x, np.float64), np.asarray(y, np.float64)

    def n_steps(self, run_time):
        return math.ceil(round(run_time / self.dt, 9))


def cells(size, dl):
    """etalon/grid/yee.py:82-84."""
    return math.ceil(round(size / dl, 9))


def time_step(courant, dx, dy):
    """etalon/grid/yee.py:87-93."""
    return courant / (C0 * math.sqrt(1.0 / dx**2 + 1.0 / dy**2))


def make_grid(sim) -> Grid:
    """etalon/grid/yee.py:96-115."""
    L = sim.boundary.layers
    n_x, n_y = cells(sim.size[0], sim.grid.dl), cells(sim.size[1], sim.grid.dl)
    dx, dy = sim.size[0] / n_x, sim.size[1] / n_y
    return Grid(n_x + 2 * L, n_y + 2 * L, dx, dy, time_step(sim.courant, dx, dy), L,
                sim.center[0] - sim.size[0] / 2.0 - L * dx,
                sim.center[1] - sim.size[1] / 2.0 - L * dy, sim.center[2])


def _inside(s, X, Y, z):
    """Boolean mask of nodes inside structure ``s`` (faces included within 1e-9 µm)."""
    t = s.type
    if t == "Box":
        (cx, cy, cz), (sx, sy, sz) = s.center, s.size
        if abs(z - cz) > sz / 2.0 + TOL_UM:
            return None
        return (np.abs(X - cx) <= sx / 2.0 + TOL_UM) & (np.abs(Y - cy) <= sy / 2.0 + TOL_UM)
    cx, cy, cz = s.center
    if t in ("Cylinder", "Ring") and s.height != 0.0 and abs(z - cz) > s.height / 2.0 + TOL_UM:
        return None
    if t == "Cylinder":
        return (X - cx) ** 2 + (Y - cy) ** 2 <= (s.radius + TOL_UM) ** 2
    if t == "Ring":
        r2 = (X - cx) ** 2 + (Y - cy) ** 2
        m = (r2 <= (s.outer_radius + TOL_UM) ** 2) & (r2 >= max(s.inner_radius - TOL_UM, 0.0) ** 2)
        span = s.angle_stop - s.angle_start
        if span < 360.0:
            rel = np.mod(np.degrees(np.arctan2(Y - cy, X - cx)) - s.angle_start, 360.0)
            m &= rel <= span + 1e-9
        return m
    raise TypeError(t)


def permittivity(sim, g: Grid):
    """ε_r on Ex and Ey nodes, last structure wins (etalon/grid/materials.py:12-35)."""
    out = {}
    for comp in ("Ex", "Ey"):
        x, y = g.coords(comp)
        X, Y = np.meshgrid(x, y, indexing="ij")
        v = np.full(X.shape, sim.medium.permittivity)
        for s in sim.structures:
            if s.type == "PixelGrid":
                (cx, cy, cz), (sx, sy, sz) = s.center, s.size
                if abs(g.z - cz) > sz / 2.0 + TOL_UM:
                    continue
                eps = np.asarray(s.permittivity, np.float64)
                npx, npy = eps.shape
                m = (np.abs(X - cx) <= sx / 2.0 + TOL_UM) & (np.abs(Y - cy) <= sy / 2.0 + TOL_UM)
                ii, jj = np.nonzero(m)
                for a, b in zip(ii, jj):
                    p = min(max(int(math.floor((X[a, b] - (cx - sx / 2.0)) / (sx / npx))), 0), npx - 1)
                    q = min(max(int(math.floor((Y[a, b] - (cy - sy / 2.0)) / (sy / npy))), 0), npy - 1)
                    v[a, b] = eps[p, q]
                continue
            if s.type in ("Cylinder", "Ring"):
                # only nodes within the shape's bounding square can be inside it: the
                # same test on that window (a speed-up for lattices of many holes)
                r = (s.radius if s.type == "Cylinder" else s.outer_radius) + 1e-6
                cx, cy, _ = s.center
                a0, a1 = np.searchsorted(x, cx - r), np.searchsorted(x, cx + r, side="right")
                b0, b1 = np.searchsorted(y, cy - r), np.searchsorted(y, cy + r, side="right")
                m = _inside(s, X[a0:a1, b0:b1], Y[a0:a1, b0:b1], g.z)
                if m is not None:
                    v[a0:a1, b0:b1][m] = s.medium.permittivity
                continue
            m = _inside(s, X, Y, g.z)
            if m is not None:
                v[m] = s.medium.permittivity
        out[comp] = v
    return out


# ------------------------------------------------------------------ CPML
def pml_rho(n, L):
    """etalon/fdtd/cpml.py:43-51."""
    re_, rh = np.zeros(n + 1), np.zeros(n)
    i = np.arange(L + 1)
    re_[: L + 1] = (L - i) / L
    i = np.arange(L)
    rh[:L] = (L - i - 0.5) / L
    return np.maximum(re_, re_[::-1]), np.maximum(rh, rh[::-1])


def _bc(rho, smax, kmax, amax, m, dt):
    """etalon/fdtd/cpml.py:54-71."""
    sigma = smax * rho**m
    kappa = 1.0 + (kmax - 1.0) * rho**m
    alpha = amax * (1.0 - rho)
    b = np.exp(-(sigma / kappa + alpha) * dt / EPS0)
    c = np.zeros_like(rho)
    inl = rho > 0.0
    s, k, a = sigma[inl], kappa[inl], alpha[inl]
    c[inl] = s * (b[inl] - 1.0) / (s * k + k**2 * a)
    b[~inl] = 1.0
    return b, c, 1.0 / kappa


def cpml(n, spacing, dt, pml, eps_r, alpha_max=0.0):
    """etalon/fdtd/cpml.py:74-92 → (b_e, c_e, ik_e, b_h, c_h, ik_h)."""
    m = pml.grading_order
    smax = pml.sigma_scale * 0.8 * (m + 1.0) / (ETA0 * spacing * UM * math.sqrt(eps_r))
    re_, rh = pml_rho(n, pml.layers)
    return (*_bc(re_, smax, pml.kappa_max, alpha_max, m, dt),
            *_bc(rh, smax, pml.kappa_max, alpha_max, m, dt))


# ------------------------------------------------------------------ sampling / sources
@dataclass(frozen=True)
class Axis:
    index: np.ndarray
    coords: np.ndarray
    weights: object
    point: object


def axis(nodes, center, size, what="m"):
    """etalon/fdtd/dft.py:39-66."""
    h = nodes[1] - nodes[0]
    if size > 0.0:
        tol = ON_NODE * h
        lo, hi = center - size / 2.0, center + size / 2.0
        idx = np.nonzero((nodes >= lo - tol) & (nodes <= hi + tol))[0].astype(np.int64)
        if idx.size == 0:
            raise ValueError(f"{what} contains no grid node along this axis (thinner than one cell)")
        return Axis(idx, nodes[idx], None, None)
    pos = (center - nodes[0]) / h
    i0 = math.floor(pos)
    fr = pos - i0
    if fr < ON_NODE:
        idx, w = np.array([i0]), np.array([1.0])
    elif fr > 1.0 - ON_NODE:
        idx, w = np.array([i0 + 1]), np.array([1.0])
    else:
        idx, w = np.array([i0, i0 + 1]), np.array([1.0 - fr, fr])
    if idx[0] < 0 or idx[-1] >= nodes.size:
        raise ValueError(f"{what} at {center} µm is outside the grid")
    idx = idx.astype(np.int64)
    return Axis(idx, nodes[idx], w, center)


@dataclass(frozen=True)
class Monitor:
    name: str
    comp: str
    wavelengths: np.ndarray
    x: Axis
    y: Axis
    flat: np.ndarray

    @property
    def freqs(self):
        return C0 / self.wavelengths

    @property
    def magnetic(self):
        return self.comp.startswith("H")


def monitors_of(mon, g: Grid):
    """etalon/fdtd/dft.py:92-107."""
    out = []
    wl = np.asarray(mon.wavelengths, np.float64)
    for comp in mon.components or ("Ex", "Ey", "Hz"):
        xn, yn = g.coords(comp)
        sx = axis(xn, mon.center[0], mon.size[0])
        sy = axis(yn, mon.center[1], mon.size[1])
        ny_c = g.shape(comp)[1]
        flat = (sx.index[:, None] * ny_c + sy.index[None, :]).ravel().astype(np.int64)
        out.append(Monitor(mon.name, comp, wl, sx, sy, flat))
    return out


def phasors(freqs, times, dt):
    """etalon/fdtd/dft.py:110-119."""
    ph = 2.0 * np.pi * times[:, None] * freqs[None, :]
    return np.asarray(np.exp(1j * ph) * dt, dtype=np.complex128)


def pulse(p, t):
    """etalon/fdtd/sources.py:26-30, with the GaussianPulse properties of
    etalon/core/source.py:48-68 (f0, σ_f, τ = 1/(2πσ_f), t0 = 5τ)."""
    f0 = C0 / p.wavelength
    f_hi = C0 / (p.wavelength - p.bandwidth / 2.0)
    f_lo = C0 / (p.wavelength + p.bandwidth / 2.0)
    tau = 1.0 / (2.0 * math.pi * ((f_hi - f_lo) / 2.0))
    s = np.asarray(t, dtype=np.float64) - 5.0 * tau
    env = np.exp(-(s**2) / (2.0 * tau**2))
    return p.amplitude * env * np.sin(2.0 * np.pi * f0 * s)


@dataclass(frozen=True)
class Injection:
    comp: str
    i: np.ndarray
    j: np.ndarray
    weights: np.ndarray
    coef: np.ndarray
    pulse: object

    @property
    def magnetic(self):
        return self.comp.startswith("H")

    def t(self, n, dt):
        return n * dt if self.magnetic else (n + 0.5) * dt

    def delta(self, n, dt):
        """etalon/fdtd/sources.py:50-56."""
        return self.coef * pulse(self.pulse, self.t(n, dt))


def injection(src, g: Grid, eps):
    """etalon/fdtd/sources.py:59-83."""
    xn, yn = g.coords(src.component)
    sx = axis(xn, src.center[0], 0.0)
    sy = axis(yn, src.center[1], 0.0)
    i = np.repeat(sx.index, sy.index.size)
    j = np.tile(sy.index, sx.index.size)
    w = np.outer(sx.weights, sy.weights).ravel()
    area = (g.dx * UM) * (g.dy * UM)
    if src.component.startswith("H"):
        upd = np.full(w.shape, g.dt / MU0)
    else:
        upd = g.dt / (EPS0 * eps[src.component][i, j])
    return Injection(src.component, i, j, w, upd * w / area, src.pulse)


# ------------------------------------------------------------------ state and updates
class State:
    """TEState (etalon/fdtd/state.py:15-99): zero fields, coefficients cast once."""

    def __init__(self, g: Grid, eps, cx, cy, dtype):
        put = lambda a: np.ascontiguousarray(a, dtype=dtype)  # noqa: E731
        z = lambda comp: put(np.zeros(g.shape(comp)))  # noqa: E731
        self.ex, self.ey, self.hz = z("Ex"), z("Ey"), z("Hz")
        self.psi_ex_y, self.psi_ey_x = z("Ex"), z("Ey")
        self.psi_hz_x, self.psi_hz_y = z("Hz"), z("Hz")
        self.cb_ex = put(g.dt / (EPS0 * eps["Ex"]))
        self.cb_ey = put(g.dt / (EPS0 * eps["Ey"]))
        (self.bx_e, self.cx_e, self.ikx_e, self.bx_h, self.cx_h, self.ikx_h) = map(put, cx)
        (self.by_e, self.cy_e, self.iky_e, self.by_h, self.cy_h, self.iky_h) = map(put, cy)
        self.db = dtype(g.dt / MU0)
        self.inv_dx = dtype(1.0 / (g.dx * UM))
        self.inv_dy = dtype(1.0 / (g.dy * UM))
        self.nx, self.ny = g.nx, g.ny


def update_h(s) -> None:
    """etalon/fdtd/backends/numpy_backend.py:41-49, same operands, same order."""
    d = (s.ex[:, 1:] - s.ex[:, :-1]) * s.inv_dy
    e = (s.ey[1:, :] - s.ey[:-1, :]) * s.inv_dx
    s.psi_hz_y[...] = s.by_h[None, :] * s.psi_hz_y + s.cy_h[None, :] * d
    s.psi_hz_x[...] = s.bx_h[:, None] * s.psi_hz_x + s.cx_h[:, None] * e
    s.hz += s.db * ((d * s.iky_h[None, :] + s.psi_hz_y) - (e * s.ikx_h[:, None] + s.psi_hz_x))


def update_e(s) -> None:
    """etalon/fdtd/backends/numpy_backend.py:51-61."""
    d = (s.hz[:, 1:] - s.hz[:, :-1]) * s.inv_dy
    p = s.psi_ex_y[:, 1:-1]
    p[...] = s.by_e[None, 1:-1] * p + s.cy_e[None, 1:-1] * d
    s.ex[:, 1:-1] += s.cb_ex[:, 1:-1] * (d * s.iky_e[None, 1:-1] + p)
    d = (s.hz[1:, :] - s.hz[:-1, :]) * s.inv_dx
    p = s.psi_ey_x[1:-1, :]
    p[...] = s.bx_e[1:-1, None] * p + s.cx_e[1:-1, None] * d
    s.ey[1:-1, :] -= s.cb_ey[1:-1, :] * (d * s.ikx_e[1:-1, None] + p)


def dft(acc, field, index, ph) -> None:
    """etalon/fdtd/backends/numpy_backend.py:63-67."""
    acc += field.ravel()[index][None, :] * ph[:, None]


# ------------------------------------------------------------------ the time loop
class Solver:
    """FDTDSolver restated (etalon/fdtd/solver.py:49-128); phasor rows one table per run."""

    def __init__(self, sim):
        self.sim = sim
        self.g = g = make_grid(sim)
        self.n_steps = g.n_steps(sim.run_time)
        self.real = np.float32 if sim.precision == "single" else np.float64
        self.cplx = np.complex64 if sim.precision == "single" else np.complex128
        self.eps = permittivity(sim, g)
        bg = sim.medium.permittivity
        self.cx = cpml(g.nx, g.dx, g.dt, sim.boundary, bg)
        self.cy = cpml(g.ny, g.dy, g.dt, sim.boundary, bg)
        self.s = State(g, self.eps, self.cx, self.cy, self.real)
        self.inj = [injection(src, g, self.eps) for src in sim.sources]
        self.mons = [m for mon in sim.monitors for m in monitors_of(mon, g)]
        self.acc = [np.zeros((m.wavelengths.size, m.flat.size), self.cplx) for m in self.mons]
        self.n = 0

    def _inject(self, magnetic):
        for inj in self.inj:
            if inj.magnetic != magnetic:
                continue
            field = getattr(self.s, inj.comp.lower())
            field[inj.i, inj.j] -= inj.delta(self.n, self.g.dt).astype(self.real)

    def _accumulate(self, magnetic):
        dt = self.g.dt
        t = np.array([self.n + (0.5 if magnetic else 1.0)], np.float64) * dt
        for m, acc in zip(self.mons, self.acc):
            if m.magnetic != magnetic:
                continue
            ph = phasors(m.freqs, t, dt).astype(self.cplx)[0]
            dft(acc, getattr(self.s, m.comp.lower()), m.flat, ph)

    def step(self):
        """etalon/fdtd/solver.py:120-128."""
        update_h(self.s)
        self._inject(True)
        self._accumulate(True)
        update_e(self.s)
        self._inject(False)
        self._accumulate(False)
        self.n += 1

    def run(self, steps=None):
        for _ in range(self.n_steps - self.n if steps is None else steps):
            self.step()
        return self

    def fields(self):
        return {"Ex": self.s.ex.copy(), "Ey": self.s.ey.copy(), "Hz": self.s.hz.copy()}

    def monitor_values(self):
        """{(monitor, component): final labelled values} as in etalon/io/result.py:54-91."""
        out = {}
        for m, acc in zip(self.mons, self.acc):
            n_f = m.wavelengths.size
            d = acc.reshape(n_f, m.x.index.size, m.y.index.size).astype(np.complex128)
            if m.x.weights is not None:
                d = np.einsum("fab,a->fb", d, m.x.weights)
            if m.y.weights is not None:
                d = d @ m.y.weights
            out[(m.name, m.comp)] = np.moveaxis(d, 0, -1).astype(acc.dtype)
        return out
