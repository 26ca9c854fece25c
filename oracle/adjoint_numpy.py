"""ORACLE — test infrastructure only; never imported by the etalon_amd package.

A float64 NumPy restatement of the discrete adjoint of the 2D TE step, for the
permittivity gradient of a DFT-monitor objective.  No reference counterpart (the
reference has no adjoint); it is pinned by the transpose identity
<A x, y> = <x, A^T y> on random states (tests/test_adjoint_oracle.py), by central
finite differences of the oracle's own forward run, and it in turn pins the HIP
adjoint.

Forward step n (te_numpy.update_h / update_e after
etalon/fdtd/backends/numpy_backend.py:41-61), linear in the state
x = (Ex, Ey, Hz, ψ_Ex,y, ψ_Ey,x, ψ_Hz,x, ψ_Hz,y):

    H:  d = Dy Ex,  e = Dx Ey,  ψy' = by ψy + cy d,  ψx' = bx ψx + cx e,
        Hz' = Hz + db (d iky + ψy' − e ikx − ψx')
    E:  a = Dy Hz', ψ' = by ψ + cy a, u = a iky + ψ',  Ex' = Ex + cb_ex u   (1 ≤ j ≤ ny−1)
        b = Dx Hz', ψ' = bx ψ + cx b, w = b ikx + ψ',  Ey' = Ey − cb_ey w   (1 ≤ i ≤ nx−1)

The objective is J = Σ over objective monitors of Σ |G|², G the monitor's labelled
values (accumulator F through the interpolation weights W of a size-0 axis):
G = W F, F_fk = Σ_n v_k^n p_f^n.  Its adjoint source at node k of step n is
∂J/∂v_k^n = 2 Re Σ_f conj(A_fk) p_f^n with A = Wᵀ G.  The reverse step applies
the transposes in reverse order (E, then H); ∂J/∂cb_ex = Σ_n λ_Ex'^n u^n and
∂J/∂cb_ey = −Σ_n λ_Ey'^n w^n, and with cb = Δt/(ε0 εr), ∂J/∂εr = −(cb/εr) ∂J/∂cb.
"""

from __future__ import annotations

import numpy as np

from oracle import te_numpy as ref

STATE = ("ex", "ey", "hz", "psi_ex_y", "psi_ey_x", "psi_hz_x", "psi_hz_y")


def step_linear(s) -> None:
    """One forward step without sources (the linear map A)."""
    ref.update_h(s)
    ref.update_e(s)


def step_transpose(s, lam: dict, grad: dict | None = None, u=None, w=None) -> None:
    """λ ← Aᵀ λ in place (λ a dict of arrays shaped like the state).  With ``grad``,
    also accumulates ∂/∂cb_ex += λ_Ex' u and ∂/∂cb_ey −= λ_Ey' w (forward u, w of
    the same step)."""
    idx, idy, db = s.inv_dx, s.inv_dy, s.db
    lex, ley, lhz = lam["ex"], lam["ey"], lam["hz"]
    # ---- E transpose (outputs Ex', Ey' → inputs Ex, Ey, Hz', ψ)
    lu = s.cb_ex[:, 1:-1] * lex[:, 1:-1]
    lp = lam["psi_ex_y"][:, 1:-1] + lu
    la = s.iky_e[None, 1:-1] * lu + s.cy_e[None, 1:-1] * lp
    lam["psi_ex_y"][:, 1:-1] = s.by_e[None, 1:-1] * lp
    lw = -s.cb_ey[1:-1, :] * ley[1:-1, :]
    lq = lam["psi_ey_x"][1:-1, :] + lw
    lb = s.ikx_e[1:-1, None] * lw + s.cx_e[1:-1, None] * lq
    lam["psi_ey_x"][1:-1, :] = s.bx_e[1:-1, None] * lq
    if grad is not None:
        grad["cb_ex"][:, 1:-1] += lex[:, 1:-1] * u
        grad["cb_ey"][1:-1, :] -= ley[1:-1, :] * w
    lhz[:, 1:] += idy * la
    lhz[:, :-1] -= idy * la
    lhz[1:, :] += idx * lb
    lhz[:-1, :] -= idx * lb
    # ---- H transpose (output Hz' → inputs Hz, Ex, Ey, ψ_Hz)
    lpy = lam["psi_hz_y"] + db * lhz
    ld = db * s.iky_h[None, :] * lhz + s.cy_h[None, :] * lpy
    lam["psi_hz_y"][...] = s.by_h[None, :] * lpy
    lpx = lam["psi_hz_x"] - db * lhz
    le = -db * s.ikx_h[:, None] * lhz + s.cx_h[:, None] * lpx
    lam["psi_hz_x"][...] = s.bx_h[:, None] * lpx
    lex[:, 1:] += idy * ld
    lex[:, :-1] -= idy * ld
    ley[1:, :] += idx * le
    ley[:-1, :] -= idx * le


def e_increments(s):
    """u (Ex nodes j = 1..ny−1) and w (Ey nodes i = 1..nx−1) of the E update about to
    run on the current state (Hz already updated), float64."""
    a = (s.hz[:, 1:] - s.hz[:, :-1]) * s.inv_dy
    u = a * s.iky_e[None, 1:-1] + (s.by_e[None, 1:-1] * s.psi_ex_y[:, 1:-1] + s.cy_e[None, 1:-1] * a)
    b = (s.hz[1:, :] - s.hz[:-1, :]) * s.inv_dx
    w = b * s.ikx_e[1:-1, None] + (s.bx_e[1:-1, None] * s.psi_ey_x[1:-1, :] + s.cx_e[1:-1, None] * b)
    return u, w


def objective(solver, names) -> float:
    """J = Σ over the named monitors (all their components) of Σ |values|²."""
    vals = solver.monitor_values()
    return float(sum(np.sum(np.abs(v.astype(np.complex128)) ** 2)
                     for (name, _), v in vals.items() if name in names))


def _weights(m):
    """Interpolation matrix W (n_out, n_nodes) of a monitor component's labelled values."""
    nxs, nys = m.x.index.size, m.y.index.size
    wx = np.eye(nxs) if m.x.weights is None else np.asarray(m.x.weights)[None, :]
    wy = np.eye(nys) if m.y.weights is None else np.asarray(m.y.weights)[None, :]
    return np.kron(wx, wy)


def gradient(sim, names, design) -> tuple[float, np.ndarray]:
    """(J, ∂J/∂ε per design pixel) in float64 by the discrete adjoint.

    ``design`` is the PixelGrid structure (the last of ``sim.structures``); its
    pixels map to the Ex/Ey nodes it paints.  The forward increments u, w that
    ∂J/∂cb needs are kept only on the bounding box of those nodes (every step's,
    for the reverse sweep): λ_Ex' and λ_Ey' of step n are λ before the transpose of
    step n, so ∂J/∂cb accumulates there, before ``step_transpose``."""
    if sim.precision != "double":
        raise ValueError("the oracle gradient runs in float64")
    fwd = ref.Solver(sim)
    s = fwd.s
    n_steps = fwd.n_steps
    nodes = pixel_map(fwd.g, design)
    # bounding boxes of the design's Ex and Ey nodes: rows [r0, r1), columns [c0, c1)
    box = {c: (int(ii.min()), int(ii.max()) + 1, int(jj.min()), int(jj.max()) + 1)
           for c, (ii, jj, _, _) in nodes.items()}
    (er0, er1, ec0, ec1), (wr0, wr1, wc0, wc1) = box["Ex"], box["Ey"]
    if ec0 < 1 or ec1 > s.ny or wr0 < 1 or wr1 > s.nx:
        raise ValueError("the design region touches the outer tangential-E edge")
    us, ws = [], []
    for _ in range(n_steps):
        ref.update_h(s)
        fwd._inject(True)
        fwd._accumulate(True)
        u, w = e_increments(s)
        # u covers Ex columns 1..ny-1 (column j at j - 1), w Ey rows 1..nx-1
        us.append(u[er0:er1, ec0 - 1:ec1 - 1].copy())
        ws.append(w[wr0 - 1:wr1 - 1, wc0:wc1].copy())
        ref.update_e(s)
        fwd._inject(False)
        fwd._accumulate(False)
        fwd.n += 1
    J = objective(fwd, names)
    # adjoint source weights A = Wᵀ G per objective monitor component
    srcs = []
    dt = fwd.g.dt
    for m, acc in zip(fwd.mons, fwd.acc):
        if m.name not in names:
            continue
        W = _weights(m)                          # (n_out, n_nodes)
        G = acc @ W.T                            # (n_f, n_out)
        A = G @ W                                # (n_f, n_nodes)
        srcs.append((m, A))
    lam = {k: np.zeros_like(getattr(s, k)) for k in STATE}
    g_ex = np.zeros((er1 - er0, ec1 - ec0))
    g_ey = np.zeros((wr1 - wr0, wc1 - wc0))
    for n in range(n_steps - 1, -1, -1):
        for m, A in srcs:
            t = (n + 0.5) * dt if m.magnetic else (n + 1.0) * dt
            p = ref.phasors(m.freqs, np.array([t]), dt)[0]          # (n_f,)
            val = 2.0 * np.real(np.conj(A).T @ p)                    # (n_nodes,)
            flat = lam[m.comp.lower()].reshape(-1)
            np.add.at(flat, m.flat, val)
        # sources of step n go to λ of its outputs (E' and Hz'), which the transpose
        # reads: E monitors before its E part, H monitors before its H part
        g_ex += lam["ex"][er0:er1, ec0:ec1] * us[n]
        g_ey -= lam["ey"][wr0:wr1, wc0:wc1] * ws[n]
        step_transpose(s, lam)
    eps = fwd.eps
    grads = {"Ex": (g_ex, er0, ec0, s.cb_ex), "Ey": (g_ey, wr0, wc0, s.cb_ey)}
    npx, npy = np.asarray(design.permittivity).shape
    out = np.zeros((npx, npy))
    for comp, (ii, jj, p, q) in nodes.items():
        g, r0, c0, cb = grads[comp]
        g_eps = -(cb[ii, jj] / eps[comp][ii, jj]) * g[ii - r0, jj - c0]
        np.add.at(out, (p, q), g_eps)
    return J, out


def pixel_map(g, design) -> dict:
    """Per E component, the nodes (i, j) a PixelGrid paints and their pixel (p, q):
    the rule of te_numpy.permittivity (inside the footprint within 1e-9 µm, pixel
    by floor of the offset over the pixel size, clipped)."""
    (cx, cy, _), (sx, sy, _) = design.center, design.size
    npx, npy = np.asarray(design.permittivity).shape
    out = {}
    for comp in ("Ex", "Ey"):
        x, y = g.coords(comp)
        X, Y = np.meshgrid(x, y, indexing="ij")
        m = (np.abs(X - cx) <= sx / 2.0 + ref.TOL_UM) & (np.abs(Y - cy) <= sy / 2.0 + ref.TOL_UM)
        ii, jj = np.nonzero(m)
        p = np.array([min(max(int(np.floor((X[a, b] - (cx - sx / 2.0)) / (sx / npx))), 0), npx - 1)
                      for a, b in zip(ii, jj)], dtype=np.int64)
        q = np.array([min(max(int(np.floor((Y[a, b] - (cy - sy / 2.0)) / (sy / npy))), 0), npy - 1)
                      for a, b in zip(ii, jj)], dtype=np.int64)
        out[comp] = (ii, jj, p, q)
    return out
