"""The float64 NumPy adjoint (oracle/adjoint_numpy.py) pinned on the CPU: the
transpose identity <A x, y> = <x, Aᵀ y> of one step on random states and CPML
profiles, and the gradient against central finite differences of the oracle's own
forward objective."""

import numpy as np

from oracle import adjoint_numpy as adj
from oracle import te_numpy as ref
from tests.adjoint_sims import design_problem, with_design


def random_state(seed, nx=23, ny=17):
    rng = np.random.default_rng(seed)
    dx, dy = 0.05, 0.04
    g = ref.Grid(nx, ny, dx, dy, ref.time_step(0.99, dx, dy), 0, 0.0, 0.0, 0.0)
    eps = {c: rng.uniform(1.0, 12.0, g.shape(c)) for c in ("Ex", "Ey")}

    def prof(n):
        return (rng.uniform(0.2, 1.0, n + 1), -rng.uniform(0, 0.8, n + 1), rng.uniform(0.3, 1.0, n + 1),
                rng.uniform(0.2, 1.0, n), -rng.uniform(0, 0.8, n), rng.uniform(0.3, 1.0, n))

    s = ref.State(g, eps, prof(nx), prof(ny), np.float64)
    return s, rng


def randomize(s, rng):
    out = {}
    for k in adj.STATE:
        v = rng.standard_normal(getattr(s, k).shape)
        setattr(s, k, v.copy())
        out[k] = v
    return out


def test_one_step_transpose_identity():
    s, rng = random_state(1)
    x = randomize(s, rng)
    adj.step_linear(s)
    ax = {k: getattr(s, k).copy() for k in adj.STATE}
    y = {k: rng.standard_normal(v.shape) for k, v in x.items()}
    lam = {k: v.copy() for k, v in y.items()}
    adj.step_transpose(s, lam)
    lhs = sum(np.vdot(ax[k], y[k]) for k in adj.STATE)
    rhs = sum(np.vdot(x[k], lam[k]) for k in adj.STATE)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0), (lhs, rhs)


def test_transpose_identity_catches_a_planted_error(monkeypatch):
    """A wrong sign in the transposed CPML coupling breaks the identity."""
    s, rng = random_state(2)
    x = randomize(s, rng)
    adj.step_linear(s)
    ax = {k: getattr(s, k).copy() for k in adj.STATE}
    y = {k: rng.standard_normal(v.shape) for k, v in x.items()}
    lam = {k: v.copy() for k, v in y.items()}
    s.cy_h = -s.cy_h  # the transpose sees the wrong coupling
    adj.step_transpose(s, lam)
    lhs = sum(np.vdot(ax[k], y[k]) for k in adj.STATE)
    rhs = sum(np.vdot(x[k], lam[k]) for k in adj.STATE)
    assert abs(lhs - rhs) > 1e-6 * abs(lhs)


def objective_of(sim):
    o = ref.Solver(sim).run()
    return adj.objective(o, {"out"})


def test_oracle_gradient_matches_central_differences():
    sim, design = design_problem("double", seed=4, steps=200)
    J, grad = adj.gradient(sim, {"out"}, design)
    assert J > 0 and np.all(np.isfinite(grad)) and np.abs(grad).max() > 0
    assert np.isclose(J, objective_of(sim), rtol=1e-14)
    eps = np.asarray(design.permittivity)
    rng = np.random.default_rng(0)
    for flat in rng.choice(eps.size, size=4, replace=False):
        p, q = divmod(int(flat), eps.shape[1])
        h = 1e-4 * eps[p, q]
        plus, minus = eps.copy(), eps.copy()
        plus[p, q] += h
        minus[p, q] -= h
        fd = (objective_of(with_design(sim, design, plus)[0]) -
              objective_of(with_design(sim, design, minus)[0])) / (2 * h)
        print(f"pixel ({p}, {q}): adjoint {grad[p, q]:.6e}  fd {fd:.6e}")
        assert abs(grad[p, q] - fd) <= 1e-5 * np.abs(grad).max(), (p, q, grad[p, q], fd)
