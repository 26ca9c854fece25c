"""HIP adjoint: the reverse step is the exact transpose of the forward step (f64
dot-product identity on random states and CPML profiles, and equality with the
NumPy transpose), the gradient equals the float64 NumPy adjoint gradient and
central finite differences, checkpoint/recompute and stored-history modes agree
bit for bit, and the torch.autograd.Function returns the same gradient."""

import ctypes as C

import numpy as np
import pytest

from etalon_amd.adjoint import AdjointProblem, objective
from etalon_amd.hip import _lib
from etalon_amd.hip.backend import zeros
from oracle import adjoint_numpy as adj
from tests.adjoint_sims import design_problem, with_design
from tests.test_adjoint_oracle import randomize, random_state
from tests.test_gpu_parity import to_hip_state

pytestmark = pytest.mark.gpu
STATE = adj.STATE


def hip_transpose(s, lam_host):
    """One HIP reverse step (no sources, no design box) on λ; returns λ' on the host."""
    hs = to_hip_state(s)
    lam = {k: zeros(v.shape, np.float64) for k, v in lam_host.items()}
    for k, v in lam_host.items():
        lam[k].set(v)
    d = _lib.EtlAdjointDesc()
    d.precision = 1
    d.nx, d.ny = s.nx, s.ny
    d.db, d.inv_dx, d.inv_dy = float(s.db), float(s.inv_dx), float(s.inv_dy)
    for k, name in enumerate(STATE):
        d.lam[k] = lam[name].ptr
    d.cb_ex, d.cb_ey = hs.cb_ex.ptr, hs.cb_ey.ptr
    for name in ("bx_e", "cx_e", "ikx_e", "bx_h", "cx_h", "ikx_h",
                 "by_e", "cy_e", "iky_e", "by_h", "cy_h", "iky_h"):
        setattr(d, name, getattr(hs, name).ptr)
    handle = C.c_void_p()
    _lib.call("etl_adjoint_create", C.byref(d), C.byref(handle))
    try:
        _lib.call("etl_adjoint_run", handle, 0, 1)
    finally:
        _lib.call("etl_adjoint_destroy", handle)
    return {k: v.get() for k, v in lam.items()}


def test_reverse_step_is_the_transpose(hip):
    s, rng = random_state(7)
    x = randomize(s, rng)
    adj.step_linear(s)  # the NumPy forward step equals the HIP one bit for bit
    ax = {k: getattr(s, k).copy() for k in STATE}
    y = {k: rng.standard_normal(v.shape) for k, v in x.items()}
    aty = hip_transpose(s, y)
    lhs = sum(np.vdot(ax[k], y[k]) for k in STATE)
    rhs = sum(np.vdot(x[k], aty[k]) for k in STATE)
    print(f"<Ax,y> = {lhs:.15e}  <x,A^T y> = {rhs:.15e}")
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    ref = {k: v.copy() for k, v in y.items()}
    adj.step_transpose(s, ref)
    for k in STATE:
        np.testing.assert_allclose(aty[k], ref[k], rtol=1e-13, atol=1e-13 * np.abs(ref[k]).max())


@pytest.fixture(scope="module")
def oracle_gradient():
    sim, design = design_problem("double", seed=4, steps=200)
    return adj.gradient(sim, {"out"}, design)


@pytest.mark.parametrize("kernel", ["tile", "plain"])
@pytest.mark.parametrize("mode", ["checkpoint", "store"])
def test_double_gradient_equals_the_numpy_adjoint(hip, oracle_gradient, mode, kernel):
    sim, design = design_problem("double", seed=4, steps=200)
    J_ref, g_ref = oracle_gradient
    prob = AdjointProblem(sim, ["out"], checkpoint_every=64, mode=mode, kernel=kernel)
    J, g = prob.run(np.asarray(design.permittivity))
    assert J == pytest.approx(J_ref, rel=1e-13)
    err = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    print(f"{mode}/{kernel}: max |g - g_numpy| / max |g_numpy| = {err:.3e}")
    assert err < 1e-10


def test_single_gradient_against_float64(hip, oracle_gradient):
    sim, design = design_problem("single", seed=4, steps=200)
    J_ref, g_ref = oracle_gradient
    J, g = AdjointProblem(sim, ["out"]).run(np.asarray(design.permittivity))
    err = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    print(f"f32 J rel err {abs(J - J_ref) / J_ref:.3e}, gradient max rel err {err:.3e}")
    assert abs(J - J_ref) / J_ref < 1e-4 and err < 1e-3


@pytest.mark.parametrize("depth", [1, 2, 4, 6])
def test_tile_and_plain_adjoint_agree(hip, depth):
    """The temporally blocked reverse steps equal the per-half-step kernels (signed
    zeros aside), with sources, design box and CPML, in float32 and any depth."""
    from etalon_amd.hip.plan import Tune

    sim, design = design_problem("single", seed=21, steps=250, pixels=(7, 5))
    eps = np.asarray(design.permittivity)
    _, g_plain = AdjointProblem(sim, ["out"], kernel="plain").run(eps)
    _, g_tile = AdjointProblem(sim, ["out"], kernel="tile", tune=Tune(temporal=depth)).run(eps)
    err = np.abs(g_tile - g_plain).max() / np.abs(g_plain).max()
    print(f"T={depth}: max |tile - plain| / max |plain| = {err:.3e}")
    assert err < 1e-6


@pytest.mark.parametrize("depth", [1, 2, 4])
def test_adjoint_launch_modes_agree_bitwise(hip, depth):
    """One kernel per step block (every tile kind) against one kernel per tile kind on
    three streams, for the forward plans and the reverse steps."""
    from etalon_amd.hip.plan import Tune

    sim, design = design_problem("single", seed=23, steps=250, pixels=(7, 5))
    eps = np.asarray(design.permittivity)
    J1, g1 = AdjointProblem(sim, ["out"], tune=Tune(temporal=depth, launch=1)).run(eps)
    J3, g3 = AdjointProblem(sim, ["out"], tune=Tune(temporal=depth, launch=3)).run(eps)
    assert J1 == J3
    assert np.array_equal(g1, g3)


def test_checkpoint_and_store_modes_agree_bitwise(hip):
    sim, design = design_problem("single", seed=9, steps=300)
    eps = np.asarray(design.permittivity)
    _, g1 = AdjointProblem(sim, ["out"], checkpoint_every=64, mode="checkpoint").run(eps)
    _, g2 = AdjointProblem(sim, ["out"], mode="store").run(eps)
    _, g3 = AdjointProblem(sim, ["out"], checkpoint_every=64, mode="checkpoint",
                           schedule="sequential").run(eps)
    assert np.array_equal(g1, g2)
    assert np.array_equal(g1, g3)


@pytest.mark.parametrize("depth", [1, 4, 6])
def test_segment_schedules_agree_bitwise(hip, depth):
    """The recompute of segment k - 1 beside the reverse steps of segment k — in one
    launch per step block (joint) or on two streams (lanes) — against one after the
    other, with 5 segments, the last one short (odd depth-1 tails)."""
    from etalon_amd.hip.plan import Tune

    sim, design = design_problem("single", seed=4, steps=301)
    eps = np.asarray(design.permittivity)
    tune = Tune(temporal=depth)
    out = {s: AdjointProblem(sim, ["out"], checkpoint_every=64, schedule=s, tune=tune).run(eps)
           for s in ("joint", "lanes", "sequential")}
    J, g = out["sequential"]
    for s in ("joint", "lanes"):
        assert out[s][0] == J, s
        assert np.array_equal(out[s][1], g), s


def test_gradient_against_central_differences(hip):
    """16 random pixels, float32 gradient against float64 central differences."""
    sim32, design = design_problem("single", seed=12, pixels=(8, 8), steps=260)
    sim64, _ = design_problem("double", seed=12, pixels=(8, 8), steps=260)
    eps = np.asarray(design.permittivity)
    _, g = AdjointProblem(sim32, ["out"]).run(eps)
    prob64 = AdjointProblem(sim64, ["out"])
    rng = np.random.default_rng(1)
    scale = np.abs(g).max()
    for flat in rng.choice(eps.size, size=16, replace=False):
        p, q = divmod(int(flat), eps.shape[1])
        h = 1e-4 * eps[p, q]
        plus, minus = eps.copy(), eps.copy()
        plus[p, q] += h
        minus[p, q] -= h
        fd = (prob64.forward(plus).J - prob64.forward(minus).J) / (2 * h)
        rel = abs(g[p, q] - fd) / scale
        print(f"pixel ({p}, {q}): adjoint {g[p, q]:.6e} fd {fd:.6e} rel {rel:.2e}")
        assert rel < 1e-3


def test_torch_autograd_function(hip):
    import torch

    sim, design = design_problem("single", seed=4, steps=120)
    prob = AdjointProblem(sim, ["out"])
    eps = torch.tensor(np.asarray(design.permittivity), dtype=torch.float64, requires_grad=True)
    J = objective(eps, prob)
    J.backward()
    J2, g = prob.run(np.asarray(design.permittivity))
    assert float(J.detach()) == pytest.approx(J2, rel=1e-12)
    np.testing.assert_allclose(eps.grad.numpy(), g, rtol=1e-12, atol=0)


def test_design_loop_reuses_the_session_bitwise(hip):
    """Gradients of a design loop on one AdjointProblem (the session's solver, plans,
    graphs and buffers kept; only cb replaced) equal fresh problems' bit for bit, for
    a new permittivity and back to the first."""
    sim, design = design_problem("single", seed=12, steps=301)
    rng = np.random.default_rng(3)
    eps1 = np.asarray(design.permittivity)
    eps2 = rng.uniform(2.085, 12.11, size=eps1.shape)
    loop = AdjointProblem(sim, ["out"], checkpoint_every=64)
    for eps in (eps1, eps2, eps1):
        J, g = loop.run(eps)
        J0, g0 = AdjointProblem(sim, ["out"], checkpoint_every=64, reuse=False).run(eps)
        assert J == J0
        assert np.array_equal(g, g0)
    assert loop._session is not None and loop._session.rev is not None
    loop.close()


def _wide_problem(precision, steps=640):
    """A 12 x 10 um grid (256 x 216 cells): the 0.6 um design box's recompute window
    (box + 66 nodes per side at K = 64) is a fraction of it, and the source lies outside
    the window."""
    return design_problem(precision, seed=6, steps=steps, size=(12.0, 10.0))


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("schedule", ["joint", "lanes", "sequential"])
def test_window_recompute_equals_the_full_grid_bitwise(hip, precision, schedule):
    """Recomputing each segment on the design box's window gives the whole grid's
    gradient bit for bit, in every schedule."""
    sim, design = _wide_problem(precision)
    eps = np.asarray(design.permittivity)
    win = AdjointProblem(sim, ["out"], checkpoint_every=64, schedule=schedule)
    full = AdjointProblem(sim, ["out"], checkpoint_every=64, schedule=schedule, window=False)
    Jw, gw = win.run(eps)
    Jf, gf = full.run(eps)
    i0, j0, ni, nj = win.box
    m, nx, ny = win.K + 2, win.grid.nx, win.grid.ny
    wx = min(nx, i0 + ni + m) - max(0, i0 - 1 - m)
    wy = min(ny, j0 + nj + m) - max(0, j0 - 1 - m)
    print(f"{precision} {schedule}: grid {nx} x {ny}, window {wx} x {wy}, "
          f"max |g| {np.abs(gf).max():.3e}", flush=True)
    assert wx * wy < nx * ny / 2
    if schedule == "joint":  # the design-loop session keeps the tape's checkpoints
        n_ck = len(win._session.starts)
        assert win._session.checkpoints["hz"].shape == (n_ck, wx * wy)
        assert win._session.checkpoints["ex"].shape == (n_ck, wx * (wy + 1))
        assert full._session.checkpoints["hz"].shape == (n_ck, nx * ny)
    assert np.abs(gf).max() > 0
    assert Jw == Jf
    assert np.array_equal(gw, gf)


def test_window_recompute_design_loop_and_batch_bitwise(hip):
    """A design loop on windows (their cb refreshed for each new permittivity) and a
    batch of candidates on windows equal fresh whole-grid problems bit for bit."""
    from etalon_amd.batch_adjoint import BatchAdjointProblem

    sim, design = _wide_problem("single")
    rng = np.random.default_rng(9)
    eps1 = np.asarray(design.permittivity)
    eps2 = rng.uniform(2.085, 12.11, size=eps1.shape)
    loop = AdjointProblem(sim, ["out"], checkpoint_every=64)
    for eps in (eps1, eps2):
        J, g = loop.run(eps)
        J0, g0 = AdjointProblem(sim, ["out"], checkpoint_every=64, window=False,
                                reuse=False).run(eps)
        assert J == J0 and np.array_equal(g, g0)
    loop.close()
    batch = BatchAdjointProblem([AdjointProblem(sim, ["out"], checkpoint_every=64)
                                 for _ in range(2)])
    for eps_list in ([eps1, eps2], [eps2, eps1]):
        Js, gs = batch.run(eps_list)
        for J, g, e in zip(Js, gs, eps_list):
            J0, g0 = AdjointProblem(sim, ["out"], checkpoint_every=64, window=False,
                                    reuse=False).run(e)
            assert J == J0 and np.array_equal(g, g0)
    batch.close()
