"""The benchmark workloads at their full sizes on the GPU, where the CPU
oracle would take hours: size-independent properties of the method stand in
for a comparison (conservation, monotone reactant, bounds), plus oracle
comparisons on the largest windows that finish in seconds.

Configurations (the workloads the bench measures):
  - 2D reactive Euler channel, 1024 x 256, order 3, stiff Newton predictor,
    reflective walls, seed 99;
  - 2D periodic Euler, 512 x 512, order 4, eight seeded Fourier modes, run to
    t = 0.5;
  - 48 independent 2D Euler grids with sides from {64, 128, 256, 512},
    four-quadrant Riemann data, order 3, transmissive walls."""
import numpy as np
import pytest

import cpu_checkers as cc
from pypde_amd import systems
from pypde_amd.solvers import Ensemble, Solver

GAMMA_EULER = 1.4


def sums(u):
    """Per-variable sums and the sums of magnitudes they are measured against."""
    V = u.shape[-1]
    f = u.reshape(-1, V)
    return f.sum(0), np.abs(f).sum(0)


def euler_pressure(u, gamma=GAMMA_EULER):
    r, ru, rv, E = (u[..., k] for k in range(4))
    return (gamma - 1.0) * (E - 0.5 * (ru * ru + rv * rv) / r)


@pytest.mark.gpu
def test_reactive_channel_full_size_invariants():
    """1024 x 256 channel, 25 steps in chunks of 5.  The walls reflect and the
    source acts on rho*lambda alone, so the sums of rho and E are conserved to
    rounding (asserted: 1e-12 of the sum of magnitudes); the source is never
    positive where rho*lambda >= 0, so total reactant never grows; lambda =
    rho*lambda / rho stays in [-1e-7, 1.005].  The scheme does not preserve
    bounds (WENO reconstructs rho*lambda and rho separately): on the oracle's
    128 x 32 window of the same channel lambda reaches 1.0012 and -2.6e-9 in
    40 steps."""
    nx, ny = 1024, 256
    u0 = systems.reactive_channel(nx, ny, seed=99)
    sol = Solver(u0, [1.0, ny / nx], systems.source("reactive2d"), order=3,
                 boundaryTypes="reflective", stiff=True, reflect=systems.REACTIVE_REFLECT,
                 newton_tol=1e-12, newton_maxit=100)
    s0, m0 = sums(u0)
    prev_rl = s0[4]
    for chunk in range(5):
        st = sol.step(5)
        u = sol.state()
        assert st["rc"] == 0 and st["steps"] == 5 * (chunk + 1)
        assert np.isfinite(u).all()
        s, _ = sums(u)
        drift = np.abs(s - s0) / m0
        print(f"step {st['steps']}: rho drift {drift[0]:.2e}, E drift {drift[3]:.2e}, "
              f"reactant {s[4]:.9e}")
        assert drift[0] <= 1e-12 and drift[3] <= 1e-12
        assert s[4] <= prev_rl * (1 + 1e-14)
        prev_rl = s[4]
        lam = u[..., 4] / u[..., 0]
        assert lam.min() >= -1e-7 and lam.max() <= 1.005, (lam.min(), lam.max())
        assert u[..., 0].min() > 0
    assert st["pred_capped"] == 0
    sol.close()


@pytest.mark.gpu
def test_reactive_channel_256x64_window_matches_oracle():
    """A quarter-width, quarter-height channel with the full channel's cell
    size (dx = 1/1024), 25 steps, against the oracle: bitwise equal."""
    nx, ny = 256, 64
    u0 = systems.reactive_channel(nx, ny, seed=99)
    L = [nx / 1024, ny / 1024]
    sol = Solver(u0, L, systems.source("reactive2d"), order=3, boundaryTypes="reflective",
                 stiff=True, reflect=systems.REACTIVE_REFLECT)
    st = sol.step(25)
    ug = sol.state()
    sol.close()
    uo, _, sto, rc = cc.oracle_solve("reactive2d", u0, 1.0, L, 2, 3, stiff=True, max_steps=25,
                                     reflect=systems.REACTIVE_REFLECT)
    assert rc == 0 and st["steps"] == sto.steps == 25
    assert st["pred_iters"] == sto.pred_iters
    assert np.array_equal(ug, uo)


@pytest.mark.gpu
def test_euler_modes_512_order4_runs_to_t_end_conserving():
    """The 2D periodic Euler configuration at full size: 512 x 512, order 4,
    eight seeded Fourier modes per field, run to t = 0.5.  Periodic walls and
    no source: every variable's sum is conserved to rounding (asserted: 1e-12
    of the sum of magnitudes, over the whole run), and density and pressure
    stay positive."""
    u0 = systems.euler2d_modes(512, seed=0)
    sol = Solver(u0, [1.0, 1.0], systems.source("euler2d"), order=4, boundaryTypes="periodic")
    st = sol.run(0.5)
    u = sol.state()
    sol.close()
    assert st["rc"] == 0 and abs(st["t"] - 0.5) <= 1e-12
    assert np.isfinite(u).all()
    s0, m0 = sums(u0)
    s, _ = sums(u)
    drift = np.abs(s - s0) / m0
    print(f"{st['steps']} steps to t = {st['t']}: drift {drift}, "
          f"Picard iterations per cell {st['pred_iters'] / st['pred_cells']:.3f}")
    assert drift.max() <= 1e-12
    assert u[..., 0].min() > 0 and euler_pressure(u).min() > 0


@pytest.mark.gpu
def test_ensemble_48_grids_full_size():
    """The bench's 48-grid ensemble (sides from {64, 128, 256, 512}, 3.1M
    cells), 20 steps: every member equals its own single-grid run bit for bit,
    density and pressure stay positive everywhere, and the 64 x 64 members
    equal the oracle."""
    shapes = systems.ensemble_sizes(48)
    grids = [systems.riemann4(nx, ny, seed=i) for i, (nx, ny) in enumerate(shapes)]
    src = systems.source("euler2d")
    ens = Ensemble(grids, [[1.0, 1.0]] * len(grids), src, order=3)
    stats, _ = ens.run(1.0, max_steps=20)
    outs = [ens.state(i) for i in range(len(grids))]
    ens.close()
    n_oracle = 0
    for i, (g, (nx, ny)) in enumerate(zip(grids, shapes)):
        assert stats[i]["steps"] == 20 and stats[i]["nan_step"] == 0, i
        u = outs[i]
        assert np.isfinite(u).all() and u[..., 0].min() > 0 and euler_pressure(u).min() > 0, i
        sol = Solver(g, [1.0, 1.0], src, order=3)
        sol.run(1.0, max_steps=20)
        single = sol.state()
        sol.close()
        assert np.array_equal(u, single), i
        if nx == 64 and ny == 64 and n_oracle < 2:
            uo, _, sto, rc = cc.oracle_solve("euler2d", g, 1.0, [1.0, 1.0], 0, 3, max_steps=20)
            assert rc == 0 and np.array_equal(u, uo), i
            n_oracle += 1
    print(f"{len(grids)} grids, {sum(nx * ny for nx, ny in shapes)} cells; "
          f"{n_oracle} checked against the oracle")
