"""Mixed precision (an FP32 predictor phase, then the FP64 predictor from its
iterate) against the FP64 oracle, per benchmark configuration.

What bounds the difference: the stiff predictor iterates in FP64 to the same
tolerance (1e-12) from a different starting point, so one step differs from
the oracle only at the tolerance's level; the Picard predictor stops at the
reference's 6e-6 tolerance (dg.cpp:187-202), so a different starting point
moves its iterate, and the step, by up to about that much.  Over many steps
the finite-difference wave speeds (NumericalDiff.h:6-44) amplify any
difference (the plain method itself moves by 5e-6 .. 4e-4 under one-ulp input
changes on Sod).  Each test prints its measured errors."""
import numpy as np
import pytest

import cpu_checkers as cc
from pypde_amd import systems
from pypde_amd.solvers import Solver


def relerr(a, b):
    V = b.shape[-1]
    sc = np.max(np.abs(b.reshape(-1, V)), axis=0)
    sc = np.maximum(sc, 1e-8 * sc.max() + 1e-300)
    return float(np.max(np.abs((a - b).reshape(-1, V)) / sc))


CONFIGS = {
    # name: (system, initial state, L, bc, order, stiff, steps, bound 1 step, bound all steps)
    # measured on the MI355X: Sod 3.6e-8 (1 step), 1.3e-4 (40 steps: within the 4e-4 that
    # the reference itself moves by under one-ulp input changes); reactive channel
    # 2.2e-10 (1 step), 7.5e-7 (12 steps)
    "sod_order3": ("euler1d", lambda: systems.sod1d(200), [1.0], 0, 3, False, 40, 1e-6, 4e-4),
    "euler2d_modes_order4": ("euler2d", lambda: systems.euler2d_modes(64, seed=0), [1.0, 1.0],
                             1, 4, False, 20, 1e-6, 1e-4),
    "reactive_channel_64x16": ("reactive2d", lambda: systems.reactive_channel(64, 16, seed=99),
                               [64 / 1024, 16 / 1024], 2, 3, True, 12, 1e-9, 1e-5),
    "riemann4_64": ("euler2d", lambda: systems.riemann4(64, 64, seed=0), [1.0, 1.0], 0, 3,
                    False, 20, 1e-6, 1e-4),
}


def _gpu(system, u0, L, bc, N, stiff, steps, mixed):
    refl = systems.REACTIVE_REFLECT if system == "reactive2d" else None
    bcn = {0: "transitive", 1: "periodic", 2: "reflective"}[bc]
    sol = Solver(u0, L, systems.source(system), order=N, boundaryTypes=bcn, stiff=stiff,
                 reflect=refl, mixed=mixed)
    st = sol.run(1.0, max_steps=steps)
    u = sol.state()
    sol.close()
    return u, st


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", sorted(CONFIGS))
def test_mixed_against_fp64_oracle(cfg):
    system, mk, L, bc, N, stiff, steps, b1, bn = CONFIGS[cfg]
    u0 = mk()
    refl = systems.REACTIVE_REFLECT if system == "reactive2d" else None
    errs = {}
    for k in (1, steps):
        uo, _, sto, rc = cc.oracle_solve(system, u0, 1.0, L, bc, N, stiff=stiff, max_steps=k,
                                         reflect=refl)
        assert rc == 0
        um, stm = _gpu(system, u0, L, bc, N, stiff, k, True)
        up, stp = _gpu(system, u0, L, bc, N, stiff, k, False)
        assert stm["rc"] == 0 and stm["steps"] == k
        assert np.array_equal(up, uo)          # the plain path is the oracle, bit for bit
        errs[k] = relerr(um, uo)
        print(f"{cfg}: {k} steps, mixed vs FP64 oracle {errs[k]:.3e}; FP32-phase iterations "
              f"per cell {stm['pred_iters_f32'] / stm['pred_cells']:.2f}, FP64 iterations per "
              f"cell {stm['pred_iters'] / stm['pred_cells']:.2f} (plain "
              f"{stp['pred_iters'] / stp['pred_cells']:.2f})")
        assert stm["pred_iters_f32"] > 0
    assert errs[1] <= b1 and errs[steps] <= bn, errs


@pytest.mark.gpu
def test_mixed_rejects_b_systems():
    from pypde_amd import _capi
    with pytest.raises(_capi.PypdeError, match="code 4"):
        Solver(np.ones((16, 16, 2)), [1.0, 1.0], systems.source("ncadvect"), order=3,
               mixed=True)
