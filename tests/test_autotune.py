"""The autotuner: every launch variant it chooses between gives bit-identical
results, and its choice is measured once per problem and then reused."""
import os

import numpy as np
import pytest

import cpu_checkers as cc
from pypde_amd import systems
from pypde_amd.solvers import Solver

# (WENO tile, flux workgroup, corrector mapping, predictor schedule) variants;
# schedules 2 and 3 are the cold/hot split of the stiff predictor, its cold pass
# built for two or three wavefronts per SIMD (the static one elsewhere)
VARIANTS = [(w, f, u, p) for w, f, u, p in
            [(0, 0, 0, 0), (1, 1, 1, 1), (2, 2, 2, 2), (3, 0, 3, 3), (0, 1, 4, 2)]]


def _run(system, u0, L, bc, N, stiff, steps, force):
    os.environ["PYPDE_TUNE_FORCE"] = ",".join(map(str, force))
    try:
        refl = systems.REACTIVE_REFLECT if system == "reactive2d" else None
        sol = Solver(u0, L, systems.source(system), order=N, boundaryTypes=bc, stiff=stiff,
                     reflect=refl)
        st = sol.run(1.0, max_steps=steps)
        tun = sol.tuning()
        u = sol.state()
        sol.close()
    finally:
        del os.environ["PYPDE_TUNE_FORCE"]
    assert tun["source"] == "forced"
    return u, st


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["reactive_stiff", "euler2d_order3", "sod_order3"])
def test_every_variant_bitwise_equal_to_oracle(case):
    system, u0, L, bc, N, stiff, steps = {
        "reactive_stiff": ("reactive2d", systems.reactive_channel(67, 29, seed=5),
                           [67 / 1024, 29 / 1024], "reflective", 3, True, 10),
        "euler2d_order3": ("euler2d", systems.riemann4(45, 70, seed=3), [1.0, 1.0],
                           "transitive", 3, False, 15),
        "sod_order3": ("euler1d", systems.sod1d(133), [1.0], "transitive", 3, False, 30),
    }[case]
    bcode = {"transitive": 0, "periodic": 1, "reflective": 2}[bc]
    uo, _, sto, rc = cc.oracle_solve(system, u0, 1.0, L, bcode, N, stiff=stiff, max_steps=steps,
                                     reflect=systems.REACTIVE_REFLECT
                                     if system == "reactive2d" else None)
    assert rc == 0
    for force in VARIANTS:
        u, st = _run(system, u0, L, bc, N, stiff, steps, force)
        assert st["steps"] == steps and st["pred_iters"] == sto.pred_iters, force
        assert np.array_equal(u, uo), force


@pytest.mark.gpu
def test_tuner_measures_once_then_reuses():
    u0 = systems.riemann4(96, 80, seed=11)
    os.environ.pop("PYPDE_TUNE_FORCE", None)
    a = Solver(u0, [1.0, 1.0], systems.source("euler2d"), order=3)
    a.run(1.0, max_steps=2)
    ta = a.tuning()
    a.close()
    b = Solver(u0, [1.0, 1.0], systems.source("euler2d"), order=3)
    b.run(1.0, max_steps=2)
    tb = b.tuning()
    b.close()
    print("tuned:", ta)
    assert ta["source"] in ("measured", "table") and tb["source"] == "table"
    assert {k: ta[k] for k in ("weno_tile", "flux_workgroup", "update", "predictor")} == \
        {k: tb[k] for k in ("weno_tile", "flux_workgroup", "update", "predictor")}
