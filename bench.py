"""Headline benchmark: 2D reactive Euler (Arrhenius, nvar 5), 1024 x 256
channel, order 3, stiff per-cell Newton predictor, reflective walls, FP64.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]

Prints one JSON line.  A step is one full ADER-WENO timestep of the whole
grid (boundaries + WENO, CFL reduction, predictor, face fluxes, corrector);
the state is resident in HBM before the timed region.  `value` is
cell-updates/s = cells x K / wall time of the K timed steps.
"""
import argparse
import ctypes as C
import json
import os
import re
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pypde_amd import systems  # noqa: E402
from pypde_amd.solvers import Solver  # noqa: E402

NX, NY, ORDER, V, D = 1024, 256, 3, 5, 2
L = [1.0, NY / NX]
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12

# Flop counts of the reactive system functions (pypde_amd/systems/reactive2d.pde),
# counted by hand: + - * / sqrt each 1, fma 2; pde_exp = 33.
FLOPS_F, FLOPS_S, FLOPS_M = 17, 46, 78


def predictor_flops_per_iteration(N=ORDER, D=D, V=V, cF=FLOPS_F, cS=FLOPS_S, cM=FLOPS_M):
    """Algorithmic FP64 flops of one stiff-predictor iteration of one cell
    (one space-time residual plus the per-node Newton correction)."""
    ND = N ** D
    deriv = 2 * N                                    # N-term dot product + divide
    f = N * ND * (D * V * deriv                      # grad q
                  + D * cF + D * V * deriv           # F and div F
                  + cS + V * (D + 4)                 # S, assembly, *W, *dt, +Ww
                  + V * (2 * N + 1))                 # residual G = rhs - K q
    node = (N * V + 2 * N * V + cM + V * V           # scaling, mean state, dS/dQ, A
            + N * (N * V * 2 * N + (N - 1) * (V * 2 * V + V))  # h_t (Horner)
            + (N - 1) * (V * V * 2 * V + V)          # P = p(A)
            + sum((V - k - 1) * (1 + 2 * (V - k - 1)) for k in range(V))  # LU
            + N * (2 * V * V)                        # triangular solves
            + 2 * N * V)                             # update + test
    return f + ND * node


def predictor_flops_epilogue(N=ORDER, D=D, V=V, cS=FLOPS_S):
    ND = N ** D
    traces = 2 * D * N * N ** (D - 1) * V * (2 * N - 1)
    volume = N * ND * (cS + (D + 1) + 2 * V) + ND * V
    return traces + volume


def gpu_run(steps, warmup, profile):
    u0 = systems.reactive_channel(NX, NY, seed=0)
    sol = Solver(u0, L, systems.source("reactive2d"), order=ORDER, boundaryTypes="reflective",
                 stiff=True, reflect=systems.REACTIVE_REFLECT, newton_tol=1e-12,
                 newton_maxit=100, profile=profile)
    w = sol.step(warmup) if warmup > 0 else None
    t0 = time.perf_counter()
    st = sol.step(steps)
    t1 = time.perf_counter()
    state = sol.state()
    sol.close()
    return st, t1 - t0, w, state


def cpu_baseline(t_target, threads):
    """The reference's own C++ core (oracle/_ref, built from its sources) on the
    same 1024 x 256 initial state for the first timesteps up to t_target.  The
    reference has no reflective walls; transmissive walls cost the same here."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import cpu_checkers as cc
    if not cc.ref_available():
        return None
    s = cc.system("reactive2d")
    u = np.ascontiguousarray(systems.reactive_channel(NX, NY, seed=0)).copy()
    nX = np.array([NX, NY], dtype=np.int32)
    dX = np.array([L[0] / NX, L[1] / NY])
    bc = np.array([0, 0], dtype=np.int32)
    ret = np.zeros(u.size)
    tf = t_target * (1 - 1e-9)
    with tempfile.TemporaryFile(mode="w+") as f:
        sys.stdout.flush()
        saved = os.dup(1)
        os.dup2(f.fileno(), 1)
        t0 = time.perf_counter()
        try:
            cc.ref().pde_solver(s.F, None, s.S, True, False, True, cc.dp(u), tf, cc.ip(nX), 2,
                                cc.dp(dX), 0.9, cc.ip(bc), True, 0, ORDER, V, 1, False,
                                cc.dp(ret), threads)
        finally:
            el = time.perf_counter() - t0
            C.CDLL(None).fflush(None)
            os.dup2(saved, 1)
            os.close(saved)
        f.seek(0)
        steps = len(re.findall(r"^t = ", f.read(), re.M))
    return dict(value=NX * NY * steps / el, unit="cell-updates/s", cores=threads,
                kind="reference",
                sample=f"reference pde_solver (oracle/_ref, -O3), {NX}x{NY} reactive channel, "
                       f"order 3, stiff Newton-Krylov (its 6e-6 tolerance), {steps} timesteps "
                       f"from t=0, {threads} threads, transmissive walls, {el:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    K, W = args.steps, args.warmup
    st, el, wst, state = gpu_run(K, W, profile=True)
    cells = NX * NY
    value = cells * K / el
    # roofline of the dominant kernel: the predictor, FP64-bound
    ms = st["ms_kernel"]
    pred_ms = ms["predictor"]
    flops = (predictor_flops_per_iteration() * st["pred_iters"]
             + predictor_flops_epilogue() * st["pred_cells"])
    achieved = flops / (pred_ms * 1e-3) if pred_ms > 0 else 0.0
    roof = dict(bound="fp64", achieved=achieved / 1e12, peak=PEAK_FP64 / 1e12, unit="TFLOP/s",
                frac=achieved / PEAK_FP64, traffic=None, kernel="k_predictor",
                avg_launch_ms=pred_ms / K)
    cpu = None
    if not args.no_cpu:
        # t after args.cpu_steps steps from t = 0, from a short GPU run
        st2, _, _, _ = gpu_run(args.cpu_steps, 0, profile=False)
        cpu = cpu_baseline(st2["t"], threads=min(16, os.cpu_count() or 1))
    out = {
        "metric": "cell-updates/s, 2D reactive Euler O3: learned+mixed vs plain FP64 @Linf<=1e-8",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": el * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded initial condition, SplitMix64 seed 0)",
        "config": {"workload": "2D reactive Euler, Arrhenius (Q=25, Ea=20, K=1e4, gamma=1.2), "
                               "1024x256 channel, order 3, Rusanov, CFL 0.9, reflective walls, "
                               "stiff per-cell Newton predictor (tol 1e-12), plain FP64",
                   "cells": cells, "order": ORDER, "nvar": V},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernel_ms_per_step": {k: v / K for k, v in ms.items()},
        "predictor_iters_per_cell": st["pred_iters"] / max(1, st["pred_cells"]),
        "predictor_max_iters": st["pred_maxit"],
        "predictor_capped": st["pred_capped"],
        "t_end": st["t"],
        "finite": bool(np.isfinite(state).all()),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
