"""HIP path vs the CPU oracle.

The kernels issue the oracle's floating-point operations in the same order, so
on identical inputs the expected difference is zero; every test states the
tolerance it enforces (at most a relative L-inf of 1e-10 for FP64, where not
bitwise).
CPU-only tests here check the C ABI itself (exports, tables, errors)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cpu_checkers as cc
from pypde_amd import _capi, systems
from pypde_amd.solvers import Solver, pde_solver, weno_solver

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = 1e-10   # max relative L-inf, per variable


def relerr(a, b):
    """max over variables of |a - b|_inf / |b_v|_inf; a variable whose scale is
    below 1e-8 of the largest one (e.g. transverse momentum of a quasi-1D
    flow, ~1e-10) is measured against that floor instead."""
    V = b.shape[-1]
    sc = np.max(np.abs(b.reshape(-1, V)), axis=0)
    sc = np.maximum(sc, 1e-8 * sc.max() + 1e-300)
    return float(np.max(np.abs((a - b).reshape(-1, V)) / sc))


# ------------------------------------------------------------ CPU checks

def test_library_exports_header_symbols():
    hdr = open(os.path.join(ROOT, "include", "pypde_hip.h")).read()
    names = re.findall(r"^\s*(?:int|void|const char|pypde_ctx)\s*\*?\s*(\w+)\s*\(", hdr, re.M)
    assert len(names) >= 12
    lib = _capi.lib()
    for n in names:
        assert hasattr(lib, n), n


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_tables_bitwise_equal_to_oracle(N):
    p, o = _capi.tables(N), cc.tables(N)
    for k in p:
        assert np.array_equal(p[k], o[k]), k


def test_argument_errors_are_reported():
    with pytest.raises(ValueError):
        pde_solver(np.ones((8, 3)), 0.1, [1.0], system=systems.source("euler1d"),
                   boundaryTypes="bogus")
    o = _capi.make_options()
    rc = _capi.lib().pypde_hip_compile(systems.source("euler2d").encode(), True, False, False,
                                       3, 3, 4, False, 0, False, C.byref(o))
    assert rc == 4 and "ndim=3" in _capi.last_error()
    rc = _capi.lib().pypde_hip_compile(b"this is not C", True, False, False, 1, 3, 3, False,
                                       0, False, C.byref(o))
    assert rc == 2 and "error" in _capi.last_error()


def test_no_cpu_fallback_without_gpu():
    """The product path has no CPU fallback: without a GPU it fails loudly."""
    _capi.lib().pypde_device_count.restype = C.c_int
    if _capi.lib().pypde_device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(_capi.PypdeError, match="GPU|hip"):
        pde_solver(systems.sod1d(32), 0.01, [1.0], system=systems.source("euler1d"), order=3,
                   ndt=1, stiff=False)


def test_flux_index_multiply_shift_is_exact():
    """The flux kernel divides item indices n < 2^31 by nf and by the faces
    per row as (n * m) >> sh, m = ceil(2^sh / d), sh = 31 + ceil(log2 d).
    Exhaustive over small divisors and n, and at the extremes of n and d
    (powers of two and their neighbours, where the error term m d - 2^sh is
    largest relative to d)."""
    def magic(d):
        L = (d - 1).bit_length()
        sh = 31 + L
        return -(-(1 << sh) // d), sh

    ds = list(range(1, 600)) + [(1 << k) + o for k in range(10, 31) for o in (-1, 0, 1)
                                if 1 <= (1 << k) + o < (1 << 31)] + [(1 << 31) - 1]
    rng = np.random.default_rng(7)
    top = np.uint64((1 << 31) - 1)
    for d in ds:
        m, sh = magic(d)
        assert m < (1 << 32)
        n = np.concatenate([np.arange(0, 3000, dtype=np.uint64),
                            top - np.arange(0, 3000, dtype=np.uint64),
                            rng.integers(0, 1 << 31, 3000, dtype=np.uint64)])
        q = (n * np.uint64(m)) >> np.uint64(sh)
        assert np.array_equal(q, n // np.uint64(d)), d


# ------------------------------------------------------------ GPU parity

def oracle_run(name, u0, tf, L, bc, N, stiff, **kw):
    return cc.oracle_solve(name, u0, tf, L, bc, N, stiff=stiff, **kw)


def gpu_run(name, u0, tf, L, bc, N, stiff, ndt=1, **kw):
    st = {}
    u = np.ascontiguousarray(u0, dtype=np.float64).copy()
    ret = pde_solver(u, tf, L, system=systems.source(name), boundaryTypes=bc, order=N, ndt=ndt,
                     stiff=stiff, stats=st, **kw)
    return u, ret, st


BCNAME = {0: "transitive", 1: "periodic", 2: "reflective"}


@pytest.mark.gpu
@pytest.mark.parametrize("ndim,N,V", [(1, 2, 3), (1, 3, 4), (2, 2, 3), (2, 3, 5), (2, 4, 4)])
def test_weno_solver_matches_oracle(ndim, N, V):
    rng = np.random.default_rng(40 + N)
    nX = [37, 29][:ndim]
    u = rng.uniform(0.5, 2.0, tuple(nX) + (V,))
    got = weno_solver(u, order=N)
    nXa = np.array(nX, dtype=np.int32)
    want = np.zeros(got.size)
    cc.oracle().oracle_weno(cc.dp(want), cc.dp(np.ascontiguousarray(u)), cc.ip(nXa), ndim, N, V)
    assert np.array_equal(got.ravel(), want)


@pytest.mark.gpu
def test_sod_order3_matches_oracle():
    """C0: 1D Sod, 200 cells, order 3, Rusanov, CFL 0.9, tf 0.2, 4 snapshots."""
    u0 = systems.sod1d()
    uo, ro, st, rc = oracle_run("euler1d", u0, 0.2, [1.0], 0, 3, False, ndt=4)
    ug, rg, sg = gpu_run("euler1d", u0, 0.2, [1.0], "transitive", 3, False, ndt=4)
    print("sod: steps", sg["steps"], "oracle", st.steps, "err", relerr(ug, uo))
    assert rc == 0 and sg["rc"] == 0
    assert sg["steps"] == st.steps == 103
    assert sg["pred_iters"] == st.pred_iters
    assert relerr(ug, uo) <= TOL
    assert relerr(rg, ro) <= TOL


@pytest.mark.gpu
@pytest.mark.parametrize("N,bc", [(2, 1), (3, 1), (4, 1), (3, 0)])
def test_euler2d_matches_oracle(N, bc):
    u0 = systems.euler2d_modes(40, seed=3)
    if bc == 0:
        u0 = u0[:, :24].copy()
    L = [1.0, u0.shape[1] / u0.shape[0]]
    uo, ro, st, rc = oracle_run("euler2d", u0, 0.04, L, bc, N, False)
    ug, rg, sg = gpu_run("euler2d", u0, 0.04, L, BCNAME[bc], N, False)
    print(f"euler2d N={N} bc={bc}: steps {sg['steps']} err {relerr(ug, uo):.3e}")
    assert sg["steps"] == st.steps and sg["pred_iters"] == st.pred_iters
    assert relerr(ug, uo) <= TOL


EULER2D_REFLECT = np.array([[1, -1, 1, 1], [1, 1, -1, 1]], dtype=np.float64)


@pytest.mark.gpu
@pytest.mark.parametrize("bcs,N", [((1, 2), 3), ((0, 1), 2), ((2, 0), 4)])
def test_mixed_boundaries_bitwise_equal_to_oracle(bcs, N):
    """One boundary type per axis (boundaries.cpp:7-53 takes them per axis;
    reflective is an addition to the reference's codes): 2D Euler on a 40 x 24 window,
    reflective walls flip the wall-normal momentum.  Tolerance: bitwise, with
    equal step and Picard iteration counts."""
    u0 = systems.euler2d_modes(40, seed=5)[:, :24].copy()
    L = [1.0, 24 / 40]
    refl = EULER2D_REFLECT if 2 in bcs else None
    uo, ro, st, rc = oracle_run("euler2d", u0, 0.04, L, list(bcs), N, False,
                                reflect=None if refl is None else refl.ravel())
    ug, rg, sg = gpu_run("euler2d", u0, 0.04, L, [BCNAME[b] for b in bcs], N, False,
                         reflect=refl)
    print(f"mixed bcs {bcs} N={N}: steps {sg['steps']} oracle {st.steps}")
    assert rc == 0 and sg["rc"] == 0
    assert sg["steps"] == st.steps and sg["pred_iters"] == st.pred_iters
    assert np.array_equal(ug, uo)


@pytest.mark.gpu
def test_flux_64bit_index_path_bitwise_equal_to_oracle(monkeypatch):
    """The flux kernels' 64-bit index path (taken when a grid's flux items or
    w-cells reach 2^31; forced here by PYPDE_FLUX_IDX64=1 on small grids) gives
    the oracle's results bitwise: 2D Euler with periodic and reflective axes,
    and 1D Sod."""
    monkeypatch.setenv("PYPDE_FLUX_IDX64", "1")
    u0 = systems.euler2d_modes(40, seed=5)[:, :24].copy()
    L = [1.0, 24 / 40]
    uo, ro, st, rc = oracle_run("euler2d", u0, 0.04, L, [1, 2], 3, False,
                                reflect=EULER2D_REFLECT.ravel())
    ug, rg, sg = gpu_run("euler2d", u0, 0.04, L, ["periodic", "reflective"], 3, False,
                         reflect=EULER2D_REFLECT)
    assert rc == 0 and sg["rc"] == 0 and sg["steps"] == st.steps
    assert np.array_equal(ug, uo)
    u0 = systems.sod1d(133)
    uo, ro, st, rc = oracle_run("euler1d", u0, 0.05, [1.0], 0, 3, False)
    ug, rg, sg = gpu_run("euler1d", u0, 0.05, [1.0], "transitive", 3, False)
    assert rc == 0 and sg["steps"] == st.steps
    assert np.array_equal(ug, uo)


def reactive_small(nx=64, ny=16, seed=0):
    return systems.reactive_channel(nx, ny, seed=seed)


@pytest.mark.gpu
@pytest.mark.parametrize("stiff", [True, False])
def test_reactive_channel_matches_oracle(stiff):
    """C2 physics on a 64 x 16 window of the 1024 x 256 grid spacing (dx = 1/1024),
    reflective walls, fixed 12-step window; stiff = per-cell Newton with the
    analytic dS/dQ."""
    u0 = reactive_small()
    L = [64 / 1024, 16 / 1024]
    refl = systems.REACTIVE_REFLECT
    uo, ro, st, rc = oracle_run("reactive2d", u0, 1.0, L, 2, 3, stiff, max_steps=12,
                                reflect=refl.ravel())
    ug, rg, sg = gpu_run("reactive2d", u0, 1.0, L, "reflective", 3, stiff, max_steps=12,
                         reflect=refl)
    print("reactive stiff", stiff, sg, "err", relerr(ug, uo))
    assert sg["steps"] == st.steps == 12
    assert sg["pred_iters"] == st.pred_iters and sg["pred_capped"] == st.pred_capped == 0
    assert relerr(ug, uo) <= TOL


@pytest.mark.gpu
@pytest.mark.parametrize("quirk", [0, 1])
def test_nonconservative_b_matches_oracle(quirk):
    x = (np.arange(24) + 0.5) / 24
    X, Y = np.meshgrid(x, x, indexing="ij")
    u0 = np.stack([np.sin(2 * np.pi * X), np.cos(2 * np.pi * Y)], -1)
    uo, ro, st, rc = oracle_run("ncadvect", u0, 0.05, [1, 1], 1, 3, False, b_quirk=quirk)
    ug, rg, sg = gpu_run("ncadvect", u0, 0.05, [1, 1], "periodic", 3, False, b_quirk=quirk)
    assert sg["steps"] == st.steps
    assert relerr(ug, uo) <= TOL


def first_step_predictor(u0, L, refl, N=3, V=5):
    """Oracle stages of the first step: ghosts, WENO, dt, converged stiff
    predictor; returns (qh, nw)."""
    s = cc.system("reactive2d")
    O = cc.oracle()
    nX = np.array(u0.shape[:2], dtype=np.int32)
    nb = nX + 2 * N
    dX = np.array([L[0] / nX[0], L[1] / nX[1]])
    ub = np.zeros(int(np.prod(nb)) * V)
    r = np.ascontiguousarray(refl.ravel())
    O.oracle_boundaries(cc.dp(ub), cc.dp(np.ascontiguousarray(u0)), cc.ip(nX), 2,
                        cc.ip(np.array([2, 2], dtype=np.int32)), cc.dp(r), N, V)
    nw = int(np.prod(nX + 2))
    wh = np.zeros(nw * N * N * V)
    O.oracle_weno(cc.dp(wh), cc.dp(ub), cc.ip(np.ascontiguousarray(nb, dtype=np.int32)), 2, N, V)
    dt = O.oracle_timestep(s.F, None, cc.dp(wh), nw, cc.dp(dX), 2, N, V, 0.9, 1.0, 0.0, 0)
    qh = np.zeros(nw * N ** 3 * V)
    opts = cc.OracleOpts()
    opts.newton_tol = 1e-12
    opts.newton_maxit = 100
    O.oracle_predictor(s.F, None, s.S, s.M, cc.dp(wh), nw, cc.dp(dX), 2, N, V, dt, 1,
                       C.byref(opts), cc.dp(qh), None)
    return qh.reshape(nw, N ** 3, V), nw


@pytest.mark.gpu
def test_learned_guess_path_matches_oracle():
    """Predictor started from supplied guesses (the initial-guess hook of
    dg.cpp:216): near-solution guesses on most cells, garbage on every third.
    GPU and oracle take the same iterations and fall back on the same cells,
    and the state agrees with the plain-guess run to the Newton tolerance."""
    u0 = reactive_small(48, 12)
    L = [48 / 1024, 12 / 1024]
    refl = systems.REACTIVE_REFLECT
    qh, nw = first_step_predictor(u0, L, refl)
    rng = np.random.default_rng(9)
    guess = qh * (1 + 1e-4 * rng.uniform(-1, 1, qh.shape))
    guess[::3] = rng.uniform(-1, 1, guess[::3].shape) * 1e3
    V = 5
    opts = cc.OracleOpts()
    opts.newton_tol = 1e-12
    opts.newton_maxit = 100
    opts.max_steps = 1
    r = np.ascontiguousarray(refl.ravel())
    opts.reflect = cc.dp(r)
    gg = np.ascontiguousarray(guess)
    opts.guess = cc.dp(gg)
    opts.guess_check = 1
    uo = np.ascontiguousarray(u0).copy()
    nX = np.array(u0.shape[:2], dtype=np.int32)
    ret = np.zeros(uo.size)
    sto = cc.OracleStats()
    s = cc.system("reactive2d")
    dX = np.array([1 / 1024, 1 / 1024])
    rc = cc.oracle().oracle_pde_solver(s.F, None, s.S, s.M, 1, 0, 1, cc.dp(uo), 1.0, cc.ip(nX), 2,
                                       cc.dp(dX), 0.9, cc.ip(np.array([2, 2], dtype=np.int32)), 1,
                                       0, 3, V, 1, 0, cc.dp(ret), C.byref(opts), C.byref(sto))
    assert rc == 0 and sto.guess_fallbacks >= nw // 3
    sol = Solver(u0, L, systems.source("reactive2d"), order=3, boundaryTypes="reflective",
                 stiff=True, reflect=refl, guess=guess)
    sg = sol.run(1.0, max_steps=1)
    ug = sol.state()
    sol.close()
    assert sg["guess_fallbacks"] == sto.guess_fallbacks
    assert sg["pred_iters"] == sto.pred_iters
    assert relerr(ug, uo) <= TOL
    up, _, stp, _ = oracle_run("reactive2d", u0, 1.0, L, 2, 3, True, max_steps=1,
                               reflect=refl.ravel())
    # both runs stop at |dq| <= 1e-12 (1 + |q|); the states agree to that scale
    assert np.max(np.abs(uo - up)) <= 1e-9 * np.max(np.abs(up))


@pytest.mark.gpu
@pytest.mark.parametrize("N,lo", [(3, 2.9), (4, 3.7)])
def test_convergence_order_on_smooth_wave_gpu(N, lo):
    """C1-style 2D periodic Euler on the GPU: an advected density wave (exact
    cell averages) converges at the scheme's order N from 32 to 128 cells per
    side, and the 32-cell run equals the oracle's."""
    errs = []
    for n in (32, 64, 128):
        u0 = systems.density_wave(n, 2)
        ug, _, sg = gpu_run("euler2d", u0, 0.1, [1.0, 1.0], "periodic", N, False)
        assert sg["rc"] == 0
        if n == 32:
            uo, _, st, rc = oracle_run("euler2d", u0, 0.1, [1.0, 1.0], 1, N, False)
            assert relerr(ug, uo) <= TOL
        ex = systems.density_wave(n, 2, t=0.1)
        errs.append(np.mean(np.abs(ug[..., 0] - ex[..., 0])))
    orders = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print(f"gpu N={N}: L1 errors {errs}, orders {orders}")
    assert orders[-1] >= lo and orders[0] >= lo - 0.5


@pytest.mark.gpu
def test_graph_replay_equals_eager():
    """A run whose timestep is one replayed HIP graph gives the eager run's
    bits (and the oracle's)."""
    u0 = systems.euler2d_modes(40, seed=5)
    L = [1.0, 1.0]
    outs = []
    for graph in (False, True):
        s = Solver(u0, L, systems.source("euler2d"), order=3, boundaryTypes="periodic",
                   use_graph=graph)
        st = s.run(0.05)
        outs.append((s.state(), st))
        s.close()
    (ue, se), (ugr, sgr) = outs
    assert se["steps"] == sgr["steps"] and np.array_equal(ue, ugr)
    uo, _, st, rc = oracle_run("euler2d", u0, 0.05, L, 1, 3, False)
    assert relerr(ugr, uo) <= TOL


@pytest.mark.gpu
@pytest.mark.parametrize("mode,cap", [("batched", None), ("batched", "1500"), ("streams", None)])
def test_ensemble_members_equal_single_runs(mode, cap, monkeypatch):
    """Ensemble: independent grids of different sizes with seeded 4-quadrant
    Riemann data, advanced together (one graph launch per step, each grid
    with its own dt), batched (one launch per stage over every grid, or over
    groups of at most 1500 reconstructed cells) or one stream per grid; every
    member equals its own single-grid run bit for bit, and the smallest ones
    equal the oracle."""
    from pypde_amd.solvers import Ensemble
    monkeypatch.setenv("PYPDE_ENSEMBLE", mode)
    if cap is not None:
        monkeypatch.setenv("PYPDE_ENSEMBLE_CELLS", cap)
    shapes = systems.ensemble_sizes(6, sizes=(16, 24, 32, 48))
    grids = [systems.riemann4(nx, ny, seed=i) for i, (nx, ny) in enumerate(shapes)]
    Ls = [[1.0, 1.0]] * len(grids)
    ens = Ensemble(grids, Ls, systems.source("euler2d"), order=3)
    stats, ms = ens.run(0.05)
    outs = [ens.state(i) for i in range(len(grids))]
    ens.close()
    print("ensemble shapes", shapes, "steps", [s["steps"] for s in stats], f"{ms:.2f} ms")
    for i, g in enumerate(grids):
        s = Solver(g, Ls[i], systems.source("euler2d"), order=3)
        st = s.run(0.05)
        ui = s.state()
        s.close()
        assert stats[i]["steps"] == st["steps"]
        assert np.array_equal(outs[i], ui), i
    small = sorted(range(len(grids)), key=lambda i: grids[i].size)[:2]
    for i in small:
        uo, _, st, rc = oracle_run("euler2d", grids[i], 0.05, Ls[i], 0, 3, False)
        assert rc == 0 and st.steps == stats[i]["steps"]
        assert relerr(outs[i], uo) <= TOL


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["reactive_stiff", "sod_1d"])
def test_batched_ensemble_stiff_and_1d(case, monkeypatch):
    """The batched ensemble kernels beyond 2D Euler: the stiff predictor's
    cold/hot split (reactive channel windows, reflective walls, each grid's
    hot cells listed and solved within the batched launches) and the 1D
    kernels (Sod tubes of different lengths).  Every member equals its own
    single-grid run bit for bit; the first also matches the oracle."""
    from pypde_amd.solvers import Ensemble
    monkeypatch.setenv("PYPDE_ENSEMBLE", "batched")
    if case == "reactive_stiff":
        shapes = [(40, 16), (24, 24), (32, 12), (16, 40)]
        grids = [systems.reactive_channel(nx, ny, seed=i + 3) for i, (nx, ny) in enumerate(shapes)]
        Ls = [[nx / 1024, ny / 1024] for nx, ny in shapes]
        kw = dict(boundaryTypes="reflective", stiff=True, reflect=systems.REACTIVE_REFLECT)
        system, N, steps, bcode = "reactive2d", 3, 8, 2
    else:
        grids = [systems.sod1d(n) for n in (50, 120, 200, 77, 150)]
        Ls = [[1.0]] * len(grids)
        kw = dict(boundaryTypes="transitive", stiff=False)
        system, N, steps, bcode = "euler1d", 3, 25, 0
    src = systems.source(system)
    ens = Ensemble(grids, Ls, src, order=N, **kw)
    stats, _ = ens.run(1.0, max_steps=steps)
    outs = [ens.state(i) for i in range(len(grids))]
    ens.close()
    for i, g in enumerate(grids):
        sol = Solver(g, Ls[i], src, order=N, **kw)
        st = sol.run(1.0, max_steps=steps)
        single = sol.state()
        sol.close()
        assert stats[i]["steps"] == st["steps"] == steps, i
        assert stats[i]["pred_iters"] == st["pred_iters"], i
        assert np.array_equal(outs[i], single), i
    refl = kw.get("reflect")
    uo, _, sto, rc = cc.oracle_solve(system, grids[0], 1.0, Ls[0], bcode, N,
                                     stiff=kw["stiff"], max_steps=steps,
                                     reflect=None if refl is None else refl.ravel())
    assert rc == 0 and sto.steps == steps and relerr(outs[0], uo) <= TOL


@pytest.mark.gpu
def test_predictor_schedules_agree():
    """The static (one wavefront per cell group), its two-wavefront register
    budget variant and the persistent work-queued predictor give the same bits
    and the same iteration statistics."""
    u0 = reactive_small(64, 16, seed=2)
    L = [64 / 1024, 16 / 1024]
    outs = []
    for sched in ("static", "static2", "queue"):
        os.environ["PYPDE_PRED_SCHEDULE"] = sched
        try:
            s = Solver(u0, L, systems.source("reactive2d"), order=3, boundaryTypes="reflective",
                       stiff=True, reflect=systems.REACTIVE_REFLECT)
        finally:
            del os.environ["PYPDE_PRED_SCHEDULE"]
        st = s.step(8)
        outs.append((s.state(), st))
        s.close()
    ua, sa = outs[0]
    for ub, sb in outs[1:]:
        assert np.array_equal(ua, ub)
        for k in ("steps", "pred_iters", "pred_cells", "pred_maxit", "pred_capped"):
            assert sa[k] == sb[k], k


@pytest.mark.gpu
@pytest.mark.parametrize("flux", ["roe", "osher"])
def test_roe_osher_match_oracle(flux):
    """Roe and Osher fluxes (fluxes.cpp:20-70): 1D Sod (C0) and a 2D periodic
    Euler field, GPU against the oracle."""
    code = {"roe": 1, "osher": 2}[flux]
    u0 = systems.sod1d()
    uo, _, st, rc = oracle_run("euler1d", u0, 0.2, [1.0], 0, 3, False, flux=code)
    ug, _, sg = gpu_run("euler1d", u0, 0.2, [1.0], "transitive", 3, False, flux=flux)
    print(f"{flux} sod: steps {sg['steps']} err {relerr(ug, uo):.3e}")
    assert rc == 0 and sg["rc"] == 0 and sg["steps"] == st.steps
    assert relerr(ug, uo) <= TOL
    u0 = systems.euler2d_modes(24, seed=4)
    uo, _, st, rc = oracle_run("euler2d", u0, 0.03, [1.0, 1.0], 1, 3, False, flux=code)
    ug, _, sg = gpu_run("euler2d", u0, 0.03, [1.0, 1.0], "periodic", 3, False, flux=flux)
    print(f"{flux} euler2d: steps {sg['steps']} err {relerr(ug, uo):.3e}")
    assert rc == 0 and sg["steps"] == st.steps
    assert relerr(ug, uo) <= TOL


@pytest.mark.gpu
def test_viscous_path_matches_oracle():
    """Second-order (viscous) F: face gradients from the predictor, viscous
    wave speeds in dt and in the flux (fv.cpp:87-122, system.cpp:53-61,
    fluxes.cpp:96-100); advection-diffusion in 1D and Navier-Stokes in 2D,
    GPU against the oracle (Rusanov; and Osher for Navier-Stokes)."""
    n = 64
    x = (np.arange(n) + 0.5) / n
    u0 = (1 + 0.5 * np.sin(2 * np.pi * x))[:, None]
    uo, _, st, rc = oracle_run("advdiff", u0, 0.2, [1.0], 1, 3, False, second=True)
    ug, _, sg = gpu_run("advdiff", u0, 0.2, [1.0], "periodic", 3, False)
    print(f"advdiff: steps {sg['steps']} err {relerr(ug, uo):.3e}")
    assert rc == 0 and sg["steps"] == st.steps and relerr(ug, uo) <= TOL
    u0 = systems.euler2d_modes(16, seed=2)
    for flux, code in (("rusanov", 0), ("osher", 2)):
        uo, _, st, rc = oracle_run("ns2d", u0, 0.02, [1.0, 1.0], 1, 3, False, second=True,
                                   flux=code)
        ug, _, sg = gpu_run("ns2d", u0, 0.02, [1.0, 1.0], "periodic", 3, False, flux=flux)
        print(f"ns2d {flux}: steps {sg['steps']} err {relerr(ug, uo):.3e}")
        assert rc == 0 and sg["steps"] == st.steps and relerr(ug, uo) <= TOL


# ------------------------------------------------ eigen routines, matrix by matrix

def _eig_cases(V, seed):
    """Matrices that reach every branch of the wave-speed routine: random,
    symmetric, Hessenberg, triangular, near-multiple clusters (as
    finite-difference Jacobians of multi-wave systems give), a lone real root
    beside a larger complex pair, a cluster next to the extreme root, Jordan
    blocks, zero, tiny, huge and NaN matrices."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(400):
        A = rng.standard_normal((V, V))
        if k % 4 == 1:
            A = A + A.T
        elif k % 4 == 2:
            A[np.tril_indices(V, -2)] = 0
        elif k % 4 == 3:
            A = np.triu(A)
        out.append(A)
    for k in range(200):
        lam = rng.uniform(-2, 2)
        D = np.diag(np.r_[np.full(V - 2, lam), rng.uniform(-4, 4, 2)])
        P = rng.standard_normal((V, V))
        out.append(P @ D @ np.linalg.inv(P) + 1e-12 * rng.standard_normal((V, V)))
    for k in range(100):
        re_, im_ = rng.uniform(-1, 1), rng.uniform(0.5, 3)
        B = np.diag(rng.uniform(-0.5, 0.5, V))
        B[V - 2:, V - 2:] = [[re_, -im_], [im_, re_]]
        P = rng.standard_normal((V, V))
        out.append(P @ B @ np.linalg.inv(P))
    for k in range(50):
        d = np.r_[np.full(V - 2, -0.9), -0.975, 0.5][:V]
        P = rng.standard_normal((V, V))
        out.append(P @ np.diag(d) @ np.linalg.inv(P) + 1e-12 * rng.standard_normal((V, V)))
    J = 2.0 * np.eye(V) + np.eye(V, k=1)
    out += [J, -J, np.zeros((V, V)), 1e-310 * rng.standard_normal((V, V)),
            1e300 * rng.standard_normal((V, V)), np.full((V, V), np.nan)]
    return np.ascontiguousarray(np.stack(out))


def _eig_ctx(name):
    """A small context whose kernel configuration is among the precompiled ones."""
    V = systems.SYSTEMS[name]["V"]
    if name == "gpr":
        return Solver(systems.gpr_riemann(16), [1.0], systems.source(name), order=2, stiff=True,
                      b_quirk=True)
    ndim = 1 if name == "euler1d" else 2
    return Solver(np.ones((8,) * ndim + (V,)), [1.0] * ndim, systems.source(name), order=3,
                  stiff=name == "reactive2d")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["euler1d", "euler2d", "reactive2d", "gpr"])
def test_max_abs_eig_batch_bitwise_equal_to_oracle(name):
    """The GPU max |lambda| equals the oracle's bit for bit on every case of
    _eig_cases, V = nvar of the system (3, 4, 5, 17)."""
    V = systems.SYSTEMS[name]["V"]
    sol = _eig_ctx(name)
    try:
        A = _eig_cases(V, 11 + V)
        n = A.shape[0] - (A.shape[0] & 1)
        A = np.ascontiguousarray(A[:n])
        want = np.array([cc.oracle().oracle_max_abs_eig(cc.dp(np.ascontiguousarray(a)), V)
                         for a in A])
        got = np.zeros(n)
        _capi.check(_capi.lib().pypde_eig_batch(sol.h, 0, _capi.dp(A), None, n,
                                                _capi.dp(got)), "eig_batch")
        assert np.array_equal(got, want, equal_nan=True), np.flatnonzero(
            ~((got == want) | (np.isnan(got) & np.isnan(want))))[:10]
        # and against LAPACK on the finite cases, as a sanity bound
        fin = np.isfinite(A).all(axis=(1, 2)) & (np.abs(A).max(axis=(1, 2)) < 1e200)
        ref = np.array([np.max(np.abs(np.linalg.eigvals(a))) for a in A[fin]])
        assert np.all(np.abs(want[fin] - ref) <= 1e-7 * ref + 1e-300)
    finally:
        sol.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["euler1d", "euler2d"])
def test_abs_apply_batch_bitwise_equal_to_oracle(name):
    """|A| x of the Roe/Osher fluxes on the GPU equals the oracle's bit for bit
    (finite cases of _eig_cases with real-diagonalisable, complex and
    defective spectra)."""
    V = systems.SYSTEMS[name]["V"]
    ndim = 1 if name == "euler1d" else 2
    u0 = np.ones((8,) * ndim + (V,))
    sol = Solver(u0, [1.0] * ndim, systems.source(name), order=3, stiff=False, flux="roe")
    try:
        A = _eig_cases(V, 21 + V)
        A = np.ascontiguousarray(A[np.isfinite(A).all(axis=(1, 2))
                                   & (np.abs(A).max(axis=(1, 2)) < 1e200)])
        n = A.shape[0]
        x = np.ascontiguousarray(np.random.default_rng(3).standard_normal((n, V)))
        want = np.zeros((n, V))
        for i in range(n):
            cc.oracle().oracle_abs_apply(cc.dp(np.ascontiguousarray(A[i])),
                                         cc.dp(np.ascontiguousarray(x[i])), V,
                                         cc.dp(want[i]))
        got = np.zeros((n, V))
        _capi.check(_capi.lib().pypde_eig_batch(sol.h, 2, _capi.dp(A), _capi.dp(x), n,
                                                _capi.dp(got)), "eig_batch")
        assert np.array_equal(got, want, equal_nan=True)
    finally:
        sol.close()


@pytest.mark.gpu
@pytest.mark.parametrize("stiff", [False, True])
def test_gpr_17_variables_bitwise_equal_to_oracle(stiff):
    """V = 17 with F, B and S (the reference's GPR model, pypde/tests/gpr/
    system.py:8-75): 1D Riemann problem, 200 cells, order 2, CFL 0.6, the
    reference's B.dQ evaluation, 30 steps, Picard and stiff Newton predictors.
    Tolerance: bitwise."""
    u0 = systems.gpr_riemann(200)
    uo, _, st, rc = cc.oracle_solve("gpr", u0, 1.0, [1.0], 0, 2, stiff=stiff, cfl=0.6, b_quirk=1,
                                    max_steps=30)
    assert rc == 0
    ug = u0.copy()
    stats = {}
    pde_solver(ug, 1.0, [1.0], system=systems.source("gpr"), order=2, ndt=1, stiff=stiff,
               cfl=0.6, b_quirk=True, max_steps=30, stats=stats)
    assert stats["steps"] == st.steps == 30
    assert np.array_equal(ug, uo)


@pytest.mark.gpu
def test_edge_cases_bitwise_equal_to_oracle():
    """Edge cases of the time loop (iterator.cpp:99-150), GPU against the
    oracle, bitwise: tf = 0 (no step; only the last snapshot row is written),
    ndt = 7 snapshot rows over 20 steps, one cell, orders 1 and 6, and a NaN
    in the initial state (the step it is reported at, as a negative return).
    Tolerance: bitwise, with equal step counts and return codes."""
    u0 = systems.sod1d(64)
    cases = [("tf0", u0, 0.0, 3, 3), ("ndt7", u0, 0.1, 3, 7), ("onecell", u0[:1], 0.05, 3, 1),
             ("order1", u0, 0.05, 1, 1), ("order6", u0, 0.05, 6, 1)]
    un = u0.copy()
    un[10, 0] = np.nan
    cases.append(("nan", un, 0.01, 3, 1))
    for name, u, tf, N, ndt in cases:
        uo, ro, st, rc = cc.oracle_solve("euler1d", u, tf, [1.0], 0, N, ndt=ndt)
        ug, rg, sg = gpu_run("euler1d", u, tf, [1.0], "transitive", N, False, ndt=ndt)
        assert sg["rc"] == rc and sg["steps"] == st.steps, name
        assert np.array_equal(ug, uo, equal_nan=True), name
        assert np.array_equal(rg, ro, equal_nan=True), name


def test_zero_cells_rejected():
    """A grid with no cells is an argument error, not a kernel launch."""
    with pytest.raises(_capi.PypdeError, match="code 1"):
        pde_solver(np.zeros((0, 3)), 0.1, [1.0], system=systems.source("euler1d"), order=3,
                   ndt=1, stiff=False)


@pytest.mark.gpu
def test_convergence_sweep_64_to_1024_order4():
    """The 2D periodic Euler configuration's convergence sweep (64 .. 1024
    cells per side, order 4): an advected density wave with exact cell
    averages, t = 0.1.  The 64-cell run equals the oracle's bit for bit.  At
    64 .. 512 cells the errors equal the reference's own (the compiled
    reference core on the CPU, tests/golden/convergence_order4.json; its
    512-cell run took 52 minutes on 6 threads) to 2 %: so the loss of
    accuracy at the wave's extrema from 512 cells on (L1 error 9.3e-10 at 256,
    1.2e-7 at 512, while near the zero crossings it keeps falling) is the
    reference method's -- the stencils' smoothness indicators there keep O(1)
    ratios however fine the grid, and the weights lambda / (p + 1e-14)^8
    (weno.cpp:26-59) turn that into a switch to one-sided stencils."""
    import json
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "convergence_order4.json")))
    errs, bands = [], []
    for n in (64, 128, 256, 512, 1024):
        u0 = systems.density_wave(n, 2)
        ug, _, sg = gpu_run("euler2d", u0, 0.1, [1.0, 1.0], "periodic", 4, False)
        assert sg["rc"] == 0 and np.isfinite(ug).all()
        if n == 64:
            uo, _, st, rc = oracle_run("euler2d", u0, 0.1, [1.0, 1.0], 1, 4, False)
            assert rc == 0 and np.array_equal(ug, uo)
        a, b, linf = cc.density_wave_errors(ug, n, 0.1)
        errs.append(a)
        bands.append(b)
        if str(n) in ref:
            r = ref[str(n)]
            print(f"n={n}: L1 {a:.4e} (reference {r['l1']:.4e}), zero-crossing band {b:.4e} "
                  f"(reference {r['l1_zero_crossings']:.4e}), Linf {linf:.4e} "
                  f"(reference {r['linf']:.4e})")
            assert abs(a - r["l1"]) <= 0.02 * r["l1"] + 1e-12, n
            assert abs(b - r["l1_zero_crossings"]) <= 0.02 * r["l1_zero_crossings"] + 1e-12, n
    orders = [np.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]
    border = [np.log2(bands[i] / bands[i + 1]) for i in range(len(bands) - 1)]
    print(f"order 4 sweep: L1 errors {errs}, orders {orders}; "
          f"zero-crossing bands {bands}, orders {border}")
    assert orders[0] >= 3.5 and orders[1] >= 2.8
    assert border[0] >= 3.0 and border[1] >= 2.5 and border[2] >= 2.0
    assert max(errs) <= 2e-7


@pytest.mark.gpu
def test_large_periodic_grid_is_shift_invariant():
    """Maximum sizes: a 8192 x 4096 periodic 2D Euler grid (33.5M cells; the
    face-trace array alone has 2.4e9 entries, beyond 32-bit indexing), two
    steps.  Each cell's arithmetic is independent of its position on a
    periodic grid and dt is an order-independent maximum, so rolling the
    initial state by (1234, 567) cells must roll the result bit for bit."""
    nx, ny, k = 8192, 4096, (1234, 567)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * (X + 2 * Y)) + 0.1 * np.cos(2 * np.pi * (3 * X - Y))
    u0 = np.ascontiguousarray(np.stack([rho, 0.5 * rho, 0.3 * rho,
                                        1.0 / 0.4 + 0.5 * rho * (0.25 + 0.09)], -1))
    del X, Y, rho
    res = []
    for u in (u0, np.ascontiguousarray(np.roll(u0, k, axis=(0, 1)))):
        ug, _, sg = gpu_run("euler2d", u, 1.0, [1.0, 0.5], "periodic", 3, False, max_steps=2)
        assert sg["rc"] == 0 and sg["steps"] == 2
        res.append(ug)
    assert np.array_equal(np.roll(res[0], k, axis=(0, 1)), res[1])
