"""Writes the golden vectors under tests/golden/ from the reference's own C++
core (oracle/_ref/libpypde_ref.so, compiled from the reference sources by
oracle/Makefile).  Run from the repository root:  python tests/golden/make_golden.py

tables.npz  basis/WENO/DG tables for N = 1..5 (poly/basis.cpp, weno_matrices.cpp,
            dg_matrices.cpp)
weno.npz    weno_solver (api.cpp:32-48) on seeded random data
solves.npz  pde_solver end states: C0 Sod (order 3; Rusanov, Roe and Osher
            fluxes) plus the reference's own sensitivity on each to 1-ulp input
            changes, 24x24 periodic 2D Euler, orders 2-4, the viscous
            path (advection-diffusion 1D, Navier-Stokes 2D) and the
            17-variable GPR Riemann problem (Picard and stiff)
            (python tests/golden/make_golden.py gpr refreshes only those,
            ... sodflux only the Roe / Osher Sod entries)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import cpu_checkers as cc  # noqa: E402
import test_oracle as to  # noqa: E402


def main():
    assert cc.ref_available(), "build oracle/_ref first (make -C oracle ref)"
    t = {}
    for n in [1, 2, 3, 4, 5]:
        for k, v in cc.ref_tables(n).items():
            t[f"{k}_{n}"] = v
    np.savez_compressed(os.path.join(HERE, "tables.npz"), **t)

    w = {}
    for ndim, N, V in [(1, 2, 3), (1, 3, 4), (2, 2, 3), (2, 3, 5), (2, 4, 4)]:
        rng = np.random.default_rng(100 + 10 * ndim + N)
        nX = np.array([12, 9][:ndim], dtype=np.int32)
        u = np.ascontiguousarray(rng.uniform(0.5, 2, tuple(nX) + (V,)))
        nout = int(np.prod(nX - 2 * (N - 1)))
        wh = np.zeros(nout * N ** ndim * V)
        cc.ref().weno_solver(cc.dp(wh), cc.dp(u.copy()), cc.ip(nX), ndim, N, V)
        w[f"u_{ndim}_{N}_{V}"] = u
        w[f"wh_{ndim}_{N}_{V}"] = wh
    np.savez_compressed(os.path.join(HERE, "weno.npz"), **w)

    s = {}
    u = to.sod1d()
    ur, _ = cc.ref_solve("euler1d", u, 0.2, [1.0], 0, 3, nthreads=4)
    # the reference's own sensitivity: one entry moved by an ulp, and every
    # entry moved by an ulp at random (two draws); the largest is kept
    noises = []
    u2 = u.copy()
    u2[37, 0] = np.nextafter(u2[37, 0], 2.0)
    ur2, _ = cc.ref_solve("euler1d", u2, 0.2, [1.0], 0, 3, nthreads=4)
    noises.append(to.relerr(ur2, ur))
    rng = np.random.default_rng(1)
    for _ in range(2):
        u2 = u.copy()
        m = rng.random(u.shape) < 0.5
        u2[m] = np.nextafter(u2[m], np.inf)
        ur2, _ = cc.ref_solve("euler1d", u2, 0.2, [1.0], 0, 3, nthreads=4)
        noises.append(to.relerr(ur2, ur))
    s["sod_final"] = ur
    s["sod_noise"] = np.array(max(noises))
    s.update(sod_flux_vectors())
    for N in [2, 3, 4]:
        ur, _ = cc.ref_solve("euler2d", to.euler2d_field(24, 24, noise=0.0), 0.05, [1, 1], 1, N,
                             nthreads=8)
        s[f"e2d_final_{N}"] = ur
    # viscous (second-order) path: advection-diffusion and 2D Navier-Stokes
    s["advdiff_final"] = to.ref_viscous("advdiff", to.advdiff_init(64), 0.2, [1.0], 1, 3)
    s["ns2d_final"] = to.ref_viscous("ns2d", to.systems.euler2d_modes(16, seed=2), 0.02,
                                     [1.0, 1.0], 1, 3)
    s.update(gpr_vectors())
    np.savez_compressed(os.path.join(HERE, "solves.npz"), **s)
    print("golden vectors written:", sorted(os.listdir(HERE)))


def sod_flux_vectors():
    """Roe (1) and Osher (2) fluxes on Sod, with the reference's own
    sensitivity to 1-ulp input changes (to.sod_flux_noise: single entries and
    all entries at random; the largest movement is kept)."""
    out = {}
    u = to.sod1d()
    for flux, name in ((1, "roe"), (2, "osher")):
        ur = to.ref_sod_flux(u, flux)
        out[f"sod_{name}_final"] = ur
        out[f"sod_{name}_noise"] = np.array(max(to.sod_flux_noise(flux, ur)))
    return out


def gpr_vectors():
    """17-variable GPR Riemann problem (1D, 200 cells, order 2, CFL 0.6) after
    44 steps, Picard and stiff predictors (the stiff run takes about a minute)."""
    u = to.systems.gpr_riemann(200)
    out = {}
    for stiff, name in ((False, "gpr_picard_final"), (True, "gpr_stiff_final")):
        ur, _ = cc.ref_solve("gpr", u, to.GPR_TF, [1.0], 0, 2, stiff=stiff, cfl=0.6, nthreads=8)
        out[name] = ur
    return out


if __name__ == "__main__":
    if sys.argv[1:] == ["sodflux"]:
        # refresh only the Roe / Osher Sod entries of solves.npz
        path = os.path.join(HERE, "solves.npz")
        s = dict(np.load(path, allow_pickle=False))
        s.update(sod_flux_vectors())
        np.savez_compressed(path, **s)
        print("Roe / Osher Sod golden vectors written")
    elif sys.argv[1:] == ["gpr"]:
        # add or refresh only the GPR entries of solves.npz
        path = os.path.join(HERE, "solves.npz")
        s = dict(np.load(path, allow_pickle=False))
        s.update(gpr_vectors())
        np.savez_compressed(path, **s)
        print("GPR golden vectors written")
    else:
        main()
