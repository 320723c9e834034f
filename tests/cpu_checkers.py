"""ctypes access to the CPU checkers (test infrastructure only).

- the oracle restatement ``oracle/_build/liboracle.so``;
- the reference's own C++ core ``oracle/_ref/libpypde_ref.so`` (present where it
  was built from the reference sources);
- C callbacks for each PDE system under ``pypde_amd/systems`` (``sys_<name>.so``).
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libpypde_ref.so")

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)
FTYPE = C.CFUNCTYPE(None, PD, PD, PD, C.c_int)
BTYPE = C.CFUNCTYPE(None, PD, PD, C.c_int)
STYPE = C.CFUNCTYPE(None, PD, PD)


def dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(PD)


def ip(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(PI)


class OracleOpts(C.Structure):
    _fields_ = [
        ("newton_tol", C.c_double),
        ("newton_maxit", C.c_int),
        ("max_steps", C.c_long),
        ("reflect", PD),
        ("guess", PD),
        ("guess_check", C.c_int),
        ("b_quirk", C.c_int),
    ]


class OracleStats(C.Structure):
    _fields_ = [
        ("steps", C.c_long),
        ("t", C.c_double),
        ("pred_iters", C.c_long),
        ("pred_cells", C.c_long),
        ("pred_maxit", C.c_long),
        ("pred_capped", C.c_long),
        ("guess_fallbacks", C.c_long),
    ]


_cache = {}


def oracle():
    if "oracle" not in _cache:
        lib = C.CDLL(ORACLE_SO)
        lib.oracle_max_abs_eig.restype = C.c_double
        lib.oracle_max_abs_eig.argtypes = [PD, C.c_int]
        lib.oracle_exp.restype = C.c_double
        lib.oracle_exp.argtypes = [C.c_double]
        lib.oracle_timestep.restype = C.c_double
        lib.oracle_timestep.argtypes = [C.c_void_p, C.c_void_p, PD, C.c_long, PD, C.c_int,
                                        C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_long]
        lib.oracle_predictor.argtypes = [C.c_void_p] * 4 + [PD, C.c_long, PD, C.c_int,
                                                            C.c_int, C.c_int, C.c_double,
                                                            C.c_int, C.c_void_p, PD,
                                                            C.c_void_p]
        lib.oracle_fv.argtypes = [C.c_void_p] * 3 + [PD, PD, PI, C.c_int, PD, C.c_int,
                                                    C.c_int, C.c_int, C.c_double, C.c_int]
        lib.oracle_pde_solver.argtypes = ([C.c_void_p] * 4 + [C.c_int] * 3 +
                                          [PD, C.c_double, PI, C.c_int, PD, C.c_double,
                                           PI, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_int, PD, C.c_void_p, C.c_void_p])
        lib.oracle_system_matrix.argtypes = [C.c_void_p, C.c_void_p, PD, PD, C.c_int, C.c_int,
                                             C.c_int, PD]
        lib.oracle_system_matrix.restype = None
        lib.oracle_boundaries.argtypes = [PD, PD, PI, C.c_int, PI, PD, C.c_int, C.c_int]
        lib.oracle_weno.argtypes = [PD, PD, PI, C.c_int, C.c_int, C.c_int]
        lib.oracle_dg_obj.argtypes = [C.c_void_p] * 3 + [PD, PD, PD, C.c_int, C.c_int,
                                                        C.c_int, C.c_double, C.c_int, PD]
        _cache["oracle"] = lib
    return _cache["oracle"]


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    if "ref" not in _cache:
        lib = C.CDLL(REF_SO)
        lib.ref_max_abs_eig.restype = C.c_double
        lib.ref_max_abs_eig.argtypes = [PD, C.c_int]
        lib.ref_timestep.restype = C.c_double
        lib.ref_timestep.argtypes = [C.c_void_p, C.c_void_p, PD, C.c_int, PD, C.c_int,
                                     C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_int, C.c_int]
        lib.ref_predictor.argtypes = [C.c_void_p] * 3 + [PD, C.c_int, PD, C.c_int, C.c_int,
                                                        C.c_int, C.c_double, C.c_int, PD]
        lib.ref_fv_apply.argtypes = [C.c_void_p] * 3 + [PD, PD, PI, C.c_int, PD, C.c_int,
                                                       C.c_int, C.c_int, C.c_double, C.c_int]
        lib.ref_dg_obj.argtypes = [C.c_void_p] * 3 + [PD, PD, PD, C.c_int, C.c_int, C.c_int,
                                                     C.c_double, PD]
        lib.pde_solver.argtypes = ([C.c_void_p] * 3 + [C.c_bool] * 3 +
                                   [PD, C.c_double, PI, C.c_int, PD, C.c_double, PI,
                                    C.c_bool, C.c_int, C.c_int, C.c_int, C.c_int, C.c_bool,
                                    PD, C.c_int])
        lib.pde_solver.restype = None
        lib.weno_solver.argtypes = [PD, PD, PI, C.c_int, C.c_int, C.c_int]
        lib.weno_solver.restype = None
        _cache["ref"] = lib
    return _cache["ref"]


class System:
    """C callbacks of one PDE system source (pypde_amd/systems/<name>.pde)."""

    def __init__(self, name):
        path = os.path.join(ROOT, "oracle", "_build", f"sys_{name}.so")
        self.lib = C.CDLL(path)
        self.name = name

        def sym(n):
            try:
                return C.cast(getattr(self.lib, n), C.c_void_p).value
            except AttributeError:
                return None

        self.F, self.B, self.S, self.M = sym("sys_F"), sym("sys_B"), sym("sys_S"), sym("sys_M")
        self.lib.sys_exp.restype = C.c_double
        self.lib.sys_exp.argtypes = [C.c_double]
        self.lib.sys_cbrt.restype = C.c_double
        self.lib.sys_cbrt.argtypes = [C.c_double]


def system(name):
    key = "sys:" + name
    if key not in _cache:
        _cache[key] = System(name)
    return _cache[key]


def tables(N):
    n = N
    out = dict(nodes=np.zeros(n), wghts=np.zeros(n), endvals=np.zeros(2 * n),
               dervals=np.zeros(n * n), weno_mat=np.zeros(4 * n * n),
               weno_inv=np.zeros(4 * n * n), sig=np.zeros(n * n), dgmat=np.zeros(n * n),
               dginv=np.zeros(n * n), fl_c=np.zeros(n), fl_m=np.zeros(n * n * n))
    rc = oracle().oracle_tables(C.c_int(n), *[dp(out[k]) for k in
                                             ["nodes", "wghts", "endvals", "dervals",
                                              "weno_mat", "weno_inv", "sig", "dgmat",
                                              "dginv", "fl_c", "fl_m"]])
    assert rc == 0
    return out


def ref_tables(N):
    n = N
    out = dict(nodes=np.zeros(n), wghts=np.zeros(n), endvals=np.zeros(2 * n),
               dervals=np.zeros(n * n), weno_mat=np.zeros(4 * n * n), sig=np.zeros(n * n),
               dgmat=np.zeros(n * n))
    ref().ref_tables(C.c_int(n), *[dp(out[k]) for k in
                                  ["nodes", "wghts", "endvals", "dervals", "weno_mat", "sig",
                                   "dgmat"]])
    return out


def oracle_solve(sysname, u0, tf, L, bc, N, stiff=False, cfl=0.9, ndt=1, max_steps=0,
                 reflect=None, newton_tol=1e-12, newton_maxit=100, useF=True, useB=None,
                 useS=None, b_quirk=0, flux=0, second=False):
    """Run oracle_pde_solver on a copy of u0 (shape nX + (V,)); returns (final, ret, stats)."""
    s = system(sysname)
    u = np.ascontiguousarray(u0, dtype=np.float64).copy()
    nX = np.array(u.shape[:-1], dtype=np.int32)
    ndim, V = len(nX), u.shape[-1]
    dX = np.array([L[i] / nX[i] for i in range(ndim)], dtype=np.float64)
    bcs = np.array(bc if isinstance(bc, (list, tuple)) else [bc] * ndim, dtype=np.int32)
    ret = np.zeros(ndt * u.size)
    opts = OracleOpts()
    opts.newton_tol = newton_tol
    opts.newton_maxit = newton_maxit
    opts.max_steps = max_steps
    opts.b_quirk = b_quirk
    refl = None
    if reflect is not None:
        refl = np.ascontiguousarray(reflect, dtype=np.float64)
        opts.reflect = dp(refl)
    st = OracleStats()
    useB = (s.B is not None) if useB is None else useB
    useS = (s.S is not None) if useS is None else useS
    rc = oracle().oracle_pde_solver(s.F, s.B, s.S, s.M, int(useF and s.F is not None),
                                    int(useB), int(useS), dp(u), tf, ip(nX), ndim, dp(dX),
                                    cfl, ip(bcs), int(stiff), int(flux), N, V, ndt, int(second),
                                    dp(ret),
                                    C.byref(opts), C.byref(st))
    return u, ret.reshape((ndt,) + u.shape), st, rc


def ref_solve(sysname, u0, tf, L, bc, N, stiff=False, cfl=0.9, ndt=1, nthreads=1, flux=0):
    """Run the reference's pde_solver (stdout of its step loop suppressed)."""
    s = system(sysname)
    u = np.ascontiguousarray(u0, dtype=np.float64).copy()
    nX = np.array(u.shape[:-1], dtype=np.int32)
    ndim, V = len(nX), u.shape[-1]
    dX = np.array([L[i] / nX[i] for i in range(ndim)], dtype=np.float64)
    bcs = np.array(bc if isinstance(bc, (list, tuple)) else [bc] * ndim, dtype=np.int32)
    ret = np.zeros(ndt * u.size)
    with quiet_stdout():
        ref().pde_solver(s.F, s.B, s.S, s.F is not None, s.B is not None, s.S is not None,
                         dp(u), tf, ip(nX), ndim, dp(dX), cfl, ip(bcs), bool(stiff), int(flux), N, V,
                         ndt, False, dp(ret), nthreads)
    return u, ret.reshape((ndt,) + u.shape)


class quiet_stdout:
    """Redirect fd 1 to /dev/null (the reference prints every step)."""

    def __enter__(self):
        import sys
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)

    def __exit__(self, *a):
        C.CDLL(None).fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.null)
        os.close(self.saved)


def density_wave_errors(u, n, t):
    """Density errors of the advected wave (systems.density_wave) at time t:
    L1 over the grid, L1 over the bands of phase within 0.05 of the zero
    crossings (farthest from the extrema), and L_inf."""
    from pypde_amd import systems
    ex = systems.density_wave(n, 2, t=t)
    e = np.abs(u[..., 0] - ex[..., 0])
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    ph = (X + Y - 2 * 0.5 * t) % 1.0        # rho - 1 = 0.2 sin(2 pi ph)
    band = (np.abs(ph - 0.5) < 0.05) | (ph < 0.05) | (ph > 0.95)
    return float(e.mean()), float(e[band].mean()), float(e.max())
