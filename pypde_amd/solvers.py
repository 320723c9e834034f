"""Python front end with the reference's call surface (pypde/solvers.py:13-244).

``pde_solver(Q0, tf, L, F, B, S, ...)`` keeps the reference's arguments,
defaults, output shape ``(ndt,) + Q0.shape`` and in-place update of a
C-contiguous float64 ``Q0``.  The physics are HIP device source instead of
numba-compiled Python callables (the reference JIT-compiles them with numba,
pypde/cfuncs.py:30-75; here hipRTC compiles them into the kernels):

    F = '''PDE_FN void F(double *out, const double *Q, const double *dQ, int d) {...}'''
    B = '''PDE_FN void B(double *out, const double *Q, int d) {...}'''   # V x V row-major
    S = '''PDE_FN void S(double *out, const double *Q) {...}'''
    M = '''PDE_FN void M(double *out, const double *Q) {...}'''          # dS/dQ (optional)

or ``system=`` one source defining several of them (see pypde_amd/systems).
"""
import ctypes as C

import numpy as np

from pypde_amd import _capi

FLUXES = {"rusanov": 0, "roe": 1, "osher": 2}
BOUNDARIES = {"transitive": 0, "periodic": 1, "reflective": 2}


def parse_boundary_types(boundaryTypes, ndim):
    """As pypde/utils.py:37-66, plus 'reflective'; raises instead of sys.exit."""
    err = ("boundaryTypes must be a string from {'transitive', 'periodic', 'reflective'} "
           "or a list of such strings, of length equal to the number of dimensions")
    if isinstance(boundaryTypes, str):
        if boundaryTypes not in BOUNDARIES:
            raise ValueError(err)
        ret = [BOUNDARIES[boundaryTypes]] * ndim
    elif isinstance(boundaryTypes, (list, tuple)):
        if len(boundaryTypes) != ndim or any(b not in BOUNDARIES for b in boundaryTypes):
            raise ValueError(err)
        ret = [BOUNDARIES[b] for b in boundaryTypes]
    else:
        raise ValueError(err)
    return np.array(ret, dtype=np.int32)


def _source(F, B, S, M, system):
    if system is not None:
        has = {k: f"#define PDE_HAS_{k}" in system for k in "FBSM"}
        return system, has["F"], has["B"], has["S"], has["M"]
    parts = [x for x in (F, B, S, M) if x]
    return "\n".join(parts), F is not None, B is not None, S is not None, M is not None


def _second_order(src, useF):
    """The reference picks its viscous path when F takes the gradient
    (three-argument F, pypde/cfuncs.py:11-27, solvers.py:196); a system
    source says so with ``#define PDE_SECOND_ORDER 1``."""
    return bool(useF and "#define PDE_SECOND_ORDER 1" in src)


def pde_solver(Q0, tf, L, F=None, B=None, S=None, boundaryTypes="transitive", cfl=0.9,
               order=2, ndt=100, flux="rusanov", stiff=True, nThreads=-1, M=None,
               system=None, reflect=None, newton_tol=1e-12, newton_maxit=100, max_steps=0,
               b_quirk=False, device=0, stats=None, mixed=False):
    """Solve dQ/dt + div F(Q) + B(Q).grad Q = S(Q) on [0, L_1] x ... x [0, L_n]
    with the ADER-WENO method on one MI355X.  Returns ``(ndt,) + Q0.shape``.

    Extra keywords (not in the reference): ``M`` / ``system`` (above),
    ``reflect`` (ndim x V multipliers for 'reflective' walls), ``newton_tol`` /
    ``newton_maxit`` (stiff predictor), ``max_steps`` (fixed step window),
    ``b_quirk`` (evaluate B.dQ as the reference's build does), ``stats`` (dict
    filled with run statistics), ``mixed`` (an FP32 predictor phase before the
    FP64 one).  ``nThreads`` is accepted and ignored."""
    Q0 = np.asarray(Q0)
    nX = np.array(Q0.shape[:-1], dtype=np.int32)
    ndim = len(nX)
    V = Q0.shape[-1]
    dX = np.array([L[i] / nX[i] for i in range(len(L))], dtype=np.float64)
    bcs = parse_boundary_types(boundaryTypes, ndim)
    if flux not in FLUXES:
        raise ValueError(f"flux must be one of {sorted(FLUXES)}")
    src, useF, useB, useS, useM = _source(F, B, S, M, system)
    secondOrder = _second_order(src, useF)
    inplace = Q0.dtype == np.float64 and Q0.flags.c_contiguous
    u = Q0.ravel() if inplace else np.ascontiguousarray(Q0, dtype=np.float64).ravel().copy()
    ret = np.zeros(ndt * u.size)
    if reflect is None and 2 in bcs:
        reflect = np.ones((ndim, V))
    opts = _capi.make_options(newton_tol=newton_tol, newton_maxit=newton_maxit,
                              max_steps=max_steps, reflect=reflect, b_quirk=int(b_quirk),
                              use_M=useM, device=device, mixed=mixed)
    st = _capi.Stats()
    st.struct_size = C.sizeof(_capi.Stats)
    rc = _capi.lib().pde_solver_hip(src.encode(), useF, useB, useS, _capi.dp(u), float(tf),
                                    _capi.ip(nX), ndim, _capi.dp(dX), float(cfl), _capi.ip(bcs),
                                    bool(stiff), FLUXES[flux], int(order), V, int(ndt),
                                    secondOrder, _capi.dp(ret), C.byref(opts), C.byref(st))
    _capi.check(rc, "pde_solver_hip")
    if stats is not None:
        stats.update(st.as_dict())
        stats["rc"] = rc
    if rc < 0:
        print("NaNs found")
    return ret.reshape((ndt,) + Q0.shape)


def weno_solver(u, order=2):
    """WENO reconstruction without ghost cells (pypde/solvers.py:217-244)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    nX = np.array(u.shape[:-1], dtype=np.int32)
    ndim = len(nX)
    V = u.shape[-1]
    nXret = nX - 2 * (order - 1)
    ret = np.zeros(int(np.prod(nXret)) * order ** ndim * V)
    rc = _capi.lib().weno_solver(_capi.dp(ret), _capi.dp(u), _capi.ip(nX), ndim, order, V)
    _capi.check(rc, "weno_solver")
    return ret.reshape(tuple(nXret) + (order,) * ndim + (V,))


class Solver:
    """Resident solver: compile once, keep the grid in HBM, step repeatedly."""

    def __init__(self, u0, L, system, ndim=None, order=3, boundaryTypes="transitive", cfl=0.9,
                 stiff=False, reflect=None, newton_tol=1e-12, newton_maxit=100, b_quirk=False,
                 profile=False, device=0, guess=None, use_graph=False, flux="rusanov",
                 mixed=False, mixed_tol=None, mixed_maxit=None):
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        self.shape = u0.shape
        self.nX = np.array(u0.shape[:-1], dtype=np.int32)
        nd = len(self.nX)
        V = u0.shape[-1]
        self.dX = np.array([L[i] / self.nX[i] for i in range(nd)], dtype=np.float64)
        self.bcs = parse_boundary_types(boundaryTypes, nd)
        src, useF, useB, useS, useM = _source(None, None, None, None, system)
        if reflect is None and 2 in self.bcs:
            reflect = np.ones((nd, V))
        self.opts = _capi.make_options(newton_tol=newton_tol, newton_maxit=newton_maxit,
                                       reflect=reflect, b_quirk=int(b_quirk), use_M=useM,
                                       profile=profile, device=device, guess=guess,
                                       use_graph=use_graph, mixed=mixed, mixed_tol=mixed_tol,
                                       mixed_maxit=mixed_maxit)
        rc = C.c_int(0)
        self.h = _capi.lib().pypde_create(src.encode(), useF, useB, useS, _capi.ip(self.nX), nd,
                                          _capi.dp(self.dX), float(cfl), _capi.ip(self.bcs),
                                          bool(stiff), FLUXES[flux], int(order), V,
                                          _second_order(src, useF),
                                          C.byref(self.opts), C.byref(rc))
        if not self.h:
            _capi.check(rc.value or 3, "pypde_create")
        self.set_state(u0)

    def set_state(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        _capi.check(_capi.lib().pypde_set_state(self.h, _capi.dp(u)), "set_state")

    def state(self):
        out = np.zeros(self.shape)
        _capi.check(_capi.lib().pypde_get_state(self.h, _capi.dp(out)), "get_state")
        return out

    def run(self, tf, max_steps=0, ndt=0):
        st = _capi.Stats()
        st.struct_size = C.sizeof(_capi.Stats)
        ret = np.zeros((max(ndt, 1),) + self.shape) if ndt else None
        rc = _capi.lib().pypde_run(self.h, float(tf), int(max_steps), int(ndt),
                                   _capi.dp(ret) if ret is not None else None, C.byref(st))
        _capi.check(rc, "run")
        d = st.as_dict()
        d["rc"] = rc
        return (ret, d) if ndt else d

    def step(self, steps, tf=1e300):
        st = _capi.Stats()
        st.struct_size = C.sizeof(_capi.Stats)
        rc = _capi.lib().pypde_step(self.h, int(steps), float(tf), C.byref(st))
        _capi.check(rc, "step")
        d = st.as_dict()
        d["rc"] = rc
        return d

    def tuning(self):
        """The autotuner's launch choices for this problem (see pypde_get_tuning)."""
        ch = np.zeros(5, dtype=np.int32)
        ms = np.zeros(4)
        _capi.check(_capi.lib().pypde_get_tuning(self.h, _capi.ip(ch), _capi.dp(ms)), "tuning")
        weno = ["8x32", "4x64", "16x16", "16x32"][ch[0]]
        flux = [256, 128, 64][ch[1]]
        upd = ["cell", "rows1", "rows2", "rows4", "rows8"][ch[2]]
        pred = ["static", "queue", "split", "split3"][ch[3]]
        src = ["defaults", "measured", "table", "forced"][ch[4]]
        return dict(weno_tile=weno, flux_workgroup=flux, update=upd, predictor=pred, source=src,
                    ms=dict(zip(["weno", "predictor", "flux", "update"], ms.tolist())))

    def close(self):
        if getattr(self, "h", None):
            _capi.lib().pypde_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Independent grids of one system advanced together: every timestep of
    every member is one HIP graph launch (members side by side on their own
    streams), each member with its own dt.  ``grids``: list of initial states,
    each of shape (n_x, [n_y,] V); ``L``: list of domain lengths per member."""

    def __init__(self, grids, L, system, order=3, boundaryTypes="transitive", cfl=0.9,
                 stiff=False, reflect=None, newton_tol=1e-12, newton_maxit=100, device=0,
                 flux="rusanov"):
        grids = [np.ascontiguousarray(g, dtype=np.float64) for g in grids]
        self.shapes = [g.shape for g in grids]
        nd = len(self.shapes[0]) - 1
        V = self.shapes[0][-1]
        if any(len(s) - 1 != nd or s[-1] != V for s in self.shapes):
            raise ValueError("ensemble members must share ndim and nvar")
        nX = np.array([s[:-1] for s in self.shapes], dtype=np.int32).ravel()
        dX = np.array([[L[i][d] / self.shapes[i][d] for d in range(nd)]
                       for i in range(len(grids))], dtype=np.float64).ravel()
        bcs = parse_boundary_types(boundaryTypes, nd)
        src, useF, useB, useS, useM = _source(None, None, None, None, system)
        if reflect is None and 2 in bcs:
            reflect = np.ones((nd, V))
        self.opts = _capi.make_options(newton_tol=newton_tol, newton_maxit=newton_maxit,
                                       reflect=reflect, use_M=useM, device=device)
        rc = C.c_int(0)
        self.h = _capi.lib().pypde_ensemble_create(len(grids), src.encode(), useF, useB, useS,
                                                   _capi.ip(nX), nd, _capi.dp(dX), float(cfl),
                                                   _capi.ip(bcs), bool(stiff), FLUXES[flux],
                                                   int(order), V, _second_order(src, useF),
                                                   C.byref(self.opts), C.byref(rc))
        if not self.h:
            _capi.check(rc.value or 3, "pypde_ensemble_create")
        for i, g in enumerate(grids):
            _capi.check(_capi.lib().pypde_ensemble_set_state(self.h, i, _capi.dp(g)), "set_state")

    def run(self, tf, max_steps=0):
        n = len(self.shapes)
        st = (_capi.Stats * n)()
        for s in st:
            s.struct_size = C.sizeof(_capi.Stats)
        ms = np.zeros(1)
        rc = _capi.lib().pypde_ensemble_run(self.h, float(tf), int(max_steps), st, _capi.dp(ms))
        _capi.check(rc, "ensemble_run")
        out = [s.as_dict() for s in st]
        return out, float(ms[0])

    def state(self, i):
        out = np.zeros(self.shapes[i])
        _capi.check(_capi.lib().pypde_ensemble_get_state(self.h, i, _capi.dp(out)), "get_state")
        return out

    def close(self):
        if getattr(self, "h", None):
            _capi.lib().pypde_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
