"""ctypes binding of the C ABI in include/pypde_hip.h (libpypde_hip.so).

The HIP library is the only compute path: if it is not built, every call
raises; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libpypde_hip.so")

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)


class Options(C.Structure):
    _fields_ = [("struct_size", C.c_int), ("newton_tol", C.c_double),
                ("newton_maxit", C.c_int), ("max_steps", C.c_long), ("reflect", PD),
                ("b_quirk", C.c_int), ("use_M", C.c_int), ("device", C.c_int),
                ("profile", C.c_int), ("guess", PD), ("use_graph", C.c_int),
                ("mixed", C.c_int), ("mixed_tol", C.c_double), ("mixed_maxit", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("struct_size", C.c_int), ("steps", C.c_long), ("t", C.c_double),
                ("pred_iters", C.c_long), ("pred_cells", C.c_long), ("pred_maxit", C.c_long),
                ("pred_capped", C.c_long), ("guess_fallbacks", C.c_long), ("nan_step", C.c_int),
                ("ms_loop", C.c_double), ("ms_kernel", C.c_double * 8),
                ("pred_node_full", C.c_long), ("pred_node_short", C.c_long),
                ("pred_iters_f32", C.c_long)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("struct_size", "ms_kernel")}
        d["ms_kernel"] = dict(zip(["weno", "dt", "guess", "predictor", "flux0", "flux1",
                                   "update", "advance"], list(self.ms_kernel)))
        return d


class PypdeError(RuntimeError):
    pass


_lib = None


def reload():
    global _lib
    _lib = None
    return lib()


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PypdeError(f"{LIB_PATH} is not built: run `python -m pypde_amd.build` "
                             "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        L.pypde_hip_last_error.restype = C.c_char_p
        L.pypde_default_options.argtypes = [C.POINTER(Options)]
        L.pde_solver_hip.argtypes = [C.c_char_p, C.c_bool, C.c_bool, C.c_bool, PD, C.c_double,
                                     PI, C.c_int, PD, C.c_double, PI, C.c_bool, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_bool, PD, C.POINTER(Options),
                                     C.POINTER(Stats)]
        L.weno_solver.argtypes = [PD, PD, PI, C.c_int, C.c_int, C.c_int]
        L.weno_solver_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, PI, C.c_void_p]
        L.pypde_get_tuning.argtypes = [C.c_void_p, PI, PD]
        L.pypde_create.restype = C.c_void_p
        L.pypde_create.argtypes = [C.c_char_p, C.c_bool, C.c_bool, C.c_bool, PI, C.c_int, PD,
                                   C.c_double, PI, C.c_bool, C.c_int, C.c_int, C.c_int, C.c_bool,
                                   C.POINTER(Options), PI]
        L.pypde_set_state.argtypes = [C.c_void_p, PD]
        L.pypde_get_state.argtypes = [C.c_void_p, PD]
        L.pypde_run.argtypes = [C.c_void_p, C.c_double, C.c_long, C.c_int, PD, C.POINTER(Stats)]
        L.pypde_step.argtypes = [C.c_void_p, C.c_long, C.c_double, C.POINTER(Stats)]
        L.pypde_device_state.restype = C.c_void_p
        L.pypde_device_state.argtypes = [C.c_void_p]
        L.pypde_stream.restype = C.c_void_p
        L.pypde_stream.argtypes = [C.c_void_p]
        L.pypde_destroy.argtypes = [C.c_void_p]
        L.pypde_hip_compile.argtypes = [C.c_char_p, C.c_bool, C.c_bool, C.c_bool, C.c_int,
                                        C.c_int, C.c_int, C.c_bool, C.c_int, C.c_bool,
                                        C.POINTER(Options)]
        L.pypde_hip_compile_ensemble.argtypes = L.pypde_hip_compile.argtypes
        L.pypde_hip_tables.argtypes = [C.c_int] + [PD] * 10
        L.pypde_ensemble_create.restype = C.c_void_p
        L.pypde_ensemble_create.argtypes = [C.c_int, C.c_char_p, C.c_bool, C.c_bool, C.c_bool, PI,
                                            C.c_int, PD, C.c_double, PI, C.c_bool, C.c_int,
                                            C.c_int, C.c_int, C.c_bool, C.POINTER(Options), PI]
        L.pypde_ensemble_set_state.argtypes = [C.c_void_p, C.c_int, PD]
        L.pypde_ensemble_get_state.argtypes = [C.c_void_p, C.c_int, PD]
        L.pypde_ensemble_run.argtypes = [C.c_void_p, C.c_double, C.c_long, C.POINTER(Stats), PD]
        L.pypde_ensemble_destroy.argtypes = [C.c_void_p]
        L.pypde_eig_batch.argtypes = [C.c_void_p, C.c_int, PD, PD, C.c_long, PD]
        _lib = L
    return _lib


def last_error():
    return lib().pypde_hip_last_error().decode()


def check(rc, what):
    if rc > 0:
        raise PypdeError(f"{what} failed (code {rc}): {last_error()}")
    return rc


def dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(PD)


def ip(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(PI)


def make_options(newton_tol=1e-12, newton_maxit=100, max_steps=0, reflect=None, b_quirk=0,
                 use_M=False, device=0, profile=False, guess=None, use_graph=False, mixed=False,
                 mixed_tol=None, mixed_maxit=None):
    o = Options()
    lib().pypde_default_options(C.byref(o))
    o.mixed = int(bool(mixed))
    if mixed_tol is not None:
        o.mixed_tol = mixed_tol
    if mixed_maxit is not None:
        o.mixed_maxit = mixed_maxit
    o.newton_tol = newton_tol
    o.newton_maxit = newton_maxit
    o.max_steps = max_steps
    o.b_quirk = b_quirk
    o.use_M = int(bool(use_M))
    o.device = device
    o.profile = int(bool(profile))
    o.use_graph = int(bool(use_graph))
    keep = []
    if reflect is not None:
        r = np.ascontiguousarray(reflect, dtype=np.float64).ravel()
        keep.append(r)
        o.reflect = dp(r)
    if guess is not None:
        g = np.ascontiguousarray(guess, dtype=np.float64).ravel()
        keep.append(g)
        o.guess = dp(g)
    o._keep = keep
    return o


def compile_only(src, useF, useB, useS, ndim, N, V, stiff=False, use_M=False, b_quirk=0, flux=0,
                 second=False, mixed=False, ensemble=False):
    """Compile (or fetch from the code-object cache) without a GPU; ensemble =
    True gives the code object of ensemble members (with the batched kernels)."""
    o = make_options(use_M=use_M, b_quirk=b_quirk, mixed=mixed)
    fn = lib().pypde_hip_compile_ensemble if ensemble else lib().pypde_hip_compile
    rc = fn(src.encode(), useF, useB, useS, ndim, N, V, stiff, int(flux), bool(second), C.byref(o))
    check(rc, f"hipRTC compile (ndim={ndim}, N={N}, V={V})")


def tables(N):
    n = N
    out = dict(nodes=np.zeros(n), wghts=np.zeros(n), endvals=np.zeros(2 * n),
               dervals=np.zeros(n * n), weno_inv=np.zeros(4 * n * n), sig=np.zeros(n * n),
               dgmat=np.zeros(n * n), dginv=np.zeros(n * n), fl_c=np.zeros(n),
               fl_m=np.zeros(n * n * n))
    check(lib().pypde_hip_tables(n, *[dp(out[k]) for k in
                                      ["nodes", "wghts", "endvals", "dervals", "weno_inv", "sig",
                                       "dgmat", "dginv", "fl_c", "fl_m"]]), "tables")
    return out
