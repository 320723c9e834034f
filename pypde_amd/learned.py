"""Learned initial guess for the stiff predictor's Newton iteration.

The reference starts every cell's space-time solve from its reconstruction held
constant in time (DGSolver::initial_guess, dg.cpp:12-20, called at dg.cpp:216).
Here a small MLP maps a cell's reconstruction and dt to a guess for the whole
space-time tensor; Newton still iterates to the unchanged tolerance and drops a
guess whose residual grows (see the predictor kernel).

Network: x = (wh (N^D V values), dt), normalised per feature ->
128 SiLU -> 128 SiLU -> y, y de-normalised = (qh - w) / dt  (N^(D+1) V values).
Training pairs come from the solver itself (pypde_harvest_step); forward and
backward run on this package's matrix-core kernels (BF16 operands, FP32
accumulation) with Adam on FP32 master weights.  Device memory comes from the
same HIP library (no second runtime in the process).
"""
import ctypes as C

import numpy as np

from pypde_amd import _capi


def _bind():
    L = _capi.lib()
    if getattr(L, "_mlp_bound", False):
        return L
    vp, i, f, l = C.c_void_p, C.c_int, C.c_float, C.c_long
    L.pypde_mlp_gemm.argtypes = [vp, i, i, vp, i, i, vp, i, i, i, i, f, f, vp]
    L.pypde_mlp_bias_act.argtypes = [vp, vp, vp, i, i, i, vp]
    L.pypde_mlp_act_grad.argtypes = [vp, vp, vp, l, vp]
    L.pypde_mlp_rowsum.argtypes = [vp, vp, i, i, vp]
    L.pypde_mlp_mse.argtypes = [vp, vp, vp, vp, i, i, vp]
    L.pypde_mlp_adam.argtypes = [vp, vp, vp, vp, l, f, f, f, f, i, vp]
    L.pypde_harvest_step.argtypes = [vp, vp, vp, l, l, C.POINTER(_capi.Stats)]
    L.pypde_set_guess_model.argtypes = [vp, vp]
    L.pypde_last_guess.argtypes = [vp, _capi.PD]
    L.pypde_geometry.argtypes = [vp, C.POINTER(C.c_long), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.pypde_dev_alloc.restype = vp
    L.pypde_dev_alloc.argtypes = [C.c_size_t]
    L.pypde_dev_free.argtypes = [vp]
    L.pypde_memcpy.argtypes = [vp, vp, C.c_size_t, i]
    L.pypde_dev_memset.argtypes = [vp, i, C.c_size_t]
    L._mlp_bound = True
    return L


class Dev:
    """A float32 device array (row-major 2D) owned by this object, or a view."""

    def __init__(self, rows, cols, ptr=None, owner=None):
        self.shape = (int(rows), int(cols))
        self.nbytes = 4 * self.shape[0] * self.shape[1]
        if ptr is None:
            p = _bind().pypde_dev_alloc(self.nbytes)
            if not p:
                raise _capi.PypdeError("device allocation failed: " + _capi.last_error())
            self.ptr, self.owned = p, True
        else:
            self.ptr, self.owned = ptr, False
        self.owner = owner

    @staticmethod
    def from_numpy(a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.ndim == 1:
            a = a[:, None]
        d = Dev(*a.shape)
        _capi.check(_bind().pypde_memcpy(d.ptr, a.ctypes.data, d.nbytes, 1), "memcpy")
        return d

    @staticmethod
    def zeros(rows, cols):
        d = Dev(rows, cols)
        _capi.check(_bind().pypde_dev_memset(d.ptr, 0, d.nbytes), "memset")
        return d

    def numpy(self):
        out = np.empty(self.shape, dtype=np.float32)
        _capi.check(_bind().pypde_memcpy(out.ctypes.data, self.ptr, self.nbytes, 2), "memcpy")
        return out

    def view_rows(self, r0, r1):
        r0, r1 = int(r0), int(r1)   # plain ints: ctypes rejects numpy integers as pointers
        return Dev(r1 - r0, self.shape[1], ptr=int(self.ptr) + 4 * r0 * self.shape[1], owner=self)

    def __del__(self):
        if getattr(self, "owned", False) and self.ptr:
            try:
                _bind().pypde_dev_free(self.ptr)
            except Exception:
                pass
            self.ptr = None

    @property
    def n(self):
        return self.shape[0] * self.shape[1]


def gemm(A, B, Cm, transA=False, transB=False, alpha=1.0, beta=0.0):
    """Cm = alpha op(A) op(B) + beta Cm on the matrix cores."""
    M, N = Cm.shape
    K = A.shape[0] if transA else A.shape[1]
    rc = _bind().pypde_mlp_gemm(A.ptr, A.shape[1], int(transA), B.ptr, B.shape[1], int(transB),
                                Cm.ptr, Cm.shape[1], M, N, K, alpha, beta, None)
    _capi.check(rc, "gemm")


class PypdeMLP(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ["W1t", "b1", "W2t", "b2", "W3t", "b3", "in_mean",
                                          "in_rstd", "out_mean", "out_std"]]


PARAMS = ["W1t", "b1", "W2t", "b2", "W3t", "b3"]
NORMS = ["in_mean", "in_rstd", "out_mean", "out_std"]


class GuessMLP:
    """3-layer SiLU MLP trained and evaluated with the package's HIP kernels."""

    def __init__(self, n_in, n_out, hidden=128, seed=0):
        self.n_in, self.n_out, self.hidden = n_in, n_out, hidden
        g = np.random.default_rng(seed)

        def init(o, i, s=1.0):
            return g.uniform(-1, 1, (o, i)) * np.sqrt(3.0 / i) * s

        host = {"W1t": init(hidden, n_in), "b1": np.zeros((hidden, 1)),
                "W2t": init(hidden, hidden), "b2": np.zeros((hidden, 1)),
                "W3t": init(n_out, hidden, 0.1), "b3": np.zeros((n_out, 1))}
        self.p = {k: Dev.from_numpy(v) for k, v in host.items()}
        self.m = {k: Dev.zeros(*v.shape) for k, v in self.p.items()}
        self.v = {k: Dev.zeros(*v.shape) for k, v in self.p.items()}
        self.norm_host = {"in_mean": np.zeros(n_in), "in_rstd": np.ones(n_in),
                          "out_mean": np.zeros(n_out), "out_std": np.ones(n_out)}
        self.norm = {k: Dev.from_numpy(v) for k, v in self.norm_host.items()}
        self.step_count = 0
        self._buf = {}

    # ---- normalisation
    def fit_normalisation(self, X, Y):
        """X, Y host arrays (features x samples)."""
        self.norm_host = {"in_mean": X.mean(1), "in_rstd": 1.0 / np.maximum(X.std(1), 1e-12),
                          "out_mean": Y.mean(1), "out_std": np.maximum(Y.std(1), 1e-12)}
        self.norm = {k: Dev.from_numpy(v) for k, v in self.norm_host.items()}

    def normalise(self, X, Y=None):
        nh = self.norm_host
        Xn = (X - nh["in_mean"][:, None]) * nh["in_rstd"][:, None]
        if Y is None:
            return Xn.astype(np.float32)
        Yn = (Y - nh["out_mean"][:, None]) / nh["out_std"][:, None]
        return Xn.astype(np.float32), Yn.astype(np.float32)

    # ---- kernels
    def _get(self, name, rows, cols):
        b = self._buf.get(name)
        if b is None or b.shape != (rows, cols):
            b = Dev(rows, cols)
            self._buf[name] = b
        return b

    def forward(self, Xn):
        L = _bind()
        B = Xn.shape[1]
        h = self.hidden
        Z1, H1 = self._get("Z1", h, B), self._get("H1", h, B)
        gemm(self.p["W1t"], Xn, Z1)
        _capi.check(L.pypde_mlp_bias_act(Z1.ptr, self.p["b1"].ptr, H1.ptr, h, B, 1, None), "act")
        Z2, H2 = self._get("Z2", h, B), self._get("H2", h, B)
        gemm(self.p["W2t"], H1, Z2)
        _capi.check(L.pypde_mlp_bias_act(Z2.ptr, self.p["b2"].ptr, H2.ptr, h, B, 1, None), "act")
        Y = self._get("Y", self.n_out, B)
        gemm(self.p["W3t"], H2, Y)
        _capi.check(L.pypde_mlp_bias_act(Y.ptr, self.p["b3"].ptr, Y.ptr, self.n_out, B, 0, None),
                    "act")
        return Y

    def gradients(self, Xn, Tn):
        """MSE loss and parameter gradients for one batch (device arrays)."""
        L = _bind()
        Y = self.forward(Xn)
        Z1, H1, Z2, H2 = (self._buf[k] for k in ("Z1", "H1", "Z2", "H2"))
        B = Xn.shape[1]
        h = self.hidden
        dY = self._get("dY", self.n_out, B)
        loss = self._get("loss", 1, 1)
        _capi.check(L.pypde_dev_memset(loss.ptr, 0, 4), "memset")
        _capi.check(L.pypde_mlp_mse(Y.ptr, Tn.ptr, dY.ptr, loss.ptr, self.n_out, B, None), "mse")
        g = {k: self._get("g" + k, *v.shape) for k, v in self.p.items()}
        gemm(dY, H2, g["W3t"], transB=True)
        _capi.check(L.pypde_mlp_rowsum(dY.ptr, g["b3"].ptr, self.n_out, B, None), "rowsum")
        dH2 = self._get("dH2", h, B)
        gemm(self.p["W3t"], dY, dH2, transA=True)
        dZ2 = self._get("dZ2", h, B)
        _capi.check(L.pypde_mlp_act_grad(dZ2.ptr, dH2.ptr, Z2.ptr, h * B, None), "act_grad")
        gemm(dZ2, H1, g["W2t"], transB=True)
        _capi.check(L.pypde_mlp_rowsum(dZ2.ptr, g["b2"].ptr, h, B, None), "rowsum")
        dH1 = self._get("dH1", h, B)
        gemm(self.p["W2t"], dZ2, dH1, transA=True)
        dZ1 = self._get("dZ1", h, B)
        _capi.check(L.pypde_mlp_act_grad(dZ1.ptr, dH1.ptr, Z1.ptr, h * B, None), "act_grad")
        gemm(dZ1, Xn, g["W1t"], transB=True)
        _capi.check(L.pypde_mlp_rowsum(dZ1.ptr, g["b1"].ptr, h, B, None), "rowsum")
        return loss, g

    def adam_step(self, g, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        L = _bind()
        self.step_count += 1
        for k, w in self.p.items():
            _capi.check(L.pypde_mlp_adam(w.ptr, g[k].ptr, self.m[k].ptr, self.v[k].ptr, w.n, lr,
                                         b1, b2, eps, self.step_count, None), "adam")

    def train(self, X, Y, epochs=10, batch=4096, lr=2e-3, seed=0, log=None):
        """X (n_in x S), Y (n_out x S): raw host pairs (float32/64)."""
        self.fit_normalisation(X, Y)
        Xn, Yn = self.normalise(X, Y)
        S = X.shape[1]
        rng = np.random.default_rng(seed)
        perm = rng.permutation(S)
        nb = S // batch
        # batch-major device copies: batch b = columns perm[b*batch:(b+1)*batch]
        Xb = Dev.from_numpy(np.concatenate([Xn[:, perm[b * batch:(b + 1) * batch]]
                                            for b in range(nb)], 0))
        Yb = Dev.from_numpy(np.concatenate([Yn[:, perm[b * batch:(b + 1) * batch]]
                                            for b in range(nb)], 0))
        hist = []
        total = epochs * nb
        k = 0
        for ep in range(epochs):
            tot = 0.0
            for b in rng.permutation(nb):
                xb = Xb.view_rows(b * self.n_in, (b + 1) * self.n_in)
                yb = Yb.view_rows(b * self.n_out, (b + 1) * self.n_out)
                loss, g = self.gradients(xb, yb)
                self.adam_step(g, lr=lr * 0.5 * (1 + np.cos(np.pi * k / total)))
                tot += float(loss.numpy()[0, 0])
                k += 1
            hist.append(tot / max(nb, 1))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4e}")
        return hist

    # ---- solver attachment
    def attach(self, solver):
        s = PypdeMLP()
        for k in PARAMS:
            setattr(s, k, self.p[k].ptr)
        for k in NORMS:
            setattr(s, k, self.norm[k].ptr)
        self._s = s
        _capi.check(_bind().pypde_set_guess_model(C.c_void_p(solver.h), C.byref(s)),
                    "set_guess_model")

    @staticmethod
    def detach(solver):
        _capi.check(_bind().pypde_set_guess_model(C.c_void_p(solver.h), None), "set_guess_model")

    def state_dict(self):
        d = {k: v.numpy() for k, v in self.p.items()}
        d.update({k: np.asarray(v) for k, v in self.norm_host.items()})
        return d

    def load_state_dict(self, d):
        for k in PARAMS:
            self.p[k] = Dev.from_numpy(d[k])
        self.norm_host = {k: np.asarray(d[k], dtype=np.float64) for k in NORMS}
        self.norm = {k: Dev.from_numpy(v) for k, v in self.norm_host.items()}


def geometry(solver):
    L = _bind()
    nwc, ni, no = C.c_long(), C.c_int(), C.c_int()
    L.pypde_geometry(C.c_void_p(solver.h), C.byref(nwc), C.byref(ni), C.byref(no))
    return nwc.value, ni.value, no.value


def harvest(solver, steps):
    """Advance `solver` by `steps` steps, collecting one (features, target) pair
    per reconstructed cell per step; returns host arrays X (n_in x S), Y."""
    L = _bind()
    nwc, ni, no = geometry(solver)
    X = Dev(ni, steps * nwc)
    Y = Dev(no, steps * nwc)
    for s in range(steps):
        st = _capi.Stats()
        st.struct_size = C.sizeof(_capi.Stats)
        rc = L.pypde_harvest_step(C.c_void_p(solver.h), X.ptr, Y.ptr, steps * nwc, s * nwc,
                                  C.byref(st))
        _capi.check(rc, "harvest_step")
    return X.numpy(), Y.numpy()
