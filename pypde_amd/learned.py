"""Learned initial guess for the stiff predictor's Newton iteration.

The reference starts every cell's space-time solve from its reconstruction held
constant in time (DGSolver::initial_guess, dg.cpp:12-20, called at dg.cpp:216).
Here a small MLP maps a cell's reconstruction and dt to a guess for the whole
space-time tensor; Newton still iterates to the unchanged tolerance and drops a
guess whose residual grows (see the predictor kernel).

Network: x = (wh (N^D V values), dt), normalised per feature ->
128 SiLU -> 128 SiLU -> y, y de-normalised = (qh - w) / dt  (N^(D+1) V values).
Training pairs come from the solver itself (pypde_harvest_step); forward and
backward run on this package's own matrix-core kernels (BF16 operands, FP32
accumulation) with Adam on FP32 master weights.  torch provides device memory
and the stream only.
"""
import ctypes as C

import numpy as np

from pypde_amd import _capi


def _torch():
    import torch
    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream():
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


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
    L._mlp_bound = True
    return L


class PypdeMLP(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ["W1t", "b1", "W2t", "b2", "W3t", "b3", "in_mean",
                                          "in_rstd", "out_mean", "out_std"]]


def gemm(A, B, Cm, transA=False, transB=False, alpha=1.0, beta=0.0):
    """Cm = alpha op(A) op(B) + beta Cm on the matrix cores (row-major fp32)."""
    L = _bind()
    M, N = Cm.shape
    K = A.shape[0] if transA else A.shape[1]
    rc = L.pypde_mlp_gemm(_ptr(A), A.shape[1], int(transA), _ptr(B), B.shape[1], int(transB),
                          _ptr(Cm), Cm.shape[1], M, N, K, alpha, beta, _stream())
    _capi.check(rc, "gemm")


class GuessMLP:
    """3-layer SiLU MLP trained and evaluated with the package's HIP kernels."""

    def __init__(self, n_in, n_out, hidden=128, seed=0, device="cuda"):
        torch = _torch()
        self.n_in, self.n_out, self.hidden = n_in, n_out, hidden
        g = np.random.default_rng(seed)

        def init(o, i):
            return torch.tensor(g.uniform(-1, 1, (o, i)) * np.sqrt(3.0 / i), dtype=torch.float32,
                                device=device)

        self.p = {"W1t": init(hidden, n_in), "b1": torch.zeros(hidden, 1, device=device),
                  "W2t": init(hidden, hidden), "b2": torch.zeros(hidden, 1, device=device),
                  "W3t": init(n_out, hidden) * 0.1, "b3": torch.zeros(n_out, 1, device=device)}
        self.norm = {"in_mean": torch.zeros(n_in, device=device),
                     "in_rstd": torch.ones(n_in, device=device),
                     "out_mean": torch.zeros(n_out, device=device),
                     "out_std": torch.ones(n_out, device=device)}
        self.m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.step_count = 0
        self.device = device

    # ---- normalisation (computed from raw harvested pairs)
    def fit_normalisation(self, X, Y):
        torch = _torch()
        self.norm["in_mean"] = X.mean(1).contiguous()
        self.norm["in_rstd"] = (1.0 / torch.clamp(X.std(1), min=1e-12)).contiguous()
        self.norm["out_mean"] = Y.mean(1).contiguous()
        self.norm["out_std"] = torch.clamp(Y.std(1), min=1e-12).contiguous()

    def normalise(self, X, Y=None):
        Xn = ((X - self.norm["in_mean"][:, None]) * self.norm["in_rstd"][:, None]).contiguous()
        if Y is None:
            return Xn
        Yn = ((Y - self.norm["out_mean"][:, None]) / self.norm["out_std"][:, None]).contiguous()
        return Xn, Yn

    # ---- forward / backward on the package's kernels
    def forward(self, Xn, keep=False):
        torch = _torch()
        L = _bind()
        B = Xn.shape[1]
        h = self.hidden
        Z1 = torch.empty(h, B, device=self.device)
        H1 = torch.empty(h, B, device=self.device)
        gemm(self.p["W1t"], Xn, Z1)
        _capi.check(L.pypde_mlp_bias_act(_ptr(Z1), _ptr(self.p["b1"]), _ptr(H1), h, B, 1,
                                         _stream()), "bias_act")
        Z2 = torch.empty(h, B, device=self.device)
        H2 = torch.empty(h, B, device=self.device)
        gemm(self.p["W2t"], H1, Z2)
        _capi.check(L.pypde_mlp_bias_act(_ptr(Z2), _ptr(self.p["b2"]), _ptr(H2), h, B, 1,
                                         _stream()), "bias_act")
        Y = torch.empty(self.n_out, B, device=self.device)
        gemm(self.p["W3t"], H2, Y)
        _capi.check(L.pypde_mlp_bias_act(_ptr(Y), _ptr(self.p["b3"]), _ptr(Y), self.n_out, B, 0,
                                         _stream()), "bias_act")
        if keep:
            self._cache = (Xn, Z1, H1, Z2, H2)
        return Y

    def gradients(self, Xn, Tn):
        """MSE loss and parameter gradients (dict) for one batch."""
        torch = _torch()
        L = _bind()
        Y = self.forward(Xn, keep=True)
        Xn, Z1, H1, Z2, H2 = self._cache
        B = Xn.shape[1]
        h = self.hidden
        dY = torch.empty_like(Y)
        loss = torch.zeros(1, device=self.device)
        _capi.check(L.pypde_mlp_mse(_ptr(Y), _ptr(Tn), _ptr(dY), _ptr(loss), self.n_out, B,
                                    _stream()), "mse")
        g = {k: torch.empty_like(v) for k, v in self.p.items()}
        gemm(dY, H2, g["W3t"], transB=True)
        _capi.check(L.pypde_mlp_rowsum(_ptr(dY), _ptr(g["b3"]), self.n_out, B, _stream()), "rowsum")
        dH2 = torch.empty(h, B, device=self.device)
        gemm(self.p["W3t"], dY, dH2, transA=True)
        dZ2 = torch.empty_like(dH2)
        _capi.check(L.pypde_mlp_act_grad(_ptr(dZ2), _ptr(dH2), _ptr(Z2), h * B, _stream()), "act")
        gemm(dZ2, H1, g["W2t"], transB=True)
        _capi.check(L.pypde_mlp_rowsum(_ptr(dZ2), _ptr(g["b2"]), h, B, _stream()), "rowsum")
        dH1 = torch.empty(h, B, device=self.device)
        gemm(self.p["W2t"], dZ2, dH1, transA=True)
        dZ1 = torch.empty_like(dH1)
        _capi.check(L.pypde_mlp_act_grad(_ptr(dZ1), _ptr(dH1), _ptr(Z1), h * B, _stream()), "act")
        gemm(dZ1, Xn, g["W1t"], transB=True)
        _capi.check(L.pypde_mlp_rowsum(_ptr(dZ1), _ptr(g["b1"]), h, B, _stream()), "rowsum")
        return loss, g

    def adam_step(self, g, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        L = _bind()
        self.step_count += 1
        for k, w in self.p.items():
            _capi.check(L.pypde_mlp_adam(_ptr(w), _ptr(g[k]), _ptr(self.m[k]), _ptr(self.v[k]),
                                         w.numel(), lr, b1, b2, eps, self.step_count, _stream()),
                        "adam")

    def train(self, X, Y, epochs=10, batch=4096, lr=2e-3, seed=0, log=None):
        """X (n_in x S), Y (n_out x S) raw fp32 device pairs."""
        torch = _torch()
        self.fit_normalisation(X, Y)
        Xn, Yn = self.normalise(X, Y)
        S = X.shape[1]
        gen = torch.Generator(device="cpu").manual_seed(seed)
        hist = []
        steps_total = epochs * max(1, S // batch)
        k = 0
        for ep in range(epochs):
            perm = torch.randperm(S, generator=gen).to(X.device)
            tot = 0.0
            nb = 0
            for s in range(0, S - batch + 1, batch):
                idx = perm[s:s + batch]
                xb = Xn[:, idx].contiguous()
                yb = Yn[:, idx].contiguous()
                loss, g = self.gradients(xb, yb)
                # cosine learning-rate decay
                lr_k = lr * 0.5 * (1 + np.cos(np.pi * k / steps_total))
                self.adam_step(g, lr=lr_k)
                tot += float(loss.item())
                nb += 1
                k += 1
            hist.append(tot / max(nb, 1))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4e}")
        return hist

    # ---- solver attachment
    def _struct(self):
        self._flat = {k: v.contiguous() for k, v in list(self.p.items()) + list(self.norm.items())}
        s = PypdeMLP()
        for k, t in self._flat.items():
            setattr(s, k, t.data_ptr())
        return s

    def attach(self, solver):
        L = _bind()
        torch = _torch()
        torch.cuda.synchronize()
        self._s = self._struct()
        _capi.check(L.pypde_set_guess_model(C.c_void_p(solver.h), C.byref(self._s)),
                    "set_guess_model")

    @staticmethod
    def detach(solver):
        L = _bind()
        _capi.check(L.pypde_set_guess_model(C.c_void_p(solver.h), None), "set_guess_model")

    def state_dict(self):
        return {k: v.detach().cpu().numpy() for k, v in list(self.p.items()) +
                list(self.norm.items())}

    def load_state_dict(self, d):
        torch = _torch()
        for k in self.p:
            self.p[k] = torch.tensor(d[k], dtype=torch.float32, device=self.device)
        for k in self.norm:
            self.norm[k] = torch.tensor(d[k], dtype=torch.float32, device=self.device)


def geometry(solver):
    L = _bind()
    nwc, ni, no = C.c_long(), C.c_int(), C.c_int()
    L.pypde_geometry(C.c_void_p(solver.h), C.byref(nwc), C.byref(ni), C.byref(no))
    return nwc.value, ni.value, no.value


def harvest(solver, steps, X=None, Y=None, col0=0):
    """Advance `solver` by `steps` steps, collecting one (features, target) pair
    per reconstructed cell per step into device arrays (allocated if None)."""
    torch = _torch()
    L = _bind()
    nwc, ni, no = geometry(solver)
    if X is None:
        X = torch.empty(ni, steps * nwc, device="cuda")
        Y = torch.empty(no, steps * nwc, device="cuda")
    ld = X.shape[1]
    for s in range(steps):
        st = _capi.Stats()
        st.struct_size = C.sizeof(_capi.Stats)
        torch.cuda.synchronize()
        rc = L.pypde_harvest_step(C.c_void_p(solver.h), _ptr(X), _ptr(Y), ld, col0 + s * nwc,
                                  C.byref(st))
        _capi.check(rc, "harvest_step")
    return X, Y
