"""Pins the CPU checker of the MLP's training kernels (the FP64 numpy
backpropagation in test_learned.ref_grads, which the GPU gradients are held
to) against central finite differences of its own loss: a dropped term, a
wrong transpose or a wrong activation derivative fails here."""
import numpy as np

from test_learned import ref_grads


def test_backprop_checker_matches_finite_differences():
    rng = np.random.default_rng(8)
    n_in, hid, n_out, B = 5, 6, 4, 7
    p = {"W1t": rng.standard_normal((hid, n_in)), "b1": rng.standard_normal((hid, 1)),
         "W2t": rng.standard_normal((hid, hid)), "b2": rng.standard_normal((hid, 1)),
         "W3t": rng.standard_normal((n_out, hid)), "b3": rng.standard_normal((n_out, 1))}
    X = rng.standard_normal((n_in, B))
    T = rng.standard_normal((n_out, B))
    _, g, _ = ref_grads(p, X, T, round_bf16=False)
    h = 1e-6
    for k, w in p.items():
        fd = np.zeros_like(w)
        for idx in np.ndindex(w.shape):
            w0 = w[idx]
            w[idx] = w0 + h
            lp = ref_grads(p, X, T, round_bf16=False)[0]
            w[idx] = w0 - h
            lm = ref_grads(p, X, T, round_bf16=False)[0]
            w[idx] = w0
            fd[idx] = (lp - lm) / (2 * h)
        err = np.max(np.abs(fd - g[k])) / max(np.max(np.abs(g[k])), 1e-30)
        assert err <= 1e-6, (k, err)
