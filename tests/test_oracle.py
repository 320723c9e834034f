"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the GPU checker.

Each oracle function is checked against something it cannot share a mistake with:
closed forms and 50-digit arithmetic (tables), numpy (eigenvalues, exp), exactness
properties of the method (polynomial data), the compiled reference core
(oracle/_ref, stage by stage and end to end) and the committed golden vectors made
from it (tests/golden/).
"""
import ctypes as C
import os

import numpy as np
import pytest

import cpu_checkers as cc
from pypde_amd import systems

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_gold(name):
    return np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)


# ---------------------------------------------------------------- tables

def _mp_tables(n):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    roots = sorted(mp.findroot(lambda z: mp.legendre(n, z),
                               mp.cos(mp.pi * (i + mp.mpf(3) / 4) / (n + mp.mpf(1) / 2)))
                   for i in range(n))
    x = [(1 + r) / 2 for r in roots]
    w = [1 / ((1 - r ** 2) * mp.diff(lambda z: mp.legendre(n, z), r) ** 2) for r in roots]

    def psi(j, s):
        p = mp.mpf(1)
        for m in range(n):
            if m != j:
                p *= (s - x[m]) / (x[j] - x[m])
        return p

    FN2, CN2 = (n - 1) // 2, n // 2
    off = [-(n - 1), 0, -CN2, -FN2]
    wm = [[[mp.quad(lambda s: psi(j, s), [i + off[k], i + off[k] + 1]) for j in range(n)]
           for i in range(n)] for k in range(4)]
    dpsi = lambda j, s: mp.diff(lambda z: psi(j, z), s)
    dg = [[psi(i, 1) * psi(j, 1) - mp.quad(lambda s: dpsi(i, s) * psi(j, s), [0, 1])
           for j in range(n)] for i in range(n)]
    f = lambda a: np.array(a, dtype=float)
    return dict(nodes=f(x), wghts=f(w), weno_mat=f(wm).ravel(), dgmat=f(dg).ravel(),
                endvals=f([[psi(j, e) for j in range(n)] for e in (0, 1)]).ravel(),
                dervals=f([[dpsi(j, x[i]) for j in range(n)] for i in range(n)]).ravel())


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_tables_match_high_precision(n):
    o = cc.tables(n)
    e = _mp_tables(n)
    for k in e:
        scale = max(1.0, np.max(np.abs(e[k])))
        assert np.max(np.abs(o[k] - e[k])) <= 2 * np.finfo(float).eps * scale, k


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_tables_identities(n):
    o = cc.tables(n)
    x, w = np.polynomial.legendre.leggauss(n)
    assert np.allclose(o["nodes"], (x + 1) / 2, rtol=0, atol=1e-15)
    assert np.allclose(o["wghts"], w / 2, rtol=0, atol=1e-15)
    # partition of unity: ENDVALS rows sum to 1, DERVALS rows to 0
    assert np.allclose(o["endvals"].reshape(2, n).sum(1), 1.0, atol=1e-13)
    assert np.allclose(o["dervals"].reshape(n, n).sum(1), 0.0, atol=1e-11)
    # stencil inverses, DG inverse
    for s in range(4):
        m = o["weno_mat"].reshape(4, n, n)[s]
        mi = o["weno_inv"].reshape(4, n, n)[s]
        assert np.allclose(m @ mi, np.eye(n), atol=1e-11)
    assert np.allclose(o["dgmat"].reshape(n, n) @ o["dginv"].reshape(n, n), np.eye(n),
                       atol=1e-12)
    # Faddeev-LeVerrier: det(xI - Tw) coefficients and adj(xI - Tw) at a sample x
    Tw = np.diag(1 / o["wghts"]) @ o["dgmat"].reshape(n, n)
    assert np.allclose(np.poly(Tw)[::-1][:n], o["fl_c"], rtol=1e-10, atol=1e-10)
    xs = 0.37
    adj = sum(o["fl_m"].reshape(n, n, n)[k] * xs ** (n - 1 - k) for k in range(n))
    assert np.allclose(adj @ (xs * np.eye(n) - Tw), np.linalg.det(xs * np.eye(n) - Tw) *
                       np.eye(n), atol=1e-9)


@pytest.mark.ref
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_tables_match_reference(n):
    # the reference builds its tables in the power basis (polynomials.cpp:6-66),
    # whose rounding grows with N; the oracle's are exact to double rounding
    o, r = cc.tables(n), cc.ref_tables(n)
    tol = 4e-15 * 10 ** max(0, n - 3)
    for k in r:
        scale = max(1.0, np.max(np.abs(r[k])))
        assert np.max(np.abs(o[k] - r[k])) <= tol * scale, k


def test_tables_match_golden():
    g = load_gold("tables")
    for n in [1, 2, 3, 4, 5]:
        o = cc.tables(n)
        tol = 4e-15 * 10 ** max(0, n - 3)
        for k in ["nodes", "wghts", "endvals", "dervals", "weno_mat", "sig", "dgmat"]:
            r = g[f"{k}_{n}"]
            assert np.max(np.abs(o[k] - r)) <= tol * max(1.0, np.max(np.abs(r))), (n, k)


# ------------------------------------------------------- eigen and exp

def test_max_abs_eig_numpy():
    rng = np.random.default_rng(0)
    for V in range(1, 9):
        for k in range(300):
            A = rng.standard_normal((V, V))
            if k % 4 == 1:
                A = A + A.T                      # real spectrum
            elif k % 4 == 2:
                A[np.tril_indices(V, -2)] = 0    # already Hessenberg
            elif k % 4 == 3 and V >= 2:
                A = np.triu(A)                   # eigenvalues on the diagonal
            A = np.ascontiguousarray(A)
            got = cc.oracle().oracle_max_abs_eig(cc.dp(A), V)
            want = np.max(np.abs(np.linalg.eigvals(A)))
            assert abs(got - want) <= 1e-12 * want, (V, k)
    # clustered spectra as finite-difference Jacobians of multi-wave systems
    # produce (a near-triple eigenvalue split by rounding-level noise)
    for k in range(300):
        V = 3 + k % 6
        lam = rng.uniform(-2, 2)
        D = np.diag(np.r_[np.full(V - 2, lam), rng.uniform(-4, 4, 2)])
        P = rng.standard_normal((V, V))
        A = np.ascontiguousarray(P @ D @ np.linalg.inv(P) + 1e-12 * rng.standard_normal((V, V)))
        got = cc.oracle().oracle_max_abs_eig(cc.dp(A), V)
        want = np.max(np.abs(np.linalg.eigvals(A)))
        assert abs(got - want) <= 1e-9 * want, (V, k)
    # special cases: zero matrix, Jordan block, pure rotation (complex pair)
    assert cc.oracle().oracle_max_abs_eig(cc.dp(np.zeros((3, 3))), 3) == 0.0
    J = np.ascontiguousarray(np.eye(4) * 2 + np.eye(4, k=1))
    assert abs(cc.oracle().oracle_max_abs_eig(cc.dp(J), 4) - 2.0) < 1e-3
    R = np.ascontiguousarray([[0.0, -3.0], [3.0, 0.0]])
    assert abs(cc.oracle().oracle_max_abs_eig(cc.dp(R), 2) - 3.0) < 1e-14
    # one real root and a larger complex pair: the Laguerre runs from both ends
    # reach the same real root, which must not pass as two roots
    for re_, im_, lr in [(0.28, 1.47, 0.448), (-0.1, 2.0, 1.0), (0.0, 1.0, -0.999)]:
        B = np.array([[lr, 0, 0], [0, re_, -im_], [0, im_, re_]])
        P = np.random.default_rng(5).standard_normal((3, 3))
        A = np.ascontiguousarray(P @ B @ np.linalg.inv(P))
        want = np.max(np.abs(np.linalg.eigvals(A)))
        assert abs(cc.oracle().oracle_max_abs_eig(cc.dp(A), 3) - want) <= 1e-12 * want
    # a simple extreme eigenvalue 0.036 from a five-fold cluster split by 1e-12
    # noise: as a polynomial root it is good to ~1e-8 only, so the accuracy
    # check must send it to QR
    D = np.diag([-0.90238406] * 5 + [-0.97495421, -0.78669644])
    P = np.random.default_rng(6).standard_normal((7, 7))
    A = np.ascontiguousarray(P @ D @ np.linalg.inv(P)
                             + 1e-12 * np.random.default_rng(7).standard_normal((7, 7)))
    want = np.max(np.abs(np.linalg.eigvals(A)))
    assert abs(cc.oracle().oracle_max_abs_eig(cc.dp(A), 7) - want) <= 1e-10 * want


def test_max_abs_eig_scales_exactly():
    """max |lambda| is homogeneous: scaling a matrix by 2^k scales it by 2^k
    bit for bit, whether the largest entry lies inside 2^+-64 (the matrix is
    reduced unscaled) or outside (scaled first), on random, real-spectrum and
    degree-17 matrices."""
    rng = np.random.default_rng(12)
    O = cc.oracle()
    for V in (3, 4, 5, 17):
        for t in range(300):
            if t % 2:
                P = rng.standard_normal((V, V))
                A = P @ np.diag(rng.uniform(-3, 3, V)) @ np.linalg.inv(P)
            else:
                A = rng.standard_normal((V, V))
            base = O.oracle_max_abs_eig(cc.dp(np.ascontiguousarray(A)), V)
            for k in (-200, -70, -30, 10, 40, 63, 90, 300):
                got = O.oracle_max_abs_eig(cc.dp(np.ascontiguousarray(A * 2.0 ** k)), V)
                assert got == base * 2.0 ** k, (V, t, k)


@pytest.mark.ref
def test_max_abs_eig_reference():
    rng = np.random.default_rng(1)
    for V in range(1, 6):
        for k in range(200):
            A = np.ascontiguousarray(rng.standard_normal((V, V)))
            a = cc.oracle().oracle_max_abs_eig(cc.dp(A), V)
            b = cc.ref().ref_max_abs_eig(cc.dp(A.copy()), V)
            assert abs(a - b) <= 1e-13 * b


@pytest.mark.parametrize("name,V", [("euler2d", 4), ("reactive2d", 5)])
def test_system_matrix_against_reference_formula(name, V):
    """The finite-difference Jacobian (system.cpp:6-26, NumericalDiff.h:6-44:
    column i = (F(q + h_i e_i) - F(q)) / h_i, h_i = max(sqrt(eps)|q_i|,
    sqrt(eps))) from the reference's formula in Python against the oracle's,
    which multiplies by reciprocals of the h_i formed with one division for all
    columns, or one division each when the product of max(|q_i|, 1) passes
    2^1000 (states of size 1e80 here): entrywise within a few ulps."""
    s = cc.system(name)
    s.lib.sys_F.argtypes = [cc.PD, cc.PD, cc.PD, C.c_int]
    s.lib.sys_F.restype = None
    eps = np.sqrt(np.finfo(float).eps)
    rng = np.random.default_rng(9)
    dq = np.zeros(2 * V)

    def F(q, d):
        out = np.zeros(V)
        s.lib.sys_F(cc.dp(out), cc.dp(np.ascontiguousarray(q)), cc.dp(dq), d)
        return out

    states = []
    for _ in range(200):
        q = rng.uniform(0.5, 2.0, V) * rng.choice([1e-3, 1.0, 10.0, 1e3])
        q[V - 1 if name == "euler2d" else 3] += 50.0 * q[0]   # energy: positive pressure
        states.append(q)
    for _ in range(20):
        states.append(rng.uniform(0.5, 2.0, V) * 1e80)        # product > 2^1000
    fallback = 0
    for q in states:
        fallback += np.sum(np.log2(np.maximum(np.abs(q), 1.0))) > 1000
        for d in (0, 1):
            want = np.zeros((V, V))
            f0 = F(q, d)
            for i in range(V):
                h = max(eps * abs(q[i]), eps)
                q2 = q.copy()
                q2[i] = q[i] + h
                want[:, i] = (F(q2, d) - f0) / h
            got = np.zeros(V * V)
            cc.oracle().oracle_system_matrix(s.F, None, cc.dp(np.ascontiguousarray(q)),
                                             cc.dp(dq), d, 2, V, cc.dp(got))
            got = got.reshape(V, V)
            assert np.all(np.abs(got - want) <= 8 * np.finfo(float).eps * np.abs(want)), (q, d)
    assert fallback == 20


def test_exp_within_one_ulp():
    rng = np.random.default_rng(2)
    xs = np.concatenate([rng.uniform(-700, 700, 20000), rng.uniform(-3, 3, 20000),
                         [0.0, 1.0, -1.0, 1e-300, -1e-300, 709.0, -708.0]])
    s = cc.system("reactive2d")
    for x in xs:
        got = cc.oracle().oracle_exp(float(x))
        assert got == s.lib.sys_exp(float(x))   # checker and system prelude agree
        want = np.exp(x)
        assert abs(got - want) <= np.spacing(want), x


def test_cbrt_within_one_ulp():
    """The system prelude's pde_cbrt (used by the GPR source term for
    det(A)^(5/3)) against 50-digit decimal arithmetic over the double range,
    plus its special values; it is the same text on host and device."""
    from decimal import Decimal, getcontext
    getcontext().prec = 50
    rng = np.random.default_rng(4)
    s = cc.system("gpr")
    xs = np.concatenate([np.exp(rng.uniform(-700, 700, 3000)), rng.uniform(0.4, 4.1, 3000),
                         [1.0, 8.0, 27.0, 0.125, 2.0 ** -1074, 1.7976931348623157e308]])
    third = Decimal(1) / Decimal(3)
    for x in xs:
        got = s.lib.sys_cbrt(float(x))
        want = Decimal(float(x)) ** third
        assert abs(Decimal(got) - want) <= Decimal(np.spacing(got)), x
    assert s.lib.sys_cbrt(0.0) == 0.0
    assert np.isnan(s.lib.sys_cbrt(-1.0)) and np.isnan(s.lib.sys_cbrt(float("nan")))
    assert s.lib.sys_cbrt(float("inf")) == float("inf")


# ----------------------------------------------------------------- WENO

def _weno(u, N):
    nX = np.array(u.shape[:-1], dtype=np.int32)
    V = u.shape[-1]
    nout = int(np.prod(nX - 2 * (N - 1)))
    wh = np.zeros(nout * N ** len(nX) * V)
    assert cc.oracle().oracle_weno(cc.dp(wh), cc.dp(np.ascontiguousarray(u)), cc.ip(nX),
                                   len(nX), N, V) == 0
    return wh.reshape(tuple(nX - 2 * (N - 1)) + (N,) * len(nX) + (V,))


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_weno_exact_on_polynomials(N):
    """Cell averages of a degree <= N-1 polynomial reconstruct it exactly at the nodes."""
    x = cc.tables(N)["nodes"]
    n = 16
    rng = np.random.default_rng(N)
    c = rng.uniform(-1, 1, N)
    P = np.polynomial.polynomial.Polynomial(c)
    Pi = P.integ()
    h = 1.0 / n
    avg = np.array([(Pi((i + 1) * h) - Pi(i * h)) / h for i in range(n)])
    u = np.ascontiguousarray(np.stack([avg, 2 * avg + 1], -1))
    wh = _weno(u, N)
    for i in range(wh.shape[0]):
        cell = i + N - 1
        want = P((cell + x) * h)
        assert np.allclose(wh[i, :, 0], want, rtol=0, atol=1e-11)
        assert np.allclose(wh[i, :, 1], 2 * want + 1, rtol=0, atol=1e-11)


@pytest.mark.parametrize("N", [2, 3, 4])
def test_weno_mean_and_constants(N):
    """Gauss mean of each reconstruction equals the cell average; constants preserved."""
    rng = np.random.default_rng(7)
    u = rng.uniform(0.5, 2.0, (11, 9, 3))
    w = cc.tables(N)["wghts"]
    wh = _weno(u, N)
    mean = np.einsum("i,j,abijv->abv", w, w, wh)
    inner = u[N - 1:u.shape[0] - N + 1, N - 1:u.shape[1] - N + 1]
    assert np.allclose(mean, inner, rtol=1e-12)
    wc = _weno(np.ones((9, 8, 2)) * 3.25, N)
    assert np.allclose(wc, 3.25, rtol=1e-14)


@pytest.mark.ref
@pytest.mark.parametrize("ndim,N,V", [(1, 2, 3), (1, 3, 4), (2, 2, 3), (2, 3, 5), (2, 4, 4)])
def test_weno_reference(ndim, N, V):
    rng = np.random.default_rng(11 + N)
    nX = np.array([12, 9][:ndim], dtype=np.int32)
    u = np.ascontiguousarray(rng.uniform(0.5, 2, tuple(nX) + (V,)))
    wo = _weno(u, N).ravel()
    wr = np.zeros_like(wo)
    cc.ref().weno_solver(cc.dp(wr), cc.dp(u.copy()), cc.ip(nX), ndim, N, V)
    # rough random data: the nonlinear weights amplify the reference's QR-solve
    # rounding (cond(M) grows with N)
    tol = {2: 1e-14, 3: 1e-12, 4: 1e-10}[N]
    assert np.max(np.abs(wo - wr)) <= tol * np.max(np.abs(wr))


def test_weno_golden():
    g = load_gold("weno")
    for key in [k for k in g.files if k.startswith("u_")]:
        tag = key[2:]
        ndim, N, V = (int(t) for t in tag.split("_"))
        wo = _weno(g[key], N).ravel()
        wr = g["wh_" + tag]
        tol = {2: 1e-14, 3: 1e-12, 4: 1e-10}[N]
        assert np.max(np.abs(wo - wr)) <= tol * np.max(np.abs(wr)), tag


# ------------------------------------------------------ stage pipeline

def euler2d_field(nx, ny, seed=0, noise=0.05):
    rng = np.random.default_rng(seed)
    g = 1.4
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    r = 1 + 0.2 * np.sin(2 * np.pi * (X + Y)) + noise * rng.uniform(-1, 1, X.shape)
    u = 0.5 + 0.1 * np.cos(2 * np.pi * X)
    v = 0.3 + 0.1 * np.sin(2 * np.pi * Y)
    p = np.ones_like(X)
    E = p / (g - 1) + 0.5 * r * (u * u + v * v)
    return np.ascontiguousarray(np.stack([r, r * u, r * v, E], -1))


@pytest.mark.ref
@pytest.mark.parametrize("N", [2, 3, 4])
def test_stages_reference(N):
    """timestep, Picard predictor and corrector, each fed identical inputs."""
    s = cc.system("euler2d")
    O, R = cc.oracle(), cc.ref()
    u = euler2d_field(14, 12)
    V = 4
    nX = np.array(u.shape[:2], dtype=np.int32)
    nw = int(np.prod(nX - 2 * (N - 1)))
    Nd = N * N
    wh = np.zeros(nw * Nd * V)
    R.weno_solver(cc.dp(wh), cc.dp(u.copy()), cc.ip(nX), 2, N, V)
    dX = np.array([1 / 14, 1 / 12])
    dto = O.oracle_timestep(s.F, None, cc.dp(wh), nw, cc.dp(dX), 2, N, V, 0.9, 1.0, 0.0, 10)
    dtr = R.ref_timestep(s.F, None, cc.dp(wh), nw, cc.dp(dX), 2, N, V, 0.9, 1.0, 0.0, 10, 0)
    # forward-difference Jacobians turn 1-ulp differences of the cell means into
    # ~1e-8 relative differences of the wave speeds (NumericalDiff.h:23-31)
    assert abs(dto - dtr) <= 3e-8 * dtr
    qo = np.zeros(nw * N * Nd * V)
    qr = np.zeros_like(qo)
    O.oracle_predictor(s.F, None, None, None, cc.dp(wh), nw, cc.dp(dX), 2, N, V, dtr, 0, None,
                       cc.dp(qo), None)
    R.ref_predictor(s.F, None, None, cc.dp(wh), nw, cc.dp(dX), 2, N, V, dtr, 0, cc.dp(qr))
    assert np.max(np.abs(qo - qr)) <= 1e-12 * np.max(np.abs(qr))
    nXi = nX - 2 * (N - 1) - 2
    ui = np.ascontiguousarray(u[N:N + nXi[0], N:N + nXi[1]]).copy()
    ur = ui.copy()
    O.oracle_fv(s.F, None, None, cc.dp(ui), cc.dp(qr), cc.ip(nXi), 2, cc.dp(dX), 0, N, V, dtr, 0)
    R.ref_fv_apply(s.F, None, None, cc.dp(ur), cc.dp(qr.copy()), cc.ip(nXi), 2, cc.dp(dX), 0,
                   N, V, dtr, 0)
    assert np.max(np.abs(ui - ur)) <= 1e-9 * np.max(np.abs(ur))


@pytest.mark.ref
def test_b_and_source_paths_reference():
    """Non-conservative B (rhs, face path integral, volume term) and a source.
    The reference evaluates B.dQ through a dimension-mismatched Eigen product
    (dg.cpp:110, fv.cpp:234); the oracle's b_quirk option reproduces that
    evaluation, which pins every other part of the B path here."""
    s = cc.system("ncadvect")
    O, R = cc.oracle(), cc.ref()
    N, V = 3, 2
    rng = np.random.default_rng(3)
    x = (np.arange(12) + 0.5) / 12
    X, Y = np.meshgrid(x, x, indexing="ij")
    u = np.ascontiguousarray(np.stack([np.sin(2 * np.pi * X) + 0.1 * rng.uniform(size=X.shape),
                                       np.cos(2 * np.pi * Y)], -1))
    nX = np.array(u.shape[:2], dtype=np.int32)
    nw = int(np.prod(nX - 2 * (N - 1)))
    wh = np.zeros(nw * N * N * V)
    R.weno_solver(cc.dp(wh), cc.dp(u.copy()), cc.ip(nX), 2, N, V)
    dX = np.array([1 / 12, 1 / 12])
    dt = R.ref_timestep(None, s.B, cc.dp(wh), nw, cc.dp(dX), 2, N, V, 0.9, 1.0, 0.0, 10, 0)
    dto = O.oracle_timestep(None, s.B, cc.dp(wh), nw, cc.dp(dX), 2, N, V, 0.9, 1.0, 0.0, 10)
    assert abs(dto - dt) <= 1e-14 * dt          # B alone: no finite differences
    qo = np.zeros(nw * N * N * N * V)
    qr = np.zeros_like(qo)
    opts = cc.OracleOpts()
    opts.b_quirk = 1
    O.oracle_predictor(None, s.B, s.S, None, cc.dp(wh), nw, cc.dp(dX), 2, N, V, dt, 0,
                       C.byref(opts), cc.dp(qo), None)
    R.ref_predictor(None, s.B, s.S, cc.dp(wh), nw, cc.dp(dX), 2, N, V, dt, 0, cc.dp(qr))
    assert np.max(np.abs(qo - qr)) <= 1e-12 * np.max(np.abs(qr))
    nXi = nX - 2 * (N - 1) - 2
    ui = np.ascontiguousarray(u[N:N + nXi[0], N:N + nXi[1]]).copy()
    ur = ui.copy()
    O.oracle_fv(None, s.B, s.S, cc.dp(ui), cc.dp(qr), cc.ip(nXi), 2, cc.dp(dX), 0, N, V, dt, 1)
    R.ref_fv_apply(None, s.B, s.S, cc.dp(ur), cc.dp(qr.copy()), cc.ip(nXi), 2, cc.dp(dX), 0, N,
                   V, dt, 0)
    assert np.max(np.abs(ui - ur)) <= 1e-13 * np.max(np.abs(ur))


def reactive_field(nx, ny, seed=0, hot=12.0):
    rng = np.random.default_rng(seed)
    g = 1.2
    x = (np.arange(nx) + 0.5) / nx
    X, _ = np.meshgrid(x, np.arange(ny), indexing="ij")
    r = np.where(X < 0.3, 1.8, 1.0) * (1 + 0.05 * rng.uniform(-1, 1, X.shape))
    p = np.where(X < 0.3, 20.0, np.where(X < 0.5, hot, 1.0))
    lam = np.where(X < 0.3, 0.0, 1.0)
    E = p / (g - 1) + 25 * r * lam
    z = np.zeros_like(r)
    return np.ascontiguousarray(np.stack([r, z, z, E, r * lam], -1))


@pytest.mark.ref
def test_stiff_predictor_solves_reference_system():
    """The stiff Newton's converged iterate zeroes the reference's own residual
    obj(q) (dg.cpp:128-154) to rounding level."""
    s = cc.system("reactive2d")
    O, R = cc.oracle(), cc.ref()
    N, V, Nd = 3, 5, 9
    u = reactive_field(14, 8)
    nX = np.array(u.shape[:2], dtype=np.int32)
    nw = int(np.prod(nX - 2 * (N - 1)))
    wh = np.zeros(nw * Nd * V)
    R.weno_solver(cc.dp(wh), cc.dp(u.copy()), cc.ip(nX), 2, N, V)
    dX = np.array([1e-3, 1e-3])
    dt = O.oracle_timestep(s.F, None, cc.dp(wh), nw, cc.dp(dX), 2, N, V, 0.9, 1.0, 0.0, 10)
    qo = np.zeros(nw * N * Nd * V)
    it = np.zeros(nw, dtype=np.int32)
    O.oracle_predictor(s.F, None, s.S, s.M, cc.dp(wh), nw, cc.dp(dX), 2, N, V, dt, 1, None,
                       cc.dp(qo), it.ctypes.data_as(C.c_void_p))
    assert it.max() < 100
    for c in range(nw):
        w = np.ascontiguousarray(wh[c * Nd * V:(c + 1) * Nd * V])
        q = np.ascontiguousarray(qo[c * N * Nd * V:(c + 1) * N * Nd * V])
        res = np.zeros_like(q)
        R.ref_dg_obj(s.F, None, s.S, cc.dp(w), cc.dp(q), cc.dp(dX), 2, N, V, dt, cc.dp(res))
        assert np.max(np.abs(res)) <= 1e-13 * max(1.0, np.max(np.abs(q)))


# ----------------------------------------------------------- end to end

def sod1d(n=200, g=1.4):
    u = np.zeros((n, 3))
    for i in range(n):
        r, p = (1.0, 1.0) if i / n < 0.5 else (0.125, 0.1)
        u[i] = [r, 0.0, p / (g - 1)]
    return u


def relerr(a, b):
    V = b.shape[-1]
    sc = np.max(np.abs(b.reshape(-1, V)), axis=0)
    return np.max(np.abs((a - b).reshape(-1, V)) / sc)


@pytest.mark.ref
def test_sod_reference_within_its_own_noise():
    """C0 (Sod, 200 cells, order 3, Rusanov, CFL 0.9, tf 0.2).  The reference's
    result moves when its input moves by one ulp (Picard stops at 6e-6,
    dg.cpp:190-200, and the wave speeds are finite differences): by ~5e-6 for
    one entry, and by up to ~4e-4 when every entry moves by an ulp (a Picard
    count near the rarefaction head flips).  Rounding-level differences in the
    wave speeds, as any other eigenvalue method gives, are perturbations of
    the second kind, so the oracle is required to stay within 3x the largest
    of these measured noises."""
    u = sod1d()
    uo, _, st, rc = cc.oracle_solve("euler1d", u, 0.2, [1.0], 0, 3)
    ur, _ = cc.ref_solve("euler1d", u, 0.2, [1.0], 0, 3, nthreads=4)
    noises = []
    u2 = u.copy()
    u2[37, 0] = np.nextafter(u2[37, 0], 2.0)
    ur2, _ = cc.ref_solve("euler1d", u2, 0.2, [1.0], 0, 3, nthreads=4)
    noises.append(relerr(ur2, ur))
    rng = np.random.default_rng(1)
    for _ in range(2):
        u2 = u.copy()
        m = rng.random(u.shape) < 0.5
        u2[m] = np.nextafter(u2[m], np.inf)
        ur2, _ = cc.ref_solve("euler1d", u2, 0.2, [1.0], 0, 3, nthreads=4)
        noises.append(relerr(ur2, ur))
    noise = max(noises)
    print("reference noise", noises, "oracle vs reference", relerr(uo, ur))
    assert rc == 0 and st.steps == 103
    assert noise > 0
    assert relerr(uo, ur) <= 3 * noise


def test_sod_golden():
    g = load_gold("solves")
    uo, _, st, rc = cc.oracle_solve("euler1d", sod1d(), 0.2, [1.0], 0, 3)
    assert rc == 0
    assert relerr(uo, g["sod_final"]) <= 3 * float(g["sod_noise"])


@pytest.mark.ref
@pytest.mark.parametrize("N", [2, 3, 4])
def test_euler2d_periodic_reference(N):
    u = euler2d_field(24, 24, noise=0.0)
    uo, _, st, rc = cc.oracle_solve("euler2d", u, 0.05, [1, 1], 1, N)
    ur, _ = cc.ref_solve("euler2d", u, 0.05, [1, 1], 1, N, nthreads=8)
    assert rc == 0
    assert relerr(uo, ur) <= 1e-9


def test_euler2d_golden():
    g = load_gold("solves")
    for N in [2, 3, 4]:
        uo, _, st, rc = cc.oracle_solve("euler2d", euler2d_field(24, 24, noise=0.0), 0.05,
                                        [1, 1], 1, N)
        assert relerr(uo, g[f"e2d_final_{N}"]) <= 1e-9


def test_periodic_conservation_and_free_stream():
    u = euler2d_field(16, 12)
    uo, _, st, rc = cc.oracle_solve("euler2d", u, 0.02, [1, 1], 1, 3)
    assert np.allclose(uo.sum((0, 1)), u.sum((0, 1)), rtol=1e-13)
    c = np.ones((10, 8, 4)) * np.array([1.3, 0.4, -0.2, 3.0])
    uo, _, st, rc = cc.oracle_solve("euler2d", c, 0.02, [1, 1], 0, 3)
    assert np.max(np.abs(uo - c)) <= 1e-13


def test_reflective_equals_mirrored_periodic():
    """Reflective walls (new; boundaries.cpp:4-5 has only transmissive/periodic)
    equal a periodic run on the domain doubled by its mirror image with the
    normal momentum negated."""
    rng = np.random.default_rng(5)
    g = 1.4
    nx, ny = 12, 10
    x = (np.arange(nx) + 0.5) / nx
    X, Y = np.meshgrid(x, (np.arange(ny) + 0.5) / ny, indexing="ij")
    r = 1 + 0.3 * np.exp(-30 * ((X - 0.3) ** 2 + (Y - 0.6) ** 2))
    u = np.stack([r, 0 * r, 0 * r, (1 + 0.5 * (r - 1)) / (g - 1)], -1)
    sign = np.array([[1, -1, 1, 1], [1, 1, -1, 1]], dtype=float)
    big = np.concatenate([u, (u * sign[0])[::-1]], 0)
    big = np.concatenate([big, (big * sign[1])[:, ::-1]], 1)
    for tf, tol in [(0.004, 0.0), (0.03, 1e-8)]:
        uh, _, st, rc = cc.oracle_solve("euler2d", u, tf, [1, 1], 2, 3, reflect=sign.ravel())
        ub, _, st2, rc2 = cc.oracle_solve("euler2d", big, tf, [2, 2], 1, 3)
        assert st.steps == st2.steps and rc == 0 and rc2 == 0
        # one step: identical bits.  Later steps: the mirrored half sums in
        # mirrored order, and the finite-difference wave speeds amplify that
        ref = ub[:nx, :ny]
        assert np.max(np.abs(uh - ref)) <= tol * np.max(np.abs(ref)), tf


@pytest.mark.parametrize("N,ndim,tf,lo", [(3, 1, 0.2, 2.9), (4, 1, 0.2, 3.7), (3, 2, 0.1, 2.9)])
def test_convergence_order_on_smooth_wave(N, ndim, tf, lo):
    """Empirical order of the ADER-WENO scheme on an advected density wave
    (exact cell averages at start and end): the L1 error of rho falls like
    dx^N between 16, 32 and 64 cells per side."""
    errs = []
    for n in (16, 32, 64):
        u0 = systems.density_wave(n, ndim)
        uf, _, st, rc = cc.oracle_solve("euler1d" if ndim == 1 else "euler2d", u0, tf,
                                        [1.0] * ndim, 1, N)
        assert rc == 0
        ex = systems.density_wave(n, ndim, t=tf)
        errs.append(np.mean(np.abs(uf[..., 0] - ex[..., 0])))
    orders = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print(f"N={N} ndim={ndim}: L1 errors {errs}, orders {orders}")
    assert orders[-1] >= lo


def test_newton_iterations_against_guess_accuracy():
    """The stiff predictor from supplied initial guesses (the hook of
    dg.cpp:216) on a reactive channel state: the converged solution as guess
    takes one iteration; guesses within 1e-4 (relative, in qh - w) take far
    fewer iterations than the plain guess; 1e-2 noise takes more.  Every run
    converges to the same solution to the Newton tolerance."""
    N, V, nx, ny = 3, 5, 48, 12
    s = cc.system("reactive2d")
    O = cc.oracle()
    refl = systems.REACTIVE_REFLECT
    u0 = systems.reactive_channel(nx, ny, seed=3)
    u, _, st, rc = cc.oracle_solve("reactive2d", u0, 1.0, [nx / 1024, ny / 1024], 2, N,
                                   stiff=True, max_steps=8, reflect=refl)
    nX = np.array([nx, ny], dtype=np.int32)
    nb = nX + 2 * N
    dX = np.array([1 / 1024, 1 / 1024])
    ub = np.zeros(int(np.prod(nb)) * V)
    r = np.ascontiguousarray(refl.ravel())
    O.oracle_boundaries(cc.dp(ub), cc.dp(np.ascontiguousarray(u)), cc.ip(nX), 2,
                        cc.ip(np.array([2, 2], dtype=np.int32)), cc.dp(r), N, V)
    nw = int(np.prod(nX + 2))
    wh = np.zeros(nw * N * N * V)
    O.oracle_weno(cc.dp(wh), cc.dp(ub), cc.ip(np.ascontiguousarray(nb, dtype=np.int32)), 2, N, V)
    dt = O.oracle_timestep(s.F, None, cc.dp(wh), nw, cc.dp(dX), 2, N, V, 0.9, 1.0, 0.0, 8)

    def solve(guess=None):
        qh = np.zeros(nw * N ** 3 * V)
        it = np.zeros(nw, dtype=np.int32)
        o = cc.OracleOpts()
        o.newton_tol = 1e-12
        o.newton_maxit = 100
        if guess is not None:
            g = np.ascontiguousarray(guess.ravel())
            o.guess = cc.dp(g)
            o.guess_check = 1
        O.oracle_predictor(s.F, None, s.S, s.M, cc.dp(wh), nw, cc.dp(dX), 2, N, V, dt, 1,
                           C.byref(o), cc.dp(qh), it.ctypes.data_as(C.POINTER(C.c_int)))
        return qh.reshape(nw, N ** 3, V), it

    qh, itp = solve()
    wt = np.tile(wh.reshape(nw, N * N, V), (1, N, 1))
    d = qh - wt
    rng = np.random.default_rng(0)
    _, it_exact = solve(qh.copy())
    q4, it4 = solve(wt + d * (1 + 1e-4 * rng.uniform(-1, 1, d.shape)))
    q2, it2 = solve(wt + d * (1 + 1e-2 * rng.uniform(-1, 1, d.shape)))
    print("mean iterations: plain", itp.mean(), "exact", it_exact.mean(), "1e-4", it4.mean(),
          "1e-2", it2.mean())
    assert np.all(it_exact == 1)
    assert it4.mean() < 0.6 * itp.mean()
    assert it2.mean() > itp.mean()
    scale = np.max(np.abs(qh), axis=(0, 1))
    for q in (q4, q2):
        assert np.max(np.abs(q - qh) / scale) <= 1e-10


def oracle_sod_flux(u, flux):
    """Oracle pde_solver on 1D Sod (200 cells, order 3, tf 0.2) with a flux code."""
    s = cc.system("euler1d")
    uo = np.ascontiguousarray(u).copy()
    nX = np.array([200], dtype=np.int32)
    dX = np.array([1 / 200])
    ret = np.zeros(u.size)
    o = cc.OracleOpts()
    st = cc.OracleStats()
    rc = cc.oracle().oracle_pde_solver(s.F, None, None, None, 1, 0, 0, cc.dp(uo), 0.2,
                                       cc.ip(nX), 1, cc.dp(dX), 0.9,
                                       cc.ip(np.array([0], dtype=np.int32)), 0, flux, 3, 3, 1, 0,
                                       cc.dp(ret), C.byref(o), C.byref(st))
    assert rc == 0
    return uo


def ref_sod_flux(u, flux):
    """The reference's pde_solver on the same problem (stdout suppressed)."""
    s = cc.system("euler1d")
    ur = np.ascontiguousarray(u).copy()
    nX = np.array([200], dtype=np.int32)
    dX = np.array([1 / 200])
    ret = np.zeros(u.size)
    with cc.quiet_stdout():
        cc.ref().pde_solver(s.F, None, None, True, False, False, cc.dp(ur), 0.2, cc.ip(nX), 1,
                            cc.dp(dX), 0.9, cc.ip(np.array([0], dtype=np.int32)), False, flux, 3,
                            3, 1, False, cc.dp(ret), 4)
    return ur


def abs_apply_numpy(A, x):
    w, R = np.linalg.eig(A)
    b = np.linalg.solve(R, x.astype(complex)) * np.abs(w)
    return (R @ b).real


def test_abs_apply_against_numpy():
    """|A| x = R |Lambda| R^-1 x (the Roe and Osher dissipation,
    fluxes.cpp:20-70) against numpy's complex eigendecomposition: real, complex,
    clustered and badly scaled spectra, n = 1..8; error relative to cond(R)."""
    O = cc.oracle()
    O.oracle_abs_apply.argtypes = [cc.PD, cc.PD, C.c_int, cc.PD]
    rng = np.random.default_rng(0)
    for trial in range(800):
        n = int(rng.integers(1, 9))
        kind = trial % 4
        if kind == 0:
            A = rng.standard_normal((n, n))
        elif kind == 1:
            R = rng.standard_normal((n, n))
            A = R @ np.diag(rng.uniform(-2, 2, n)) @ np.linalg.inv(R)
        elif kind == 2:
            R = rng.standard_normal((n, n))
            lam = np.full(n, 0.3) + 1e-8 * rng.standard_normal(n)
            lam[0] = 2.0
            if n > 1:
                lam[-1] = -1.5
            A = R @ np.diag(lam) @ np.linalg.inv(R)
        else:
            A = rng.standard_normal((n, n)) * rng.uniform(1e-3, 1e3)
        x = rng.standard_normal(n)
        y = np.zeros(n)
        O.oracle_abs_apply(cc.dp(np.ascontiguousarray(A.ravel())), cc.dp(x), n, cc.dp(y))
        want = abs_apply_numpy(A, x)
        cond = max(np.linalg.cond(np.linalg.eig(A)[1]), 1.0)
        assert np.max(np.abs(y - want)) <= 1e-13 * cond * max(np.max(np.abs(want)), 1e-300), trial


def sod_flux_noise(flux, ur):
    """The reference's own movement on Sod with flux `flux` under 1-ulp input
    changes: one entry at a time at four places across the domain, and every
    entry at random (two draws).  A single entry can flip a Picard iteration
    count somewhere (Osher: 3.6e-4 from entry (150, 2)), which dwarfs the
    smooth part of the sensitivity (about 2e-6)."""
    u = sod1d()
    noises = []
    for idx in ((37, 0), (99, 2), (150, 2), (175, 1)):
        u2 = u.copy()
        u2[idx] = np.nextafter(u2[idx], np.inf)
        noises.append(relerr(ref_sod_flux(u2, flux), ur))
    rng = np.random.default_rng(1)
    for _ in range(2):
        u2 = u.copy()
        m = rng.random(u.shape) < 0.5
        u2[m] = np.nextafter(u2[m], np.inf)
        noises.append(relerr(ref_sod_flux(u2, flux), ur))
    return noises


@pytest.mark.ref
@pytest.mark.parametrize("flux", [1, 2])
def test_roe_osher_reference_within_its_own_noise(flux):
    """Roe (flux 1) and Osher (flux 2) on 1D Sod against the reference, within
    3x the reference's own movement under 1-ulp input changes."""
    u = sod1d()
    uo = oracle_sod_flux(u, flux)
    ur = ref_sod_flux(u, flux)
    noises = sod_flux_noise(flux, ur)
    print("flux", flux, "oracle vs reference", relerr(uo, ur), "reference noise", noises)
    assert relerr(uo, ur) <= 3 * max(noises)


@pytest.mark.parametrize("flux,name", [(1, "roe"), (2, "osher")])
def test_roe_osher_golden(flux, name):
    g = load_gold("solves")
    uo = oracle_sod_flux(sod1d(), flux)
    assert relerr(uo, g[f"sod_{name}_final"]) <= 3 * float(g[f"sod_{name}_noise"])


def advdiff_init(n):
    x = (np.arange(n) + 0.5) / n
    return (1 + 0.5 * np.sin(2 * np.pi * x))[:, None]


def oracle_viscous(name, u0, tf, L, bc, N, flux=0):
    """Oracle pde_solver with secondOrder (F depends on dQ)."""
    s = cc.system(name)
    u = np.ascontiguousarray(u0).copy()
    nX = np.array(u.shape[:-1], dtype=np.int32)
    nd, V = len(nX), u.shape[-1]
    dX = np.array([L[i] / nX[i] for i in range(nd)])
    ret = np.zeros(u.size)
    o = cc.OracleOpts()
    st = cc.OracleStats()
    rc = cc.oracle().oracle_pde_solver(s.F, None, None, None, 1, 0, 0, cc.dp(u), tf, cc.ip(nX),
                                       nd, cc.dp(dX), 0.9,
                                       cc.ip(np.array([bc] * nd, dtype=np.int32)), 0, flux, N, V,
                                       1, 1, cc.dp(ret), C.byref(o), C.byref(st))
    assert rc == 0
    return u, st


def ref_viscous(name, u0, tf, L, bc, N, flux=0):
    """The reference's pde_solver with secondOrder = true."""
    s = cc.system(name)
    u = np.ascontiguousarray(u0).copy()
    nX = np.array(u.shape[:-1], dtype=np.int32)
    nd, V = len(nX), u.shape[-1]
    dX = np.array([L[i] / nX[i] for i in range(nd)])
    ret = np.zeros(u.size)
    with cc.quiet_stdout():
        cc.ref().pde_solver(s.F, None, None, True, False, False, cc.dp(u), tf, cc.ip(nX), nd,
                            cc.dp(dX), 0.9, cc.ip(np.array([bc] * nd, dtype=np.int32)), False,
                            flux, N, V, 1, True, cc.dp(ret), 4)
    return u


@pytest.mark.ref
def test_viscous_path_reference():
    """Second-order (viscous) F: the gradient traces at faces (fv.cpp:87-122),
    the viscous wave speed in dt and in the flux (system.cpp:53-61,
    stepper.cpp:58-59, fluxes.cpp:96-100) against the reference."""
    uo, _ = oracle_viscous("advdiff", advdiff_init(64), 0.2, [1.0], 1, 3)
    ur = ref_viscous("advdiff", advdiff_init(64), 0.2, [1.0], 1, 3)
    assert relerr(uo, ur) <= 1e-10
    u0 = systems.euler2d_modes(16, seed=2)
    uo, _ = oracle_viscous("ns2d", u0, 0.02, [1.0, 1.0], 1, 3)
    ur = ref_viscous("ns2d", u0, 0.02, [1.0, 1.0], 1, 3)
    assert relerr(uo, ur) <= 1e-8


def test_viscous_golden():
    g = load_gold("solves")
    uo, _ = oracle_viscous("advdiff", advdiff_init(64), 0.2, [1.0], 1, 3)
    assert relerr(uo, g["advdiff_final"]) <= 1e-10
    uo, _ = oracle_viscous("ns2d", systems.euler2d_modes(16, seed=2), 0.02, [1.0, 1.0], 1, 3)
    assert relerr(uo, g["ns2d_final"]) <= 1e-8


def test_advection_diffusion_converges_to_exact():
    """Viscous path against the exact solution of advection-diffusion, u = 1 +
    0.5 sin(2 pi (x - t)) exp(-nu (2 pi)^2 t) (cell averages).  At order 3 the
    reference's viscous treatment (gradient traces plus the 2 (N + 1) nu / dX
    penalty, fluxes.cpp:96-100) converges at about second order; the reference
    gives the same errors (L1 4.04e-3, 8.63e-4, 1.95e-4 at 16, 32, 64 cells)."""
    nu, tf = 0.02, 0.2
    errs = []
    for n in (16, 32, 64):
        e = np.arange(n + 1) / n - tf
        k = 2 * np.pi
        avg = lambda ee: 1 + 0.5 * (-(np.cos(k * ee[1:]) - np.cos(k * ee[:-1])) / k * n)  # noqa
        u0 = avg(np.arange(n + 1) / n)[:, None]
        exact = 1 + (avg(e) - 1) * np.exp(-nu * k * k * tf)
        uo, _ = oracle_viscous("advdiff", u0, tf, [1.0], 1, 3)
        errs.append(np.mean(np.abs(uo[:, 0] - exact)))
    orders = [np.log2(errs[i] / errs[i + 1]) for i in range(2)]
    print("advection-diffusion L1 errors", errs, "orders", orders)
    assert np.allclose(errs, [4.039684724e-3, 8.632007961e-4, 1.949539833e-4], rtol=1e-6)
    assert min(orders) >= 2.0


# ------------------------------------------------------------ 17-variable GPR

GPR_TF = 0.005  # 44 steps of the 200-cell Riemann problem at CFL 0.6


def gpr_relerr(a, b):
    """relerr with each variable's scale floored at 1e-8 of the largest: in
    this 1D problem the transverse momenta and most of A and J stay at zero
    or at 1e-18 rounding residue."""
    V = b.shape[-1]
    sc = np.max(np.abs(b.reshape(-1, V)), axis=0)
    sc = np.maximum(sc, 1e-8 * sc.max())
    return float(np.max(np.abs((a - b).reshape(-1, V)) / sc))


def oracle_gpr(stiff):
    """GPR Riemann problem (1D, 200 cells, order 2, CFL 0.6, transmissive,
    Rusanov) as the reference's GPR example runs it (pypde/tests/gpr/
    __init__.py:32-36: F, B and S; order and flux at pde_solver's defaults),
    with B.dQ evaluated the reference's way (dg.cpp:110, fv.cpp:234)."""
    uo, _, st, rc = cc.oracle_solve("gpr", systems.gpr_riemann(200), GPR_TF, [1.0], 0, 2,
                                    stiff=stiff, cfl=0.6, b_quirk=1)
    assert rc == 0 and st.steps == 44
    return uo


@pytest.mark.ref
def test_gpr_picard_reference():
    """V = 17 with B != 0 and S (system.py:8-75): with the Picard predictor the
    oracle and the reference run the same iteration, and agree to rounding
    (3e-10 measured; the eigenvalue path differs: QR / polynomial here,
    Spectra's Arnoldi there, system.cpp:37-42)."""
    ur, _ = cc.ref_solve("gpr", systems.gpr_riemann(200), GPR_TF, [1.0], 0, 2, stiff=False,
                         cfl=0.6, nthreads=8)
    assert gpr_relerr(oracle_gpr(False), ur) <= 1e-8


def test_gpr_golden():
    """Against the reference's end states (tests/golden/make_golden.py).  The
    stiff predictors differ by design: the reference's Newton-Krylov stops at
    an absolute residual of 6e-6 (newton_krylov.cpp:54-76), the oracle's
    per-node Newton at 1e-12, which over 44 steps moves the state by 2.5e-5
    (relative L-inf); the bound is 1e-4."""
    g = load_gold("solves")
    assert gpr_relerr(oracle_gpr(False), g["gpr_picard_final"]) <= 1e-8
    assert gpr_relerr(oracle_gpr(True), g["gpr_stiff_final"]) <= 1e-4


def test_gpr_b_quirk_matters():
    """The reference's B.dQ evaluation changes the GPR solution at the percent
    level (5 % in rho u at 44 steps), so the comparison above depends on it."""
    g = load_gold("solves")
    uo, _, _, _ = cc.oracle_solve("gpr", systems.gpr_riemann(200), GPR_TF, [1.0], 0, 2,
                                  stiff=False, cfl=0.6, b_quirk=0)
    assert gpr_relerr(uo, g["gpr_picard_final"]) > 1e-3
