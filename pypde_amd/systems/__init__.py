"""PDE systems shipped as single-source HIP/C text (*.pde), plus the initial
conditions of the workloads this package is measured on.

A system source defines F, B, S and optionally M (see pde_prelude.h).  The
same text is what a user hands to ``pde_solver`` for their own physics.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# name: (V, ndim it is written for, has F, has B, has S, has M)
SYSTEMS = {
    "euler1d": dict(V=3, F=True, B=False, S=False, M=False),
    "euler2d": dict(V=4, F=True, B=False, S=False, M=False),
    "reactive2d": dict(V=5, F=True, B=False, S=True, M=True),
    "advect": dict(V=1, F=True, B=False, S=False, M=False),
    "ncadvect": dict(V=2, F=False, B=True, S=True, M=False),
    "advdiff": dict(V=1, F=True, B=False, S=False, M=False, second=True),
    "ns2d": dict(V=4, F=True, B=False, S=False, M=False, second=True),
    "gpr": dict(V=17, F=True, B=True, S=True, M=False),
}


def source(name):
    with open(os.path.join(HERE, name + ".pde")) as f:
        return f.read()


def info(name):
    return dict(SYSTEMS[name])


def standard_configs(full=False):
    """(system, ndim, N, stiff) combinations used by the tests, smoke and bench."""
    combos = [("euler1d", 1, 3, False), ("euler2d", 2, 2, False), ("euler2d", 2, 3, False),
              ("euler2d", 2, 4, False), ("reactive2d", 2, 3, True)]
    if full:
        combos += [("euler1d", 1, 2, False), ("euler1d", 1, 1, False), ("euler1d", 1, 6, False),
                   ("advect", 1, 3, False), ("advect", 2, 3, False),
                   ("ncadvect", 2, 3, False), ("reactive2d", 2, 3, False),
                   ("euler2d", 2, 3, True), ("ncadvect", 2, 3, True)]
    out = []
    for name, ndim, N, stiff in combos:
        s = SYSTEMS[name]
        out.append(dict(src=source(name), useF=s["F"], useB=s["B"], useS=s["S"], ndim=ndim, N=N,
                        V=s["V"], stiff=stiff, use_M=s["M"], flux=0))
    # Roe (1) and Osher (2) fluxes
    for name, ndim, N in ([("euler1d", 1, 3), ("euler2d", 2, 3)] if full else []):
        s = SYSTEMS[name]
        for flux in (1, 2):
            out.append(dict(src=source(name), useF=s["F"], useB=s["B"], useS=s["S"], ndim=ndim,
                            N=N, V=s["V"], stiff=False, use_M=s["M"], flux=flux))
    # viscous (second-order) systems
    for name, ndim, N in ([("advdiff", 1, 3), ("ns2d", 2, 3)] if full else []):
        s = SYSTEMS[name]
        out.append(dict(src=source(name), useF=s["F"], useB=s["B"], useS=s["S"], ndim=ndim, N=N,
                        V=s["V"], stiff=False, use_M=s["M"], flux=0, second=True))
    # mixed precision (FP32 predictor phase): the reactive channel, the 2D Euler
    # configurations and Sod
    for name, ndim, N, stiff in ([("reactive2d", 2, 3, True), ("euler2d", 2, 3, False),
                                  ("euler2d", 2, 4, False), ("euler1d", 1, 3, False)]
                                 if full else []):
        s = SYSTEMS[name]
        out.append(dict(src=source(name), useF=s["F"], useB=s["B"], useS=s["S"], ndim=ndim, N=N,
                        V=s["V"], stiff=stiff, use_M=s["M"], flux=0, mixed=True))
    # ensemble members (code objects with the batched ensemble kernels): 2D Euler,
    # the stiff reactive channel and 1D Euler, order 3
    for name, ndim, N, stiff in ([("euler2d", 2, 3, False), ("reactive2d", 2, 3, True),
                                  ("euler1d", 1, 3, False)] if full else []):
        s = SYSTEMS[name]
        out.append(dict(src=source(name), useF=s["F"], useB=s["B"], useS=s["S"], ndim=ndim, N=N,
                        V=s["V"], stiff=stiff, use_M=s["M"], flux=0, ensemble=True))
    # 17-variable GPR (1D, order 2, the reference's B.dQ evaluation); each takes
    # a few minutes to compile
    for stiff in ((False, True) if full else ()):
        s = SYSTEMS["gpr"]
        out.append(dict(src=source("gpr"), useF=True, useB=True, useS=True, ndim=1, N=2, V=17,
                        stiff=stiff, use_M=False, flux=0, b_quirk=1))
    return out


# ------------------------------------------------------------ workloads

def sod1d(n=200, gamma=1.4):
    """Reference Sod initial state (pypde/tests/euler/__init__.py:25-44) in 1D:
    (rho, u, p) = (1, 0, 1) left of x = 0.5, (0.125, 0, 0.1) right."""
    u = np.zeros((n, 3))
    for i in range(n):
        r, p = (1.0, 1.0) if i / n < 0.5 else (0.125, 0.1)
        u[i] = [r, 0.0, p / (gamma - 1)]
    return u


def density_wave(n, ndim, t=0.0, amp=0.2, vel=0.5, p=1.0, gamma=1.4):
    """Smooth Euler solution for convergence studies: rho = 1 + amp sin(2 pi
    (sum_d x_d - ndim vel t)) advected at constant velocity vel (every axis)
    and pressure p on the periodic unit square/line.  Returns exact cell
    averages (n^ndim cells) of (rho, rho u_d..., E)."""
    k = 2 * np.pi
    e = (np.arange(n + 1) / n) - vel * t      # shifted cell edges along each axis
    if ndim == 1:
        s = -(np.cos(k * e[1:]) - np.cos(k * e[:-1])) / k * n
    else:
        a, b = e[:, None], e[None, :]
        S = lambda x, y: np.sin(k * (x + y))  # noqa: E731
        s = -(S(a[1:], b[:, 1:]) - S(a[:-1], b[:, 1:]) - S(a[1:], b[:, :-1])
              + S(a[:-1], b[:, :-1])) / k ** 2 * n * n
    rho = 1.0 + amp * s
    comps = [rho] + [vel * rho] * ndim + [p / (gamma - 1) + 0.5 * ndim * vel * vel * rho]
    return np.ascontiguousarray(np.stack(comps, -1))


def splitmix64(seed, count):
    """`count` uniforms in [0, 1) from the SplitMix64 stream of `seed`:
    x_i = seed + (i + 1) * 0x9E3779B97F4A7C15, mixed, top 53 bits scaled."""
    with np.errstate(over="ignore"):
        i = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def reactive_channel(nx=1024, ny=256, seed=0, gamma=1.2, Qc=25.0):
    """2D reactive Euler channel: burnt high-pressure slab (p=20, rho=1.8,
    lambda=0) for x < 0.1 over quiescent unburnt gas (rho=1, p=1, lambda=1),
    density scaled per cell by 1 + 0.05 U(-1, 1) (SplitMix64 stream `seed`).
    Domain [0,1] x [0, ny/nx]; conserved (rho, rho u, rho v, E, rho lambda)."""
    noise = (2.0 * splitmix64(seed, nx * ny) - 1.0).reshape(nx, ny)
    x = (np.arange(nx) + 0.5) / nx
    X = np.broadcast_to(x[:, None], (nx, ny))
    burnt = X < 0.1
    rho = np.where(burnt, 1.8, 1.0) * (1.0 + 0.05 * noise)
    p = np.where(burnt, 20.0, 1.0)
    lam = np.where(burnt, 0.0, 1.0)
    E = p / (gamma - 1.0) + Qc * rho * lam
    z = np.zeros_like(rho)
    return np.ascontiguousarray(np.stack([rho, z, z, E, rho * lam], -1))


def riemann4(nx, ny, seed, gamma=1.4):
    """Four-quadrant 2D Riemann problem on the unit square (x, y split at 1/2):
    per quadrant rho ~ U(0.1, 2), u, v ~ U(-1, 1), p ~ U(0.1, 2) from the
    SplitMix64 stream `seed`.  Conserved (rho, rho u, rho v, E)."""
    r = splitmix64(seed, 16).reshape(4, 4)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    qd = (x[:, None] >= 0.5).astype(int) * 2 + (y[None, :] >= 0.5).astype(int)
    rho = 0.1 + 1.9 * r[qd, 0]
    u = -1 + 2 * r[qd, 1]
    v = -1 + 2 * r[qd, 2]
    p = 0.1 + 1.9 * r[qd, 3]
    E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
    return np.ascontiguousarray(np.stack([rho, rho * u, rho * v, E], -1))


def ensemble_sizes(n=48, sizes=(64, 128, 256, 512), seed=7):
    """n grid shapes (n_x, n_y), each side drawn uniformly from `sizes`."""
    r = splitmix64(seed, 2 * n).reshape(n, 2)
    k = np.minimum((r * len(sizes)).astype(int), len(sizes) - 1)
    return [(sizes[a], sizes[b]) for a, b in k]


def gpr_state(rho, p, v, gamma=1.4):
    """Conserved GPR state for density rho, pressure p, velocity v (3), with
    A = rho^(1/3) I and J = 0 (pypde/tests/gpr/misc/utils.py:26-35): then
    dev(A^T A) = 0 and E = p / ((gamma - 1) rho) + |v|^2 / 2."""
    Q = np.zeros(17)
    Q[0] = rho
    Q[1] = rho * (p / (rho * (gamma - 1)) + np.dot(v, v) / 2)
    Q[2:5] = rho * np.asarray(v, dtype=np.float64)
    Q[5:14] = (rho ** (1 / 3) * np.eye(3)).ravel()
    return Q


def gpr_riemann(n=200, left=(1.0, 1.0, (0.0, 0.0, 0.0)), right=(0.125, 0.1, (0.0, 0.0, 0.0))):
    """1D GPR Riemann problem (pypde/tests/gpr/misc/utils.py:38-50): cells with
    i / n < 1/2 take the left state, the rest the right one."""
    QL, QR = gpr_state(*left), gpr_state(*right)
    return np.ascontiguousarray(np.stack([QL if i / n < 0.5 else QR for i in range(n)]))


REACTIVE_REFLECT = np.array([[1, -1, 1, 1, 1], [1, 1, -1, 1, 1]], dtype=np.float64)


def euler2d_modes(n=512, seed=0, gamma=1.4, amp=0.2, vamp=0.1):
    """2D periodic Euler: rho = 1 + amp * P_rho, u, v = 0.5 + vamp * P_{u,v}, p = 1,
    each P a sum of 8 seeded Fourier modes (wavenumbers 1..4, amplitudes
    U(0,1), phases U(0, 2 pi)) normalised to max |P| = 1."""
    r = iter(splitmix64(seed, 3 * 8 * 4))
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")

    def field():
        P = np.zeros_like(X)
        for _ in range(8):
            kx = 1 + int(next(r) * 4)
            ky = 1 + int(next(r) * 4)
            a = next(r)
            ph = 2 * np.pi * next(r)
            P += a * np.sin(2 * np.pi * (kx * X + ky * Y) + ph)
        return P / np.max(np.abs(P))

    rho = 1 + amp * field()
    u = 0.5 + vamp * field()
    v = 0.5 + vamp * field()
    p = np.ones_like(X)
    E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v)
    return np.ascontiguousarray(np.stack([rho, rho * u, rho * v, E], -1))
