"""Seeded generator of RBF-FD shallow-water inputs in the reference's binary
format (swe_code/io/input.c:189-414, written in the order of
swe_code/io/nc2bin.c:32-80).

The reference ships one input (icos2562_tc5_input.bin) built offline by a
MATLAB generator that is not in its tree.  This module rebuilds that
construction so that any mesh size can be produced:

* nodes: icosahedron refined by recursive edge bisection, points pushed to
  the unit sphere; N = 10*4**level + 2 (level 4 -> 2562 ... level 8 ->
  655362);
* stencils: the n = 31 nearest nodes in Euclidean distance, the node itself
  first (the RCM ordering relies on that, swe_code/rcm/rcm.c:37, :115);
* weights: Gaussian RBF phi(r) = exp(-(eps r)^2) augmented with a constant,
  eps = 0.035*sqrt(N) - 0.1, for the sphere-projected gradient
  P grad (rows Dx, Dy, Dz) and the 4th power of the tangent-plane Laplacian
  (L), unit-sphere units (the reader scales D by 1/a and L by gamma).  On the
  shipped 2562 mesh these reproduce the shipped weights to ~1e-5;
* state: Williamson test case 5 style solid-body zonal flow u0, geostrophic
  h = -(a*Omega*u0 + u0^2/2) z^2, mean depth gh0 = g*depth and a cosine-bell
  mountain ghm, gradghm = D ghm / a.

The weights are computed in float64 with numpy/scipy (host tooling, not the
hot path).  LAPACK results can differ in the last bits between machines, so
bitwise comparisons are always made on one machine's bytes.
"""
from __future__ import annotations

import hashlib
import math
import os
import struct
from dataclasses import dataclass, field

import numpy as np

A_EARTH = 6371220.0
OMEGA = 7.292e-5
GRAV = 9.80616
NNBR = 31

_LEVEL_OF_N = {10 * 4 ** k + 2: k for k in range(0, 10)}


def level_of(n: int) -> int:
    if n not in _LEVEL_OF_N:
        raise ValueError(f"N={n} is not an icosahedral node count 10*4**k+2")
    return _LEVEL_OF_N[n]


def icosahedral_nodes(level: int) -> np.ndarray:
    """Unit vectors of the level-`level` bisected icosahedron, (N, 3).

    The base icosahedron has the orientation and first twelve nodes of the
    shipped 2562 input, (0, +-t, +-1), (+-1, 0, +-t), (+-t, +-1, 0)
    normalised (t the golden ratio); the refined node set equals that file's
    node set."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[0, t, 1], [0, -t, 1], [0, t, -1], [0, -t, -1],
                  [1, 0, t], [-1, 0, t], [1, 0, -t], [-1, 0, -t],
                  [t, 1, 0], [-t, 1, 0], [t, -1, 0], [-t, -1, 0]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    edge = np.linalg.norm(v[0] - v[4])
    d = np.linalg.norm(v[:, None, :] - v[None, :, :], axis=2)
    adj = np.abs(d - edge) < 1e-9
    f = np.array([(a, b, c) for a in range(12) for b in range(a + 1, 12)
                  for c in range(b + 1, 12) if adj[a, b] and adj[b, c] and adj[a, c]],
                 dtype=np.int64)
    for _ in range(level):
        nv = len(v)
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * nv + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // nv, uk % nv
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid])
        nf = len(f)
        m01, m12, m20 = (nv + inv[:nf], nv + inv[nf:2 * nf], nv + inv[2 * nf:])
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1), np.stack([m01, f[:, 1], m12], 1),
            np.stack([m20, m12, f[:, 2]], 1), np.stack([m01, m12, m20], 1)])
    return np.ascontiguousarray(v)


def stencils(X: np.ndarray, n: int = NNBR) -> np.ndarray:
    """n nearest nodes of every node, nearest first (self at column 0)."""
    from scipy.spatial import cKDTree
    _, ii = cKDTree(X).query(X, k=n)
    ii = ii.astype(np.int32)
    self_col = np.arange(len(X), dtype=np.int32)
    # Coincident points never occur on these meshes, so column 0 is self;
    # enforce it for safety.
    bad = ii[:, 0] != self_col
    if bad.any():
        raise ValueError("stencil search did not put the node itself first")
    return np.ascontiguousarray(ii)


def shape_parameter(n_nodes: int) -> float:
    return 0.035 * math.sqrt(n_nodes) - 0.1


def _laguerre(k: int, t: np.ndarray) -> np.ndarray:
    from scipy.special import eval_laguerre
    return eval_laguerre(k, t)


def rbffd_weights(X: np.ndarray, idx: np.ndarray, eps: float, k: int = 4,
                  chunk: int = 256) -> tuple[np.ndarray, ...]:
    """Dx, Dy, Dz (projected gradient) and L (Laplacian**k) weights,
    unit sphere, each (N, n).

    For node i with stencil points x_j and r = |x - x_j|:
      P grad phi(x_i) = phi'(r)/r * (x_i (x_i . x_j) - x_j),  phi'(r)/r = -2 eps^2 phi
      Lap^k phi       = (-4 eps^2)^k k! Lag_k(eps^2 r^2) phi        (2-D Gaussian)
    and the weights solve [[A, 1], [1^T, 0]] [w; mu] = [b; 0].
    """
    N, n = idx.shape
    out = np.empty((4, N, n), dtype=np.float64)
    kfac = ((-4.0 * eps * eps) ** k) * math.factorial(k)
    for s in range(0, N, chunk):
        e = min(N, s + chunk)
        P = X[idx[s:e]]                                   # (c, n, 3)
        xi = X[s:e]                                       # (c, 3)
        diff = P[:, :, None, :] - P[:, None, :, :]
        R2 = np.einsum('cjkd,cjkd->cjk', diff, diff)
        c = e - s
        A = np.ones((c, n + 1, n + 1))
        A[:, :n, :n] = np.exp(-eps * eps * R2)
        A[:, n, n] = 0.0
        dx = P - xi[:, None, :]
        r2 = np.einsum('cjd,cjd->cj', dx, dx)
        phi = np.exp(-eps * eps * r2)
        g = -2.0 * eps * eps * phi
        proj = xi[:, None, :] * np.einsum('cjd,cd->cj', P, xi)[:, :, None] - P
        B = np.zeros((c, n + 1, 4))
        B[:, :n, 0] = g * proj[:, :, 0]
        B[:, :n, 1] = g * proj[:, :, 1]
        B[:, :n, 2] = g * proj[:, :, 2]
        B[:, :n, 3] = kfac * _laguerre(k, eps * eps * r2) * phi
        W = np.linalg.solve(A, B)[:, :n, :]
        out[:, s:e, :] = np.moveaxis(W, 2, 0)
    return out[0], out[1], out[2], out[3]


def projection_rows(X: np.ndarray) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Rows of P = I - x x^T: p_u, p_v, p_w, each (N, 3)."""
    P = np.eye(3)[None, :, :] - X[:, :, None] * X[:, None, :]
    return (np.ascontiguousarray(P[:, 0, :]), np.ascontiguousarray(P[:, 1, :]),
            np.ascontiguousarray(P[:, 2, :]))


def default_dt(n_nodes: int) -> float:
    """900 s at N=2562, halved with every refinement level (Courant-scaled)."""
    return 900.0 * 2.0 ** (4 - level_of(n_nodes))


# Hyperviscosity strength: gamma is set so that dt * rho(gamma * L) = KAPPA,
# where rho is the spectral radius of the assembled L (the grid-scale decay
# rate).  The shipped 2562 file has gamma = -0.05 N**-4, i.e. dt*rho ~ 0.014,
# too weak to hold its run past ~300 steps; with the shape parameter above,
# rho(L) does not scale exactly as N**4, so a fixed gamma*N**4 is unstable
# at some N.  RK4's stability interval on the negative real axis is 2.785.
DEFAULT_KAPPA = 1.5


def spectral_radius(W: np.ndarray, idx: np.ndarray, iters: int = 300) -> float:
    """Estimate of the largest |eigenvalue| of the sparse matrix with rows W
    at columns idx: geometric-mean growth of |M^k v| over the last third of
    `iters` power iterations from a seeded start vector."""
    import scipy.sparse as sp
    N, n = idx.shape
    M = sp.csr_matrix((W.ravel(), (np.repeat(np.arange(N), n), idx.ravel())), shape=(N, N))
    v = np.random.default_rng(12345).standard_normal(N)
    v /= np.linalg.norm(v)
    logs = []
    for _ in range(iters):
        w = M @ v
        nrm = float(np.linalg.norm(w))
        logs.append(math.log(nrm))
        v = w / nrm
    tail = logs[2 * iters // 3:]
    return math.exp(sum(tail) / len(tail))


@dataclass
class SweCase:
    n_nodes: int
    nnbr: int
    a: float
    gh0: float
    dt: float
    gamma: float
    X: np.ndarray            # (N, 3)
    f: np.ndarray
    ghm: np.ndarray
    p_u: np.ndarray
    p_v: np.ndarray
    p_w: np.ndarray
    gradghm: np.ndarray      # (N, 3)
    Dx: np.ndarray           # (N, n) unit sphere
    Dy: np.ndarray
    Dz: np.ndarray
    L: np.ndarray
    idx: np.ndarray          # (N, n) 0-based
    H_init: np.ndarray       # (N, 4)
    H_final: np.ndarray      # (N, 4) known answer slot (H_100)
    params: dict = field(default_factory=dict)


def tc5_state(X: np.ndarray, u0: float, a: float = A_EARTH) -> np.ndarray:
    """Solid-body zonal flow and its balanced geopotential perturbation."""
    H = np.zeros((len(X), 4))
    H[:, 0] = -u0 * X[:, 1]
    H[:, 1] = u0 * X[:, 0]
    H[:, 3] = -(a * OMEGA * u0 + 0.5 * u0 * u0) * X[:, 2] ** 2
    return H


def cosine_bell(X: np.ndarray, lat0: float, lon0: float, height: float,
                radius: float = math.pi / 9.0) -> np.ndarray:
    c = np.array([math.cos(lat0) * math.cos(lon0), math.cos(lat0) * math.sin(lon0),
                  math.sin(lat0)])
    r = np.arccos(np.clip(X @ c, -1.0, 1.0))
    hm = np.where(r < radius, 0.5 * height * (1.0 + np.cos(math.pi * r / radius)), 0.0)
    return hm


def make_case(n_nodes: int, seed: int = 1, dt: float | None = None,
              kappa: float = DEFAULT_KAPPA, nnbr: int = NNBR) -> SweCase:
    rng = np.random.default_rng(seed)
    u0 = float(rng.uniform(15.0, 25.0))
    depth = float(rng.uniform(5000.0, 6000.0))
    hill = float(rng.uniform(1000.0, 2500.0))
    lat0 = float(rng.uniform(math.pi / 12, math.pi / 4))
    lon0 = float(rng.uniform(0.0, 2.0 * math.pi))
    X = icosahedral_nodes(level_of(n_nodes))
    idx = stencils(X, nnbr)
    eps = shape_parameter(n_nodes)
    Dx, Dy, Dz, L = rbffd_weights(X, idx, eps)
    p_u, p_v, p_w = projection_rows(X)
    ghm = GRAV * cosine_bell(X, lat0, lon0, hill)
    nb = ghm[idx]
    gradghm = np.stack([(Dx * nb).sum(1), (Dy * nb).sum(1), (Dz * nb).sum(1)], 1) / A_EARTH
    H0 = tc5_state(X, u0)
    if dt is None:
        dt = default_dt(n_nodes)
    rho = spectral_radius(L, idx)
    gamma = -kappa / (dt * rho)
    params = dict(seed=seed, u0=u0, depth=depth, hill=hill, lat0=lat0, lon0=lon0,
                  eps=eps, dt=dt, kappa=kappa, rho_L=rho)
    return SweCase(n_nodes=n_nodes, nnbr=nnbr, a=A_EARTH, gh0=GRAV * depth, dt=float(dt),
                   gamma=gamma, X=X, f=2.0 * OMEGA * X[:, 2], ghm=ghm, p_u=p_u, p_v=p_v,
                   p_w=p_w, gradghm=gradghm, Dx=Dx, Dy=Dy, Dz=Dz, L=L, idx=idx,
                   H_init=H0, H_final=np.zeros_like(H0), params=params)


def perturb_members(case: SweCase, members: int, seed: int,
                    sigma: float = 1e-3) -> np.ndarray:
    """Ensemble initial states (members, N, 4): h and (u,v,w) scaled by
    (1 + sigma*xi), xi ~ N(0,1) per node and member, velocity re-projected
    onto the tangent plane."""
    rng = np.random.default_rng(seed)
    X = case.X
    out = np.empty((members,) + case.H_init.shape)
    for m in range(members):
        H = case.H_init.copy()
        H[:, :3] *= 1.0 + sigma * rng.standard_normal((len(X), 1))
        H[:, 3] *= 1.0 + sigma * rng.standard_normal(len(X))
        vel = H[:, :3]
        H[:, :3] = vel - X * np.einsum('nd,nd->n', vel, X)[:, None]
        out[m] = H
    return out


# ---------------------------------------------------------------- file I/O --

def write_bin(path: str, case: SweCase) -> None:
    N, n = case.idx.shape
    with open(path + ".tmp%d" % os.getpid(), "wb") as fp:
        fp.write(struct.pack("<ii", N, n))
        fp.write(struct.pack("<4d", case.a, case.gh0, case.dt, case.gamma))
        for arr in (case.X[:, 0], case.X[:, 1], case.X[:, 2], case.f, case.ghm):
            fp.write(np.ascontiguousarray(arr, dtype="<f8").tobytes())
        for arr in (case.p_u, case.p_v, case.p_w, case.gradghm):
            fp.write(np.ascontiguousarray(arr, dtype="<f8").tobytes())
        for arr in (case.Dx, case.Dy, case.Dz, case.L):
            fp.write(np.ascontiguousarray(arr, dtype="<f8").tobytes())
        fp.write(np.ascontiguousarray(case.idx + 1, dtype="<i4").tobytes())
        fp.write(np.ascontiguousarray(case.H_init, dtype="<f8").tobytes())
        fp.write(np.ascontiguousarray(case.H_final, dtype="<f8").tobytes())
    os.replace(path + ".tmp%d" % os.getpid(), path)


def read_bin(path: str) -> SweCase:
    """Raw (unscaled, 0-based idx) contents of an input file."""
    b = np.fromfile(path, dtype=np.uint8)
    N, n = (int(v) for v in b[:8].view("<i4"))
    off = 8
    c = b[off:off + 32].view("<f8")
    off += 32

    def f64(cnt):
        nonlocal off
        a = b[off:off + 8 * cnt].view("<f8")
        off += 8 * cnt
        return a.copy()
    x, y, z, f, ghm = (f64(N) for _ in range(5))
    p_u, p_v, p_w, gg = (f64(3 * N).reshape(N, 3) for _ in range(4))
    Dx, Dy, Dz, L = (f64(N * n).reshape(N, n) for _ in range(4))
    idx = b[off:off + 4 * N * n].view("<i4").reshape(N, n) - 1
    off += 4 * N * n
    H0 = f64(4 * N).reshape(N, 4)
    H1 = f64(4 * N).reshape(N, 4)
    return SweCase(n_nodes=N, nnbr=n, a=float(c[0]), gh0=float(c[1]), dt=float(c[2]),
                   gamma=float(c[3]), X=np.stack([x, y, z], 1), f=f, ghm=ghm, p_u=p_u,
                   p_v=p_v, p_w=p_w, gradghm=gg, Dx=Dx, Dy=Dy, Dz=Dz, L=L,
                   idx=idx.astype(np.int32), H_init=H0, H_final=H1)


def cache_dir() -> str:
    d = os.environ.get("SWE_CACHE_DIR") or os.path.join(
        os.environ.get("TMPDIR", "/tmp"), "swe_cases")
    os.makedirs(d, exist_ok=True)
    return d


def case_path(n_nodes: int, seed: int = 1, dt: float | None = None,
              kappa: float = DEFAULT_KAPPA) -> str:
    """Generate (once, cached on disk) and return the input file path."""
    if dt is None:
        dt = default_dt(n_nodes)
    tag = hashlib.sha1(f"v2:{n_nodes}:{seed}:{dt!r}:{kappa!r}".encode()).hexdigest()[:12]
    path = os.path.join(cache_dir(), f"icos{n_nodes}_s{seed}_{tag}.bin")
    if not os.path.exists(path):
        write_bin(path, make_case(n_nodes, seed=seed, dt=dt, kappa=kappa))
    return path


def read_history(path: str):
    """History file of swe_run (SWE_OUTPUT_FILE / SWE_HIST_INT): returns
    (steps, times, states[nsnap][N][4]) in the file node order."""
    b = np.fromfile(path, dtype=np.uint8)
    N = int(b[:4].view("<i4")[0])
    rec = 4 + 8 + 32 * N
    nsnap = (len(b) - 4) // rec
    steps, times, states = [], [], []
    for k in range(nsnap):
        o = 4 + k * rec
        steps.append(int(b[o:o + 4].view("<i4")[0]))
        times.append(float(b[o + 4:o + 12].view("<f8")[0]))
        states.append(b[o + 12:o + rec].view("<f8").reshape(N, 4).copy())
    return steps, times, np.array(states)
