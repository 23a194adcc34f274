"""The oracle against what the mathematics fixes, independent of the
reference build (which pins it bit for bit in test_reference_parity.py):

* one RK4 step of the linear system dH/dt = L H (gradient weights, Coriolis
  and orography switched off) equals the textbook RK4 polynomial
  sum_k (dt L)^k / k! H for k <= 4 (rk4_rbffd_swe.c:30-186: the stage
  constants dt/2, dt/2, dt and the D weights 1, 2, 2, 1 with dt/6);
* the right-hand side (rk4_rbffd_swe.c:189-362) against dense stencil
  operators Dx, Dy, Dz, L built from the stencil lists, on states that
  switch parts of the equations off: zero velocity (pressure gradient and
  hyperviscosity only), zero depth field (advection, Coriolis, continuity),
  and a constant state, whose RHS has a closed form once the weights' zero
  row sums are used (F_uvw = -f r x U, F_h = U . grad(ghm));
* tangency: x . (F_uvw - L U) = 0 at every node (the projection
  P = I - x x^T; the hyperviscosity term is added unprojected);
* verify_output (main.c:439-477) counts planted differences and NaNs.

A wrong stencil index, a transposed operand, a dropped or sign-flipped term,
or a wrong stage constant fails one of these.
"""
import dataclasses
import math
import os

import numpy as np
import pytest

N = 162   # the smallest icosahedral mesh with 31-point stencils


@pytest.fixture(scope="module")
def case(pkg):
    return pkg.mesh.read_bin(pkg.mesh.case_path(N, seed=3))


def _problem(pkg, raw, tmp_path, name, **changes):
    c = dataclasses.replace(raw, **changes)
    path = os.path.join(str(tmp_path), name + ".bin")
    pkg.mesh.write_bin(path, c)
    return pkg.host.Problem(path, use_rcm=False, layout="cfdl", simd=4)


def _dense(prob):
    """Dense N x N operators from the prepared (scaled) stencil arrays."""
    n, W = prob.n_nodes, prob.padded_nnbr
    idx = prob.idx.reshape(prob.padded, W)[:n, :prob.nnbr]
    ops = []
    for nm in ("Dx", "Dy", "Dz", "L"):
        w = getattr(prob, nm).reshape(prob.padded, W)[:n, :prob.nnbr]
        M = np.zeros((n, n))
        np.add.at(M, (np.repeat(np.arange(n), prob.nnbr), idx.ravel()), w.ravel())
        ops.append(M)
    return ops


def _geom(prob):
    n = prob.n_nodes
    X = np.stack([prob.x[:n], prob.y[:n], prob.z[:n]], 1)
    P = np.stack([prob.p_u.reshape(-1, 3)[:n], prob.p_v.reshape(-1, 3)[:n],
                  prob.p_w.reshape(-1, 3)[:n]], 1)           # [node][row][col]
    return X, P, prob.f[:n], prob.ghm[:n], prob.gradghm.reshape(-1, 3)[:n]


def _rhs(pyoracle, prob, S):
    """Oracle RHS of the node-major state S [N][4] (cfdl host layout)."""
    K = np.zeros((prob.patch_size, 4))
    K[:prob.n_nodes] = S
    K[prob.n_nodes:] = S[-1]   # padding replicates the last node
    F = pyoracle.Oracle(prob, "strict").eval_rhs(K.ravel())
    return F.reshape(-1, 4)[:prob.n_nodes]


def _close(a, b, rel=1e-11):
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= rel * scale, (np.abs(a - b).max(), scale)


def test_rk4_step_of_the_linear_system_is_the_textbook_polynomial(pkg, pyoracle, case,
                                                                  tmp_path):
    z = np.zeros_like(case.Dx)
    prob = _problem(pkg, case, tmp_path, "linear", Dx=z, Dy=z, Dz=z, f=np.zeros(N),
                    gradghm=np.zeros((N, 3)))
    _, _, _, L = _dense(prob)
    H0 = case.H_init.copy()
    H = np.zeros((prob.patch_size, 4))
    H[:N] = H0
    H[N:] = H0[-1]
    got = pyoracle.Oracle(prob, "strict").run(H.ravel(), 1).reshape(-1, 4)[:N]
    A = prob.dt * L
    want, term = H0.copy(), H0.copy()
    for k in range(1, 5):
        term = A @ term / k
        want = want + term
    _close(got, want, 1e-12)
    # a planted mistake in the stage constants would show: the Euler step and
    # the second-order polynomial both differ from the oracle's step
    assert np.abs(got - (H0 + A @ H0)).max() > 1e3 * np.abs(got - want).max()


def test_rhs_zero_velocity_is_pressure_gradient_and_hyperviscosity(pkg, pyoracle, case,
                                                                  tmp_path):
    prob = _problem(pkg, case, tmp_path, "base")
    Dx, Dy, Dz, L = _dense(prob)
    X, P, f, ghm, gg = _geom(prob)
    S = case.H_init.copy()
    S[:, :3] = 0.0
    h = S[:, 3]
    F = _rhs(pyoracle, prob, S)
    grad = np.stack([Dx @ h, Dy @ h, Dz @ h], 1)
    _close(F[:, :3], -np.einsum("nij,nj->ni", P, grad))
    _close(F[:, 3], L @ h)


def test_rhs_zero_depth_field_is_advection_coriolis_continuity(pkg, pyoracle, case,
                                                              tmp_path):
    prob = _problem(pkg, case, tmp_path, "base")
    Dx, Dy, Dz, L = _dense(prob)
    X, P, f, ghm, gg = _geom(prob)
    S = case.H_init.copy()
    S[:, 3] = 0.0
    U = S[:, :3]
    F = _rhs(pyoracle, prob, S)
    adv = np.stack([U[:, 0] * (Dx @ U[:, c]) + U[:, 1] * (Dy @ U[:, c])
                    + U[:, 2] * (Dz @ U[:, c]) for c in range(3)], 1)
    cor = f[:, None] * np.cross(X, U)
    want_uvw = -np.einsum("nij,nj->ni", P, adv + cor) + (L @ U)
    div = Dx @ U[:, 0] + Dy @ U[:, 1] + Dz @ U[:, 2]
    want_h = np.einsum("ni,ni->n", U, gg) - (prob.gh0 - ghm) * div
    _close(F[:, :3], want_uvw)
    _close(F[:, 3], want_h)


def test_rhs_constant_state_closed_form(pkg, pyoracle, case, tmp_path):
    """Zero row sums (RBF-FD weights reproduce constants) leave only the
    Coriolis force, tangent already, and the orography term."""
    prob = _problem(pkg, case, tmp_path, "base")
    X, P, f, ghm, gg = _geom(prob)
    U = np.array([12.0, -7.0, 3.0])
    S = np.tile(np.r_[U, 850.0], (N, 1))
    F = _rhs(pyoracle, prob, S)
    _close(F[:, :3], -f[:, None] * np.cross(X, U[None, :]), 1e-9)
    _close(F[:, 3], gg @ U, 1e-9)


def test_rhs_is_tangent_up_to_hyperviscosity(pkg, pyoracle, case, tmp_path):
    prob = _problem(pkg, case, tmp_path, "base")
    *_, L = _dense(prob)
    X = _geom(prob)[0]
    S = case.H_init.copy()
    F = _rhs(pyoracle, prob, S)
    radial = np.einsum("ni,ni->n", X, F[:, :3] - L @ S[:, :3])
    assert np.abs(radial).max() <= 1e-11 * np.abs(F[:, :3]).max()
    # the hyperviscosity part itself is not tangent, so the check is not empty
    assert np.abs(np.einsum("ni,ni->n", X, L @ S[:, :3])).max() > 1e3 * np.abs(radial).max()


def test_verify_counts_planted_differences(pkg, pyoracle, case, tmp_path):
    prob = _problem(pkg, case, tmp_path, "base")
    orc = pyoracle.Oracle(prob, "strict")
    H = prob.H_init.copy()
    assert orc.verify(H, H, 0.0) == 0
    bad = H.copy()
    rng = np.random.default_rng(5)
    picks = rng.choice(4 * prob.compute_size, 7, replace=False)
    bad[picks[:6]] += 1e-6
    bad[picks[6]] = math.nan
    assert orc.verify(H, bad, 1e-9) == 7
    assert orc.verify(H, bad, 1e-5) == 1   # only the NaN exceeds a looser tolerance
