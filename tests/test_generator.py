"""The seeded input generator: closed forms, the reference's shipped mesh
and weights, and stability of the generated cases."""
import math

import numpy as np
import pytest

from conftest import REFERENCE_BIN, have_reference


def test_icosahedral_counts_and_unit_norm(pkg):
    for lev in range(0, 5):
        X = pkg.mesh.icosahedral_nodes(lev)
        assert len(X) == 10 * 4 ** lev + 2
        assert np.allclose(np.linalg.norm(X, axis=1), 1.0, atol=1e-15)
        assert len(np.unique(np.round(X, 12), axis=0)) == len(X)


def test_stencils_self_first_and_nearest(pkg):
    X = pkg.mesh.icosahedral_nodes(3)
    idx = pkg.mesh.stencils(X, 31)
    assert (idx[:, 0] == np.arange(len(X))).all()
    d = np.linalg.norm(X[idx] - X[:, None, :], axis=2)
    assert (np.diff(d, axis=1) >= -1e-15).all()
    # the pentagonal base vertices have 5 nearest neighbours at one distance
    assert np.allclose(d[0, 1:6], d[0, 1])


def test_closed_forms(pkg):
    c = pkg.mesh.make_case(642, seed=2)
    X = c.X
    assert np.array_equal(c.f, 2.0 * pkg.mesh.OMEGA * X[:, 2])
    P = np.stack([c.p_u, c.p_v, c.p_w], 1)
    assert np.allclose(P, np.eye(3) - X[:, :, None] * X[:, None, :], atol=1e-15)
    vel = c.H_init[:, :3]
    assert np.abs((vel * X).sum(1)).max() < 1e-12        # tangent initial flow
    u0 = c.params["u0"]
    assert np.allclose(c.H_init[:, 3],
                       -(pkg.mesh.A_EARTH * pkg.mesh.OMEGA * u0 + 0.5 * u0 * u0) * X[:, 2] ** 2)
    nb = c.ghm[c.idx]
    assert np.array_equal(c.gradghm[:, 0], (c.Dx * nb).sum(1) / pkg.mesh.A_EARTH)
    # constant augmentation: rows of every operator annihilate constants
    for W in (c.Dx, c.Dy, c.Dz):
        assert np.abs(W.sum(1)).max() < 1e-9 * np.abs(W).max()
    assert np.abs(c.L.sum(1)).max() < 1e-9 * np.abs(c.L).max()
    assert c.gamma < 0 and c.dt == 1800.0


def test_gradient_weights_are_accurate(pkg):
    """P grad of the smooth function g = x + y^2 + z^3 on the sphere."""
    c = pkg.mesh.make_case(2562, seed=1)
    X = c.X
    g = X[:, 0] + X[:, 1] ** 2 + X[:, 2] ** 3
    grad = np.stack([np.ones(len(X)), 2 * X[:, 1], 3 * X[:, 2] ** 2], 1)
    Pg = grad - X * (grad * X).sum(1, keepdims=True)
    approx = np.stack([(W * g[c.idx]).sum(1) for W in (c.Dx, c.Dy, c.Dz)], 1)
    assert np.abs(approx - Pg).max() < 1e-4 * np.abs(Pg).max()


def test_gaussian_laplacian_power_formula(pkg):
    """Lap^k exp(-(eps r)^2) in 2-D == (-4 eps^2)^k k! Lag_k(eps^2 r^2) phi,
    checked for k = 1, 2 against a 5-point finite difference of the k-1 form."""
    from scipy.special import eval_laguerre
    eps = 1.7

    def lap_k(k, x, y):
        r2 = x * x + y * y
        return ((-4 * eps * eps) ** k) * math.factorial(k) * eval_laguerre(k, eps * eps * r2) \
            * np.exp(-eps * eps * r2)
    h = 1e-3
    for k in (1, 2):
        for (x, y) in [(0.1, 0.2), (0.5, -0.3), (0.0, 0.0)]:
            fd = (lap_k(k - 1, x + h, y) + lap_k(k - 1, x - h, y) + lap_k(k - 1, x, y + h)
                  + lap_k(k - 1, x, y - h) - 4 * lap_k(k - 1, x, y)) / (h * h)
            assert abs(fd - lap_k(k, x, y)) < 1e-4 * max(1.0, abs(lap_k(k, x, y)))


@pytest.mark.skipif(not have_reference(), reason="reference tree not present")
def test_reproduces_reference_mesh_and_weights(pkg):
    ref = pkg.mesh.read_bin(REFERENCE_BIN)
    X = pkg.mesh.icosahedral_nodes(4)
    from scipy.spatial import cKDTree
    d, _ = cKDTree(X).query(ref.X)
    assert d.max() < 1e-14                      # same node set
    eps = pkg.mesh.shape_parameter(ref.n_nodes)
    rows = np.arange(0, ref.n_nodes, 97)
    Dx, Dy, Dz, L = pkg.mesh.rbffd_weights(ref.X, ref.idx, eps)
    for mine, theirs in ((Dx, ref.Dx), (Dy, ref.Dy), (Dz, ref.Dz), (L, ref.L)):
        err = np.abs(mine[rows] - theirs[rows]).max(1) / np.abs(theirs[rows]).max(1)
        assert err.max() < 1e-3
    assert np.allclose(ref.gamma * ref.n_nodes ** 4, -0.05, rtol=1e-3)


def test_generated_case_is_stable(pkg, pyoracle):
    """300 RK4 steps at N=2562, dt=900 s stay bounded (the reference's own
    input is unstable past ~300 steps at dt=197 s)."""
    path = pkg.mesh.case_path(2562, seed=1)
    prob = pkg.host.Problem(path, use_rcm=True, layout="cfdl")
    H = pyoracle.Oracle(prob, "fast").run(prob.H_init, 300)
    Hn = prob.to_node_major(H)
    speed = np.sqrt((Hn[:, :3] ** 2).sum(1))
    assert np.isfinite(Hn).all() and speed.max() < 60.0


def test_read_write_round_trip(pkg, tmp_path):
    c = pkg.mesh.make_case(162, seed=5)
    p = str(tmp_path / "c.bin")
    pkg.mesh.write_bin(p, c)
    r = pkg.mesh.read_bin(p)
    for nm in ("X", "f", "ghm", "p_u", "gradghm", "Dx", "L", "idx", "H_init"):
        assert np.array_equal(getattr(r, nm), getattr(c, nm)), nm
    import os
    N, W = c.idx.shape
    assert os.path.getsize(p) == 40 + 200 * N + 36 * N * W


def test_members_perturbation(pkg):
    c = pkg.mesh.make_case(642, seed=1)
    E = pkg.mesh.perturb_members(c, 3, seed=9)
    assert E.shape == (3, 642, 4)
    assert np.abs((E[:, :, :3] * c.X[None]).sum(2)).max() < 1e-12
    nz = np.abs(c.H_init[:, 3]) > 0
    rel = np.abs(E[:, nz, 3] / c.H_init[nz, 3] - 1)
    assert 1e-5 < rel.max() < 1e-2
    assert not np.array_equal(E[0], E[1])
