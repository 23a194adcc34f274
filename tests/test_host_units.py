"""C host units: index functions, layout transforms, reader errors, RCM."""
import ctypes as C
import os
import struct

import numpy as np
import pytest


def test_index_functions_closed_form(pkg):
    h = pkg.host
    L = h.lib()
    for cf, sf, S in [(0, 0, 4), (1, 0, 8), (0, 1, 4), (1, 1, 32), (0, 1, 3)]:
        lay = h.Layout(cf, sf, S)
        for i in [0, 1, S - 1, S, 3 * S + 2, 1000]:
            for j in [0, 1, 30]:
                want = (i // S) * S * 32 + j * S + i % S if sf else i * 32 + j
                assert L.swe_dm_lin_id(C.byref(lay), i, j, 32) == want
            for c in range(3):
                want = (i // S) * S * 3 + c * S + i % S if sf else i * 3 + c
                assert L.swe_vect_lin_id(C.byref(lay), i, c) == want
            for v in range(4):
                want = 4 * i + v if cf else (i // S) * S * 4 + v * S + i % S
                assert L.swe_svm_lin_id(C.byref(lay), i, v) == want


def test_svm_layout_is_a_permutation(pkg, tmp_path):
    path = pkg.mesh.case_path(642, seed=1)
    for lay in ("default", "cfdl", "sfdl", "cfdl_sfdl"):
        prob = pkg.host.Problem(path, use_rcm=False, layout=lay, simd=4)
        idx = prob.svm_index().ravel()
        assert np.array_equal(np.sort(idx), np.arange(4 * prob.patch_size))
        # padded rows replicate the last real node (matrix_transformations.c:20)
        Hn = prob.to_node_major(prob.H_init)
        assert (Hn[prob.n_nodes:] == Hn[prob.n_nodes - 1]).all()


def test_reader_scaling_and_padding(pkg):
    path = pkg.mesh.case_path(642, seed=1)
    raw = pkg.mesh.read_bin(path)
    prob = pkg.host.Problem(path, use_rcm=False, layout="cfdl", simd=8)
    N, W, Wp = raw.n_nodes, raw.nnbr, prob.padded_nnbr
    assert Wp == 32 and prob.padded == ((N - 1) // 8 + 1) * 8
    Dx = prob.Dx.reshape(prob.padded, Wp)
    L = prob.L.reshape(prob.padded, Wp)
    idx = prob.idx.reshape(prob.padded, Wp)
    assert np.array_equal(Dx[:N, :W], raw.Dx / raw.a)
    assert np.array_equal(L[:N, :W], raw.L * raw.gamma)
    assert np.array_equal(idx[:N, :W], raw.idx)
    # the padded stencil column replicates the last one (matrix_transformations.c:21)
    assert np.array_equal(Dx[:N, W], Dx[:N, W - 1])
    assert (idx[:, 0][:N] == np.arange(N)).all()


def _write(path, blob):
    with open(path, "wb") as fp:
        fp.write(blob)


def test_reader_errors(pkg, tmp_path):
    h = pkg.host
    g = h.Gsmd()
    assert h.lib().swe_get_gsmd_bin(str(tmp_path / "missing.bin").encode(), C.byref(g)) == -1
    good = open(pkg.mesh.case_path(642, seed=1), "rb").read()
    trunc = tmp_path / "trunc.bin"
    _write(trunc, good[: len(good) // 2])
    assert h.lib().swe_get_gsmd_bin(str(trunc).encode(), C.byref(g)) == -2
    # an out-of-range stencil id is rejected instead of read out of bounds
    N, W = struct.unpack("<ii", good[:8])
    off = 8 + 32 + 8 * (5 * N + 12 * N + 4 * N * W)
    bad = bytearray(good)
    bad[off:off + 4] = struct.pack("<i", N + 5)
    badp = tmp_path / "bad.bin"
    _write(badp, bytes(bad))
    assert h.lib().swe_get_gsmd_bin(str(badp).encode(), C.byref(g)) == -2
    with pytest.raises(RuntimeError):
        h.Problem(str(tmp_path / "missing.bin"))


def test_rcm_is_permutation_and_narrows_band(pkg):
    path = pkg.mesh.case_path(2562, seed=1)
    nat = pkg.host.Problem(path, use_rcm=False, layout="cfdl")
    rcm = pkg.host.Problem(path, use_rcm=True, layout="cfdl")
    N = rcm.n_nodes
    assert np.array_equal(np.sort(rcm.mapping), np.arange(N))

    def band(p):
        idx = p.idx.reshape(p.padded, p.padded_nnbr)[:N, : p.nnbr]
        return int((idx.max(1) - idx.min(1)).max())
    assert band(rcm) < band(nat) // 3


def test_rcm_even_n_middle_pair_not_swapped(pkg):
    """rcm.c:51 reverses with i < (N-1)/2: for even N the two middle entries
    keep their BFS order.  On a path graph 0-1-2-3 (stencil = self + the two
    neighbours, clamped), BFS from node 0 gives 0,1,2,3 and the reversal
    gives 3,1,2,0."""
    h = pkg.host
    N, W = 4, 3
    idx = np.array([[0, 1, 1], [1, 0, 2], [2, 1, 3], [3, 2, 2]], dtype=np.int32)
    g = h.Gsmd()
    g.Nnodes, g.Nnbr = N, W
    g.idx = idx.ctypes.data_as(C.POINTER(C.c_int))
    out = C.POINTER(C.c_int)()
    assert h.lib().swe_rcm_mapping(C.byref(g), C.byref(out)) == 0
    got = [out[i] for i in range(N)]
    assert got == [3, 1, 2, 0]


def test_rcm_disconnected_graph_is_an_error(pkg):
    h = pkg.host
    idx = np.array([[0, 1], [1, 0], [2, 3], [3, 2]], dtype=np.int32)
    g = h.Gsmd()
    g.Nnodes, g.Nnbr = 4, 2
    g.idx = idx.ctypes.data_as(C.POINTER(C.c_int))
    out = C.POINTER(C.c_int)()
    assert h.lib().swe_rcm_mapping(C.byref(g), C.byref(out)) == -5


@pytest.mark.parametrize("seed", range(6))
def test_rcm_matches_reference_on_random_graphs(pkg, pyoracle, seed):
    """Product RCM == the reference's rcm_mapping on kNN graphs of random
    points (odd and even N, varied widths)."""
    from conftest import have_reference
    if not (have_reference() and pyoracle.ref_available("cfdl")):
        pytest.skip("reference build missing")
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    N = int(rng.integers(50, 400))
    W = int(rng.integers(5, 20))
    X = rng.standard_normal((N, 3))
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    _, idx = cKDTree(X).query(X, k=W)
    idx = np.ascontiguousarray(idx.astype(np.int32))
    idx[:, 0] = np.arange(N)
    h = pkg.host
    g = h.Gsmd()
    g.Nnodes, g.Nnbr = N, W
    g.idx = idx.ctypes.data_as(C.POINTER(C.c_int))
    out = C.POINTER(C.c_int)()
    rc = h.lib().swe_rcm_mapping(C.byref(g), C.byref(out))
    if rc == -5:
        pytest.skip("random kNN graph not connected")
    mine = np.array([out[i] for i in range(N)])
    assert np.array_equal(mine, pyoracle.ref_rcm(idx))
