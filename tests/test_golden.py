"""Committed fixtures made by the reference build (tests/golden/make_golden.py):
the host + oracle must reproduce them bit for bit on any machine."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

CASE = os.path.join(GOLDEN, "icos642_case.bin")
REFV = os.path.join(GOLDEN, "icos642_ref.npz")


def natural(prob, H):
    Hn = prob.to_node_major(H)[: prob.n_nodes]
    out = np.empty_like(Hn)
    out[prob.mapping] = Hn
    return out


@pytest.fixture(scope="module")
def golden(pkg):
    prob = pkg.host.Problem(CASE, use_rcm=True, layout="cfdl", simd=4)
    return prob, np.load(REFV, allow_pickle=False)


def test_rcm_matches_reference(golden):
    prob, g = golden
    assert np.array_equal(prob.mapping, g["rcm"])


def test_oracle_rhs_matches_reference(golden, pyoracle):
    prob, g = golden
    F = pyoracle.Oracle(prob).eval_rhs(prob.H_init)
    assert np.array_equal(natural(prob, F), g["F0"])


@pytest.mark.parametrize("steps", [1, 10])
def test_oracle_steps_match_reference(golden, pyoracle, steps):
    prob, g = golden
    H = pyoracle.Oracle(prob).run(prob.H_init, steps)
    assert np.array_equal(natural(prob, H), g[f"H{steps}"])


def test_oracle_known_answer_100_steps(golden, pyoracle):
    prob, _ = golden
    orc = pyoracle.Oracle(prob)
    H = orc.run(prob.H_init, 100)
    assert orc.verify(H, prob.H_100, 0.0) == 0       # bitwise: the slot holds the reference's state
    assert np.array_equal(H, prob.H_100)


def test_other_layouts_same_numbers(pkg, pyoracle, golden):
    prob0, g = golden
    for lay in ("default", "sfdl", "cfdl_sfdl"):
        prob = pkg.host.Problem(CASE, use_rcm=True, layout=lay, simd=8)
        H = pyoracle.Oracle(prob).run(prob.H_init, 10)
        assert np.array_equal(natural(prob, H), g["H10"]), lay


REFV32 = os.path.join(GOLDEN, "icos642_ref32.npz")


@pytest.mark.parametrize("layout", ["cfdl", "sfdl"])
def test_fp32_oracle_matches_reference_fixtures(pkg, pyoracle, layout):
    """The FP32 oracle against the vectors the reference's shipped FP32 build
    wrote (RHS, 1, 10 and 100 steps, float32, natural order), and the FP32
    known-answer check against the file's H_100 at 4e-2."""
    g = np.load(REFV32, allow_pickle=False)
    prob = pkg.host.Problem(CASE, use_rcm=True, layout=layout, simd=4, fp32=True)
    orc = pyoracle.Oracle(prob, "strict32")
    assert np.array_equal(natural(prob, orc.eval_rhs(prob.H_init)), g["F0"])
    for k in (1, 10, 100):
        assert np.array_equal(natural(prob, orc.run(prob.H_init, k)), g[f"H{k}"]), k
    assert orc.verify(orc.run(prob.H_init, 100), prob.H_100, 4e-2) == 0
