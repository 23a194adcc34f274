"""Writes the committed golden fixtures from the reference build
(oracle/_ref, compiled from the reference's own sources by
oracle/build_ref.sh).  Run in a container that has the reference tree:

    python tests/golden/make_golden.py

Outputs (all data, no reference source):
  icos642_case.bin  a generated N=642 input in the reference's binary
                    format; its known-answer slot (H_100) holds the
                    reference build's state after 100 RK4 steps (natural
                    node order), so the reference's own verify_output check
                    applies to it
  icos642_ref.npz   reference-build vectors on that input, natural node
                    order: F0 = eval_RHS_swe(H_init), H1/H10 = states after
                    1 and 10 steps, rcm = rcm_mapping(), plus the
                    reference's 2562 input SHA-256 and its H_100 check count
  icos642_ref32.npz the same vectors (float32) from the reference built with
                    its header as shipped (RUN_32BIT, fType = float), plus
                    H100, on the committed icos642_case.bin
                    (python tests/golden/make_golden.py fp32 writes only this)
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
import swe_pkg  # noqa: E402

pkg = swe_pkg.pkg
REF_BIN = "/root/reference/input_files/icos2562_tc5_input.bin"


def natural(ref, H):
    """Reference patch state (cfdl layout, RCM order, padded) -> [N][4] in
    the file's node order."""
    Hn = np.asarray(H).reshape(-1, 4)[: ref.n_nodes]
    out = np.empty_like(Hn)
    out[ref.mapping] = Hn
    return out


def main():
    case = pkg.mesh.make_case(642, seed=3)
    tmp = os.path.join(HERE, "icos642_case.bin")
    pkg.mesh.write_bin(tmp, case)
    ref = pyoracle.RefBuild(tmp, use_rcm=True, layout="cfdl")
    F0 = natural(ref, ref.eval_rhs(ref.H_init))
    H1 = natural(ref, ref.run(ref.H_init, 1))
    H10 = natural(ref, ref.run(ref.H_init, 10))
    H100 = natural(ref, ref.run(ref.H_init, 100))
    case.H_final = H100
    pkg.mesh.write_bin(tmp, case)
    sha = hashlib.sha256(open(REF_BIN, "rb").read()).hexdigest() if os.path.exists(REF_BIN) else ""
    np.savez_compressed(os.path.join(HERE, "icos642_ref.npz"), F0=F0, H1=H1, H10=H10,
                        rcm=ref.mapping.astype(np.int32), ref2562_sha256=np.array(sha),
                        case_params=np.array(repr(case.params)))
    print("wrote", tmp, "max|u| after 100 steps", np.abs(H100[:, :3]).max())


def main_fp32():
    path = os.path.join(HERE, "icos642_case.bin")
    ref = pyoracle.RefBuild(path, use_rcm=True, layout="cfdl", fp32=True)
    out = {"F0": natural(ref, ref.eval_rhs(ref.H_init))}
    for k in (1, 10, 100):
        out[f"H{k}"] = natural(ref, ref.run(ref.H_init, k))
    assert all(v.dtype == np.float32 for v in out.values())
    np.savez_compressed(os.path.join(HERE, "icos642_ref32.npz"),
                        rcm=ref.mapping.astype(np.int32), **out)
    print("wrote icos642_ref32.npz, max|u| after 100 steps", np.abs(out["H100"][:, :3]).max())


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "fp32":
        main_fp32()
    else:
        main()
