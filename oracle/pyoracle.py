"""ctypes bindings of the CPU oracle (liboracle.so / liboracle_fast.so) and of
the reference build (_ref/libswe_ref_<layout>.so) -- TEST INFRASTRUCTURE
ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg, never by the product."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OraclePsmd(C.Structure):
    _fields_ = [("compute_size", C.c_int), ("patch_size", C.c_int),
                ("compute_pid_s", C.c_int), ("Nnbr", C.c_int), ("padded_Nnbr", C.c_int),
                ("cfdl", C.c_int), ("sfdl", C.c_int), ("simd", C.c_int),
                ("gh0", C.c_double), ("dt", C.c_double),
                ("x", _dp), ("y", _dp), ("z", _dp), ("f", _dp), ("ghm", _dp),
                ("p_u", _dp), ("p_v", _dp), ("p_w", _dp), ("gradghm", _dp),
                ("idx", _ip), ("Dx", _dp), ("Dy", _dp), ("Dz", _dp), ("L", _dp)]


class OracleTimers(C.Structure):
    _fields_ = [("t_eval_rhs", C.c_double), ("t_eval_K", C.c_double),
                ("t_update_D", C.c_double), ("t_update_H", C.c_double)]


_libs: dict = {}


def load(flavour: str = "strict") -> C.CDLL:
    name = {"strict": "liboracle.so", "fast": "liboracle_fast.so"}[flavour]
    if name not in _libs:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (make -C oracle)")
        L = C.CDLL(path)
        P = C.POINTER(OraclePsmd)
        T = C.POINTER(OracleTimers)
        L.oracle_eval_rhs.argtypes = [P, _dp, _dp, T]
        L.oracle_rk_substep.argtypes = [P, _dp, _dp, _dp, _dp, C.c_int, T]
        L.oracle_run.argtypes = [P, _dp, _dp, _dp, _dp, C.c_int, T]
        L.oracle_verify.argtypes = [P, _dp, _dp, C.c_double]
        _libs[name] = L
    return _libs[name]


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


class Oracle:
    """The oracle bound to one prepared problem (rbf-fd-swe_amd host.Problem
    or any object with the same attributes, in the reference host layout)."""

    def __init__(self, prob, flavour: str = "strict"):
        self.lib = load(flavour)
        self.prob = prob
        lay = prob.layout
        self._keep = []

        def keep(a, dtype=np.float64):
            a = np.ascontiguousarray(a, dtype=dtype)
            self._keep.append(a)
            return a
        m = OraclePsmd()
        m.compute_size, m.patch_size = prob.compute_size, prob.patch_size
        m.compute_pid_s = prob.compute_pid_s
        m.Nnbr, m.padded_Nnbr = prob.nnbr, prob.padded_nnbr
        m.cfdl, m.sfdl, m.simd = lay.cfdl, lay.sfdl, lay.simd
        m.gh0, m.dt = prob.gh0, prob.dt
        for nm in ("x", "y", "z", "f", "ghm", "p_u", "p_v", "p_w", "gradghm",
                   "Dx", "Dy", "Dz", "L"):
            setattr(m, nm, _ptr(keep(getattr(prob, nm))))
        m.idx = _ptr(keep(prob.idx, np.int32), _ip)
        self.m = m
        n4c, n4p = 4 * prob.compute_size, 4 * prob.patch_size
        self.K = np.zeros(n4p)
        self.F = np.zeros(n4c)
        self.D = np.zeros(n4c)
        self.timers = OracleTimers()

    def eval_rhs(self, K: np.ndarray) -> np.ndarray:
        K = np.ascontiguousarray(K, dtype=np.float64)
        F = np.zeros(4 * self.prob.compute_size)
        self.lib.oracle_eval_rhs(C.byref(self.m), _ptr(K), _ptr(F), C.byref(self.timers))
        return F

    def rk_substep(self, H: np.ndarray, s: int) -> None:
        """In place on H (and on the oracle's K, F, D scratch)."""
        assert H.dtype == np.float64 and H.flags.c_contiguous
        self.lib.oracle_rk_substep(C.byref(self.m), _ptr(H), _ptr(self.K), _ptr(self.F),
                                   _ptr(self.D), int(s), C.byref(self.timers))

    def run(self, H: np.ndarray, nsteps: int) -> np.ndarray:
        H = np.array(H, dtype=np.float64, copy=True)
        self.lib.oracle_run(C.byref(self.m), _ptr(H), _ptr(self.K), _ptr(self.F),
                            _ptr(self.D), int(nsteps), C.byref(self.timers))
        return H

    def verify(self, H1, H2, tol=1e-9) -> int:
        H1 = np.ascontiguousarray(H1, dtype=np.float64)
        H2 = np.ascontiguousarray(H2, dtype=np.float64)
        return self.lib.oracle_verify(C.byref(self.m), _ptr(H1), _ptr(H2), tol)


# ------------------------------------------------------------ reference build

def ref_available(layout: str = "cfdl") -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", f"libswe_ref_{layout}.so"))


class RefBuild:
    """The reference's own sources (built by build_ref.sh) driven through
    ref_harness.c.  Global state lives in the library, so one instance per
    layout is used at a time."""

    def __init__(self, path: str, use_rcm: bool = True, layout: str = "cfdl",
                 nattempts: int = 1):
        so = os.path.join(HERE, "_ref", f"libswe_ref_{layout}.so")
        if not os.path.exists(so):
            raise RuntimeError(f"{so} not built")
        L = C.CDLL(so)
        L.ref_setup.argtypes = [C.c_char_p, C.c_int, C.c_int]
        L.ref_info.argtypes = [_ip]
        L.ref_scalars.argtypes = [_dp]
        L.ref_arrays.argtypes = [C.POINTER(_dp), C.POINTER(_ip)]
        L.ref_H_init.restype = _dp
        L.ref_H_100.restype = _dp
        L.ref_mapping.restype = _ip
        L.ref_rk_substep.argtypes = [_dp, _dp, _dp, _dp, C.c_int]
        L.ref_eval_rhs.argtypes = [_dp, _dp]
        if L.ref_setup(path.encode(), int(bool(use_rcm)), nattempts) != 0:
            raise RuntimeError("ref_setup failed")
        info = (C.c_int * 10)()
        L.ref_info(info)
        (self.n_nodes, self.padded, self.compute_size, self.patch_size, self.compute_pid_s,
         self.nnbr, self.padded_nnbr, self.simd, self.cfdl, self.sfdl) = list(info)
        sc = (C.c_double * 2)()
        L.ref_scalars(sc)
        self.gh0, self.dt = sc[0], sc[1]
        fp = (_dp * 13)()
        ip = _ip()
        L.ref_arrays(fp, C.byref(ip))
        Np, W = self.compute_size, self.padded_nnbr
        names = ["x", "y", "z", "f", "ghm", "p_u", "p_v", "p_w", "gradghm", "Dx", "Dy", "Dz", "L"]
        sizes = [Np] * 5 + [3 * Np] * 4 + [Np * W] * 4
        for nm, p, n in zip(names, fp, sizes):
            setattr(self, nm, np.ctypeslib.as_array(p, shape=(n,)).copy())
        self.idx = np.ctypeslib.as_array(ip, shape=(Np * W,)).copy()
        self.H_init = np.ctypeslib.as_array(L.ref_H_init(), shape=(4 * self.patch_size,)).copy()
        self.H_100 = np.ctypeslib.as_array(L.ref_H_100(), shape=(4 * self.patch_size,)).copy()
        mp = L.ref_mapping()
        self.mapping = (np.ctypeslib.as_array(mp, shape=(self.n_nodes,)).copy()
                        if mp else None)
        self.lib = L

    def rk_substep(self, H, K, F, D, s):
        self.lib.ref_rk_substep(_ptr(H), _ptr(K), _ptr(F), _ptr(D), int(s))

    def run(self, H, nsteps):
        H = np.array(H, dtype=np.float64, copy=True)
        K = np.zeros(4 * self.patch_size)
        F = np.zeros(4 * self.compute_size)
        D = np.zeros(4 * self.compute_size)
        for _ in range(nsteps):
            for s in range(4):
                self.rk_substep(H, K, F, D, s)
        return H

    def eval_rhs(self, K):
        K = np.ascontiguousarray(K, dtype=np.float64)
        F = np.zeros(4 * self.compute_size)
        self.lib.ref_eval_rhs(_ptr(K), _ptr(F))
        return F


def ref_rcm(idx: np.ndarray, layout: str = "cfdl") -> np.ndarray:
    """The reference's rcm_mapping on a stencil graph (N, W), 0-based ids."""
    so = os.path.join(HERE, "_ref", f"libswe_ref_{layout}.so")
    L = C.CDLL(so)
    L.ref_rcm_of.argtypes = [C.c_int, C.c_int, _ip, _ip]
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    out = np.empty(idx.shape[0], dtype=np.int32)
    L.ref_rcm_of(idx.shape[0], idx.shape[1], idx.ctypes.data_as(_ip), out.ctypes.data_as(_ip))
    return out


def ref_profiling(phase_times: dict, nsteps: int, layout: str = "cfdl") -> str:
    """The reference's process_profiling_results report (its stdout text) for
    per-attempt phase times (seconds; keys t_main, t_eval_rhs, t_eval_K,
    t_update_D, t_update_H)."""
    so = os.path.join(HERE, "_ref", f"libswe_ref_{layout}.so")
    L = C.CDLL(so)
    L.ref_profiling.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_char_p, C.c_int]
    keys = ("t_main", "t_eval_rhs", "t_eval_K", "t_update_D", "t_update_H")
    arrs = [np.ascontiguousarray(phase_times[k], dtype=np.float64) for k in keys]
    buf = C.create_string_buffer(1 << 16)
    n = L.ref_profiling(len(arrs[0]), int(nsteps), *[_ptr(a) for a in arrs], buf, 1 << 16)
    if n < 0:
        raise RuntimeError("ref_profiling could not capture the report")
    return buf.value.decode()
