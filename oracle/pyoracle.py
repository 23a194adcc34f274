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


def _psmd_fields(real, rp):
    return [("compute_size", C.c_int), ("patch_size", C.c_int),
            ("compute_pid_s", C.c_int), ("Nnbr", C.c_int), ("padded_Nnbr", C.c_int),
            ("cfdl", C.c_int), ("sfdl", C.c_int), ("simd", C.c_int),
            ("gh0", real), ("dt", real),
            ("x", rp), ("y", rp), ("z", rp), ("f", rp), ("ghm", rp),
            ("p_u", rp), ("p_v", rp), ("p_w", rp), ("gradghm", rp),
            ("idx", _ip), ("Dx", rp), ("Dy", rp), ("Dz", rp), ("L", rp)]


_fp = C.POINTER(C.c_float)


class OraclePsmd(C.Structure):
    _fields_ = _psmd_fields(C.c_double, _dp)


class OraclePsmd32(C.Structure):
    """The FP32 build's struct (oracle_real = float)."""
    _fields_ = _psmd_fields(C.c_float, _fp)


class OracleTimers(C.Structure):
    _fields_ = [("t_eval_rhs", C.c_double), ("t_eval_K", C.c_double),
                ("t_update_D", C.c_double), ("t_update_H", C.c_double)]


_libs: dict = {}


FLAVOURS = {"strict": "liboracle.so", "fast": "liboracle_fast.so", "strict32": "liboracle32.so"}


def load(flavour: str = "strict") -> C.CDLL:
    name = FLAVOURS[flavour]
    if name not in _libs:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (make -C oracle)")
        L = C.CDLL(path)
        fp32 = flavour == "strict32"
        P = C.POINTER(OraclePsmd32 if fp32 else OraclePsmd)
        T = C.POINTER(OracleTimers)
        rp = _fp if fp32 else _dp
        L.oracle_eval_rhs.argtypes = [P, rp, rp, T]
        L.oracle_rk_substep.argtypes = [P, rp, rp, rp, rp, C.c_int, T]
        L.oracle_run.argtypes = [P, rp, rp, rp, rp, C.c_int, T]
        L.oracle_verify.argtypes = [P, rp, rp, C.c_double]
        if L.oracle_real_size() != (4 if fp32 else 8):
            raise RuntimeError(f"{path}: unexpected model precision")
        _libs[name] = L
    return _libs[name]


def _ptr(a, t=_dp):
    if a.dtype == np.float32 and t is _dp:
        t = _fp
    return a.ctypes.data_as(t)


class Oracle:
    """The oracle bound to one prepared problem (rbf-fd-swe_amd host.Problem
    or any object with the same attributes, in the reference host layout)."""

    def __init__(self, prob, flavour: str = "strict"):
        self.lib = load(flavour)
        self.prob = prob
        self.real = np.float32 if flavour == "strict32" else np.float64
        lay = prob.layout
        self._keep = []

        def keep(a, dtype=None):
            dtype = self.real if dtype is None else dtype
            if dtype == np.float32 and np.asarray(a).dtype == np.float64:
                a32 = np.asarray(a, dtype=np.float32)
                if not np.array_equal(a32.astype(np.float64), np.asarray(a)):
                    raise ValueError("FP32 oracle: array is not FP32-valued "
                                     "(load the problem with fp32=True)")
            a = np.ascontiguousarray(a, dtype=dtype)
            self._keep.append(a)
            return a
        m = OraclePsmd32() if self.real == np.float32 else OraclePsmd()
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
        self.K = np.zeros(n4p, dtype=self.real)
        self.F = np.zeros(n4c, dtype=self.real)
        self.D = np.zeros(n4c, dtype=self.real)
        self.timers = OracleTimers()

    def eval_rhs(self, K: np.ndarray) -> np.ndarray:
        K = np.ascontiguousarray(K, dtype=self.real)
        F = np.zeros(4 * self.prob.compute_size, dtype=self.real)
        self.lib.oracle_eval_rhs(C.byref(self.m), _ptr(K), _ptr(F), C.byref(self.timers))
        return F

    def rk_substep(self, H: np.ndarray, s: int) -> None:
        """In place on H (and on the oracle's K, F, D scratch)."""
        assert H.dtype == self.real and H.flags.c_contiguous
        self.lib.oracle_rk_substep(C.byref(self.m), _ptr(H), _ptr(self.K), _ptr(self.F),
                                   _ptr(self.D), int(s), C.byref(self.timers))

    def run(self, H: np.ndarray, nsteps: int) -> np.ndarray:
        H = np.array(H, dtype=self.real, copy=True)
        self.lib.oracle_run(C.byref(self.m), _ptr(H), _ptr(self.K), _ptr(self.F),
                            _ptr(self.D), int(nsteps), C.byref(self.timers))
        return H

    def verify(self, H1, H2, tol=1e-9) -> int:
        H1 = np.ascontiguousarray(H1, dtype=self.real)
        H2 = np.ascontiguousarray(H2, dtype=self.real)
        return self.lib.oracle_verify(C.byref(self.m), _ptr(H1), _ptr(H2), tol)


# ------------------------------------------------------------ reference build

def ref_available(layout: str = "cfdl", fp32: bool = False) -> bool:
    tag = "ref32" if fp32 else "ref"
    return os.path.exists(os.path.join(HERE, "_ref", f"libswe_{tag}_{layout}.so"))


class RefBuild:
    """The reference's own sources (built by build_ref.sh) driven through
    ref_harness.c.  Global state lives in the library, so one instance per
    layout is used at a time."""

    def __init__(self, path: str, use_rcm: bool = True, layout: str = "cfdl",
                 nattempts: int = 1, fp32: bool = False):
        tag = "ref32" if fp32 else "ref"
        so = os.path.join(HERE, "_ref", f"libswe_{tag}_{layout}.so")
        if not os.path.exists(so):
            raise RuntimeError(f"{so} not built")
        L = C.CDLL(so)
        if L.ref_ftype_size() != (4 if fp32 else 8):
            raise RuntimeError(f"{so}: unexpected model precision")
        self.real = np.float32 if fp32 else np.float64
        rp = _fp if fp32 else _dp
        L.ref_setup.argtypes = [C.c_char_p, C.c_int, C.c_int]
        L.ref_info.argtypes = [_ip]
        L.ref_scalars.argtypes = [_dp]
        L.ref_arrays.argtypes = [C.POINTER(rp), C.POINTER(_ip)]
        L.ref_H_init.restype = rp
        L.ref_H_100.restype = rp
        L.ref_mapping.restype = _ip
        L.ref_rk_substep.argtypes = [rp, rp, rp, rp, C.c_int]
        L.ref_eval_rhs.argtypes = [rp, rp]
        if L.ref_setup(path.encode(), int(bool(use_rcm)), nattempts) != 0:
            raise RuntimeError("ref_setup failed")
        info = (C.c_int * 10)()
        L.ref_info(info)
        (self.n_nodes, self.padded, self.compute_size, self.patch_size, self.compute_pid_s,
         self.nnbr, self.padded_nnbr, self.simd, self.cfdl, self.sfdl) = list(info)
        sc = (C.c_double * 2)()
        L.ref_scalars(sc)
        self.gh0, self.dt = sc[0], sc[1]
        fp = (rp * 13)()
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
        H = np.array(H, dtype=self.real, copy=True)
        K = np.zeros(4 * self.patch_size, dtype=self.real)
        F = np.zeros(4 * self.compute_size, dtype=self.real)
        D = np.zeros(4 * self.compute_size, dtype=self.real)
        for _ in range(nsteps):
            for s in range(4):
                self.rk_substep(H, K, F, D, s)
        return H

    def eval_rhs(self, K):
        K = np.ascontiguousarray(K, dtype=self.real)
        F = np.zeros(4 * self.compute_size, dtype=self.real)
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
