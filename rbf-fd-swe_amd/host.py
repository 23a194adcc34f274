"""ctypes view of the C host library (libswe_host.so, include/swe_host.h).

`Problem` runs the reference driver's init pipeline (binary read, optional
RCM, padding, cfdl/sfdl layout, serial patch; swe_code/main/main.c:94-190)
in C and keeps numpy copies of the prepared arrays.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Layout(C.Structure):
    _fields_ = [("cfdl", C.c_int), ("sfdl", C.c_int), ("simd", C.c_int)]


class Gsmd(C.Structure):
    _fields_ = [("Nnodes", C.c_int), ("padded_Nnodes", C.c_int),
                ("Nnbr", C.c_int), ("padded_Nnbr", C.c_int),
                ("a", C.c_double), ("gh0", C.c_double), ("dt", C.c_double),
                ("gamma", C.c_double),
                ("x", _dp), ("y", _dp), ("z", _dp), ("f", _dp), ("ghm", _dp),
                ("p_u", _dp), ("p_v", _dp), ("p_w", _dp), ("gradghm", _dp),
                ("idx", _ip), ("Dx", _dp), ("Dy", _dp), ("Dz", _dp), ("L", _dp)]


class Patch(C.Structure):
    _fields_ = [("patch_size", C.c_int), ("compute_size", C.c_int),
                ("compute_pid_s", C.c_int), ("compute_pid_e", C.c_int),
                ("halo_pid_s", C.c_int), ("halo_pid_e", C.c_int),
                ("Nnbr", C.c_int), ("padded_Nnbr", C.c_int),
                ("gh0", C.c_double), ("dt", C.c_double),
                ("x", _dp), ("y", _dp), ("z", _dp), ("f", _dp), ("ghm", _dp),
                ("p_u", _dp), ("p_v", _dp), ("p_w", _dp), ("gradghm", _dp),
                ("idx", _ip), ("Dx", _dp), ("Dy", _dp), ("Dz", _dp), ("L", _dp)]


class ProblemStruct(C.Structure):
    _fields_ = [("g", Gsmd), ("layout", Layout), ("patch", Patch),
                ("H_init", _dp), ("H_100", _dp), ("mapping", _ip)]


class Timers(C.Structure):
    _fields_ = [("nattempts", C.c_int), ("attempt", C.c_int),
                ("t_init", C.c_double), ("t_dev", C.c_double),
                ("t_main", _dp), ("t_eval_rhs", _dp), ("t_eval_K", _dp),
                ("t_update_D", _dp), ("t_update_H", _dp)]


PHASES = ("t_main", "t_eval_rhs", "t_eval_K", "t_update_D", "t_update_H")


def process_profiling(phase_times: dict, nsteps: int) -> np.ndarray:
    """swe_process_profiling over per-attempt phase times (seconds; keys
    PHASES, equal lengths): [5 phases][avg, min, max, stddev] s/step."""
    arrs = [np.ascontiguousarray(phase_times[k], dtype=np.float64) for k in PHASES]
    t = Timers()
    t.nattempts = len(arrs[0])
    for k, a in zip(PHASES, arrs):
        setattr(t, k, a.ctypes.data_as(_dp))
    out = np.zeros(20)
    lib().swe_process_profiling(C.byref(t), int(nsteps), 0, out.ctypes.data_as(_dp))
    return out.reshape(5, 4)


LAYOUTS = {"default": (0, 0), "cfdl": (1, 0), "sfdl": (0, 1), "cfdl_sfdl": (1, 1)}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libswe_host.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        L.swe_problem_load.argtypes = [C.c_char_p, C.c_int, C.POINTER(Layout),
                                       C.POINTER(ProblemStruct)]
        L.swe_problem_load_ex.argtypes = [C.c_char_p, C.c_int, C.POINTER(Layout), C.c_int,
                                          C.POINTER(ProblemStruct)]
        L.swe_problem_load_fmt.argtypes = [C.c_char_p, C.c_int, C.POINTER(Layout), C.c_int,
                                           C.c_int, C.POINTER(ProblemStruct)]
        L.swe_problem_free.argtypes = [C.POINTER(ProblemStruct)]
        L.swe_get_gsmd_nc_ex.argtypes = [C.c_char_p, C.c_int, C.POINTER(Gsmd)]
        L.swe_get_state_nc_ex.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(_dp)]
        L.swe_nc2bin.argtypes = [C.c_char_p, C.c_char_p]
        L.swe_verify_output.argtypes = [C.POINTER(Patch), C.POINTER(Layout), _dp, _dp, C.c_double]
        L.swe_rcm_mapping.argtypes = [C.POINTER(Gsmd), C.POINTER(_ip)]
        L.swe_get_gsmd_bin.argtypes = [C.c_char_p, C.POINTER(Gsmd)]
        L.swe_gsmd_free.argtypes = [C.POINTER(Gsmd)]
        L.swe_svm_lin_id.argtypes = [C.POINTER(Layout), C.c_int, C.c_int]
        L.swe_svm_lin_id.restype = C.c_long
        L.swe_dm_lin_id.argtypes = [C.POINTER(Layout), C.c_int, C.c_int, C.c_int]
        L.swe_dm_lin_id.restype = C.c_long
        L.swe_vect_lin_id.argtypes = [C.POINTER(Layout), C.c_int, C.c_int]
        L.swe_vect_lin_id.restype = C.c_long
        L.swe_stencil_bandwidth.argtypes = [C.POINTER(Gsmd), _ip, _ip]
        L.swe_process_profiling.argtypes = [C.POINTER(Timers), C.c_int, C.c_int, _dp]
        L.swe_process_profiling.restype = None
        _lib = L
    return _lib


HOST_ERRORS = {-1: "EOPEN", -2: "EREAD", -3: "EARG", -4: "ENOMEM", -5: "EGRAPH",
               -6: "EFORMAT", -7: "ENOVAR", -8: "ERANGE"}


class HostError(RuntimeError):
    """A negative SWE_HOST_E* code of the C host (include/swe_host.h)."""

    def __init__(self, code: int, what: str):
        self.code = int(code)
        self.name = HOST_ERRORS.get(self.code, str(self.code))
        super().__init__(f"{what} failed with SWE_HOST_{self.name} ({self.code})")


def nc2bin(nc_path: str, bin_path: str) -> None:
    """The reference's nc2bin (swe_code/io/nc2bin.c) through the C host."""
    err = lib().swe_nc2bin(nc_path.encode(), bin_path.encode())
    if err:
        raise HostError(err, f"swe_nc2bin({nc_path})")


def make_layout(name: str = "cfdl", simd: int = 4) -> Layout:
    cf, sf = LAYOUTS[name]
    return Layout(cf, sf, simd)


def _copy(ptr, count, dtype=np.float64):
    if count == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(count,)).view(dtype).copy()


class Problem:
    """Prepared serial patch of one input file (numpy copies of the arrays the
    C host prepared; the C allocation is released before returning)."""

    def __init__(self, path: str, use_rcm: bool = True, layout: str = "cfdl",
                 simd: int = 4, fp32: bool = False, netcdf: bool = False):
        self._s = ProblemStruct()
        self.layout_name = layout
        self.layout = make_layout(layout, simd)
        self.fp32 = bool(fp32)   # values rounded as the reference's FP32 build reads them
        self.netcdf = bool(netcdf)   # input is a NetCDF classic file (SWE_USE_NETCDF=1)
        err = lib().swe_problem_load_fmt(path.encode(), int(bool(use_rcm)),
                                         C.byref(self.layout), int(self.fp32),
                                         int(self.netcdf), C.byref(self._s))
        if err:
            raise HostError(err, f"swe_problem_load({path})")
        g, p = self._s.g, self._s.patch
        self.n_nodes = g.Nnodes
        self.padded = g.padded_Nnodes
        self.nnbr, self.padded_nnbr = g.Nnbr, g.padded_Nnbr
        self.a, self.gh0, self.dt, self.gamma = g.a, g.gh0, g.dt, g.gamma
        self.compute_size, self.patch_size = p.compute_size, p.patch_size
        self.compute_pid_s = p.compute_pid_s
        Np, Wp = self.padded, self.padded_nnbr
        for nm in ("x", "y", "z", "f", "ghm"):
            setattr(self, nm, _copy(getattr(g, nm), Np))
        for nm in ("p_u", "p_v", "p_w", "gradghm"):
            setattr(self, nm, _copy(getattr(g, nm), 3 * Np))
        for nm in ("Dx", "Dy", "Dz", "L"):
            setattr(self, nm, _copy(getattr(g, nm), Np * Wp))
        self.idx = _copy(g.idx, Np * Wp, np.int32)
        self.H_init = _copy(self._s.H_init, 4 * self.patch_size)
        self.H_100 = _copy(self._s.H_100, 4 * self.patch_size)
        self.mapping = (_copy(self._s.mapping, self.n_nodes, np.int32)
                        if self._s.mapping else None)
        self._patch = Patch()
        C.memmove(C.byref(self._patch), C.byref(p), C.sizeof(Patch))
        lib().swe_problem_free(C.byref(self._s))
        self._s = None
        for nm, _t in Patch._fields_[10:]:
            setattr(self._patch, nm, None)

    @property
    def patch(self) -> Patch:
        """Patch sizes (array pointers cleared; the arrays are attributes)."""
        return self._patch

    def svm_index(self) -> np.ndarray:
        """(patch_size, 4) host-layout index of every (node, variable)."""
        S, n = self.layout.simd, self.patch_size
        pid = np.arange(n)[:, None]
        sv = np.arange(4)[None, :]
        if self.layout.cfdl:
            return 4 * pid + sv
        return (pid // S) * S * 4 + sv * S + pid % S

    def to_node_major(self, H: np.ndarray) -> np.ndarray:
        """Host-layout state -> (patch_size, 4)."""
        return np.asarray(H)[self.svm_index()]

    def from_node_major(self, Hn: np.ndarray) -> np.ndarray:
        out = np.empty(4 * self.patch_size)
        out[self.svm_index()] = Hn
        return out

    def verify(self, H1: np.ndarray, H2: np.ndarray, tol: float = 1e-9) -> int:
        H1 = np.ascontiguousarray(H1, dtype=np.float64)
        H2 = np.ascontiguousarray(H2, dtype=np.float64)
        return lib().swe_verify_output(C.byref(self._patch), C.byref(self.layout),
                                       H1.ctypes.data_as(_dp), H2.ctypes.data_as(_dp), tol)
