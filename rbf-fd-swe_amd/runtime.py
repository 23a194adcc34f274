"""ctypes binding of libswe_hip.so (include/swe_hip.h).

There is no CPU fallback: if the library or a HIP device is missing, every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

SWE_HIP_EXACT = 1
SWE_HIP_FP32 = 2
_fp = C.POINTER(C.c_float)
KERNELS = {"auto": 0, "quad": 1, "lane": 2, "quadw": 3, "pair": 4, "tile": 5, "oct": 6,
           "resident": 7}
ORDERS = {"host": 0, "morton": 1, "hilbert": 2}

# Every function include/swe_hip.h declares.
EXPORTS = (
    "swe_hip_create", "swe_hip_destroy", "swe_hip_set_state", "swe_hip_get_state",
    "swe_hip_rk_substep", "swe_hip_rk_steps", "swe_hip_eval_rhs", "swe_hip_set_timing",
    "swe_hip_get_timers", "swe_hip_synchronize", "swe_hip_info",
    "swe_hip_state_device_ptr", "swe_hip_stream", "swe_hip_device_bytes",
    "swe_hip_strerror", "swe_hip_last_error", "swe_hip_create_ex", "swe_hip_query",
    "swe_hip_set_active", "swe_hip_move_member", "swe_hip_static_bytes",
    "swe_hip_tune_member_block",
    "swe_hip_tune_ensemble", "swe_hip_rk_steps_noalloc",
    "swe_hip_create32", "swe_hip_set_state32", "swe_hip_get_state32", "swe_hip_eval_rhs32",
    "swe_hip_tune_ex",
)


class SweHipStatic(C.Structure):
    _fields_ = [("compute_size", C.c_int), ("patch_size", C.c_int),
                ("compute_pid_s", C.c_int), ("Nnbr", C.c_int), ("padded_Nnbr", C.c_int),
                ("cfdl", C.c_int), ("sfdl", C.c_int), ("simd", C.c_int),
                ("gh0", C.c_double), ("dt", C.c_double),
                ("x", _dp), ("y", _dp), ("z", _dp), ("f", _dp), ("ghm", _dp),
                ("p_u", _dp), ("p_v", _dp), ("p_w", _dp), ("gradghm", _dp),
                ("idx", _ip), ("Dx", _dp), ("Dy", _dp), ("Dz", _dp), ("L", _dp)]


class SweHipStatic32(C.Structure):
    _fields_ = [("compute_size", C.c_int), ("patch_size", C.c_int),
                ("compute_pid_s", C.c_int), ("Nnbr", C.c_int), ("padded_Nnbr", C.c_int),
                ("cfdl", C.c_int), ("sfdl", C.c_int), ("simd", C.c_int),
                ("gh0", C.c_float), ("dt", C.c_float),
                ("x", _fp), ("y", _fp), ("z", _fp), ("f", _fp), ("ghm", _fp),
                ("p_u", _fp), ("p_v", _fp), ("p_w", _fp), ("gradghm", _fp),
                ("idx", _ip), ("Dx", _fp), ("Dy", _fp), ("Dz", _fp), ("L", _fp)]


class SweHipOptions(C.Structure):
    _fields_ = [("members", C.c_int), ("flags", C.c_uint), ("device", C.c_int),
                ("member_block", C.c_int), ("share", C.c_void_p), ("kernel", C.c_int),
                ("prefetch", C.c_int), ("node_order", C.c_int), ("tiled", C.c_int),
                ("slot_block", C.c_int), ("tile_nodes", C.c_int), ("tile_rows", C.c_int)]


class SweHipTimers(C.Structure):
    _fields_ = [("t_eval_rhs", C.c_double), ("t_eval_K", C.c_double),
                ("t_update_D", C.c_double), ("t_update_H", C.c_double)]


_lib = None


def library_path() -> str:
    return os.path.join(_HERE, "libswe_hip.so")


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.swe_hip_create.argtypes = [C.POINTER(SweHipStatic), C.c_int, C.c_uint, C.c_int,
                                     C.POINTER(vp)]
        L.swe_hip_create_ex.argtypes = [C.POINTER(SweHipStatic), C.POINTER(SweHipOptions),
                                        C.POINTER(vp)]
        L.swe_hip_tune_member_block.argtypes = [vp, C.c_int, C.c_uint, C.c_int,
                                                C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.swe_hip_tune_ensemble.argtypes = [vp, C.c_int, C.c_uint, C.c_int, C.POINTER(C.c_int),
                                            C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.swe_hip_tune_ex.argtypes = [vp, C.c_int, C.c_uint, C.c_int, C.POINTER(SweHipOptions),
                                      C.POINTER(C.c_double)]
        L.swe_hip_query.argtypes = [vp]
        L.swe_hip_set_active.argtypes = [vp, C.c_int]
        L.swe_hip_move_member.argtypes = [vp, C.c_int, C.c_int]
        L.swe_hip_static_bytes.argtypes = [vp]
        L.swe_hip_static_bytes.restype = C.c_longlong
        L.swe_hip_destroy.argtypes = [vp]
        L.swe_hip_destroy.restype = None
        L.swe_hip_set_state.argtypes = [vp, C.c_int, _dp]
        L.swe_hip_get_state.argtypes = [vp, C.c_int, _dp]
        L.swe_hip_rk_substep.argtypes = [vp, C.c_int]
        L.swe_hip_rk_steps.argtypes = [vp, C.c_int]
        L.swe_hip_rk_steps_noalloc.argtypes = [vp, vp, C.c_int]
        L.swe_hip_create32.argtypes = [C.POINTER(SweHipStatic32), C.c_int, C.c_uint, C.c_int,
                                       C.POINTER(vp)]
        L.swe_hip_set_state32.argtypes = [vp, C.c_int, _fp]
        L.swe_hip_get_state32.argtypes = [vp, C.c_int, _fp]
        L.swe_hip_eval_rhs32.argtypes = [vp, C.c_int, _fp, _fp]
        L.swe_hip_eval_rhs.argtypes = [vp, C.c_int, _dp, _dp]
        L.swe_hip_set_timing.argtypes = [vp, C.c_int]
        L.swe_hip_get_timers.argtypes = [vp, C.POINTER(SweHipTimers), C.c_int]
        L.swe_hip_synchronize.argtypes = [vp]
        L.swe_hip_info.argtypes = [vp, _ip, _ip, _ip, C.POINTER(C.c_uint)]
        L.swe_hip_state_device_ptr.argtypes = [vp]
        L.swe_hip_state_device_ptr.restype = vp
        L.swe_hip_stream.argtypes = [vp]
        L.swe_hip_stream.restype = vp
        L.swe_hip_device_bytes.argtypes = [vp]
        L.swe_hip_device_bytes.restype = C.c_longlong
        L.swe_hip_strerror.argtypes = [C.c_int]
        L.swe_hip_strerror.restype = C.c_char_p
        L.swe_hip_last_error.restype = C.c_char_p
        _lib = L
    return _lib


class SweHipError(RuntimeError):
    def __init__(self, code: int, where: str):
        L = lib()
        msg = L.swe_hip_strerror(code).decode()
        detail = L.swe_hip_last_error().decode()
        super().__init__(f"{where}: {msg} ({code}): {detail}")
        self.code = code


def _chk(rc: int, where: str) -> None:
    if rc != 0:
        raise SweHipError(rc, where)


def static_from_problem(prob) -> tuple[SweHipStatic, list]:
    """swe_hip_static over the arrays of a host.Problem (kept alive by the
    returned list)."""
    keep = []

    def p(a, dtype=np.float64, t=_dp):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return a.ctypes.data_as(t)
    s = SweHipStatic()
    s.compute_size, s.patch_size = prob.compute_size, prob.patch_size
    s.compute_pid_s = prob.compute_pid_s
    s.Nnbr, s.padded_Nnbr = prob.nnbr, prob.padded_nnbr
    s.cfdl, s.sfdl, s.simd = prob.layout.cfdl, prob.layout.sfdl, prob.layout.simd
    s.gh0, s.dt = prob.gh0, prob.dt
    for nm in ("x", "y", "z", "f", "ghm", "p_u", "p_v", "p_w", "gradghm", "Dx", "Dy", "Dz", "L"):
        setattr(s, nm, p(getattr(prob, nm)))
    s.idx = p(prob.idx, np.int32, _ip)
    return s, keep


def static32_from_problem(prob) -> tuple[SweHipStatic32, list]:
    """swe_hip_static32 over a host.Problem read with fp32=True (its arrays
    are floats held as doubles; converted exactly)."""
    keep = []

    def p(a, dtype=np.float32, t=_fp):
        a32 = np.ascontiguousarray(np.asarray(a).astype(dtype))
        if dtype == np.float32 and not np.array_equal(a32.astype(np.float64), np.asarray(a)):
            raise ValueError("FP32 context: static data are not FP32-valued "
                             "(read the problem with fp32=True)")
        keep.append(a32)
        return a32.ctypes.data_as(t)
    s = SweHipStatic32()
    s.compute_size, s.patch_size = prob.compute_size, prob.patch_size
    s.compute_pid_s = prob.compute_pid_s
    s.Nnbr, s.padded_Nnbr = prob.nnbr, prob.padded_nnbr
    s.cfdl, s.sfdl, s.simd = prob.layout.cfdl, prob.layout.sfdl, prob.layout.simd
    s.gh0, s.dt = prob.gh0, prob.dt
    for nm in ("x", "y", "z", "f", "ghm", "p_u", "p_v", "p_w", "gradghm", "Dx", "Dy", "Dz", "L"):
        setattr(s, nm, p(getattr(prob, nm)))
    s.idx = p(prob.idx, np.int32, _ip)
    return s, keep


class HipContext:
    """One device context: static data of one mesh + `members` state slots.
    fp32=True: the reference's FP32 model type (states are float32 arrays)."""

    def __init__(self, prob, members: int = 1, exact: bool = False, device: int = 0,
                 member_block: int = 0, share: "HipContext | None" = None,
                 kernel: str = "auto", prefetch: int = 0, order: str = "host",
                 tiled: int = 0, slot_block: int = 0, tile_nodes: int = 0,
                 tile_rows: int = 0, fp32: bool = False):
        self.prob = prob
        self.members = int(members)
        self.exact = bool(exact)
        self.fp32 = bool(fp32)
        self.n_state = 4 * prob.patch_size
        if self.fp32 and kernel == "auto" and share is None:   # the FP32 drop-in entry point
            s, keep = static32_from_problem(prob)
            h = C.c_void_p()
            _chk(lib().swe_hip_create32(C.byref(s), self.members,
                                        SWE_HIP_EXACT if exact else 0, device, C.byref(h)),
                 "swe_hip_create32")
            self._share = None
            self._h = h
            return
        # FP32 with options: the double entry point with SWE_HIP_FP32 (the
        # problem's values are already the FP32 type's, exactly representable)
        flags = (SWE_HIP_EXACT if exact else 0) | (SWE_HIP_FP32 if self.fp32 else 0)
        opts = SweHipOptions(self.members, flags, device,
                             int(member_block), share.handle if share is not None else None,
                             KERNELS[kernel], int(prefetch), ORDERS[order], int(tiled),
                             int(slot_block), int(tile_nodes), int(tile_rows))
        h = C.c_void_p()
        if share is not None:
            rc = lib().swe_hip_create_ex(None, C.byref(opts), C.byref(h))
        else:
            s, keep = static_from_problem(prob)
            rc = lib().swe_hip_create_ex(C.byref(s), C.byref(opts), C.byref(h))
        _chk(rc, "swe_hip_create_ex")
        self._share = share            # keeps the static data's owner alive
        self._h = h

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise RuntimeError("context is closed")
        return self._h

    def set_state(self, H: np.ndarray, member: int = 0) -> None:
        if self.fp32:
            H = np.ascontiguousarray(H, dtype=np.float32)
        else:
            H = np.ascontiguousarray(H, dtype=np.float64)
        if H.size != self.n_state:
            raise ValueError(f"state has {H.size} values, expected {self.n_state}")
        if self.fp32:
            _chk(lib().swe_hip_set_state32(self.handle, member, H.ctypes.data_as(_fp)),
                 "swe_hip_set_state32")
        else:
            _chk(lib().swe_hip_set_state(self.handle, member, H.ctypes.data_as(_dp)),
                 "swe_hip_set_state")

    def get_state(self, member: int = 0) -> np.ndarray:
        if self.fp32:
            H = np.empty(self.n_state, dtype=np.float32)
            _chk(lib().swe_hip_get_state32(self.handle, member, H.ctypes.data_as(_fp)),
                 "swe_hip_get_state32")
            return H
        H = np.empty(self.n_state)
        _chk(lib().swe_hip_get_state(self.handle, member, H.ctypes.data_as(_dp)),
             "swe_hip_get_state")
        return H

    def rk_substep(self, s: int) -> None:
        _chk(lib().swe_hip_rk_substep(self.handle, int(s)), "swe_hip_rk_substep")

    def rk_steps(self, nsteps: int) -> None:
        _chk(lib().swe_hip_rk_steps(self.handle, int(nsteps)), "swe_hip_rk_steps")

    def rk_steps_noalloc(self, H_dev: int, nsteps: int) -> None:
        """nsteps fused RK4 steps of the device state at address H_dev
        (this context's device layout) with no device allocation."""
        _chk(lib().swe_hip_rk_steps_noalloc(self.handle, C.c_void_p(H_dev), int(nsteps)),
             "swe_hip_rk_steps_noalloc")

    def eval_rhs(self, K: np.ndarray, member: int = 0) -> np.ndarray:
        if self.fp32:
            K = np.ascontiguousarray(K, dtype=np.float32)
            F = np.empty(4 * self.prob.compute_size, dtype=np.float32)
            _chk(lib().swe_hip_eval_rhs32(self.handle, member, K.ctypes.data_as(_fp),
                                          F.ctypes.data_as(_fp)), "swe_hip_eval_rhs32")
            return F
        K = np.ascontiguousarray(K, dtype=np.float64)
        F = np.empty(4 * self.prob.compute_size)
        _chk(lib().swe_hip_eval_rhs(self.handle, member, K.ctypes.data_as(_dp),
                                    F.ctypes.data_as(_dp)), "swe_hip_eval_rhs")
        return F

    def set_timing(self, on: bool) -> None:
        _chk(lib().swe_hip_set_timing(self.handle, int(bool(on))), "swe_hip_set_timing")

    def timers(self, reset: bool = False) -> dict:
        t = SweHipTimers()
        _chk(lib().swe_hip_get_timers(self.handle, C.byref(t), int(bool(reset))),
             "swe_hip_get_timers")
        return {k: getattr(t, k) for k, _ in SweHipTimers._fields_}

    def synchronize(self) -> None:
        _chk(lib().swe_hip_synchronize(self.handle), "swe_hip_synchronize")

    def state_device_ptr(self) -> int:
        return lib().swe_hip_state_device_ptr(self.handle) or 0

    def stream(self) -> int:
        return lib().swe_hip_stream(self.handle) or 0

    def device_bytes(self) -> int:
        return int(lib().swe_hip_device_bytes(self.handle))

    def static_bytes(self) -> int:
        return int(lib().swe_hip_static_bytes(self.handle))

    def query(self) -> bool:
        rc = lib().swe_hip_query(self.handle)
        if rc < 0:
            _chk(rc, "swe_hip_query")
        return rc == 1

    def tune_member_block(self, members: int, steps: int = 5, exact: bool = False):
        """(best member block, ms per step) for a `members`-slot ensemble on
        this context's mesh."""
        mb = C.c_int()
        ms = C.c_double()
        _chk(lib().swe_hip_tune_member_block(self.handle, int(members),
                                             SWE_HIP_EXACT if exact else 0, int(steps),
                                             C.byref(mb), C.byref(ms)),
             "swe_hip_tune_member_block")
        return mb.value, ms.value

    def tune_ensemble(self, members: int, steps: int = 5, exact: bool = False):
        """(best member block, tiled, ms per step) over the member-block and
        LDS-tiled ensemble kernels on this context's mesh and node order."""
        mb = C.c_int()
        tiled = C.c_int()
        ms = C.c_double()
        _chk(lib().swe_hip_tune_ensemble(self.handle, int(members),
                                         SWE_HIP_EXACT if exact else 0, int(steps),
                                         C.byref(mb), C.byref(tiled), C.byref(ms)),
             "swe_hip_tune_ensemble")
        return mb.value, bool(tiled.value), ms.value

    def tune(self, members: int = 1, steps: int = 5, exact: bool = False):
        """(options dict for HipContext(..., share=self, **opts), ms per step):
        the fastest configuration swe_hip_tune_ex finds on this mesh."""
        o = SweHipOptions()
        ms = C.c_double()
        flags = (SWE_HIP_EXACT if exact else 0) | (SWE_HIP_FP32 if self.fp32 else 0)
        _chk(lib().swe_hip_tune_ex(self.handle, int(members), flags, int(steps), C.byref(o),
                                   C.byref(ms)), "swe_hip_tune_ex")
        kname = {v: k for k, v in KERNELS.items()}[o.kernel]
        opts = dict(members=o.members, member_block=o.member_block, kernel=kname,
                    prefetch=o.prefetch, tiled=o.tiled, tile_nodes=o.tile_nodes,
                    tile_rows=o.tile_rows)
        return opts, ms.value

    def set_active(self, active: int) -> None:
        _chk(lib().swe_hip_set_active(self.handle, int(active)), "swe_hip_set_active")

    def move_member(self, src: int, dst: int) -> None:
        _chk(lib().swe_hip_move_member(self.handle, int(src), int(dst)), "swe_hip_move_member")

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().swe_hip_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
