"""ctypes binding of the continuous-batching forecast server (libswe_serve.so,
include/swe_serve.h) and a driver that replays timed request arrivals."""
from __future__ import annotations

import ctypes as C
import os
import time
from dataclasses import dataclass, field

import numpy as np

from . import runtime

_HERE = os.path.dirname(os.path.abspath(__file__))
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

EXPORTS = ("swe_srv_default_config", "swe_srv_create", "swe_srv_destroy", "swe_srv_add_mesh",
           "swe_srv_submit", "swe_srv_pump", "swe_srv_wait", "swe_srv_completed",
           "swe_srv_get_stats", "swe_srv_add_mesh_ex", "swe_srv_submit_ex")


class SrvConfig(C.Structure):
    _fields_ = [("hbm_cap_bytes", C.c_longlong), ("slots_per_batch", C.c_int),
                ("member_block", C.c_int), ("chunk_steps", C.c_int),
                ("max_batches_per_mesh", C.c_int), ("flags", C.c_uint), ("device", C.c_int),
                ("tiled", C.c_int)]


class SrvStats(C.Structure):
    _fields_ = [("member_steps_done", C.c_longlong), ("requests_done", C.c_longlong),
                ("requests_pending", C.c_longlong), ("hbm_in_use", C.c_longlong),
                ("hbm_peak", C.c_longlong), ("batches", C.c_int), ("meshes_resident", C.c_int),
                ("chunks_launched", C.c_longlong), ("evictions", C.c_longlong)]


_lib = None


def library_path() -> str:
    return os.path.join(_HERE, "libswe_serve.so")


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        runtime.lib()   # libswe_hip first (rpath $ORIGIN also finds it)
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.swe_srv_default_config.argtypes = [C.POINTER(SrvConfig)]
        L.swe_srv_default_config.restype = None
        L.swe_srv_create.argtypes = [C.POINTER(SrvConfig), C.POINTER(vp)]
        L.swe_srv_destroy.argtypes = [vp]
        L.swe_srv_destroy.restype = None
        L.swe_srv_add_mesh.argtypes = [vp, C.POINTER(runtime.SweHipStatic), C.POINTER(C.c_int)]
        L.swe_srv_submit.argtypes = [vp, C.c_int, C.c_int, C.POINTER(_dp), C.POINTER(_dp),
                                     C.c_int, C.c_longlong, C.POINTER(C.c_int)]
        L.swe_srv_add_mesh_ex.argtypes = [vp, C.POINTER(runtime.SweHipStatic), _ip, C.c_int,
                                          C.POINTER(C.c_int)]
        L.swe_srv_submit_ex.argtypes = [vp, C.c_int, C.c_int, C.POINTER(_dp), C.POINTER(_dp),
                                        C.c_int, C.c_longlong, C.c_char_p, C.POINTER(C.c_int)]
        L.swe_srv_pump.argtypes = [vp, C.POINTER(C.c_int)]
        L.swe_srv_wait.argtypes = [vp]
        L.swe_srv_completed.argtypes = [vp, C.POINTER(C.c_longlong), _dp, _dp, C.c_int,
                                        C.POINTER(C.c_int)]
        L.swe_srv_get_stats.argtypes = [vp, C.POINTER(SrvStats)]
        _lib = L
    return _lib


def _chk(rc, where):
    if rc != 0:
        raise runtime.SweHipError(rc, where)


@dataclass
class Request:
    """One forecast: initial states (host layout) of its members."""
    mesh: int
    inits: list
    nsteps: int
    tag: int
    arrival_s: float = 0.0
    results: list = field(default_factory=list)
    latency_s: float = float("nan")
    queue_s: float = float("nan")
    result_path: str | None = None   # result file written by the server at completion


class Server:
    def __init__(self, hbm_cap_bytes: int = 200 << 30, slots: int = 64, member_block: int = 16,
                 chunk_steps: int = 16, max_batches: int = 64, exact: bool = False,
                 device: int = 0, tiled: bool = False):
        cfg = SrvConfig()
        lib().swe_srv_default_config(C.byref(cfg))
        cfg.hbm_cap_bytes = int(hbm_cap_bytes)
        cfg.slots_per_batch = int(slots)
        cfg.member_block = int(member_block)
        cfg.chunk_steps = int(chunk_steps)
        cfg.max_batches_per_mesh = int(max_batches)
        cfg.flags = runtime.SWE_HIP_EXACT if exact else 0
        cfg.device = int(device)
        cfg.tiled = int(bool(tiled))
        h = C.c_void_p()
        _chk(lib().swe_srv_create(C.byref(cfg), C.byref(h)), "swe_srv_create")
        self._h = h
        self._keep = []          # host arrays borrowed by the server
        self._live = {}          # req id -> Request
        self.problems = []

    def add_mesh(self, prob) -> int:
        s, keep = runtime.static_from_problem(prob)
        mapping = None
        if prob.mapping is not None:
            mapping = np.ascontiguousarray(prob.mapping, dtype=np.int32)
            keep.append(mapping)
        self._keep.append((s, keep))
        mid = C.c_int()
        _chk(lib().swe_srv_add_mesh_ex(self._h, C.byref(s),
                                       mapping.ctypes.data_as(_ip) if mapping is not None
                                       else None, int(prob.n_nodes), C.byref(mid)),
             "swe_srv_add_mesh_ex")
        self.problems.append(prob)
        return mid.value

    def submit(self, req: Request) -> int:
        prob = self.problems[req.mesh]
        n = 4 * prob.patch_size
        inits = [np.ascontiguousarray(a, dtype=np.float64) for a in req.inits]
        for a in inits:
            if a.size != n:
                raise ValueError("initial state size does not match the mesh")
        req.results = [np.empty(n) for _ in inits]
        pin = (_dp * len(inits))(*[a.ctypes.data_as(_dp) for a in inits])
        pout = (_dp * len(inits))(*[a.ctypes.data_as(_dp) for a in req.results])
        rid = C.c_int()
        path = req.result_path.encode() if req.result_path else None
        _chk(lib().swe_srv_submit_ex(self._h, req.mesh, len(inits), pin, pout, int(req.nsteps),
                                     int(req.tag), path, C.byref(rid)), "swe_srv_submit_ex")
        self._live[req.tag] = (req, inits, pin, pout)
        return rid.value

    def pump(self) -> int:
        p = C.c_int()
        _chk(lib().swe_srv_pump(self._h, C.byref(p)), "swe_srv_pump")
        return p.value

    def wait(self) -> None:
        _chk(lib().swe_srv_wait(self._h), "swe_srv_wait")

    def drain(self) -> list:
        out = []
        tags = (C.c_longlong * 256)()
        lat = (C.c_double * 256)()
        que = (C.c_double * 256)()
        cnt = C.c_int()
        while True:
            _chk(lib().swe_srv_completed(self._h, tags, lat, que, 256, C.byref(cnt)),
                 "swe_srv_completed")
            for i in range(cnt.value):
                req = self._live.pop(tags[i])[0]
                req.latency_s, req.queue_s = lat[i], que[i]
                out.append(req)
            if cnt.value < 256:
                return out

    def stats(self) -> dict:
        st = SrvStats()
        _chk(lib().swe_srv_get_stats(self._h, C.byref(st)), "swe_srv_get_stats")
        return {k: getattr(st, k) for k, _ in SrvStats._fields_}

    def run(self, requests: list, on_done=None) -> dict:
        """Replays `requests` (sorted by arrival_s, seconds from now); returns
        when all are done.  on_done(req) is called as each completes (e.g.
        to check and drop its result arrays)."""
        order = sorted(requests, key=lambda r: r.arrival_s)
        t0 = time.perf_counter()
        nxt = 0
        finished = 0
        while finished < len(order):
            now = time.perf_counter() - t0
            while nxt < len(order) and order[nxt].arrival_s <= now:
                self.submit(order[nxt])
                nxt += 1
            progress = self.pump()
            for req in self.drain():
                finished += 1
                if on_done:
                    on_done(req)
            if not progress:
                if nxt < len(order) and not self._live:
                    time.sleep(max(0.0, min(0.002, order[nxt].arrival_s - now)))
                else:
                    time.sleep(20e-6)
        makespan = time.perf_counter() - t0
        return {"makespan_s": makespan, **self.stats()}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            lib().swe_srv_destroy(self._h)
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


def member_pool(pkg_mesh, prob, count: int, seed: int) -> list:
    """`count` seeded ensemble initial states of one mesh, host layout of
    `prob` (perturbed TC5-like states from the generator)."""
    case = pkg_mesh.read_bin(prob.path)
    E = pkg_mesh.perturb_members(case, count, seed=seed)
    out = []
    for Hn in E:
        if prob.mapping is not None:
            Hn = Hn[prob.mapping]
        full = np.concatenate([Hn, np.repeat(Hn[-1:], prob.padded - prob.n_nodes, 0)])
        out.append(prob.from_node_major(full))
    return out


def poisson_mix(pools: list, n_requests: int, rate_per_s: float, seed: int,
                members=(1, 32), steps=(50, 500)) -> list:
    """Requests with Poisson arrivals: mesh uniform over the pools, members
    and step count uniform in the given inclusive ranges; member states
    drawn from the mesh's pool."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for tag in range(n_requests):
        t += float(rng.exponential(1.0 / rate_per_s))
        mid = int(rng.integers(len(pools)))
        B = int(rng.integers(members[0], members[1] + 1))
        pool = pools[mid]
        first = int(rng.integers(len(pool)))
        inits = [pool[(first + m) % len(pool)] for m in range(B)]
        reqs.append(Request(mesh=mid, inits=inits, nsteps=int(rng.integers(steps[0], steps[1] + 1)),
                            tag=tag, arrival_s=t))
    return reqs
