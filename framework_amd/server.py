"""Python side of the mdmc request loop (mdmc_server_* in include/mdmc.h).

A Server owns a device-memory budget, admits extraction requests by their
footprint, batches them over a few streams and keeps results resident until
they are released.  `run_stream` replays a timed arrival stream against it
(the serving benchmark); `run_plain_loop` runs the same requests one at a time
through the single-volume entry points (the baseline it is compared with).
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from .mdmc import MDMCError, _check, _default_box, _DeviceArray, lib

_VP = ctypes.c_void_p

FIELD_DEVICE, FIELD_BLOBS, FIELD_GYROID = 0, 1, 2

# server option: requests of at least this many cells run alone through the
# single-volume kernels (0: every request may join a batch)
OPT_SOLO_CELLS = 100


class ServerConfig(ctypes.Structure):
    _fields_ = [("memory_budget", ctypes.c_uint64), ("max_batch", ctypes.c_uint32),
                ("num_streams", ctypes.c_uint32)]


class Request(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_uint32 * 3), ("iso", ctypes.c_float), ("xyz_min", ctypes.c_float * 3),
                ("xyz_max", ctypes.c_float * 3), ("field_kind", ctypes.c_int), ("field", _VP),
                ("seed", ctypes.c_uint64), ("nblobs", ctypes.c_uint32), ("sigma_lo", ctypes.c_double),
                ("sigma_hi", ctypes.c_double), ("amp_lo", ctypes.c_double), ("amp_hi", ctypes.c_double),
                ("period", ctypes.c_double), ("noise_amp", ctypes.c_double), ("tag", ctypes.c_uint64)]


class _Counts(ctypes.Structure):
    _fields_ = [("num_active", ctypes.c_uint32), ("num_vertices", ctypes.c_uint32),
                ("num_quads", ctypes.c_uint32), ("status", ctypes.c_uint32)]


class Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("tag", ctypes.c_uint64), ("status", ctypes.c_int), ("counts", _Counts),
                ("vertices", _VP), ("quads", _VP), ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double),
                ("t_done", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("queued", ctypes.c_uint64), ("in_flight", ctypes.c_uint64),
                ("memory_in_use", ctypes.c_uint64), ("memory_peak", ctypes.c_uint64),
                ("memory_budget", ctypes.c_uint64), ("busy_seconds", ctypes.c_double)]


def _bind():
    L = lib()
    if getattr(L, "_server_bound", False):
        return L
    L.mdmc_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(_VP)]
    L.mdmc_server_create.restype = ctypes.c_int
    L.mdmc_server_destroy.argtypes = [_VP]
    L.mdmc_server_destroy.restype = None
    L.mdmc_server_submit.argtypes = [_VP, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
    L.mdmc_server_submit.restype = ctypes.c_int
    L.mdmc_server_poll.argtypes = [_VP, ctypes.c_int, ctypes.POINTER(Result), ctypes.c_uint32,
                                   ctypes.POINTER(ctypes.c_uint32)]
    L.mdmc_server_poll.restype = ctypes.c_int
    L.mdmc_server_copy_result.argtypes = [_VP, ctypes.c_uint64, _VP, _VP]
    L.mdmc_server_copy_result.restype = ctypes.c_int
    L.mdmc_server_release.argtypes = [_VP, ctypes.c_uint64]
    L.mdmc_server_release.restype = ctypes.c_int
    L.mdmc_server_get_stats.argtypes = [_VP, ctypes.POINTER(ServerStats)]
    L.mdmc_server_get_stats.restype = ctypes.c_int
    L.mdmc_server_set_option.argtypes = [_VP, ctypes.c_int, ctypes.c_int]
    L.mdmc_server_set_option.restype = ctypes.c_int
    L.mdmc_server_now.argtypes = [_VP]
    L.mdmc_server_now.restype = ctypes.c_double
    L._server_bound = True
    return L


def make_request(dims_xyz, iso, kind="blobs", seed=0, nblobs=16, sigma=(8.0, 32.0), amp=(-1.0, 1.0),
                 period=24.0, noise_amp=0.05, field=None, box=None, tag=0) -> Request:
    """A request for a generated field ("blobs" / "gyroid") or a caller's device field (torch tensor)."""
    r = Request()
    X, Y, Z = dims_xyz
    r.dims[0], r.dims[1], r.dims[2] = X, Y, Z
    r.iso = float(iso)
    mn, mx = box if box is not None else _default_box((X, Y, Z))
    for k in range(3):
        r.xyz_min[k], r.xyz_max[k] = float(mn[k]), float(mx[k])
    if field is not None:
        r.field_kind = FIELD_DEVICE
        r.field = field.data_ptr()
    else:
        r.field_kind = {"blobs": FIELD_BLOBS, "gyroid": FIELD_GYROID}[kind]
    r.seed = int(seed)
    r.nblobs = int(nblobs)
    r.sigma_lo, r.sigma_hi = float(sigma[0]), float(sigma[1])
    r.amp_lo, r.amp_hi = float(amp[0]), float(amp[1])
    r.period = float(period)
    r.noise_amp = float(noise_amp)
    r.tag = int(tag)
    return r


class Server:
    def __init__(self, memory_budget: int, max_batch: int = 64, num_streams: int = 2):
        L = _bind()
        cfg = ServerConfig(int(memory_budget), int(max_batch), int(num_streams))
        h = _VP()
        _check(L.mdmc_server_create(ctypes.byref(cfg), ctypes.byref(h)), "mdmc_server_create")
        self._h = h
        self._buf = (Result * 1024)()

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().mdmc_server_destroy(self._h)
            self._h = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, req: Request) -> int:
        rid = ctypes.c_uint64()
        _check(lib().mdmc_server_submit(self._h, ctypes.byref(req), ctypes.byref(rid)), "mdmc_server_submit")
        return rid.value

    def poll(self, wait: bool = False) -> list:
        n = ctypes.c_uint32()
        _check(lib().mdmc_server_poll(self._h, int(wait), self._buf, len(self._buf), ctypes.byref(n)),
               "mdmc_server_poll")
        out = []
        for i in range(n.value):
            r = self._buf[i]
            out.append({"id": r.id, "tag": r.tag, "status": r.status, "num_active": r.counts.num_active,
                        "num_vertices": r.counts.num_vertices, "num_quads": r.counts.num_quads,
                        "t_submit": r.t_submit, "t_start": r.t_start, "t_done": r.t_done,
                        "vertices_ptr": r.vertices or 0, "quads_ptr": r.quads or 0})
        return out

    def copy_result(self, res: dict):
        V, Q = res["num_vertices"], res["num_quads"]
        v = np.empty((V, 3), np.float32)
        q = np.empty((Q, 4), np.int32)
        _check(lib().mdmc_server_copy_result(self._h, res["id"], _VP(v.ctypes.data if V else None),
                                             _VP(q.ctypes.data if Q else None)), "mdmc_server_copy_result")
        return v, q

    def result_tensors(self, res: dict):
        """(vertices (V,3), quads (Q,4)) torch tensors viewing the result's
        device memory in place (valid until the result is released)."""
        import torch

        V, Q = res["num_vertices"], res["num_quads"]
        v = (torch.as_tensor(_DeviceArray(res["vertices_ptr"], (V, 3), "<f4"), device="cuda") if V
             else torch.zeros((0, 3), dtype=torch.float32, device="cuda"))
        q = (torch.as_tensor(_DeviceArray(res["quads_ptr"], (Q, 4), "<i4"), device="cuda") if Q
             else torch.zeros((0, 4), dtype=torch.int32, device="cuda"))
        return v, q

    def release(self, rid: int) -> None:
        _check(lib().mdmc_server_release(self._h, int(rid)), "mdmc_server_release")

    def stats(self) -> dict:
        st = ServerStats()
        _check(lib().mdmc_server_get_stats(self._h, ctypes.byref(st)), "mdmc_server_get_stats")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def now(self) -> float:
        return lib().mdmc_server_now(self._h)

    def set_options(self, options: dict) -> None:
        """Launch options ({MDMC_OPT_* id: value}) for every lane, e.g. an
        autotuner's batch choice, or server options (OPT_SOLO_CELLS)."""
        for opt, val in options.items():
            _check(lib().mdmc_server_set_option(self._h, int(opt), int(val)), "mdmc_server_set_option")


def run_stream(server: Server, requests: list, arrivals: np.ndarray, keep: set | None = None,
               on_result=None) -> dict:
    """Submits requests[i] at arrivals[i] seconds (server clock, from now) and
    collects results until all have finished.  Results whose index is in `keep`
    are copied to the host before release; on_result(i, result) is called for
    every result before its release.  Returns latencies, timings and the kept
    meshes."""
    n = len(requests)
    t0 = server.now()
    done, lat, kept, status = 0, np.zeros(n), {}, np.zeros(n, np.int32)
    idx_of = {}
    i = 0
    while done < n:
        now = server.now() - t0
        while i < n and arrivals[i] <= now:
            idx_of[server.submit(requests[i])] = i
            i += 1
        res = server.poll(wait=False)
        if not res and i < n:
            # idle until the next arrival or a completion, without spinning hard
            time.sleep(min(max(arrivals[i] - (server.now() - t0), 0.0), 0.00005))
            continue
        if not res and i >= n:
            res = server.poll(wait=True)
            if not res:
                raise RuntimeError(f"server went idle with {n - done} request(s) unfinished "
                                   "(their results were all released, so none can start)")
        for r in res:
            k = idx_of[r["id"]]
            lat[k] = r["t_done"] - r["t_submit"]
            status[k] = r["status"]
            if keep is not None and k in keep and r["status"] == 0:
                kept[k] = server.copy_result(r)
            if on_result is not None:
                on_result(k, r)
            server.release(r["id"])
            done += 1
    t_end = server.now() - t0
    return {"latency": lat, "status": status, "wall": t_end, "meshes": kept}


def run_plain_loop(ctx, requests: list, gen_ctx=None) -> dict:
    """The same requests, one at a time through the single-volume entry points
    (generate the field, count, read the counts, allocate, emit, wait)."""
    import torch

    gen = gen_ctx or ctx
    t0 = time.perf_counter()
    lat = np.zeros(len(requests))
    for k, r in enumerate(requests):
        ts = time.perf_counter()
        X, Y, Z = r.dims[0], r.dims[1], r.dims[2]
        if r.field_kind == FIELD_BLOBS:
            f = gen.field_blobs((X, Y, Z), r.seed, r.nblobs, (r.sigma_lo, r.sigma_hi), (r.amp_lo, r.amp_hi))
        else:
            f = gen.field_gyroid((X, Y, Z), r.period, r.seed, r.noise_amp)
        v, q, c = ctx.extract(f, r.iso, tuple(r.xyz_min), tuple(r.xyz_max))
        del v, q, f
        lat[k] = time.perf_counter() - ts
    torch.cuda.synchronize()
    return {"latency": lat, "wall": time.perf_counter() - t0}


__all__ = ["Server", "Request", "make_request", "run_stream", "run_plain_loop", "FIELD_BLOBS", "FIELD_GYROID",
           "FIELD_DEVICE", "MDMCError"]
