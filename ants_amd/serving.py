"""Batched and served SyN registration over the library's batch scheduler.

``Batch`` wraps ants_batch_* (include/ants_syn.h): `slots` registrations in
flight, each on its own stream, admitted in order against a device-memory
budget, with resident fixed templates.  ``serve`` replays a stream of
requests arriving at given times and reports per-request latency.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib as L
from .registration import SyNStage, stage_params


class BatchConfigC(C.Structure):
    _fields_ = [("slots", C.c_int), ("memory_budget", C.c_int64)]


class RequestC(C.Structure):
    _fields_ = [
        ("params", L.SynParamsC),
        ("fixed", L.ImageDesc),
        ("host_fixed", C.c_void_p),
        ("moving", L.ImageDesc),
        ("host_moving", C.c_void_p),
        ("host_moving_labels", C.c_void_p),
        ("use_template", C.c_int),
        ("template_index", C.c_int),
        ("host_out_fwd", C.c_void_p),
        ("host_out_inv", C.c_void_p),
        ("host_out_labels", C.c_void_p),
        ("user", C.c_void_p),
        ("moving_initial", C.POINTER(L.TransformC)),
        ("n_moving_initial", C.c_int),
        ("fixed_initial", C.POINTER(L.TransformC)),
        ("n_fixed_initial", C.c_int),
    ]


class ResultC(C.Structure):
    _fields_ = [
        ("id", C.c_int64),
        ("status", C.c_int),
        ("slot", C.c_int),
        ("t_submit", C.c_double),
        ("t_start", C.c_double),
        ("t_end", C.c_double),
        ("voxel_iterations", C.c_double),
        ("algorithmic_bytes", C.c_double),
        ("footprint_bytes", C.c_int64),
        ("user", C.c_void_p),
        ("error", C.c_char * 256),
    ]


def _bind():
    lib = L.lib()
    if getattr(lib, "_batch_bound", False):
        return lib
    vp = C.c_void_p
    lib.ants_batch_create.argtypes = [C.POINTER(BatchConfigC), C.POINTER(vp)]
    lib.ants_batch_destroy.argtypes = [vp]
    lib.ants_batch_destroy.restype = None
    lib.ants_batch_set_template.argtypes = [vp, C.c_int, C.POINTER(L.ImageDesc), vp]
    lib.ants_batch_submit.argtypes = [vp, C.POINTER(RequestC), C.POINTER(C.c_int64)]
    lib.ants_batch_poll.argtypes = [vp, C.POINTER(ResultC), C.c_int, C.POINTER(C.c_int)]
    lib.ants_batch_wait.argtypes = [vp, C.POINTER(ResultC), C.c_int, C.POINTER(C.c_int)]
    lib.ants_batch_pending.argtypes = [vp]
    lib.ants_batch_pending.restype = C.c_int64
    lib.ants_batch_cancel.argtypes = [vp, C.c_int64]
    lib.ants_batch_device_bytes.argtypes = [vp]
    lib.ants_batch_device_bytes.restype = C.c_int64
    lib.ants_batch_clock.restype = C.c_double
    lib.ants_syn_footprint.argtypes = [C.POINTER(L.SynParamsC), C.POINTER(L.ImageDesc), C.POINTER(L.ImageDesc)]
    lib.ants_syn_footprint.restype = C.c_int64
    lib._batch_bound = True
    return lib




@dataclass
class Request:
    stage: SyNStage
    moving: np.ndarray                     # numpy (z,y,x) float32
    fixed: Optional[np.ndarray] = None     # None: use a resident template
    template_index: int = 0
    moving_labels: Optional[np.ndarray] = None
    spacing: Sequence[float] = (1.0, 1.0, 1.0)
    precision: str = "fp32"
    want_fields: bool = False
    tag: str = ""
    moving_initial: Sequence = ()          # apply_transforms.AffineTransform list (-r / preceding linear stages)
    fixed_initial: Sequence = ()


@dataclass
class Result:
    id: int
    request: Request
    status: int
    slot: int
    t_submit: float
    t_start: float
    t_end: float
    voxel_iterations: float
    algorithmic_bytes: float
    footprint_bytes: int
    error: str
    fwd: Optional[np.ndarray] = None
    inv: Optional[np.ndarray] = None
    labels: Optional[np.ndarray] = None

    @property
    def latency(self) -> float:
        return self.t_end - self.t_submit


class Batch:
    def __init__(self, slots: int = 4, memory_budget: int = 0):
        self.lib = _bind()
        cfg = BatchConfigC(slots, memory_budget)
        h = C.c_void_p()
        L.check(self.lib.ants_batch_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.templates: Dict[int, tuple] = {}
        self.live: Dict[int, tuple] = {}  # id -> (Request, buffers kept alive, RequestC)

    def close(self):
        if self.h:
            self.lib.ants_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def clock(self) -> float:
        return self.lib.ants_batch_clock()

    def set_template(self, index: int, image: np.ndarray, spacing=None, origin=None):
        img = np.ascontiguousarray(image, dtype=np.float32)
        d = L.desc(tuple(reversed(img.shape)), spacing, origin)
        L.check(self.lib.ants_batch_set_template(self.h, index, C.byref(d), img.ctypes.data_as(C.c_void_p)))
        self.templates[index] = (img, d)

    def submit(self, req: Request) -> int:
        dim = req.moving.ndim
        r = RequestC()
        r.params = stage_params(req.stage, dim, req.precision)
        mov = np.ascontiguousarray(req.moving, dtype=np.float32)
        keep = [mov]
        r.moving = L.desc(tuple(reversed(mov.shape)), req.spacing)
        r.host_moving = mov.ctypes.data
        if req.fixed is None:
            r.use_template = 1
            r.template_index = req.template_index
            fdesc = self.templates[req.template_index][1]
        else:
            fix = np.ascontiguousarray(req.fixed, dtype=np.float32)
            keep.append(fix)
            r.fixed = L.desc(tuple(reversed(fix.shape)), req.spacing)
            r.host_fixed = fix.ctypes.data
            fdesc = r.fixed
        # outputs live on the last level grid of the fixed image
        out = L.ImageDesc()
        L.check(self.lib.ants_syn_level_geometry(C.byref(r.params), C.byref(fdesc), r.params.n_levels - 1,
                                                 C.byref(out)))
        oshape = tuple(int(out.size[k]) for k in reversed(range(dim)))
        fwd = inv = lab = None
        if req.want_fields:
            fwd = np.empty(oshape + (dim,), np.float32)
            inv = np.empty(oshape + (dim,), np.float32)
            r.host_out_fwd = fwd.ctypes.data
            r.host_out_inv = inv.ctypes.data
        if req.moving_labels is not None:
            ml = np.ascontiguousarray(req.moving_labels, dtype=np.int32)
            keep.append(ml)
            lab = np.empty(oshape, np.int32)
            r.host_moving_labels = ml.ctypes.data
            r.host_out_labels = lab.ctypes.data
        if req.moving_initial or req.fixed_initial:
            from .apply_transforms import transform_array
            mx, keep_m = transform_array(list(req.moving_initial), dim)
            fx, keep_f = transform_array(list(req.fixed_initial), dim)
            keep += [mx, fx] + keep_m + keep_f  # field buffers stay alive until the result returns
            r.moving_initial, r.n_moving_initial = mx, len(req.moving_initial)
            r.fixed_initial, r.n_fixed_initial = fx, len(req.fixed_initial)
        rid = C.c_int64()
        L.check(self.lib.ants_batch_submit(self.h, C.byref(r), C.byref(rid)))
        self.live[rid.value] = (req, keep, r, fwd, inv, lab)
        return rid.value

    def _collect(self, fn, max_results=64) -> List[Result]:
        arr = (ResultC * max_results)()
        n = C.c_int()
        L.check(fn(self.h, arr, max_results, C.byref(n)))
        out = []
        for i in range(n.value):
            a = arr[i]
            req, keep, r, fwd, inv, lab = self.live.pop(a.id)
            out.append(Result(a.id, req, a.status, a.slot, a.t_submit, a.t_start, a.t_end, a.voxel_iterations,
                              a.algorithmic_bytes, a.footprint_bytes, a.error.decode(errors="replace"), fwd, inv,
                              lab))
        return out

    def poll(self) -> List[Result]:
        return self._collect(self.lib.ants_batch_poll)

    def wait(self) -> List[Result]:
        return self._collect(self.lib.ants_batch_wait)

    def pending(self) -> int:
        return int(self.lib.ants_batch_pending(self.h))

    def cancel(self, rid: int):
        """Cancel a queued or running request (its result comes back with
        status ANTS_ERR_CANCELLED)."""
        L.check(self.lib.ants_batch_cancel(self.h, int(rid)))

    def device_bytes(self) -> int:
        """Device bytes held or reserved: templates, slot buffers, running requests."""
        return int(self.lib.ants_batch_device_bytes(self.h))

    def drain(self) -> List[Result]:
        out = []
        while self.pending() > 0:
            out += self.wait()
        return out


def run_all(batch: Batch, requests: Sequence[Request]) -> List[Result]:
    """Submit every request at once and wait for all of them."""
    for r in requests:
        batch.submit(r)
    return sorted(batch.drain(), key=lambda x: x.id)


def serve(batch: Batch, requests: Sequence[Request], arrival_times: Sequence[float]) -> List[Result]:
    """Submit each request at its arrival time (seconds from the start) and
    collect results as they complete (continuous batching: the scheduler
    admits queued requests whenever a slot and memory free up)."""
    t0 = batch.clock()
    results: List[Result] = []
    for req, ta in zip(requests, arrival_times):
        while True:
            now = batch.clock() - t0
            if now >= ta:
                break
            results += batch.poll()
            time.sleep(min(0.002, max(ta - now, 0.0)))
        batch.submit(req)
        results += batch.poll()
    results += batch.drain()
    return sorted(results, key=lambda x: x.id)


def latency_stats(results: Sequence[Result]) -> dict:
    lat = np.array([r.latency for r in results])
    t_first = min(r.t_submit for r in results)
    t_last = max(r.t_end for r in results)
    return {"n": len(results), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "pairs_per_s": len(results) / (t_last - t_first),
            "wall_s": t_last - t_first}
