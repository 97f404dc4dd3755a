"""Render server: continuous batching of render requests (resident mesh +
camera + resolution) over the binning pipeline, with FIFO admission under an
HBM budget for frame arenas and framebuffers.  Thin wrapper of the
axr_server_* C ABI; the scheduler and the device lanes are native
(serve_core.hpp, serve.hip).

    srv = Server(ctx, lanes=3, hbm_budget=64 << 30)
    t = srv.submit(mesh, 1280, 720, params)
    srv.wait(t)
    bgra, depth, _ = srv.read(t)
    srv.release(t)

`replay` drives an open-loop arrival schedule natively (submission at the
scheduled instants, results released on completion) and returns each
request's latency.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from .render import Color, Context, Mesh

QUEUED, RUNNING, DONE, FAILED = N.REQ_QUEUED, N.REQ_RUNNING, N.REQ_DONE, N.REQ_FAILED


def request(mesh: Mesh, width: int, height: int, params: N.AxrDrawParams) -> N.AxrServeRequest:
    q = N.AxrServeRequest()
    q.mesh = mesh.handle.value
    q.width = width
    q.height = height
    q.params = params
    return q


class Server:
    def __init__(self, ctx: Context, lanes: int = 3, max_batch_views: int = 64, max_batch_faces: int = 8 << 20,
                 max_batch_tiles: int = 1 << 20, hbm_budget: int = 0, clear: Optional[Color] = None,
                 with_ids: bool = False, max_batch_pixels: int = 0):
        self.ctx = ctx
        self._lib = ctx._lib
        cfg = N.AxrServerConfig()
        cfg.lanes = lanes
        cfg.max_batch_views = max_batch_views
        cfg.max_batch_faces = max_batch_faces
        cfg.max_batch_tiles = max_batch_tiles
        cfg.max_batch_pixels = max_batch_pixels
        cfg.hbm_budget = hbm_budget
        cfg.clear_bgra = (clear or Color(0, 0, 0, 255)).packed_bgra()
        cfg.with_ids = 1 if with_ids else 0
        self.with_ids = with_ids
        h = C.c_void_p()
        ctx.check(self._lib.axr_server_create(ctx.handle, C.byref(cfg), C.byref(h)))
        self.handle = h
        self._sizes = {}

    def check(self, code: int) -> None:
        if code != N.AXR_OK:
            raw = self._lib.axr_server_last_error(self.handle)
            raise N.AxrError(code, raw.decode() if raw else "")

    def submit(self, mesh: Mesh, width: int, height: int, params: N.AxrDrawParams) -> int:
        q = request(mesh, width, height, params)
        t = C.c_uint64()
        self.check(self._lib.axr_server_submit(self.handle, C.byref(q), C.byref(t)))
        self._sizes[t.value] = (width, height)
        return t.value

    def poll(self, ticket: int) -> int:
        st = C.c_int()
        self.check(self._lib.axr_server_poll(self.handle, ticket, C.byref(st)))
        return st.value

    def wait(self, ticket: int, timeout_s: float = 60.0) -> None:
        self.check(self._lib.axr_server_wait(self.handle, ticket, timeout_s))

    def read(self, ticket: int) -> Tuple[np.ndarray, np.ndarray, Optional[np.ndarray]]:
        w, h = self._sizes[ticket]
        bgra = np.empty((h, w, 4), np.uint8)
        depth = np.empty((h, w), np.float32)
        ids = np.empty((h, w), np.uint32) if self.with_ids else None
        self.check(self._lib.axr_server_read(self.handle, ticket, N.u8ptr(bgra), N.fptr(depth),
                                             N.u32ptr(ids) if ids is not None else None))
        return bgra, depth, ids

    def release(self, ticket: int) -> None:
        self.check(self._lib.axr_server_release(self.handle, ticket))
        self._sizes.pop(ticket, None)

    def pause(self, on: bool = True) -> None:
        """Queue without admitting (on) / admit the queue in order (off)."""
        self.check(self._lib.axr_server_pause(self.handle, 1 if on else 0))

    def stats(self) -> dict:
        s = N.AxrServerStats()
        self.check(self._lib.axr_server_get_stats(self.handle, C.byref(s)))
        return s.as_dict()

    def replay(self, requests: Sequence[N.AxrServeRequest], arrival_s: np.ndarray) -> np.ndarray:
        n = len(requests)
        arr = (N.AxrServeRequest * n)(*requests)
        arrival = np.ascontiguousarray(arrival_s, dtype=np.float64)
        lat = np.empty(n, np.float64)
        self.check(self._lib.axr_server_replay(self.handle, n, arr, arrival.ctypes.data_as(C.POINTER(C.c_double)),
                                               lat.ctypes.data_as(C.POINTER(C.c_double))))
        return lat

    def close(self) -> None:
        if self.handle:
            self._lib.axr_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
