"""Render serving: resident volumes, continuous batching, HBM admission.

The reference serves one request at a time: each endpoint of its demo server
mutates the process-wide DataManager and renders synchronously
(examples/server.py:303-318 /rotatevr, 497-513 /mprdata, 516-530
/rotatestretchedcpr; core/kernel.cu:471-568 cu_render allocates, launches and
copies back per frame). Here every volume, mask and TF stays resident in its
own ``Context`` and requests flow through ``RenderServer``:

* requests wait in a FIFO queue;
* whenever a launch slot is free, the longest FIFO prefix that fits the
  slot's limits (items, 16x16 tiles, device arena bytes = the HBM admission
  budget split over the slots) becomes one batch: CPR geometry is copied into
  the arena, ``mk_batch_submit`` replays the slot's graph on the slot's
  stream, and the outputs are copied back into the slot's pinned staging
  buffer on the same stream;
* slots run on separate streams, so a batch formed from requests that arrived
  while another batch runs overlaps it on the device;
* a request that can never fit (larger than a slot's arena or tile limit) is
  rejected at admission with an error; it never blocks the queue.

``plan_batch`` holds the admission rule and runs without a GPU.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _capi
from ._capi import (MK_ITEM_CPR, MK_ITEM_PLANE, MK_ITEM_RENDER, MK_SAMPLE_LINEAR, BatchItem, PlaneParams,
                    RenderParams)

TILE = 16
ALIGN = 256

KIND_RENDER, KIND_PLANE, KIND_CPR = "render", "plane", "cpr"


@dataclass(eq=False)
class Request:
    """One request. kind: "render" (VR / MIP / Surface by params.render_type),
    "plane" (slab MPR) or "cpr". volume: index of the server's context."""
    kind: str
    volume: int
    render: RenderParams | None = None
    plane: PlaneParams | None = None
    cpr_width: int = 0
    cpr_points: np.ndarray | None = None   # (h, 3) float64, voxel units
    cpr_dirs: np.ndarray | None = None     # (h, 3) float64
    cpr_invert_z: bool = False
    cpr_sampling: int = MK_SAMPLE_LINEAR
    t_arrival: float = 0.0
    tag: object = None
    solo: bool = False   # set when a batch holding it failed validation: retried in a batch of its own
    # filled by the server
    result: np.ndarray | None = None
    error: str | None = None
    t_admit: float = 0.0
    t_rendered: float = 0.0   # encode mode: when the host saw the texts' lengths
    t_done: float = 0.0

    @property
    def size(self) -> tuple:
        if self.kind == KIND_RENDER:
            return self.render.width, self.render.height
        if self.kind == KIND_PLANE:
            return self.plane.width, self.plane.height
        return self.cpr_width, int(self.cpr_points.shape[0])

    def out_bytes(self) -> int:
        w, h = self.size
        return w * h * (3 if self.kind == KIND_RENDER else 2)

    def geo_bytes(self) -> int:
        return 2 * self.size[1] * 3 * 8 if self.kind == KIND_CPR else 0

    def tiles(self) -> int:
        w, h = self.size
        return ((w + TILE - 1) // TILE) * ((h + TILE - 1) // TILE)

    def encode_dims(self) -> tuple:
        """(width, height, channels) of the PNG the reference makes of this
        result: the RGB frame, or the int16 image read as RGBA of width W/2."""
        w, h = self.size
        return (w, h, 3) if self.kind == KIND_RENDER else (w // 2, h, 4)

    def b64_bytes(self) -> int:
        w, h, ch = self.encode_dims()
        lib = _capi.load()
        return int(lib.mk_base64_max_bytes(lib.mk_png_max_bytes(w, h, ch)))

    def arena_bytes(self, encode: bool = False) -> int:
        n = _align(self.out_bytes()) + _align(self.geo_bytes())
        return n + (_align(self.b64_bytes()) if encode else 0)


def _align(n: int) -> int:
    return (n + ALIGN - 1) // ALIGN * ALIGN


def _render_key(p: RenderParams) -> bytes:
    return bytes(p.alpha_wwwl)


def plan_batch(queue: list, max_items: int, max_tiles: int, arena_bytes: int, encode: bool = False):
    """Admission rule. Returns (batch, rejected): batch is the longest prefix
    of ``queue`` that fits one slot (items, tiles, arena bytes) and whose render
    items agree per volume on step_scale and per (volume, render type) on the
    label alpha/window table (the batch kernels share those per-volume tables);
    rejected holds requests that could never fit a slot. A request marked
    ``solo`` goes in a batch of its own. With ``encode`` a request also
    holds the bound of its base64 PNG text in the arena. Neither list is
    removed from ``queue``."""
    batch, rejected = [], []
    tiles = nbytes = 0
    steps, tables = {}, {}
    for r in queue:
        need = r.arena_bytes(encode)
        unencodable = encode and ((r.kind != KIND_RENDER and r.size[0] % 2) or r.b64_bytes() == 0)
        if r.tiles() > max_tiles or need > arena_bytes or unencodable:
            rejected.append(r)
            continue
        if batch and (r.solo or batch[0].solo):
            break
        if len(batch) == max_items or tiles + r.tiles() > max_tiles or nbytes + need > arena_bytes:
            break
        if r.kind == KIND_RENDER:
            p = r.render
            s = steps.get(r.volume)
            t = tables.get((r.volume, p.render_type))
            if (s is not None and s != p.step_scale) or (t is not None and t != _render_key(p)):
                break
            steps[r.volume] = p.step_scale
            tables[(r.volume, p.render_type)] = _render_key(p)
        batch.append(r)
        tiles += r.tiles()
        nbytes += need
    return batch, rejected


@dataclass
class _Slot:
    arena: object            # torch uint8 device tensor
    host: object             # torch uint8 pinned tensor
    stream: object           # torch.cuda.Stream
    event: object = None     # torch.cuda.Event after the copy back
    batch: list = field(default_factory=list)
    offsets: list = field(default_factory=list)
    # encoding: 1 = waiting for the text lengths, 2 = waiting for the texts
    phase: int = 0
    b64_offsets: list = field(default_factory=list)
    lens_dev: object = None
    lens_host: object = None
    png: object = None
    workspace: object = None


class RenderServer:
    """Continuous batching over ``contexts`` (one resident volume each).

    hbm_budget: device bytes for request outputs and CPR geometry, split
    evenly over ``nslots`` launch slots; a batch never exceeds its slot's
    share, and a request larger than a share is rejected.

    encode: results come back as the reference's wire format, base64 text of
    a PNG (RGB frames; int16 images as RGBA of width W/2), built on the device
    after the batch (mk_png_encode_async + mk_base64_encode_async); only the
    texts are copied back, in two steps (their lengths, then exactly their
    bytes)."""

    def __init__(self, contexts: list, hbm_budget: int = 256 << 20, nslots: int = 2, max_items: int = 64,
                 max_tiles: int = 32768, encode: bool = False):
        import torch

        if nslots < 1:
            raise ValueError("nslots must be >= 1")
        self.torch = torch
        self.contexts = list(contexts)
        self.batcher = _capi.Batcher(self.contexts, max_items, max_tiles, nslots)
        self.max_items, self.max_tiles = max_items, max_tiles
        self.slot_bytes = hbm_budget // nslots // ALIGN * ALIGN
        if self.slot_bytes <= 0:
            raise ValueError("hbm_budget too small")
        self.slots = [
            _Slot(arena=torch.empty(self.slot_bytes, dtype=torch.uint8, device="cuda"),
                  host=torch.empty(self.slot_bytes, dtype=torch.uint8, pin_memory=True),
                  stream=torch.cuda.Stream())
            for _ in range(nslots)
        ]
        self.encode = encode
        self.lib = _capi.load()
        if encode:
            for sl in self.slots:
                sl.lens_dev = torch.zeros(2 * max_items, dtype=torch.int64, device="cuda")
                sl.lens_host = torch.zeros(2 * max_items, dtype=torch.int64, pin_memory=True)
        self.queue: list = []
        self.done: list = []
        self.batches = 0
        self.batch_sizes: list = []

    def close(self) -> None:
        for s in self.slots:
            if s.event is not None:
                s.event.synchronize()
        self.batcher.close()

    # -- request flow -------------------------------------------------------
    def submit(self, r: Request) -> None:
        self.queue.append(r)

    def busy(self) -> bool:
        return bool(self.queue) or any(s.batch for s in self.slots)

    def _grow(self, s: _Slot, name: str, nbytes: int) -> None:
        t = getattr(s, name)
        if t is None or t.numel() < nbytes:
            setattr(s, name, self.torch.empty((nbytes + 255) // 256 * 256, dtype=self.torch.uint8, device="cuda"))

    def _encode_batch(self, s: _Slot, batch: list, offs: list, b64_offs: list) -> None:
        """PNG + base64 of every result on the slot's stream, then the text
        lengths to pinned memory."""
        lib = self.lib
        dims = [r.encode_dims() for r in batch]
        self._grow(s, "png", max(int(lib.mk_png_max_bytes(*d)) for d in dims))
        self._grow(s, "workspace", max(int(lib.mk_png_workspace_bytes(*d)) for d in dims))
        st = s.stream.cuda_stream
        lens = s.lens_dev.data_ptr()
        base = s.arena.data_ptr()
        for i, (r, off, bo, d) in enumerate(zip(batch, offs, b64_offs, dims)):
            cap = int(lib.mk_png_max_bytes(*d))
            ws = int(lib.mk_png_workspace_bytes(*d))
            st_ = lib.mk_png_encode_async(base + off, d[0], d[1], d[2], s.workspace.data_ptr(), ws, s.png.data_ptr(),
                                          cap, lens + 16 * i, st)
            if st_ == 0:
                st_ = lib.mk_base64_encode_async(s.png.data_ptr(), lens + 16 * i, cap, base + bo,
                                                 int(lib.mk_base64_max_bytes(cap)), lens + 16 * i + 8, st)
            if st_ != 0:
                raise _capi.MkError(st_, lib.mk_last_error(None).decode())
        s.lens_host[:2 * len(batch)].copy_(s.lens_dev[:2 * len(batch)], non_blocking=True)

    def _texts_back(self, s: _Slot) -> None:
        """Phase 2 of an encoded batch: copy exactly the texts back."""
        torch = self.torch
        now = time.perf_counter()
        for r in s.batch:
            r.t_rendered = now
        lens = s.lens_host.numpy()
        with torch.cuda.stream(s.stream):
            for i, bo in enumerate(s.b64_offsets):
                n = int(lens[2 * i + 1])
                s.host[bo:bo + n].copy_(s.arena[bo:bo + n], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(s.stream)
        s.event, s.phase = ev, 2

    def _finish(self, s: _Slot) -> None:
        host = s.host.numpy()
        now = time.perf_counter()
        if self.encode:
            lens = s.lens_host.numpy()
            for i, (r, bo) in enumerate(zip(s.batch, s.b64_offsets)):
                n = int(lens[2 * i + 1])
                r.result = host[bo:bo + n].tobytes().decode("ascii")
                r.t_done = now
                self.done.append(r)
            s.batch, s.offsets, s.b64_offsets, s.event, s.phase = [], [], [], None, 0
            return
        for r, off in zip(s.batch, s.offsets):
            w, h = r.size
            n = r.out_bytes()
            raw = host[off:off + n]
            if r.kind == KIND_RENDER:
                r.result = raw.reshape(h, w, 3).copy()
            else:
                r.result = raw.view(np.int16).reshape(h, w).copy()
            r.t_done = now
            self.done.append(r)
        s.batch, s.offsets, s.event = [], [], None

    def _launch(self, si: int, batch: list) -> None:
        torch = self.torch
        s = self.slots[si]
        offs, items, geo = [], [], []
        cur = 0
        for r in batch:
            offs.append(cur)
            cur += _align(r.out_bytes())
        out_end = cur
        host = s.host.numpy()
        for r, off in zip(batch, offs):
            it = BatchItem()
            it.context = r.volume
            it.out_offset = off
            if r.kind == KIND_RENDER:
                it.kind = MK_ITEM_RENDER
                it.render = r.render
            elif r.kind == KIND_PLANE:
                it.kind = MK_ITEM_PLANE
                it.plane = r.plane
            else:
                it.kind = MK_ITEM_CPR
                h = r.size[1]
                pts = np.ascontiguousarray(r.cpr_points, np.float64).reshape(h, 3)
                drs = np.ascontiguousarray(r.cpr_dirs, np.float64).reshape(h, 3)
                nb = h * 24
                host[cur:cur + nb] = pts.view(np.uint8).reshape(-1)
                host[cur + nb:cur + 2 * nb] = drs.view(np.uint8).reshape(-1)
                it.cpr_width, it.cpr_height = r.cpr_width, h
                it.cpr_invert_z, it.cpr_sampling = int(r.cpr_invert_z), r.cpr_sampling
                it.cpr_points_offset, it.cpr_dirs_offset = cur, cur + nb
                cur = _align(cur + 2 * nb)
            items.append(it)
        geo_end = cur
        b64_offs = []
        if self.encode:
            for r in batch:
                b64_offs.append(cur)
                cur = _align(cur + r.b64_bytes())
        now = time.perf_counter()
        for r in batch:
            r.t_admit = now
        with torch.cuda.stream(s.stream):
            if geo_end > out_end:   # CPR geometry; the text region is written on the device
                s.arena[out_end:geo_end].copy_(s.host[out_end:geo_end], non_blocking=True)
            try:
                self.batcher.submit(si, items, s.arena, self.slot_bytes, stream=s.stream.cuda_stream)
            except _capi.MkError as e:
                if e.status == _capi.MK_ERR_HIP or e.status == _capi.MK_ERR_NOMEM:
                    for r in batch:
                        r.error = str(e)
                        r.t_done = time.perf_counter()
                        self.done.append(r)
                    raise
                if len(batch) > 1:
                    # validation failed for some item: every request of the
                    # batch is retried alone, so only the bad one fails
                    for r in batch:
                        r.solo = True
                    self.queue[:0] = batch
                else:
                    batch[0].error = str(e)
                    batch[0].t_done = time.perf_counter()
                    self.done.append(batch[0])
                return
            if self.encode:
                self._encode_batch(s, batch, offs, b64_offs)
            else:
                s.host[:out_end].copy_(s.arena[:out_end], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(s.stream)
        s.event, s.batch, s.offsets, s.b64_offsets = ev, list(batch), offs, b64_offs
        s.phase = 1 if self.encode else 0
        self.batches += 1
        self.batch_sizes.append(len(batch))

    def poll(self) -> None:
        """Retires finished batches, then fills every free slot from the queue."""
        for s in self.slots:
            if s.batch and s.event.query():
                if s.phase == 1:
                    self._texts_back(s)
                else:
                    self._finish(s)
        for si, s in enumerate(self.slots):
            if s.batch or not self.queue:
                continue
            batch, rejected = plan_batch(self.queue, self.max_items, self.max_tiles, self.slot_bytes, self.encode)
            for r in rejected:
                r.error = (f"request needs {r.arena_bytes(self.encode)} arena bytes / {r.tiles()} tiles; a slot "
                           f"holds {self.slot_bytes} bytes / {self.max_tiles} tiles"
                           + ("; encoded images need an even plane width and rows of at most 24576 bytes"
                              if self.encode else ""))
                r.t_done = time.perf_counter()
                self.done.append(r)
            gone = {id(r) for r in batch} | {id(r) for r in rejected}
            self.queue = [r for r in self.queue if id(r) not in gone]
            if batch:
                self._launch(si, batch)

    def run(self, requests: list, t0: float | None = None) -> list:
        """Open-loop run: request i enters the queue at t0 + its t_arrival
        (seconds). Returns the requests with results, errors and times
        (t_arrival made absolute)."""
        pending = sorted(requests, key=lambda r: r.t_arrival)
        t0 = time.perf_counter() if t0 is None else t0
        for r in pending:
            r.t_arrival += t0
        i = 0
        while i < len(pending) or self.busy():
            now = time.perf_counter()
            while i < len(pending) and pending[i].t_arrival <= now:
                self.submit(pending[i])
                i += 1
            self.poll()
        return requests


def render_one(ctx, r: Request) -> np.ndarray:
    """The one-request plain path (mk_render / mk_render_plane / mk_render_cpr),
    synchronous, as the reference's server calls it."""
    if r.kind == KIND_RENDER:
        return ctx.render(r.render)
    if r.kind == KIND_PLANE:
        return ctx.render_plane(r.plane)
    return ctx.render_cpr(r.cpr_width, r.cpr_points, r.cpr_dirs, r.cpr_invert_z, r.cpr_sampling)


def run_sequential(contexts: list, requests: list, t0: float | None = None) -> list:
    """One request at a time through the plain path, in arrival order, with
    the same open-loop arrivals as RenderServer.run."""
    pending = sorted(requests, key=lambda r: r.t_arrival)
    t0 = time.perf_counter() if t0 is None else t0
    for r in pending:
        r.t_arrival += t0
        while time.perf_counter() < r.t_arrival:
            pass
        r.t_admit = time.perf_counter()
        r.result = render_one(contexts[r.volume], r)
        r.t_done = time.perf_counter()
    return requests


def latency_stats(requests: list) -> dict:
    ok = [r for r in requests if r.error is None]
    lat = np.array([r.t_done - r.t_arrival for r in ok]) * 1e3
    if lat.size == 0:
        return {"n": 0}
    span = max(r.t_done for r in ok) - min(r.t_arrival for r in ok)
    return {
        "n": int(lat.size),
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "max_ms": float(lat.max()),
        "mean_ms": float(lat.mean()),
        "frames_per_s": float(lat.size / span) if span > 0 else float("nan"),
    }
