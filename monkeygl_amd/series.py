"""Batch series of slab MPR images (GetBatchData) on the device.

The reference declares ``GetBatchData(std::vector<short*>&, BatchInfo)``
(core/HelloMonkey.cpp:415-420) and ``BatchInfo`` (core/BatchInfo.h:29-54) but
its renderer returns false (core/Render.cpp:117-120). Here a series is a list
of plane images (``mpr.batch_plane_params``) rendered by the batcher
(``mk_batch_submit`` with MK_ITEM_PLANE items) straight into one device
buffer: image k at byte offset k * width * height * 2, so the buffer is the
series as an (num, height, width) int16 array. Each image is the same as
``mk_render_plane`` of its parameters alone (same kernel code and flags).
"""
from __future__ import annotations

from . import _capi
from ._capi import MK_ITEM_PLANE, BatchItem


class SeriesRenderer:
    """Renders plane series over one context with a batcher of two slots;
    series longer than a batch are cut into batches of at most max_items
    images and max_tiles 16x16 tiles."""

    def __init__(self, ctx, max_items: int = 256, max_tiles: int = 1 << 20):
        self.ctx = ctx
        self.batcher = _capi.Batcher([ctx], max_items, max_tiles, nslots=2)
        self.max_items, self.max_tiles = max_items, max_tiles

    def close(self) -> None:
        for s in range(2):
            self.batcher.wait(s)
        self.batcher.close()

    def render_async(self, params: list, dev_out, stream: int = 0) -> None:
        """Images of ``params`` (equal sizes) into dev_out (int16, at least
        len(params) * width * height elements) on ``stream``. Returns once the
        last batch is queued; a slot is reused only after its batch finished,
        so a long series waits on the host for all but its last two batches."""
        if not params:
            return
        w, h = params[0].width, params[0].height
        if any(p.width != w or p.height != h for p in params):
            raise ValueError("the images of a series share one size")
        img_bytes = w * h * 2
        need = len(params) * img_bytes
        have = _capi._nbytes(dev_out)
        if have < need:
            raise ValueError(f"output holds {have} bytes; the series needs {need}")
        tiles_per = ((w + 15) // 16) * ((h + 15) // 16)
        if tiles_per > self.max_tiles:
            raise ValueError(f"an image needs {tiles_per} tiles; the batcher holds {self.max_tiles}")
        per_batch = max(1, min(self.max_items, self.max_tiles // tiles_per))
        slot = 0
        for k0 in range(0, len(params), per_batch):
            items = []
            for k in range(k0, min(len(params), k0 + per_batch)):
                it = BatchItem()
                it.kind = MK_ITEM_PLANE
                it.context = 0
                it.out_offset = k * img_bytes
                it.plane = params[k]
                items.append(it)
            self.batcher.wait(slot)
            self.batcher.submit(slot, items, dev_out, need, stream=stream)
            slot ^= 1

    def render(self, params: list):
        """The series as a host (num, height, width) int16 array, rendered on
        the context's stream into a buffer of the context's device."""
        import numpy as np

        n = len(params) * params[0].width * params[0].height
        out = _capi.DeviceBuffer(self.ctx, 2 * n)
        try:
            self.render_async(params, out, stream=self.ctx.stream)
            return out.to_host(np.int16, n).reshape(len(params), params[0].height, params[0].width)
        finally:
            out.close()
