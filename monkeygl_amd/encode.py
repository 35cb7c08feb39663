"""PNG + base64 encoding of rendered images on the device.

The reference turns each served image into a base64 PNG string on the host:
``GetVRData_pngString`` (RGB frame) and ``GetPlaneData_pngString`` (int16
plane read as an RGBA image of width W/2), core/HelloMonkey.cpp:212-259,
333-413. ``Encoder`` does the same on the GPU through ``mk_png_encode_async``
and ``mk_base64_encode_async`` (include/monkeygl_hip.h): the image never
leaves HBM until its (smaller) base64 text is copied back.
"""
from __future__ import annotations

import numpy as np

from . import _capi
from ._capi import MK_OK, MkError


def _check(lib, st: int) -> None:
    if st != MK_OK:
        raise MkError(st, lib.mk_last_error(None).decode())


class Encoder:
    """Device buffers for images up to max_width x max_height x channels.

    With a context (``ctx``) the buffers are the context's device memory
    (mk_device_alloc) and the text comes back through mk_copy_to_host: no
    device-array library is needed, and callers encode on the context's
    stream. Without one they are torch tensors on ``device``."""

    def __init__(self, max_width: int, max_height: int, channels: int, device: str = "cuda", ctx=None):
        self.lib = _capi.load()
        self.ctx = ctx
        self.device = device
        self.channels = channels
        self.png_cap = int(self.lib.mk_png_max_bytes(max_width, max_height, channels))
        ws = int(self.lib.mk_png_workspace_bytes(max_width, max_height, channels))
        if self.png_cap == 0 or ws == 0:
            raise ValueError(f"unsupported image {max_width}x{max_height}x{channels}")
        self.b64_cap = int(self.lib.mk_base64_max_bytes(self.png_cap))
        self.workspace = self._buffer(ws)
        self.png = self._buffer((self.png_cap + 255) // 256 * 256)
        self.b64 = self._buffer(self.b64_cap)
        self.lens = self._buffer(16)   # png length, base64 length (int64 each)
        if ctx is not None:
            self.lens.from_host(np.zeros(2, np.int64))
        else:
            self.lens.zero_()
        self.max_width, self.max_height = max_width, max_height

    def _buffer(self, nbytes: int):
        if self.ctx is not None:
            return _capi.DeviceBuffer(self.ctx, nbytes)
        import torch

        return torch.empty(nbytes, dtype=torch.uint8, device=self.device)

    def _host(self, buf, n: int) -> bytes:
        if self.ctx is not None:
            return buf.to_host(np.uint8, n).tobytes()
        return bytes(buf[:n].cpu().numpy())

    def encode_async(self, dev_pixels, width: int, height: int, channels: int | None = None, stream: int = 0,
                     base64: bool = True) -> None:
        ch = self.channels if channels is None else channels
        if width > self.max_width or height > self.max_height or ch != self.channels:
            raise ValueError("image larger than the encoder's buffers")
        cap = int(self.lib.mk_png_max_bytes(width, height, ch))
        ws = int(self.lib.mk_png_workspace_bytes(width, height, ch))
        lens = self.lens.data_ptr()
        _check(self.lib, self.lib.mk_png_encode_async(_capi._ptr(dev_pixels), width, height, ch,
                                                      self.workspace.data_ptr(), ws, self.png.data_ptr(), cap,
                                                      lens, stream))
        if base64:
            _check(self.lib, self.lib.mk_base64_encode_async(self.png.data_ptr(), lens, cap, self.b64.data_ptr(),
                                                             int(self.lib.mk_base64_max_bytes(cap)), lens + 8,
                                                             stream))

    def lengths(self) -> tuple:
        """(PNG bytes, base64 bytes) of the last encode (synchronises)."""
        if self.ctx is not None:
            a = self.lens.to_host(np.int64, 2)
        else:
            import torch

            a = self.lens.view(torch.int64).cpu().numpy()
        return int(a[0]), int(a[1])

    def png_bytes(self) -> bytes:
        """The last PNG (synchronises)."""
        return self._host(self.png, self.lengths()[0])

    def base64_text(self) -> str:
        """The last base64 text (synchronises)."""
        return self._host(self.b64, self.lengths()[1]).decode("ascii")


def encode_vr_png_base64(enc: Encoder, dev_rgb, width: int, height: int, stream: int = 0) -> str:
    """GetVRData_pngString: RGB24 frame -> base64 PNG."""
    enc.encode_async(dev_rgb, width, height, 3, stream=stream)
    return enc.base64_text()


def encode_plane_png_base64(enc: Encoder, dev_int16, width: int, height: int, stream: int = 0) -> str:
    """GetPlaneData_pngString: int16 plane (width even) read as RGBA of width/2 -> base64 PNG."""
    if width % 2:
        raise ValueError("plane width must be even (the reference pads it)")
    enc.encode_async(dev_int16, width // 2, height, 4, stream=stream)
    return enc.base64_text()
