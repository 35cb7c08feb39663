"""PNG + base64 encoding of rendered images on the device.

The reference turns each served image into a base64 PNG string on the host:
``GetVRData_pngString`` (RGB frame) and ``GetPlaneData_pngString`` (int16
plane read as an RGBA image of width W/2), core/HelloMonkey.cpp:212-259,
333-413. ``Encoder`` does the same on the GPU through ``mk_png_encode_async``
and ``mk_base64_encode_async`` (include/monkeygl_hip.h): the image never
leaves HBM until its (smaller) base64 text is copied back.
"""
from __future__ import annotations

import ctypes as C

from . import _capi
from ._capi import MK_OK, MkError


def _check(lib, st: int) -> None:
    if st != MK_OK:
        raise MkError(st, lib.mk_last_error(None).decode())


class Encoder:
    """Device buffers for images up to max_width x max_height x channels."""

    def __init__(self, max_width: int, max_height: int, channels: int, device: str = "cuda"):
        import torch

        self.lib = _capi.load()
        self.torch = torch
        self.channels = channels
        self.png_cap = int(self.lib.mk_png_max_bytes(max_width, max_height, channels))
        ws = int(self.lib.mk_png_workspace_bytes(max_width, max_height, channels))
        if self.png_cap == 0 or ws == 0:
            raise ValueError(f"unsupported image {max_width}x{max_height}x{channels}")
        self.b64_cap = int(self.lib.mk_base64_max_bytes(self.png_cap))
        self.workspace = torch.empty(ws, dtype=torch.uint8, device=device)
        self.png = torch.empty((self.png_cap + 255) // 256 * 256, dtype=torch.uint8, device=device)
        self.b64 = torch.empty(self.b64_cap, dtype=torch.uint8, device=device)
        self.lens = torch.zeros(2, dtype=torch.int64, device=device)   # png length, base64 length
        self.max_width, self.max_height = max_width, max_height

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

    def png_bytes(self) -> bytes:
        """The last PNG (synchronises)."""
        n = int(self.lens[0].item())
        return bytes(self.png[:n].cpu().numpy())

    def base64_text(self) -> str:
        """The last base64 text (synchronises)."""
        n = int(self.lens[1].item())
        return bytes(self.b64[:n].cpu().numpy()).decode("ascii")


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
