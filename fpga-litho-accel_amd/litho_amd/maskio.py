"""Mask and image files: PNG <-> BIN (f32 row-major) and 8-bit mask levels.

The reference converts its mask PNGs (input/mask/*.png, the ICCAD 2013 T1-T10
set among them) to the f32 BIN layout the host flow reads with
tool/png_bin_converter.py: ``png_to_bin`` takes channel 0 of RGB(A) images
and divides by 255 (L164-183); ``bin_to_png`` min-max normalises to 0..255
(L104-140).  The same entry points are kept here with a self-contained PNG
codec (zlib from the standard library; no imaging package needed), so the
GPU host flow can be fed straight from the reference's PNG masks.

``mask_levels`` gives the 8-bit form the matrix-core batch path reads
(mask = level / 255, exact for binary and 8-bit grey masks).
"""
import os
import struct
import zlib

import numpy as np

__all__ = ["read_png", "write_png", "read_bin", "write_bin", "png_to_bin", "bin_to_png",
           "load_mask", "mask_levels"]

_SIG = b"\x89PNG\r\n\x1a\n"
_CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _unfilter(raw, height, stride, bpp):
    """Undo the five PNG scanline filters (None, Sub, Up, Average, Paeth)."""
    out = np.zeros((height, stride), np.uint8)
    prev = np.zeros(stride, np.int64)
    pos = 0
    for y in range(height):
        ftype = raw[pos]
        line = np.frombuffer(raw, np.uint8, stride, pos + 1).astype(np.int64)
        pos += stride + 1
        if ftype == 0:
            cur = line
        elif ftype == 1:  # Sub: running sum along each byte lane of the pixel
            cur = line.copy()
            for k in range(bpp):
                cur[k::bpp] = np.cumsum(line[k::bpp]) & 0xFF
        elif ftype == 2:  # Up
            cur = (line + prev) & 0xFF
        elif ftype in (3, 4):
            cur = line.copy()
            lv, pv = cur.tolist(), prev.tolist()
            for i in range(stride):
                a = lv[i - bpp] if i >= bpp else 0
                b = pv[i]
                if ftype == 3:
                    lv[i] = (lv[i] + ((a + b) >> 1)) & 0xFF
                else:
                    c = pv[i - bpp] if i >= bpp else 0
                    lv[i] = (lv[i] + _paeth(a, b, c)) & 0xFF
            cur = np.array(lv, np.int64)
        else:
            raise ValueError(f"PNG: bad filter type {ftype} in row {y}")
        out[y] = cur
        prev = cur
    return out


def read_png(path):
    """Decode a non-interlaced PNG to uint8: (H, W) for grey / palette images,
    (H, W, C) for grey+alpha, RGB and RGBA.  16-bit samples keep their high
    byte; 1/2/4-bit grey is scaled to 0..255, palette indices are looked up."""
    data = open(path, "rb").read()
    if data[:8] != _SIG:
        raise ValueError(f"{path}: not a PNG file")
    pos, idat, plte, hdr = 8, [], None, None
    while pos < len(data):
        n, tag = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        if zlib.crc32(tag + body) & 0xFFFFFFFF != struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]:
            raise ValueError(f"{path}: CRC mismatch in {tag!r}")
        if tag == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif tag == b"PLTE":
            plte = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif tag == b"IDAT":
            idat.append(body)
        elif tag == b"IEND":
            break
        pos += 12 + n
    if hdr is None:
        raise ValueError(f"{path}: no IHDR")
    w, h, depth, ctype, _, _, interlace = hdr
    if interlace:
        raise ValueError(f"{path}: interlaced PNG is not supported")
    if ctype not in _CHANNELS:
        raise ValueError(f"{path}: colour type {ctype}")
    ch = _CHANNELS[ctype]
    bits = ch * depth
    stride = (w * bits + 7) // 8
    bpp = max(1, bits // 8)
    rows = _unfilter(zlib.decompress(b"".join(idat)), h, stride, bpp)
    if depth == 8:
        px = rows.reshape(h, w, ch)
    elif depth == 16:
        px = rows.reshape(h, w, ch, 2)[..., 0]
    else:  # 1, 2, 4 bits: grey or palette, one channel
        per = 8 // depth
        shifts = np.arange(per - 1, -1, -1) * depth
        v = (rows[:, :, None] >> shifts) & ((1 << depth) - 1)
        px = v.reshape(h, -1)[:, :w, None].astype(np.uint8)
        if ctype == 0:
            px = (px.astype(np.uint32) * 255 // ((1 << depth) - 1)).astype(np.uint8)
    if ctype == 3:
        if plte is None:
            raise ValueError(f"{path}: palette image without PLTE")
        px = plte[px[..., 0]]
        return np.ascontiguousarray(px)
    return np.ascontiguousarray(px[..., 0] if ch == 1 else px)


def _chunk(tag, body):
    return struct.pack(">I", len(body)) + tag + body + struct.pack(
        ">I", zlib.crc32(tag + body) & 0xFFFFFFFF)


def write_png(path, img, filters=0):
    """Write an (H, W) uint8 image as 8-bit grey PNG.  `filters` is one filter
    type 0-4 for every row or a sequence giving one per row."""
    img = np.ascontiguousarray(img, np.uint8)
    if img.ndim != 2:
        raise ValueError("write_png: 2-D grey image expected")
    h, w = img.shape
    fl = [filters] * h if np.isscalar(filters) else list(filters)
    out = bytearray()
    prev = np.zeros(w, np.int64)
    for y in range(h):
        cur = img[y].astype(np.int64)
        left = np.concatenate([[0], cur[:-1]])
        upleft = np.concatenate([[0], prev[:-1]])
        f = fl[y]
        if f == 0:
            line = cur
        elif f == 1:
            line = cur - left
        elif f == 2:
            line = cur - prev
        elif f == 3:
            line = cur - ((left + prev) >> 1)
        elif f == 4:
            pred = np.array([_paeth(a, b, c) for a, b, c in zip(left, prev, upleft)], np.int64)
            line = cur - pred
        else:
            raise ValueError(f"write_png: filter {f}")
        out.append(f)
        out += (line & 0xFF).astype(np.uint8).tobytes()
        prev = cur
    png = _SIG + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 0, 0, 0, 0))
    png += _chunk(b"IDAT", zlib.compress(bytes(out), 6)) + _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)


def read_bin(path, shape=None):
    """f32 little-endian row-major BIN file; square shape inferred if None."""
    a = np.fromfile(path, dtype="<f4")
    if shape is None:
        s = int(round(np.sqrt(a.size)))
        if s * s != a.size:
            raise ValueError(f"{path}: {a.size} values is not a square; give the shape")
        shape = (s, s)
    return a.reshape(shape).astype(np.float32)


def write_bin(path, arr):
    np.ascontiguousarray(arr, dtype="<f4").tofile(path)


def png_to_bin(png_path, bin_path, normalize_to_01=True):
    """PNG mask -> f32 BIN: channel 0 of colour images, / 255 when normalising
    (tool/png_bin_converter.py L164-183).  Returns the array written."""
    px = read_png(png_path)
    if px.ndim == 3:
        px = px[..., 0]
    arr = px.astype(np.float32)
    if normalize_to_01:
        arr = arr / np.float32(255.0)
    write_bin(bin_path, arr)
    return arr


def bin_to_png(bin_path, png_path, shape=None, normalize=True):
    """f32 BIN -> 8-bit grey PNG, min-max normalised to 0..255 (truncating, as
    tool/png_bin_converter.py L104-140) or clipped when not normalising."""
    arr = read_bin(bin_path, shape)
    if normalize:
        lo, hi = float(arr.min()), float(arr.max())
        u8 = ((arr - lo) / (hi - lo) * 255).astype(np.uint8) if hi != lo else np.zeros_like(
            arr, np.uint8)
    else:
        u8 = np.clip(arr, 0, 255).astype(np.uint8)
    write_png(png_path, u8)
    return u8


def load_mask(path):
    """A mask from .png (/255, channel 0) or .bin (f32) as float32 (H, W)."""
    if os.path.splitext(path)[1].lower() == ".png":
        px = read_png(path)
        if px.ndim == 3:
            px = px[..., 0]
        return px.astype(np.float32) / np.float32(255.0)
    return read_bin(path)


def mask_levels(mask):
    """float mask in [0, 1] -> uint8 levels (round(255 m)); exact for masks
    that came from 8-bit PNGs or are binary."""
    m = np.asarray(mask, np.float32)
    if m.min() < 0 or m.max() > 1:
        raise ValueError("mask_levels: values outside [0, 1]")
    return np.rint(m * np.float32(255.0)).astype(np.uint8)
