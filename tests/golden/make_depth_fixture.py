"""Writes tests/golden/move_dragon_depth.npz from the reference's own sample data.

The reference ships a recorded RGB-D sequence (data/move_dragon/cam-00/
frame-NNNNNN.depth.png: 640 x 480, 16-bit depth in mm) with its camera in
data/move_dragon/context.json ("intrinsic" fx fy cx cy, clip_near, clip_far,
downsample_scale) and system.json (surfel_radius_scale).  This script decodes
two frames (the configured start frame and the one after it) with a small
PNG reader of its own (16-bit greyscale, non-interlaced: zlib + the five
scanline filters) and stores them with the camera, so the GPU tests can run
the measurement maps on real depth without the reference present.

    python tests/golden/make_depth_fixture.py [/root/reference]
"""
import json
import struct
import sys
import zlib
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def read_png_u16(path: Path) -> np.ndarray:
    data = path.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n", "not a PNG"
    pos, idat, hdr = 8, [], None
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat.append(body)
        pos += 12 + n
    w, h, depth, color, _, _, interlace = hdr
    assert depth == 16 and color == 0 and interlace == 0, "expects 16-bit greyscale, non-interlaced"
    raw = zlib.decompress(b"".join(idat))
    bpp, stride = 2, 2 * w
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    p = 0
    for y in range(h):
        f = raw[p]
        line = np.frombuffer(raw, np.uint8, stride, p + 1).astype(np.int32)
        p += 1 + stride
        cur = np.zeros(stride, np.int32)
        if f == 0:
            cur[:] = line
        elif f == 2:
            cur[:] = (line + prev) & 255
        else:
            for i in range(stride):
                a = cur[i - bpp] if i >= bpp else 0
                b = prev[i]
                c = prev[i - bpp] if i >= bpp else 0
                if f == 1:
                    pred = a
                elif f == 3:
                    pred = (a + b) >> 1
                else:  # Paeth
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    pred = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
                cur[i] = (line[i] + pred) & 255
        out[y] = cur
        prev = cur
    return out.reshape(h, w, 2).astype(np.uint16)[:, :, 0] * 256 + out.reshape(h, w, 2)[:, :, 1]


def main(ref_root: str = "/root/reference") -> None:
    root = Path(ref_root) / "data" / "move_dragon"
    ctx = json.loads((root / "context.json").read_text())["cam-00"]
    system = json.loads((root / "system.json").read_text())
    start = int(system["start_frame"])
    frames = [read_png_u16(root / "cam-00" / f"frame-{i:06d}.depth.png") for i in (start, start + 1)]
    fx, fy, cx, cy = ctx["intrinsic"]
    np.savez_compressed(
        HERE / "move_dragon_depth.npz", depth_mm=np.stack(frames), frame_index=np.array([start, start + 1]),
        intrinsic=np.array([fx, fy, cx, cy], np.float32),
        clip=np.array([ctx["clip_near"], ctx["clip_far"]], np.float32),
        downsample_scale=np.float32(ctx["downsample_scale"]),
        surfel_radius_scale=np.float32(system["surfel_radius_scale"]))
    print("wrote", HERE / "move_dragon_depth.npz", frames[0].shape, "nonzero", int((frames[0] > 0).sum()))


if __name__ == "__main__":
    main(*sys.argv[1:])
