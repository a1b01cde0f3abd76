"""Multilayer OpenEXR output of render passes.

`write_multilayer_exr` mirrors psycles::io::write_multilayer_exr
(src/io/openexr.cpp:146-260): one single-part scanline file, 32-bit float
channels named `<view layer>.<pass name>.<component>` (R G B A; X Y Z for
Normal; Z for Depth; X Y for one- and two-channel passes), ZIP compression,
and the scene-linear Rec.709 colour-space attributes Cycles writes.  The file
layout follows the OpenEXR file-format specification: magic and version,
the attribute header, a table of chunk offsets, then chunks of 16 scanlines
whose channels are stored in name order, each chunk's bytes split into even
and odd halves, delta-encoded and deflated.

`read_exr` decodes scanline files, single- or multi-part, with NONE, ZIPS
or ZIP compression and HALF, FLOAT or UINT channels, which covers what this
writer produces, what the reference writes through OpenImageIO and the
multi-part files Blender writes for Cycles.  `ExrOutputSink` collects the
passes of a render and writes them when the render ends.
"""
from __future__ import annotations

import struct
import zlib
from dataclasses import dataclass

import numpy as np

from . import abi
from .render import OutputSink, PassTile

# Cycles / Blender pass names (the names in Cycles' multilayer EXR files); the
# albedo pass is "DiffCol" as the reference's Blender example names it
# (examples/render_blender_scene.cpp:430-433), distinct from Denoising Albedo
PASS_NAMES = {
    abi.PASS_COMBINED: "Combined", abi.PASS_DEPTH: "Depth", abi.PASS_NORMAL: "Normal",
    abi.PASS_POSITION: "Position", abi.PASS_ALBEDO: "DiffCol", abi.PASS_EMISSION: "Emit",
    abi.PASS_DIFFUSE_DIRECT: "DiffDir", abi.PASS_DIFFUSE_INDIRECT: "DiffInd",
    abi.PASS_GLOSSY_DIRECT: "GlossDir", abi.PASS_GLOSSY_INDIRECT: "GlossInd",
    abi.PASS_TRANSMISSION_DIRECT: "TransDir", abi.PASS_TRANSMISSION_INDIRECT: "TransInd",
    abi.PASS_VOLUME_DIRECT: "VolumeDir", abi.PASS_VOLUME_INDIRECT: "VolumeInd",
    abi.PASS_DENOISING_NORMAL: "Denoising Normal", abi.PASS_DENOISING_ALBEDO: "Denoising Albedo",
    abi.PASS_SAMPLE_COUNT: "Debug Sample Count", abi.PASS_GLOSSY_COLOR: "GlossCol",
    abi.PASS_TRANSMISSION_COLOR: "TransCol", abi.PASS_ENVIRONMENT: "Env",
}

_MAGIC = 20000630
_HALF, _FLOAT, _UINT = 1, 2, 0
_NONE, _RLE, _ZIPS, _ZIP = 0, 1, 2, 3
_LINES = {_NONE: 1, _RLE: 1, _ZIPS: 1, _ZIP: 16}


@dataclass
class PassImage:
    """One pass: kind, name and an (H, W, C) float32 image
    (include/psycles/io/image.h PassImage)."""

    kind: int
    name: str
    pixels: np.ndarray


def component_name(kind: int, channels: int, component: int) -> str:
    """src/io/openexr.cpp:33-53."""
    if kind == abi.PASS_NORMAL and component < 3:
        return "XYZ"[component]
    if kind == abi.PASS_DEPTH and component == 0:
        return "Z"
    if channels <= 4:
        return "XYZ"[component] if channels <= 2 else "RGBA"[component]
    return f"C{component}"


def _attribute(name: str, kind: str, value: bytes) -> bytes:
    return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(value)) + value


def _string(value: str) -> bytes:
    return value.encode()


def _predict_and_interleave(raw: bytes) -> bytes:
    """The ZIP preprocessing: bytes split into even / odd halves, then each
    byte replaced by its difference from the previous one, offset by 128."""
    b = np.frombuffer(raw, dtype=np.uint8)
    t = np.concatenate([b[0::2], b[1::2]]).astype(np.int16)
    d = t.copy()
    d[1:] = (t[1:] - t[:-1] + 128) & 0xFF
    return d.astype(np.uint8).tobytes()


def _undo_predict_and_interleave(data: bytes) -> bytes:
    d = np.frombuffer(data, dtype=np.uint8).astype(np.int64)
    t = np.empty_like(d)
    if d.size:
        t[0] = d[0]
        t[1:] = d[1:] - 128
        t = np.cumsum(t) & 0xFF
    t = t.astype(np.uint8)
    half = (t.size + 1) // 2
    out = np.empty_like(t)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def write_multilayer_exr(images, path, view_layer: str = "ViewLayer") -> None:
    """Write `images` (PassImage, all the same width and height) as one
    multilayer EXR; raises ValueError for the cases the reference reports
    as errors (src/io/openexr.cpp:146-240)."""
    if not view_layer:
        raise ValueError("the OpenEXR view-layer name is empty")
    if not images:
        raise ValueError("no passes to write")
    h, w = images[0].pixels.shape[:2]
    columns = []   # (channel name, (H, W) float32)
    seen = set()
    for img in images:
        px = np.asarray(img.pixels, dtype=np.float32)
        if px.ndim == 2:
            px = px[..., None]
        if px.shape[:2] != (h, w):
            raise ValueError("OpenEXR passes differ in size")
        c = px.shape[2]
        if c == 0:
            raise ValueError("an OpenEXR pass has no channels")
        if not img.name:
            raise ValueError("an OpenEXR pass has an empty name")
        for k in range(c):
            name = f"{view_layer}.{img.name}.{component_name(img.kind, c, k)}"
            if name in seen:
                raise ValueError("duplicate OpenEXR channel name: " + name)
            seen.add(name)
            columns.append((name, px[..., k]))
    columns.sort(key=lambda x: x[0].encode())   # channels are stored in name order
    chlist = b"".join(n.encode() + b"\0" + struct.pack("<iBBBBii", _FLOAT, 0, 0, 0, 0, 1, 1) for n, _ in columns)
    chlist += b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    long_names = any(len(n) > 31 for n, _ in columns)
    header = struct.pack("<ii", _MAGIC, 2 | (0x400 if long_names else 0))
    header += _attribute("channels", "chlist", chlist)
    header += _attribute("colorInteropID", "string", _string("lin_rec709_scene"))
    header += _attribute("compression", "compression", bytes([_ZIP]))
    header += _attribute("dataWindow", "box2i", box)
    header += _attribute("displayWindow", "box2i", box)
    header += _attribute("lineOrder", "lineOrder", bytes([0]))
    header += _attribute("oiio:ColorSpace", "string", _string("lin_rec709_scene"))
    header += _attribute("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    header += _attribute("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
    header += _attribute("screenWindowWidth", "float", struct.pack("<f", 1.0))
    header += b"\0"
    planes = np.stack([col for _, col in columns], axis=1)   # (H, C, W)
    chunks = []
    for y0 in range(0, h, _LINES[_ZIP]):
        raw = np.ascontiguousarray(planes[y0:y0 + _LINES[_ZIP]]).astype("<f4").tobytes()
        packed = zlib.compress(_predict_and_interleave(raw))
        data = packed if len(packed) < len(raw) else raw
        chunks.append(struct.pack("<ii", y0, len(data)) + data)
    offsets = []
    pos = len(header) + 8 * len(chunks)
    for c in chunks:
        offsets.append(pos)
        pos += len(c)
    with open(path, "wb") as f:
        f.write(header)
        f.write(struct.pack(f"<{len(offsets)}Q", *offsets))
        for c in chunks:
            f.write(c)


def _parse_attributes(data: bytes, pos: int):
    attrs = {}
    while data[pos] != 0:
        e = data.index(b"\0", pos)
        name = data[pos:e].decode()
        pos = e + 1
        e = data.index(b"\0", pos)
        kind = data[pos:e].decode()
        pos = e + 1
        size = struct.unpack_from("<i", data, pos)[0]
        pos += 4
        attrs[name] = (kind, data[pos:pos + size])
        pos += size
    return attrs, pos + 1


def _channels(attrs):
    chl = attrs["channels"][1]
    channels, p = [], 0
    while chl[p] != 0:
        e = chl.index(b"\0", p)
        name = chl[p:e].decode()
        ptype, _, _, _, _, xs, ys = struct.unpack_from("<iBBBBii", chl, e + 1)
        if xs != 1 or ys != 1:
            raise ValueError("subsampled channels are not read")
        channels.append((name, ptype))
        p = e + 1 + 16
    return channels


def read_exr(path):
    """Returns (channels, attributes): channels maps each channel name to an
    (H, W) float32 image; attributes maps names to (type, raw bytes) of the
    (first) part.  Multi-part files (as Blender writes them) are read part
    by part; a channel of a named part is returned as `<part>.<channel>`
    unless its name already starts with the part name."""
    data = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", data, 0)
    if magic != _MAGIC:
        raise ValueError("not an OpenEXR file")
    if version & 0x200 or version & 0x800:
        raise ValueError("tiled and deep OpenEXR files are not read")
    multipart = (version & 0x1000) != 0
    pos, parts = 8, []
    while True:
        attrs, pos = _parse_attributes(data, pos)
        parts.append(attrs)
        if not multipart or data[pos] == 0:
            break
    if multipart:
        pos += 1
    tables = []
    for attrs in parts:
        x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
        compression = attrs["compression"][1][0]
        if compression not in (_NONE, _ZIPS, _ZIP):
            raise ValueError(f"unsupported OpenEXR compression {compression}")
        lines = _LINES[compression]
        n_chunks = (y1 - y0 + lines) // lines
        if "chunkCount" in attrs:
            n_chunks = struct.unpack("<i", attrs["chunkCount"][1])[0]
        tables.append(struct.unpack_from(f"<{n_chunks}Q", data, pos))
        pos += 8 * n_chunks
    sizes = {_HALF: 2, _FLOAT: 4, _UINT: 4}
    dtypes = {_HALF: "<f2", _FLOAT: "<f4", _UINT: "<u4"}
    out = {}
    for index, (attrs, offsets) in enumerate(zip(parts, tables)):
        channels = _channels(attrs)
        compression = attrs["compression"][1][0]
        lines = _LINES[compression]
        x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
        w, h = x1 - x0 + 1, y1 - y0 + 1
        part = attrs["name"][1].decode() if multipart and "name" in attrs else ""
        full = [n if not part or n.startswith(part + ".") else f"{part}.{n}" for n, _ in channels]
        planes = {n: np.zeros((h, w), dtype=np.float32) for n in full}
        for off in offsets:
            q0 = off + (4 if multipart else 0)
            y, size = struct.unpack_from("<ii", data, q0)
            payload = data[q0 + 8:q0 + 8 + size]
            rows = min(lines, y1 - y + 1)
            expect = rows * w * sum(sizes[t] for _, t in channels)
            if compression != _NONE and size < expect:
                payload = _undo_predict_and_interleave(zlib.decompress(payload))
            q = 0
            for r in range(rows):
                for n, (_, t) in zip(full, channels):
                    k = w * sizes[t]
                    planes[n][y - y0 + r] = np.frombuffer(payload[q:q + k], dtype=dtypes[t]).astype(np.float32)
                    q += k
        out.update(planes)
    return out, parts[0]


class ExrOutputSink(OutputSink):
    """Collects the passes a render delivers and writes them to `path` as a
    multilayer EXR when the render ends (not when it is cancelled)."""

    def __init__(self, path, view_layer: str = "ViewLayer", names: dict | None = None):
        self.path = path
        self.view_layer = view_layer
        self.names = dict(PASS_NAMES, **(names or {}))
        self.passes: dict[int, np.ndarray] = {}

    def begin(self, settings) -> None:
        self.passes = {}

    def write(self, tile: PassTile) -> None:
        self.passes[tile.pass_kind] = tile.pixels

    def end(self, cancelled: bool) -> None:
        if cancelled or not self.passes:
            return
        images = [PassImage(k, self.names.get(k, f"Pass{k}"), px) for k, px in self.passes.items()]
        write_multilayer_exr(images, self.path, self.view_layer)
