"""Multilayer OpenEXR writer and reader.

Pinned by tests/test_openexr.cpp (a 2 x 2 Combined + Normal pair written as
`ViewLayer.<pass>.<component>` float channels with the lin_rec709_scene
colour-space attributes, read back exactly) and by two EXR files from the
reference's validation set: the reference's own output (one part) and
Blender's (multi part) decode to the same image, and the reference's light
passes compose to its Combined pass, as Cycles defines them:
Combined = (DiffDir + DiffInd) DiffCol + (GlossDir + GlossInd) GlossCol +
(TransDir + TransInd) TransCol + Emit + Env + VolumeDir + VolumeInd.
"""
from pathlib import Path

import numpy as np
import pytest

from psycles_amd import abi, exr
from psycles_amd.render import PassTile

GOLDEN = Path(__file__).parent / "golden"


def test_reference_pair_round_trips(tmp_path):
    """tests/test_openexr.cpp:27-151."""
    combined = np.array([[[1, 2, 3, 0.25], [4, 5, 6, 0.5]], [[7, 8, 9, 0.75], [10, 11, 12, 1.0]]], np.float32)
    normal = np.array([[[-1, 0, 1], [-0.5, 0.5, 1]], [[0.25, 0.5, 0.75], [1, 0, -1]]], np.float32)
    path = tmp_path / "pair.exr"
    exr.write_multilayer_exr([exr.PassImage(abi.PASS_COMBINED, "Combined", combined),
                              exr.PassImage(abi.PASS_NORMAL, "Normal", normal)], path, "ViewLayer")
    ch, attrs = exr.read_exr(path)
    assert sorted(ch) == ["ViewLayer.Combined.A", "ViewLayer.Combined.B", "ViewLayer.Combined.G",
                          "ViewLayer.Combined.R", "ViewLayer.Normal.X", "ViewLayer.Normal.Y", "ViewLayer.Normal.Z"]
    assert attrs["oiio:ColorSpace"][1] == b"lin_rec709_scene"
    assert attrs["colorInteropID"][1] == b"lin_rec709_scene"
    assert attrs["compression"][1] == bytes([3])
    got = np.stack([ch[f"ViewLayer.Combined.{c}"] for c in "RGBA"] + [ch[f"ViewLayer.Normal.{c}"] for c in "XYZ"], -1)
    assert np.array_equal(got, np.concatenate([combined, normal], -1))


def test_large_multipass_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(3)
    imgs = [exr.PassImage(abi.PASS_COMBINED, "Combined", rng.normal(size=(37, 53, 4)).astype(np.float32)),
            exr.PassImage(abi.PASS_DIFFUSE_DIRECT, "DiffDir", rng.uniform(size=(37, 53, 3)).astype(np.float32)),
            exr.PassImage(abi.PASS_DEPTH, "Depth", rng.uniform(size=(37, 53, 1)).astype(np.float32)),
            exr.PassImage(abi.PASS_SAMPLE_COUNT, "Debug Sample Count", np.full((37, 53, 1), 7.0, np.float32))]
    path = tmp_path / "multi.exr"
    exr.write_multilayer_exr(imgs, path, "A Long View Layer Name")
    ch, _ = exr.read_exr(path)
    assert np.array_equal(ch["A Long View Layer Name.Depth.Z"], imgs[2].pixels[..., 0])
    assert np.array_equal(ch["A Long View Layer Name.Debug Sample Count.X"], imgs[3].pixels[..., 0])
    for k, c in enumerate("RGBA"):
        assert np.array_equal(ch[f"A Long View Layer Name.Combined.{c}"], imgs[0].pixels[..., k])


def test_writer_rejects_what_the_reference_rejects(tmp_path):
    px = np.zeros((2, 2, 3), np.float32)
    with pytest.raises(ValueError):
        exr.write_multilayer_exr([exr.PassImage(abi.PASS_COMBINED, "Combined", px)], tmp_path / "a.exr", "")
    with pytest.raises(ValueError):
        exr.write_multilayer_exr([exr.PassImage(abi.PASS_COMBINED, "", px)], tmp_path / "b.exr")
    with pytest.raises(ValueError):
        exr.write_multilayer_exr([exr.PassImage(abi.PASS_COMBINED, "X", px),
                                  exr.PassImage(abi.PASS_ALBEDO, "X", px)], tmp_path / "c.exr")


def rgb(ch, name):
    return np.stack([ch[f"ViewLayer.{name}.{c}"] for c in "RGB"], -1).astype(np.float64)


def test_reference_files_decode_and_passes_compose():
    ours, attrs = exr.read_exr(GOLDEN / "glass_transport_psycles.exr")
    assert len(ours) == 46 and attrs["compression"][1] == bytes([3])
    combined = rgb(ours, "Combined")
    composed = ((rgb(ours, "DiffDir") + rgb(ours, "DiffInd")) * rgb(ours, "DiffCol") +
                (rgb(ours, "GlossDir") + rgb(ours, "GlossInd")) * rgb(ours, "GlossCol") +
                (rgb(ours, "TransDir") + rgb(ours, "TransInd")) * rgb(ours, "TransCol") +
                rgb(ours, "Emit") + rgb(ours, "Env") + rgb(ours, "VolumeDir") + rgb(ours, "VolumeInd"))
    assert np.abs(composed - combined).max() < 1e-5
    assert np.all(ours["ViewLayer.Combined.A"] == 1.0)
    cycles, _ = exr.read_exr(GOLDEN / "glass_transport_cycles.exr")
    assert "ViewLayer.Combined.R" in cycles and "ViewLayer.Depth.Z" in cycles
    assert abs(rgb(cycles, "Combined").mean() / combined.mean() - 1.0) < 1e-3


def test_sink_writes_the_render_passes(tmp_path):
    sink = exr.ExrOutputSink(tmp_path / "sink.exr")
    sink.begin(None)
    px = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)
    sink.write(PassTile(abi.PASS_COMBINED, 4, (0, 0, 3, 2), (3, 2), px))
    sink.write(PassTile(abi.PASS_EMISSION, 3, (0, 0, 3, 2), (3, 2), px[..., :3]))
    sink.end(False)
    ch, _ = exr.read_exr(tmp_path / "sink.exr")
    assert np.array_equal(ch["ViewLayer.Emit.G"], px[..., 1])
    assert np.array_equal(ch["ViewLayer.Combined.A"], px[..., 3])


def test_albedo_and_denoising_albedo_are_distinct_channels(tmp_path):
    """The albedo pass is DiffCol (examples/render_blender_scene.cpp:430-433),
    so a render requesting both albedo passes writes both."""
    sink = exr.ExrOutputSink(tmp_path / "both.exr")
    px = np.full((2, 2, 3), 0.5, np.float32)
    sink.begin(None)
    sink.write(PassTile(abi.PASS_ALBEDO, 3, (0, 0, 2, 2), (2, 2), px))
    sink.write(PassTile(abi.PASS_DENOISING_ALBEDO, 3, (0, 0, 2, 2), (2, 2), px * 2))
    sink.end(False)
    ch, _ = exr.read_exr(tmp_path / "both.exr")
    assert np.all(ch["ViewLayer.DiffCol.R"] == 0.5)
    assert np.all(ch["ViewLayer.Denoising Albedo.R"] == 1.0)


def test_sink_exceptions_reach_the_caller():
    """An exception raised inside a sink callback is held and re-raised after
    the C call, not printed and dropped by ctypes."""
    import ctypes as C

    from psycles_amd.render import OutputSink, make_sink

    class Failing(OutputSink):
        def begin(self, settings):
            pass

        def write(self, tile):
            raise ValueError("disk full")

        def end(self, cancelled):
            raise AssertionError("end must not run after a failed write")

    sink, callbacks = make_sink(Failing(), None)
    pixels = (C.c_float * 4)(1, 2, 3, 4)
    sink.begin(None, None)
    sink.write(None, abi.PASS_COMBINED, 4, 0, 0, 1, 1, 1, 1, pixels)
    sink.end(None, 0)
    with pytest.raises(ValueError, match="disk full"):
        callbacks.reraise()
    callbacks.reraise()  # raised once
