"""DEN and NRRD files (io/den, io/nrrd) through the C ABI; no device needed.

Pinned by the reference's own fixtures: testData/header.nrrd with the values its test expects
(nrrdfileiotest.cpp:21-60: sizes 256 256 199, raw, little, float, version 4, header offset 118)
and testData/tubeGeo.den / cArmGeo.den, read independently with numpy. The rest are round trips
and the header text the reference's writer produces (nrrdfileio.tpp:250-337).
"""
from pathlib import Path

import numpy as np
import pytest

from ctl_amd import ProjectionData, VoxelVolume, _lib
from ctl_amd import fileio as io
from ctl_amd._lib import CtlProjError

GOLDEN = Path(__file__).parent / "golden"


def test_reference_nrrd_header_fixture():
    info, kv = io.nrrd_info(GOLDEN / "header.nrrd")
    assert info.version == 4 and info.ndim == 3
    assert [info.sizes[k] for k in range(3)] == [256, 256, 199]
    assert info.elem_bytes == 4  # "type: float"
    assert info.header_bytes == 118  # nrrdfileiotest.cpp: "nrrd header offset" 118
    assert kv == {"myKey": "myValue"}  # comments "# asd1", "#asd2", "##asd3" skipped
    assert info.data_complete == 0  # the fixture holds a header only
    with pytest.raises(CtlProjError) as ei:
        io.load_nrrd(GOLDEN / "header.nrrd")
    assert ei.value.status == _lib.ERR_IO


@pytest.mark.parametrize("name,modules,views", [("tubeGeo.den", 40, 10), ("cArmGeo.den", 40, 30)])
def test_reference_den_geometry_fixtures(name, modules, views):
    rows, cols, count, dt = io.den_info(GOLDEN / name)
    assert (rows, cols, count, dt) == (3, 4, modules * views, np.float64)
    P = io.load_geometry_den(GOLDEN / name, modules)
    raw = (GOLDEN / name).read_bytes()
    ref = np.frombuffer(raw[6:], "<f8").reshape(views, modules, 3, 4)  # independent reader
    assert P.shape == (views, modules, 3, 4) and np.array_equal(P, ref)


def test_den_round_trips(tmp_path):
    rng = np.random.default_rng(5)
    vol = VoxelVolume(rng.random((7, 5, 3), dtype=np.float32), (0.5, 0.5, 1.0))
    io.save_volume_den(tmp_path / "v.den", vol)
    assert io.den_info(tmp_path / "v.den") == (5, 3, 7, np.float32)  # rows=Y cols=X count=Z
    assert (tmp_path / "v.den").stat().st_size == 6 + 7 * 5 * 3 * 4
    assert np.array_equal(io.load_volume_den(tmp_path / "v.den").data, vol.data)
    proj = ProjectionData(rng.random((3, 2, 4, 6), dtype=np.float32))
    io.save_projections_den(tmp_path / "p.den", proj)
    assert io.den_info(tmp_path / "p.den")[:3] == (4, 6, 6)  # count = modules x views
    assert np.array_equal(io.load_projections_den(tmp_path / "p.den", modules=2).data, proj.data)
    pm = rng.random((4, 3, 3, 4))
    io.save_geometry_den(tmp_path / "g.den", pm)
    assert np.array_equal(io.load_geometry_den(tmp_path / "g.den", 3), pm)
    for dt in (np.uint8, np.uint16):
        a = (rng.random((2, 3, 4)) * 200).astype(dt)
        io.save_den(tmp_path / "u.den", a)
        b = io.load_den(tmp_path / "u.den")
        assert b.dtype == dt and np.array_equal(a, b)


def test_den_errors(tmp_path):
    with pytest.raises(CtlProjError) as ei:  # den_header.h isOutOfBounds
        io.save_den(tmp_path / "big.den", np.zeros((1, 1, 70000), np.float32))
    assert ei.value.status == _lib.ERR_INVALID_ARGUMENT
    (tmp_path / "bad.den").write_bytes(np.array([2, 2, 1], "<u2").tobytes() + b"\0" * 12)  # 3 B/elem
    with pytest.raises(CtlProjError) as ei:
        io.den_info(tmp_path / "bad.den")
    assert ei.value.status == _lib.ERR_IO
    with pytest.raises(CtlProjError) as ei:
        io.den_info(tmp_path / "missing.den")
    assert ei.value.status == _lib.ERR_IO


def test_nrrd_volume_header_and_round_trip(tmp_path):
    rng = np.random.default_rng(6)
    vol = VoxelVolume(rng.random((4, 3, 5), dtype=np.float32), (0.25, 0.5, 1.0), (1.5, -2.0, 0.0))
    io.save_volume_nrrd(tmp_path / "v.nrrd", vol)
    text = (tmp_path / "v.nrrd").read_bytes().split(b"\n\n", 1)[0].decode().splitlines()
    assert text == ["NRRD0004", "type: float", "dimension: 3", "sizes: 5 3 4",
                    'labels: "num vox x" "num vox y" "num vox z"', "encoding: raw", "endian: little",
                    "spacings: 0.25 0.5 1", 'units: "mm" "mm" "mm"', "space: scanner-xyz",
                    "space origin: (1.5,-2,0)"]
    back = io.load_volume_nrrd(tmp_path / "v.nrrd")
    assert np.array_equal(back.data, vol.data)
    assert back.voxel_size == vol.voxel_size and back.offset == vol.offset


def test_nrrd_projections_and_keyvals(tmp_path):
    rng = np.random.default_rng(7)
    proj = ProjectionData(rng.random((2, 3, 4, 5), dtype=np.float32))
    io.save_projections_nrrd(tmp_path / "p.nrrd", proj)
    info, _ = io.nrrd_info(tmp_path / "p.nrrd")
    assert info.ndim == 4 and [info.sizes[k] for k in range(4)] == [5, 4, 3, 2]  # C R M V
    assert np.array_equal(io.load_projections_nrrd(tmp_path / "p.nrrd").data, proj.data)
    a = np.arange(6, dtype=np.float64).reshape(2, 3)
    io.save_nrrd(tmp_path / "k.nrrd", a, keyvals={"scanner": "c-arm", "mu unit": "1/mm"})
    info, kv = io.nrrd_info(tmp_path / "k.nrrd")
    assert kv == {"scanner": "c-arm", "mu unit": "1/mm"} and info.elem_bytes == 8
    assert np.array_equal(io.load_nrrd(tmp_path / "k.nrrd"), a)


@pytest.mark.parametrize("header,what", [
    (b"NRRD04\ntype: float\n\n", "magic"),
    (b"NRRD0004\ntype: float\ndimension: 1\nsizes: 2\nendian: big\n\n", "endian"),
    (b"NRRD0004\ntype: float\ndimension: 1\nsizes: 2\nencoding: gzip\n\n", "encoding"),
    (b"NRRD0004\ntype: float\ndimension: 2\nsizes: 2\n\n", "sizes"),
    (b"NRRD0004\ntype: float\ndimension: 1\nsizes: 2\n", "end"),
])
def test_nrrd_malformed(tmp_path, header, what):
    (tmp_path / "m.nrrd").write_bytes(header + b"\0" * 8)
    with pytest.raises(CtlProjError) as ei:
        io.nrrd_info(tmp_path / "m.nrrd")
    assert ei.value.status == _lib.ERR_IO, what
