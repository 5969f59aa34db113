"""Extraction (.xtr v5) and offset (.off) files: the header bytes the
reference's LocalPropertyOutputTests expect, records decoded independently
with numpy, and the distribution checkpoint round trip."""
import json
import os

import numpy as np
import pytest

from hemelb_amd import _native as N
from hemelb_amd import extraction as X

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _dummy_grid():
    # DummyDataSource.h:22-40: 4x4x4 sites, i slowest
    return np.array([(i, j, k) for i in range(4) for j in range(4) for k in range(4)], np.int32)


def test_headers_match_reference_test_bytes(tmp_path):
    g = json.load(open(os.path.join(GOLDEN, "xtr_headers.json")))
    path = str(tmp_path / "simple.xtr")
    w = X.ExtractionWriter(path, g["voxel_size"], g["origin"], _dummy_grid(), [
        X.OutputField("Pressure", "float", 1, (g["reference_pressure_mmHg"],)),
        X.OutputField("Velocity", "float", 3)])
    w.close()
    raw = open(path, "rb").read()
    assert raw[:60].hex() == g["main_header"]
    assert raw[60:112].hex() == g["field_header"]
    assert len(raw) == 112


def _decode_records(raw, header_len, n, site_dtype):
    rec = np.dtype([("t", ">u8"), ("sites", site_dtype, (n,))])
    body = np.frombuffer(raw[header_len:], dtype=np.uint8)
    assert body.size % rec.itemsize == 0
    return np.frombuffer(body.tobytes(), dtype=rec)


def test_records_and_offset_file(tmp_path):
    """LocalPropertyOutputTests.cc CheckDataWriting: 8 + 28 bytes per site,
    coordinates then pressure - reference pressure then velocity, as XDR."""
    grid = _dummy_grid()
    rng = np.random.default_rng(1358)
    path = str(tmp_path / "simple.xtr")
    w = X.ExtractionWriter(path, 0.3e-3, (0.034, 0.001, 0.074), grid, [
        X.OutputField("Pressure", "float", 1, (8.0,)), X.OutputField("Velocity", "float", 3)])
    data = []
    for t in (0, 100):
        p = 80 + 2 * rng.random(64)
        u = 0.01 * rng.random((64, 3))
        w.append(t, [p, u])
        data.append((t, p, u))
    w.close()
    raw = open(path, "rb").read()
    site = np.dtype([("x", ">u4", (3,)), ("p", ">f4"), ("u", ">f4", (3,))])
    assert site.itemsize == 28
    recs = _decode_records(raw, 112, 64, site)
    assert len(recs) == 2
    for r, (t, p, u) in zip(recs, data):
        assert r["t"] == t
        np.testing.assert_array_equal(r["sites"]["x"], grid)
        np.testing.assert_array_equal(r["sites"]["p"], np.float32(p - 8.0))
        np.testing.assert_array_equal(r["sites"]["u"], np.float32(u))
    off = open(str(tmp_path / "simple.off"), "rb").read()
    hdr = np.frombuffer(off[:16], ">u4")
    assert hdr.tolist() == [0x686c6221, 0x6f666604, 1, 1]
    assert np.frombuffer(off[16:], ">u8").tolist() == [112, 112 + 8 + 28 * 64]
    # the reference never overwrites a file (MPI_MODE_EXCL)
    with pytest.raises(N.HemeLBError, match="exists"):
        X.ExtractionWriter(path, 1.0, (0, 0, 0), grid, [X.OutputField("a")])


class _FakeLBM:
    """The parts of LBM the checkpoint helpers use."""

    def __init__(self, coords, Q, f, t):
        from types import SimpleNamespace
        self.domain = SimpleNamespace(coords=coords, n=coords.shape[0])
        self.Q, self.f, self.t = Q, f, t

    def get_distributions(self):
        return self.f.copy()

    def set_distributions(self, f):
        self.f = f.copy()

    def GetTimeStep(self):
        return self.t

    def SetTimeStep(self, t):
        self.t = t


def test_checkpoint_round_trip_and_lookup(tmp_path):
    rng = np.random.default_rng(3)
    coords = rng.permutation(np.array([(x, y, z) for x in range(5) for y in range(4)
                                       for z in range(3)], np.int32))
    Q = 19
    a = _FakeLBM(coords, Q, rng.random((len(coords), Q)), 11)
    path = str(tmp_path / "cp.xtr")
    w = X.checkpoint_writer(a, path, 5e-5, (0.0, 0.0, 0.0))
    X.write_checkpoint(w, a)
    first = a.f.copy()
    a.f, a.t = rng.random((len(coords), Q)), 42
    X.write_checkpoint(w, a)
    w.close()
    # binary64 values and coordinates decoded independently
    site = np.dtype([("x", ">u4", (3,)), ("f", ">f8", (Q,))])
    raw = open(path, "rb").read()
    hl = 60 + 4 + 16 + 12            # "distributions": 13 chars padded to 16
    recs = _decode_records(raw, hl, len(coords), site)
    assert recs["t"].tolist() == [11, 42]
    np.testing.assert_array_equal(recs[0]["sites"]["f"], first)
    # last record by default; sites matched by coordinates into another order
    perm = rng.permutation(len(coords))
    b = _FakeLBM(coords[perm], Q, np.zeros((len(coords), Q)), 0)
    assert X.load_checkpoint(b, path) == 42 and b.t == 42
    np.testing.assert_array_equal(b.f, a.f[perm])
    assert X.load_checkpoint(b, path, 11) == 11
    np.testing.assert_array_equal(b.f, first[perm])
    with pytest.raises(N.HemeLBError, match="not found"):
        X.load_checkpoint(b, path, 20)
    c = _FakeLBM(coords[:-1], Q, np.zeros((len(coords) - 1, Q)), 0)
    with pytest.raises(N.HemeLBError, match="sites"):
        X.load_checkpoint(c, path)
    moved = coords.copy()
    moved[0] += 100
    d = _FakeLBM(moved, Q, np.zeros((len(coords), Q)), 0)
    with pytest.raises(N.HemeLBError, match="not a site"):
        X.load_checkpoint(d, path)
    e = _FakeLBM(coords, 15, np.zeros((len(coords), 15)), 0)
    with pytest.raises(N.HemeLBError, match="binary64"):
        X.load_checkpoint(e, path)


class _FakeFlow:
    """The parts of LBM the property writer uses: fOld's density, velocity
    and wall shear stress, the domain and the time step."""

    def __init__(self, coords, rho, u, wss, wall):
        from types import SimpleNamespace
        self.domain = SimpleNamespace(coords=coords, n=coords.shape[0],
                                      wall_mask=wall.astype(np.uint32))
        self.rho, self.u, self.wss = rho, u, wss

    def macroscopic(self):
        return self.rho.copy(), self.u.copy()

    def wall_shear_stress(self):
        return self.wss.copy()

    def GetTimeStep(self):
        return 7


def test_property_file_pressure_offset_is_the_reference_pressure(tmp_path):
    """The header's pressure offset is the UnitConverter's reference
    pressure, so stored value + offset gives ConvertPressureToPhysicalUnits
    of the density (LocalPropertyOutput.cc, LbDataSourceIterator.cc:34-38)."""
    from hemelb_amd import lb as L
    grid = _dummy_grid()
    rng = np.random.default_rng(5)
    n = len(grid)
    wall = rng.random(n) < 0.3
    flow = _FakeFlow(grid, 1.0 + 1e-3 * rng.standard_normal(n), 1e-3 * rng.random((n, 3)),
                     np.where(wall, 1e-4 * rng.random(n), X.NO_VALUE), wall)
    units = L.UnitConverter(6.25e-5, 5e-5, 1000.0, 80.0)
    path = str(tmp_path / "props.xtr")
    w = X.property_writer(flow, path, units)
    assert X.write_properties(w, flow, units) == 7
    w.close()
    raw = open(path, "rb").read()
    # field header of "pressure": name (4 + 8 bytes), n, type, n_offsets,
    # then the offset as an XDR float
    fh = raw[60:]
    assert np.frombuffer(fh[:4], ">u4")[0] == 8 and fh[4:12] == b"pressure"
    assert np.frombuffer(fh[12:24], ">u4").tolist() == [1, 0, 1]
    offset = float(np.frombuffer(fh[24:28], ">f4")[0])
    assert offset == 80.0
    hlen = 60 + int(np.frombuffer(raw[56:60], ">u4")[0])
    site = np.dtype([("x", ">u4", (3,)), ("p", ">f4"), ("u", ">f4", (3,)), ("s", ">f4")])
    rec = _decode_records(raw, hlen, n, site)[0]
    want = units.ConvertPressureToPhysicalUnits(flow.rho * L.Cs2)
    got = rec["sites"]["p"].astype(np.float64) + offset
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-5 * np.abs(want).max())
