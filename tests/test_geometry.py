"""Geometry input: generators, cut distances, .gmy v4 reading/writing and
the per-lattice projection in the reference's site order."""
import math
import os

import numpy as np
import pytest

from hemelb_amd import _native as N
from hemelb_amd.geometry import MOORE, Geometry
from tests import cases as K

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _fluid_set(coords):
    return {tuple(c) for c in coords}


def _check_link_consistency(g):
    """A link is cut exactly when its neighbour is not a fluid site, and cut
    distances lie in [0, 1] (fraction of the link vector)."""
    coords, lt, li, ld, nrm, hn = g.links()
    fluid = _fluid_set(coords)
    for i, x in enumerate(coords):
        for k, c in enumerate(MOORE):
            nb = tuple(x + c)
            if nb in fluid:
                assert lt[i, k] == N.HLBG_CUT_NONE, (x, c)
            else:
                assert lt[i, k] != N.HLBG_CUT_NONE, (x, c)
                assert 0.0 <= ld[i, k] <= 1.0
    # wall normals exactly where wall links exist (GeometryReader.cc:633-641)
    has_wall = (lt == N.HLBG_CUT_WALL).any(1)
    np.testing.assert_array_equal(hn.astype(bool), has_wall)
    return coords, lt, li, ld


def test_cylinder_sites_and_links():
    R, L = 6.0, 20
    g = Geometry.cylinder(R, L)
    coords, lt, li, ld = _check_link_consistency(g)
    assert abs(g.n - math.pi * R * R * L) / (math.pi * R * R * L) < 0.05
    inl, outl = g.iolets(False), g.iolets(True)
    assert len(inl) == 1 and len(outl) == 1
    zmin, zmax = coords[:, 2].min(), coords[:, 2].max()
    assert inl[0].position[2] == zmin - 0.5 and outl[0].position[2] == zmax + 0.5
    # exact ray-cylinder wall distances (CylinderGenerator.cpp:112-150),
    # recomputed independently in numpy
    cx, cy = inl[0].position[0], inl[0].position[1]
    wall = np.argwhere(lt == N.HLBG_CUT_WALL)
    for i, k in wall[::7]:
        p = coords[i].astype(float)
        c = MOORE[k].astype(float)
        ax, ay = p[0] - cx, p[1] - cy
        A, B, Cc = c[0] ** 2 + c[1] ** 2, 2 * (ax * c[0] + ay * c[1]), ax * ax + ay * ay - R * R
        t = (-B + math.sqrt(B * B - 4 * A * Cc)) / (2 * A)
        assert abs(ld[i, k] - np.float32(t)) <= 1e-6
    # iolet links cross the planes half a voxel away
    iol = np.argwhere(lt >= N.HLBG_CUT_INLET)
    for i, k in iol:
        assert abs(ld[i, k] - 0.5) < 1e-6
        assert lt[i, k] == (N.HLBG_CUT_INLET if coords[i, 2] == zmin else N.HLBG_CUT_OUTLET)


def test_stenosed_cylinder_area_reduction():
    R, L, sev = 10.0, 60, 0.5
    g = Geometry.stenosed_cylinder(R, L, sev, 30.0, 8.0)
    coords, *_ = _check_link_consistency(g)
    z = coords[:, 2]
    counts = np.bincount(z - z.min())
    assert abs(counts.min() / counts.max() - (1 - sev)) < 0.08


def test_vessel_tree_is_seeded_and_consistent():
    a = Geometry.vessel_tree(7, 3, 6.0)
    b = Geometry.vessel_tree(7, 3, 6.0)
    c = Geometry.vessel_tree(8, 3, 6.0)
    assert a.n == b.n and a.n != c.n
    assert len(a.iolets(False)) == 1 and len(a.iolets(True)) == 4
    _check_link_consistency(a)
    np.testing.assert_array_equal(a.links()[0], b.links()[0])


@pytest.mark.parametrize("Q", [15, 19])
def test_projection_follows_reference_site_order(Q):
    """Sites grouped by collision type 0..5 (Domain.cc:187-207), then blocks
    in Morton order (LookupTree.h:92-97), then z fastest in a block; masks
    and iolet ids follow SiteData (SiteDataBare.cc:23-75)."""
    g = Geometry.vessel_tree(3, 3, 5.0)
    d = g.project(Q)
    ct = d.collision_type()
    assert (np.diff(ct) >= 0).all()
    bs = 8

    def spread(v):
        out = 0
        for b in range(16):
            out |= ((v >> b) & 1) << (3 * b)
        return out
    key = [(int(t), (spread(int(x) // bs) << 2) | (spread(int(y) // bs) << 1) | spread(int(z) // bs),
            ((x % bs) * bs + (y % bs)) * bs + (z % bs)) for t, (x, y, z) in zip(ct, d.coords)]
    assert key == sorted(key)
    lat = K.L.LATTICES[Q]
    coords, lt, li, ld, *_ = g.links()
    index = {tuple(c): i for i, c in enumerate(coords)}
    moore = {tuple(c): k for k, c in enumerate(MOORE)}
    for s in range(0, d.n, 37):
        i = index[tuple(d.coords[s])]
        for dd in range(1, Q):
            k = moore[tuple(lat.VECTORS[dd])]
            assert bool((d.wall_mask[s] >> (dd - 1)) & 1) == (lt[i, k] == N.HLBG_CUT_WALL)
            assert bool((d.iolet_mask[s] >> (dd - 1)) & 1) == (lt[i, k] >= N.HLBG_CUT_INLET)
            assert d.cut[s, dd - 1] == ld[i, k]


def test_gmy_roundtrip(tmp_path):
    g = Geometry.vessel_tree(5, 3, 5.0)
    p = str(tmp_path / "tree.gmy")
    g.write_gmy(p)
    h = Geometry.read_gmy(p)
    assert h.n == g.n
    assert h.block_info()[0].tolist() == g.block_info()[0].tolist()
    for a, b in zip(g.links(), h.links()):
        np.testing.assert_array_equal(a, b)


def test_reads_the_reference_four_cube_gmy():
    """Code/tests/geometry/GeometryReaderTests.cc:41-76 (TestSameAsFourCube):
    the reference's four_cube.gmy holds the programmatic FourCube: the same
    wall/iolet links (D3Q15) and wall normals exactly at the wall sites."""
    g = Geometry.read_gmy(os.path.join(GOLDEN, "four_cube.gmy"))
    assert g.n == 64
    blocks, bs = g.block_info()
    assert blocks.tolist() == [1, 1, 1] and bs == 6
    d = g.project(15)
    fc = K.four_cube(15)
    got = {tuple(c): (int(st), int(wm), int(im)) for c, st, wm, im in
           zip(d.coords, d.site_type, d.wall_mask, d.iolet_mask)}
    want = {tuple(c): (int(st), int(wm), int(im)) for c, st, wm, im in
            zip(fc.coords, fc.site_type, fc.wall_mask, fc.iolet_mask)}
    assert got == want
    coords, lt, li, ld, nrm, hn = g.links()
    for c, h in zip(coords, hn):
        is_wall = c[0] in (1, 4) or c[1] in (1, 4)
        assert bool(h) == is_wall


def test_read_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.gmy"
    p.write_bytes(b"\x00" * 64)
    with pytest.raises(N.HemeLBError, match="magic"):
        Geometry.read_gmy(str(p))
    with pytest.raises(N.HemeLBError):
        Geometry.read_gmy(str(tmp_path / "missing.gmy"))


def test_read_rejects_block_count_mismatch(tmp_path):
    """A block whose fluid sites differ from the count its header entry
    declares is malformed (GeometryReader.cc:248-281 reads the counts)."""
    raw = bytearray(open(os.path.join(GOLDEN, "four_cube.gmy"), "rb").read())
    # preamble: 3 magic/version words, 3 block counts, block size, padding
    nblocks = int(np.prod(np.frombuffer(bytes(raw[12:24]), ">u4")))
    hdr = np.frombuffer(bytes(raw[32:32 + 12 * nblocks]), ">u4").reshape(nblocks, 3)
    b = int(np.flatnonzero(hdr[:, 0])[0])
    bad = bytearray(raw)
    bad[32 + 12 * b:32 + 12 * b + 4] = int(hdr[b, 0] + 1).to_bytes(4, "big")
    p = tmp_path / "bad_count.gmy"
    p.write_bytes(bytes(bad))
    with pytest.raises(N.HemeLBError, match="fluid sites"):
        Geometry.read_gmy(str(p))


def test_split_domain_ghosts_are_the_pull_sources_across_the_cut():
    """Every site a subdomain pulls from along a fluid link is either its
    own or one of its ghosts, and every ghost is such a source (checked by
    brute force over all sites and directions)."""
    from hemelb_amd.geometry import split_domain
    from hemelb_amd.lb import LATTICES
    d = Geometry.vessel_tree(1, 3, 6.0).project(19)
    c = LATTICES[19].VECTORS
    pieces = split_domain(d, 3)
    assert sum(p[0].n for p in pieces) == d.n
    assert np.array_equal(np.sort(np.concatenate([p[1] for p in pieces])), np.arange(d.n))
    where = {tuple(x): k for k, x in enumerate(d.coords)}
    for own, idx, ghosts in pieces:
        mine = set(map(tuple, own.coords))
        need = set()
        for x in own.coords:
            for e in range(1, 19):
                y = tuple(x - c[e])
                j = where.get(y)
                if j is None or y in mine:
                    continue
                if not ((d.wall_mask[j] | d.iolet_mask[j]) >> (e - 1)) & 1:
                    need.add(y)
        assert need == set(map(tuple, ghosts.coords))
        assert need and not (need & mine)
