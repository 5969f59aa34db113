"""Geometry input: synthetic vessels with exact cut distances, .gmy v4 files,
and the per-lattice ``Domain`` view (the reference's ``geometry::Domain``,
Code/geometry/Domain.cc:69-357) handed to the solver.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


@dataclass
class IoletPlane:
    position: np.ndarray  # lattice units
    normal: np.ndarray    # unit, pointing into the fluid
    radius: float


@dataclass
class Domain:
    """Site arrays of one lattice, in the reference's local site order."""
    Q: int
    coords: np.ndarray      # int32 [n,3]
    site_type: np.ndarray   # uint32 [n]
    wall_mask: np.ndarray   # uint32 [n]
    iolet_mask: np.ndarray  # uint32 [n]
    iolet_id: np.ndarray    # int32 [n]
    cut: np.ndarray         # float32 [n, Q-1]
    normal: np.ndarray      # float32 [n,3]
    inlets: list
    outlets: list

    @property
    def n(self) -> int:
        return int(self.coords.shape[0])

    def collision_type(self) -> np.ndarray:
        """0..5: FLUID, WALL, INLET, OUTLET, INLET|WALL, OUTLET|WALL."""
        w = self.wall_mask != 0
        t = self.site_type
        ct = np.where(t == N.HLB_SITE_FLUID, w.astype(np.int32),
                      np.where(t == N.HLB_SITE_INLET, 2 + 2 * w, 3 + 2 * w))
        return ct.astype(np.int32)

    def as_c(self):
        """hlb_domain referencing this object's arrays (keep self alive)."""
        d = N.hlb_domain()
        d.n_sites = self.n
        d.coords = _ptr(self.coords)
        d.site_type = _ptr(self.site_type)
        d.wall_mask = _ptr(self.wall_mask)
        d.iolet_mask = _ptr(self.iolet_mask)
        d.iolet_id = _ptr(self.iolet_id)
        d.cut_distance = _ptr(self.cut)
        d.wall_normal = _ptr(self.normal)
        return d


class Geometry:
    """Owner of an ``hlbg_geometry`` (fluid sites plus their 26 links)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            N.lib().hlbg_free(self._h)
            self._h = C.c_void_p()

    def signed_distance(self, origin, dims) -> np.ndarray:
        """Signed distance field of the crop [origin, origin + dims) in voxels:
        positive at fluid sites (distance to the nearest non-fluid voxel),
        negative elsewhere (distance to the nearest fluid site); [x][y][z]."""
        o = np.ascontiguousarray(origin, np.int32)
        d = np.ascontiguousarray(dims, np.int32)
        out = np.empty(tuple(int(v) for v in d), np.float32)
        N.check(N.lib().hlbg_signed_distance(self._h, _ptr(o), _ptr(d), _ptr(out)), True)
        return out

    # ---- constructors ----
    @classmethod
    def cylinder(cls, radius: float, length: int, block_size: int = 8) -> "Geometry":
        h = C.c_void_p()
        N.check(N.lib().hlbg_cylinder(radius, length, block_size, C.byref(h)), True)
        return cls(h.value)

    @classmethod
    def stenosed_cylinder(cls, radius, length, severity, z0, width,
                          block_size=8) -> "Geometry":
        h = C.c_void_p()
        N.check(N.lib().hlbg_stenosed_cylinder(radius, length, severity, z0, width,
                                               block_size, C.byref(h)), True)
        return cls(h.value)

    @classmethod
    def vessel_tree(cls, seed: int, generations: int, root_radius: float,
                    block_size: int = 8) -> "Geometry":
        h = C.c_void_p()
        N.check(N.lib().hlbg_vessel_tree(seed, generations, root_radius, block_size,
                                         C.byref(h)), True)
        return cls(h.value)

    @classmethod
    def read_gmy(cls, path: str) -> "Geometry":
        h = C.c_void_p()
        N.check(N.lib().hlbg_read_gmy(path.encode(), C.byref(h)), True)
        return cls(h.value)

    def write_gmy(self, path: str) -> None:
        N.check(N.lib().hlbg_write_gmy(self._h, path.encode()), True)

    # ---- queries ----
    @property
    def n(self) -> int:
        return int(N.lib().hlbg_fluid_site_count(self._h))

    def block_info(self):
        b = np.zeros(3, np.int32)
        bs = np.zeros(1, np.int32)
        N.check(N.lib().hlbg_block_info(self._h, _ptr(b), _ptr(bs)), True)
        return b, int(bs[0])

    def iolets(self, outlet: bool) -> list:
        out = []
        for i in range(N.lib().hlbg_iolet_count(self._h, int(outlet))):
            io = N.hlbg_iolet()
            N.check(N.lib().hlbg_get_iolet(self._h, int(outlet), i, C.byref(io)), True)
            out.append(IoletPlane(np.array(io.position[:]), np.array(io.normal[:]),
                                  float(io.radius)))
        return out

    def set_iolets(self, inlets, outlets) -> None:
        def arr(lst):
            a = (N.hlbg_iolet * max(1, len(lst)))()
            for i, p in enumerate(lst):
                a[i].position[:] = list(map(float, p.position))
                a[i].normal[:] = list(map(float, p.normal))
                a[i].radius = float(p.radius)
            return a
        N.check(N.lib().hlbg_set_iolets(self._h, arr(inlets), len(inlets),
                                        arr(outlets), len(outlets)), True)

    def links(self):
        n = self.n
        coords = np.zeros((n, 3), np.int32)
        lt = np.zeros((n, 26), np.uint8)
        li = np.zeros((n, 26), np.int32)
        ld = np.zeros((n, 26), np.float32)
        nrm = np.zeros((n, 3), np.float32)
        hn = np.zeros(n, np.uint8)
        N.check(N.lib().hlbg_get_links(self._h, _ptr(coords), _ptr(lt), _ptr(li),
                                       _ptr(ld), _ptr(nrm), _ptr(hn)), True)
        return coords, lt, li, ld, nrm, hn

    def project(self, Q: int) -> Domain:
        n = self.n
        coords = np.zeros((n, 3), np.int32)
        st = np.zeros(n, np.uint32)
        wm = np.zeros(n, np.uint32)
        im = np.zeros(n, np.uint32)
        iid = np.zeros(n, np.int32)
        cut = np.zeros((n, Q - 1), np.float32)
        nrm = np.zeros((n, 3), np.float32)
        N.check(N.lib().hlbg_project(self._h, Q, _ptr(coords), _ptr(st), _ptr(wm),
                                     _ptr(im), _ptr(iid), _ptr(cut), _ptr(nrm)), True)
        return Domain(Q, coords, st, wm, im, iid, cut, nrm,
                      self.iolets(False), self.iolets(True))


# Moore neighbourhood in .gmy order (Code/io/formats/geometry.h:120-156)
MOORE = np.array([(x, y, z) for x in (-1, 0, 1) for y in (-1, 0, 1) for z in (-1, 0, 1)
                  if (x, y, z) != (0, 0, 0)], np.int32)


# ---------------------------------------------------------------------------
# Decomposition into subdomains with ghost sites
# ---------------------------------------------------------------------------
def _morton_block_key(coords: np.ndarray) -> np.ndarray:
    """Morton key of each site's 8^3 block (x, y, z interleaved, x highest)."""
    b = (coords - coords.min(0)).astype(np.uint64) >> np.uint64(3)
    key = np.zeros(len(coords), np.uint64)
    for bit in range(21):
        for a in range(3):
            key |= ((b[:, a] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit + 2 - a)
    return key


def _subset(d: Domain, sel: np.ndarray) -> Domain:
    return Domain(d.Q, d.coords[sel], d.site_type[sel], d.wall_mask[sel], d.iolet_mask[sel],
                  d.iolet_id[sel], d.cut[sel], d.normal[sel], d.inlets, d.outlets)


def split_domain(d: Domain, parts: int = 2):
    """Cuts `d` into `parts` subdomains of contiguous 8^3 blocks in Morton
    order with about equal site counts (the reference balances ParMETIS
    partitions; whole blocks keep each subdomain's layout dense).  Returns
    per part (owned sites, their indices into d, ghost sites): the sites of
    the other parts from which a site of this part pulls along a fluid link
    (the reference's shared distributions, Domain.cc:166-179)."""
    from .lb import LATTICES
    c = LATTICES[d.Q].VECTORS
    bkey = _morton_block_key(d.coords)
    order = np.argsort(bkey, kind="stable")
    ub, first = np.unique(bkey[order], return_index=True)
    cum = np.append(first, d.n)
    part_of_block = np.minimum(
        (np.searchsorted(cum, (np.arange(1, parts) * d.n) / parts)[None, :] <=
         np.arange(len(ub))[:, None]).sum(1), parts - 1)
    part = np.empty(d.n, np.int64)
    part[order] = part_of_block[np.searchsorted(ub, bkey[order])]
    # site lookup by coordinate key
    lo = d.coords.min(0) - 1
    span = d.coords.max(0) - lo + 2
    ck = lambda x: ((x[:, 0] - lo[0]) * span[1] + (x[:, 1] - lo[1])) * span[2] + (x[:, 2] - lo[2])  # noqa: E731
    keys = ck(d.coords.astype(np.int64))
    ko = np.argsort(keys)
    ks = keys[ko]
    cut = d.wall_mask | d.iolet_mask
    ghost = [np.zeros(d.n, bool) for _ in range(parts)]
    for e in range(1, d.Q):
        t = d.coords.astype(np.int64) + c[e]
        tk = ck(t)
        pos = np.clip(np.searchsorted(ks, tk), 0, d.n - 1)
        hit = (ks[pos] == tk) & (((cut >> np.uint32(e - 1)) & 1) == 0)
        src = np.flatnonzero(hit)
        dst = ko[pos[hit]]
        across = part[src] != part[dst]
        for p in range(parts):
            ghost[p][src[across & (part[dst] == p)]] = True
    out = []
    for p in range(parts):
        idx = np.flatnonzero(part == p)
        out.append((_subset(d, idx), idx, _subset(d, np.flatnonzero(ghost[p]))))
    return out
