// Geometry input for the MI355X HemeLB timestep: synthetic generators with
// exact cut distances, .gmy v4 reading/writing, and projection onto a lattice
// in the reference's local site order.  Host code only (built with g++).
//
// Reference behaviour followed:
//   .gmy layout            Code/io/formats/geometry.h:26-157,
//                          Code/geometry/GeometryReader.cc:188-281,556-653
//   link -> mask/type      Code/geometry/SiteDataBare.cc:23-75
//   local site order       Code/geometry/Domain.cc:134-228,288-357,
//                          Code/geometry/LookupTree.h:92-97 (Morton order)
//   cut distances          geometry-tool .../CylinderGenerator.cpp:91-203
//                          (fraction of the link vector, closest hit wins)
#include "hemelb_geometry.h"

#include <zlib.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;

struct V3 {
  double x, y, z;
  V3 operator+(const V3& o) const { return {x + o.x, y + o.y, z + o.z}; }
  V3 operator-(const V3& o) const { return {x - o.x, y - o.y, z - o.z}; }
  V3 operator*(double s) const { return {x * s, y * s, z * s}; }
};
double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(const V3& a, const V3& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
double norm(const V3& a) { return std::sqrt(dot(a, a)); }
V3 unit(const V3& a) { return a * (1.0 / norm(a)); }

// Moore neighbourhood in .gmy order (io/formats/geometry.h:120-156).
struct Moore {
  int c[26][3];
  Moore() {
    int k = 0;
    for (int x = -1; x <= 1; ++x)
      for (int y = -1; y <= 1; ++y)
        for (int z = -1; z <= 1; ++z) {
          if (!x && !y && !z) continue;
          c[k][0] = x;
          c[k][1] = y;
          c[k][2] = z;
          ++k;
        }
  }
  int index(int x, int y, int z) const {
    int i = (x + 1) * 9 + (y + 1) * 3 + (z + 1);
    return i > 13 ? i - 1 : i;
  }
};
const Moore kMoore;

struct LinkSet {
  uint8_t type[26];
  int32_t iolet[26];
  float dist[26];
};

}  // namespace

struct hlbg_geometry {
  int32_t block_size = 8;
  int32_t blocks[3] = {0, 0, 0};
  // Fluid sites in .gmy order (blocks z-fastest, then sites z-fastest).
  std::vector<int32_t> coords;        // [n][3]
  std::vector<int32_t> link_index;    // [n], -1 when no link is cut
  std::vector<LinkSet> links;         // one per site with cut links
  std::vector<float> normal;          // [n][3]
  std::vector<uint8_t> has_normal;    // [n]
  std::vector<hlbg_iolet> inlets, outlets;
  int64_t size() const { return (int64_t)has_normal.size(); }
};

namespace {

// ---------------------------------------------------------------------------
// Sparse occupancy on a block grid.
// ---------------------------------------------------------------------------
struct Occupancy {
  int32_t bs, nb[3];
  std::vector<int32_t> block_id;             // dense over the block grid
  std::vector<std::vector<uint64_t>> bits;   // per allocated block
  Occupancy(int32_t bs_, const int32_t ext[3]) : bs(bs_) {
    for (int a = 0; a < 3; ++a) nb[a] = (ext[a] + bs - 1) / bs;
    block_id.assign((size_t)nb[0] * nb[1] * nb[2], -1);
  }
  size_t bidx(int bx, int by, int bz) const {
    return ((size_t)bx * nb[1] + by) * nb[2] + bz;
  }
  bool in_box(int x, int y, int z) const {
    return x >= 0 && y >= 0 && z >= 0 && x < nb[0] * bs && y < nb[1] * bs &&
           z < nb[2] * bs;
  }
  void set(int x, int y, int z) {
    size_t b = bidx(x / bs, y / bs, z / bs);
    if (block_id[b] < 0) {
      block_id[b] = (int32_t)bits.size();
      bits.emplace_back(((size_t)bs * bs * bs + 63) / 64, 0ull);
    }
    int l = ((x % bs) * bs + (y % bs)) * bs + (z % bs);
    bits[block_id[b]][l >> 6] |= 1ull << (l & 63);
  }
  bool get(int x, int y, int z) const {
    if (!in_box(x, y, z)) return false;
    int32_t id = block_id[bidx(x / bs, y / bs, z / bs)];
    if (id < 0) return false;
    int l = ((x % bs) * bs + (y % bs)) * bs + (z % bs);
    return (bits[id][l >> 6] >> (l & 63)) & 1ull;
  }
};

// ---------------------------------------------------------------------------
// Shapes.  A shape provides inside(p) and, for a link from a fluid point p to
// a non-fluid point p+c, the parameter t in (0,1] of the first boundary hit
// along the link together with its cut type and iolet id.
// ---------------------------------------------------------------------------
struct IoletPlane {
  V3 c, n;       // centre, inward unit normal
  double r;      // radius of the vessel at the plane
  double reach;  // cut region: radial <= reach, depth <= reach behind plane
  int type;      // HLBG_CUT_INLET / HLBG_CUT_OUTLET
  int id;
  bool cuts(const V3& p) const {
    V3 d = p - c;
    double s = dot(d, n);
    if (s >= 0.0 || s < -reach) return false;
    V3 rad = d - n * s;
    return dot(rad, rad) <= reach * reach;
  }
  // Crossing of the segment p + t c, t in (0, 1].
  bool hit(const V3& p, const V3& cv, double* t) const {
    double den = dot(cv, n);
    if (std::fabs(den) < 1e-14) return false;
    if (dot(p - c, n) < 0.0) return false;
    double tt = dot(c - p, n) / den;
    if (!(tt >= 0.0 && tt <= 1.0 + 1e-9)) return false;
    V3 x = p + cv * tt;
    V3 d = x - c;
    V3 rad = d - n * dot(d, n);
    if (dot(rad, rad) > reach * reach) return false;
    *t = tt;
    return true;
  }
};

struct Capsule {
  V3 a, b;
  double r;
  V3 lo, hi;  // bounding box
  void finish() {
    lo = {std::min(a.x, b.x) - r, std::min(a.y, b.y) - r, std::min(a.z, b.z) - r};
    hi = {std::max(a.x, b.x) + r, std::max(a.y, b.y) + r, std::max(a.z, b.z) + r};
  }
  double dist2(const V3& p) const {
    V3 ab = b - a;
    double h2 = dot(ab, ab);
    double s = h2 > 0 ? dot(p - a, ab) / h2 : 0.0;
    s = std::min(1.0, std::max(0.0, s));
    V3 q = a + ab * s;
    V3 d = p - q;
    return dot(d, d);
  }
  bool inside(const V3& p) const { return dist2(p) < r * r; }
  // Interval of t (on the infinite line p + t c) inside the capsule.
  bool interval(const V3& p, const V3& cv, double* t0, double* t1) const {
    double lo_t = 1e300, hi_t = -1e300;
    auto quad = [&](double A, double B, double C, double* r0, double* r1) {
      // A t^2 + B t + C <= 0
      if (std::fabs(A) < 1e-300) return false;
      double disc = B * B - 4 * A * C;
      if (disc <= 0) return false;
      double sq = std::sqrt(disc);
      *r0 = (-B - sq) / (2 * A);
      *r1 = (-B + sq) / (2 * A);
      return true;
    };
    // spheres at both ends
    for (const V3* e : {&a, &b}) {
      V3 w = p - *e;
      double r0, r1;
      if (quad(dot(cv, cv), 2 * dot(w, cv), dot(w, w) - r * r, &r0, &r1)) {
        lo_t = std::min(lo_t, r0);
        hi_t = std::max(hi_t, r1);
      }
    }
    // cylinder between the end planes
    V3 ab = b - a;
    double h = norm(ab);
    if (h > 0) {
      V3 u = ab * (1.0 / h);
      V3 w = p - a;
      double wu = dot(w, u), cu = dot(cv, u);
      V3 rho0 = w - u * wu, rho1 = cv - u * cu;
      double A = dot(rho1, rho1), B = 2 * dot(rho0, rho1),
             C = dot(rho0, rho0) - r * r;
      double c0, c1;
      bool ok;
      if (A < 1e-300) {
        ok = C < 0;
        c0 = -1e300;
        c1 = 1e300;
      } else {
        ok = quad(A, B, C, &c0, &c1);
      }
      if (ok) {
        // slab 0 <= wu + t cu <= h
        double s0 = -1e300, s1 = 1e300;
        if (std::fabs(cu) > 1e-300) {
          double ta = (0 - wu) / cu, tb = (h - wu) / cu;
          s0 = std::min(ta, tb);
          s1 = std::max(ta, tb);
        } else if (!(wu >= 0 && wu <= h)) {
          s0 = 1;
          s1 = 0;
        }
        double i0 = std::max(c0, s0), i1 = std::min(c1, s1);
        if (i0 < i1) {
          lo_t = std::min(lo_t, i0);
          hi_t = std::max(hi_t, i1);
        }
      }
    }
    if (lo_t >= hi_t) return false;
    *t0 = lo_t;
    *t1 = hi_t;
    return true;
  }
  V3 outward_normal(const V3& x) const {
    V3 ab = b - a;
    double h2 = dot(ab, ab);
    double s = h2 > 0 ? dot(x - a, ab) / h2 : 0.0;
    s = std::min(1.0, std::max(0.0, s));
    V3 d = x - (a + ab * s);
    double nn = norm(d);
    return nn > 0 ? d * (1.0 / nn) : V3{0, 0, 0};
  }
};

struct Shape {
  virtual ~Shape() = default;
  virtual bool inside(const V3& p) const = 0;
  // Wall exit along p + t c (p fluid); returns t (may exceed 1) and normal.
  virtual bool wall_exit(const V3& p, const V3& cv, double* t, V3* nrm) const = 0;
  std::vector<IoletPlane> iolets;
  bool cut_by_iolet(const V3& p) const {
    for (auto& io : iolets)
      if (io.cuts(p)) return true;
    return false;
  }
  bool fluid(const V3& p) const { return inside(p) && !cut_by_iolet(p); }
};

struct CapsuleUnion : Shape {
  std::vector<Capsule> caps;
  bool inside(const V3& p) const override {
    for (auto& c : caps) {
      if (p.x < c.lo.x || p.y < c.lo.y || p.z < c.lo.z || p.x > c.hi.x ||
          p.y > c.hi.y || p.z > c.hi.z)
        continue;
      if (c.inside(p)) return true;
    }
    return false;
  }
  bool wall_exit(const V3& p, const V3& cv, double* t, V3* nrm) const override {
    struct Iv { double a, b; int k; };
    Iv iv[64];
    int n = 0;
    for (int k = 0; k < (int)caps.size() && n < 64; ++k) {
      const Capsule& c = caps[k];
      if (p.x < c.lo.x - 2 || p.y < c.lo.y - 2 || p.z < c.lo.z - 2 ||
          p.x > c.hi.x + 2 || p.y > c.hi.y + 2 || p.z > c.hi.z + 2)
        continue;
      double t0, t1;
      if (c.interval(p, cv, &t0, &t1)) iv[n++] = {t0, t1, k};
    }
    double cur = 0.0;
    int who = -1;
    bool grew = true;
    while (grew) {
      grew = false;
      for (int i = 0; i < n; ++i)
        if (iv[i].a <= cur + 1e-12 && iv[i].b > cur) {
          cur = iv[i].b;
          who = iv[i].k;
          grew = true;
        }
    }
    if (who < 0) return false;
    *t = cur;
    *nrm = caps[who].outward_normal(p + cv * cur);
    return true;
  }
};

// Axisymmetric tube along z: fluid where r < R(z); R constant or stenosed.
struct Tube : Shape {
  double cx, cy, R, delta, z0, w;
  double Rz(double z) const {
    if (delta == 0.0) return R;
    double s = (z - z0) / w;
    return R * (1.0 - delta * std::exp(-0.5 * s * s));
  }
  double g(const V3& p) const {
    double dx = p.x - cx, dy = p.y - cy, rz = Rz(p.z);
    return dx * dx + dy * dy - rz * rz;
  }
  bool inside(const V3& p) const override { return g(p) < 0.0; }
  bool wall_exit(const V3& p, const V3& cv, double* t, V3* nrm) const override {
    V3 q = p + cv;
    double tt;
    if (delta == 0.0) {
      // exact ray / infinite-cylinder intersection (CylinderGenerator.cpp:112-150)
      double ax = p.x - cx, ay = p.y - cy;
      double A = cv.x * cv.x + cv.y * cv.y;
      if (A < 1e-300) return false;
      double B = 2 * (ax * cv.x + ay * cv.y);
      double C = ax * ax + ay * ay - R * R;
      double disc = B * B - 4 * A * C;
      if (disc <= 0) return false;
      tt = (-B + std::sqrt(disc)) / (2 * A);
      if (!(tt > 0)) return false;
    } else {
      if (g(q) < 0.0) return false;
      double lo = 0.0, hi = 1.0;
      for (int it = 0; it < 200 && hi - lo > 1e-16; ++it) {
        double mid = 0.5 * (lo + hi);
        if (g(p + cv * mid) < 0.0) lo = mid; else hi = mid;
      }
      tt = 0.5 * (lo + hi);
    }
    *t = tt;
    V3 x = p + cv * tt;
    V3 rad = {x.x - cx, x.y - cy, 0.0};
    double nn = norm(rad);
    *nrm = nn > 0 ? rad * (1.0 / nn) : V3{0, 0, 0};
    return true;
  }
};

// splitmix64 stream: the generator's only source of randomness, so that a
// seed gives the same tree on every platform.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform(double a, double b) {
    return a + (b - a) * ((double)(next() >> 11) * (1.0 / 9007199254740992.0));
  }
};

// Voxelise `shape` over the integer points of [lo, hi] (lattice coords) and
// fill `g` with the fluid sites and their 26 links.
void voxelise(const Shape& shape, const int32_t lo[3], const int32_t hi[3],
              int32_t block_size, const std::vector<std::pair<V3, V3>>& cap_boxes,
              hlbg_geometry* g) {
  // Lattice coordinates are shifted so that the box starts at 1 (one solid
  // layer at 0), as the sites at coordinate 0 never need to be fluid.
  const int32_t ext[3] = {hi[0] - lo[0] + 3, hi[1] - lo[1] + 3, hi[2] - lo[2] + 3};
  Occupancy occ(block_size, ext);
  auto mark_box = [&](int32_t x0, int32_t x1, int32_t y0, int32_t y1, int32_t z0,
                      int32_t z1) {
    for (int32_t x = x0; x <= x1; ++x)
      for (int32_t y = y0; y <= y1; ++y)
        for (int32_t z = z0; z <= z1; ++z) {
          V3 p{(double)x, (double)y, (double)z};
          if (shape.fluid(p)) occ.set(x - lo[0] + 1, y - lo[1] + 1, z - lo[2] + 1);
        }
  };
  if (cap_boxes.empty()) {
    mark_box(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]);
  } else {
    for (auto& bx : cap_boxes)
      mark_box(std::max(lo[0], (int32_t)std::floor(bx.first.x)),
               std::min(hi[0], (int32_t)std::ceil(bx.second.x)),
               std::max(lo[1], (int32_t)std::floor(bx.first.y)),
               std::min(hi[1], (int32_t)std::ceil(bx.second.y)),
               std::max(lo[2], (int32_t)std::floor(bx.first.z)),
               std::min(hi[2], (int32_t)std::ceil(bx.second.z)));
  }
  g->block_size = block_size;
  for (int a = 0; a < 3; ++a) g->blocks[a] = occ.nb[a];
  // Shift the iolet descriptions into the same frame.
  const V3 shift{(double)(1 - lo[0]), (double)(1 - lo[1]), (double)(1 - lo[2])};
  for (auto& io : shape.iolets) {
    hlbg_iolet d;
    V3 c = io.c + shift;
    d.position[0] = c.x; d.position[1] = c.y; d.position[2] = c.z;
    d.normal[0] = io.n.x; d.normal[1] = io.n.y; d.normal[2] = io.n.z;
    d.radius = io.r;
    (io.type == HLBG_CUT_INLET ? g->inlets : g->outlets).push_back(d);
  }
  // Fluid sites in block order (block-major, z fastest inside a block).
  const int bs = block_size;
  std::vector<int32_t> sites;
  for (int bx = 0; bx < occ.nb[0]; ++bx)
    for (int by = 0; by < occ.nb[1]; ++by)
      for (int bz = 0; bz < occ.nb[2]; ++bz) {
        if (occ.block_id[occ.bidx(bx, by, bz)] < 0) continue;
        for (int lx = 0; lx < bs; ++lx)
          for (int ly = 0; ly < bs; ++ly)
            for (int lz = 0; lz < bs; ++lz) {
              int x = bx * bs + lx, y = by * bs + ly, z = bz * bs + lz;
              if (!occ.get(x, y, z)) continue;
              sites.push_back(x);
              sites.push_back(y);
              sites.push_back(z);
            }
      }
  const int64_t n = (int64_t)sites.size() / 3;
  // Links of every site: independent per site, computed by a pool of host
  // threads over contiguous chunks (the cut searches dominate set-up time).
  std::vector<LinkSet> ls_all((size_t)n);
  std::vector<uint8_t> any_all((size_t)n), wall_all((size_t)n);
  std::vector<float> nrm_all((size_t)n * 3);
  auto work = [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i) {
      const int32_t x = sites[3 * i], y = sites[3 * i + 1], z = sites[3 * i + 2];
      V3 p{(double)(x - 1 + lo[0]), (double)(y - 1 + lo[1]), (double)(z - 1 + lo[2])};
      LinkSet& ls = ls_all[(size_t)i];
      bool any = false, wall = false;
      V3 nrm_acc{0, 0, 0};
      for (int k = 0; k < 26; ++k) {
        ls.type[k] = HLBG_CUT_NONE;
        ls.iolet[k] = -1;
        ls.dist[k] = -1.0f;
        const int* c = kMoore.c[k];
        if (occ.get(x + c[0], y + c[1], z + c[2])) continue;
        any = true;
        V3 cv{(double)c[0], (double)c[1], (double)c[2]};
        double best = 1e300;
        int type = HLBG_CUT_WALL, id = -1;
        V3 nrm{0, 0, 0};
        double tw;
        V3 nw;
        if (shape.wall_exit(p, cv, &tw, &nw) && tw > 0.0 && tw <= 1.0 + 1e-9) {
          best = tw;
          nrm = nw;
        }
        for (auto& io : shape.iolets) {
          double ti;
          if (io.hit(p, cv, &ti) && ti < best) {
            best = ti;
            type = io.type;
            id = io.id;
          }
        }
        if (best > 1.0) {
          // No hit inside the link (round-off at a tangency): as the
          // geometry tool does, place the cut at the far end.
          best = 1.0;
          if (type == HLBG_CUT_WALL && !shape.wall_exit(p, cv, &tw, &nrm)) nrm = V3{0, 0, 0};
        }
        ls.type[k] = (uint8_t)type;
        ls.iolet[k] = id;
        ls.dist[k] = (float)best;
        if (type == HLBG_CUT_WALL) {
          wall = true;
          nrm_acc = nrm_acc + nrm;
        }
      }
      any_all[(size_t)i] = any ? 1 : 0;
      wall_all[(size_t)i] = wall ? 1 : 0;
      V3 u{0, 0, 0};
      if (wall) {
        double nn = norm(nrm_acc);
        u = nn > 0 ? nrm_acc * (1.0 / nn) : V3{0, 0, 1};
      }
      nrm_all[3 * i] = (float)u.x;
      nrm_all[3 * i + 1] = (float)u.y;
      nrm_all[3 * i + 2] = (float)u.z;
    }
  };
  const int64_t chunk = 4096;
  const int nthreads = (int)std::max<int64_t>(
      1, std::min<int64_t>({(int64_t)std::max(1u, std::thread::hardware_concurrency()), 16,
                            (n + chunk - 1) / chunk}));
  if (nthreads == 1) {
    work(0, n);
  } else {
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
      pool.emplace_back([&] {
        for (int64_t i0; (i0 = next.fetch_add(chunk)) < n;) work(i0, std::min(n, i0 + chunk));
      });
    for (auto& th : pool) th.join();
  }
  g->coords = std::move(sites);
  g->link_index.reserve((size_t)n);
  g->has_normal.reserve((size_t)n);
  g->normal = std::move(nrm_all);
  for (int64_t i = 0; i < n; ++i) {
    if (any_all[(size_t)i]) {
      g->link_index.push_back((int32_t)g->links.size());
      g->links.push_back(ls_all[(size_t)i]);
    } else {
      g->link_index.push_back(-1);
    }
    g->has_normal.push_back(wall_all[(size_t)i]);
  }
}

int fail(const std::string& msg) {
  g_err = msg;
  return -1;
}

// ---------------------------------------------------------------------------
// XDR (big-endian) helpers for .gmy
// ---------------------------------------------------------------------------
void put_u32(std::vector<unsigned char>& b, uint32_t v) {
  b.push_back((unsigned char)(v >> 24));
  b.push_back((unsigned char)(v >> 16));
  b.push_back((unsigned char)(v >> 8));
  b.push_back((unsigned char)v);
}
void put_f32(std::vector<unsigned char>& b, float f) {
  uint32_t v;
  std::memcpy(&v, &f, 4);
  put_u32(b, v);
}
struct XdrIn {
  const unsigned char* p;
  size_t n, i = 0;
  uint32_t u32() {
    if (i + 4 > n) throw std::runtime_error("truncated XDR data");
    uint32_t v = ((uint32_t)p[i] << 24) | ((uint32_t)p[i + 1] << 16) |
                 ((uint32_t)p[i + 2] << 8) | (uint32_t)p[i + 3];
    i += 4;
    return v;
  }
  float f32() {
    uint32_t v = u32();
    float f;
    std::memcpy(&f, &v, 4);
    return f;
  }
};

constexpr uint32_t kHemeLbMagic = 0x686c6221;  // "hlb!" (io/formats/formats.h)
constexpr uint32_t kGmyMagic = 0x676d7904;     // io/formats/geometry.h:33
constexpr uint32_t kGmyVersion = 4;            // io/formats/geometry.h:36

uint64_t spread3(uint32_t v) {  // LookupTree.h bitspread_one
  uint64_t x = v & 0xffff;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

}  // namespace

extern "C" {

const char* hlbg_last_error(void) { return g_err.c_str(); }

namespace {
// Squared Euclidean distance transform of a sampled function along one line
// (lower envelope of parabolas; Felzenszwalb & Huttenlocher, "Distance
// Transforms of Sampled Functions", Theory of Computing 8, 2012).
void edt_line(const double* f, int n, double* d, int* v, double* z) {
  int k = 0;
  v[0] = 0;
  z[0] = -1e300;
  z[1] = 1e300;
  for (int q = 1; q < n; ++q) {
    double s = ((f[q] + (double)q * q) - (f[v[k]] + (double)v[k] * v[k])) / (2.0 * (q - v[k]));
    while (s <= z[k]) {  // z[0] = -inf ends the loop at k = 0
      --k;
      s = ((f[q] + (double)q * q) - (f[v[k]] + (double)v[k] * v[k])) / (2.0 * (q - v[k]));
    }
    ++k;
    v[k] = q;
    z[k] = s;
    z[k + 1] = 1e300;
  }
  k = 0;
  for (int q = 0; q < n; ++q) {
    while (z[k + 1] < q) ++k;
    const double dq = q - v[k];
    d[q] = dq * dq + f[v[k]];
  }
}

// Squared distance from every voxel to the nearest voxel with seed != 0.
void edt3(const std::vector<uint8_t>& seed, const int32_t* dims, std::vector<double>& out) {
  const int64_t nx = dims[0], ny = dims[1], nz = dims[2];
  const double inf = 1e20;
  out.assign((size_t)(nx * ny * nz), 0.0);
  for (size_t i = 0; i < out.size(); ++i) out[i] = seed[i] ? 0.0 : inf;
  const int m = (int)std::max(nx, std::max(ny, nz));
  std::vector<double> f(m), d(m), z(m + 1);
  std::vector<int> v(m);
  auto pass = [&](int64_t n, int64_t stride, int64_t count, auto base_of) {
    for (int64_t c = 0; c < count; ++c) {
      const int64_t b = base_of(c);
      for (int64_t i = 0; i < n; ++i) f[i] = out[b + i * stride];
      edt_line(f.data(), (int)n, d.data(), v.data(), z.data());
      for (int64_t i = 0; i < n; ++i) out[b + i * stride] = d[i];
    }
  };
  // index = (x * ny + y) * nz + z
  pass(nz, 1, nx * ny, [&](int64_t c) { return c * nz; });
  pass(ny, nz, nx * nz, [&](int64_t c) { return (c / nz) * ny * nz + (c % nz); });
  pass(nx, ny * nz, ny * nz, [&](int64_t c) { return c; });
}
}  // namespace

int hlbg_signed_distance(const hlbg_geometry* g, const int32_t origin[3],
                         const int32_t dims[3], float* out) {
  if (!g || !origin || !dims || !out) return fail("null argument");
  for (int a = 0; a < 3; ++a)
    if (dims[a] < 1 || dims[a] > 4096) return fail("crop dimensions must be 1..4096");
  const int64_t nx = dims[0], ny = dims[1], nz = dims[2];
  std::vector<uint8_t> fluid((size_t)(nx * ny * nz), 0), solid;
  for (int64_t s = 0; s < g->size(); ++s) {
    const int64_t x = g->coords[3 * s] - origin[0], y = g->coords[3 * s + 1] - origin[1],
                  z = g->coords[3 * s + 2] - origin[2];
    if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) continue;
    fluid[(size_t)((x * ny + y) * nz + z)] = 1;
  }
  solid.resize(fluid.size());
  for (size_t i = 0; i < fluid.size(); ++i) solid[i] = !fluid[i];
  std::vector<double> to_solid, to_fluid;
  edt3(solid, dims, to_solid);
  edt3(fluid, dims, to_fluid);
  for (size_t i = 0; i < fluid.size(); ++i)
    out[i] = fluid[i] ? (float)std::sqrt(to_solid[i]) : -(float)std::sqrt(to_fluid[i]);
  return 0;
}

void hlbg_free(hlbg_geometry* g) { delete g; }

int64_t hlbg_fluid_site_count(const hlbg_geometry* g) { return g ? g->size() : -1; }

int hlbg_block_info(const hlbg_geometry* g, int32_t* blocks, int32_t* bs) {
  if (!g) return fail("null geometry");
  for (int a = 0; a < 3; ++a) blocks[a] = g->blocks[a];
  *bs = g->block_size;
  return 0;
}

int32_t hlbg_iolet_count(const hlbg_geometry* g, int32_t outlet) {
  if (!g) return -1;
  return (int32_t)(outlet ? g->outlets.size() : g->inlets.size());
}

int hlbg_get_iolet(const hlbg_geometry* g, int32_t outlet, int32_t i,
                   hlbg_iolet* out) {
  if (!g) return fail("null geometry");
  const auto& v = outlet ? g->outlets : g->inlets;
  if (i < 0 || i >= (int32_t)v.size()) return fail("iolet index out of range");
  *out = v[i];
  return 0;
}

int hlbg_set_iolets(hlbg_geometry* g, const hlbg_iolet* in, int32_t n_in,
                    const hlbg_iolet* outl, int32_t n_out) {
  if (!g) return fail("null geometry");
  g->inlets.assign(in, in + n_in);
  g->outlets.assign(outl, outl + n_out);
  return 0;
}

int hlbg_cylinder(double radius, int32_t length, int32_t block_size,
                  hlbg_geometry** out) {
  return hlbg_stenosed_cylinder(radius, length, 0.0, 0.0, 1.0, block_size, out);
}

int hlbg_stenosed_cylinder(double radius, int32_t length, double severity,
                           double z0, double width, int32_t block_size,
                           hlbg_geometry** out) {
  if (!(radius > 0) || length < 2 || block_size < 2 || !(width > 0) ||
      severity < 0 || severity >= 1)
    return fail("invalid cylinder parameters");
  try {
    Tube t;
    t.R = radius;
    t.cx = t.cy = 0.0;
    t.delta = 1.0 - std::sqrt(1.0 - severity);
    t.z0 = z0;
    t.w = width;
    // fluid z in [0, length-1]; iolet planes half a voxel outside
    IoletPlane in{{0, 0, -0.5}, {0, 0, 1}, radius, radius + 1.5, HLBG_CUT_INLET, 0};
    IoletPlane outp{{0, 0, length - 0.5}, {0, 0, -1}, radius, radius + 1.5,
                    HLBG_CUT_OUTLET, 0};
    t.iolets = {in, outp};
    const int32_t R = (int32_t)std::ceil(radius) + 1;
    // Axis at a half-integer offset so no site lies on it.
    t.cx = t.cy = 0.5;
    t.iolets[0].c = {0.5, 0.5, -0.5};
    t.iolets[1].c = {0.5, 0.5, length - 0.5};
    const int32_t lo[3] = {-R, -R, -1}, hi[3] = {R + 1, R + 1, length};
    auto g = std::make_unique<hlbg_geometry>();
    voxelise(t, lo, hi, block_size, {}, g.get());
    *out = g.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e.what());
  }
}

int hlbg_vessel_tree(uint64_t seed, int32_t generations, double root_radius,
                     int32_t block_size, hlbg_geometry** out) {
  if (generations < 1 || generations > 8 || !(root_radius > 1) || block_size < 2)
    return fail("invalid vessel tree parameters");
  try {
    Rng rng(seed);
    CapsuleUnion u;
    struct Seg { V3 a, dir; double r; int gen; };
    // The root axis runs through (0.5, 0.5, z) so that no lattice site lies on
    // it, and the inlet plane sits half a voxel below the first fluid layer.
    std::vector<Seg> stack{{{0.5, 0.5, -0.5}, {0, 0, 1}, root_radius, 0}};
    std::vector<IoletPlane> outlets;
    int n_out = 0;
    const double deg = 3.14159265358979323846 / 180.0;
    while (!stack.empty()) {
      Seg s = stack.back();
      stack.pop_back();
      double len = 8.0 * s.r * rng.uniform(0.8, 1.2);
      Capsule c{s.a, s.a + s.dir * len, s.r, {}, {}};
      c.finish();
      u.caps.push_back(c);
      if (s.gen + 1 < generations) {
        double frac = rng.uniform(0.4, 0.6);
        double r1 = s.r * std::cbrt(frac), r2 = s.r * std::cbrt(1.0 - frac);
        double th1 = rng.uniform(25.0, 45.0) * deg, th2 = rng.uniform(25.0, 45.0) * deg;
        double phi = rng.uniform(0.0, 2.0 * 3.14159265358979323846);
        // Orthonormal frame around the parent direction.
        V3 ref = std::fabs(s.dir.x) < 0.9 ? V3{1, 0, 0} : V3{0, 1, 0};
        V3 e1 = unit(cross(s.dir, ref));
        V3 e2 = cross(s.dir, e1);
        V3 side = e1 * std::cos(phi) + e2 * std::sin(phi);
        V3 d1 = unit(s.dir * std::cos(th1) + side * std::sin(th1));
        V3 d2 = unit(s.dir * std::cos(th2) - side * std::sin(th2));
        stack.push_back({c.b, d2, r2, s.gen + 1});
        stack.push_back({c.b, d1, r1, s.gen + 1});
      } else {
        outlets.push_back({c.b, s.dir * -1.0, s.r, s.r + 1.5, HLBG_CUT_OUTLET, n_out++});
      }
    }
    u.iolets.push_back({{0.5, 0.5, -0.5}, {0, 0, 1}, root_radius, root_radius + 1.5,
                        HLBG_CUT_INLET, 0});
    for (auto& o : outlets) u.iolets.push_back(o);
    V3 lo{1e300, 1e300, 1e300}, hi{-1e300, -1e300, -1e300};
    std::vector<std::pair<V3, V3>> boxes;
    for (auto& c : u.caps) {
      lo = {std::min(lo.x, c.lo.x), std::min(lo.y, c.lo.y), std::min(lo.z, c.lo.z)};
      hi = {std::max(hi.x, c.hi.x), std::max(hi.y, c.hi.y), std::max(hi.z, c.hi.z)};
      boxes.push_back({c.lo, c.hi});
    }
    const int32_t ilo[3] = {(int32_t)std::floor(lo.x) - 1, (int32_t)std::floor(lo.y) - 1,
                            (int32_t)std::floor(lo.z) - 1};
    const int32_t ihi[3] = {(int32_t)std::ceil(hi.x) + 1, (int32_t)std::ceil(hi.y) + 1,
                            (int32_t)std::ceil(hi.z) + 1};
    auto g = std::make_unique<hlbg_geometry>();
    voxelise(u, ilo, ihi, block_size, boxes, g.get());
    *out = g.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e.what());
  }
}

int hlbg_get_links(const hlbg_geometry* g, int32_t* coords, uint8_t* link_type,
                   int32_t* link_iolet, float* link_dist, float* normal,
                   uint8_t* has_normal) {
  if (!g) return fail("null geometry");
  const int64_t n = g->size();
  if (coords) std::memcpy(coords, g->coords.data(), sizeof(int32_t) * 3 * n);
  if (normal) std::memcpy(normal, g->normal.data(), sizeof(float) * 3 * n);
  if (has_normal) std::memcpy(has_normal, g->has_normal.data(), n);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t li = g->link_index[i];
    for (int k = 0; k < 26; ++k) {
      if (link_type) link_type[i * 26 + k] = li < 0 ? 0 : g->links[li].type[k];
      if (link_iolet) link_iolet[i * 26 + k] = li < 0 ? -1 : g->links[li].iolet[k];
      if (link_dist) link_dist[i * 26 + k] = li < 0 ? -1.0f : g->links[li].dist[k];
    }
  }
  return 0;
}

int hlbg_project(const hlbg_geometry* g, int32_t Q, int32_t* coords,
                 uint32_t* site_type, uint32_t* wall_mask, uint32_t* iolet_mask,
                 int32_t* iolet_id, float* cut, float* normal) {
  if (!g) return fail("null geometry");
  static const int C15[15][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},
                                 {0, -1, 0}, {0, 0, 1},   {0, 0, -1}, {1, 1, 1},
                                 {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1},
                                 {-1, 1, -1}, {1, -1, -1}, {-1, 1, 1}};
  static const int C19[19][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},
                                 {0, -1, 0}, {0, 0, 1},  {0, 0, -1}, {1, 1, 0},
                                 {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1},
                                 {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {0, 1, 1},
                                 {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
  const int(*C)[3];
  if (Q == 15) C = C15;
  else if (Q == 19) C = C19;
  else return fail("unsupported lattice (Q must be 15 or 19)");
  const int64_t n = g->size();
  struct Key { int32_t group; uint64_t oct; int32_t local; int64_t src; };
  std::vector<Key> keys(n);
  std::vector<uint32_t> wm(n), im(n), st(n);
  std::vector<int32_t> id(n);
  const int bs = g->block_size;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t li = g->link_index[i];
    uint32_t w = 0, io = 0;
    int32_t iid = -1;
    bool had_in = false, had_out = false;
    if (li >= 0) {
      const LinkSet& L = g->links[li];
      // SiteData::SiteData (SiteDataBare.cc:41-72)
      for (int d = 1; d < Q; ++d) {
        int k = kMoore.index(C[d][0], C[d][1], C[d][2]);
        switch (L.type[k]) {
          case HLBG_CUT_WALL: w |= 1u << (d - 1); break;
          case HLBG_CUT_INLET:
          case HLBG_CUT_OUTLET:
            iid = L.iolet[k];
            io |= 1u << (d - 1);
            if (L.type[k] == HLBG_CUT_INLET) had_in = true; else had_out = true;
            break;
          default: break;
        }
      }
    }
    wm[i] = w;
    im[i] = io;
    id[i] = iid;
    st[i] = had_in ? HLBG_SITE_INLET : (had_out ? HLBG_SITE_OUTLET : HLBG_SITE_FLUID);
    // Collision-type group 0..5 (Domain.cc:187-207)
    int grp;
    if (st[i] == HLBG_SITE_FLUID) grp = w ? 1 : 0;
    else if (st[i] == HLBG_SITE_INLET) grp = w ? 4 : 2;
    else grp = w ? 5 : 3;
    const int32_t x = g->coords[3 * i], y = g->coords[3 * i + 1], z = g->coords[3 * i + 2];
    const uint64_t oct = (spread3(x / bs) << 2) ^ (spread3(y / bs) << 1) ^ spread3(z / bs);
    keys[i] = {grp, oct, ((x % bs) * bs + (y % bs)) * bs + (z % bs), i};
  }
  std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& b) {
    if (a.group != b.group) return a.group < b.group;
    if (a.oct != b.oct) return a.oct < b.oct;
    return a.local < b.local;
  });
  for (int64_t o = 0; o < n; ++o) {
    const int64_t i = keys[o].src;
    if (coords)
      for (int a = 0; a < 3; ++a) coords[3 * o + a] = g->coords[3 * i + a];
    if (site_type) site_type[o] = st[i];
    if (wall_mask) wall_mask[o] = wm[i];
    if (iolet_mask) iolet_mask[o] = im[i];
    if (iolet_id) iolet_id[o] = id[i];
    if (normal)
      for (int a = 0; a < 3; ++a) normal[3 * o + a] = g->normal[3 * i + a];
    if (cut) {
      const int32_t li = g->link_index[i];
      for (int d = 1; d < Q; ++d) {
        float v = -1.0f;
        if (li >= 0) v = g->links[li].dist[kMoore.index(C[d][0], C[d][1], C[d][2])];
        cut[o * (Q - 1) + d - 1] = v;
      }
    }
  }
  return 0;
}

int hlbg_write_gmy(const hlbg_geometry* g, const char* path) {
  if (!g) return fail("null geometry");
  const int bs = g->block_size;
  const int64_t nblocks = (int64_t)g->blocks[0] * g->blocks[1] * g->blocks[2];
  const int64_t per = (int64_t)bs * bs * bs;
  // Sites are stored in .gmy order, so each block's fluid sites are a
  // contiguous run; index them by block and local site id.
  std::vector<int64_t> first(nblocks + 1, 0);
  std::vector<int64_t> bid(g->size());
  for (int64_t i = 0; i < g->size(); ++i) {
    const int32_t x = g->coords[3 * i], y = g->coords[3 * i + 1], z = g->coords[3 * i + 2];
    bid[i] = ((int64_t)(x / bs) * g->blocks[1] + y / bs) * g->blocks[2] + z / bs;
    first[bid[i] + 1]++;
  }
  for (int64_t b = 0; b < nblocks; ++b) first[b + 1] += first[b];
  std::vector<unsigned char> header, body;
  put_u32(header, kHemeLbMagic);
  put_u32(header, kGmyMagic);
  put_u32(header, kGmyVersion);
  for (int a = 0; a < 3; ++a) put_u32(header, (uint32_t)g->blocks[a]);
  put_u32(header, (uint32_t)bs);
  put_u32(header, 0);
  std::vector<unsigned char> raw, comp;
  for (int64_t b = 0; b < nblocks; ++b) {
    const int64_t s0 = first[b], s1 = first[b + 1];
    if (s0 == s1) {
      put_u32(header, 0);
      put_u32(header, 0);
      put_u32(header, 0);
      continue;
    }
    raw.clear();
    int64_t s = s0;
    for (int64_t l = 0; l < per; ++l) {
      bool fluid = false;
      if (s < s1) {
        const int32_t x = g->coords[3 * s], y = g->coords[3 * s + 1], z = g->coords[3 * s + 2];
        fluid = (((x % bs) * bs + (y % bs)) * bs + (z % bs)) == l;
      }
      if (!fluid) {
        put_u32(raw, 0);
        continue;
      }
      put_u32(raw, 1);
      const int32_t li = g->link_index[s];
      for (int k = 0; k < 26; ++k) {
        const uint32_t t = li < 0 ? 0 : g->links[li].type[k];
        put_u32(raw, t);
        if (t == HLBG_CUT_WALL) {
          put_f32(raw, g->links[li].dist[k]);
        } else if (t != HLBG_CUT_NONE) {
          put_u32(raw, (uint32_t)g->links[li].iolet[k]);
          put_f32(raw, g->links[li].dist[k]);
        }
      }
      put_u32(raw, g->has_normal[s]);
      if (g->has_normal[s])
        for (int a = 0; a < 3; ++a) put_f32(raw, g->normal[3 * s + a]);
      ++s;
    }
    uLongf clen = compressBound(raw.size());
    comp.resize(clen);
    if (compress2(comp.data(), &clen, raw.data(), raw.size(), Z_DEFAULT_COMPRESSION) != Z_OK)
      return fail("zlib compression failed");
    put_u32(header, (uint32_t)(s1 - s0));
    put_u32(header, (uint32_t)clen);
    put_u32(header, (uint32_t)raw.size());
    body.insert(body.end(), comp.begin(), comp.begin() + clen);
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(std::string("cannot open for writing: ") + path);
  bool ok = std::fwrite(header.data(), 1, header.size(), f) == header.size() &&
            std::fwrite(body.data(), 1, body.size(), f) == body.size();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : fail("write failed");
}

int hlbg_read_gmy(const char* path, hlbg_geometry** out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(std::string("cannot open: ") + path);
  std::vector<unsigned char> data;
  unsigned char buf[1 << 16];
  size_t r;
  while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) data.insert(data.end(), buf, buf + r);
  std::fclose(f);
  try {
    XdrIn pre{data.data(), data.size()};
    // GeometryReader::ReadPreamble (GeometryReader.cc:188-240)
    if (pre.u32() != kHemeLbMagic) return fail("not a HemeLB file (magic number)");
    if (pre.u32() != kGmyMagic) return fail("not a geometry file (magic number)");
    if (pre.u32() != kGmyVersion) return fail("unsupported geometry version");
    auto g = std::make_unique<hlbg_geometry>();
    for (int a = 0; a < 3; ++a) g->blocks[a] = (int32_t)pre.u32();
    g->block_size = (int32_t)pre.u32();
    pre.u32();
    const int bs = g->block_size;
    const int64_t nblocks = (int64_t)g->blocks[0] * g->blocks[1] * g->blocks[2];
    std::vector<uint32_t> nfl(nblocks), clen(nblocks), ulen(nblocks);
    for (int64_t b = 0; b < nblocks; ++b) {  // ReadHeader (GeometryReader.cc:248-281)
      nfl[b] = pre.u32();
      clen[b] = pre.u32();
      ulen[b] = pre.u32();
    }
    size_t off = pre.i;
    std::vector<unsigned char> raw;
    for (int64_t b = 0; b < nblocks; ++b) {
      if (nfl[b] == 0) continue;
      if (off + clen[b] > data.size()) return fail("truncated block data");
      raw.resize(ulen[b]);
      uLongf ul = ulen[b];
      if (uncompress(raw.data(), &ul, data.data() + off, clen[b]) != Z_OK || ul != ulen[b])
        return fail("block decompression failed");
      off += clen[b];
      XdrIn in{raw.data(), raw.size()};
      const int bx = (int)(b / ((int64_t)g->blocks[1] * g->blocks[2]));
      const int by = (int)((b / g->blocks[2]) % g->blocks[1]);
      const int bz = (int)(b % g->blocks[2]);
      uint32_t fluid = 0;
      for (int l = 0; l < bs * bs * bs; ++l) {  // ParseSite (GeometryReader.cc:556-653)
        uint32_t type = in.u32();
        if (type == 0) continue;
        if (type != 1) return fail("invalid site type");
        ++fluid;
        const int lx = l / (bs * bs), ly = (l / bs) % bs, lz = l % bs;
        g->coords.push_back(bx * bs + lx);
        g->coords.push_back(by * bs + ly);
        g->coords.push_back(bz * bs + lz);
        LinkSet ls;
        bool any = false, wall = false;
        for (int k = 0; k < 26; ++k) {
          uint32_t t = in.u32();
          ls.type[k] = (uint8_t)t;
          ls.iolet[k] = -1;
          ls.dist[k] = -1.0f;
          if (t == HLBG_CUT_WALL) {
            ls.dist[k] = in.f32();
            wall = true;
          } else if (t == HLBG_CUT_INLET || t == HLBG_CUT_OUTLET) {
            ls.iolet[k] = (int32_t)in.u32();
            ls.dist[k] = in.f32();
          } else if (t != HLBG_CUT_NONE) {
            return fail("invalid cut type");
          }
          any = any || t != HLBG_CUT_NONE;
        }
        uint32_t has = in.u32();
        if ((has != 0) != wall)
          return fail("malformed geometry: wall normal availability does not match wall links");
        g->has_normal.push_back(has ? 1 : 0);
        for (int a = 0; a < 3; ++a) g->normal.push_back(has ? in.f32() : 0.0f);
        if (any) {
          g->link_index.push_back((int32_t)g->links.size());
          g->links.push_back(ls);
        } else {
          g->link_index.push_back(-1);
        }
      }
      // the header declares each block's fluid-site count and uncompressed
      // length; a block that disagrees is malformed
      if (fluid != nfl[b])
        return fail("block " + std::to_string(b) + " holds " + std::to_string(fluid) +
                    " fluid sites, its header declares " + std::to_string(nfl[b]));
      if (in.i != (size_t)ulen[b])
        return fail("block " + std::to_string(b) + " has " +
                    std::to_string((size_t)ulen[b] - in.i) + " unread bytes");
    }
    *out = g.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e.what());
  }
}

}  // extern "C"
