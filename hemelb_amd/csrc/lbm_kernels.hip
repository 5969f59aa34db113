// HIP kernels of the MI355X HemeLB timestep (gfx950).
//
// One thread updates one fluid site.  The state holds post-collision
// distributions f*(t) structure-of-arrays (f[d*np + x]); a step pulls the
// values streaming into every slot of a site (the reference pushes them,
// Code/lb/streamers/BulkStreamer.h:31-39), resolves wall and iolet links in
// registers with the reference's link priority (iolet first, then wall,
// StreamerTypeFactory.h:65-78), applies the BFL PostStep correction
// (BouzidiFirdaousLallemand.h:68-91) locally, collides (LBGK.h:140-174 or
// MRT.h:56-105) and writes f*(t+1) to its own slots.  Binary64 arithmetic is
// evaluated in the reference's SSE3 order without contraction, so the plain
// method reproduces the reference bit for bit (up to the sign of zero).
//
// The bandwidth-bound path uses no MFMA.  By default consecutive workgroups
// (dispatched round-robin over the XCDs) cover consecutive x planes of the
// Morton-ordered domain; an XCD-contiguous remap is one of the launch shapes
// the tuner measures.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "lbm_internal.h"

// Static scratch of the allocation-free monitor: kMonitorGroups * 5
// workgroup partials, then (rho, u) of up to kStaticIoletSites iolet sites.
// Host code never names this array (its name is only a host-side stand-in);
// it uses the per-device address from hlb::scratch_address.
static __device__ __align__(256) double scratch[hlb::kStaticScratchDoubles];

namespace hlb {
namespace {

// ---------------------------------------------------------------------------
// Lattice constants (Code/lb/lattices/D3Q15.h:15-56, D3Q19.h:14-49)
// ---------------------------------------------------------------------------
template <int Q> struct Lat;
template <> struct Lat<15> {
  static constexpr int c[15][3] = {
      {0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
      {0, 0, 1},   {0, 0, -1},  {1, 1, 1},  {-1, -1, -1}, {1, 1, -1},
      {-1, -1, 1}, {1, -1, 1},  {-1, 1, -1}, {1, -1, -1}, {-1, 1, 1}};
  static constexpr double w[15] = {
      2.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
      1.0 / 9.0,  1.0 / 9.0,  1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0,
      1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0};
  static constexpr int nmom = 11;
  // DHumieresD3Q15MRTBasis::REDUCED_MOMENT_BASIS (DHumieresD3Q15MRTBasis.h:41-53)
  static constexpr int M[11][15] = {
      {-2, -1, -1, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1},
      {16, -4, -4, -4, -4, -4, -4, 1, 1, 1, 1, 1, 1, 1, 1},
      {0, -4, 4, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1},
      {0, 0, 0, -4, 4, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1},
      {0, 0, 0, 0, 0, -4, 4, 1, -1, -1, 1, 1, -1, -1, 1},
      {0, 2, 2, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0},
      {0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1},
      {0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, -1, -1, 1, 1},
      {0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1},
      {0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1}};
};
template <> struct Lat<19> {
  static constexpr int c[19][3] = {
      {0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
      {0, 0, 1},  {0, 0, -1},  {1, 1, 0},  {-1, -1, 0}, {1, -1, 0},
      {-1, 1, 0}, {1, 0, 1},   {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
      {0, 1, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
  static constexpr double w[19] = {
      1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
      1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
  static constexpr int nmom = 15;
  // DHumieresD3Q19MRTBasis::REDUCED_MOMENT_BASIS (DHumieresD3Q19MRTBasis.h:38-53)
  static constexpr int M[15][19] = {
      {-30, -11, -11, -11, -11, -11, -11, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8},
      {12, -4, -4, -4, -4, -4, -4, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
      {0, -4, 4, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0},
      {0, 0, 0, -4, 4, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1},
      {0, 0, 0, 0, 0, -4, 4, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1},
      {0, 2, 2, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2},
      {0, -4, -4, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2},
      {0, 0, 0, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0},
      {0, 0, 0, -2, -2, 2, 2, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1},
      {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0, 0, -1, 1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1},
      {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1}};
};

// Inverse directions: both lattices pair d with d+1 for odd d
// (detail::compute_inverses, Lattice.h:49-69).
__host__ __device__ constexpr int inv_dir(int d) {
  return d == 0 ? 0 : ((d & 1) ? d + 1 : d - 1);
}

// Link masks hold direction d in bit d-1, so a pair (d, d+1) sits in bits
// (d-1, d): bit d-1 of pair_swap(x) is bit inv_dir(d)-1 of x.
__host__ __device__ constexpr uint32_t pair_swap(uint32_t x) {
  return ((x & 0x55555555u) << 1) | ((x >> 1) & 0x55555555u);
}

constexpr double kCs2 = 1.0 / 3.0;  // constants.h:41

template <int I, int N, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    sfor<I + 1, N>(f);
  }
}

// acc + C*v for a lattice component C in {-1,0,1}; the zero term is dropped
// (it can only change the sign of an exact zero).
template <int C, typename T>
__device__ __forceinline__ T addc(T acc, T v) {
  if constexpr (C == 0) return acc;
  else if constexpr (C > 0) return acc + v;
  else return acc - v;
}

// (C0*a + C1*b) + C2*c as the reference evaluates the projection of the
// momentum on a lattice vector (Lattice.h:276-285).
template <int C0, int C1, int C2, typename T>
__device__ __forceinline__ T dotc(T a, T b, T c) {
  constexpr bool h0 = C0 != 0, h1 = C1 != 0, h2 = C2 != 0;
  const T s0 = C0 > 0 ? a : -a;
  const T s1 = C1 > 0 ? b : -b;
  const T s2 = C2 > 0 ? c : -c;
  if constexpr (h0 && h1 && h2) return (s0 + s1) + s2;
  else if constexpr (h0 && h1) return s0 + s1;
  else if constexpr (h0 && h2) return s0 + s2;
  else if constexpr (h1 && h2) return s1 + s2;
  else if constexpr (h0) return s0;
  else if constexpr (h1) return s1;
  else if constexpr (h2) return s2;
  else return T(0);
}

// a*b + c.  Binary32 (the reduced-precision forms, which follow no reference
// rounding) fuses it into one FMA; binary64 keeps the reference's separately
// rounded product and sum.
__device__ __forceinline__ float mad(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ double mad(double a, double b, double c) { return a * b + c; }

// 1/x in binary32: the hardware reciprocal (1 ulp).  (Binary64 formulas
// divide.)
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }

// ---------------------------------------------------------------------------
// Storage: binary64 holds f; binary32/binary16 hold f - w_i*rho0.
// ---------------------------------------------------------------------------
template <typename S> struct Store;
template <> struct Store<double> {
  using T = double;
  static __device__ __forceinline__ double ld(const double* p) { return *p; }
  static __device__ __forceinline__ void st(double* p, double v) { *p = v; }
};
template <> struct Store<float> {
  using T = float;
  static __device__ __forceinline__ float ld(const float* p) { return *p; }
  static __device__ __forceinline__ void st(float* p, float v) { *p = v; }
};
template <> struct Store<_Float16> {
  using T = float;
  static __device__ __forceinline__ float ld(const _Float16* p) { return (float)*p; }
  static __device__ __forceinline__ void st(_Float16* p, float v) { *p = (_Float16)v; }
};

template <typename S> constexpr bool kShifted = !std::is_same<S, double>::value;

// ---------------------------------------------------------------------------
// Moments and equilibrium
// ---------------------------------------------------------------------------
// Lattice::CalculateDensityAndMomentum, SSE3 order (Lattice.h:119-178):
// lane 0 from 0 over even directions, lane 1 from the last (odd Q) direction
// over odd directions; result lane0 + lane1.  For shifted storage the same
// sums give rho - rho0 and the momentum.
template <int Q, typename T>
__device__ __forceinline__ void moments(const T (&f)[Q], T& rho, T& mx, T& my,
                                        T& mz) {
  static_assert(Q & 1, "odd lattices only");
  using L = Lat<Q>;
  T r0 = T(0), r1 = f[Q - 1];
  T x0 = T(0), y0 = T(0), z0 = T(0);
  T x1 = addc<L::c[Q - 1][0]>(T(0), f[Q - 1]);
  T y1 = addc<L::c[Q - 1][1]>(T(0), f[Q - 1]);
  T z1 = addc<L::c[Q - 1][2]>(T(0), f[Q - 1]);
  sfor<0, (Q - 1) / 2>([&](auto p) {
    constexpr int d = 2 * decltype(p)::value;
    r0 = r0 + f[d];
    r1 = r1 + f[d + 1];
    x0 = addc<L::c[d][0]>(x0, f[d]);
    x1 = addc<L::c[d + 1][0]>(x1, f[d + 1]);
    y0 = addc<L::c[d][1]>(y0, f[d]);
    y1 = addc<L::c[d + 1][1]>(y1, f[d + 1]);
    z0 = addc<L::c[d][2]>(z0, f[d]);
    z1 = addc<L::c[d + 1][2]>(z1, f[d + 1]);
  });
  rho = r0 + r1;
  mx = x0 + x1;
  my = y0 + y1;
  mz = z0 + z1;
}

// Common factors of Lattice::CalculateFeq, SSE3 variant (Lattice.h:236-270).
template <typename T>
struct FeqCoef {
  T t1, nine, mx, my, mz;
  __device__ __forceinline__ FeqCoef(T rho, T mx_, T my_, T mz_)
      : mx(mx_), my(my_), mz(mz_) {
    const T three_halves = T(3. / 2.) * ((mx * mx + my * my) + mz * mz);
    t1 = rho - three_halves / rho;
    const T rho_1 = T(1.) / rho;
    nine = T(9. / 2.);
    nine *= rho_1;
  }
  // f_eq[i] = w_i * ((t1 + nine*(e*e)) + 3*e)  (Lattice.h:271-320)
  template <int Q, int I>
  __device__ __forceinline__ T eq() const {
    using L = Lat<Q>;
    const T e = dotc<L::c[I][0], L::c[I][1], L::c[I][2]>(mx, my, mz);
    return T(L::w[I]) * ((t1 + nine * (e * e)) + T(3.) * e);
  }
};

// Shifted equilibrium feq_i - w_i*rho0 for reduced-precision storage, with
// drho = rho - rho0 carried separately (no cancellation against rho0).
template <typename T>
struct FeqShifted {
  T drho, q, nine, mx, my, mz;
  __device__ __forceinline__ FeqShifted(T drho_, T rho, T mx_, T my_, T mz_)
      : drho(drho_), mx(mx_), my(my_), mz(mz_) {
    const T inv = recip(rho);
    q = mad(T(-1.5) * mad(mz, mz, mad(my, my, mx * mx)), inv, drho);
    nine = T(4.5) * inv;
  }
  template <int Q, int I>
  __device__ __forceinline__ T eq() const {
    using L = Lat<Q>;
    const T e = dotc<L::c[I][0], L::c[I][1], L::c[I][2]>(mx, my, mz);
    return T(L::w[I]) * mad(T(3.), e, mad(nine, e * e, q));
  }
};

// ---------------------------------------------------------------------------
// Collision
// ---------------------------------------------------------------------------
template <int Q, typename T>
struct Hydro {
  T rho, mx, my, mz;  // rho is the full density (rho0 added back if shifted)
};

// d'Humieres relaxation rates (DHumieresD3Q15MRTBasis.cc:10-26,
// DHumieresD3Q19MRTBasis.cc:11-36); 0 marks the moments relaxed with 1/tau.
__host__ __device__ constexpr double mrt_rate(int Q, int k) {
  constexpr double r15[11] = {1.6, 1.2, 1.6, 1.6, 1.6, 0, 0, 0, 0, 0, 1.2};
  constexpr double r19[15] = {1.19, 1.4, 1.2, 1.2, 1.2, 0, 1.4, 0,
                              1.4,  0,   0,   0,   1.98, 1.98, 1.98};
  return Q == 15 ? r15[k] : r19[k];
}
// BASIS_TIMES_BASIS_TRANSPOSED (basis_helpers.h:14-26), exact in binary64.
template <int Q>
__host__ __device__ constexpr double mrt_diag(int k) {
  double acc = 0.0;
  for (int i = 0; i < Q; ++i) acc = acc + (double)Lat<Q>::M[k][i] * (double)Lat<Q>::M[k][i];
  return acc;
}

// Parity of MRT moment k under inversion of the directions, which come in
// opposite pairs (d, d+1), d odd: 1 when row k of the basis takes the same
// value on both directions of every pair, -1 when opposite values (and 0 on
// the rest direction), 0 when neither.
template <int Q>
__host__ __device__ constexpr int mrt_parity(int k) {
  bool even = true, odd = Lat<Q>::M[k][0] == 0;
  for (int d = 1; d < Q; d += 2) {
    if (Lat<Q>::M[k][d] != Lat<Q>::M[k][d + 1]) even = false;
    if (Lat<Q>::M[k][d] != -Lat<Q>::M[k][d + 1]) odd = false;
  }
  return even ? 1 : (odd ? -1 : 0);
}
template <int Q>
__host__ __device__ constexpr bool mrt_rows_have_parity() {
  for (int k = 0; k < Lat<Q>::nmom; ++k)
    if (mrt_parity<Q>(k) == 0) return false;
  return true;
}

// Sums over the binary32-rounded weights (exact, in binary64): sum_i w_i and
// sum_i w_i c_ix^2 (the same for y and z on both lattices).
template <int Q>
__host__ __device__ constexpr double rounded_weight_sum() {
  double s = 0.0;
  for (int i = 0; i < Q; ++i) s += (double)(float)Lat<Q>::w[i];
  return s;
}
template <int Q>
__host__ __device__ constexpr double rounded_weight_c2() {
  double s = 0.0;
  for (int i = 0; i < Q; ++i) s += (double)(float)Lat<Q>::w[i] * Lat<Q>::c[i][0] * Lat<Q>::c[i][0];
  return s;
}
// sum_i fl(3 w_i) c_ix^2 with the binary32 coefficients of the linear
// equilibrium term
template <int Q>
__host__ __device__ constexpr double rounded_3w_c2() {
  double s = 0.0;
  for (int i = 0; i < Q; ++i)
    s += (double)(float)(3.0 * Lat<Q>::w[i]) * Lat<Q>::c[i][0] * Lat<Q>::c[i][0];
  return s;
}
template <int Q>
__host__ __device__ constexpr bool pairs_opposite() {
  for (int i = 1; i < Q; i += 2)
    for (int k = 0; k < 3; ++k)
      if (Lat<Q>::c[i + 1][k] != -Lat<Q>::c[i][k]) return false;
  return true;
}

// LBGK of the shifted storage forms in binary32 (these follow no reference
// rounding).  Directions come in opposite pairs (d, d+1), so the moments are
// sums of pair sums and differences, and a pair's equilibria are A +- B with
// A = w (q + 9/2 e^2 / rho) and B = 3 w e, e = c_d . m.
// Conservation: with binary32 coefficients, sum_i w_i = W, sum_i w_i c_i c_i
// = G I and sum_i fl(3 w_i) c_i c_i = K I differ from 1, 1/3 and 1 by 1e-8 to
// 3e-8, so the equilibrium's mass and momentum are W q + (9/2) G m^2/rho and
// K m instead of drho and m.  Relaxing towards it would act as a body force
// of ~1e-8 m per step (a 2e-5 to 3e-5 velocity bias on the vessel tree).
// Two corrections of f_neq before the relaxation, both applied to f_neq,
// whose magnitude keeps them above the rounding of binary32:
// * EXACT (binary32 storage): the mass and momentum of the computed f_neq
//   are removed, n_d -= w_d (s0 + 3 c_d . s) with s0, s its sums; this also
//   removes the rounding noise of each collision (velocity 1e-7 from
//   binary64 on the tree after 10000 steps);
// * otherwise (binary16 storage, whose own rounding is 1e-3): only the
//   systematic amounts above, the mass one on the rest direction and the
//   momentum one split over the axis pairs (1e-6 in binary32 arithmetic),
//   at a fraction of the projection's instructions.
template <int Q, typename T, bool EXACT>
__device__ __forceinline__ Hydro<Q, T> collide_lbgk_shifted(T (&f)[Q], const StepArgs& a) {
  using L = Lat<Q>;
  static_assert(pairs_opposite<Q>(), "directions d and d+1 must be opposite");
  constexpr double W = rounded_weight_sum<Q>(), G = rounded_weight_c2<Q>();
  constexpr double K = rounded_3w_c2<Q>();
  T r = f[0], mx = T(0), my = T(0), mz = T(0);
  sfor<0, (Q - 1) / 2>([&](auto p) {
    constexpr int i = 2 * decltype(p)::value + 1;
    const T dd = f[i] - f[i + 1];
    r = r + (f[i] + f[i + 1]);
    mx = addc<L::c[i][0]>(mx, dd);
    my = addc<L::c[i][1]>(my, dd);
    mz = addc<L::c[i][2]>(mz, dd);
  });
  Hydro<Q, T> h;
  h.rho = T(a.rho0) + r;
  h.mx = mx;
  h.my = my;
  h.mz = mz;
  const T inv = recip(h.rho);
  const T m2 = mad(mz, mz, mad(my, my, mx * mx));
  const T q = mad(T(-1.5) * m2, inv, r);
  const T nine = T(4.5) * inv;
  const T om = T(a.omega_f);
  T n[Q];  // f_neq
  n[0] = f[0] - T(L::w[0]) * q;
  sfor<0, (Q - 1) / 2>([&](auto p) {
    constexpr int i = 2 * decltype(p)::value + 1;
    const T e = dotc<L::c[i][0], L::c[i][1], L::c[i][2]>(mx, my, mz);
    const T A = mad(T(L::w[i]) * nine, e * e, T(L::w[i]) * q);
    const T B = T(3.0 * L::w[i]) * e;
    n[i] = (f[i] - A) - B;
    n[i + 1] = (f[i + 1] - A) + B;
  });
  if constexpr (EXACT) {
    T s0 = n[0], sx = T(0), sy = T(0), sz = T(0);
    sfor<0, (Q - 1) / 2>([&](auto p) {
      constexpr int i = 2 * decltype(p)::value + 1;
      const T dd = n[i] - n[i + 1];
      s0 = s0 + (n[i] + n[i + 1]);
      sx = addc<L::c[i][0]>(sx, dd);
      sy = addc<L::c[i][1]>(sy, dd);
      sz = addc<L::c[i][2]>(sz, dd);
    });
    sx = T(3) * sx;
    sy = T(3) * sy;
    sz = T(3) * sz;
    n[0] = mad(T(-L::w[0]), s0, n[0]);
    sfor<0, (Q - 1) / 2>([&](auto p) {
      constexpr int i = 2 * decltype(p)::value + 1;
      const T e = dotc<L::c[i][0], L::c[i][1], L::c[i][2]>(sx, sy, sz);
      n[i] = mad(T(-L::w[i]), s0 + e, n[i]);
      n[i + 1] = mad(T(-L::w[i + 1]), s0 - e, n[i + 1]);
    });
  } else {
    // systematic mass of f_neq: drho - (W q + (9/2) G m^2/rho)
    n[0] = n[0] - mad(T(1.0 - W), r, T(1.5 * W - 4.5 * G) * (m2 * inv));
    // systematic momentum of f_neq: (1 - K) m, half on each axis direction
    constexpr double kHalf = 0.5 * (1.0 - K);
    sfor<0, (Q - 1) / 2>([&](auto p) {
      constexpr int i = 2 * decltype(p)::value + 1;
      constexpr int cx = L::c[i][0], cy = L::c[i][1], cz = L::c[i][2];
      if constexpr (cx * cx + cy * cy + cz * cz == 1) {
        const T k = T(kHalf) * (cx != 0 ? mx : cy != 0 ? my : mz) * T(cx + cy + cz);
        n[i] = n[i] - k;
        n[i + 1] = n[i + 1] + k;
      }
    });
  }
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    f[d] = mad(om, n[d], f[d]);
  });
  return h;
}

// MRT relaxation of the shifted storage forms in binary32 (these follow no
// reference rounding): the same m_neq = M f_neq and f* = f - sum_k (S_k N_kd)
// m_neq_k as the binary64 form, with the rows of the d'Humieres basis
// grouped by their parity under inversion.  An even row needs the rest value
// and the pair sums of f_neq, an odd row the pair differences; the
// relaxation of pair (d, d+1) is E + O on d and E - O on d+1, with E the
// even moments' part and O the odd moments'.  About half the operations of
// the row-by-column form.
template <int Q, typename T>
__device__ __forceinline__ void mrt_relax_pairs(const T (&fneq)[Q], const StepArgs& a,
                                                T (&f)[Q]) {
  using L = Lat<Q>;
  static_assert(pairs_opposite<Q>() && mrt_rows_have_parity<Q>(),
                "MRT pair form: opposite direction pairs and rows of definite parity");
  constexpr int NM = L::nmom;
  T sg[Q], df[Q];  // pair sums and differences, at the pair's first direction
  sfor<1, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    if constexpr (d % 2 == 1) {
      sg[d] = fneq[d] + fneq[d + 1];
      df[d] = fneq[d] - fneq[d + 1];
    }
  });
  T mneq[NM];
  sfor<0, NM>([&](auto ki) {
    constexpr int k = decltype(ki)::value;
    constexpr bool even = mrt_parity<Q>(k) == 1;
    T acc = T(0.);
    if constexpr (even && L::M[k][0] != 0) acc = T(L::M[k][0]) * fneq[0];
    sfor<1, Q>([&](auto di) {
      constexpr int d = decltype(di)::value;
      constexpr int m = L::M[k][d];
      if constexpr (d % 2 == 1 && m != 0) {
        const T v = even ? sg[d] : df[d];
        if constexpr (m == 1) acc = acc + v;
        else if constexpr (m == -1) acc = acc - v;
        else acc = mad(T(m), v, acc);
      }
    });
    mneq[k] = acc;
  });
  // S_k N_kd as in the binary64 form (the 1/tau moments from the runtime rate)
  const double s_tau = a.mrt_s_tau;
  const auto coef = [&](auto ki, auto di) {
    constexpr int k = decltype(ki)::value, d = decltype(di)::value;
    constexpr double N = (double)L::M[k][d] / mrt_diag<Q>(k);
    constexpr double rate = mrt_rate(Q, k);
    if constexpr (rate != 0.0) {
      constexpr double cc = rate * N;
      return T(cc);
    } else {
      return T(s_tau * N);
    }
  };
  {
    T e0 = T(0.);
    sfor<0, NM>([&](auto ki) {
      constexpr int k = decltype(ki)::value;
      if constexpr (mrt_parity<Q>(k) == 1 && L::M[k][0] != 0)
        e0 = mad(coef(ki, std::integral_constant<int, 0>{}), mneq[k], e0);
    });
    f[0] = f[0] - e0;
  }
  sfor<1, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    if constexpr (d % 2 == 1) {
      T ev = T(0.), od = T(0.);
      sfor<0, NM>([&](auto ki) {
        constexpr int k = decltype(ki)::value;
        if constexpr (L::M[k][d] != 0) {
          if constexpr (mrt_parity<Q>(k) == 1) ev = mad(coef(ki, di), mneq[k], ev);
          else od = mad(coef(ki, di), mneq[k], od);
        }
      });
      f[d] = f[d] - (ev + od);
      f[d + 1] = f[d + 1] - (ev - od);
    }
  });
}

// Collide in place: f (post-stream, possibly shifted) -> f* ; returns the
// moments the streamers of the next step need.  Binary64 LBGK and MRT in
// the reference's evaluation order, MRT of the shifted forms below; LBGK of
// the shifted forms in collide_lbgk_shifted.
template <int Q, int KERNEL, typename S, typename T = typename Store<S>::T>
__device__ __forceinline__ Hydro<Q, T> collide(T (&f)[Q], const StepArgs& a) {
  constexpr bool SHIFT = kShifted<S>;
  if constexpr (SHIFT && KERNEL == 0) {
    return collide_lbgk_shifted<Q, T, std::is_same<S, float>::value>(f, a);
  } else {
    using L = Lat<Q>;
    Hydro<Q, T> h;
    T r;
    moments<Q, T>(f, r, h.mx, h.my, h.mz);
    T fneq[Q];
    if constexpr (SHIFT) {
      const T rho = T(a.rho0) + r;
      h.rho = rho;
      FeqShifted<T> fe(r, rho, h.mx, h.my, h.mz);
      sfor<0, Q>([&](auto di) {
        constexpr int d = decltype(di)::value;
        fneq[d] = f[d] - fe.template eq<Q, d>();
      });
    } else {
      h.rho = r;
      FeqCoef<T> fe(r, h.mx, h.my, h.mz);
      sfor<0, Q>([&](auto di) {
        constexpr int d = decltype(di)::value;
        fneq[d] = f[d] - fe.template eq<Q, d>();
      });
    }
    if constexpr (KERNEL == 0) {
      // LBGK::Collide: f + f_neq * omega (LBGK.h:167-174)
      const T omega = std::is_same<T, double>::value ? T(a.omega) : T(a.omega_f);
      sfor<0, Q>([&](auto di) {
        constexpr int d = decltype(di)::value;
        f[d] = mad(fneq[d], omega, f[d]);
      });
    } else if constexpr (SHIFT) {
      mrt_relax_pairs<Q, T>(fneq, a, f);
    } else {
      // MRT: m_neq = M f_neq (MRT.h:123-135); f* = f - sum_k (S_k N_kd) m_neq_k
      // with N = M / diag(M M^T) (MRT.h:139-152).  S_k N_kd is a compile-time
      // binary64 product except for the 1/tau moments, where it is formed from
      // the runtime 1/tau exactly as the reference forms it.
      constexpr int NM = L::nmom;
      T mneq[NM];
      sfor<0, NM>([&](auto ki) {
        constexpr int k = decltype(ki)::value;
        T acc = T(0.);
        sfor<0, Q>([&](auto ii) {
          constexpr int i = decltype(ii)::value;
          constexpr int m = L::M[k][i];
          if constexpr (m == 1) acc = acc + fneq[i];
          else if constexpr (m == -1) acc = acc - fneq[i];
          else if constexpr (m != 0) acc = acc + T(m) * fneq[i];
        });
        mneq[k] = acc;
      });
      const double s_tau = a.mrt_s_tau;
      sfor<0, Q>([&](auto di) {
        constexpr int d = decltype(di)::value;
        T coll = T(0.);
        sfor<0, NM>([&](auto ki) {
          constexpr int k = decltype(ki)::value;
          if constexpr (L::M[k][d] != 0) {
            constexpr double N = (double)L::M[k][d] / mrt_diag<Q>(k);
            constexpr double rate = mrt_rate(Q, k);
            double c;
            if constexpr (rate != 0.0) {
              constexpr double cc = rate * N;
              c = cc;
            } else {
              c = s_tau * N;
            }
            coll = mad(T(c), mneq[k], coll);
          }
        });
        f[d] = f[d] - coll;
      });
    }
    return h;
  }
}

// ---------------------------------------------------------------------------
// Layouts: where a thread's site lives and where its neighbours are.
// ---------------------------------------------------------------------------
struct SiteRef {
  int64_t x;      // index into every per-direction array
  uint32_t cut;   // wall|iolet link mask of the site (0: bulk site)
  uint32_t rec;   // boundary record (valid when cut != 0)
  uint32_t w;     // block layout: wave index (block * 8 + x plane)
  uint32_t lane;  // block layout: lane = y * 8 + z inside the plane
  int32_t tb[27]; // block layout: neighbour blocks of this wave's block
};

__device__ __forceinline__ uint32_t logical_workgroup() {
  // XCD-aware remap: hardware dispatches workgroup b to XCD b % 8; give
  // every XCD one contiguous range of logical workgroups.
  const uint32_t nwg = gridDim.x, b = blockIdx.x;
  const uint32_t xcd = b & 7u, k = b >> 3;
  const uint32_t base = nwg >> 3, rem = nwg & 7u;
  const uint32_t start = xcd * base + (xcd < rem ? xcd : rem);
  return start + k;
}

// Logical workgroup of the step kernels.  remap 0: the hardware order
// (consecutive workgroups on different XCDs); 1: one contiguous range per
// XCD; run >= 2: XCD x takes runs of `run` consecutive logical workgroups
// and the eight XCDs work on eight adjacent runs at a time, so the planes
// of a block and of its Morton neighbours stay in one XCD's L2 while the
// whole chip sweeps one region.  The grid is a multiple of 8 * run (the
// launcher pads it), which makes the map a bijection.
__device__ __forceinline__ uint32_t step_workgroup(uint32_t remap) {
  if (remap == 0) return blockIdx.x;
  if (remap == 1) return logical_workgroup();
  const uint32_t b = blockIdx.x, x = b & 7u, j = b >> 3;
  const uint32_t r = j / remap, o = j - r * remap;
  return (r * 8u + x) * remap + o;
}

// Compact layout: sites packed in Morton-block order.  The 32-bit pull
// table holds, per direction e, the offset (inside array e) of the value
// that streams into slot e: the source site for a fluid link, or the site's
// own slot of the inverse direction (x + (inv(e) - e) * np) for a cut link.
struct SparseLayout {
  static __device__ __forceinline__ bool locate(const DevDomain& D, SiteRef& r) {
    const int64_t x = (int64_t)logical_workgroup() * blockDim.x + threadIdx.x;
    if (x >= D.n) return false;
    at(D, x, r);
    return true;
  }
  // Compact site x (x < D.n).
  static __device__ __forceinline__ void at(const DevDomain& D, int64_t x, SiteRef& r) {
    r.x = x;
    const uint32_t bi = D.binfo[x];
    r.cut = bi != 0 ? 1u : 0u;  // only "is boundary" is needed here
    r.rec = bi - 1;
  }
  // State index of direction d of slot x, and of the value streaming into
  // slot E of the lane's site.
  template <int Q>
  static __device__ __forceinline__ int64_t at(const DevDomain& D, int d, int64_t x) {
    return (int64_t)d * D.np + x;
  }
  template <int Q, int E, bool M>
  static __device__ __forceinline__ int64_t index(const DevDomain& D, const SiteRef& r) {
    return (int64_t)E * D.np + (int64_t)D.src[(int64_t)(E - 1) * D.np + r.x];
  }
  static __device__ __forceinline__ int64_t compact(const DevDomain&, const SiteRef& r) {
    return r.x;
  }
  static __device__ __forceinline__ void relax(const DevDomain&, const SiteRef&, StepArgs&) {}
};

// In-plane order of the block layouts (the lane of a site inside its x
// plane, slot & 63):
// * binary64 storage: the bits of (y, z) inside the block interleaved, z in
//   bits 0, 2, 4 and y in bits 1, 3, 5 (a 2-D Morton order).  A 32-byte
//   sector then holds a 2 x 2 tile of the plane, so the column of a
//   neighbour block that a pull across a z face reads touches 4 sectors
//   instead of 8; the step reads 4 % fewer bytes and runs 2 % faster.
// * binary32 and binary16: rows of z (y in bits 3-5, z in bits 0-2).  The
//   interleaved order saved them no bytes, and binary16 ran 6 % slower with
//   it (its shifted gathers touch more distinct lines per quarter wave).
template <typename S> constexpr bool kMortonLanes = std::is_same<S, double>::value;
template <bool M> struct LaneBits {
  static constexpr uint32_t Y = M ? 0x2Au : 0x38u, Z = M ? 0x15u : 0x07u;
  static constexpr uint32_t UY = M ? 2u : 8u;  // lowest bit of Y
};
// Lane of (y - CY, z - CZ) mod 8: the decrement or increment of one
// coordinate's bits, with the carry or borrow run through the other's.
template <int CY, int CZ, bool M>
__device__ __forceinline__ uint32_t lane_shift(uint32_t lane) {
  using B = LaneBits<M>;
  uint32_t y = lane & B::Y, z = lane & B::Z;
  if constexpr (CY > 0) y = (y - B::UY) & B::Y;
  if constexpr (CY < 0) y = ((y | B::Z) + B::UY) & B::Y;
  if constexpr (CZ > 0) z = (z - 1u) & B::Z;
  if constexpr (CZ < 0) z = ((z | B::Y) + 1u) & B::Z;
  return y | z;
}
// Whether x - c leaves the block across its y (z) face: y == 0 for CY > 0,
// y == 7 for CY < 0.
template <int CY, bool M>
__device__ __forceinline__ bool lane_cross_y(uint32_t lane) {
  if constexpr (CY == 0) return false;
  else return (lane & LaneBits<M>::Y) == (CY > 0 ? 0u : LaneBits<M>::Y);
}
template <int CZ, bool M>
__device__ __forceinline__ bool lane_cross_z(uint32_t lane) {
  if constexpr (CZ == 0) return false;
  else return (lane & LaneBits<M>::Z) == (CZ > 0 ? 0u : LaneBits<M>::Z);
}

// Block layout: every 8^3 block holding fluid owns 512 slots (x-plane major,
// lanes in the storage's in-plane order above) and its Q directions are stored together, block after block
// ([block][direction][512 slots]), so that one block's distributions are
// one contiguous run of memory; a wave covers one x plane of one block.  Neighbours come from
// slot arithmetic and the block's 27-entry neighbour table, which each wave
// loads once; per-wave 64-bit masks say which lanes are fluid and which carry
// a boundary record, whose wall and iolet masks mark the cut links.
struct BlockLayout {
  static __device__ __forceinline__ bool locate(const DevDomain& D, SiteRef& r) {
    const int64_t slot = (int64_t)logical_workgroup() * blockDim.x + threadIdx.x;
    return locate_wave(D, __builtin_amdgcn_readfirstlane((uint32_t)(slot >> 6)), r);
  }
  // The lane's site in wave (x plane) w; false for lanes without a fluid
  // site and for w beyond the domain.
  static __device__ __forceinline__ bool locate_wave(const DevDomain& D, uint32_t w,
                                                     SiteRef& r) {
    r.w = w;
    r.lane = threadIdx.x & 63u;
    r.x = ((int64_t)r.w << 6) + r.lane;
    if (r.w >= D.nwaves) return false;
    const uint64_t fm = D.wave_fluid[r.w];
    const uint64_t bm = D.wave_bnd[r.w];
    const uint32_t rb = D.wave_rec[r.w];
    if (!((fm >> r.lane) & 1ull)) return false;
    const int32_t* nbt = D.bnbr + (int64_t)(r.w >> 3) * 27;
    sfor<0, 27>([&](auto ii) {
      constexpr int i = decltype(ii)::value;
      r.tb[i] = nbt[i];
    });
    // A lane's cut-link mask comes from its boundary record (the gather
    // reads the same two words again for the wall and iolet formulas).
    r.rec = rb + (uint32_t)__popcll(bm & ((1ull << r.lane) - 1ull));
    r.cut = ((bm >> r.lane) & 1ull) ? (D.b_wall[r.rec] | D.b_iolet[r.rec]) : 0u;
    return true;
  }
  template <int Q>
  static __device__ __forceinline__ int64_t at(const DevDomain&, int d, int64_t slot) {
    return ((slot >> 9) * Q + d) * 512 + (slot & 511);
  }
  template <int Q, int E, bool M>
  static __device__ __forceinline__ int64_t index(const DevDomain& D, const SiteRef& r) {
    constexpr int CX = Lat<Q>::c[E][0], CY = Lat<Q>::c[E][1], CZ = Lat<Q>::c[E][2];
    const int lx = (int)(r.w & 7u);
    // source = x - c_E; crossing a block face moves to the neighbour block
    // at offset -c_E along that axis.
    const bool bx = CX != 0 && (CX > 0 ? lx == 0 : lx == 7);  // wave-uniform
    const bool by = lane_cross_y<CY, M>(r.lane);
    const bool bz = lane_cross_z<CZ, M>(r.lane);
    constexpr int i000 = 13, i100 = 13 - 9 * CX, i010 = 13 - 3 * CY, i001 = 13 - CZ;
    constexpr int i110 = 13 - 9 * CX - 3 * CY, i101 = 13 - 9 * CX - CZ;
    constexpr int i011 = 13 - 3 * CY - CZ, i111 = 13 - 9 * CX - 3 * CY - CZ;
    const int32_t n00 = bx ? r.tb[i100] : r.tb[i000];
    const int32_t n10 = bx ? r.tb[i110] : r.tb[i010];
    const int32_t n01 = bx ? r.tb[i101] : r.tb[i001];
    const int32_t n11 = bx ? r.tb[i111] : r.tb[i011];
    const int32_t nb = by ? (bz ? n11 : n10) : (bz ? n01 : n00);
    const int64_t src = ((int64_t)nb * Q + E) * 512 +
                        ((((lx - CX) & 7) << 6) | lane_shift<CY, CZ, M>(r.lane));
    constexpr int DP = inv_dir(E);
    const int64_t own = at<Q>(D, DP, r.x);
    return ((r.cut >> (DP - 1)) & 1u) ? own : src;
  }
  static __device__ __forceinline__ int64_t compact(const DevDomain& D, const SiteRef& r) {
    const uint64_t fm = D.wave_fluid[r.w];
    return (int64_t)D.wave_cbase[r.w] + __popcll(fm & ((1ull << r.lane) - 1ull));
  }
  // Ensembles relax each block with its instance's rates (one scalar load
  // per wave); otherwise the StepArgs values hold everywhere.
  static __device__ __forceinline__ void relax(const DevDomain& D, const SiteRef& r,
                                               StepArgs& a) {
    if (D.blk_omega) {
      const uint32_t b = r.w >> 3;
      a.omega = D.blk_omega[b];
      a.omega_f = D.blk_omega_f[b];
      a.mrt_s_tau = D.blk_s_tau[b];
    }
  }
};

// Compact-block layout: the waves, masks and neighbour-block table of the
// block layout, but the distributions are stored densely by compact site
// (Morton block order, then planes, then the in-plane order), so that no cache line carries empty
// slots.  The site of lane s of plane w is wave_cbase[w] + popcount of the
// plane's fluid mask below s; a neighbour's index comes from the same count
// on the source plane, whose mask and base are wave-uniform scalar loads.
struct CompactBlockLayout {
  static __device__ __forceinline__ int64_t site_of(uint64_t mask, uint32_t base,
                                                    uint32_t lane) {
    return (int64_t)base + __popcll(mask & ((1ull << lane) - 1ull));
  }
  static __device__ __forceinline__ bool locate(const DevDomain& D, SiteRef& r) {
    const int64_t slot = (int64_t)logical_workgroup() * blockDim.x + threadIdx.x;
    return locate_wave(D, __builtin_amdgcn_readfirstlane((uint32_t)(slot >> 6)), r);
  }
  static __device__ __forceinline__ bool locate_wave(const DevDomain& D, uint32_t w,
                                                     SiteRef& r) {
    r.w = w;
    r.lane = threadIdx.x & 63u;
    if (r.w >= D.nwaves) return false;
    const uint64_t fm = D.wave_fluid[r.w];
    const uint64_t bm = D.wave_bnd[r.w];
    const uint32_t rb = D.wave_rec[r.w];
    if (!((fm >> r.lane) & 1ull)) return false;
    r.x = site_of(fm, D.wave_cbase[r.w], r.lane);
    const int32_t* nbt = D.bnbr + (int64_t)(r.w >> 3) * 27;
    sfor<0, 27>([&](auto ii) {
      constexpr int i = decltype(ii)::value;
      r.tb[i] = nbt[i];
    });
    r.rec = rb + (uint32_t)__popcll(bm & ((1ull << r.lane) - 1ull));
    r.cut = ((bm >> r.lane) & 1ull) ? (D.b_wall[r.rec] | D.b_iolet[r.rec]) : 0u;
    return true;
  }
  // Fluid mask and compact base of x plane px of block nb (nb < 0: absent,
  // and then no lane reads it).
  static __device__ __forceinline__ void plane(const DevDomain& D, int32_t nb, int px,
                                               uint64_t& m, uint32_t& b) {
    if (nb >= 0) {
      const uint32_t pw = ((uint32_t)nb << 3) | (uint32_t)px;
      m = D.wave_fluid[pw];
      b = D.wave_cbase[pw];
    } else {
      m = 0;
      b = 0;
    }
  }
  template <int Q>
  static __device__ __forceinline__ int64_t at(const DevDomain& D, int d, int64_t x) {
    return (int64_t)d * D.np + x;
  }
  template <int Q, int E, bool M>
  static __device__ __forceinline__ int64_t index(const DevDomain& D, const SiteRef& r) {
    constexpr int CX = Lat<Q>::c[E][0], CY = Lat<Q>::c[E][1], CZ = Lat<Q>::c[E][2];
    const int lx = (int)(r.w & 7u);
    const bool bx = CX != 0 && (CX > 0 ? lx == 0 : lx == 7);  // wave-uniform
    const bool by = lane_cross_y<CY, M>(r.lane);
    const bool bz = lane_cross_z<CZ, M>(r.lane);
    constexpr int i000 = 13, i100 = 13 - 9 * CX, i010 = 13 - 3 * CY, i001 = 13 - CZ;
    constexpr int i110 = 13 - 9 * CX - 3 * CY, i101 = 13 - 9 * CX - CZ;
    constexpr int i011 = 13 - 3 * CY - CZ, i111 = 13 - 9 * CX - 3 * CY - CZ;
    const int px = (lx - CX) & 7;
    uint64_t m00, m10 = 0, m01 = 0, m11 = 0;
    uint32_t b00, b10 = 0, b01 = 0, b11 = 0;
    plane(D, bx ? r.tb[i100] : r.tb[i000], px, m00, b00);
    if constexpr (CY != 0) plane(D, bx ? r.tb[i110] : r.tb[i010], px, m10, b10);
    if constexpr (CZ != 0) plane(D, bx ? r.tb[i101] : r.tb[i001], px, m01, b01);
    if constexpr (CY != 0 && CZ != 0) plane(D, bx ? r.tb[i111] : r.tb[i011], px, m11, b11);
    const uint64_t m = by ? (bz ? m11 : m10) : (bz ? m01 : m00);
    const uint32_t b = by ? (bz ? b11 : b10) : (bz ? b01 : b00);
    const uint32_t sl = lane_shift<CY, CZ, M>(r.lane);
    const int64_t src = site_of(m, b, sl);
    constexpr int DP = inv_dir(E);
    return ((r.cut >> (DP - 1)) & 1u) ? at<Q>(D, DP, r.x) : at<Q>(D, E, src);
  }
  static __device__ __forceinline__ int64_t compact(const DevDomain&, const SiteRef& r) {
    return r.x;
  }
  static __device__ __forceinline__ void relax(const DevDomain& D, const SiteRef& r,
                                               StepArgs& a) {
    BlockLayout::relax(D, r, a);
  }
};

// The block layout addressed by 32-bit byte offsets from the state buffer's
// base, for states of every direction below 4 GiB.  A load or store is then
// the kernel argument's 64-bit base (a scalar register pair) plus one 32-bit
// lane offset, which the global instructions add themselves: the wave-uniform
// parts (block, direction, plane, and the neighbour blocks' offsets) are
// scalar arithmetic, and a gather costs a select and an add per lane instead
// of 64-bit multiply-adds.  Same elements, same order as BlockLayout.
struct BlockLayout32 : BlockLayout {
  // byte offset of slot `lane` of direction d in this wave's plane
  template <int Q, typename P>
  static __device__ __forceinline__ uint32_t own_off(const SiteRef& r, uint32_t d) {
    constexpr uint32_t ES = sizeof(P);
    return (((r.w >> 3) * (uint32_t)Q + d) * 512u + ((r.w & 7u) << 6)) * ES + r.lane * ES;
  }
  template <int Q, typename P>
  static __device__ __forceinline__ P* own_ptr(const SiteRef& r, P* base, int d) {
    constexpr uint32_t ES = sizeof(P);
    // wave-uniform part first (scalar), then the lane's offset
    const uint32_t uni = (((r.w >> 3) * (uint32_t)Q + (uint32_t)d) * 512u + ((r.w & 7u) << 6)) * ES;
    using C = std::conditional_t<std::is_const<P>::value, const char, char>;
    return (P*)(((C*)base + uni) + r.lane * ES);
  }
  template <int Q, int E, typename P>
  static __device__ __forceinline__ P* src_ptr(const SiteRef& r, P* base) {
    constexpr int CX = Lat<Q>::c[E][0], CY = Lat<Q>::c[E][1], CZ = Lat<Q>::c[E][2];
    constexpr uint32_t ES = sizeof(P), BS = (uint32_t)Q * 512u * ES;
    constexpr bool M = kMortonLanes<std::remove_const_t<P>>;
    const int lx = (int)(r.w & 7u);
    const bool bx = CX != 0 && (CX > 0 ? lx == 0 : lx == 7);  // wave-uniform
    const bool by = lane_cross_y<CY, M>(r.lane);
    const bool bz = lane_cross_z<CZ, M>(r.lane);
    constexpr int i000 = 13, i100 = 13 - 9 * CX, i010 = 13 - 3 * CY, i001 = 13 - CZ;
    constexpr int i110 = 13 - 9 * CX - 3 * CY, i101 = 13 - 9 * CX - CZ;
    constexpr int i011 = 13 - 3 * CY - CZ, i111 = 13 - 9 * CX - 3 * CY - CZ;
    // wave-uniform: direction E's array in each candidate block, source plane
    const uint32_t uni = ((uint32_t)E * 512u + ((uint32_t)((lx - CX) & 7) << 6)) * ES;
    // (readfirstlane keeps these products scalar: without it the compiler
    // sinks the multiply below the per-lane select, into a vector multiply)
    const auto blk = [&](int32_t t) {
      return __builtin_amdgcn_readfirstlane((uint32_t)t * BS + uni);
    };
    const uint32_t n00 = blk(bx ? r.tb[i100] : r.tb[i000]);
    const uint32_t n10 = blk(bx ? r.tb[i110] : r.tb[i010]);
    const uint32_t n01 = blk(bx ? r.tb[i101] : r.tb[i001]);
    const uint32_t n11 = blk(bx ? r.tb[i111] : r.tb[i011]);
    const uint32_t nb = by ? (bz ? n11 : n10) : (bz ? n01 : n00);
    const uint32_t yz = lane_shift<CY, CZ, M>(r.lane) * ES;
    constexpr int DP = inv_dir(E);
    // a cut link reads the site's own slot of the inverse direction
    // (both offsets are formed first so that the choice is a select, not a
    // divergent branch)
    const uint32_t own = own_off<Q, P>(r, DP), src = nb + yz;
    const uint32_t off = ((r.cut >> (DP - 1)) & 1u) ? own : src;
    using C = std::conditional_t<std::is_const<P>::value, const char, char>;
    return (P*)((C*)base + off);
  }
  // The value at x - c_E whether or not link E is cut (the gather then
  // replaces a cut link's value): a neighbour block that does not exist is
  // replaced by the own block, so that the address stays inside the state.
  template <int Q, int E, typename P>
  static __device__ __forceinline__ P* src_any_ptr(const SiteRef& r, P* base) {
    constexpr int CX = Lat<Q>::c[E][0], CY = Lat<Q>::c[E][1], CZ = Lat<Q>::c[E][2];
    constexpr uint32_t ES = sizeof(P), BS = (uint32_t)Q * 512u * ES;
    constexpr bool M = kMortonLanes<std::remove_const_t<P>>;
    const int lx = (int)(r.w & 7u);
    const bool bx = CX != 0 && (CX > 0 ? lx == 0 : lx == 7);  // wave-uniform
    const bool by = lane_cross_y<CY, M>(r.lane);
    const bool bz = lane_cross_z<CZ, M>(r.lane);
    constexpr int i000 = 13, i100 = 13 - 9 * CX, i010 = 13 - 3 * CY, i001 = 13 - CZ;
    constexpr int i110 = 13 - 9 * CX - 3 * CY, i101 = 13 - 9 * CX - CZ;
    constexpr int i011 = 13 - 3 * CY - CZ, i111 = 13 - 9 * CX - 3 * CY - CZ;
    const uint32_t uni = ((uint32_t)E * 512u + ((uint32_t)((lx - CX) & 7) << 6)) * ES;
    const int32_t self = (int32_t)(r.w >> 3);
    const auto blk = [&](int32_t t) {
      return __builtin_amdgcn_readfirstlane((uint32_t)(t < 0 ? self : t) * BS + uni);
    };
    const uint32_t n00 = blk(bx ? r.tb[i100] : r.tb[i000]);
    const uint32_t n10 = blk(bx ? r.tb[i110] : r.tb[i010]);
    const uint32_t n01 = blk(bx ? r.tb[i101] : r.tb[i001]);
    const uint32_t n11 = blk(bx ? r.tb[i111] : r.tb[i011]);
    const uint32_t nb = by ? (bz ? n11 : n10) : (bz ? n01 : n00);
    using C = std::conditional_t<std::is_const<P>::value, const char, char>;
    return (P*)((C*)base + (nb + lane_shift<CY, CZ, M>(r.lane) * ES));
  }
  // Streamed store of slot d of the lane's site: a buffer store whose scalar
  // offset carries the wave-uniform part and whose vector offset is the
  // lane's (the same register for every direction), with the non-temporal
  // bit, as the 64-bit layouts' non-temporal stores.
  template <int Q, typename S>
  static __device__ __forceinline__ void store(const DevDomain& D, const SiteRef& r, S* base,
                                               int d, S v) {
    constexpr uint32_t ES = sizeof(S);
    const uint32_t uni = (((r.w >> 3) * (uint32_t)Q + (uint32_t)d) * 512u + ((r.w & 7u) << 6)) * ES;
    const uint32_t bytes = (uint32_t)D.np * (uint32_t)Q * ES;  // below 4 GiB
    // word 3 of a raw buffer resource on gfx9 parts: 32-bit data format
    const auto rs = __builtin_amdgcn_make_buffer_rsrc((void*)base, 0, (int)bytes, 0x00020000);
    constexpr int kNonTemporal = 2;
    const int voff = (int)(r.lane * ES);
    if constexpr (ES == 2)
      __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), rs, voff,
                                            (int)uni, kNonTemporal);
    else if constexpr (ES == 4)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), rs, voff,
                                            (int)uni, kNonTemporal);
    else {
      using U2 = unsigned int __attribute__((ext_vector_type(2)));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), rs, voff, (int)uni,
                                            kNonTemporal);
    }
  }
};

template <class Lay>
constexpr bool kWaveLayout = std::is_same<Lay, BlockLayout>::value ||
                             std::is_same<Lay, BlockLayout32>::value ||
                             std::is_same<Lay, CompactBlockLayout>::value;

// Own slot E of the lane's site where pick_e, else its own slot DP (a
// per-lane choice between two directions).  BlockLayout32: the wave-uniform
// offset of the lower direction plus a per-lane choice of two non-negative
// constants.
template <int Q, int E, int DP, class Lay, typename P>
__device__ __forceinline__ P* own_pick_ptr(const DevDomain& D, const SiteRef& r, P* base,
                                           bool pick_e) {
  if constexpr (std::is_same<Lay, BlockLayout32>::value) {
    constexpr uint32_t ES = sizeof(P);
    constexpr int LO = E < DP ? E : DP;
    constexpr uint32_t dE = (uint32_t)(E - LO) * 512u * ES, dP = (uint32_t)(DP - LO) * 512u * ES;
    const uint32_t uni = (((r.w >> 3) * (uint32_t)Q + (uint32_t)LO) * 512u + ((r.w & 7u) << 6)) * ES;
    using C = std::conditional_t<std::is_const<P>::value, const char, char>;
    return (P*)(((C*)base + uni) + (r.lane * ES + (pick_e ? dE : dP)));
  } else {
    return base + Lay::template at<Q>(D, pick_e ? E : DP, r.x);
  }
}

// Addresses of state elements: slot d of the lane's own site, and the value
// streaming into slot E of it (its source, or its own inverse slot across a
// cut link).  Element indices for the 64-bit layouts, byte offsets for
// BlockLayout32.
template <int Q, class Lay, typename P>
__device__ __forceinline__ P* own_ptr(const DevDomain& D, const SiteRef& r, P* base, int d) {
  if constexpr (std::is_same<Lay, BlockLayout32>::value)
    return Lay::template own_ptr<Q>(r, base, d);
  else
    return base + Lay::template at<Q>(D, d, r.x);
}
template <int Q, int E, class Lay, typename P>
__device__ __forceinline__ P* src_ptr(const DevDomain& D, const SiteRef& r, P* base) {
  if constexpr (std::is_same<Lay, BlockLayout32>::value)
    return Lay::template src_ptr<Q, E>(r, base);
  else
    return base + Lay::template index<Q, E, kMortonLanes<std::remove_const_t<P>>>(D, r);
}

// ---------------------------------------------------------------------------
// Streaming (pull) with boundary links
// ---------------------------------------------------------------------------
// Ghost equilibrium of the Nash zeroth-order pressure link
// (NashZerothOrderPressure.h:27-60): density of the iolet, momentum =
// (n_float * (m.n_float / rho)) * rho_ghost, one component of
// Lattice::CalculateFeq.  Always binary64.
template <int Q, int E>
__device__ __forceinline__ double nash_value(const DevDomain& D, const StepArgs& A,
                                             int io, int k) {
  const double rho_g = A.bv[2 * k];
  const double comp = D.io_stash[io];
  const double gx = (double)D.iolet_nf[3 * k + 0] * comp * rho_g;
  const double gy = (double)D.iolet_nf[3 * k + 1] * comp * rho_g;
  const double gz = (double)D.iolet_nf[3 * k + 2] * comp * rho_g;
  FeqCoef<double> fe(rho_g, gx, gy, gz);
  return fe.template eq<Q, E>();
}

// Correction of the Ladd link (LaddIolet.h:29-66): 2 w_d (rho u_w . c_d)/Cs2
// with u_w the iolet velocity half way along the link.
template <int Q, int DP>
__device__ __forceinline__ double ladd_correction(const DevDomain& D,
                                                  const StepArgs& A, int io,
                                                  int k) {
  using L = Lat<Q>;
  const double rho = D.io_stash[io];
  const double* lc = D.io_ladd + ((int64_t)io * (Q - 1) + DP - 1) * 2;
  const double nx = D.iolet_n[3 * k], ny = D.iolet_n[3 * k + 1],
               nz = D.iolet_n[3 * k + 2];
  double ux, uy, uz;
  if (D.iolet_kind[k] == 2) {
    // InOutLetWomersleyVelocity::GetVelocity: Re[A e^{i w t}], u = n * -vm
    const double vm = lc[0] * A.bv[2 * k] - lc[1] * A.bv[2 * k + 1];
    ux = nx * -vm;
    uy = ny * -vm;
    uz = nz * -vm;
  } else {
    // InOutLetParabolicVelocity::GetVelocity: n * (max * (1 - r^2/R^2))
    const double s = A.bv[2 * k] * lc[0];
    ux = nx * s;
    uy = ny * s;
    uz = nz * s;
  }
  ux *= rho;
  uy *= rho;
  uz *= rho;
  const double dot = dotc<L::c[DP][0], L::c[DP][1], L::c[DP][2]>(ux, uy, uz);
  return 2. * L::w[DP] * dot / kCs2;
}

// Cut distances of links 1..Q-1 from 16-bit fixed-point records (the
// reduced-precision forms): QS16 = Q - 1 rounded up to a multiple of 8.
template <int Q> constexpr int kQS16 = (Q - 1 + 7) / 8 * 8;
template <int Q>
__device__ __forceinline__ void decode_q16(const uint4 (&v)[kQS16<Q> / 8], float (&q)[kQS16<Q>]) {
  sfor<0, kQS16<Q> / 8>([&](auto ii) {
    constexpr int i = decltype(ii)::value;
    const uint32_t w[4] = {v[i].x, v[i].y, v[i].z, v[i].w};
    sfor<0, 4>([&](auto jj) {
      constexpr int j = decltype(jj)::value;
      q[8 * i + 2 * j] = (float)(w[j] & 0xFFFFu) * (1.0f / 65535.0f);
      q[8 * i + 2 * j + 1] = (float)(w[j] >> 16) * (1.0f / 65535.0f);
    });
  });
}

// BouzidiFirdaousLallemandLink::StreamLink (BouzidiFirdaousLallemand.h:41-66):
// with q >= 0.5 and a fluid opposite link the slot gets
// (f*_d + (2q - 1) f*_inv(d)) / (2q); both values are the site's own
// (opp[e] = f*_e of the site).  Slot e is a candidate (bit e-1 of cand) when
// the inverse link is a wall link and not an iolet link, and link e itself
// is not a wall link.  A link outside `plane` (the union of cand over the
// wave's lanes, a scalar) is passed over by one scalar branch.
template <int Q, typename S, typename T>
__device__ __forceinline__ void bfl_stream_links(const float (&q)[kQS16<Q>], uint32_t cand,
                                                 const T (&opp)[Q], T (&f)[Q],
                                                 uint32_t plane = ~0u) {
  sfor<1, Q>([&](auto ei) {
    constexpr int e = decltype(ei)::value;
    constexpr int dp = inv_dir(e);
    if (!((plane >> (e - 1)) & 1u)) return;
    const T qd = (T)(double)q[dp - 1];
    if (((cand >> (e - 1)) & 1u) && !(qd < T(0.5))) {
      if constexpr (kShifted<S>)
        f[e] = mad(T(2.0) * qd - T(1), opp[e], f[e]) * recip(T(2.0) * qd);
      else
        f[e] = (f[e] + (T(2.0) * qd - T(1)) * opp[e]) / (T(2.0) * qd);
    }
  });
}

// Iolet links of boundary record `rec`: the Nash ghost equilibrium or the
// Ladd correction of the bounce-back value.
template <int Q, typename S, typename T>
__device__ __forceinline__ void iolet_links(const DevDomain& D, const StepArgs& A, uint32_t rec,
                                            uint32_t im, T (&f)[Q]) {
  const int io = D.b_io[rec];
  const int k = D.io_iolet[io];
  const int bc = D.io_bc[io];
  sfor<1, Q>([&](auto ei) {
    constexpr int e = decltype(ei)::value;
    constexpr int dp = inv_dir(e);
    if ((im >> (dp - 1)) & 1u) {
      if (bc == 0) {
        double v = nash_value<Q, e>(D, A, io, k);
        if constexpr (kShifted<S>) v -= Lat<Q>::w[e] * A.rho0;
        f[e] = (T)v;
      } else {
        f[e] = f[e] - (T)ladd_correction<Q, dp>(D, A, io, k);
      }
    }
  });
}

// BouzidiFirdaousLallemandLink::PostStepLink: links d with a wall and no
// wall on the inverse link, q < 0.5 (`plane`: the union of those links over
// the wave's lanes, as for bfl_stream_links)
template <int Q, typename S, typename T>
__device__ __forceinline__ void bfl_post_links(const float (&q)[kQS16<Q>], uint32_t wm,
                                               T (&f)[Q], uint32_t plane = ~0u) {
  const uint32_t post = wm & ~pair_swap(wm);
  sfor<1, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    constexpr int id = inv_dir(d);
    if (!((plane >> (d - 1)) & 1u)) return;
    const T qd = (T)(double)q[d - 1];
    if (((post >> (d - 1)) & 1u) && qd < T(0.5)) {
      if constexpr (kShifted<S>)
        f[id] = mad(T(2.0) * qd, f[id] - f[d], f[d]);
      else
        f[id] = T(2.0) * qd * f[id] + (T(1.0) - T(2.0) * qd) * f[d];
    }
  });
}

// Binary16 with 32-bit offsets: every load of a site is issued in one batch
// after the plane's tables, none waits for the boundary record.  The pulls
// read x - c_e whether or not link e is cut (an absent neighbour block is
// replaced by the own block, so the address stays in the state), and lanes
// with a boundary record also load their own Q-1 values and cut distances;
// the bounce-back value of a cut link and the BFL value f*_e are then picked
// from the own values.  The same values and arithmetic as gather's other
// path: binary16 loads every own value that path picks from anyway.
template <int Q, typename S, typename T>
__device__ __forceinline__ void gather_early(const DevDomain& D, const StepArgs& A,
                                             const S* __restrict__ fin, const SiteRef& r,
                                             T (&f)[Q]) {
  using St = Store<S>;
  using Lay = BlockLayout32;
  const bool bnd = (D.wave_bnd[r.w] >> r.lane) & 1ull;
  f[0] = St::ld(Lay::template own_ptr<Q>(r, fin, 0));
  sfor<1, Q>([&](auto ei) {
    constexpr int e = decltype(ei)::value;
    f[e] = St::ld(Lay::template src_any_ptr<Q, e>(r, fin));
  });
  T own[Q];
  uint4 qv[kQS16<Q> / 8];
  if (bnd) {
    sfor<1, Q>([&](auto ei) {
      constexpr int e = decltype(ei)::value;
      own[e] = St::ld(Lay::template own_ptr<Q>(r, fin, e));
    });
    if (A.wall == 1) {
      const uint4* q8 = reinterpret_cast<const uint4*>(D.b_q16 + (int64_t)r.rec * kQS16<Q>);
      sfor<0, kQS16<Q> / 8>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        qv[i] = q8[i];
      });
    }
  }
  if (!bnd) return;
  sfor<1, Q>([&](auto ei) {
    constexpr int e = decltype(ei)::value;
    constexpr int dp = inv_dir(e);
    if ((r.cut >> (dp - 1)) & 1u) f[e] = own[dp];
  });
  const uint32_t rec = r.rec;
  const uint32_t wm = D.b_wall[rec], im = D.b_iolet[rec];
  // the plane's BFL links: formulas of links no lane has are skipped
  const uint64_t wl = D.wave_links[r.w];
  float q[kQS16<Q>];
  if (A.wall == 1 && wm) {
    decode_q16<Q>(qv, q);
    bfl_stream_links<Q, S, T>(q, pair_swap(wm) & ~pair_swap(im) & ~wm, own, f, (uint32_t)wl);
  }
  if (im) iolet_links<Q, S, T>(D, A, rec, im, f);
  if (A.wall == 1 && wm) bfl_post_links<Q, S, T>(q, wm, f, (uint32_t)(wl >> 32));
}

// Post-stream distributions of a site (the reference's fNew after
// StreamAndCollide and PostStep) from the post-collision state `fin`.
// All Q loads are issued first: a cut link reads the site's own value of the
// inverse direction (the bounce-back value); the wall and iolet formulas are
// applied afterwards.
template <int Q, typename S, typename T, class Lay, bool BULK = false>
__device__ __forceinline__ void gather(const DevDomain& D, const StepArgs& A,
                                       const S* __restrict__ fin, const SiteRef& r,
                                       T (&f)[Q]) {
  if constexpr (!BULK && sizeof(S) == 2 && std::is_same<Lay, BlockLayout32>::value) {
    gather_early<Q, S, T>(D, A, fin, r, f);
    return;
  }
  using St = Store<S>;
  f[0] = St::ld(own_ptr<Q, Lay>(D, r, fin, 0));
  sfor<1, Q>([&](auto ei) {
    constexpr int e = decltype(ei)::value;
    f[e] = St::ld(src_ptr<Q, e, Lay>(D, r, fin));
  });
  if constexpr (BULK) return;  // a wave without boundary sites
  if (!r.cut) return;
  const uint32_t rec = r.rec;
  const uint32_t wm = D.b_wall[rec], im = D.b_iolet[rec];
  // cut distances of links 1..Q-1 (only the wall links' are read): 16-bit
  // fixed-point records for the shifted forms, binary32 values of the wall
  // links alone, packed, for binary64 storage
  float q[kQS16<Q>];
  if (A.wall == 1 && wm) {
    if constexpr (kShifted<S>) {
      const uint4* q8 = reinterpret_cast<const uint4*>(D.b_q16 + (int64_t)rec * kQS16<Q>);
      uint4 qv[kQS16<Q> / 8];
      sfor<0, kQS16<Q> / 8>([&](auto ii) {
        constexpr int i = decltype(ii)::value;
        qv[i] = q8[i];
      });
      decode_q16<Q>(qv, q);
    } else {
      // 4.6 wall links per wall site on the vessel tree: 18 bytes instead
      // of an 80-byte record of every link
      const float* qp = D.b_q + D.b_qoff[rec];
      sfor<0, Q - 1>([&](auto kk) {
        constexpr int k = decltype(kk)::value;
        q[k] = -1.0f;
        if ((wm >> k) & 1u) q[k] = qp[__popc(wm & ((1u << k) - 1u))];
      });
    }
    // The links that take the interpolated form need the site's own f*_e.
    // Binary64 and binary32: only those lanes load (the rest are masked
    // off; measured 0.3-0.6 % faster on the vessel tree).  Binary16 (64-bit
    // addressing): every lane loads, the unused ones their already-fetched
    // bounce-back slot (predicated 16-bit loads measured 10 % slower).
    const uint32_t cand = pair_swap(wm) & ~pair_swap(im) & ~wm;
    T opp[Q];
    sfor<1, Q>([&](auto ei) {
      constexpr int e = decltype(ei)::value;
      constexpr int dp = inv_dir(e);
      const bool link = (cand >> (e - 1)) & 1u;
      if constexpr (sizeof(S) >= 4) {
        opp[e] = T(0);
        if (link && !((T)(double)q[dp - 1] < T(0.5)))
          opp[e] = St::ld(own_ptr<Q, Lay>(D, r, fin, e));
      } else {
        opp[e] = St::ld(own_pick_ptr<Q, e, dp, Lay>(D, r, fin, link));
      }
    });
    bfl_stream_links<Q, S, T>(q, cand, opp, f);
  }
  if (im) iolet_links<Q, S, T>(D, A, rec, im, f);
  if (A.wall == 1 && wm) bfl_post_links<Q, S, T>(q, wm, f);
}

template <int Q, typename T, bool BULK = false>
__device__ __forceinline__ void update_stash(const DevDomain& D, const SiteRef& r,
                                             const Hydro<Q, T>& h) {
  if constexpr (BULK) return;
  if (!r.cut) return;
  const int io = D.b_io[r.rec];
  if (io < 0) return;
  const double rho = (double)h.rho;
  if (D.io_bc[io] == 0) {
    // component = Dot(momentum, n_float) / density (NashZerothOrderPressure.h:43)
    const int k = D.io_iolet[io];
    double acc = 0.0;
    acc = acc + (double)h.mx * (double)D.iolet_nf[3 * k];
    acc = acc + (double)h.my * (double)D.iolet_nf[3 * k + 1];
    acc = acc + (double)h.mz * (double)D.iolet_nf[3 * k + 2];
    D.io_stash[io] = acc / rho;
  } else {
    D.io_stash[io] = rho;  // LaddIolet.h:55-58 (compressible lattice)
  }
}

template <typename S>
__device__ __forceinline__ void store_out(S* p, typename Store<S>::T v) {
  // Written once per step and read only by the next launch: stream it past
  // the caches.
  __builtin_nontemporal_store((S)v, p);
}

template <int Q, int KERNEL, typename S, class Lay, bool BULK = false>
__device__ __forceinline__ void step_site(const DevDomain& D, const StepArgs& A,
                                          const S* __restrict__ fin,
                                          S* __restrict__ fout, const SiteRef& r) {
  using T = typename Store<S>::T;
  T f[Q];
  gather<Q, S, T, Lay, BULK>(D, A, fin, r, f);
  // Stability (every f > 0) is tested on the state by the monitor's
  // reduction, as the reference's StabilityTester tests fNew
  // (StabilityTester.h:97-113), not inside the step.
  StepArgs a = A;
  Lay::relax(D, r, a);
  const Hydro<Q, T> h = collide<Q, KERNEL, S>(f, a);
  update_stash<Q, T, BULK>(D, r, h);
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    if constexpr (std::is_same<Lay, BlockLayout32>::value)
      Lay::template store<Q, S>(D, r, fout, d, (S)f[d]);
    else
      store_out<S>(own_ptr<Q, Lay>(D, r, fout, d), f[d]);
  });
}

// One timestep, one site per thread.  Block layout: one x plane per wave,
// workgroups of A.threads, optionally remapped to XCD-contiguous ranges.
template <int Q, int KERNEL, typename S, class Lay>
__device__ __forceinline__ void step_wave(const DevDomain& D, const StepArgs& A,
                                          const S* __restrict__ fin, S* __restrict__ fout) {
  SiteRef r;
  if constexpr (kWaveLayout<Lay>) {
    const uint32_t g = step_workgroup(A.xcd_remap);
    const uint32_t w = __builtin_amdgcn_readfirstlane(g * (blockDim.x >> 6) + (threadIdx.x >> 6));
    // (A wave-uniform branch to the boundary-free code for planes without
    // boundary sites cuts their instructions by a third but measured slower:
    // the kernel then needs more registers and the two paths more
    // instruction cache.)
    if (!Lay::locate_wave(D, w, r)) return;
  } else {
    if (!Lay::locate(D, r)) return;
  }
  step_site<Q, KERNEL, S, Lay>(D, A, fin, fout, r);
}
template <int Q, int KERNEL, typename S, class Lay>
__global__ __launch_bounds__(256) void k_step(DevDomain D, StepArgs A,
                                              const S* __restrict__ fin,
                                              S* __restrict__ fout) {
  step_wave<Q, KERNEL, S, Lay>(D, A, fin, fout);
}
// The same step for binary16 storage, held to 7 waves per SIMD (at most 72
// VGPRs instead of the 78 it would take, 6 waves): binary16 is bound by the
// issue and latency of its vector instructions rather than by bytes, and
// the seventh wave measured 3 % faster on the vessel tree (binary32 1 %
// slower, binary64 spills).
template <int Q, int KERNEL, typename S, class Lay>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7)))
void k_step_w7(DevDomain D, StepArgs A, const S* __restrict__ fin, S* __restrict__ fout) {
  step_wave<Q, KERNEL, S, Lay>(D, A, fin, fout);
}

// The timestep over a list of waves (x planes).  The solver splits the
// planes of the block layouts into those without any boundary site, run by
// the BULK instantiation (no boundary code, so far fewer registers and more
// waves in flight), and those with one, run by the full one; together the
// two launches update every site exactly as k_step does.
template <int Q, int KERNEL, typename S, class Lay, bool BULK>
__global__ __launch_bounds__(256) void k_step_list(DevDomain D, StepArgs A,
                                                   const S* __restrict__ fin,
                                                   S* __restrict__ fout,
                                                   const uint32_t* __restrict__ waves,
                                                   uint32_t count) {
  const uint32_t g = step_workgroup(A.xcd_remap);
  const uint32_t i = __builtin_amdgcn_readfirstlane(g * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (i >= count) return;
  const uint32_t w = __builtin_amdgcn_readfirstlane(waves[i]);
  SiteRef r;
  if (!Lay::locate_wave(D, w, r)) return;
  if constexpr (BULK) r.cut = 0;  // known: no cut links in these planes
  step_site<Q, KERNEL, S, Lay, BULK>(D, A, fin, fout, r);
}

// The same timestep for the block layout with a wave loop: each wave takes
// A.planes consecutive x planes and a grid of A.groups workgroups strides
// over the domain (a persistent launch).  The loop keeps lane-derived
// offsets live across iterations, so it needs more registers than k_step;
// the launch-shape tuner measures both.
template <int Q, int KERNEL, typename S, class Lay>
__global__ __launch_bounds__(256) void k_step_loop(DevDomain D, StepArgs A,
                                                   const S* __restrict__ fin,
                                                   S* __restrict__ fout) {
  const uint32_t wpg = blockDim.x >> 6;
  const uint32_t g = step_workgroup(A.xcd_remap);
  const uint32_t gw = __builtin_amdgcn_readfirstlane(g * wpg + (threadIdx.x >> 6));
  const uint32_t stride = gridDim.x * wpg * A.planes;
  for (uint32_t base = gw * A.planes; base < D.nwaves; base += stride) {
    for (uint32_t i = 0; i < A.planes; ++i) {
      SiteRef r;
      if (Lay::locate_wave(D, base + i, r))
        step_site<Q, KERNEL, S, Lay>(D, A, fin, fout, r);
    }
  }
}

template <int Q, int KERNEL, typename S, class Lay>
__global__ __launch_bounds__(256) void k_collide_only(DevDomain D, StepArgs A,
                                                      const S* __restrict__ fin,
                                                      S* __restrict__ fout) {
  using T = typename Store<S>::T;
  SiteRef r;
  if (!Lay::locate(D, r)) return;
  T f[Q];
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    f[d] = Store<S>::ld(fin + Lay::template at<Q>(D, d, r.x));
  });
  StepArgs a = A;
  Lay::relax(D, r, a);
  const Hydro<Q, T> h = collide<Q, KERNEL, S>(f, a);
  update_stash<Q, T>(D, r, h);
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    Store<S>::st(fout + Lay::template at<Q>(D, d, r.x), f[d]);
  });
}

// Post-stream distributions in binary64, compact SoA [Q][npc].
template <int Q, typename S, class Lay>
__global__ __launch_bounds__(256) void k_stream_only(DevDomain D, StepArgs A,
                                                     const S* __restrict__ fin,
                                                     double* __restrict__ out,
                                                     int from_post_stream) {
  using T = typename Store<S>::T;
  SiteRef r;
  if (!Lay::locate(D, r)) return;
  T f[Q];
  if (from_post_stream) {
    sfor<0, Q>([&](auto di) {
      constexpr int d = decltype(di)::value;
      f[d] = Store<S>::ld(fin + Lay::template at<Q>(D, d, r.x));
    });
  } else {
    gather<Q, S, T, Lay>(D, A, fin, r, f);
  }
  const int64_t c = Lay::compact(D, r);
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    double v = (double)f[d];
    if constexpr (kShifted<S>) v += Lat<Q>::w[d] * A.rho0;
    out[(int64_t)d * D.npc + c] = v;
  });
}

// Device-copy ceiling of the step's traffic: the Q state arrays (stride
// `stride` words of type W) copied from `in` to `out`, `words` per array,
// one word per thread and array, all loads issued before the streamed
// stores, workgroups mapped to XCDs as the step maps them.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int Q, typename W>
__global__ __launch_bounds__(256) void k_copy_arrays(int64_t stride, int64_t words, int64_t run,
                                                     const W* __restrict__ in,
                                                     W* __restrict__ out, uint32_t remap) {
  const int64_t i = (int64_t)step_workgroup(remap) * blockDim.x + threadIdx.x;
  if (i >= words) return;
  // run > 0: the block layout's order, runs of `run` words per direction
  // and block; else arrays of `stride` words per direction
  const int64_t base = run ? (i / run) * Q * run + (i % run) : i;
  const int64_t step = run ? run : stride;
  W v[Q];
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    v[d] = __builtin_nontemporal_load(in + base + d * step);
  });
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    __builtin_nontemporal_store(v[d], out + base + d * step);
  });
}

// Halo exchange between subdomains: stored values copied as raw words
// (E = the storage's 8-, 4- or 2-byte word), from state to a buffer (pack)
// or from a buffer to state (unpack), at precomputed state indices.
template <typename E>
__global__ void k_halo_pack(int64_t count, const int64_t* __restrict__ idx,
                            const E* __restrict__ state, E* __restrict__ buf) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) buf[k] = state[idx[k]];
}
template <typename E>
__global__ void k_halo_unpack(int64_t count, const int64_t* __restrict__ idx,
                              const E* __restrict__ buf, E* __restrict__ state) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) state[idx[k]] = buf[k];
}

// Caller-order AoS [n][Q] -> state, perm[compact] = caller index.
template <int Q, typename S, class Lay>
__global__ __launch_bounds__(256) void k_import(DevDomain D, const double* aos,
                                                const int64_t* perm, S* state,
                                                double rho0) {
  SiteRef r;
  if (!Lay::locate(D, r)) return;
  const int64_t c = perm[Lay::compact(D, r)];
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    double v = aos[c * Q + d];
    if constexpr (kShifted<S>) v -= Lat<Q>::w[d] * rho0;
    Store<S>::st(state + Lay::template at<Q>(D, d, r.x), (typename Store<S>::T)v);
  });
}

// Equilibrium of a per-site (density, momentum) field, caller's site order
// (perm maps the layout's compact index to it), with the collision's own
// Lattice::CalculateFeq arithmetic.
template <int Q, typename S, class Lay>
__global__ __launch_bounds__(256) void k_import_equilibrium(DevDomain D, const double* rho,
                                                            const double* mom,
                                                            const int64_t* perm, S* state,
                                                            double rho0) {
  SiteRef r;
  if (!Lay::locate(D, r)) return;
  const int64_t c = perm[Lay::compact(D, r)];
  const FeqCoef<double> fe(rho[c], mom[3 * c], mom[3 * c + 1], mom[3 * c + 2]);
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    double v = fe.template eq<Q, d>();
    if constexpr (kShifted<S>) v -= Lat<Q>::w[d] * rho0;
    Store<S>::st(state + Lay::template at<Q>(D, d, r.x), (typename Store<S>::T)v);
  });
}

template <int Q, typename S>
__global__ void k_fill(int64_t nslots, int64_t np, int blockwise, const double* vals,
                       S* state, double rho0) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nslots) return;
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    double v = vals[d];
    if constexpr (kShifted<S>) v -= Lat<Q>::w[d] * rho0;
    const int64_t i = blockwise ? ((x >> 9) * Q + d) * 512 + (x & 511) : (int64_t)d * np + x;
    Store<S>::st(state + i, (typename Store<S>::T)v);
  });
}

__global__ void k_soa_to_aos(int Q, int64_t n, int64_t np, const double* soa,
                             const int64_t* perm, double* aos) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const int64_t c = perm[x];
  for (int d = 0; d < Q; ++d) aos[c * Q + d] = soa[(int64_t)d * np + x];
}

template <int Q>
__global__ void k_moments(int64_t n, int64_t np, const double* soa, double* rho,
                          double* u) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  double f[Q];
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    f[d] = soa[(int64_t)d * np + x];
  });
  double r, mx, my, mz;
  moments<Q, double>(f, r, mx, my, mz);
  rho[x] = r;
  u[x] = mx / r;
  u[np + x] = my / r;
  u[2 * np + x] = mz / r;
}

// Per-block binary64 reductions in a fixed tree order: [mass, min rho,
// max rho, max |u|, unstable] per block.
template <int Q>
__global__ __launch_bounds__(256) void k_reduce(int64_t n, int64_t np,
                                                const double* soa,
                                                const double* rho,
                                                const double* u,
                                                double* partials) {
  __shared__ double s_sum[256], s_min[256], s_max[256], s_u[256], s_bad[256];
  const int t = threadIdx.x;
  double sum = 0.0, mn = 1.7976931348623157e308, mx = -1.7976931348623157e308,
         um = 0.0, bad = 0.0;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + t; x < n;
       x += (int64_t)gridDim.x * blockDim.x) {
    const double r = rho[x];
    sum += r;
    mn = r < mn ? r : mn;
    mx = r > mx ? r : mx;
    const double a = u[x], b = u[np + x], c = u[2 * np + x];
    const double m = sqrt(((0.0 + a * a) + b * b) + c * c);
    um = m > um ? m : um;
    for (int d = 0; d < Q; ++d)
      if (!(soa[(int64_t)d * np + x] > 0.0)) bad = 1.0;
  }
  s_sum[t] = sum; s_min[t] = mn; s_max[t] = mx; s_u[t] = um; s_bad[t] = bad;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      s_sum[t] += s_sum[t + w];
      s_min[t] = s_min[t + w] < s_min[t] ? s_min[t + w] : s_min[t];
      s_max[t] = s_max[t + w] > s_max[t] ? s_max[t + w] : s_max[t];
      s_u[t] = s_u[t + w] > s_u[t] ? s_u[t + w] : s_u[t];
      s_bad[t] = s_bad[t + w] > s_bad[t] ? s_bad[t + w] : s_bad[t];
    }
    __syncthreads();
  }
  if (t == 0) {
    double* p = partials + 5 * blockIdx.x;
    p[0] = s_sum[0]; p[1] = s_min[0]; p[2] = s_max[0]; p[3] = s_u[0]; p[4] = s_bad[0];
  }
}

// The monitor's statistics straight from the state, without an exported
// copy: every site's fOld (the state itself when it is post-stream, else the
// pull of k_stream_only), its density and velocity as k_moments computes
// them, and per workgroup [mass, min rho, max rho, max |u|, unstable] as
// k_reduce combines them.  Workgroup g visits the x planes g*W + k*W*grid
// (W waves per workgroup) of the block layouts, or the compact sites
// g*256 + t + k*256*grid of the pull table, so the partials depend on the
// grid size only.  The density and velocity of every iolet site go to
// io_vals[4*record..] (rho, u_x, u_y, u_z) for the iolet fluxes.
template <int Q, typename S, class Lay>
__global__ __launch_bounds__(256) void k_monitor(DevDomain D, StepArgs A,
                                                 const S* __restrict__ fin,
                                                 int from_post_stream,
                                                 double* __restrict__ partials,
                                                 double* __restrict__ io_vals) {
  using T = typename Store<S>::T;
  __shared__ double s_sum[256], s_min[256], s_max[256], s_u[256], s_bad[256];
  const int t = threadIdx.x;
  double sum = 0.0, mn = 1.7976931348623157e308, mx = -1.7976931348623157e308,
         um = 0.0, bad = 0.0;
  auto visit = [&](const SiteRef& r) {
    T f[Q];
    if (from_post_stream) {
      sfor<0, Q>([&](auto di) {
        constexpr int d = decltype(di)::value;
        f[d] = Store<S>::ld(fin + Lay::template at<Q>(D, d, r.x));
      });
    } else {
      gather<Q, S, T, Lay>(D, A, fin, r, f);
    }
    double g[Q];
    sfor<0, Q>([&](auto di) {
      constexpr int d = decltype(di)::value;
      double v = (double)f[d];
      if constexpr (kShifted<S>) v += Lat<Q>::w[d] * A.rho0;
      g[d] = v;
      if (!(v > 0.0)) bad = 1.0;
    });
    double rho, m0, m1, m2;
    moments<Q, double>(g, rho, m0, m1, m2);
    const double ux = m0 / rho, uy = m1 / rho, uz = m2 / rho;
    sum += rho;
    mn = rho < mn ? rho : mn;
    mx = rho > mx ? rho : mx;
    const double m = sqrt(((0.0 + ux * ux) + uy * uy) + uz * uz);
    um = m > um ? m : um;
    if (r.cut) {
      const int io = D.b_io[r.rec];
      if (io >= 0) {
        io_vals[4 * (int64_t)io] = rho;
        io_vals[4 * (int64_t)io + 1] = ux;
        io_vals[4 * (int64_t)io + 2] = uy;
        io_vals[4 * (int64_t)io + 3] = uz;
      }
    }
  };
  if constexpr (kWaveLayout<Lay>) {
    const uint32_t wpg = blockDim.x >> 6;
    for (uint32_t base = blockIdx.x * wpg; base < D.nwaves; base += gridDim.x * wpg) {
      const uint32_t w = __builtin_amdgcn_readfirstlane(base + ((uint32_t)t >> 6));
      SiteRef r;
      if (Lay::locate_wave(D, w, r)) visit(r);
    }
  } else {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + t; x < D.n;
         x += (int64_t)gridDim.x * blockDim.x) {
      SiteRef r;
      Lay::at(D, x, r);
      visit(r);
    }
  }
  s_sum[t] = sum; s_min[t] = mn; s_max[t] = mx; s_u[t] = um; s_bad[t] = bad;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      s_sum[t] += s_sum[t + w];
      s_min[t] = s_min[t + w] < s_min[t] ? s_min[t + w] : s_min[t];
      s_max[t] = s_max[t + w] > s_max[t] ? s_max[t + w] : s_max[t];
      s_u[t] = s_u[t + w] > s_u[t] ? s_u[t + w] : s_u[t];
      s_bad[t] = s_bad[t + w] > s_bad[t] ? s_bad[t + w] : s_bad[t];
    }
    __syncthreads();
  }
  if (t == 0) {
    double* p = partials + 5 * blockIdx.x;
    p[0] = s_sum[0]; p[1] = s_min[0]; p[2] = s_max[0]; p[3] = s_u[0]; p[4] = s_bad[0];
  }
}

// The same statistics per chunk of sites [range[2b], range[2b+1]) (one
// workgroup per chunk; chunks never mix instances of an ensemble).
template <int Q>
__global__ __launch_bounds__(256) void k_reduce_ranges(int64_t np, const int64_t* range,
                                                       const double* soa,
                                                       const double* rho,
                                                       const double* u,
                                                       double* partials) {
  __shared__ double s_sum[256], s_min[256], s_max[256], s_u[256], s_bad[256];
  const int t = threadIdx.x;
  double sum = 0.0, mn = 1.7976931348623157e308, mx = -1.7976931348623157e308,
         um = 0.0, bad = 0.0;
  const int64_t end = range[2 * blockIdx.x + 1];
  for (int64_t x = range[2 * blockIdx.x] + t; x < end; x += blockDim.x) {
    const double r = rho[x];
    sum += r;
    mn = r < mn ? r : mn;
    mx = r > mx ? r : mx;
    const double a = u[x], b = u[np + x], c = u[2 * np + x];
    const double m = sqrt(((0.0 + a * a) + b * b) + c * c);
    um = m > um ? m : um;
    for (int d = 0; d < Q; ++d)
      if (!(soa[(int64_t)d * np + x] > 0.0)) bad = 1.0;
  }
  s_sum[t] = sum; s_min[t] = mn; s_max[t] = mx; s_u[t] = um; s_bad[t] = bad;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      s_sum[t] += s_sum[t + w];
      s_min[t] = s_min[t + w] < s_min[t] ? s_min[t + w] : s_min[t];
      s_max[t] = s_max[t + w] > s_max[t] ? s_max[t + w] : s_max[t];
      s_u[t] = s_u[t + w] > s_u[t] ? s_u[t + w] : s_u[t];
      s_bad[t] = s_bad[t + w] > s_bad[t] ? s_bad[t + w] : s_bad[t];
    }
    __syncthreads();
  }
  if (t == 0) {
    double* p = partials + 5 * blockIdx.x;
    p[0] = s_sum[0]; p[1] = s_min[0]; p[2] = s_max[0]; p[3] = s_u[0]; p[4] = s_bad[0];
  }
}

__global__ __launch_bounds__(256) void k_velocity_change_ranges(
    int64_t np, const int64_t* range, const double* u, const double* u0,
    const double* inv_ref, double* partials) {
  __shared__ double s[256];
  const int t = threadIdx.x;
  double m = 0.0;
  const double ir = inv_ref[blockIdx.x];
  const int64_t end = range[2 * blockIdx.x + 1];
  for (int64_t x = range[2 * blockIdx.x] + t; x < end; x += blockDim.x) {
    const double a = u[x] - u0[x], b = u[np + x] - u0[np + x],
                 c = u[2 * np + x] - u0[2 * np + x];
    const double v = sqrt(((0.0 + a * a) + b * b) + c * c) * ir;
    m = v > m ? v : m;
  }
  s[t] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) s[t] = s[t + w] > s[t] ? s[t + w] : s[t];
    __syncthreads();
  }
  if (t == 0) partials[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_velocity_change(int64_t n, int64_t np,
                                                         const double* u,
                                                         const double* u0,
                                                         double inv_ref,
                                                         double* partials) {
  __shared__ double s[256];
  const int t = threadIdx.x;
  double m = 0.0;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + t; x < n;
       x += (int64_t)gridDim.x * blockDim.x) {
    const double a = u[x] - u0[x], b = u[np + x] - u0[np + x],
                 c = u[2 * np + x] - u0[2 * np + x];
    const double v = sqrt(((0.0 + a * a) + b * b) + c * c) * inv_ref;
    m = v > m ? v : m;
  }
  s[t] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) s[t] = s[t + w] > s[t] ? s[t + w] : s[t];
    __syncthreads();
  }
  if (t == 0) partials[blockIdx.x] = s[0];
}

// Wall shear stress magnitude at the wall sites (Lattice.h:652-727) from the
// post-stream binary64 distributions: f_neq = f - f_eq with the SSE3-order
// equilibrium; Pi_ab = sum_l f_neq_l c_la c_lb accumulated in direction
// order (terms with c_la c_lb = 0 add nothing); t_a = sum_b (Pi_ab n_b) temp;
// result sqrt(|t|^2 - (t . n)^2).
template <int Q, int A, int B>
__device__ __forceinline__ double second_moment(const double (&fneq)[Q]) {
  using L = Lat<Q>;
  double acc = 0.0;
  sfor<0, Q>([&](auto li) {
    constexpr int l = decltype(li)::value;
    constexpr int c = L::c[l][A] * L::c[l][B];
    if constexpr (c == 1) acc = acc + fneq[l];
    else if constexpr (c == -1) acc = acc - fneq[l];
  });
  return acc;
}

template <int Q>
__global__ __launch_bounds__(256) void k_wall_shear_stress(
    int64_t nw, int64_t np, const double* __restrict__ soa,
    const int64_t* __restrict__ site, const float* __restrict__ nrm,
    const double* __restrict__ temp_of, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  const int64_t x = site[i];
  double f[Q];
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    f[d] = soa[(int64_t)d * np + x];
  });
  double r, mx, my, mz;
  moments<Q, double>(f, r, mx, my, mz);
  const FeqCoef<double> fe(r, mx, my, mz);
  double fneq[Q];
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    fneq[d] = f[d] - fe.template eq<Q, d>();
  });
  const double p00 = second_moment<Q, 0, 0>(fneq), p10 = second_moment<Q, 1, 0>(fneq),
               p11 = second_moment<Q, 1, 1>(fneq), p20 = second_moment<Q, 2, 0>(fneq),
               p21 = second_moment<Q, 2, 1>(fneq), p22 = second_moment<Q, 2, 2>(fneq);
  const double pi[3][3] = {{p00, p10, p20}, {p10, p11, p21}, {p20, p21, p22}};
  const double n[3] = {(double)nrm[3 * i], (double)nrm[3 * i + 1], (double)nrm[3 * i + 2]};
  const double temp = temp_of[i];
  double sq = 0.0, nn = 0.0;
  for (int a = 0; a < 3; ++a) {
    double t = 0.0;
    for (int b = 0; b < 3; ++b) t += pi[a][b] * n[b] * temp;
    sq += t * t;
    nn += t * n[a];
  }
  out[i] = sqrt(sq - nn * nn);
}

// The other stress outputs of the reference's property cache (Lattice.h):
// f_neq = f - f_eq (SSE3 order), Pi_ab = sum_l f_neq_l c_la c_lb in direction
// order.  Stress tensor (CalculateStressTensor, 630-645): Pi * (1 - 1/(2 tau))
// plus (rho - 1) Cs2 on the diagonal, upper triangle xx xy xz yy yz zz;
// von Mises stress (508-552); shear rate (728-753, 890-907).  The tau of
// each site is tau_site[x] when given.
template <int Q>
__global__ __launch_bounds__(256) void k_stress_fields(
    int64_t n, int64_t np, const double* __restrict__ soa, double tau,
    const double* __restrict__ tau_site, double* __restrict__ t6,
    double* __restrict__ vm, double* __restrict__ sr) {
  using L = Lat<Q>;
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  const double ts = tau_site ? tau_site[x] : tau;
  double f[Q];
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    f[d] = soa[(int64_t)d * np + x];
  });
  double r, mx, my, mz;
  moments<Q, double>(f, r, mx, my, mz);
  const FeqCoef<double> fe(r, mx, my, mz);
  double fneq[Q];
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    fneq[d] = f[d] - fe.template eq<Q, d>();
  });
  const double p00 = second_moment<Q, 0, 0>(fneq), p10 = second_moment<Q, 1, 0>(fneq),
               p11 = second_moment<Q, 1, 1>(fneq), p20 = second_moment<Q, 2, 0>(fneq),
               p21 = second_moment<Q, 2, 1>(fneq), p22 = second_moment<Q, 2, 2>(fneq);
  if (t6) {
    const double fac = 1 - 1 / (2 * ts);
    const double pr = (r - 1) * kCs2;
    t6[x] = p00 * fac + pr;
    t6[np + x] = p10 * fac;
    t6[2 * np + x] = p20 * fac;
    t6[3 * np + x] = p11 * fac + pr;
    t6[4 * np + x] = p21 * fac;
    t6[5 * np + x] = p22 * fac + pr;
  }
  if (vm) {
    // sums of f_neq (cx^2 - cy^2) etc. in direction order
    double xx_yy = 0.0, yy_zz = 0.0, xx_zz = 0.0, xy = 0.0, xz = 0.0, yz = 0.0;
    sfor<0, Q>([&](auto li) {
      constexpr int l = decltype(li)::value;
      constexpr int cx = L::c[l][0], cy = L::c[l][1], cz = L::c[l][2];
      constexpr int a = cx * cx - cy * cy, b = cy * cy - cz * cz, c = cx * cx - cz * cz;
      if constexpr (a != 0) xx_yy += fneq[l] * (double)a;
      if constexpr (b != 0) yy_zz += fneq[l] * (double)b;
      if constexpr (c != 0) xx_zz += fneq[l] * (double)c;
      if constexpr (cx * cy != 0) xy += fneq[l] * (double)cx * (double)cy;
      if constexpr (cx * cz != 0) xz += fneq[l] * (double)cx * (double)cz;
      if constexpr (cy * cz != 0) yz += fneq[l] * (double)cy * (double)cz;
    });
    const double a = xx_yy * xx_yy + yy_zz * yy_zz + xx_zz * xx_zz;
    const double b = xy * xy + xz * xz + yz * yz;
    const double sp = (1.0 - 1.0 / (2.0 * ts)) / sqrt(2.0);
    vm[x] = sp * sqrt(a + 6.0 * b);
  }
  if (sr) {
    const double k = -1.0 / (2.0 * ts * r * kCs2);
    const double e[6] = {p00 * k, p10 * k, p20 * k, p11 * k, p21 * k, p22 * k};
    // rows: (0,0) (0,1) (0,2) (1,1) (1,2) (2,2)
    double acc = 0.0;
    acc += e[0] * e[0];
    acc += 2 * e[1] * e[1];
    acc += 2 * e[2] * e[2];
    acc += e[3] * e[3];
    acc += 2 * e[4] * e[4];
    acc += e[5] * e[5];
    sr[x] = sqrt(2 * acc);
  }
}

// Traction sigma n and its tangential projection t - n (t . n) at the wall
// sites (CalculateTractionOnAPoint / CalculateTangentialProjectionTraction,
// Lattice.h:560-602); tau_of[i] is the tau of wall site i.
template <int Q>
__global__ __launch_bounds__(256) void k_traction(
    int64_t nw, int64_t np, const double* __restrict__ soa,
    const int64_t* __restrict__ site, const float* __restrict__ nrm,
    const double* __restrict__ tau_of, double* __restrict__ tr, double* __restrict__ tg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  const int64_t x = site[i];
  double f[Q];
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    f[d] = soa[(int64_t)d * np + x];
  });
  double r, mx, my, mz;
  moments<Q, double>(f, r, mx, my, mz);
  const FeqCoef<double> fe(r, mx, my, mz);
  double fneq[Q];
  sfor<0, Q>([&](auto di) {
    constexpr int d = decltype(di)::value;
    fneq[d] = f[d] - fe.template eq<Q, d>();
  });
  const double fac = 1 - 1 / (2 * tau_of[i]);
  const double pr = (r - 1) * kCs2;
  const double p00 = second_moment<Q, 0, 0>(fneq), p10 = second_moment<Q, 1, 0>(fneq),
               p11 = second_moment<Q, 1, 1>(fneq), p20 = second_moment<Q, 2, 0>(fneq),
               p21 = second_moment<Q, 2, 1>(fneq), p22 = second_moment<Q, 2, 2>(fneq);
  const double sg[3][3] = {{p00 * fac + pr, p10 * fac, p20 * fac},
                           {p10 * fac, p11 * fac + pr, p21 * fac},
                           {p20 * fac, p21 * fac, p22 * fac + pr}};
  const double n[3] = {(double)nrm[3 * i], (double)nrm[3 * i + 1], (double)nrm[3 * i + 2]};
  double t[3];
  for (int a = 0; a < 3; ++a) {
    t[a] = 0.0;
    for (int b = 0; b < 3; ++b) t[a] += sg[a][b] * n[b];
  }
  double tn = 0.0;
  for (int a = 0; a < 3; ++a) tn = tn + t[a] * n[a];
  for (int a = 0; a < 3; ++a) {
    tr[3 * i + a] = t[a];
    tg[3 * i + a] = t[a] - n[a] * tn;
  }
}

template <typename S>
__global__ void k_gather_pairs(int64_t count, int64_t np, int Q, int blockwise,
                               const int64_t* slot, const S* state, double rho0,
                               const double* w, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int64_t sidx = slot[i];
  double v = (double)Store<S>::ld(state + sidx);
  const int d = blockwise ? (int)((sidx >> 9) % Q) : (int)(sidx / np);
  if constexpr (kShifted<S>) v += w[d] * rho0;
  out[i] = v;
}

// Halo unpack: state[slot[i]] = in[i] (shifted storage stores in - w rho0).
template <typename S>
__global__ void k_scatter_pairs(int64_t count, int64_t np, int Q, int blockwise,
                                const int64_t* slot, const double* in, double rho0,
                                const double* w, S* state) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int64_t sidx = slot[i];
  double v = in[i];
  const int d = blockwise ? (int)((sidx >> 9) % Q) : (int)(sidx / np);
  if constexpr (kShifted<S>) v -= w[d] * rho0;
  Store<S>::st(state + sidx, (typename Store<S>::T)v);
}

inline dim3 grid_for(int64_t n, int threads = 256) {
  return dim3((unsigned)((n + threads - 1) / threads));
}

// Threads launched for a per-site kernel: compact sites or all slots.
inline int64_t thread_count(const DevDomain& d) {
  return d.layout == 0 ? d.n : (int64_t)d.nwaves * 64;  // sites, or lanes of all planes
}

// (layout, lattice, kernel, storage) -> instantiation
template <typename Fn>
int dispatch(const DevDomain& d, const StepArgs& a, Fn&& fn) {
#define HLB_CASE(L_, Q_, K_, S_, SC_)                                        \
  if (d.layout == L_ && a.Q == Q_ && a.kernel == K_ && a.storage == SC_)    \
    return fn(std::integral_constant<int, Q_>{}, std::integral_constant<int, K_>{}, \
              (S_*)nullptr,                                                  \
              (std::conditional_t<L_ == 0, SparseLayout,                     \
                                  std::conditional_t<L_ == 1, BlockLayout,    \
                                                     CompactBlockLayout>>*)nullptr);
#define HLB_CASES(L_)                                                        \
  HLB_CASE(L_, 15, 0, double, 0) HLB_CASE(L_, 15, 1, double, 0)             \
  HLB_CASE(L_, 19, 0, double, 0) HLB_CASE(L_, 19, 1, double, 0)             \
  HLB_CASE(L_, 15, 0, float, 1) HLB_CASE(L_, 15, 1, float, 1)               \
  HLB_CASE(L_, 19, 0, float, 1) HLB_CASE(L_, 19, 1, float, 1)               \
  HLB_CASE(L_, 15, 0, _Float16, 2) HLB_CASE(L_, 15, 1, _Float16, 2)         \
  HLB_CASE(L_, 19, 0, _Float16, 2) HLB_CASE(L_, 19, 1, _Float16, 2)
  HLB_CASES(0)
  HLB_CASES(1)
  HLB_CASES(2)
#undef HLB_CASES
#undef HLB_CASE
  return (int)hipErrorInvalidValue;
}

}  // namespace

int launch_fill(const DevDomain& d, const StepArgs& a, const double* vals, void* state,
                void* stream) {
  return dispatch(d, a, [&](auto q, auto, auto* s, auto*) {
    using S = std::remove_pointer_t<decltype(s)>;
    constexpr int Q = decltype(q)::value;
    hipLaunchKernelGGL((k_fill<Q, S>), grid_for(d.np), dim3(256), 0, (hipStream_t)stream,
                       d.np, d.np, d.layout == 1 ? 1 : 0, vals, (S*)state, a.rho0);
    return (int)hipGetLastError();
  });
}

namespace {

// One step on the launch shape in `a` with layout Lay.
template <int Q, int K, typename S, class Lay>
int step_launch(const DevDomain& d, const StepArgs& a, const void* fin, void* fout,
                void* stream) {
  if constexpr (kWaveLayout<Lay>) {
    // launch shape: a.threads per workgroup, a.planes per wave, a grid of
    // a.groups workgroups (0: enough to cover every plane once)
    const uint32_t wpg = a.threads / 64u;
    const uint64_t need = ((uint64_t)d.nwaves + (uint64_t)wpg * a.planes - 1) /
                          ((uint64_t)wpg * a.planes);
    uint64_t groups = a.groups ? std::min<uint64_t>(a.groups, need) : need;
    const bool one_pass = a.planes == 1 && groups == need;
    if (one_pass && a.split && d.bulk_waves && d.bnd_waves) {
      // bulk planes first, then the planes with boundary sites
      const uint32_t counts[2] = {d.n_bulk_waves, d.n_bnd_waves};
      const uint32_t* lists[2] = {d.bulk_waves, d.bnd_waves};
      for (int part = 0; part < 2; ++part) {
        if (!counts[part]) continue;
        uint64_t gp = ((uint64_t)counts[part] + wpg - 1) / wpg;
        if (a.xcd_remap >= 2) {
          const uint64_t m = 8ull * a.xcd_remap;
          gp = (gp + m - 1) / m * m;
        }
        const dim3 gridp((unsigned)gp), blockp(a.threads);
        if (part == 0)
          hipLaunchKernelGGL((k_step_list<Q, K, S, Lay, true>),
                             gridp, blockp, 0, (hipStream_t)stream, d, a, (const S*)fin,
                             (S*)fout, lists[part], counts[part]);
        else
          hipLaunchKernelGGL((k_step_list<Q, K, S, Lay, false>),
                             gridp, blockp, 0, (hipStream_t)stream, d, a, (const S*)fin,
                             (S*)fout, lists[part], counts[part]);
      }
      return (int)hipGetLastError();
    }
    if (a.xcd_remap >= 2) {
      // runs of a.xcd_remap workgroups per XCD: pad the grid to whole
      // groups of 8 runs; padded workgroups see planes >= nwaves and exit
      const uint64_t m = 8ull * a.xcd_remap;
      groups = (groups + m - 1) / m * m;
    }
    const dim3 grid((unsigned)std::max<uint64_t>(groups, 1)), block(a.threads);
    if (one_pass) {
      if constexpr (std::is_same<S, _Float16>::value)
        hipLaunchKernelGGL((k_step_w7<Q, K, S, Lay>), grid,
                           block, 0, (hipStream_t)stream, d, a, (const S*)fin, (S*)fout);
      else
        hipLaunchKernelGGL((k_step<Q, K, S, Lay>), grid,
                           block, 0, (hipStream_t)stream, d, a, (const S*)fin, (S*)fout);
    } else
      hipLaunchKernelGGL((k_step_loop<Q, K, S, Lay>), grid,
                         block, 0, (hipStream_t)stream, d, a, (const S*)fin, (S*)fout);
  } else {
    hipLaunchKernelGGL((k_step<Q, K, S, Lay>),
                       grid_for(thread_count(d)), dim3(256), 0, (hipStream_t)stream, d, a,
                       (const S*)fin, (S*)fout);
  }
  return (int)hipGetLastError();
}

// Whether every state element of the block layout is reachable by a 32-bit
// byte offset (and the launch shape does not ask for 64-bit addressing).
inline bool narrow_offsets(const DevDomain& d, const StepArgs& a, size_t esize) {
  return !a.wide && (uint64_t)a.Q * (uint64_t)d.np * esize <= 0xFFFFFFFFull;
}

}  // namespace

int launch_step(const DevDomain& d, const StepArgs& a, const void* fin, void* fout,
                void* stream) {
  return dispatch(d, a, [&](auto q, auto k, auto* s, auto* lay) {
    using S = std::remove_pointer_t<decltype(s)>;
    using Lay = std::remove_pointer_t<decltype(lay)>;
    constexpr int Q = decltype(q)::value, K = decltype(k)::value;
    if constexpr (std::is_same<Lay, BlockLayout>::value) {
      if (narrow_offsets(d, a, sizeof(S)))
        return step_launch<Q, K, S, BlockLayout32>(d, a, fin, fout, stream);
    }
    return step_launch<Q, K, S, Lay>(d, a, fin, fout, stream);
  });
}

int launch_collide_only(const DevDomain& d, const StepArgs& a, const void* fin,
                        void* fout, void* stream) {
  return dispatch(d, a, [&](auto q, auto k, auto* s, auto* lay) {
    using S = std::remove_pointer_t<decltype(s)>;
    using Lay = std::remove_pointer_t<decltype(lay)>;
    hipLaunchKernelGGL((k_collide_only<decltype(q)::value, decltype(k)::value, S, Lay>),
                       grid_for(thread_count(d)), dim3(256), 0, (hipStream_t)stream, d, a,
                       (const S*)fin, (S*)fout);
    return (int)hipGetLastError();
  });
}

int launch_stream_only(const DevDomain& d, const StepArgs& a, const void* fin,
                       double* s_out, int from_post_stream, void* stream) {
  return dispatch(d, a, [&](auto q, auto, auto* s, auto* lay) {
    using S = std::remove_pointer_t<decltype(s)>;
    using Lay = std::remove_pointer_t<decltype(lay)>;
    hipLaunchKernelGGL((k_stream_only<decltype(q)::value, S, Lay>), grid_for(thread_count(d)),
                       dim3(256), 0, (hipStream_t)stream, d, a, (const S*)fin, s_out,
                       from_post_stream);
    return (int)hipGetLastError();
  });
}

int launch_import(const DevDomain& d, const StepArgs& a, const double* aos,
                  const int64_t* perm, void* state, void* stream) {
  return dispatch(d, a, [&](auto q, auto, auto* s, auto* lay) {
    using S = std::remove_pointer_t<decltype(s)>;
    using Lay = std::remove_pointer_t<decltype(lay)>;
    hipLaunchKernelGGL((k_import<decltype(q)::value, S, Lay>), grid_for(thread_count(d)),
                       dim3(256), 0, (hipStream_t)stream, d, aos, perm, (S*)state, a.rho0);
    return (int)hipGetLastError();
  });
}

int launch_import_equilibrium(const DevDomain& d, const StepArgs& a, const double* rho,
                              const double* mom, const int64_t* perm, void* state,
                              void* stream) {
  return dispatch(d, a, [&](auto q, auto, auto* s, auto* lay) {
    using S = std::remove_pointer_t<decltype(s)>;
    using Lay = std::remove_pointer_t<decltype(lay)>;
    hipLaunchKernelGGL((k_import_equilibrium<decltype(q)::value, S, Lay>),
                       grid_for(thread_count(d)), dim3(256), 0, (hipStream_t)stream, d, rho,
                       mom, perm, (S*)state, a.rho0);
    return (int)hipGetLastError();
  });
}

int launch_soa_to_aos(int Q, int64_t n, int64_t np, const double* soa,
                      const int64_t* perm, double* aos, void* stream) {
  hipLaunchKernelGGL(k_soa_to_aos, grid_for(n), dim3(256), 0, (hipStream_t)stream,
                     Q, n, np, soa, perm, aos);
  return (int)hipGetLastError();
}

int launch_moments(int Q, int64_t n, int64_t np, const double* soa, double* rho,
                   double* u, void* stream) {
  if (Q == 15)
    hipLaunchKernelGGL(k_moments<15>, grid_for(n), dim3(256), 0,
                       (hipStream_t)stream, n, np, soa, rho, u);
  else
    hipLaunchKernelGGL(k_moments<19>, grid_for(n), dim3(256), 0,
                       (hipStream_t)stream, n, np, soa, rho, u);
  return (int)hipGetLastError();
}

int launch_wall_shear_stress(int Q, int64_t nw, int64_t np, const double* soa,
                             const int64_t* site, const float* normal, const double* temp,
                             double* out, void* stream) {
  if (nw == 0) return 0;
  if (Q == 15)
    hipLaunchKernelGGL(k_wall_shear_stress<15>, grid_for(nw), dim3(256), 0,
                       (hipStream_t)stream, nw, np, soa, site, normal, temp, out);
  else
    hipLaunchKernelGGL(k_wall_shear_stress<19>, grid_for(nw), dim3(256), 0,
                       (hipStream_t)stream, nw, np, soa, site, normal, temp, out);
  return (int)hipGetLastError();
}

int launch_reduce(int Q, int64_t n, int64_t np, const double* soa,
                  const double* rho, const double* u, double* partials,
                  int nblocks, void* stream) {
  if (Q == 15)
    hipLaunchKernelGGL(k_reduce<15>, dim3(nblocks), dim3(256), 0,
                       (hipStream_t)stream, n, np, soa, rho, u, partials);
  else
    hipLaunchKernelGGL(k_reduce<19>, dim3(nblocks), dim3(256), 0,
                       (hipStream_t)stream, n, np, soa, rho, u, partials);
  return (int)hipGetLastError();
}

int launch_copy_arrays(int Q, int64_t stride_bytes, int64_t bytes, int64_t run_bytes,
                       int width, const void* in, void* out, uint32_t threads, uint32_t remap,
                       void* stream) {
  if (width != 4 && width != 8 && width != 16) return (int)hipErrorInvalidValue;
  if (stride_bytes % width || bytes % width || run_bytes % width)
    return (int)hipErrorInvalidValue;
  if (run_bytes && bytes % run_bytes) return (int)hipErrorInvalidValue;
  const int64_t run = run_bytes / width;
  if (threads != 64 && threads != 128 && threads != 256) return (int)hipErrorInvalidValue;
  const int64_t words = bytes / width, stride = stride_bytes / width;
  uint64_t groups = ((uint64_t)words + threads - 1) / threads;
  if (remap >= 2) {
    const uint64_t m = 8ull * remap;
    groups = (groups + m - 1) / m * m;
  }
  const dim3 grid((unsigned)std::max<uint64_t>(groups, 1)), block(threads);
  auto go = [&](auto q, auto* w) {
    using W = std::remove_pointer_t<decltype(w)>;
    hipLaunchKernelGGL((k_copy_arrays<decltype(q)::value, W>), grid, block, 0,
                       (hipStream_t)stream, stride, words, run, (const W*)in, (W*)out, remap);
  };
  if (Q == 15) {
    if (width == 4) go(std::integral_constant<int, 15>{}, (unsigned int*)nullptr);
    else if (width == 8) go(std::integral_constant<int, 15>{}, (u32x2*)nullptr);
    else go(std::integral_constant<int, 15>{}, (u32x4*)nullptr);
  } else if (Q == 19) {
    if (width == 4) go(std::integral_constant<int, 19>{}, (unsigned int*)nullptr);
    else if (width == 8) go(std::integral_constant<int, 19>{}, (u32x2*)nullptr);
    else go(std::integral_constant<int, 19>{}, (u32x4*)nullptr);
  } else {
    return (int)hipErrorInvalidValue;
  }
  return (int)hipGetLastError();
}

int launch_halo(int64_t count, const int64_t* idx, int esize, const void* from, void* to,
                int pack, void* stream) {
  if (count == 0) return 0;
  const dim3 grid = grid_for(count), block(256);
  auto go = [&](auto* e) {
    using E = std::remove_pointer_t<decltype(e)>;
    if (pack)
      hipLaunchKernelGGL(k_halo_pack<E>, grid, block, 0, (hipStream_t)stream, count, idx,
                         (const E*)from, (E*)to);
    else
      hipLaunchKernelGGL(k_halo_unpack<E>, grid, block, 0, (hipStream_t)stream, count, idx,
                         (const E*)from, (E*)to);
  };
  if (esize == 8) go((uint64_t*)nullptr);
  else if (esize == 4) go((uint32_t*)nullptr);
  else if (esize == 2) go((uint16_t*)nullptr);
  else return (int)hipErrorInvalidValue;
  return (int)hipGetLastError();
}

int scratch_address(int device, double** out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static double* addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    // the caller has made `device` current: this is its copy of the array
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    addr[device] = (double*)p;
  }
  *out = addr[device];
  return 0;
}

int launch_monitor(const DevDomain& d, const StepArgs& a, const void* fin,
                   int from_post_stream, double* partials, double* io_vals, void* stream) {
  return dispatch(d, a, [&](auto q, auto, auto* s, auto* lay) {
    using S = std::remove_pointer_t<decltype(s)>;
    using Lay = std::remove_pointer_t<decltype(lay)>;
    hipLaunchKernelGGL((k_monitor<decltype(q)::value, S, Lay>), dim3(kMonitorGroups),
                       dim3(256), 0, (hipStream_t)stream, d, a, (const S*)fin,
                       from_post_stream, partials, io_vals);
    return (int)hipGetLastError();
  });
}

int launch_stress_fields(int Q, int64_t n, int64_t np, const double* soa, double tau,
                         const double* tau_site, double* t6, double* vm, double* sr,
                         void* stream) {
  if (n == 0) return 0;
  if (Q == 15)
    hipLaunchKernelGGL(k_stress_fields<15>, grid_for(n), dim3(256), 0, (hipStream_t)stream, n,
                       np, soa, tau, tau_site, t6, vm, sr);
  else
    hipLaunchKernelGGL(k_stress_fields<19>, grid_for(n), dim3(256), 0, (hipStream_t)stream, n,
                       np, soa, tau, tau_site, t6, vm, sr);
  return (int)hipGetLastError();
}

int launch_traction(int Q, int64_t nw, int64_t np, const double* soa, const int64_t* site,
                    const float* normal, const double* tau_of, double* tr, double* tg,
                    void* stream) {
  if (nw == 0) return 0;
  if (Q == 15)
    hipLaunchKernelGGL(k_traction<15>, grid_for(nw), dim3(256), 0, (hipStream_t)stream, nw, np,
                       soa, site, normal, tau_of, tr, tg);
  else
    hipLaunchKernelGGL(k_traction<19>, grid_for(nw), dim3(256), 0, (hipStream_t)stream, nw, np,
                       soa, site, normal, tau_of, tr, tg);
  return (int)hipGetLastError();
}

int launch_reduce_ranges(int Q, int64_t nchunks, int64_t np, const int64_t* range,
                         const double* soa, const double* rho, const double* u,
                         double* partials, void* stream) {
  if (nchunks == 0) return 0;
  if (Q == 15)
    hipLaunchKernelGGL(k_reduce_ranges<15>, dim3((unsigned)nchunks), dim3(256), 0,
                       (hipStream_t)stream, np, range, soa, rho, u, partials);
  else
    hipLaunchKernelGGL(k_reduce_ranges<19>, dim3((unsigned)nchunks), dim3(256), 0,
                       (hipStream_t)stream, np, range, soa, rho, u, partials);
  return (int)hipGetLastError();
}

int launch_velocity_change_ranges(int64_t nchunks, int64_t np, const int64_t* range,
                                  const double* u, const double* u_ref,
                                  const double* inv_ref, double* partials, void* stream) {
  if (nchunks == 0) return 0;
  hipLaunchKernelGGL(k_velocity_change_ranges, dim3((unsigned)nchunks), dim3(256), 0,
                     (hipStream_t)stream, np, range, u, u_ref, inv_ref, partials);
  return (int)hipGetLastError();
}

int launch_velocity_change(int64_t n, int64_t np, const double* u,
                           const double* u_ref, double inv_ref, double* partials,
                           int nblocks, void* stream) {
  hipLaunchKernelGGL(k_velocity_change, dim3(nblocks), dim3(256), 0,
                     (hipStream_t)stream, n, np, u, u_ref, inv_ref, partials);
  return (int)hipGetLastError();
}

int launch_gather_pairs(int64_t count, int64_t np, int Q, int blockwise,
                        const int64_t* slot_index, const void* state, int storage, double rho0,
                        const double* w, double* out, void* stream) {
  if (storage == 0)
    hipLaunchKernelGGL(k_gather_pairs<double>, grid_for(count), dim3(256), 0,
                       (hipStream_t)stream, count, np, Q, blockwise, slot_index,
                       (const double*)state, rho0, w, out);
  else if (storage == 1)
    hipLaunchKernelGGL(k_gather_pairs<float>, grid_for(count), dim3(256), 0,
                       (hipStream_t)stream, count, np, Q, blockwise, slot_index,
                       (const float*)state, rho0, w, out);
  else
    hipLaunchKernelGGL(k_gather_pairs<_Float16>, grid_for(count), dim3(256), 0,
                       (hipStream_t)stream, count, np, Q, blockwise, slot_index,
                       (const _Float16*)state, rho0, w, out);
  return (int)hipGetLastError();
}

int launch_scatter_pairs(int64_t count, int64_t np, int Q, int blockwise,
                         const int64_t* slot_index, const double* in, int storage, double rho0,
                         const double* w, void* state, void* stream) {
  if (storage == 0)
    hipLaunchKernelGGL(k_scatter_pairs<double>, grid_for(count), dim3(256), 0,
                       (hipStream_t)stream, count, np, Q, blockwise, slot_index, in, rho0, w,
                       (double*)state);
  else if (storage == 1)
    hipLaunchKernelGGL(k_scatter_pairs<float>, grid_for(count), dim3(256), 0,
                       (hipStream_t)stream, count, np, Q, blockwise, slot_index, in, rho0, w,
                       (float*)state);
  else
    hipLaunchKernelGGL(k_scatter_pairs<_Float16>, grid_for(count), dim3(256), 0,
                       (hipStream_t)stream, count, np, Q, blockwise, slot_index, in, rho0, w,
                       (_Float16*)state);
  return (int)hipGetLastError();
}

}  // namespace hlb
