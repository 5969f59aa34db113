// Host side of the MI355X HemeLB timestep behind the C ABI in hemelb_amd.h.
// Built with g++ (-ffp-contract=off) so that the boundary values computed on
// the host (cosine densities, Womersley profiles via the complex Bessel
// series) use the same compiler runtime and libm as the reference's g++
// build and its CPU oracle.
#include "hemelb_amd.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "lbm_internal.h"

namespace {

thread_local std::string g_err;

constexpr double kPi = 3.14159265358979323846264338327950288;  // constants.h:18
constexpr double kCs2 = 1.0 / 3.0;                             // constants.h:41
constexpr double kMmHgToPa = 133.3223874;                      // constants.h:20

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_OR_FAIL(expr)                                                    \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess)                                                    \
      return fail(HLB_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// Lattice tables (Code/lb/lattices/D3Q15.h:15-56, D3Q19.h:14-49).
const int kC15[15][3] = {{0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},
                         {0, -1, 0},  {0, 0, 1},   {0, 0, -1}, {1, 1, 1},
                         {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1},
                         {-1, 1, -1}, {1, -1, -1}, {-1, 1, 1}};
const double kW15[15] = {2.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
                         1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 72.0,
                         1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0,
                         1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0};
const int kC19[19][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},
                         {0, -1, 0}, {0, 0, 1},  {0, 0, -1}, {1, 1, 0},
                         {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1},
                         {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {0, 1, 1},
                         {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
const double kW19[19] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                         1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0,
                         1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                         1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                         1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
// d'Humieres reduced moment bases (DHumieresD3Q15MRTBasis.h:41-53,
// DHumieresD3Q19MRTBasis.h:38-53) and relaxation rates (*.cc SetUpCollisionMatrix).
const int kM15[11][15] = {
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
const int kM19[15][19] = {
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

struct LatticeHost {
  int Q;
  int c[19][3];
  double w[19];
  explicit LatticeHost(int q) : Q(q) {
    for (int i = 0; i < Q; ++i) {
      for (int a = 0; a < 3; ++a) c[i][a] = Q == 15 ? kC15[i][a] : kC19[i][a];
      w[i] = Q == 15 ? kW15[i] : kW19[i];
    }
  }
  static int inv(int d) { return d == 0 ? 0 : ((d & 1) ? d + 1 : d - 1); }
};

// Lattice::CalculateFeq, SSE3 variant (Lattice.h:236-321).
void feq_host(const LatticeHost& L, double rho, const double m[3], double* feq) {
  const double mx = m[0], my = m[1], mz = m[2];
  const double three_halves = (3. / 2.) * ((mx * mx + my * my) + mz * mz);
  const double t1 = rho - three_halves / rho;
  const double rho_1 = 1. / rho;
  double nine = (9. / 2.);
  nine *= rho_1;
  for (int i = 0; i < L.Q; ++i) {
    const double e = ((double)L.c[i][0] * mx + (double)L.c[i][1] * my) +
                     (double)L.c[i][2] * mz;
    feq[i] = L.w[i] * ((t1 + nine * (e * e)) + 3. * e);
  }
}

// util::Dot over Vector3D<double> (Vector3D.h:583-588).
double dot3(const double* a, const double* b) {
  return std::inner_product(a, a + 3, b, 0.0);
}

// util::BesselJ0ComplexArgument (Code/util/Bessel.cc:14-35).
std::complex<double> bessel_j0(const std::complex<double>& z, double tolSq) {
  std::complex<double> sum(1.0, 0.0);
  double fact = 1;
  std::complex<double> zSqOver4_pow(1.0, 0.0);
  std::complex<double> term(1.0, 0.0);
  for (unsigned i = 1; term.real() * term.real() + term.imag() * term.imag() > tolSq;
       ++i) {
    fact *= i;
    zSqOver4_pow *= -0.25 * z * z;
    term = zSqOver4_pow / (fact * fact);
    sum += term;
  }
  return sum;
}

struct IoletHost {
  hlb_iolet cfg;        // normal normalised
  double min_sim_density = 0;
  int group = 0;        // 0 inlet, 1 outlet
  int bc = 0;           // hlb_iolet_bc of its group
};

// InOutLetCosine::GetDensity (InOutLetCosine.cc:26-43).
double cosine_density(const IoletHost& io, uint64_t time_step) {
  const double w = 2.0 * kPi / io.cfg.period;
  const double target = io.cfg.density_mean +
                        io.cfg.density_amp * std::cos(w * (double)time_step + io.cfg.phase);
  if ((int64_t)time_step >= io.cfg.warmup) return target;
  const double f = ((double)time_step) / ((double)io.cfg.warmup);
  return f * target + (1. - f) * io.min_sim_density;
}

// Time-independent factor of InOutLetWomersleyVelocity::GetVelocity at x
// (InOutLetWomersleyVelocity.cc:21-38):
//   A = G / (rho * omega) * (1 - J0(i^{3/2} a r / R) / J0(i^{3/2} a)),
// so that the velocity magnitude is Re(A * exp(i omega t)).
std::complex<double> womersley_coef(const IoletHost& io, const double* x) {
  static const std::complex<double> i(0, 1);
  static const std::complex<double> iPowThreeHalves = pow(i, 1.5);
  const double displ[3] = {x[0] - io.cfg.position[0], x[1] - io.cfg.position[1],
                           x[2] - io.cfg.position[2]};
  const double z = dot3(displ, io.cfg.normal);
  const double r = std::sqrt(dot3(displ, displ) - z * z);
  const double omega = 2.0 * kPi / io.cfg.period;
  const double density = 1.0;
  const std::complex<double> besselNumer =
      bessel_j0(iPowThreeHalves * io.cfg.womersley_number * r / io.cfg.radius, 1e-12);
  const std::complex<double> besselDenom =
      bessel_j0(iPowThreeHalves * io.cfg.womersley_number, 1e-12);
  return io.cfg.pressure_gradient_amplitude / (density * omega) *
         (1.0 - besselNumer / besselDenom);
}

// 1 - r^2/R^2 of InOutLetParabolicVelocity::GetVelocity (.cc:23-43).
double parabolic_coef(const IoletHost& io, const double* x) {
  const double displ[3] = {x[0] - io.cfg.position[0], x[1] - io.cfg.position[1],
                           x[2] - io.cfg.position[2]};
  const double z = dot3(displ, io.cfg.normal);
  const double rSq = (dot3(displ, displ) - z * z) / (io.cfg.radius * io.cfg.radius);
  return 1. - rSq;
}

uint64_t spread3(uint64_t v) {
  uint64_t x = v & 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}

// Slot of a site inside its 8^3 block: x plane major, then the in-plane
// lane, the order the block-layout kernels address (lbm_kernels.hip
// kMortonLanes): with `morton` (binary64 storage) the bits of (y, z)
// interleaved, z in bits 0, 2, 4 and y in bits 1, 3, 5; otherwise rows of z
// (y * 8 + z).
inline uint32_t plane_lane(uint64_t y, uint64_t z, bool morton) {
  if (!morton) return (uint32_t)((y << 3) | z);
  uint32_t j = 0;
  for (int b = 0; b < 3; ++b)
    j |= (uint32_t)(((z >> b) & 1u) << (2 * b)) | (uint32_t)(((y >> b) & 1u) << (2 * b + 1));
  return j;
}
inline uint32_t block_slot(uint64_t x, uint64_t y, uint64_t z, bool morton) {
  return (uint32_t)((x & 7) << 6) | plane_lane(y & 7, z & 7, morton);
}
// (y, z) inside the block of in-plane lane j
inline void lane_yz(uint32_t j, bool morton, int& y, int& z) {
  if (!morton) {
    y = (int)(j >> 3);
    z = (int)(j & 7);
    return;
  }
  y = z = 0;
  for (int b = 0; b < 3; ++b) {
    z |= (int)((j >> (2 * b)) & 1u) << b;
    y |= (int)((j >> (2 * b + 1)) & 1u) << b;
  }
}

// Absolute coordinate key of a site (21 bits per axis, offset 2^20).
inline uint64_t coord_key(const int32_t* c) {
  return ((uint64_t)((int64_t)c[0] + (1 << 20)) << 42) |
         ((uint64_t)((int64_t)c[1] + (1 << 20)) << 21) | (uint64_t)((int64_t)c[2] + (1 << 20));
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t count = 0;
  ~DevBuf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t n) {
    count = n;
    return hipMalloc((void**)&p, sizeof(T) * (n ? n : 1));
  }
  hipError_t upload(const std::vector<T>& v) {
    hipError_t e = alloc(v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
  }
  size_t bytes() const { return p ? sizeof(T) * (count ? count : 1) : 0; }
};

// Iolets per solver (an ensemble has two per instance); hlb_monitor
// reports the fluxes of the first 64.
constexpr int kMaxIolets = 4096;

size_t storage_bytes(int storage) { return storage == 0 ? 8 : storage == 1 ? 4 : 2; }

}  // namespace

struct hlb_solver {
  hlb_params p{};
  int Q = 19;
  int64_t n = 0, np = 0, npc = 0;
  std::vector<int64_t> perm;      // internal -> caller
  std::vector<int64_t> internal;  // caller -> internal
  std::vector<IoletHost> iolets;  // inlets then outlets
  int n_inlets = 0;
  // device data
  DevBuf<int32_t> src;
  DevBuf<uint32_t> binfo, b_wall, b_iolet;
  DevBuf<int32_t> b_io, io_iolet, io_bc, io_site, iolet_kind;
  DevBuf<float> b_q, iolet_nf;
  DevBuf<uint32_t> b_qoff;
  DevBuf<uint16_t> b_q16;
  DevBuf<double> io_stash, io_ladd, iolet_n, weights;
  DevBuf<int32_t> bnbr;
  DevBuf<uint64_t> wave_fluid, wave_bnd, wave_links;
  DevBuf<uint32_t> wave_rec, wave_cbase, bulk_waves, bnd_waves;
  std::vector<int64_t> slot_of;   // compact -> state index (slot or compact)
  DevBuf<int64_t> d_perm;
  void* state[2] = {nullptr, nullptr};
  int cur = 0;
  DevBuf<double> bv[2];
  double* bv_host[2] = {nullptr, nullptr};
  hipEvent_t bv_done[2] = {nullptr, nullptr};
  int bv_cap = 0;
  int bv_which = 0;
  // scratch for exports and monitors, allocated on first use
  DevBuf<double> s_soa, s_aos, rho, u, u_snap, partials, gathered, mon_io;
  // this device's copy of the library's static scratch (hlb_reduce_noalloc)
  double* scratch = nullptr;
  DevBuf<int64_t> slots;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hlb::DevDomain dom;
  hlb::StepArgs args;
  // ensemble: instances, per-block relaxation, runs of one instance in the
  // internal order
  struct Run { int64_t start, end; int32_t inst; };
  int32_t n_inst = 1;
  std::vector<double> inst_tau;
  std::vector<Run> runs;
  DevBuf<double> blk_omega, blk_s_tau;
  DevBuf<float> blk_omega_f;
  std::vector<double> wall_temp, wall_tau;
  std::vector<int32_t> site_inst;  // instance of each internal site (ensembles)
  DevBuf<double> d_wall_temp, d_wall_tau, d_site_tau, d_t6, d_vm, d_sr, d_tr, d_tg;
  // wall sites (internal compact index) and their normals, for the wall
  // shear stress output
  bool has_normals = false;
  std::vector<int64_t> wall_sites;
  std::vector<float> wall_normals;
  DevBuf<int64_t> d_wall_sites;
  DevBuf<float> d_wall_normals;
  DevBuf<double> d_wss;
  // subdomains (hlb_create_subdomain): ghost sites in internal order n..,
  // their cut-link masks and coordinates, and every owned or ghost site by
  // absolute coordinate key (sorted) for the halo plans
  bool subdomain = false;
  std::vector<uint32_t> ghost_cut;
  std::vector<int32_t> ghost_coords;
  std::vector<std::pair<uint64_t, int64_t>> site_by_key;
  std::vector<int64_t> io_sites_internal;  // per iolet-site record
  std::vector<int32_t> io_sites_iolet;
  uint64_t t_ref = 1;
  bool post_stream = true;
  bool initialised = false;

  ~hlb_solver() {
    for (auto& s : state)
      if (s) (void)hipFree(s);
    for (auto& h : bv_host)
      if (h) (void)hipHostFree(h);
    for (auto& e : bv_done)
      if (e) (void)hipEventDestroy(e);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int64_t device_bytes() const {
    int64_t b = 2 * (int64_t)storage_bytes(p.storage) * Q * np;
    b += src.bytes() + binfo.bytes() + b_wall.bytes() + b_iolet.bytes() +
         b_io.bytes() + io_iolet.bytes() + io_bc.bytes() + io_site.bytes() +
         iolet_kind.bytes() + b_q.bytes() + b_qoff.bytes() + b_q16.bytes() + iolet_nf.bytes() + io_stash.bytes() +
         io_ladd.bytes() + iolet_n.bytes() + weights.bytes() + d_perm.bytes() +
         bnbr.bytes() + wave_fluid.bytes() + wave_bnd.bytes() + wave_links.bytes() +
         wave_rec.bytes() +
         wave_cbase.bytes() + bulk_waves.bytes() + bnd_waves.bytes() +
         s_soa.bytes() + s_aos.bytes() + rho.bytes() + u.bytes() + u_snap.bytes();
    return b;
  }

  // Boundary values of reference step t (1-indexed) for every iolet: the
  // Nash density of Get0IndexedTimeStep() = t-1 (BoundaryValues.cc:162-165)
  // or the Ladd velocity factor at GetTimeStep() = t (LaddIolet.h:52).
  void boundary_values(uint64_t t, double* out) const {
    for (size_t k = 0; k < iolets.size(); ++k) {
      const IoletHost& io = iolets[k];
      double a = 0, b = 0;
      if (io.bc == HLB_IOLET_NASH) {
        a = io.cfg.kind == HLB_IOLET_COSINE ? cosine_density(io, t - 1) : 1.0;
      } else if (io.cfg.kind == HLB_IOLET_WOMERSLEY) {
        static const std::complex<double> i(0, 1);
        const unsigned int ts = (unsigned int)t;  // BoundaryValues::GetTimeStep
        const double omega = 2.0 * kPi / io.cfg.period;
        const std::complex<double> e = exp(i * omega * double(ts));
        a = e.real();
        b = e.imag();
      } else {
        const unsigned int ts = (unsigned int)t;
        double mx = io.cfg.max_speed;
        if ((int64_t)ts < io.cfg.warmup) mx *= ts / double(io.cfg.warmup);
        a = mx;
      }
      out[2 * k] = a;
      out[2 * k + 1] = b;
    }
  }

  // The block layout stores each block's directions together
  // ([block][direction][512 slots]); the others one array per direction.
  int blockwise() const { return p.layout == HLB_LAYOUT_BLOCKS ? 1 : 0; }
  int64_t state_index(int d, int64_t slot) const {
    return blockwise() ? ((slot >> 9) * Q + d) * 512 + (slot & 511) : d * np + slot;
  }

  int ensure_scratch_soa() {
    if (!s_soa.p) HIP_OR_FAIL(s_soa.alloc((size_t)Q * npc));
    return HLB_OK;
  }
  int ensure_moments() {
    if (!rho.p) HIP_OR_FAIL(rho.alloc((size_t)npc));
    if (!u.p) HIP_OR_FAIL(u.alloc((size_t)3 * npc));
    return HLB_OK;
  }

  // The reference's fOld as binary64 SoA in s_soa.
  int post_stream_soa() {
    int r = ensure_scratch_soa();
    if (r) return r;
    int e = hlb::launch_stream_only(dom, args_for_stream(), state[cur], s_soa.p,
                                    post_stream ? 1 : 0, stream);
    if (e) return fail(HLB_E_DEVICE, "stream-only kernel launch failed");
    return HLB_OK;
  }

  hlb::StepArgs args_for_stream() {
    hlb::StepArgs a = args;
    // the pending stream belongs to reference step t_ref - 1
    if (!post_stream && !iolets.empty()) {
      // stage through the first pinned and device bv buffers, synchronously
      // (exports and monitors are rare); once the stream is idle no step's
      // copy still reads the pinned buffer
      (void)hipStreamSynchronize(stream);
      boundary_values(t_ref - 1, bv_host[0]);
      (void)hipMemcpy(bv[0].p, bv_host[0], sizeof(double) * 2 * iolets.size(),
                      hipMemcpyHostToDevice);
      a.bv = bv[0].p;
    }
    return a;
  }
};

extern "C" {

int hlb_abi_version(void) { return HLB_ABI_VERSION; }
const char* hlb_last_error(void) { return g_err.c_str(); }

double hlb_tau_from_physical(double dt, double dx, double rho, double eta) {
  return 0.5 + (dt * eta / rho) / (kCs2 * dx * dx);  // LbmParameters.h:159
}

static double lattice_pressure(double dt, double dx, double rho) {
  const double mass = rho * dx * dx * dx;  // UnitConverter.cc:14-22
  return mass / (dx * dt * dt);
}

double hlb_pressure_difference_to_lattice(double dp, double dt, double dx, double rho) {
  return dp * kMmHgToPa / lattice_pressure(dt, dx, rho);
}

double hlb_pressure_to_lattice(double p, double ref, double dt, double dx, double rho) {
  return kCs2 + hlb_pressure_difference_to_lattice(p - ref, dt, dx, rho);
}

double hlb_pressure_gradient_to_lattice(double pg, double dt, double dx, double rho) {
  return pg * kMmHgToPa * dx / lattice_pressure(dt, dx, rho);
}

// Builds a solver; `instance` (optional, [n]) assigns every site to one of
// n_inst independent vessels relaxed with inst_tau[k] (an ensemble); `gh`
// (optional) lists ghost sites: sites of neighbouring subdomains whose
// post-collision values this subdomain's sites pull across the cut.  Ghost
// sites get state slots that the pull reads and the halo exchange fills;
// the step never updates them.
static int create_impl(const hlb_params* params, const hlb_domain* dm, const hlb_domain* gh,
                       const int32_t* instance, int32_t n_inst, const double* inst_tau,
                       const hlb_iolet* inlets, int32_t n_inlets,
                       const hlb_iolet* outlets, int32_t n_outlets, hlb_solver** out) {
  if (!params || !dm || !out) return fail(HLB_E_INVALID, "null argument");
  *out = nullptr;
  const hlb_params& P = *params;
  if (P.lattice != 15 && P.lattice != 19)
    return fail(HLB_E_INVALID, "lattice must be HLB_D3Q15 or HLB_D3Q19");
  if (P.kernel != HLB_LBGK && P.kernel != HLB_MRT)
    return fail(HLB_E_INVALID, "unknown collision kernel");
  if (P.wall != HLB_WALL_SBB && P.wall != HLB_WALL_BFL)
    return fail(HLB_E_INVALID, "unknown wall streamer");
  for (int bc : {P.inlet_bc, P.outlet_bc})
    if (bc != HLB_IOLET_NASH && bc != HLB_IOLET_LADD)
      return fail(HLB_E_INVALID, "unknown iolet streamer");
  if (P.storage < 0 || P.storage > 2) return fail(HLB_E_INVALID, "unknown storage");
  if (P.layout != HLB_LAYOUT_BLOCKS && P.layout != HLB_LAYOUT_PULL_TABLE &&
      P.layout != HLB_LAYOUT_COMPACT_BLOCKS)
    return fail(HLB_E_INVALID, "unknown layout");
  if (!(P.tau > 0.0) || !std::isfinite(P.tau)) return fail(HLB_E_INVALID, "tau must be positive");
  if (P.storage != HLB_STORE_FP64 && !(P.rho0 > 0.0))
    return fail(HLB_E_INVALID, "rho0 must be positive for reduced-precision storage");
  const int64_t n = dm->n_sites;
  if (n <= 0 || n >= (int64_t)0xFFFFFFF0ll) return fail(HLB_E_INVALID, "n_sites out of range");
  if (!dm->coords || !dm->site_type || !dm->wall_mask || !dm->iolet_mask || !dm->iolet_id)
    return fail(HLB_E_INVALID, "domain arrays missing");
  if (P.wall == HLB_WALL_BFL && !dm->cut_distance)
    return fail(HLB_E_INVALID, "BFL walls need cut distances");
  if (n_inlets < 0 || n_outlets < 0 || n_inlets + n_outlets > kMaxIolets)
    return fail(HLB_E_INVALID, "between 0 and " + std::to_string(kMaxIolets) +
                                   " iolets are supported");
  if (instance) {
    if (P.layout == HLB_LAYOUT_PULL_TABLE)
      return fail(HLB_E_INVALID, "ensembles need a block layout");
    if (n_inst < 1 || !inst_tau) return fail(HLB_E_INVALID, "ensemble needs instances and their tau");
    for (int32_t k = 0; k < n_inst; ++k)
      if (!(inst_tau[k] > 0.0) || !std::isfinite(inst_tau[k]))
        return fail(HLB_E_INVALID, "instance tau must be positive");
    for (int64_t i = 0; i < dm->n_sites; ++i)
      if (instance[i] < 0 || instance[i] >= n_inst)
        return fail(HLB_E_INVALID, "site " + std::to_string(i) + ": instance out of range");
  }

  const int64_t ng = gh ? gh->n_sites : 0;
  if (gh) {
    if (ng < 0 || ng >= (int64_t)0x7FFFFFF0ll - n) return fail(HLB_E_INVALID, "ghost count out of range");
    if (ng && (!gh->coords || !gh->wall_mask || !gh->iolet_mask))
      return fail(HLB_E_INVALID, "ghost arrays missing");
    if (instance) return fail(HLB_E_INVALID, "ensembles take no ghost sites");
    if (P.layout == HLB_LAYOUT_COMPACT_BLOCKS)
      return fail(HLB_E_INVALID, "ghost sites need the block or pull-table layout");
  }

  auto s = std::make_unique<hlb_solver>();
  s->p = P;
  const int Q = P.lattice;
  s->Q = Q;
  s->n = n;
  s->npc = (n + 255) / 256 * 256;
  // Both block layouts build the wave tables; the compact one stores the
  // distributions by compact site instead of by block slot.
  const bool blocks_layout = P.layout == HLB_LAYOUT_BLOCKS || P.layout == HLB_LAYOUT_COMPACT_BLOCKS;
  const bool compact_store = P.layout != HLB_LAYOUT_BLOCKS;
  LatticeHost L(Q);
  const uint32_t all = (1u << (Q - 1)) - 1u;

  // Iolets: normalised normals (InOutLet::SetNormal, InOutLet.h:160-163) and
  // the minimum density over all of them (LBM::PrepareBoundaryObjects).
  double min_density = 1.7976931348623157e308;
  for (int g = 0; g < 2; ++g) {
    const hlb_iolet* src = g ? outlets : inlets;
    const int cnt = g ? n_outlets : n_inlets;
    if (cnt && !src) return fail(HLB_E_INVALID, "iolet array missing");
    for (int k = 0; k < cnt; ++k) {
      IoletHost io;
      io.cfg = src[k];
      io.group = g;
      io.bc = g ? P.outlet_bc : P.inlet_bc;
      if (io.cfg.kind < 0 || io.cfg.kind > 2) return fail(HLB_E_INVALID, "unknown iolet kind");
      const double mag = std::sqrt(dot3(src[k].normal, src[k].normal));
      if (!(mag > 0)) return fail(HLB_E_INVALID, "iolet normal is zero");
      for (int a = 0; a < 3; ++a) io.cfg.normal[a] = src[k].normal[a] / mag;
      if (io.bc == HLB_IOLET_LADD && io.cfg.kind == HLB_IOLET_COSINE)
        return fail(HLB_E_INVALID, "Ladd iolets need a velocity boundary (parabolic or Womersley)");
      if (io.cfg.kind == HLB_IOLET_COSINE && !(io.cfg.period != 0))
        return fail(HLB_E_INVALID, "cosine iolet period must be non-zero");
      if (io.cfg.kind != HLB_IOLET_COSINE && !(io.cfg.radius > 0))
        return fail(HLB_E_INVALID, "velocity iolet radius must be positive");
      if (io.cfg.kind == HLB_IOLET_WOMERSLEY && !(io.cfg.period != 0))
        return fail(HLB_E_INVALID, "Womersley period must be non-zero");
      const double dmin = io.cfg.kind == HLB_IOLET_COSINE
                              ? io.cfg.density_mean - io.cfg.density_amp
                              : 1.0;
      min_density = std::min(min_density, dmin);
      s->iolets.push_back(io);
    }
  }
  for (auto& io : s->iolets) io.min_sim_density = min_density;
  s->n_inlets = n_inlets;

  // Validate per-site data (SiteData semantics, SiteDataBare.cc:23-75).
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t t = dm->site_type[i], wm = dm->wall_mask[i], im = dm->iolet_mask[i];
    if (t < 1 || t > 3) return fail(HLB_E_INVALID, "site " + std::to_string(i) + ": bad site type");
    if ((wm | im) & ~all)
      return fail(HLB_E_INVALID, "site " + std::to_string(i) + ": mask has bits beyond Q-1");
    if ((t == HLB_SITE_FLUID) != (im == 0))
      return fail(HLB_E_INVALID, "site " + std::to_string(i) +
                                     ": iolet links must match an INLET/OUTLET site type");
    if (t != HLB_SITE_FLUID) {
      const int cnt = t == HLB_SITE_INLET ? n_inlets : n_outlets;
      if (dm->iolet_id[i] < 0 || dm->iolet_id[i] >= cnt)
        return fail(HLB_E_INVALID, "site " + std::to_string(i) + ": iolet id out of range");
    }
    if (P.wall == HLB_WALL_BFL && wm)
      for (int d = 1; d < Q; ++d)
        if ((wm >> (d - 1)) & 1u) {
          const float q = dm->cut_distance[i * (Q - 1) + d - 1];
          // BFL divides by 2q only for q >= 0.5 (BouzidiFirdaousLallemand.h:56-66)
          if (!std::isfinite(q))
            return fail(HLB_E_INVALID, "site " + std::to_string(i) +
                                           ": BFL wall link needs a finite cut distance");
        }
  }

  // Internal order: 8^3 blocks in Morton order; inside a block x planes,
  // then the storage's in-plane lane order (block_slot).
  const bool morton = params->storage == HLB_STORE_FP64;
  int32_t lo[3];
  for (int a = 0; a < 3; ++a) lo[a] = dm->coords[a];
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) lo[a] = std::min(lo[a], dm->coords[3 * i + a]);
  for (int64_t g = 0; g < ng; ++g)
    for (int a = 0; a < 3; ++a) lo[a] = std::min(lo[a], gh->coords[3 * g + a]);
  auto key_of = [&](int64_t i) {
    const uint64_t x = (uint64_t)(dm->coords[3 * i] - lo[0]),
                   y = (uint64_t)(dm->coords[3 * i + 1] - lo[1]),
                   z = (uint64_t)(dm->coords[3 * i + 2] - lo[2]);
    const uint64_t oct = (spread3(x >> 3) << 2) | (spread3(y >> 3) << 1) | spread3(z >> 3);
    return (oct << 9) | block_slot(x, y, z, morton);
  };
  std::vector<std::pair<uint64_t, int64_t>> keys(n);
  for (int64_t i = 0; i < n; ++i) keys[i] = {key_of(i), i};
  std::sort(keys.begin(), keys.end());
  s->perm.resize(n);
  s->internal.resize(n);
  for (int64_t o = 0; o < n; ++o) {
    if (o && keys[o].first == keys[o - 1].first)
      return fail(HLB_E_INVALID, "two sites share the same coordinates");
    s->perm[o] = keys[o].second;
    s->internal[keys[o].second] = o;
  }
  s->has_normals = dm->wall_normal != nullptr;
  for (int64_t o = 0; o < n; ++o) {
    const int64_t i = s->perm[o];
    if (!dm->wall_mask[i]) continue;
    s->wall_sites.push_back(o);
    if (s->has_normals)
      for (int a = 0; a < 3; ++a) s->wall_normals.push_back(dm->wall_normal[3 * i + a]);
  }
  // Ghost sites in Morton order, internal indices n .. n+ng-1.
  std::vector<std::pair<uint64_t, int64_t>> gkeys(ng);
  for (int64_t g = 0; g < ng; ++g) {
    const uint64_t x = (uint64_t)(gh->coords[3 * g] - lo[0]),
                   y = (uint64_t)(gh->coords[3 * g + 1] - lo[1]),
                   z = (uint64_t)(gh->coords[3 * g + 2] - lo[2]);
    const uint64_t oct = (spread3(x >> 3) << 2) | (spread3(y >> 3) << 1) | spread3(z >> 3);
    gkeys[g] = {(oct << 9) | block_slot(x, y, z, morton), g};
  }
  std::sort(gkeys.begin(), gkeys.end());
  for (int64_t g = 1; g < ng; ++g)
    if (gkeys[g].first == gkeys[g - 1].first)
      return fail(HLB_E_INVALID, "two ghost sites share the same coordinates");
  s->ghost_cut.resize(ng);
  s->ghost_coords.resize(3 * ng);
  for (int64_t g = 0; g < ng; ++g) {
    const int64_t i = gkeys[g].second;
    s->ghost_cut[g] = gh->wall_mask[i] | gh->iolet_mask[i];
    for (int a = 0; a < 3; ++a) s->ghost_coords[3 * g + a] = gh->coords[3 * i + a];
  }
  // Site lookup by coordinates: block key -> internal index of slot 0..511.
  std::unordered_map<uint64_t, std::vector<int32_t>> blocks;
  blocks.reserve(n / 64 + 16);
  auto bkey = [](int64_t bx, int64_t by, int64_t bz) {
    return ((uint64_t)(bx & 0x1fffff) << 42) | ((uint64_t)(by & 0x1fffff) << 21) |
           (uint64_t)(bz & 0x1fffff);
  };
  for (int64_t o = 0; o < n; ++o) {
    const int64_t i = s->perm[o];
    const int64_t x = dm->coords[3 * i] - lo[0], y = dm->coords[3 * i + 1] - lo[1],
                  z = dm->coords[3 * i + 2] - lo[2];
    auto& b = blocks[bkey(x >> 3, y >> 3, z >> 3)];
    if (b.empty()) b.assign(512, -1);
    b[block_slot(x, y, z, morton)] = (int32_t)o;
  }
  for (int64_t g = 0; g < ng; ++g) {
    const int64_t x = s->ghost_coords[3 * g] - lo[0], y = s->ghost_coords[3 * g + 1] - lo[1],
                  z = s->ghost_coords[3 * g + 2] - lo[2];
    auto& b = blocks[bkey(x >> 3, y >> 3, z >> 3)];
    if (b.empty()) b.assign(512, -1);
    int32_t& slot = b[block_slot(x, y, z, morton)];
    if (slot >= 0) return fail(HLB_E_INVALID, "a ghost site is also a site of the domain");
    slot = (int32_t)(n + g);
  }
  auto lookup = [&](int64_t x, int64_t y, int64_t z) -> int64_t {
    if (x < 0 || y < 0 || z < 0) return -1;
    auto it = blocks.find(bkey(x >> 3, y >> 3, z >> 3));
    if (it == blocks.end()) return -1;
    return it->second[block_slot(x, y, z, morton)];
  };

  // Blocks of the compact order (it is sorted by Morton block key).
  std::vector<int64_t> blk_of(n);
  std::vector<uint64_t> blk_keys;
  for (int64_t o = 0; o < n; ++o) {
    const uint64_t k = keys[o].first >> 9;
    if (blk_keys.empty() || blk_keys.back() != k) blk_keys.push_back(k);
    blk_of[o] = (int64_t)blk_keys.size() - 1;
  }
  const int64_t nblocks = (int64_t)blk_keys.size();
  // ghost sites outside the domain's blocks live in ghost blocks after them
  std::vector<int64_t> gblk_of(ng);
  std::vector<uint64_t> gblk_keys;
  for (int64_t g = 0; g < ng; ++g) {
    const uint64_t k = gkeys[g].first >> 9;
    auto it = std::lower_bound(blk_keys.begin(), blk_keys.end(), k);
    if (it != blk_keys.end() && *it == k) {
      gblk_of[g] = it - blk_keys.begin();
    } else {
      if (gblk_keys.empty() || gblk_keys.back() != k) gblk_keys.push_back(k);
      gblk_of[g] = nblocks + (int64_t)gblk_keys.size() - 1;
    }
  }
  const int64_t nblocks_all = nblocks + (int64_t)gblk_keys.size();
  s->np = compact_store ? (n + ng + 255) / 256 * 256 : nblocks_all * 512;
  const int64_t np = s->np;
  s->slot_of.resize(n + ng);
  for (int64_t o = 0; o < n; ++o)
    s->slot_of[o] = compact_store ? o : ((blk_of[o] << 9) | (int64_t)(keys[o].first & 511));
  for (int64_t g = 0; g < ng; ++g)
    s->slot_of[n + g] =
        compact_store ? n + g : ((gblk_of[g] << 9) | (int64_t)(gkeys[g].first & 511));
  if (gh) {
    s->subdomain = true;
    s->site_by_key.reserve(n + ng);
    for (int64_t o = 0; o < n; ++o)
      s->site_by_key.push_back({coord_key(&dm->coords[3 * s->perm[o]]), o});
    for (int64_t g = 0; g < ng; ++g)
      s->site_by_key.push_back({coord_key(&s->ghost_coords[3 * g]), n + g});
    std::sort(s->site_by_key.begin(), s->site_by_key.end());
  }

  // Pull table, boundary and iolet records.
  std::vector<int32_t> h_src;
  std::vector<uint32_t> h_binfo;
  if (!blocks_layout) {
    if (np >= ((int64_t)1 << 30))
      return fail(HLB_E_INVALID, "too many sites for the pull-table layout");
    h_src.assign((size_t)(Q - 1) * np, 0);
    h_binfo.assign(np, 0);
  }
  std::vector<uint32_t> h_bwall, h_biolet;
  std::vector<int32_t> h_bio, h_io_iolet, h_io_bc, h_io_site;
  std::vector<float> h_bq;        // binary64: the wall links' cut distances, packed
  std::vector<uint32_t> h_bqoff;  // binary64: first entry of each record in h_bq
  std::vector<uint16_t> h_bq16;  // reduced-precision storage: q * 65535
  std::vector<double> h_ladd;
  for (int64_t o = 0; o < n; ++o) {
    const int64_t i = s->perm[o];
    const int64_t x = dm->coords[3 * i] - lo[0], y = dm->coords[3 * i + 1] - lo[1],
                  z = dm->coords[3 * i + 2] - lo[2];
    const uint32_t wm = dm->wall_mask[i], im = dm->iolet_mask[i];
    for (int e = 1; e < Q; ++e) {
      const int dp = LatticeHost::inv(e);
      const bool own = ((wm | im) >> (dp - 1)) & 1u;
      const int64_t j = lookup(x - L.c[e][0], y - L.c[e][1], z - L.c[e][2]);
      // The source pushes into this slot unless its link e is cut.
      bool pushes = false;
      if (j >= n) {
        pushes = !((s->ghost_cut[j - n] >> (e - 1)) & 1u);
      } else if (j >= 0) {
        const int64_t ci = s->perm[j];
        pushes = !(((dm->wall_mask[ci] | dm->iolet_mask[ci]) >> (e - 1)) & 1u);
        if (pushes && instance && instance[ci] != instance[i])
          return fail(HLB_E_INVALID, "sites of different instances are linked");
      }
      if (own == pushes) {
        return fail(HLB_E_INVALID,
                    "site (" + std::to_string(dm->coords[3 * i]) + "," +
                        std::to_string(dm->coords[3 * i + 1]) + "," +
                        std::to_string(dm->coords[3 * i + 2]) + ") direction " +
                        std::to_string(e) +
                        (own ? ": slot written both by a cut link and by a fluid neighbour"
                             : ": slot has no writer (non-fluid neighbour without a cut link)"));
      }
      // offset inside array e: the source site, or the own slot of the
      // inverse direction (x + (inv(e) - e) * np) for a cut link
      if (!blocks_layout)
        h_src[(size_t)(e - 1) * np + o] = own ? (int32_t)(o + (int64_t)(dp - e) * np) : (int32_t)j;
    }
    if (wm | im) {
      const uint32_t rec = (uint32_t)h_bwall.size();
      if (!blocks_layout) h_binfo[o] = rec + 1;
      h_bwall.push_back(wm);
      h_biolet.push_back(im);
      // Cut distances (the BFL links read those of wall links only):
      // * binary64 storage: the reference's float q, bit for bit, of the
      //   wall links alone, packed in direction order from h_bqoff[rec];
      // * reduced-precision forms: q * 65535 rounded to 16 bits for every
      //   link, in 16-byte aligned records (a wall position error below 8e-6
      //   lattice units).
      const bool reduced = P.storage != HLB_STORE_FP64;
      if (reduced) {
        for (int d = 1; d <= (Q - 1 + 7) / 8 * 8; ++d) {
          const float q = d < Q && dm->cut_distance ? dm->cut_distance[i * (Q - 1) + d - 1]
                                                    : -1.0f;
          h_bq16.push_back((uint16_t)std::lround((double)std::min(1.0f, std::max(0.0f, q)) *
                                                 65535.0));
        }
      } else {
        h_bqoff.push_back((uint32_t)h_bq.size());
        for (int d = 1; d < Q; ++d)
          if ((wm >> (d - 1)) & 1u)
            h_bq.push_back(dm->cut_distance ? dm->cut_distance[i * (Q - 1) + d - 1] : -1.0f);
      }
      if (im) {
        const int g = dm->site_type[i] == HLB_SITE_OUTLET ? 1 : 0;
        const int k = (g ? n_inlets : 0) + dm->iolet_id[i];
        const int32_t io_rec = (int32_t)h_io_iolet.size();
        h_bio.push_back(io_rec);
        h_io_iolet.push_back(k);
        h_io_bc.push_back(s->iolets[k].bc);
        h_io_site.push_back((int32_t)o);
        s->io_sites_internal.push_back(o);
        s->io_sites_iolet.push_back(k);
        for (int d = 1; d < Q; ++d) {
          double c0 = 0, c1 = 0;
          if (s->iolets[k].bc == HLB_IOLET_LADD && ((im >> (d - 1)) & 1u)) {
            // LaddIoletLink: half-way point x + c_d / 2 (LaddIolet.h:122-125)
            double half[3];
            for (int a = 0; a < 3; ++a)
              half[a] = (double)dm->coords[3 * i + a] + (double)L.c[d][a] * 0.5;
            if (s->iolets[k].cfg.kind == HLB_IOLET_WOMERSLEY) {
              const std::complex<double> A = womersley_coef(s->iolets[k], half);
              c0 = A.real();
              c1 = A.imag();
            } else {
              c0 = parabolic_coef(s->iolets[k], half);
            }
          }
          h_ladd.push_back(c0);
          h_ladd.push_back(c1);
        }
      } else {
        h_bio.push_back(-1);
      }
    }
  }

  // Block layout tables: neighbour blocks and per-wave masks.
  std::vector<int32_t> h_bnbr;
  std::vector<uint64_t> h_wf, h_wb, h_wl;
  std::vector<uint32_t> h_bulkw, h_bndw;
  std::vector<uint32_t> h_wr, h_wc;
  if (blocks_layout) {
    // Block neighbour table (offsets (dx,dy,dz) in {-1,0,1}^3, index 13 = self)
    std::unordered_map<uint64_t, int32_t> ordinal;
    ordinal.reserve(nblocks * 2);
    std::vector<int64_t> bxyz(3 * nblocks);
    for (int64_t o = 0; o < n; ++o) {
      const int64_t i = s->perm[o];
      const int64_t b = blk_of[o];
      bxyz[3 * b] = (dm->coords[3 * i] - lo[0]) >> 3;
      bxyz[3 * b + 1] = (dm->coords[3 * i + 1] - lo[1]) >> 3;
      bxyz[3 * b + 2] = (dm->coords[3 * i + 2] - lo[2]) >> 3;
    }
    for (int64_t b = 0; b < nblocks; ++b)
      ordinal[bkey(bxyz[3 * b], bxyz[3 * b + 1], bxyz[3 * b + 2])] = (int32_t)b;
    for (int64_t g = 0; g < ng; ++g)
      if (gblk_of[g] >= nblocks)
        ordinal[bkey((s->ghost_coords[3 * g] - lo[0]) >> 3,
                     (s->ghost_coords[3 * g + 1] - lo[1]) >> 3,
                     (s->ghost_coords[3 * g + 2] - lo[2]) >> 3)] = (int32_t)gblk_of[g];
    h_bnbr.assign((size_t)nblocks * 27, -1);
    for (int64_t b = 0; b < nblocks; ++b)
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dz = -1; dz <= 1; ++dz) {
            const int64_t X = bxyz[3 * b] + dx, Y = bxyz[3 * b + 1] + dy, Z = bxyz[3 * b + 2] + dz;
            if (X < 0 || Y < 0 || Z < 0) continue;
            auto it = ordinal.find(bkey(X, Y, Z));
            if (it != ordinal.end())
              h_bnbr[(size_t)b * 27 + (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)] = it->second;
          }
    const int64_t nwaves = nblocks * 8;
    if (nwaves >= (int64_t)0xFFFFFFFFll) return fail(HLB_E_INVALID, "domain too large");
    h_wf.assign(nwaves, 0);
    h_wb.assign(nwaves, 0);
    h_wl.assign(nwaves, 0);
    h_wr.assign(nwaves, 0);
    h_wc.assign(nwaves, 0);
    uint32_t rec = 0;
    int64_t o = 0;
    // block slot of internal site o (x plane major, then block_slot's lanes)
    auto bslot = [&](int64_t q) { return (blk_of[q] << 9) | (int64_t)(keys[q].first & 511); };
    for (int64_t w = 0; w < nwaves; ++w) {
      h_wr[w] = rec;
      h_wc[w] = (uint32_t)o;
      while (o < n && (bslot(o) >> 6) == w) {
        const int lane = (int)(bslot(o) & 63);
        const int64_t i = s->perm[o];
        h_wf[w] |= 1ull << lane;
        const uint32_t wm = dm->wall_mask[i], im = dm->iolet_mask[i];
        if (wm | im) {
          h_wb[w] |= 1ull << lane;
          // the plane's BFL links (the gather's masks; directions are
          // numbered in inverse pairs (1,2), (3,4), ...)
          const auto swap = [](uint32_t v) {
            return ((v & 0x55555555u) << 1) | ((v >> 1) & 0x55555555u);
          };
          h_wl[w] |= (uint64_t)(swap(wm) & ~swap(im) & ~wm) |
                     ((uint64_t)(wm & ~swap(wm)) << 32);
          ++rec;
        }
        ++o;
      }
    }
    s->dom.nwaves = (uint32_t)nwaves;
    // planes without boundary sites and planes with them, for the split step
    for (int64_t w = 0; w < nwaves; ++w)
      if (h_wf[w]) (h_wb[w] ? h_bndw : h_bulkw).push_back((uint32_t)w);
    // Check, on the host, every neighbour the block-layout kernel will
    // compute (BlockLayout::src in lbm_kernels.hip): each bulk pull of each
    // fluid site must land on the slot of its source site.
    for (int64_t o2 = 0; o2 < n; ++o2) {
      const int64_t i = s->perm[o2];
      const uint32_t cutm = dm->wall_mask[i] | dm->iolet_mask[i];
      const int64_t slot = bslot(o2);
      const int64_t w = slot >> 6, lane = slot & 63;
      const int lx = (int)(w & 7);
      int ly, lz;
      lane_yz((uint32_t)lane, morton, ly, lz);
      const int64_t x = dm->coords[3 * i] - lo[0], y = dm->coords[3 * i + 1] - lo[1],
                    z = dm->coords[3 * i + 2] - lo[2];
      for (int e = 1; e < Q; ++e) {
        if ((cutm >> (LatticeHost::inv(e) - 1)) & 1u) continue;
        const int CX = L.c[e][0], CY = L.c[e][1], CZ = L.c[e][2];
        const int sx = lx - CX;
        const int ox = sx < 0 ? -1 : (sx > 7 ? 1 : 0);
        const bool cy = CY != 0 && (CY > 0 ? ly == 0 : ly == 7);
        const bool cz = CZ != 0 && (CZ > 0 ? lz == 0 : lz == 7);
        const int32_t* nbt = &h_bnbr[(size_t)(w >> 3) * 27];
        const int base = 13 + 9 * ox;
        const int32_t nb = cy ? (cz ? nbt[base - 3 * CY - CZ] : nbt[base - 3 * CY])
                              : (cz ? nbt[base - CZ] : nbt[base]);
        const int sl =
            (int)plane_lane((uint64_t)((ly - CY) & 7), (uint64_t)((lz - CZ) & 7), morton);
        int64_t src = ((int64_t)nb << 9) + (((sx & 7) << 6) | sl);
        if (compact_store && nb >= 0) {
          // CompactBlockLayout::offset: plane base + fluid lanes below
          const int64_t pw = ((int64_t)nb << 3) | (sx & 7);
          const uint64_t below = sl ? (h_wf[pw] & ((1ull << sl) - 1ull)) : 0ull;
          src = (int64_t)h_wc[pw] + (int64_t)__builtin_popcountll(below);
        }
        const int64_t j = lookup(x - CX, y - CY, z - CZ);
        if (nb < 0 || j < 0 || src != s->slot_of[j])
          return fail(HLB_E_INVALID, "internal error: block neighbour table inconsistent");
      }
    }
  }

  // Ensemble: every 8^3 block belongs to one instance; contiguous runs of
  // one instance in the internal order for the per-instance reductions.
  std::vector<double> h_bomega, h_bstau;
  std::vector<float> h_bomega_f;
  if (instance) {
    std::vector<int32_t> binst(nblocks, -1);
    for (int64_t o = 0; o < n; ++o) {
      const int32_t k = instance[s->perm[o]];
      int32_t& b = binst[blk_of[o]];
      if (b >= 0 && b != k) return fail(HLB_E_INVALID, "instances must not share an 8^3 block");
      b = k;
    }
    h_bomega.resize(nblocks);
    h_bomega_f.resize(nblocks);
    h_bstau.resize(nblocks);
    for (int64_t b = 0; b < nblocks; ++b) {
      h_bomega[b] = -1.0 / inst_tau[binst[b]];
      h_bomega_f[b] = (float)h_bomega[b];
      h_bstau[b] = 1.0 / inst_tau[binst[b]];
    }
    s->n_inst = n_inst;
    s->inst_tau.assign(inst_tau, inst_tau + n_inst);
    for (int64_t o = 0; o < n; ++o) {
      const int32_t k = instance[s->perm[o]];
      if (s->runs.empty() || s->runs.back().inst != k || s->runs.back().end != o)
        s->runs.push_back({o, o + 1, k});
      else
        s->runs.back().end = o + 1;
    }
  } else {
    s->n_inst = 1;
    s->inst_tau.assign(1, P.tau);
    s->runs.push_back({0, n, 0});
  }
  // wall shear stress factor -(1 - 1/(2 tau)) per wall site
  for (int64_t o : s->wall_sites) {
    const double tau = instance ? inst_tau[instance[s->perm[o]]] : P.tau;
    const double stress_parameter = (1.0 - 1.0 / (2.0 * tau)) / std::sqrt(2.0);
    s->wall_temp.push_back(stress_parameter * (-std::sqrt(2.0)));
    s->wall_tau.push_back(tau);
  }
  if (instance) {
    s->site_inst.resize(n);
    for (int64_t o = 0; o < n; ++o) s->site_inst[o] = instance[s->perm[o]];
  }

  // Device set-up.
  HIP_OR_FAIL(hipSetDevice(P.device));
  if (hlb::scratch_address(P.device, &s->scratch))
    return fail(HLB_E_DEVICE, "static scratch of the device not found");
  HIP_OR_FAIL(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  HIP_OR_FAIL(hipEventCreate(&s->ev0));
  HIP_OR_FAIL(hipEventCreate(&s->ev1));
  const size_t sb = storage_bytes(P.storage);
  for (auto& st : s->state) HIP_OR_FAIL(hipMalloc(&st, sb * Q * np));
  HIP_OR_FAIL(s->src.upload(h_src));
  HIP_OR_FAIL(s->binfo.upload(h_binfo));
  HIP_OR_FAIL(s->bnbr.upload(h_bnbr));
  HIP_OR_FAIL(s->wave_fluid.upload(h_wf));
  HIP_OR_FAIL(s->wave_bnd.upload(h_wb));
  HIP_OR_FAIL(s->wave_links.upload(h_wl));
  HIP_OR_FAIL(s->wave_rec.upload(h_wr));
  HIP_OR_FAIL(s->wave_cbase.upload(h_wc));
  HIP_OR_FAIL(s->bulk_waves.upload(h_bulkw));
  HIP_OR_FAIL(s->bnd_waves.upload(h_bndw));
  HIP_OR_FAIL(s->b_wall.upload(h_bwall));
  HIP_OR_FAIL(s->b_iolet.upload(h_biolet));
  HIP_OR_FAIL(s->b_io.upload(h_bio));
  HIP_OR_FAIL(s->b_q.upload(h_bq));
  HIP_OR_FAIL(s->b_qoff.upload(h_bqoff));
  HIP_OR_FAIL(s->b_q16.upload(h_bq16));
  HIP_OR_FAIL(s->io_iolet.upload(h_io_iolet));
  HIP_OR_FAIL(s->io_bc.upload(h_io_bc));
  HIP_OR_FAIL(s->io_site.upload(h_io_site));
  HIP_OR_FAIL(s->io_ladd.upload(h_ladd));
  HIP_OR_FAIL(s->io_stash.alloc(h_io_iolet.size()));
  HIP_OR_FAIL(hipMemset(s->io_stash.p, 0, sizeof(double) * (h_io_iolet.empty() ? 1 : h_io_iolet.size())));
  std::vector<float> h_nf;
  std::vector<double> h_n;
  std::vector<int32_t> h_kind;
  for (auto& io : s->iolets) {
    for (int a = 0; a < 3; ++a) {
      h_nf.push_back((float)io.cfg.normal[a]);
      h_n.push_back(io.cfg.normal[a]);
    }
    h_kind.push_back(io.cfg.kind);
  }
  HIP_OR_FAIL(s->iolet_nf.upload(h_nf));
  HIP_OR_FAIL(s->iolet_n.upload(h_n));
  HIP_OR_FAIL(s->iolet_kind.upload(h_kind));
  HIP_OR_FAIL(s->weights.upload(std::vector<double>(L.w, L.w + Q)));
  HIP_OR_FAIL(s->d_perm.upload(s->perm));
  if (instance) {
    HIP_OR_FAIL(s->blk_omega.upload(h_bomega));
    HIP_OR_FAIL(s->blk_omega_f.upload(h_bomega_f));
    HIP_OR_FAIL(s->blk_s_tau.upload(h_bstau));
  }
  s->bv_cap = 512;
  const size_t nio = std::max<size_t>(1, s->iolets.size());
  for (int b = 0; b < 2; ++b) {
    HIP_OR_FAIL(s->bv[b].alloc((size_t)s->bv_cap * 2 * nio));
    HIP_OR_FAIL(hipHostMalloc((void**)&s->bv_host[b], sizeof(double) * s->bv_cap * 2 * nio));
    HIP_OR_FAIL(hipEventCreateWithFlags(&s->bv_done[b], hipEventDisableTiming));
    HIP_OR_FAIL(hipEventRecord(s->bv_done[b], s->stream));
  }

  hlb::DevDomain& D = s->dom;
  D.layout = P.layout == HLB_LAYOUT_BLOCKS ? 1 : (P.layout == HLB_LAYOUT_COMPACT_BLOCKS ? 2 : 0);
  D.n = n;
  D.np = np;
  D.npc = s->npc;
  D.bnbr = s->bnbr.p;
  D.wave_fluid = s->wave_fluid.p;
  D.wave_bnd = s->wave_bnd.p;
  D.wave_links = s->wave_links.p;
  D.wave_rec = s->wave_rec.p;
  D.wave_cbase = s->wave_cbase.p;
  if (blocks_layout) {
    D.bulk_waves = s->bulk_waves.p;
    D.bnd_waves = s->bnd_waves.p;
    D.n_bulk_waves = (uint32_t)h_bulkw.size();
    D.n_bnd_waves = (uint32_t)h_bndw.size();
  }
  D.src = s->src.p;
  D.binfo = s->binfo.p;
  D.nb = (int64_t)h_bwall.size();
  D.b_wall = s->b_wall.p;
  D.b_iolet = s->b_iolet.p;
  D.b_io = s->b_io.p;
  D.b_q = s->b_q.p;
  D.b_qoff = s->b_qoff.p;
  D.b_q16 = s->b_q16.p;
  D.ni = (int64_t)h_io_iolet.size();
  D.io_iolet = s->io_iolet.p;
  D.io_bc = s->io_bc.p;
  D.io_site = s->io_site.p;
  D.io_stash = s->io_stash.p;
  D.io_ladd = s->io_ladd.p;
  D.n_iolets = (int32_t)s->iolets.size();
  D.iolet_nf = s->iolet_nf.p;
  D.iolet_n = s->iolet_n.p;
  D.iolet_kind = s->iolet_kind.p;
  D.blk_omega = s->blk_omega.p;
  D.blk_omega_f = s->blk_omega_f.p;
  D.blk_s_tau = s->blk_s_tau.p;

  hlb::StepArgs& A = s->args;
  A.Q = Q;
  A.kernel = P.kernel;
  A.wall = P.wall;
  A.storage = P.storage;
  A.omega = -1.0 / P.tau;  // LbmParameters::CalcDerivedParams
  A.omega_f = (float)A.omega;
  A.rho0 = P.storage == HLB_STORE_FP64 ? 0.0 : P.rho0;
  // MRT relaxation rate of the viscous moments (DHumieres*MRTBasis.cc)
  A.mrt_s_tau = 1.0 / P.tau;
  // Default launch shape: the workgroup size and workgroup-to-XCD order
  // that measured fastest on the vessel tree for each storage (two sweeps
  // of every shape): 128-thread workgroups in runs of 2 (one half block)
  // per XCD for binary64 (0.7 % ahead of runs of 8) and in runs of 8 (two
  // blocks) for binary32; 64-thread workgroups in runs of 16 (two blocks)
  // for binary16 (1-2 % ahead of 128 threads in the hardware order).
  A.threads = P.storage == HLB_STORE_FP16 ? 64u : 128u;
  A.xcd_remap = P.storage == HLB_STORE_FP64 ? 2u : P.storage == HLB_STORE_FP32 ? 8u : 16u;
  s->initialised = false;
  *out = s.release();
  return HLB_OK;
}

int hlb_create(const hlb_params* params, const hlb_domain* domain,
               const hlb_iolet* inlets, int32_t n_inlets,
               const hlb_iolet* outlets, int32_t n_outlets, hlb_solver** out) {
  return create_impl(params, domain, nullptr, nullptr, 0, nullptr, inlets, n_inlets, outlets,
                     n_outlets, out);
}

int hlb_create_subdomain(const hlb_params* params, const hlb_domain* domain,
                         const hlb_domain* ghosts, const hlb_iolet* inlets, int32_t n_inlets,
                         const hlb_iolet* outlets, int32_t n_outlets, hlb_solver** out) {
  if (!ghosts) return fail(HLB_E_INVALID, "null ghost domain");
  return create_impl(params, domain, ghosts, nullptr, 0, nullptr, inlets, n_inlets, outlets,
                     n_outlets, out);
}

int hlb_create_ensemble(const hlb_params* params, const hlb_domain* domain,
                        const int32_t* instance, int32_t n_instances, const double* tau,
                        const hlb_iolet* inlets, int32_t n_inlets,
                        const hlb_iolet* outlets, int32_t n_outlets, hlb_solver** out) {
  if (!instance) return fail(HLB_E_INVALID, "null instance array");
  return create_impl(params, domain, nullptr, instance, n_instances, tau, inlets, n_inlets,
                     outlets, n_outlets, out);
}

void hlb_destroy(hlb_solver* s) { delete s; }

int hlb_set_equilibrium(hlb_solver* s, double density, const double momentum[3]) {
  if (!s) return fail(HLB_E_INVALID, "null solver");
  LatticeHost L(s->Q);
  double m[3] = {0, 0, 0};
  if (momentum) std::copy(momentum, momentum + 3, m);
  double feq[19];
  feq_host(L, density, m, feq);
  DevBuf<double> vals;
  HIP_OR_FAIL(vals.upload(std::vector<double>(feq, feq + s->Q)));
  if (hlb::launch_fill(s->dom, s->args, vals.p, s->state[s->cur], s->stream))
    return fail(HLB_E_DEVICE, "fill kernel launch failed");
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  s->post_stream = true;
  s->t_ref = 1;
  s->initialised = true;
  return HLB_OK;
}

int hlb_set_time(hlb_solver* s, uint64_t t) {
  if (!s) return fail(HLB_E_INVALID, "null solver");
  if (!s->post_stream) return fail(HLB_E_STATE, "time can only be set before stepping");
  if (t < 1) return fail(HLB_E_INVALID, "time steps start at 1");
  s->t_ref = t;
  return HLB_OK;
}

uint64_t hlb_get_time(const hlb_solver* s) { return s ? s->t_ref : 0; }

int hlb_step(hlb_solver* s, int64_t nsteps) {
  if (!s) return fail(HLB_E_INVALID, "null solver");
  if (!s->initialised) return fail(HLB_E_STATE, "state not initialised");
  if (nsteps < 0) return fail(HLB_E_INVALID, "negative step count");
  while (nsteps > 0) {
    if (s->post_stream) {
      // Collision of step t_ref from the post-stream state; its stream is
      // carried out by the next pull.
      if (hlb::launch_collide_only(s->dom, s->args, s->state[s->cur],
                                   s->state[s->cur ^ 1], s->stream))
        return fail(HLB_E_DEVICE, "collide kernel launch failed");
      s->cur ^= 1;
      s->post_stream = false;
      s->t_ref += 1;
      --nsteps;
      continue;
    }
    const int64_t batch = std::min<int64_t>(nsteps, s->bv_cap);
    const size_t nio = s->iolets.size();
    const int b = s->bv_which;
    s->bv_which ^= 1;
    if (nio) {
      HIP_OR_FAIL(hipEventSynchronize(s->bv_done[b]));
      for (int64_t k = 0; k < batch; ++k)
        s->boundary_values(s->t_ref - 1 + k, s->bv_host[b] + k * 2 * nio);
      HIP_OR_FAIL(hipMemcpyAsync(s->bv[b].p, s->bv_host[b], sizeof(double) * batch * 2 * nio,
                                 hipMemcpyHostToDevice, s->stream));
    }
    for (int64_t k = 0; k < batch; ++k) {
      hlb::StepArgs a = s->args;
      a.bv = nio ? s->bv[b].p + k * 2 * nio : nullptr;
      if (hlb::launch_step(s->dom, a, s->state[s->cur], s->state[s->cur ^ 1], s->stream))
        return fail(HLB_E_DEVICE, "step kernel launch failed");
      s->cur ^= 1;
    }
    HIP_OR_FAIL(hipEventRecord(s->bv_done[b], s->stream));
    s->t_ref += batch;
    nsteps -= batch;
  }
  return HLB_OK;
}

int hlb_time_steps(hlb_solver* s, int64_t n, double* ms) {
  if (!s || !ms) return fail(HLB_E_INVALID, "null argument");
  HIP_OR_FAIL(hipEventRecord(s->ev0, s->stream));
  int r = hlb_step(s, n);
  if (r) return r;
  HIP_OR_FAIL(hipEventRecord(s->ev1, s->stream));
  HIP_OR_FAIL(hipEventSynchronize(s->ev1));
  float f = 0;
  HIP_OR_FAIL(hipEventElapsedTime(&f, s->ev0, s->ev1));
  *ms = f;
  return HLB_OK;
}

int hlb_get_distributions(hlb_solver* s, double* f_out) {
  if (!s || !f_out) return fail(HLB_E_INVALID, "null argument");
  if (!s->initialised) return fail(HLB_E_STATE, "state not initialised");
  int r = s->post_stream_soa();
  if (r) return r;
  if (!s->s_aos.p) HIP_OR_FAIL(s->s_aos.alloc((size_t)s->Q * s->n));
  if (hlb::launch_soa_to_aos(s->Q, s->n, s->npc, s->s_soa.p, s->d_perm.p, s->s_aos.p, s->stream))
    return fail(HLB_E_DEVICE, "permute kernel launch failed");
  HIP_OR_FAIL(hipMemcpyAsync(f_out, s->s_aos.p, sizeof(double) * s->Q * s->n,
                             hipMemcpyDeviceToHost, s->stream));
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  return HLB_OK;
}

int hlb_set_distributions(hlb_solver* s, const double* f_in) {
  if (!s || !f_in) return fail(HLB_E_INVALID, "null argument");
  if (!s->s_aos.p) HIP_OR_FAIL(s->s_aos.alloc((size_t)s->Q * s->n));
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  HIP_OR_FAIL(hipMemcpy(s->s_aos.p, f_in, sizeof(double) * s->Q * s->n, hipMemcpyHostToDevice));
  if (hlb::launch_import(s->dom, s->args, s->s_aos.p, s->d_perm.p, s->state[s->cur],
                         s->stream))
    return fail(HLB_E_DEVICE, "import kernel launch failed");
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  s->post_stream = true;
  s->initialised = true;
  return HLB_OK;
}

int hlb_set_equilibrium_fields(hlb_solver* s, const double* density, const double* momentum,
                               int32_t device) {
  if (!s || !density || !momentum) return fail(HLB_E_INVALID, "null argument");
  const double* rho = density;
  const double* mom = momentum;
  DevBuf<double> hr, hm;
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  if (!device) {
    HIP_OR_FAIL(hr.alloc((size_t)s->n));
    HIP_OR_FAIL(hm.alloc((size_t)3 * s->n));
    HIP_OR_FAIL(hipMemcpy(hr.p, density, sizeof(double) * s->n, hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(hm.p, momentum, sizeof(double) * 3 * s->n, hipMemcpyHostToDevice));
    rho = hr.p;
    mom = hm.p;
  }
  if (hlb::launch_import_equilibrium(s->dom, s->args, rho, mom, s->d_perm.p,
                                     s->state[s->cur], s->stream))
    return fail(HLB_E_DEVICE, "equilibrium import kernel launch failed");
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  s->post_stream = true;
  s->initialised = true;
  return HLB_OK;
}

static int compute_moments(hlb_solver* s) {
  int r = s->post_stream_soa();
  if (r) return r;
  r = s->ensure_moments();
  if (r) return r;
  if (hlb::launch_moments(s->Q, s->n, s->npc, s->s_soa.p, s->rho.p, s->u.p, s->stream))
    return fail(HLB_E_DEVICE, "moments kernel launch failed");
  return HLB_OK;
}

int hlb_get_macroscopic(hlb_solver* s, double* density, double* velocity) {
  if (!s) return fail(HLB_E_INVALID, "null solver");
  if (!s->initialised) return fail(HLB_E_STATE, "state not initialised");
  int r = compute_moments(s);
  if (r) return r;
  std::vector<double> rho(s->npc), u(3 * s->npc);
  HIP_OR_FAIL(hipMemcpyAsync(rho.data(), s->rho.p, sizeof(double) * s->npc,
                             hipMemcpyDeviceToHost, s->stream));
  HIP_OR_FAIL(hipMemcpyAsync(u.data(), s->u.p, sizeof(double) * 3 * s->npc,
                             hipMemcpyDeviceToHost, s->stream));
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  for (int64_t o = 0; o < s->n; ++o) {
    const int64_t c = s->perm[o];
    if (density) density[c] = rho[o];
    if (velocity)
      for (int a = 0; a < 3; ++a) velocity[3 * c + a] = u[(size_t)a * s->npc + o];
  }
  return HLB_OK;
}

int hlb_get_wall_shear_stress(hlb_solver* s, double* stress) {
  if (!s || !stress) return fail(HLB_E_INVALID, "null argument");
  if (!s->initialised) return fail(HLB_E_STATE, "state not initialised");
  if (!s->has_normals)
    return fail(HLB_E_STATE, "wall normals were not given to hlb_create");
  int r = s->post_stream_soa();
  if (r) return r;
  const int64_t nw = (int64_t)s->wall_sites.size();
  std::vector<double> w(nw);
  if (nw) {
    if (!s->d_wall_sites.p) {
      HIP_OR_FAIL(s->d_wall_sites.upload(s->wall_sites));
      HIP_OR_FAIL(s->d_wall_normals.upload(s->wall_normals));
      // per site: -sqrt(2) * the stress parameter of LbmParameters
      // (LbmParameters.h:34-39) as CalculateWallShearStressMagnitude uses it
      // (Lattice.h:669), from the site's tau
      HIP_OR_FAIL(s->d_wall_temp.upload(s->wall_temp));
      HIP_OR_FAIL(s->d_wss.alloc((size_t)nw));
    }
    if (hlb::launch_wall_shear_stress(s->Q, nw, s->npc, s->s_soa.p, s->d_wall_sites.p,
                                      s->d_wall_normals.p, s->d_wall_temp.p, s->d_wss.p,
                                      s->stream))
      return fail(HLB_E_DEVICE, "wall shear stress kernel launch failed");
    HIP_OR_FAIL(hipMemcpyAsync(w.data(), s->d_wss.p, sizeof(double) * nw,
                               hipMemcpyDeviceToHost, s->stream));
  }
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  std::fill(stress, stress + s->n, std::numeric_limits<double>::max());
  for (int64_t k = 0; k < nw; ++k) stress[s->perm[s->wall_sites[k]]] = w[k];
  return HLB_OK;
}

int hlb_get_stress(hlb_solver* s, double* tensor, double* von_mises, double* shear_rate,
                   double* traction, double* tangential) {
  if (!s) return fail(HLB_E_INVALID, "null solver");
  if (!s->initialised) return fail(HLB_E_STATE, "state not initialised");
  if ((traction || tangential) && !s->has_normals)
    return fail(HLB_E_STATE, "wall normals were not given to hlb_create");
  int r = s->post_stream_soa();
  if (r) return r;
  const int64_t npc = s->npc, n = s->n;
  if (!s->site_inst.empty() && !s->d_site_tau.p) {
    std::vector<double> t(n);
    for (int64_t o = 0; o < n; ++o) t[o] = s->inst_tau[s->site_inst[o]];
    HIP_OR_FAIL(s->d_site_tau.upload(t));
  }
  if (tensor && !s->d_t6.p) HIP_OR_FAIL(s->d_t6.alloc((size_t)6 * npc));
  if (von_mises && !s->d_vm.p) HIP_OR_FAIL(s->d_vm.alloc((size_t)npc));
  if (shear_rate && !s->d_sr.p) HIP_OR_FAIL(s->d_sr.alloc((size_t)npc));
  if (tensor || von_mises || shear_rate) {
    if (hlb::launch_stress_fields(s->Q, n, npc, s->s_soa.p, s->p.tau, s->d_site_tau.p,
                                  tensor ? s->d_t6.p : nullptr, von_mises ? s->d_vm.p : nullptr,
                                  shear_rate ? s->d_sr.p : nullptr, s->stream))
      return fail(HLB_E_DEVICE, "stress kernel launch failed");
  }
  const int64_t nw = (int64_t)s->wall_sites.size();
  std::vector<double> h6, hvm, hsr, htr, htg;
  if (tensor) h6.resize(6 * npc);
  if (von_mises) hvm.resize(npc);
  if (shear_rate) hsr.resize(npc);
  if ((traction || tangential) && nw) {
    if (!s->d_wall_sites.p) {
      HIP_OR_FAIL(s->d_wall_sites.upload(s->wall_sites));
      HIP_OR_FAIL(s->d_wall_normals.upload(s->wall_normals));
      HIP_OR_FAIL(s->d_wall_temp.upload(s->wall_temp));
      HIP_OR_FAIL(s->d_wss.alloc((size_t)nw));
    }
    if (!s->d_wall_tau.p) {
      HIP_OR_FAIL(s->d_wall_tau.upload(s->wall_tau));
      HIP_OR_FAIL(s->d_tr.alloc((size_t)3 * nw));
      HIP_OR_FAIL(s->d_tg.alloc((size_t)3 * nw));
    }
    if (hlb::launch_traction(s->Q, nw, npc, s->s_soa.p, s->d_wall_sites.p, s->d_wall_normals.p,
                             s->d_wall_tau.p, s->d_tr.p, s->d_tg.p, s->stream))
      return fail(HLB_E_DEVICE, "traction kernel launch failed");
    htr.resize(3 * nw);
    htg.resize(3 * nw);
    HIP_OR_FAIL(hipMemcpyAsync(htr.data(), s->d_tr.p, sizeof(double) * 3 * nw,
                               hipMemcpyDeviceToHost, s->stream));
    HIP_OR_FAIL(hipMemcpyAsync(htg.data(), s->d_tg.p, sizeof(double) * 3 * nw,
                               hipMemcpyDeviceToHost, s->stream));
  }
  if (tensor)
    HIP_OR_FAIL(hipMemcpyAsync(h6.data(), s->d_t6.p, sizeof(double) * 6 * npc,
                               hipMemcpyDeviceToHost, s->stream));
  if (von_mises)
    HIP_OR_FAIL(hipMemcpyAsync(hvm.data(), s->d_vm.p, sizeof(double) * npc,
                               hipMemcpyDeviceToHost, s->stream));
  if (shear_rate)
    HIP_OR_FAIL(hipMemcpyAsync(hsr.data(), s->d_sr.p, sizeof(double) * npc,
                               hipMemcpyDeviceToHost, s->stream));
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  for (int64_t o = 0; o < n; ++o) {
    const int64_t c = s->perm[o];
    if (tensor)
      for (int k = 0; k < 6; ++k) tensor[6 * c + k] = h6[(size_t)k * npc + o];
    if (von_mises) von_mises[c] = hvm[o];
    if (shear_rate) shear_rate[c] = hsr[o];
  }
  // traction is zero away from the walls (Common.h:107-121)
  if (traction) std::fill(traction, traction + 3 * n, 0.0);
  if (tangential) std::fill(tangential, tangential + 3 * n, 0.0);
  for (int64_t k = 0; k < (int64_t)htr.size() / 3; ++k) {
    const int64_t c = s->perm[s->wall_sites[k]];
    for (int a = 0; a < 3; ++a) {
      if (traction) traction[3 * c + a] = htr[3 * k + a];
      if (tangential) tangential[3 * c + a] = htg[3 * k + a];
    }
  }
  return HLB_OK;
}

// The monitor's statistics (hlb_reduce, hlb_reduce_noalloc): one launch of
// k_monitor over the state into `partials` and `io_vals`.
static int launch_monitor_into(hlb_solver* s, double* partials, double* io_vals) {
  const hlb::StepArgs a = s->args_for_stream();
  if (hlb::launch_monitor(s->dom, a, s->state[s->cur], s->post_stream ? 1 : 0, partials,
                          io_vals, s->stream))
    return fail(HLB_E_DEVICE, "monitor kernel launch failed");
  return HLB_OK;
}

// Workgroup partials [mass, min rho, max rho, max |u|, unstable] combined
// in workgroup order.
static void combine_partials(const hlb_solver* s, const double* part, hlb_monitor& m) {
  m = hlb_monitor{};
  m.density_min = 1.7976931348623157e308;
  m.density_max = -1.7976931348623157e308;
  bool bad = false;
  for (int b = 0; b < hlb::kMonitorGroups; ++b) {
    const double* p = &part[5 * b];
    m.mass += p[0];
    m.density_min = std::min(m.density_min, p[1]);
    m.density_max = std::max(m.density_max, p[2]);
    m.velocity_max = std::max(m.velocity_max, p[3]);
    bad = bad || p[4] != 0.0;
  }
  m.stable = bad ? 0 : 1;
  m.n_iolets = (int32_t)s->iolets.size();
}

// Iolet fluxes: sum over the iolet sites (records k0 .. k0+count-1, in
// record order) of rho (u . n); vals[4*(k-k0)..] = rho, u of record k.
static void add_fluxes(const hlb_solver* s, const double* vals, int64_t k0, int64_t count,
                       hlb_monitor& m) {
  for (int64_t j = 0; j < count; ++j) {
    const int io = s->io_sites_iolet[k0 + j];
    const double* nrm = s->iolets[io].cfg.normal;
    const double* v = vals + 4 * j;
    const double un = ((0.0 + v[1] * nrm[0]) + v[2] * nrm[1]) + v[3] * nrm[2];
    if (io < 64) m.iolet_flux[io] += v[0] * un;
  }
}

int hlb_reduce(hlb_solver* s, hlb_monitor* out) {
  if (!s || !out) return fail(HLB_E_INVALID, "null argument");
  if (!s->initialised) return fail(HLB_E_STATE, "state not initialised");
  const int nb = hlb::kMonitorGroups;
  const int64_t ni = (int64_t)s->io_sites_internal.size();
  if (!s->partials.p) HIP_OR_FAIL(s->partials.alloc(5 * nb));
  if (!s->mon_io.p) HIP_OR_FAIL(s->mon_io.alloc(4 * std::max<int64_t>(ni, 1)));
  int r = launch_monitor_into(s, s->partials.p, s->mon_io.p);
  if (r) return r;
  std::vector<double> part(5 * nb), g(4 * std::max<int64_t>(ni, 1));
  HIP_OR_FAIL(hipMemcpyAsync(part.data(), s->partials.p, sizeof(double) * 5 * nb,
                             hipMemcpyDeviceToHost, s->stream));
  if (ni)
    HIP_OR_FAIL(hipMemcpyAsync(g.data(), s->mon_io.p, sizeof(double) * 4 * ni,
                               hipMemcpyDeviceToHost, s->stream));
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  hlb_monitor m;
  combine_partials(s, part.data(), m);
  add_fluxes(s, g.data(), 0, ni, m);
  *out = m;
  return HLB_OK;
}

int hlb_reduce_noalloc(hlb_solver* s, hlb_monitor* out) {
  if (!s || !out) return fail(HLB_E_INVALID, "null argument");
  if (!s->initialised) return fail(HLB_E_STATE, "state not initialised");
  const int64_t ni = (int64_t)s->io_sites_internal.size();
  if (ni > hlb::kStaticIoletSites)
    return fail(HLB_E_INVALID, "hlb_reduce_noalloc: " + std::to_string(ni) +
                                   " iolet sites exceed the static scratch limit of " +
                                   std::to_string(hlb::kStaticIoletSites) +
                                   "; use hlb_reduce");
  double* partials = s->scratch;
  double* io_vals = s->scratch + 5 * hlb::kMonitorGroups;
  int r = launch_monitor_into(s, partials, io_vals);
  if (r) return r;
  // host copies into fixed buffers on the stack: no allocation anywhere
  double part[5 * hlb::kMonitorGroups];
  HIP_OR_FAIL(hipMemcpyAsync(part, partials, sizeof(part), hipMemcpyDeviceToHost, s->stream));
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  hlb_monitor m;
  combine_partials(s, part, m);
  constexpr int64_t kChunk = 512;
  double vals[4 * kChunk];
  for (int64_t k0 = 0; k0 < ni; k0 += kChunk) {
    const int64_t cnt = std::min<int64_t>(kChunk, ni - k0);
    HIP_OR_FAIL(hipMemcpyAsync(vals, io_vals + 4 * k0, sizeof(double) * 4 * cnt,
                               hipMemcpyDeviceToHost, s->stream));
    HIP_OR_FAIL(hipStreamSynchronize(s->stream));
    add_fluxes(s, vals, k0, cnt, m);
  }
  *out = m;
  return HLB_OK;
}

int hlb_snapshot_velocity(hlb_solver* s) {
  if (!s) return fail(HLB_E_INVALID, "null solver");
  if (!s->initialised) return fail(HLB_E_STATE, "state not initialised");
  int r = compute_moments(s);
  if (r) return r;
  if (!s->u_snap.p) HIP_OR_FAIL(s->u_snap.alloc(3 * s->npc));
  HIP_OR_FAIL(hipMemcpyAsync(s->u_snap.p, s->u.p, sizeof(double) * 3 * s->npc,
                             hipMemcpyDeviceToDevice, s->stream));
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  return HLB_OK;
}

int hlb_velocity_change(hlb_solver* s, double ref, double* out) {
  if (!s || !out) return fail(HLB_E_INVALID, "null argument");
  if (!s->u_snap.p) return fail(HLB_E_STATE, "no velocity snapshot");
  if (!(ref > 0)) return fail(HLB_E_INVALID, "reference speed must be positive");
  int r = compute_moments(s);
  if (r) return r;
  const int nb = 1024;
  if (!s->partials.p) HIP_OR_FAIL(s->partials.alloc(5 * nb));
  if (hlb::launch_velocity_change(s->n, s->npc, s->u.p, s->u_snap.p, 1.0 / ref, s->partials.p,
                                  nb, s->stream))
    return fail(HLB_E_DEVICE, "velocity change kernel launch failed");
  std::vector<double> part(nb);
  HIP_OR_FAIL(hipMemcpyAsync(part.data(), s->partials.p, sizeof(double) * nb,
                             hipMemcpyDeviceToHost, s->stream));
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  double m = 0.0;
  for (double v : part) m = std::max(m, v);
  *out = m;
  return HLB_OK;
}

// Chunks of at most kChunk sites inside the runs of one instance (the
// per-instance reductions combine their partials in chunk order).
static void instance_chunks(const hlb_solver* s, std::vector<int64_t>& range,
                            std::vector<int32_t>& inst) {
  constexpr int64_t kChunk = 1 << 16;
  range.clear();
  inst.clear();
  for (const auto& r : s->runs)
    for (int64_t a = r.start; a < r.end; a += kChunk) {
      range.push_back(a);
      range.push_back(std::min(r.end, a + kChunk));
      inst.push_back(r.inst);
    }
}

int32_t hlb_instance_count(const hlb_solver* s) { return s ? s->n_inst : 0; }

int hlb_reduce_instances(hlb_solver* s, hlb_instance_monitor* out) {
  if (!s || !out) return fail(HLB_E_INVALID, "null argument");
  if (!s->initialised) return fail(HLB_E_STATE, "state not initialised");
  int r = compute_moments(s);
  if (r) return r;
  std::vector<int64_t> range;
  std::vector<int32_t> inst;
  instance_chunks(s, range, inst);
  const int64_t nc = (int64_t)inst.size();
  DevBuf<int64_t> d_range;
  DevBuf<double> d_part;
  HIP_OR_FAIL(d_range.upload(range));
  HIP_OR_FAIL(d_part.alloc((size_t)5 * std::max<int64_t>(nc, 1)));
  if (hlb::launch_reduce_ranges(s->Q, nc, s->npc, d_range.p, s->s_soa.p, s->rho.p, s->u.p,
                                d_part.p, s->stream))
    return fail(HLB_E_DEVICE, "reduction kernel launch failed");
  std::vector<double> part(5 * nc);
  HIP_OR_FAIL(hipMemcpyAsync(part.data(), d_part.p, sizeof(double) * 5 * nc,
                             hipMemcpyDeviceToHost, s->stream));
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  for (int32_t k = 0; k < s->n_inst; ++k) {
    hlb_instance_monitor& m = out[k];
    m = hlb_instance_monitor{};
    m.density_min = 1.7976931348623157e308;
    m.density_max = -1.7976931348623157e308;
    m.stable = 1;
  }
  for (int64_t c = 0; c < nc; ++c) {
    hlb_instance_monitor& m = out[inst[c]];
    const double* p = &part[5 * c];
    m.mass += p[0];
    m.density_min = std::min(m.density_min, p[1]);
    m.density_max = std::max(m.density_max, p[2]);
    m.velocity_max = std::max(m.velocity_max, p[3]);
    if (p[4] != 0.0) m.stable = 0;
    m.sites += range[2 * c + 1] - range[2 * c];
  }
  return HLB_OK;
}

int hlb_velocity_change_instances(hlb_solver* s, const double* reference_speed, double* out) {
  if (!s || !reference_speed || !out) return fail(HLB_E_INVALID, "null argument");
  if (!s->u_snap.p) return fail(HLB_E_STATE, "no velocity snapshot");
  for (int32_t k = 0; k < s->n_inst; ++k)
    if (!(reference_speed[k] > 0)) return fail(HLB_E_INVALID, "reference speed must be positive");
  int r = compute_moments(s);
  if (r) return r;
  std::vector<int64_t> range;
  std::vector<int32_t> inst;
  instance_chunks(s, range, inst);
  const int64_t nc = (int64_t)inst.size();
  std::vector<double> inv(nc);
  for (int64_t c = 0; c < nc; ++c) inv[c] = 1.0 / reference_speed[inst[c]];
  DevBuf<int64_t> d_range;
  DevBuf<double> d_inv, d_part;
  HIP_OR_FAIL(d_range.upload(range));
  HIP_OR_FAIL(d_inv.upload(inv));
  HIP_OR_FAIL(d_part.alloc((size_t)std::max<int64_t>(nc, 1)));
  if (hlb::launch_velocity_change_ranges(nc, s->npc, d_range.p, s->u.p, s->u_snap.p, d_inv.p,
                                         d_part.p, s->stream))
    return fail(HLB_E_DEVICE, "velocity change kernel launch failed");
  std::vector<double> part(nc);
  HIP_OR_FAIL(hipMemcpyAsync(part.data(), d_part.p, sizeof(double) * nc,
                             hipMemcpyDeviceToHost, s->stream));
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  std::fill(out, out + s->n_inst, 0.0);
  for (int64_t c = 0; c < nc; ++c) out[inst[c]] = std::max(out[inst[c]], part[c]);
  return HLB_OK;
}

static int check_shape(const hlb_solver* s, const hlb_launch_shape& sh) {
  if (s->p.layout == HLB_LAYOUT_PULL_TABLE)
    return fail(HLB_E_INVALID, "launch shapes apply to the block layout");
  if (sh.threads != 64 && sh.threads != 128 && sh.threads != 256)
    return fail(HLB_E_INVALID, "threads per workgroup must be 64, 128 or 256");
  if (sh.planes < 1 || sh.planes > 8) return fail(HLB_E_INVALID, "planes per wave must be 1..8");
  if (sh.groups < 0) return fail(HLB_E_INVALID, "workgroup count must not be negative");
  if (sh.xcd_remap < 0 || sh.xcd_remap > 4096)
    return fail(HLB_E_INVALID, "xcd_remap must be in 0..4096");
  if (sh.split != 0 && sh.split != 1) return fail(HLB_E_INVALID, "split must be 0 or 1");
  if (sh.wide != 0 && sh.wide != 1) return fail(HLB_E_INVALID, "wide must be 0 or 1");
  return HLB_OK;
}

int hlb_set_launch_shape(hlb_solver* s, const hlb_launch_shape* shape) {
  if (!s || !shape) return fail(HLB_E_INVALID, "null argument");
  int r = check_shape(s, *shape);
  if (r) return r;
  s->args.threads = (uint32_t)shape->threads;
  s->args.planes = (uint32_t)shape->planes;
  s->args.groups = (uint32_t)shape->groups;
  s->args.xcd_remap = (uint32_t)shape->xcd_remap;
  s->args.split = (uint32_t)shape->split;
  s->args.wide = (uint32_t)shape->wide;
  return HLB_OK;
}

int hlb_get_launch_shape(const hlb_solver* s, hlb_launch_shape* shape) {
  if (!s || !shape) return fail(HLB_E_INVALID, "null argument");
  shape->threads = (int32_t)s->args.threads;
  shape->planes = (int32_t)s->args.planes;
  shape->groups = (int32_t)s->args.groups;
  shape->xcd_remap = (int32_t)s->args.xcd_remap;
  shape->split = (int32_t)s->args.split;
  shape->wide = (int32_t)s->args.wide;
  return HLB_OK;
}

int hlb_autotune(hlb_solver* s, int32_t steps, hlb_launch_shape* best, double* best_ms) {
  if (!s) return fail(HLB_E_INVALID, "null solver");
  if (!s->initialised) return fail(HLB_E_STATE, "state not initialised");
  if (s->p.layout == HLB_LAYOUT_PULL_TABLE)
    return fail(HLB_E_INVALID, "the launch-shape tuner applies to the block layouts");
  if (steps < 1) return fail(HLB_E_INVALID, "steps per trial must be positive");
  // The trials advance the state: keep a device copy of everything a step
  // changes (both state buffers, the iolet stash) and the time bookkeeping.
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  const size_t sbytes = storage_bytes(s->p.storage) * s->Q * s->np;
  DevBuf<unsigned char> saved[2];  // freed on every return path
  for (int b = 0; b < 2; ++b) {
    HIP_OR_FAIL(saved[b].alloc(sbytes));
    HIP_OR_FAIL(hipMemcpyAsync(saved[b].p, s->state[b], sbytes, hipMemcpyDeviceToDevice,
                               s->stream));
  }
  const size_t ni = std::max<size_t>(1, s->io_sites_internal.size());
  DevBuf<double> stash;
  HIP_OR_FAIL(stash.alloc(ni));
  HIP_OR_FAIL(hipMemcpyAsync(stash.p, s->io_stash.p, sizeof(double) * ni,
                             hipMemcpyDeviceToDevice, s->stream));
  const int cur = s->cur;
  const uint64_t t_ref = s->t_ref;
  const bool post_stream = s->post_stream;
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->p.device);
  std::vector<hlb_launch_shape> cand;
  for (int split : {1, 0}) {
    for (int th : {256, 128, 64})
      for (int xcd : {1, 0}) cand.push_back({th, 1, 0, xcd, split});
    for (int run : {2, 16, 64}) cand.push_back({256, 1, 0, run, split});
    cand.push_back({128, 1, 0, 4, split});
    cand.push_back({128, 1, 0, 8, split});
  }
  cand.push_back({256, 2, 0, 1, 0});
  cand.push_back({128, 2, 0, 1, 0});
  for (int th : {256, 128}) cand.push_back({th, 1, cus * 24 * 64 / th, 1, 0});
  cand.push_back({256, 2, cus * 24 * 64 / 256, 1, 0});
  hlb_launch_shape keep{};
  hlb_get_launch_shape(s, &keep);
  for (auto& c : cand) c.wide = keep.wide;  // the addressing stays the caller's
  double best_t = 1e300;
  hlb_launch_shape best_s = keep;
  int rc = HLB_OK;
  for (const auto& c : cand) {
    hlb_set_launch_shape(s, &c);
    double ms = 0;
    rc = hlb_step(s, 2);  // warm
    if (!rc) rc = hlb_time_steps(s, steps, &ms);
    if (rc) break;
    if (ms < best_t) {
      best_t = ms;
      best_s = c;
    }
  }
  // restore
  for (int b = 0; b < 2; ++b) {
    if (hipMemcpyAsync(s->state[b], saved[b].p, sbytes, hipMemcpyDeviceToDevice, s->stream) !=
        hipSuccess)
      rc = fail(HLB_E_DEVICE, "state restore failed");
  }
  if (hipMemcpyAsync(s->io_stash.p, stash.p, sizeof(double) * ni, hipMemcpyDeviceToDevice,
                     s->stream) != hipSuccess)
    rc = fail(HLB_E_DEVICE, "state restore failed");
  (void)hipStreamSynchronize(s->stream);
  s->cur = cur;
  s->t_ref = t_ref;
  s->post_stream = post_stream;
  if (rc) {
    hlb_set_launch_shape(s, &keep);
    return rc;
  }
  hlb_set_launch_shape(s, &best_s);
  if (best) *best = best_s;
  if (best_ms) *best_ms = best_t / steps;
  return HLB_OK;
}

int hlb_copy_ceiling(hlb_solver* s, int32_t reps, hlb_copy_result* out) {
  if (!s || !out) return fail(HLB_E_INVALID, "null argument");
  if (reps < 1) return fail(HLB_E_INVALID, "reps must be positive");
  const int64_t esize = (int64_t)storage_bytes(s->p.storage);
  const int64_t bytes = s->np * esize;
  // the other state buffer is free between steps (the next step overwrites it)
  const void* src = s->state[s->cur];
  void* dst = s->state[s->cur ^ 1];
  hlb_copy_result best{};
  for (int order : {0, 1})
  for (int width : {(int)std::max<int64_t>(esize, 4), 16}) {
    // order 1: the block layout's order (runs of 512 slots per direction and
    // block); order 0: one array per direction
    const int64_t run = order ? 512 * esize : 0;
    if (bytes % width || (order && (s->np % 512 || run % width))) continue;
    for (uint32_t threads : {128u, 256u}) {
      for (uint32_t remap : {0u, 8u}) {
        auto launch = [&]() {
          return hlb::launch_copy_arrays(s->Q, bytes, bytes, run, width, src, dst, threads,
                                         remap, s->stream);
        };
        if (launch()) return fail(HLB_E_DEVICE, "copy kernel launch failed");
        HIP_OR_FAIL(hipEventRecord(s->ev0, s->stream));
        for (int r = 0; r < reps; ++r)
          if (launch()) return fail(HLB_E_DEVICE, "copy kernel launch failed");
        HIP_OR_FAIL(hipEventRecord(s->ev1, s->stream));
        HIP_OR_FAIL(hipEventSynchronize(s->ev1));
        float ms = 0;
        HIP_OR_FAIL(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        const double per = (double)ms / reps;
        const double rate = 2.0 * s->Q * (double)bytes / (per * 1e-3);
        if (rate > best.bytes_per_s) {
          best.bytes_per_s = rate;
          best.ms = per;
          best.bytes = 2 * (int64_t)s->Q * bytes;
          best.threads = (int32_t)threads;
          best.xcd_remap = (int32_t)remap;
          best.width = width;
          best.order = order;
        }
      }
    }
  }
  *out = best;
  return HLB_OK;
}

int hlb_halo_pack(hlb_solver* s, const int64_t* pairs, int64_t count, double* host_buffer) {
  if (!s || (count && (!pairs || !host_buffer))) return fail(HLB_E_INVALID, "null argument");
  if (s->post_stream)
    return fail(HLB_E_STATE, "no post-collision state yet (step at least once)");
  if (count == 0) return HLB_OK;
  std::vector<int64_t> slots(count);
  for (int64_t k = 0; k < count; ++k) {
    const int64_t site = pairs[2 * k], d = pairs[2 * k + 1];
    if (site < 0 || site >= s->n || d < 0 || d >= s->Q)
      return fail(HLB_E_INVALID, "halo pair out of range");
    slots[k] = s->state_index(d, s->slot_of[s->internal[site]]);
  }
  DevBuf<int64_t> d_slots;
  DevBuf<double> d_out;
  HIP_OR_FAIL(d_slots.upload(slots));
  HIP_OR_FAIL(d_out.alloc(count));
  if (hlb::launch_gather_pairs(count, s->np, s->Q, s->blockwise(), d_slots.p,
                               s->state[s->cur], s->p.storage,
                               s->args.rho0, s->weights.p, d_out.p, s->stream))
    return fail(HLB_E_DEVICE, "halo pack kernel launch failed");
  HIP_OR_FAIL(hipMemcpyAsync(host_buffer, d_out.p, sizeof(double) * count,
                             hipMemcpyDeviceToHost, s->stream));
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  return HLB_OK;
}

int hlb_halo_unpack(hlb_solver* s, const int64_t* pairs, int64_t count,
                    const double* host_buffer) {
  if (!s || (count && (!pairs || !host_buffer))) return fail(HLB_E_INVALID, "null argument");
  if (s->post_stream)
    return fail(HLB_E_STATE, "no post-collision state yet (step at least once)");
  if (count == 0) return HLB_OK;
  std::vector<int64_t> slots(count);
  for (int64_t k = 0; k < count; ++k) {
    const int64_t site = pairs[2 * k], d = pairs[2 * k + 1];
    if (site < 0 || site >= s->n || d < 0 || d >= s->Q)
      return fail(HLB_E_INVALID, "halo pair out of range");
    slots[k] = s->state_index(d, s->slot_of[s->internal[site]]);
  }
  {
    std::vector<int64_t> sorted(slots);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
      return fail(HLB_E_INVALID, "a (site, direction) pair appears twice");
  }
  DevBuf<int64_t> d_slots;
  DevBuf<double> d_in;
  HIP_OR_FAIL(d_slots.upload(slots));
  HIP_OR_FAIL(d_in.upload(std::vector<double>(host_buffer, host_buffer + count)));
  if (hlb::launch_scatter_pairs(count, s->np, s->Q, s->blockwise(), d_slots.p, d_in.p,
                                s->p.storage, s->args.rho0,
                                s->weights.p, s->state[s->cur], s->stream))
    return fail(HLB_E_DEVICE, "halo unpack kernel launch failed");
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  return HLB_OK;
}

// ---- Halo plans between subdomains ---------------------------------------
// The values a subdomain's sites pull across the cut are the neighbouring
// subdomain's post-collision values of its edge sites: the pairs (site x,
// direction e) with x + c_e in the other subdomain (the reference's shared
// distributions, Domain.cc:166-179 and 247-284, sent by
// FieldData::SendAndReceive and placed by CopyReceived, FieldData.cc:27-49).
// A plan holds their state indices on both sides, computed once; every
// exchange is two kernels and an event, with no allocation and no host
// synchronisation.
}  // extern "C"

struct hlb_halo_plan {
  hlb_solver* src = nullptr;
  hlb_solver* dst = nullptr;
  int64_t count = 0;
  int esize = 8;
  DevBuf<int64_t> src_idx, dst_idx;
  DevBuf<unsigned char> buf;
  hipEvent_t packed = nullptr, ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bool timed = false;
  ~hlb_halo_plan() {
    if (packed) (void)hipEventDestroy(packed);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

extern "C" {

static int64_t find_site(const hlb_solver* s, uint64_t key) {
  auto it = std::lower_bound(s->site_by_key.begin(), s->site_by_key.end(),
                             std::make_pair(key, (int64_t)-1));
  return (it != s->site_by_key.end() && it->first == key) ? it->second : -1;
}

int hlb_halo_plan_create(hlb_solver* src, hlb_solver* dst, hlb_halo_plan** out) {
  if (!src || !dst || !out) return fail(HLB_E_INVALID, "null argument");
  *out = nullptr;
  if (!src->subdomain || !dst->subdomain)
    return fail(HLB_E_INVALID, "halo plans join solvers made by hlb_create_subdomain");
  if (src == dst) return fail(HLB_E_INVALID, "a halo plan joins two different subdomains");
  if (src->Q != dst->Q || src->p.storage != dst->p.storage || src->p.device != dst->p.device)
    return fail(HLB_E_INVALID, "subdomains differ in lattice, storage or device");
  if (src->p.storage != HLB_STORE_FP64 && src->p.rho0 != dst->p.rho0)
    return fail(HLB_E_INVALID, "subdomains differ in the reduced-precision shift rho0");
  LatticeHost L(dst->Q);
  std::vector<std::pair<int64_t, int64_t>> pairs;
  const int64_t ng = (int64_t)dst->ghost_cut.size();
  for (int64_t g = 0; g < ng; ++g) {
    const int32_t* c = &dst->ghost_coords[3 * g];
    const int64_t o_src = find_site(src, coord_key(c));
    if (o_src < 0 || o_src >= src->n) continue;  // a ghost of another subdomain
    for (int e = 1; e < dst->Q; ++e) {
      if ((dst->ghost_cut[g] >> (e - 1)) & 1u) continue;  // cut link: nothing streams
      const int32_t t[3] = {c[0] + L.c[e][0], c[1] + L.c[e][1], c[2] + L.c[e][2]};
      const int64_t o_dst = find_site(dst, coord_key(t));
      if (o_dst < 0 || o_dst >= dst->n) continue;
      pairs.push_back({src->state_index(e, src->slot_of[o_src]),
                       dst->state_index(e, dst->slot_of[dst->n + g])});
    }
  }
  std::sort(pairs.begin(), pairs.end());
  auto plan = std::make_unique<hlb_halo_plan>();
  plan->src = src;
  plan->dst = dst;
  plan->count = (int64_t)pairs.size();
  plan->esize = (int)storage_bytes(src->p.storage);
  std::vector<int64_t> si(pairs.size()), di(pairs.size());
  for (size_t k = 0; k < pairs.size(); ++k) {
    si[k] = pairs[k].first;
    di[k] = pairs[k].second;
  }
  HIP_OR_FAIL(hipSetDevice(src->p.device));
  HIP_OR_FAIL(plan->src_idx.upload(si));
  HIP_OR_FAIL(plan->dst_idx.upload(di));
  HIP_OR_FAIL(plan->buf.alloc((size_t)plan->count * plan->esize));
  HIP_OR_FAIL(hipEventCreateWithFlags(&plan->packed, hipEventDisableTiming));
  for (auto& e : plan->ev) HIP_OR_FAIL(hipEventCreate(&e));
  *out = plan.release();
  return HLB_OK;
}

void hlb_halo_plan_destroy(hlb_halo_plan* plan) { delete plan; }

int64_t hlb_halo_plan_count(const hlb_halo_plan* plan) { return plan ? plan->count : -1; }

static int halo_check(const hlb_halo_plan* plan) {
  if (!plan) return fail(HLB_E_INVALID, "null plan");
  if (!plan->src->initialised || !plan->dst->initialised)
    return fail(HLB_E_STATE, "state not initialised");
  if (plan->src->post_stream || plan->dst->post_stream)
    return fail(HLB_E_STATE, "no post-collision state yet (step at least once)");
  return HLB_OK;
}

int hlb_halo_pack_device(hlb_halo_plan* plan, void* buffer) {
  int r = halo_check(plan);
  if (r) return r;
  if (plan->count && !buffer) return fail(HLB_E_INVALID, "null buffer");
  hlb_solver* s = plan->src;
  if (hlb::launch_halo(plan->count, plan->src_idx.p, plan->esize, s->state[s->cur], buffer, 1,
                       s->stream))
    return fail(HLB_E_DEVICE, "halo pack kernel launch failed");
  return HLB_OK;
}

int hlb_halo_unpack_device(hlb_halo_plan* plan, const void* buffer) {
  int r = halo_check(plan);
  if (r) return r;
  if (plan->count && !buffer) return fail(HLB_E_INVALID, "null buffer");
  hlb_solver* d = plan->dst;
  if (hlb::launch_halo(plan->count, plan->dst_idx.p, plan->esize, buffer, d->state[d->cur], 0,
                       d->stream))
    return fail(HLB_E_DEVICE, "halo unpack kernel launch failed");
  return HLB_OK;
}

int hlb_halo_exchange(hlb_halo_plan* plan) {
  int r = halo_check(plan);
  if (r) return r;
  hlb_solver* s = plan->src;
  hlb_solver* d = plan->dst;
  HIP_OR_FAIL(hipEventRecord(plan->ev[0], s->stream));
  r = hlb_halo_pack_device(plan, plan->buf.p);
  if (r) return r;
  HIP_OR_FAIL(hipEventRecord(plan->ev[1], s->stream));
  HIP_OR_FAIL(hipEventRecord(plan->packed, s->stream));
  HIP_OR_FAIL(hipStreamWaitEvent(d->stream, plan->packed, 0));
  HIP_OR_FAIL(hipEventRecord(plan->ev[2], d->stream));
  r = hlb_halo_unpack_device(plan, plan->buf.p);
  if (r) return r;
  HIP_OR_FAIL(hipEventRecord(plan->ev[3], d->stream));
  plan->timed = true;
  return HLB_OK;
}

int hlb_halo_last_times(hlb_halo_plan* plan, double* pack_ms, double* unpack_ms) {
  if (!plan || !pack_ms || !unpack_ms) return fail(HLB_E_INVALID, "null argument");
  if (!plan->timed) return fail(HLB_E_STATE, "no exchange yet");
  HIP_OR_FAIL(hipEventSynchronize(plan->ev[3]));
  float a = 0, b = 0;
  HIP_OR_FAIL(hipEventElapsedTime(&a, plan->ev[0], plan->ev[1]));
  HIP_OR_FAIL(hipEventElapsedTime(&b, plan->ev[2], plan->ev[3]));
  *pack_ms = a;
  *unpack_ms = b;
  return HLB_OK;
}

int64_t hlb_site_count(const hlb_solver* s) { return s ? s->n : -1; }
int32_t hlb_lattice_q(const hlb_solver* s) { return s ? s->Q : -1; }
int64_t hlb_device_bytes(const hlb_solver* s) { return s ? s->device_bytes() : -1; }
void* hlb_stream(const hlb_solver* s) { return s ? (void*)s->stream : nullptr; }

int hlb_synchronize(hlb_solver* s) {
  if (!s) return fail(HLB_E_INVALID, "null solver");
  HIP_OR_FAIL(hipStreamSynchronize(s->stream));
  return HLB_OK;
}

}  // extern "C"
