// Internal interface between the host solver (solver.cpp, built with g++)
// and the HIP kernels (lbm_kernels.hip, built with hipcc).  Plain structs of
// device pointers; no HIP types so that the host side needs no device
// compiler.
#pragma once

#include <cstdint>

namespace hlb {

constexpr int kMaxMoments = 15;
constexpr int kMaxQ = 19;

// Device-side description of one solver's domain in the library's own
// layout.  Sites are ordered by 8^3 block in Morton order, z-fastest inside a
// block ("compact" order).  Distributions are structure-of-arrays with stride
// `np`, indexed by compact site with a 32-bit pull table (layout 0), by slot
// in dense 512-slot blocks with neighbours by arithmetic (layout 1), or by
// compact site with neighbours from the blocks' per-plane fluid masks
// (layout 2).
struct DevDomain {
  int layout = 1;   // 0 = pull table, 1 = dense blocks, 2 = compact blocks
  int64_t n = 0;    // fluid sites
  int64_t np = 0;   // stride of every per-direction state array
  int64_t npc = 0;  // stride of compact binary64 exports (n padded)
  // Block layout: 8 waves (x planes) per block.
  uint32_t nwaves = 0;
  const int32_t* bnbr = nullptr;       // [blocks][27] neighbour block or -1
  const uint64_t* wave_fluid = nullptr;  // [nwaves] lane is a fluid site
  const uint64_t* wave_bnd = nullptr;    // [nwaves] lane has a boundary record
  const uint32_t* wave_rec = nullptr;    // [nwaves] first boundary record
  const uint32_t* wave_cbase = nullptr;  // [nwaves] compact index of first fluid lane
  // [nwaves] OR over the plane's boundary sites of their BFL candidate links
  // (bits 0-31: inverse link a wall link, not an iolet link, link itself no
  // wall link) and of their post-step links (bits 32-63: wall link whose
  // inverse is no wall link)
  const uint64_t* wave_links = nullptr;
  // Block layouts: the planes holding fluid split into those without any
  // boundary site and those with one (Morton order within each list).
  const uint32_t* bulk_waves = nullptr;
  const uint32_t* bnd_waves = nullptr;
  uint32_t n_bulk_waves = 0, n_bnd_waves = 0;
  // Pull table (layout 0): src[(d-1)*np + x] is the offset inside array d
  // of the value streaming into slot d of site x: the site at x - c_d, or
  // x + (inv(d) - d) * np (the site's own inverse slot) for a cut link.
  const int32_t* src = nullptr;
  // Pull-table layout: 0 for a site with no wall/iolet link, else its
  // boundary record + 1.  Block layouts find the record from the wave's
  // boundary mask.
  const uint32_t* binfo = nullptr;
  int64_t nb = 0;                  // boundary records
  const uint32_t* b_wall = nullptr;   // [nb] wall link mask
  const uint32_t* b_iolet = nullptr;  // [nb] iolet link mask
  const int32_t* b_io = nullptr;      // [nb] iolet-site record or -1
  const float* b_q = nullptr;         // binary64 storage: the cut distances
                                      // of each record's wall links, packed
                                      // in direction order
  const uint32_t* b_qoff = nullptr;   // [nb] first entry of record r in b_q
  const uint16_t* b_q16 = nullptr;    // [nb][QS16] the same as q * 65535,
                                      // QS16 = Q-1 rounded up to a multiple of 8
                                      // (reduced-precision storage)
  int64_t ni = 0;                  // iolet-site records
  const int32_t* io_iolet = nullptr;  // [ni] global iolet index (inlets first)
  const int32_t* io_bc = nullptr;     // [ni] hlb_iolet_bc of the site's group
  const int32_t* io_site = nullptr;   // [ni] internal site index
  double* io_stash = nullptr;         // [ni] Nash: m.n/rho ; Ladd: rho
  const double* io_ladd = nullptr;    // [ni][Q-1][2] per-link velocity coeffs
  int32_t n_iolets = 0;
  const float* iolet_nf = nullptr;    // [n_iolets][3] normal rounded to float
  const double* iolet_n = nullptr;    // [n_iolets][3] normal
  const int32_t* iolet_kind = nullptr;  // [n_iolets] hlb_iolet_kind
  // Per-block relaxation of an ensemble (block layout; null otherwise):
  // -1/tau, its binary32 rounding and the MRT rate 1/tau of the block's
  // instance.
  const double* blk_omega = nullptr;
  const float* blk_omega_f = nullptr;
  const double* blk_s_tau = nullptr;
};

struct StepArgs {
  int Q = 19;
  int kernel = 0;     // hlb_kernel
  int wall = 0;       // hlb_wall
  int storage = 0;    // hlb_storage
  double omega = 0;   // -1/tau
  double rho0 = 1;    // shift for reduced-precision storage
  float omega_f = 0;
  double mrt_s_tau = 0;  // MRT: the relaxation rate 1/tau

  // Boundary values of the step being streamed: per iolet, two doubles
  // (Nash: ghost density; Ladd Womersley: e^{i w t}; Ladd parabolic: speed).
  const double* bv = nullptr;

  // Launch shape of the step kernel in the block layout: threads per
  // workgroup (64, 128 or 256), x planes per wave, workgroups in the grid
  // (0: one pass over the domain; fewer: a persistent grid that strides over
  // it) and the workgroup-to-XCD remap (0: hardware round-robin order; 1:
  // one contiguous range per XCD; K = 2^k: runs of K workgroups per XCD,
  // the eight XCDs on adjacent runs).  Default (set per storage by the
  // solver): short runs of workgroups per XCD (128 threads in runs of 2 for
  // binary64 and of 8 for binary32, 64 threads in runs of 16 for
  // binary16), so that the whole chip sweeps one region of the domain at a
  // time; one range per XCD measured 4-16 % slower on the vessel tree.
  uint32_t threads = 128;
  uint32_t planes = 1;
  uint32_t groups = 0;
  uint32_t xcd_remap = 8;
  // One-pass shapes: run the planes without boundary sites and those with
  // them as two launches of separately compiled kernels (1) or all planes
  // in one launch (0).  The bulk kernel needs far fewer registers (FP64
  // 84 instead of 110 VGPRs, FP32/FP16 45-48 instead of 76-78), but on the
  // vessel tree the two serial launches measured slower in every form
  // (FP64 7 %, FP32 13 %, FP16 1.5 %), so one launch is the default and the
  // split is a tuner candidate.
  uint32_t split = 0;
  // Block layout: 1 addresses the state with 64-bit element indices even
  // where 32-bit byte offsets reach every element (below 4 GiB per state
  // buffer, where the step uses them by default).
  uint32_t wide = 0;
};

// Launches; every function returns 0 or a HIP error code.
// Every site gets the post-stream distributions vals[Q] (binary64).
int launch_fill(const DevDomain& d, const StepArgs& a, const double* vals,
                void* state, void* stream);
int launch_step(const DevDomain& d, const StepArgs& a, const void* fin,
                void* fout, void* stream);
int launch_collide_only(const DevDomain& d, const StepArgs& a, const void* fin,
                        void* fout, void* stream);
// Post-stream distributions (the reference's fOld) as binary64 compact SoA
// [Q][npc].  `from_post_stream` copies fin when the state already holds them.
int launch_stream_only(const DevDomain& d, const StepArgs& a, const void* fin,
                       double* s_out, int from_post_stream, void* stream);
// Compact SoA [Q][np] binary64 <-> caller-order AoS [n][Q] through perm
// (perm[compact] = caller index).
int launch_soa_to_aos(int Q, int64_t n, int64_t np, const double* soa,
                      const int64_t* perm, double* aos, void* stream);
int launch_import(const DevDomain& d, const StepArgs& a, const double* aos,
                  const int64_t* perm, void* state, void* stream);
// Equilibrium of caller-order fields rho[n], momentum[n][3] into the state.
int launch_import_equilibrium(const DevDomain& d, const StepArgs& a, const double* rho,
                              const double* mom, const int64_t* perm, void* state,
                              void* stream);
// Moments of a binary64 SoA state: rho[np], u[3][np] (internal order).
int launch_moments(int Q, int64_t n, int64_t np, const double* soa, double* rho,
                   double* u, void* stream);
// Wall shear stress magnitude at nw wall sites (compact index site[i],
// normal[3*i..]) of a binary64 SoA state; temp[i] = -sqrt(2) * the stress
// parameter of the site's relaxation time.
int launch_wall_shear_stress(int Q, int64_t nw, int64_t np, const double* soa,
                             const int64_t* site, const float* normal, const double* temp,
                             double* out, void* stream);
// Fixed-order binary64 reductions over moments; partial results per block.
int launch_reduce(int Q, int64_t n, int64_t np, const double* soa,
                  const double* rho, const double* u, double* partials,
                  int nblocks, void* stream);
// Device-copy ceiling: Q arrays of `bytes` bytes copied from `in` to `out`
// in words of `width` bytes (4, 8 or 16), one word per thread and array,
// with the step's workgroup size and workgroup-to-XCD remap.  run_bytes 0:
// array d at byte offset d * stride_bytes; else the block layout's order,
// runs of run_bytes per direction and block.
int launch_copy_arrays(int Q, int64_t stride_bytes, int64_t bytes, int64_t run_bytes,
                       int width, const void* in, void* out, uint32_t threads, uint32_t remap,
                       void* stream);
// Halo exchange of a subdomain pair: pack (pack != 0) copies
// from[idx[k]] to to[k], unpack copies from[k] to to[idx[k]], for k < count,
// as raw stored words of esize bytes (8, 4 or 2).
int launch_halo(int64_t count, const int64_t* idx, int esize, const void* from, void* to,
                int pack, void* stream);
// Monitor statistics of the reference's fOld read straight from the state
// `fin` (post-stream when from_post_stream, else pulled as
// launch_stream_only does): partials[5*g..] = mass, min rho, max rho,
// max |u|, unstable of workgroup g < kMonitorGroups, and io_vals[4*r..] =
// rho, u of iolet-site record r < d.ni.
constexpr int kMonitorGroups = 1024;
int launch_monitor(const DevDomain& d, const StepArgs& a, const void* fin,
                   int from_post_stream, double* partials, double* io_vals, void* stream);

// The library's static device scratch (a file-scope __device__ array of
// kStaticScratchDoubles binary64 values, 8 MiB, one copy per GPU held until
// the process ends): the monitor partials followed by the values of up to
// kStaticIoletSites iolet sites, for the allocation-free monitor.
// scratch_address gives its address on `device`, which must be the current
// device on the first call for that device (hipGetSymbolAddress, cached per
// device); returns 0 or a HIP error code.
constexpr int64_t kStaticIoletSites = int64_t(1) << 18;
constexpr int64_t kStaticScratchDoubles = 5 * kMonitorGroups + 4 * kStaticIoletSites;
int scratch_address(int device, double** out);
// Stress tensor [6][np] (xx xy xz yy yz zz), von Mises [np] and shear rate
// [np] of every site of a binary64 SoA state (any output may be null); the
// tau of site x is tau_site[x] when given, else tau.
int launch_stress_fields(int Q, int64_t n, int64_t np, const double* soa, double tau,
                         const double* tau_site, double* t6, double* vm, double* sr,
                         void* stream);
// Traction and tangential traction [nw][3] at the wall sites.
int launch_traction(int Q, int64_t nw, int64_t np, const double* soa, const int64_t* site,
                    const float* normal, const double* tau_of, double* tr, double* tg,
                    void* stream);
// Per-chunk statistics of sites [range[2c], range[2c+1]): partials[5c..]
// = mass, min rho, max rho, max |u|, unstable; velocity change per chunk
// with its own 1/reference speed.
int launch_reduce_ranges(int Q, int64_t nchunks, int64_t np, const int64_t* range,
                         const double* soa, const double* rho, const double* u,
                         double* partials, void* stream);
int launch_velocity_change_ranges(int64_t nchunks, int64_t np, const int64_t* range,
                                  const double* u, const double* u_ref,
                                  const double* inv_ref, double* partials, void* stream);
int launch_velocity_change(int64_t n, int64_t np, const double* u,
                           const double* u_ref, double inv_ref, double* partials,
                           int nblocks, void* stream);
// Halo unpack: state[slot_index[i]] = in[i] (storage conversion applied);
// the direction of a state index is index / np, or (index / 512) % Q for the
// block layout (blockwise != 0).
int launch_scatter_pairs(int64_t count, int64_t np, int Q, int blockwise,
                         const int64_t* slot_index, const double* in, int storage, double rho0,
                         const double* w, void* state, void* stream);
int launch_gather_pairs(int64_t count, int64_t np, int Q, int blockwise,
                        const int64_t* slot_index, const void* state, int storage, double rho0,
                        const double* w, double* out, void* stream);

}  // namespace hlb
