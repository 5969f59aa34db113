/*
 * hemelb_amd.h -- C ABI of the MI355X lattice-Boltzmann timestep.
 *
 * This library replaces the reference's compile-time LB hot path: the
 * hemelb::lb::LBM<Traits> IteratedAction (Code/lb/lb.h:30-142,
 * Code/lb/lb.hpp:74-309) with its collision kernels (LBGK, MRT), streamers
 * (bulk, simple bounce-back, BFL, Nash zeroth-order pressure, Ladd) and
 * lattices (D3Q15, D3Q19), the FieldData distributions it updates
 * (Code/geometry/FieldData.h:27-214) and the BoundaryValues / InOutLet
 * objects that feed the iolet streamers (Code/lb/iolets/BoundaryValues.cc,
 * InOutLetCosine.cc, InOutLetParabolicVelocity.cc,
 * InOutLetWomersleyVelocity.cc).
 *
 * Conventions
 *  - Plain C types only; arrays are caller-owned host memory unless stated.
 *  - Every function returns HLB_OK (0) or a negative HLB_E* code; no C++
 *    exception crosses the boundary.  hlb_last_error() gives a message for
 *    the last failure on the calling thread.
 *  - A solver is used from one host thread at a time; all of its device work
 *    is ordered on its own HIP stream.
 *  - After HLB_E_DEVICE from a call that advances or overwrites the state
 *    (hlb_step, hlb_time_steps, the set/import calls), the state and time
 *    step are unspecified: destroy the solver, or set the distributions and
 *    time step again before further steps.
 *  - Site arrays describe the reference's geometry::Domain for one rank
 *    (Code/geometry/Domain.h:53-531): n sites in the caller's local order,
 *    with wall/iolet link bitmasks (bit d-1 for lattice direction d,
 *    Code/geometry/SiteDataBare.cc:147-163), one iolet id per site and cut
 *    distances per lattice direction (Domain::GetCutDistance, Domain.h:405).
 *  - Distributions crossing the ABI are in the reference's FieldData layout:
 *    double f[n][Q] in the caller's site order.
 */
#ifndef HEMELB_AMD_H
#define HEMELB_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HLB_ABI_VERSION 1

enum hlb_status {
  HLB_OK = 0,
  HLB_E_INVALID = -1,   /* invalid argument or inconsistent geometry */
  HLB_E_DEVICE = -2,    /* HIP runtime failure */
  HLB_E_NOMEM = -3,     /* allocation failure */
  HLB_E_STATE = -4      /* call not valid in the solver's current state */
};

/* Lattices (Code/lb/lattices/D3Q15.h, D3Q19.h). */
enum hlb_lattice { HLB_D3Q15 = 15, HLB_D3Q19 = 19 };
/* Collision kernels (Code/lb/kernels/LBGK.h, MRT.h with the d'Humieres
 * bases DHumieresD3Q15MRTBasis / DHumieresD3Q19MRTBasis). */
enum hlb_kernel { HLB_LBGK = 0, HLB_MRT = 1 };
/* Wall link streamers (Code/lb/streamers/SimpleBounceBack.h,
 * BouzidiFirdaousLallemand.h). */
enum hlb_wall { HLB_WALL_SBB = 0, HLB_WALL_BFL = 1 };
/* Iolet link streamers (Code/lb/streamers/NashZerothOrderPressure.h,
 * LaddIolet.h). */
enum hlb_iolet_bc { HLB_IOLET_NASH = 0, HLB_IOLET_LADD = 1 };
/* Boundary-value objects (Code/lb/iolets/InOutLetCosine.h,
 * InOutLetParabolicVelocity.h, InOutLetWomersleyVelocity.h). */
enum hlb_iolet_kind { HLB_IOLET_COSINE = 0, HLB_IOLET_PARABOLIC = 1,
                      HLB_IOLET_WOMERSLEY = 2 };
/* Site types (geometry::SiteType). */
enum hlb_site_type { HLB_SITE_FLUID = 1, HLB_SITE_INLET = 2, HLB_SITE_OUTLET = 3 };
/* Storage of the distributions in HBM.  HLB_STORE_FP64 is the plain method
 * (all arithmetic in binary64, bit-compatible with the reference).  FP32 and
 * FP16 store f_i - w_i*rho0 and collide in binary32, with BFL cut distances
 * held as 16-bit fixed point (q*65535); iolet boundary values and all
 * reductions stay binary64. */
enum hlb_storage { HLB_STORE_FP64 = 0, HLB_STORE_FP32 = 1, HLB_STORE_FP16 = 2 };
/* Layout of the distributions in HBM.  HLB_LAYOUT_BLOCKS gives every 8^3
 * block that holds fluid 512 slots (Morton block order, z fastest), stores a
 * block's directions together ([block][direction][slot]) and finds
 * neighbours by slot arithmetic; HLB_LAYOUT_PULL_TABLE packs the fluid sites
 * and reads a 32-bit pull index per site and direction;
 * HLB_LAYOUT_COMPACT_BLOCKS packs the fluid sites in the same order and finds
 * neighbours from the per-plane fluid masks of the 8^3 blocks (no index
 * table, no empty slots).  Results are identical. */
enum hlb_layout { HLB_LAYOUT_BLOCKS = 0, HLB_LAYOUT_PULL_TABLE = 1,
                  HLB_LAYOUT_COMPACT_BLOCKS = 2 };

typedef struct hlb_solver hlb_solver;

typedef struct {
  int32_t lattice;     /* hlb_lattice */
  int32_t kernel;      /* hlb_kernel */
  int32_t wall;        /* hlb_wall */
  int32_t inlet_bc;    /* hlb_iolet_bc used for sites of type INLET */
  int32_t outlet_bc;   /* hlb_iolet_bc used for sites of type OUTLET */
  int32_t storage;     /* hlb_storage */
  double tau;          /* LbmParameters::GetTau() (LbmParameters.h:158-163) */
  double rho0;         /* reference density for FP32/FP16 storage (ignored
                          for FP64) */
  int32_t device;      /* HIP device ordinal */
  int32_t layout;      /* hlb_layout */
} hlb_params;

typedef struct {
  int64_t n_sites;
  const int32_t* coords;       /* [n][3] global site coordinates            */
  const uint32_t* site_type;   /* [n] hlb_site_type                         */
  const uint32_t* wall_mask;   /* [n] bit d-1: link d crosses a wall        */
  const uint32_t* iolet_mask;  /* [n] bit d-1: link d crosses an iolet      */
  const int32_t* iolet_id;     /* [n] index into inlets (INLET sites) or
                                  outlets (OUTLET sites); ignored otherwise */
  const float* cut_distance;   /* [n][Q-1] cut distance of link d (as read
                                  from the .gmy file); required for BFL     */
  const float* wall_normal;    /* [n][3] optional (wall shear stress)       */
} hlb_domain;

typedef struct {
  int32_t kind;                /* hlb_iolet_kind */
  int32_t reserved;
  double position[3];          /* lattice units (InOutLet::SetPosition) */
  double normal[3];            /* normalised by the library as
                                  InOutLet::SetNormal does */
  /* cosine pressure: density = mean + amp*cos(2*pi*t/period + phase),
   * linear warm-up from the minimum simulation density over `warmup`
   * steps (InOutLetCosine.cc:26-43) */
  double density_mean, density_amp, phase, period;
  int64_t warmup;
  /* velocity iolets */
  double radius;               /* lattice units */
  double max_speed;            /* parabolic: centreline speed */
  double pressure_gradient_amplitude;  /* Womersley (period shared) */
  double womersley_number;
} hlb_iolet;

/* Reductions over the current state (the reference's fOld), all binary64:
 * mass = sum of density; min/max density; max |u|; `stable` is 0 if any
 * distribution is not > 0 (StabilityTester.h:91-118 criterion);
 * iolet_flux[k] = sum over the sites touching iolet k of (momentum . normal),
 * inlets first then outlets (the first 64 iolets; n_iolets counts all). */
typedef struct {
  double mass;
  double density_min, density_max;
  double velocity_max;
  int32_t stable;
  int32_t n_iolets;
  double iolet_flux[64];
} hlb_monitor;

int hlb_abi_version(void);
const char* hlb_last_error(void);

/* Build a solver for one rank's domain.  The iolet arrays are the inlet and
 * outlet BoundaryValues (Code/lb/iolets/BoundaryValues.cc:13-58); the
 * minimum simulation density passed to each iolet is computed as
 * LBM::PrepareBoundaryObjects does (Code/lb/lb.hpp:127-152).  The state is
 * undefined until hlb_set_equilibrium or hlb_set_distributions. */
int hlb_create(const hlb_params* params, const hlb_domain* domain,
               const hlb_iolet* inlets, int32_t n_inlets,
               const hlb_iolet* outlets, int32_t n_outlets,
               hlb_solver** out);
void hlb_destroy(hlb_solver* s);

/* EquilibriumInitialCondition::SetFs (Code/lb/InitialCondition.hpp:39-52):
 * every site gets feq(density, momentum). Resets the time to 1. */
int hlb_set_equilibrium(hlb_solver* s, double density, const double momentum[3]);
/* SimulationState time step (1-indexed, SimulationState.cc:15-49). */
int hlb_set_time(hlb_solver* s, uint64_t time_step);
uint64_t hlb_get_time(const hlb_solver* s);

/* Advance n time steps (n x SimulationMaster::DoTimeStep for the LB actor:
 * stream-and-collide, PostStep, swap). */
int hlb_step(hlb_solver* s, int64_t n);

/* fOld in the caller's order, double[n][Q] (checkpoint export/import,
 * Code/extraction/LbDataSourceIterator.cc:76-80,
 * Code/extraction/LocalDistributionInput.cc:38-80). */
int hlb_get_distributions(hlb_solver* s, double* f_out);
int hlb_set_distributions(hlb_solver* s, const double* f_in);
/* fOld = feq(density[i], momentum[i]) at every site i, in the caller's order
 * (density [n], momentum [n][3]), with the arithmetic of
 * Lattice::CalculateFeq (Lattice.h:236-320): the equilibrium initial
 * condition of InitialCondition.hpp:39-52 generalised from one uniform
 * value to a field (e.g. a predicted flow field).  `device` != 0: the
 * arrays are device pointers, read on the solver's stream; the work that
 * writes them must be complete before the call.  The time is left
 * unchanged. */
int hlb_set_equilibrium_fields(hlb_solver* s, const double* density, const double* momentum,
                               int32_t device);

/* Density [n] and velocity [n][3] of fOld in the caller's order (either may
 * be NULL), as MacroscopicPropertyCache / LbDataSourceIterator compute them. */
int hlb_get_macroscopic(hlb_solver* s, double* density, double* velocity);

/* Wall shear stress magnitude [n] of fOld in the caller's order: at sites
 * with wall links, Lattice::CalculateWallShearStressMagnitude
 * (Lattice.h:652-689) with the stress parameter of LbmParameters
 * (LbmParameters.h:34-39), as UpdateCachePostCollision fills
 * wallShearStressMagnitudeCache (Code/lb/streamers/Common.h:37-50); DBL_MAX
 * (the reference's NO_VALUE, constants.h:47) elsewhere.  Needs
 * hlb_domain.wall_normal at create. */
int hlb_get_wall_shear_stress(hlb_solver* s, double* stress);

/* The other stress outputs of the property cache (Code/lb/streamers/
 * Common.h:52-124) of fOld in the caller's order, each optional (NULL):
 * stress tensor [n][6] (xx, xy, xz, yy, yz, zz; Lattice::CalculateStressTensor,
 * Lattice.h:630-645), von Mises stress [n] (508-552), shear rate [n]
 * (728-753), traction [n][3] and its tangential projection [n][3] at sites
 * with wall links, zero elsewhere (560-602; needs wall normals). */
int hlb_get_stress(hlb_solver* s, double* tensor, double* von_mises, double* shear_rate,
                   double* traction, double* tangential);

int hlb_reduce(hlb_solver* s, hlb_monitor* out);

/* hlb_reduce without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernel as hlb_reduce fills, instead of solver-owned buffers, the library's
 * static device scratch (8 MiB per GPU, held until the process ends), so the
 * results are identical to hlb_reduce's.  Limit: at most 262144 iolet sites
 * (HLB_E_INVALID beyond).  Every solver on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (two host threads, or two solvers whose
 * streams run concurrently) race; the call synchronises the solver's stream
 * before it returns. */
int hlb_reduce_noalloc(hlb_solver* s, hlb_monitor* out);

/* ---- Batched ensemble --------------------------------------------------
 * No counterpart in the reference, which runs one vessel per job.  Several
 * independent vessels share one solver: instance[i] in [0, n_instances)
 * gives the vessel of site i and tau[k] its relaxation time.  The sites of
 * different instances must not be linked and must not share an 8^3 block
 * (keep one empty block between vessels); iolet ids index the common inlet
 * and outlet arrays.  Needs a block layout.  One launch per step updates
 * every instance. */
int hlb_create_ensemble(const hlb_params* params, const hlb_domain* domain,
                        const int32_t* instance, int32_t n_instances, const double* tau,
                        const hlb_iolet* inlets, int32_t n_inlets,
                        const hlb_iolet* outlets, int32_t n_outlets, hlb_solver** out);
int32_t hlb_instance_count(const hlb_solver* s);

typedef struct {
  double mass;
  double density_min, density_max;
  double velocity_max;
  int32_t stable;
  int32_t reserved;
  int64_t sites;
} hlb_instance_monitor;

/* hlb_reduce's statistics for every instance (out[n_instances]); a plain
 * solver is one instance. */
int hlb_reduce_instances(hlb_solver* s, hlb_instance_monitor* out);
/* hlb_velocity_change per instance, each against its own reference speed. */
int hlb_velocity_change_instances(hlb_solver* s, const double* reference_speed,
                                  double* out);

/* Launch shape of the step kernel (block layouts): threads per workgroup
 * (64, 128, 256), x planes (64-site tiles) per wave, workgroups in the grid
 * (0: one pass over the domain; a positive count gives a persistent grid
 * that strides over it) and the workgroup-to-XCD remap: 0 the hardware's
 * round-robin order, 1 one contiguous range of the domain per XCD, K in
 * 2..4096 runs of K consecutive workgroups per XCD with the eight XCDs on
 * adjacent runs (the grid is padded to a multiple of 8K).
 * hlb_autotune times a set of shapes for `steps` steps each on the current
 * state, restores the state and time, and keeps the fastest. */
typedef struct {
  int32_t threads;
  int32_t planes;
  int32_t groups;
  int32_t xcd_remap;
  int32_t split;       /* one-pass shapes: 1 runs the planes without boundary
                          sites and those with them as two launches of
                          separately compiled kernels, 0 one launch */
  int32_t wide;        /* 1: 64-bit element addressing even where 32-bit
                          byte offsets reach the whole state (below 4 GiB
                          per state buffer, where 0 uses them) */
} hlb_launch_shape;
int hlb_set_launch_shape(hlb_solver* s, const hlb_launch_shape* shape);
int hlb_get_launch_shape(const hlb_solver* s, hlb_launch_shape* shape);
int hlb_autotune(hlb_solver* s, int32_t steps, hlb_launch_shape* best,
                 double* best_ms_per_step);

/* Convergence monitor: hlb_snapshot_velocity stores the current velocity
 * field on the device; hlb_velocity_change returns
 * max over sites of |u - u_snapshot| / reference_speed
 * (StabilityTester::ComputeRelativeDifference, StabilityTester.h:160-183). */
int hlb_snapshot_velocity(hlb_solver* s);
int hlb_velocity_change(hlb_solver* s, double reference_speed, double* out);

/* Device-copy ceiling of the step's traffic (a measurement aid; no
 * counterpart in the reference): the solver's Q state arrays copied from the
 * current state buffer into the other one (free between steps, so the state
 * is kept), reps times per variant, one word per thread and array with the
 * loads issued before streamed stores.  Variants: the words of one array per
 * direction (order 0) or in the block layout's order (order 1: runs of 512
 * slots per direction and block), words of the storage's element size (4
 * bytes at least) or 16 bytes, 128- or 256-thread workgroups, the hardware
 * workgroup order or runs of 8 workgroups per XCD.  Reports the fastest:
 * read + write bytes per second, milliseconds per copy, bytes per copy and
 * its launch. */
typedef struct {
  double bytes_per_s;
  double ms;
  int64_t bytes;
  int32_t threads;
  int32_t xcd_remap;
  int32_t width;
  int32_t order;
} hlb_copy_result;
int hlb_copy_ceiling(hlb_solver* s, int32_t reps, hlb_copy_result* out);

/* Halo exchange for a subdomain (Code/geometry/FieldData.cc:27-49): pack the
 * pending post-collision distributions of the given (site, direction) pairs
 * (the values the reference sends to the neighbouring rank) into a host
 * buffer.  `pairs` is int64[count][2] of (caller site index, direction). */
int hlb_halo_pack(hlb_solver* s, const int64_t* pairs, int64_t count,
                  double* host_buffer);
/* The inverse of hlb_halo_pack: the values streaming in from a neighbouring
 * subdomain replace the pending post-collision values of the given (site,
 * direction) pairs, which must be distinct (the receive side of
 * FieldData::SendAndReceive / CopyReceived, FieldData.cc:27-49). */
int hlb_halo_unpack(hlb_solver* s, const int64_t* pairs, int64_t count,
                    const double* host_buffer);

/* ---- Subdomains and halo plans ------------------------------------------
 * One rank's subdomain with ghost sites (the reference's decomposed Domain:
 * edge sites and the shared distributions of Domain.cc:166-179 and 247-284).
 * `ghosts` lists the sites of neighbouring subdomains that this domain's
 * sites pull from across the cut (coords, wall_mask and iolet_mask are read;
 * the other arrays may be NULL).  Ghost sites get state slots that the step
 * reads and a halo exchange fills; the step never updates them.  The block
 * and pull-table layouts support ghosts; ensembles do not. */
int hlb_create_subdomain(const hlb_params* params, const hlb_domain* domain,
                         const hlb_domain* ghosts, const hlb_iolet* inlets, int32_t n_inlets,
                         const hlb_iolet* outlets, int32_t n_outlets, hlb_solver** out);

/* A halo plan carries src's post-collision values into dst's ghost slots:
 * every (ghost site g of dst owned by src, direction e) with g + c_e a site
 * of dst and link e of g not cut -- what FieldData::SendAndReceive sends and
 * CopyReceived places (FieldData.cc:27-49) in the reference's PreSend /
 * PreReceive / PostReceive sequence (lb.hpp:175-251).  Both solvers are
 * subdomains on one device with the same lattice and storage.  The indices
 * are computed once; an exchange allocates nothing and does not block the
 * host.  Run it after every step of both solvers (the first step included)
 * and before the next step or export of dst. */
typedef struct hlb_halo_plan hlb_halo_plan;
int hlb_halo_plan_create(hlb_solver* src, hlb_solver* dst, hlb_halo_plan** out);
void hlb_halo_plan_destroy(hlb_halo_plan* plan);
/* Number of values one exchange moves (each one stored word: 8, 4 or 2
 * bytes for FP64, FP32, FP16 storage). */
int64_t hlb_halo_plan_count(const hlb_halo_plan* plan);
/* pack on src's stream into the plan's device buffer, an event, unpack on
 * dst's stream. */
int hlb_halo_exchange(hlb_halo_plan* plan);
/* The two halves for a caller that moves the buffer itself (e.g. between
 * GPUs): pack into / unpack from a device buffer of count stored words, on
 * src's / dst's stream; the caller orders the unpack after the pack. */
int hlb_halo_pack_device(hlb_halo_plan* plan, void* device_buffer);
int hlb_halo_unpack_device(hlb_halo_plan* plan, const void* device_buffer);
/* Device milliseconds of the last hlb_halo_exchange's pack and unpack. */
int hlb_halo_last_times(hlb_halo_plan* plan, double* pack_ms, double* unpack_ms);

/* Device timing of n steps with HIP events on the solver's stream
 * (milliseconds, launch gaps included). */
int hlb_time_steps(hlb_solver* s, int64_t n, double* elapsed_ms);

/* Introspection. */
int64_t hlb_site_count(const hlb_solver* s);
int32_t hlb_lattice_q(const hlb_solver* s);
/* Bytes of device memory held by the solver. */
int64_t hlb_device_bytes(const hlb_solver* s);
/* The HIP stream (hipStream_t) the solver launches on. */
void* hlb_stream(const hlb_solver* s);
int hlb_synchronize(hlb_solver* s);

/* Stand-alone helpers mirroring the reference's parameter set-up. */
/* LbmParameters::CalcDerivedParams (LbmParameters.h:158-163). */
double hlb_tau_from_physical(double time_step_s, double voxel_size_m,
                             double fluid_density_kgm3, double viscosity_pas);
/* util::UnitConverter (Code/util/UnitConverter.cc:14-48). */
double hlb_pressure_to_lattice(double p_mmhg, double reference_mmhg,
                               double time_step_s, double voxel_size_m,
                               double fluid_density_kgm3);
double hlb_pressure_difference_to_lattice(double dp_mmhg, double time_step_s,
                                          double voxel_size_m,
                                          double fluid_density_kgm3);
double hlb_pressure_gradient_to_lattice(double pg_mmhg_per_m,
                                        double time_step_s,
                                        double voxel_size_m,
                                        double fluid_density_kgm3);

#ifdef __cplusplus
}
#endif
#endif
