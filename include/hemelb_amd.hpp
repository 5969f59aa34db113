/*
 * hemelb_amd.hpp -- header-only C++ layer over the C ABI (hemelb_amd.h) that
 * keeps the reference's compile-time LB API, so that a HemeLB build selects
 * the GPU timestep with the same type names it already uses:
 *
 *   hemelb::Traits<Lattice, KERNEL, COLLISION, STREAMER, WALL_BOUNDARY,
 *                  INLET_BOUNDARY, OUTLET_BOUNDARY>        (Code/Traits.h:17-39)
 *   hemelb::lb::LBM<Traits>, an IteratedAction            (Code/lb/lb.h:30-142)
 *   hemelb::lb::D3Q15, D3Q19                               (Code/lb/lattices/)
 *   hemelb::lb::LBGK<L>, MRT<DHumieres...MRTBasis>         (Code/lb/kernels/)
 *   hemelb::lb::Normal<K>                                  (Code/lb/collisions/Normal.h)
 *   hemelb::lb::StreamerTypeFactory<WallLink, IoletLink>,
 *     BounceBackLink, BouzidiFirdaousLallemandLink,
 *     NashZerothOrderPressureLink, LaddIoletLink, NullLink (Code/lb/streamers/)
 *
 * The template arguments are tags: each maps to the hlb_* codes of the
 * compiled GPU variant, and the mapping is checked at compile time (a
 * combination the library does not provide does not compile).  Errors of
 * the library are thrown as hemelb::Exception, as the reference reports
 * failures (Code/Exception.h).  The whole step (collide, stream, PostStep,
 * swap) runs in PreSend; on one rank the other IteratedAction phases have
 * nothing to do.
 */
#ifndef HEMELB_AMD_HPP
#define HEMELB_AMD_HPP

#include "hemelb_amd.h"
#include "hemelb_geometry.h"

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace hemelb {

using site_t = int64_t;
using distribn_t = double;

// Failures of the library and of the layer (Code/Exception.h).
class Exception : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

namespace lb {

// ---- lattices (Code/lb/lattices/D3Q15.h, D3Q19.h) ----
struct D3Q15 {
  static constexpr int NUMVECTORS = 15;
};
struct D3Q19 {
  static constexpr int NUMVECTORS = 19;
};

// ---- collision kernels (Code/lb/kernels/LBGK.h:18-19, MRT.h:39-40) ----
struct DHumieresD3Q15MRTBasis {
  using Lattice = D3Q15;
};
struct DHumieresD3Q19MRTBasis {
  using Lattice = D3Q19;
};
template <class L>
struct LBGK {
  using LatticeType = L;
  static constexpr int32_t code = HLB_LBGK;
};
template <class MomentBasis>
struct MRT {
  using LatticeType = typename MomentBasis::Lattice;
  static constexpr int32_t code = HLB_MRT;
};
// MRT with the d'Humieres basis of lattice L: the kernel the reference's
// build picks for KERNEL=MRT (Code/lb/Kernels.h, get_default_kernel).
template <class L> struct DHumieresBasis;
template <> struct DHumieresBasis<D3Q15> { using type = DHumieresD3Q15MRTBasis; };
template <> struct DHumieresBasis<D3Q19> { using type = DHumieresD3Q19MRTBasis; };
template <class L>
using DHumieresMRT = MRT<typename DHumieresBasis<L>::type>;

// ---- collision (Code/lb/collisions/Normal.h:16-17) ----
template <class K>
struct Normal {
  using KernelType = K;
};

// ---- link streamers (Code/lb/streamers/*.h); -1: no link of that kind ----
template <class C>
struct NullLink {
  static constexpr int32_t wall = -1, iolet = -1;
};
template <class C>
struct BounceBackLink {
  static constexpr int32_t wall = HLB_WALL_SBB, iolet = -1;
};
template <class C>
struct BouzidiFirdaousLallemandLink {
  static constexpr int32_t wall = HLB_WALL_BFL, iolet = -1;
};
template <class C>
struct NashZerothOrderPressureLink {
  static constexpr int32_t wall = -1, iolet = HLB_IOLET_NASH;
};
template <class C>
struct LaddIoletLink {
  static constexpr int32_t wall = -1, iolet = HLB_IOLET_LADD;
};

// A streamer built from a wall link and an iolet link
// (Code/lb/streamers/StreamerTypeFactory.h:48-108).
template <class WallLink, class IoletLink>
struct StreamerTypeFactory {
  static constexpr int32_t wall = WallLink::wall;
  static constexpr int32_t iolet = IoletLink::iolet;
};
// The streamers the reference's build options name (Code/lb/Streamers.h:22-67).
template <class C>
using BulkStreamer = StreamerTypeFactory<NullLink<C>, NullLink<C>>;
template <class C>
using SimpleBounceBack = StreamerTypeFactory<BounceBackLink<C>, NullLink<C>>;
template <class C>
using BouzidiFirdaousLallemand = StreamerTypeFactory<BouzidiFirdaousLallemandLink<C>, NullLink<C>>;
template <class C>
using NashZerothOrderPressure = StreamerTypeFactory<NullLink<C>, NashZerothOrderPressureLink<C>>;
template <class C>
using LaddIolet = StreamerTypeFactory<NullLink<C>, LaddIoletLink<C>>;

}  // namespace lb

// Compile-time choices of a build (Code/Traits.h:17-39); the defaults are
// the reference's CMake defaults (D3Q15, LBGK, simple bounce-back, Nash
// zeroth-order pressure iolets).
template <class LATTICE = lb::D3Q15, template <class> class KERNEL = lb::LBGK,
          template <class> class COLLISION = lb::Normal,
          template <class> class STREAMER = lb::BulkStreamer,
          template <class> class WALL_BOUNDARY = lb::SimpleBounceBack,
          template <class> class INLET_BOUNDARY = lb::NashZerothOrderPressure,
          template <class> class OUTLET_BOUNDARY = lb::NashZerothOrderPressure>
struct Traits {
  using Lattice = LATTICE;
  using Kernel = KERNEL<Lattice>;
  using Collision = COLLISION<Kernel>;
  using Streamer = STREAMER<Collision>;
  using WallBoundary = WALL_BOUNDARY<Collision>;
  using InletBoundary = INLET_BOUNDARY<Collision>;
  using OutletBoundary = OUTLET_BOUNDARY<Collision>;
};

namespace lb {

// The hlb_params codes of a Traits, with the combinations the library
// provides checked at compile time.
template <class T>
struct TraitsCodes {
  using L = typename T::Lattice;
  static_assert(std::is_same<L, D3Q15>::value || std::is_same<L, D3Q19>::value,
                "the GPU timestep provides D3Q15 and D3Q19");
  static_assert(std::is_same<typename T::Kernel::LatticeType, L>::value,
                "the kernel is built for another lattice");
  static_assert(std::is_same<typename T::Collision, Normal<typename T::Kernel>>::value,
                "the GPU timestep provides the Normal collision");
  static_assert(T::Streamer::wall == -1 && T::Streamer::iolet == -1,
                "bulk sites use the bulk streamer");
  static_assert(T::WallBoundary::wall >= 0 && T::WallBoundary::iolet == -1,
                "WALL_BOUNDARY must be a wall streamer (SBB or BFL)");
  static_assert(T::InletBoundary::iolet >= 0 && T::InletBoundary::wall == -1,
                "INLET_BOUNDARY must be an iolet streamer (Nash or Ladd)");
  static_assert(T::OutletBoundary::iolet >= 0 && T::OutletBoundary::wall == -1,
                "OUTLET_BOUNDARY must be an iolet streamer (Nash or Ladd)");
  static constexpr int32_t lattice = L::NUMVECTORS;
  static constexpr int32_t kernel = T::Kernel::code;
  static constexpr int32_t wall = T::WallBoundary::wall;
  static constexpr int32_t inlet = T::InletBoundary::iolet;
  static constexpr int32_t outlet = T::OutletBoundary::iolet;
};

inline void check(int rc, const char* what) {
  if (rc != HLB_OK) throw Exception(std::string(what) + ": " + hlb_last_error());
}

// tau from the physical time step and voxel size (LbmParameters.h:158-163).
class LbmParameters {
 public:
  explicit LbmParameters(double tau) : tau_(tau) {}
  LbmParameters(double time_step_s, double voxel_size_m, double fluid_density_kgm3,
                double viscosity_pas)
      : tau_(hlb_tau_from_physical(time_step_s, voxel_size_m, fluid_density_kgm3,
                                   viscosity_pas)) {}
  double GetTau() const { return tau_; }
  double GetOmega() const { return -1.0 / tau_; }

 private:
  double tau_;
};

// ---- boundary-value objects (Code/lb/iolets/) as hlb_iolet records ----
inline hlb_iolet iolet_base(int32_t kind, const std::array<double, 3>& position,
                            const std::array<double, 3>& normal) {
  hlb_iolet s{};
  s.kind = kind;
  for (int a = 0; a < 3; ++a) {
    s.position[a] = position[a];
    s.normal[a] = normal[a];
  }
  return s;
}
// Cosine pressure (InOutLetCosine.cc:26-43).
inline hlb_iolet InOutLetCosine(const std::array<double, 3>& position,
                                const std::array<double, 3>& normal, double density_mean,
                                double density_amp, double period, double phase = 0.0,
                                int64_t warmup = 0) {
  hlb_iolet s = iolet_base(HLB_IOLET_COSINE, position, normal);
  s.density_mean = density_mean;
  s.density_amp = density_amp;
  s.period = period;
  s.phase = phase;
  s.warmup = warmup;
  return s;
}
// Parabolic velocity (InOutLetParabolicVelocity.cc).
inline hlb_iolet InOutLetParabolicVelocity(const std::array<double, 3>& position,
                                           const std::array<double, 3>& normal,
                                           double radius, double max_speed,
                                           int64_t warmup = 0) {
  hlb_iolet s = iolet_base(HLB_IOLET_PARABOLIC, position, normal);
  s.radius = radius;
  s.max_speed = max_speed;
  s.warmup = warmup;
  return s;
}
// Womersley velocity (InOutLetWomersleyVelocity.cc).
inline hlb_iolet InOutLetWomersleyVelocity(const std::array<double, 3>& position,
                                           const std::array<double, 3>& normal,
                                           double radius, double pressure_gradient_amplitude,
                                           double period, double womersley_number) {
  hlb_iolet s = iolet_base(HLB_IOLET_WOMERSLEY, position, normal);
  s.radius = radius;
  s.pressure_gradient_amplitude = pressure_gradient_amplitude;
  s.period = period;
  s.womersley_number = womersley_number;
  return s;
}

// One rank's geometry::Domain as the site arrays of hlb_domain, with the
// iolet planes the geometry carries (lattice projection of a .gmy file or of
// a generated vessel).
struct DomainArrays {
  int Q = 0;
  std::vector<int32_t> coords;
  std::vector<uint32_t> site_type, wall_mask, iolet_mask;
  std::vector<int32_t> iolet_id;
  std::vector<float> cut, normal;
  std::vector<hlbg_iolet> inlets, outlets;

  site_t size() const { return (site_t)site_type.size(); }

  static DomainArrays Project(const hlbg_geometry* g, int Q) {
    DomainArrays d;
    d.Q = Q;
    const int64_t n = hlbg_fluid_site_count(g);
    if (n <= 0) throw Exception("geometry has no fluid sites");
    d.coords.resize(3 * n);
    d.site_type.resize(n);
    d.wall_mask.resize(n);
    d.iolet_mask.resize(n);
    d.iolet_id.resize(n);
    d.cut.resize((size_t)(Q - 1) * n);
    d.normal.resize(3 * n);
    if (hlbg_project(g, Q, d.coords.data(), d.site_type.data(), d.wall_mask.data(),
                     d.iolet_mask.data(), d.iolet_id.data(), d.cut.data(),
                     d.normal.data()) != 0)
      throw Exception(std::string("hlbg_project: ") + hlbg_last_error());
    for (int outlet = 0; outlet < 2; ++outlet)
      for (int32_t i = 0; i < hlbg_iolet_count(g, outlet); ++i) {
        hlbg_iolet io{};
        if (hlbg_get_iolet(g, outlet, i, &io) != 0)
          throw Exception(std::string("hlbg_get_iolet: ") + hlbg_last_error());
        (outlet ? d.outlets : d.inlets).push_back(io);
      }
    return d;
  }

  hlb_domain view() const {
    hlb_domain v{};
    v.n_sites = size();
    v.coords = coords.data();
    v.site_type = site_type.data();
    v.wall_mask = wall_mask.data();
    v.iolet_mask = iolet_mask.data();
    v.iolet_id = iolet_id.data();
    v.cut_distance = cut.data();
    v.wall_normal = normal.empty() ? nullptr : normal.data();
    return v;
  }
};

enum class Storage : int32_t { FP64 = HLB_STORE_FP64, FP32 = HLB_STORE_FP32, FP16 = HLB_STORE_FP16 };

// The LB actor of one rank on one GPU (Code/lb/lb.h:30-142).  Owns the
// device state; single-thread use, all work on the solver's HIP stream.
template <class TRAITS = hemelb::Traits<>>
class LBM {
 public:
  using Traits = TRAITS;
  using LatticeType = typename Traits::Lattice;
  using Codes = TraitsCodes<Traits>;

  LBM(const LbmParameters& params, const DomainArrays& domain,
      const std::vector<hlb_iolet>& inlets, const std::vector<hlb_iolet>& outlets,
      Storage storage = Storage::FP64, int32_t device = 0, double rho0 = 1.0)
      : params_(params), n_(domain.size()) {
    if (domain.Q != LatticeType::NUMVECTORS)
      throw Exception("domain was projected for another lattice");
    hlb_params p{};
    p.lattice = Codes::lattice;
    p.kernel = Codes::kernel;
    p.wall = Codes::wall;
    p.inlet_bc = Codes::inlet;
    p.outlet_bc = Codes::outlet;
    p.storage = (int32_t)storage;
    p.tau = params.GetTau();
    p.rho0 = rho0;
    p.device = device;
    p.layout = HLB_LAYOUT_BLOCKS;
    const hlb_domain d = domain.view();
    check(hlb_create(&p, &d, inlets.data(), (int32_t)inlets.size(), outlets.data(),
                     (int32_t)outlets.size(), &s_),
          "hlb_create");
  }
  ~LBM() {
    if (s_) hlb_destroy(s_);
  }
  LBM(const LBM&) = delete;
  LBM& operator=(const LBM&) = delete;
  LBM(LBM&& o) noexcept : params_(o.params_), n_(o.n_), s_(std::exchange(o.s_, nullptr)) {}

  // IteratedAction phases (Code/net/IteratedAction.h): the GPU step runs
  // whole in PreSend; one rank has no communication to request or receive.
  void RequestComms() {}
  void PreSend() { check(hlb_step(s_, 1), "hlb_step"); }
  void PreReceive() {}
  void PostReceive() {}
  void EndIteration() {}

  // EquilibriumInitialCondition (Code/lb/InitialCondition.hpp:39-52).
  void SetInitialConditions(double density, const std::array<double, 3>& momentum = {0, 0, 0}) {
    check(hlb_set_equilibrium(s_, density, momentum.data()), "hlb_set_equilibrium");
  }
  // SimulationState time step (1-indexed).
  uint64_t GetTimeStep() const { return hlb_get_time(s_); }
  void SetTimeStep(uint64_t t) { check(hlb_set_time(s_, t), "hlb_set_time"); }
  void Step(int64_t n) { check(hlb_step(s_, n), "hlb_step"); }

  site_t TotalFluidSiteCount() const { return n_; }
  const LbmParameters& GetLbmParams() const { return params_; }

  // fOld in FieldData layout, double[n][Q] in the domain's site order.
  std::vector<distribn_t> GetDistributions() {
    std::vector<distribn_t> f((size_t)n_ * LatticeType::NUMVECTORS);
    check(hlb_get_distributions(s_, f.data()), "hlb_get_distributions");
    return f;
  }
  void SetDistributions(const std::vector<distribn_t>& f) {
    if (f.size() != (size_t)n_ * LatticeType::NUMVECTORS)
      throw Exception("distributions have the wrong size");
    check(hlb_set_distributions(s_, f.data()), "hlb_set_distributions");
  }
  // Density [n] and velocity [n][3] (MacroscopicPropertyCache).
  void GetMacroscopic(std::vector<double>& density, std::vector<double>& velocity) {
    density.resize((size_t)n_);
    velocity.resize(3 * (size_t)n_);
    check(hlb_get_macroscopic(s_, density.data(), velocity.data()), "hlb_get_macroscopic");
  }
  hlb_monitor Monitor() {
    hlb_monitor m{};
    check(hlb_reduce(s_, &m), "hlb_reduce");
    return m;
  }
  // Periodic-state monitor: SnapshotVelocity keeps the current velocity
  // field on the device; VelocityChange is max |u - u_kept| / reference_speed.
  void SnapshotVelocity() { check(hlb_snapshot_velocity(s_), "hlb_snapshot_velocity"); }
  double VelocityChange(double reference_speed) {
    double v = 0.0;
    check(hlb_velocity_change(s_, reference_speed, &v), "hlb_velocity_change");
    return v;
  }
  hlb_solver* Handle() { return s_; }

 private:
  LbmParameters params_;
  site_t n_;
  hlb_solver* s_ = nullptr;
};

// Inlet cycles of `period` steps run until the cycle-to-cycle change
// max |u(t) - u(t - T)| / max |u| is below `tolerance` or, with
// stop_on_plateau, a cycle no longer halves it (a reduced-precision form's
// rounding floor).  Appends every cycle's change to `changes`; returns the
// cycles run when it stopped for either reason, 0 after max_cycles.
template <class T>
int CyclesToPeriodic(LBM<T>& lbm, int64_t period, double tolerance, int max_cycles,
                     bool stop_on_plateau, std::vector<double>& changes) {
  lbm.SnapshotVelocity();
  const size_t first = changes.size();
  for (int cycles = 1; cycles <= max_cycles; ++cycles) {
    lbm.Step(period);
    const double umax = lbm.Monitor().velocity_max;
    const double change = lbm.VelocityChange(umax > 1e-30 ? umax : 1e-30);
    changes.push_back(change);
    lbm.SnapshotVelocity();
    const bool plateau = stop_on_plateau && changes.size() > first + 1 &&
                         change > 0.5 * changes[changes.size() - 2];
    if (change < tolerance || plateau) return cycles;
  }
  return 0;
}

// The periodic state of the last solver of `stages` with the transient run
// in the earlier, faster (reduced-precision) ones, all of the same domain
// and iolets.  Each earlier stage runs until its change is below
// stage_tolerance or stops halving; its distributions (exact in binary64)
// and time step then start the next.  The last stage runs until its own
// change is below `tolerance`.  `cycles[k]` and `changes[k]` report stage
// k; returns whether the last stage converged.
template <class T>
bool PeriodicStateStaged(const std::vector<LBM<T>*>& stages, int64_t period, double tolerance,
                         double stage_tolerance, int max_cycles, std::vector<int>& cycles,
                         std::vector<std::vector<double>>& changes) {
  cycles.assign(stages.size(), 0);
  changes.assign(stages.size(), {});
  for (size_t k = 0; k < stages.size(); ++k) {
    const bool last = k + 1 == stages.size();
    if (k) {
      stages[k]->SetDistributions(stages[k - 1]->GetDistributions());
      stages[k]->SetTimeStep(stages[k - 1]->GetTimeStep());
    }
    cycles[k] = CyclesToPeriodic(*stages[k], period, last ? tolerance : stage_tolerance,
                                 max_cycles, !last, changes[k]);
  }
  return cycles.back() > 0;
}

}  // namespace lb
}  // namespace hemelb

#endif
