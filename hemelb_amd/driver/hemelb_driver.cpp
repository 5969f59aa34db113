// hemelb_driver: a SimulationMaster-like driver of the GPU timestep written
// against the C++ layer (include/hemelb_amd.hpp), the way a HemeLB build
// drives its LB actor: the lattice, kernel and streamers are compile-time
// Traits, the actor is hemelb::lb::LBM<Traits>, and every step runs the
// IteratedAction phases (Code/SimulationMaster.impl.h:168-220,
// Code/net/phased/StepManager.cc).
//
// Geometry: a generated straight cylinder (--cylinder R L) or a .gmy file
// (--gmy path, iolet planes from --inlet/--outlet are not supported: the
// file's own planes are used when present).  Iolets: Nash zeroth-order
// pressure with a cosine inlet density 1 + A cos(2 pi t / T) and outlets at
// density 1.
//
// Output: one JSON line with the run's monitor; --out writes fOld as raw
// binary64 [n][Q] in the domain's site order.
#include "hemelb_amd.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

using namespace hemelb;

namespace {

struct Options {
  double radius = 6.0;
  int32_t length = 24;
  std::string gmy;
  int lattice = 15;
  std::string kernel = "LBGK", wall = "SBB", storage = "fp64";
  double tau = 0.8, amplitude = 1e-3, period = 100.0;
  int64_t steps = 100, monitor_every = 0;
  std::string out;
  // periodic-state mode: run to the periodic state instead of `steps`,
  // optionally with the transient in faster storage forms (`stages`)
  double to_periodic = 0.0, stage_tol = 1e-2;
  int max_cycles = 12;
  std::string stages;
};

[[noreturn]] void usage() {
  std::fprintf(stderr,
               "usage: hemelb_driver [--cylinder R L | --gmy file] [--lattice 15|19]\n"
               "       [--kernel LBGK|MRT] [--wall SBB|BFL] [--storage fp64|fp32|fp16]\n"
               "       [--tau t] [--amplitude A] [--period T] [--steps n]\n"
               "       [--monitor-every k] [--out file]\n"
               "       [--to-periodic tol [--stages fp16,fp32,fp64] [--stage-tol t]\n"
               "        [--max-cycles n]]\n");
  std::exit(2);
}

Options parse(int argc, char** argv) {
  Options o;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) usage();
      return argv[++i];
    };
    if (a == "--cylinder") {
      o.radius = std::atof(next().c_str());
      o.length = std::atoi(next().c_str());
    } else if (a == "--gmy") {
      o.gmy = next();
    } else if (a == "--lattice") {
      o.lattice = std::atoi(next().c_str());
    } else if (a == "--kernel") {
      o.kernel = next();
    } else if (a == "--wall") {
      o.wall = next();
    } else if (a == "--storage") {
      o.storage = next();
    } else if (a == "--tau") {
      o.tau = std::atof(next().c_str());
    } else if (a == "--amplitude") {
      o.amplitude = std::atof(next().c_str());
    } else if (a == "--period") {
      o.period = std::atof(next().c_str());
    } else if (a == "--steps") {
      o.steps = std::atoll(next().c_str());
    } else if (a == "--monitor-every") {
      o.monitor_every = std::atoll(next().c_str());
    } else if (a == "--out") {
      o.out = next();
    } else if (a == "--to-periodic") {
      o.to_periodic = std::atof(next().c_str());
    } else if (a == "--stages") {
      o.stages = next();
    } else if (a == "--stage-tol") {
      o.stage_tol = std::atof(next().c_str());
    } else if (a == "--max-cycles") {
      o.max_cycles = std::atoi(next().c_str());
    } else {
      usage();
    }
  }
  return o;
}

lb::Storage storage_of(const std::string& s) {
  if (s == "fp64") return lb::Storage::FP64;
  if (s == "fp32") return lb::Storage::FP32;
  if (s == "fp16") return lb::Storage::FP16;
  throw Exception("unknown storage " + s);
}

std::array<double, 3> a3(const double* v) { return {v[0], v[1], v[2]}; }

void write_f(const std::string& path, const std::vector<distribn_t>& f) {
  std::unique_ptr<FILE, int (*)(FILE*)> fp(std::fopen(path.c_str(), "wb"), &std::fclose);
  if (!fp || std::fwrite(f.data(), sizeof(double), f.size(), fp.get()) != f.size())
    throw Exception("cannot write " + path);
}

// From rest to the periodic state (cycle-to-cycle change below
// o.to_periodic), the transient in the earlier storage forms of o.stages
// and the state handed over exactly between them (PeriodicStateStaged).
template <class T>
int run_periodic(const Options& o, const lb::DomainArrays& dom,
                 const std::vector<hlb_iolet>& inlets, const std::vector<hlb_iolet>& outlets) {
  std::vector<std::string> forms;
  const std::string list = o.stages.empty() ? o.storage : o.stages;
  for (size_t b = 0; b <= list.size();) {
    const size_t e = std::min(list.find(',', b), list.size());
    forms.push_back(list.substr(b, e - b));
    b = e + 1;
  }
  std::vector<std::unique_ptr<lb::LBM<T>>> owned;
  std::vector<lb::LBM<T>*> stages;
  for (const auto& form : forms) {
    owned.push_back(std::make_unique<lb::LBM<T>>(lb::LbmParameters(o.tau), dom, inlets,
                                                 outlets, storage_of(form)));
    owned.back()->SetInitialConditions(1.0);
    stages.push_back(owned.back().get());
  }
  const int64_t period = (int64_t)std::llround(o.period);
  std::vector<int> cycles;
  std::vector<std::vector<double>> changes;
  const auto t0 = std::chrono::steady_clock::now();
  const bool converged = lb::PeriodicStateStaged(stages, period, o.to_periodic, o.stage_tol,
                                                 o.max_cycles, cycles, changes);
  lb::LBM<T>& last = *stages.back();
  const hlb_monitor m = last.Monitor();
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (!m.stable) throw Exception("simulation unstable");
  if (!o.out.empty()) write_f(o.out, last.GetDistributions());
  std::string sc;
  for (size_t k = 0; k < forms.size(); ++k)
    sc += (k ? ", " : "") + std::string("[\"") + forms[k] + "\", " + std::to_string(cycles[k]) + "]";
  std::printf(
      "{\"sites\": %lld, \"Q\": %d, \"time_step\": %llu, \"seconds\": %.6f, "
      "\"converged\": %d, \"stage_cycles\": [%s], \"last_change\": %.6g, "
      "\"mass\": %.17g, \"velocity_max\": %.17g, \"stable\": %d}\n",
      (long long)last.TotalFluidSiteCount(), T::Lattice::NUMVECTORS,
      (unsigned long long)last.GetTimeStep(), secs, converged ? 1 : 0, sc.c_str(),
      changes.back().empty() ? 0.0 : changes.back().back(), m.mass, m.velocity_max, m.stable);
  return converged ? 0 : 3;
}

template <class T>
int run(const Options& o, const lb::DomainArrays& dom) {
  std::vector<hlb_iolet> inlets, outlets;
  for (const auto& io : dom.inlets)
    inlets.push_back(lb::InOutLetCosine(a3(io.position), a3(io.normal), 1.0, o.amplitude,
                                        o.period));
  for (const auto& io : dom.outlets)
    outlets.push_back(lb::InOutLetCosine(a3(io.position), a3(io.normal), 1.0, 0.0, o.period));
  if (o.to_periodic > 0.0) return run_periodic<T>(o, dom, inlets, outlets);
  lb::LBM<T> lbm(lb::LbmParameters(o.tau), dom, inlets, outlets, storage_of(o.storage));
  lbm.SetInitialConditions(1.0);
  const auto t0 = std::chrono::steady_clock::now();
  for (int64_t step = 1; step <= o.steps; ++step) {
    // one pass of the step manager over the actor's phases
    lbm.RequestComms();
    lbm.PreSend();
    lbm.PreReceive();
    lbm.PostReceive();
    lbm.EndIteration();
    if (o.monitor_every > 0 && step % o.monitor_every == 0) {
      const hlb_monitor m = lbm.Monitor();
      // an unstable run ends the simulation (SimulationMaster.impl.h:247-255)
      if (!m.stable) throw Exception("simulation unstable at step " + std::to_string(step));
      std::fprintf(stderr, "[driver] step %lld: mass %.12g, max |u| %.6g\n",
                   (long long)step, m.mass, m.velocity_max);
    }
  }
  const hlb_monitor m = lbm.Monitor();
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (!o.out.empty()) write_f(o.out, lbm.GetDistributions());
  std::printf(
      "{\"sites\": %lld, \"Q\": %d, \"steps\": %lld, \"time_step\": %llu, \"seconds\": %.6f, "
      "\"mass\": %.17g, \"density_min\": %.17g, \"density_max\": %.17g, "
      "\"velocity_max\": %.17g, \"stable\": %d}\n",
      (long long)lbm.TotalFluidSiteCount(), T::Lattice::NUMVECTORS, (long long)o.steps,
      (unsigned long long)lbm.GetTimeStep(), secs, m.mass, m.density_min, m.density_max,
      m.velocity_max, m.stable);
  return 0;
}

// Runtime options -> the compile-time Traits of the build.
template <class L, template <class> class K>
int run_wall(const Options& o, const lb::DomainArrays& dom) {
  if (o.wall == "SBB")
    return run<Traits<L, K, lb::Normal, lb::BulkStreamer, lb::SimpleBounceBack>>(o, dom);
  if (o.wall == "BFL")
    return run<Traits<L, K, lb::Normal, lb::BulkStreamer, lb::BouzidiFirdaousLallemand>>(o, dom);
  throw Exception("unknown wall streamer " + o.wall);
}

template <class L>
int run_kernel(const Options& o, const lb::DomainArrays& dom) {
  if (o.kernel == "LBGK") return run_wall<L, lb::LBGK>(o, dom);
  if (o.kernel == "MRT") return run_wall<L, lb::DHumieresMRT>(o, dom);
  throw Exception("unknown kernel " + o.kernel);
}

}  // namespace

int main(int argc, char** argv) {
  const Options o = parse(argc, argv);
  try {
    hlbg_geometry* g = nullptr;
    const int rc = o.gmy.empty() ? hlbg_cylinder(o.radius, o.length, 8, &g)
                                 : hlbg_read_gmy(o.gmy.c_str(), &g);
    if (rc != 0) throw Exception(std::string("geometry: ") + hlbg_last_error());
    std::unique_ptr<hlbg_geometry, void (*)(hlbg_geometry*)> geo(g, &hlbg_free);
    if (o.lattice != 15 && o.lattice != 19) throw Exception("lattice must be 15 or 19");
    const lb::DomainArrays dom = lb::DomainArrays::Project(g, o.lattice);
    return o.lattice == 15 ? run_kernel<lb::D3Q15>(o, dom) : run_kernel<lb::D3Q19>(o, dom);
  } catch (const Exception& e) {
    std::fprintf(stderr, "hemelb_driver: %s\n", e.what());
    return 1;
  }
}
