/*
 * CPU oracle for HemeLB's lattice-Boltzmann timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against; it is loaded by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, and by
 * nothing else.  The product path never links or calls it.
 *
 * It restates, in plain C with IEEE double arithmetic and no FMA contraction,
 * the reference's push-scheme timestep (file:line citations are relative to
 * the HemeLB source tree, Code/...):
 *   - lattice constants and SSE3-order moment/equilibrium arithmetic
 *     (lb/lattices/Lattice.h, D3Q15.h, D3Q19.h);
 *   - LBGK and MRT collisions (lb/kernels/LBGK.h, MRT.h, DHumieres*MRTBasis.*);
 *   - bulk, simple bounce-back, BFL, Nash zeroth-order pressure and Ladd
 *     link streamers and the StreamerTypeFactory link priority
  *     (lb/streamers/StreamerTypeFactory.h and the link headers);
 *   - cosine pressure, parabolic and Womersley velocity iolets and the
 *     complex-argument Bessel J0 series (lb/iolets/InOutLet*.cc, util/Bessel.cc);
 *   - the LBM step order: stream-and-collide over all sites, then PostStep,
 *     then swap (lb/lb.hpp, SimulationMaster.impl.h).
 */
#ifndef HLB_ORACLE_H
#define HLB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LBGK = 0, ORC_MRT = 1 };
enum { ORC_WALL_SBB = 0, ORC_WALL_BFL = 1 };
enum { ORC_IOLET_NASH = 0, ORC_IOLET_LADD = 1 };
enum { ORC_COSINE = 0, ORC_PARABOLIC = 1, ORC_WOMERSLEY = 2 };
/* geometry::SiteData site types (io/formats/geometry.h, geometry/SiteDataBare.cc:68-72) */
enum { ORC_SITE_FLUID = 1, ORC_SITE_INLET = 2, ORC_SITE_OUTLET = 3 };

typedef struct {
  int32_t kind;                 /* ORC_COSINE / ORC_PARABOLIC / ORC_WOMERSLEY */
  int32_t pad_;
  double position[3];           /* lattice units (InOutLet::SetPosition)     */
  double normal[3];             /* normalised at create (InOutLet::SetNormal) */
  /* cosine pressure (InOutLetCosine) */
  double density_mean, density_amp, phase, period;
  int64_t warmup;
  /* velocity (InOutLetParabolicVelocity / InOutLetWomersleyVelocity) */
  double radius, max_speed;
  double pressure_gradient_amplitude, womersley_number;
  double min_sim_density;       /* set at create (LBM::PrepareBoundaryObjects) */
} orc_iolet;

typedef struct orc_sim orc_sim;

/* ---- building blocks (each pinned separately by the tests) ---- */
void orc_lattice(int Q, int* cx, int* cy, int* cz, double* w, int* inv);
void orc_density_momentum(int Q, const double* f, double* rho, double* m);
void orc_feq(int Q, double rho, const double* m, double* feq);
/* LBGK collision: f -> fpc with omega = -1/tau; also returns the moments. */
void orc_collide_lbgk(int Q, double tau, const double* f, double* fpc,
                      double* rho, double* m, double* feq);
/* MRT collision with the d'Humieres relaxation rates for tau (s may be NULL),
 * or with explicit relaxation rates s[nmom]. */
void orc_collide_mrt(int Q, double tau, const double* s, const double* f,
                     double* fpc);
int orc_mrt_basis(int Q, double* M, double* diag, double* s_for_tau, double tau);
/* Wall shear stress magnitude of a site with distributions f (fOld) and
 * wall normal `normal` (Lattice.h:652-689). */
double orc_wall_shear_stress(int Q, double tau, const double* f,
                             const float* normal);
/* Stress outputs of a site (stress tensor upper triangle, traction and its
 * tangential projection for a wall normal, von Mises stress, shear rate);
 * any output may be NULL, normal NULL skips the traction outputs. */
void orc_stress_fields(int Q, double tau, const double* f, const float* normal,
                       double* tensor6, double* traction, double* tangential,
                       double* von_mises, double* shear_rate);
void orc_bessel_j0(double re, double im, double tolsq, double* out);
void orc_ipow_three_halves(double* out);
double orc_cosine_density(const orc_iolet* io, uint64_t t0);
void orc_iolet_velocity(const orc_iolet* io, const double* x, uint64_t t,
                        double* u);
double orc_tau(double dt, double dx, double rho_phys, double eta);
double orc_pressure_to_lattice(double p_mmhg, double ref_mmhg, double dt,
                               double dx, double rho_phys);
double orc_pressure_diff_to_lattice(double dp_mmhg, double dt, double dx,
                                    double rho_phys);
double orc_pressure_gradient_to_lattice(double pg, double dt, double dx,
                                        double rho_phys);

/* ---- whole-domain push-scheme solver ---- */
orc_sim* orc_create(int Q, int kernel, int wall, int inlet_bc, int outlet_bc,
                    double tau, int64_t n_sites, const int32_t* coords,
                    const uint32_t* site_type, const uint32_t* wall_mask,
                    const uint32_t* iolet_mask, const int32_t* iolet_id,
                    const float* cut, const orc_iolet* inlets, int n_inlets,
                    const orc_iolet* outlets, int n_outlets);
void orc_destroy(orc_sim* s);
void orc_set_equilibrium(orc_sim* s, double rho, const double* m);
void orc_set_time(orc_sim* s, uint64_t t);
uint64_t orc_get_time(const orc_sim* s);
/* Runs n timesteps; returns 1 if the StabilityTester would flag the last
 * step unstable (some fNew not > 0), else 0. */
int orc_step(orc_sim* s, int64_t n);
/* Host threads of orc_step (OpenMP over sites; results do not depend on it). */
int orc_threads(void);
void orc_set_threads(int n);
void orc_get_f(const orc_sim* s, double* out);
void orc_set_f(orc_sim* s, const double* in);
void orc_macro(const orc_sim* s, double* rho, double* u);
/* Per-site wall shear stress of fOld; DBL_MAX (NO_VALUE, constants.h:47) at
 * sites without wall links (Common.h:37-50). */
void orc_wall_shear_stress_all(const orc_sim* s, const float* normal,
                               double* out);

#ifdef __cplusplus
}
#endif
#endif
