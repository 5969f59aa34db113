// extern "C" entry points over the parts of the HemeLB reference that compile
// from their own sources without any external dependency:
//   Code/lb/lattices/{Lattice,D3Q15,D3Q19}.h   (header only)
//   Code/lb/kernels/DHumieresD3Q{15,19}MRTBasis.{h,cc}
//   Code/util/Bessel.cc, Code/util/UnitConverter.cc, Code/util/Matrix3D.cc
// TEST INFRASTRUCTURE: built by oracle/build_oracle.py into oracle/_ref/ only
// when the reference tree is present, and used by tests/ to pin the C oracle
// bit for bit.  The streamers, HydroVars and kernels need geometry::Domain
// (boost flat_map) and are not built.
#include <complex>
#include <cstring>

#include "lb/kernels/DHumieresD3Q15MRTBasis.h"
#include "lb/kernels/DHumieresD3Q19MRTBasis.h"
#include "lb/lattices/D3Q15.h"
#include "lb/lattices/D3Q19.h"
#include "util/Bessel.h"
#include "util/UnitConverter.h"

using namespace hemelb;

namespace {
template <class L>
void dmf(const double* f, double* rho, double* m, double* u, double* feq) {
  distribn_t density;
  util::Vector3D<distribn_t> mom;
  util::Vector3D<distribn_t> vel;
  std::array<distribn_t, L::NUMVECTORS> fa, fe;
  std::memcpy(fa.data(), f, sizeof(double) * L::NUMVECTORS);
  L::CalculateDensityMomentumFEq(typename L::const_span(fa), density, mom, vel,
                                 typename L::mut_span(fe));
  *rho = density;
  for (int a = 0; a < 3; ++a) {
    m[a] = mom[a];
    u[a] = vel[a];
  }
  std::memcpy(feq, fe.data(), sizeof(double) * L::NUMVECTORS);
}

template <class L>
void fq(double rho, const double* m, double* feq) {
  std::array<distribn_t, L::NUMVECTORS> fe;
  L::CalculateFeq(rho, util::Vector3D<distribn_t>(m[0], m[1], m[2]),
                  typename L::mut_span(fe));
  std::memcpy(feq, fe.data(), sizeof(double) * L::NUMVECTORS);
}

template <class B>
int basis(double* M, double* diag, double* s, double tau) {
  constexpr int Q = B::Lattice::NUMVECTORS;
  auto S = B::SetUpCollisionMatrix(tau);
  for (unsigned k = 0; k < B::NUMMOMENTS; ++k) {
    for (int i = 0; i < Q; ++i) M[k * Q + i] = B::REDUCED_MOMENT_BASIS[k][i];
    diag[k] = B::BASIS_TIMES_BASIS_TRANSPOSED[k];
    s[k] = S[k];
  }
  return B::NUMMOMENTS;
}

// f_neq = f - f_eq (LBGK.h:140-150, MRT.h:55-66), then
// CalculateWallShearStressMagnitude with LbmParameters' stress parameter
// (LbmParameters.h:34-39) and the float wall normal widened to double.
template <class L>
double wss(double tau, const double* f, const float* nrm) {
  std::array<distribn_t, L::NUMVECTORS> fa, fe, fneq;
  std::memcpy(fa.data(), f, sizeof(double) * L::NUMVECTORS);
  distribn_t density;
  util::Vector3D<distribn_t> mom, vel;
  L::CalculateDensityMomentumFEq(typename L::const_span(fa), density, mom, vel,
                                 typename L::mut_span(fe));
  for (unsigned i = 0; i < L::NUMVECTORS; ++i) fneq[i] = fa[i] - fe[i];
  const double stress_parameter = (1.0 - 1.0 / (2.0 * tau)) / std::sqrt(2.0);
  const util::Vector3D<float> n(nrm[0], nrm[1], nrm[2]);
  distribn_t stress;
  L::CalculateWallShearStressMagnitude(density, typename L::const_span(fneq),
                                       util::Vector3D<double>(n), stress,
                                       stress_parameter);
  return stress;
}

// The other stress outputs of Lattice.h from the same f_neq.
template <class L>
void stress_fields(double tau, const double* f, const float* nrm, double* t6, double* tr,
                   double* tg, double* vm, double* sr) {
  std::array<distribn_t, L::NUMVECTORS> fa, fe, fneq;
  std::memcpy(fa.data(), f, sizeof(double) * L::NUMVECTORS);
  distribn_t density;
  util::Vector3D<distribn_t> mom, vel;
  L::CalculateDensityMomentumFEq(typename L::const_span(fa), density, mom, vel,
                                 typename L::mut_span(fe));
  for (unsigned i = 0; i < L::NUMVECTORS; ++i) fneq[i] = fa[i] - fe[i];
  typename L::const_span span(fneq);
  util::Matrix3D sigma;
  L::CalculateStressTensor(density, tau, span, sigma);
  t6[0] = sigma[0][0]; t6[1] = sigma[0][1]; t6[2] = sigma[0][2];
  t6[3] = sigma[1][1]; t6[4] = sigma[1][2]; t6[5] = sigma[2][2];
  const util::Vector3D<double> n(util::Vector3D<float>(nrm[0], nrm[1], nrm[2]));
  util::Vector3D<LatticeStress> t, g;
  L::CalculateTractionOnAPoint(density, tau, span, n, t);
  L::CalculateTangentialProjectionTraction(density, tau, span, n, g);
  for (int a = 0; a < 3; ++a) {
    tr[a] = t[a];
    tg[a] = g[a];
  }
  const double stress_parameter = (1.0 - 1.0 / (2.0 * tau)) / std::sqrt(2.0);
  L::CalculateVonMisesStress(span, *vm, stress_parameter);
  *sr = L::CalculateShearRate(tau, span, density);
}

template <class L>
void lattice(int* cx, int* cy, int* cz, double* w, int* inv) {
  for (unsigned i = 0; i < L::NUMVECTORS; ++i) {
    cx[i] = L::CX[i];
    cy[i] = L::CY[i];
    cz[i] = L::CZ[i];
    w[i] = L::EQMWEIGHTS[i];
    inv[i] = L::INVERSEDIRECTIONS[i];
  }
}
}  // namespace

extern "C" {
int ref_lattice(int Q, int* cx, int* cy, int* cz, double* w, int* inv) {
  if (Q == 15) lattice<lb::D3Q15>(cx, cy, cz, w, inv);
  else if (Q == 19) lattice<lb::D3Q19>(cx, cy, cz, w, inv);
  else return -1;
  return 0;
}

int ref_density_momentum_feq(int Q, const double* f, double* rho, double* m,
                             double* u, double* feq) {
  if (Q == 15) dmf<lb::D3Q15>(f, rho, m, u, feq);
  else if (Q == 19) dmf<lb::D3Q19>(f, rho, m, u, feq);
  else return -1;
  return 0;
}

int ref_feq(int Q, double rho, const double* m, double* feq) {
  if (Q == 15) fq<lb::D3Q15>(rho, m, feq);
  else if (Q == 19) fq<lb::D3Q19>(rho, m, feq);
  else return -1;
  return 0;
}

double ref_wall_shear_stress(int Q, double tau, const double* f, const float* normal) {
  if (Q == 15) return wss<lb::D3Q15>(tau, f, normal);
  if (Q == 19) return wss<lb::D3Q19>(tau, f, normal);
  return -1.0;
}

int ref_stress_fields(int Q, double tau, const double* f, const float* normal, double* t6,
                      double* traction, double* tangential, double* von_mises,
                      double* shear_rate) {
  if (Q == 15) stress_fields<lb::D3Q15>(tau, f, normal, t6, traction, tangential, von_mises, shear_rate);
  else if (Q == 19) stress_fields<lb::D3Q19>(tau, f, normal, t6, traction, tangential, von_mises, shear_rate);
  else return -1;
  return 0;
}

int ref_mrt_basis(int Q, double* M, double* diag, double* s, double tau) {
  if (Q == 15) return basis<lb::DHumieresD3Q15MRTBasis>(M, diag, s, tau);
  if (Q == 19) return basis<lb::DHumieresD3Q19MRTBasis>(M, diag, s, tau);
  return -1;
}

void ref_bessel_j0(double re, double im, double tolsq, double* out) {
  std::complex<double> r = util::BesselJ0ComplexArgument({re, im}, tolsq);
  out[0] = r.real();
  out[1] = r.imag();
}

// The expression InOutLetWomersleyVelocity.cc:12-13 evaluates with the same
// standard library (that translation unit itself needs SimConfig.h).
void ref_ipow_three_halves(double* out) {
  const std::complex<double> r = pow(std::complex<double>(0, 1), 1.5);
  out[0] = r.real();
  out[1] = r.imag();
}

double ref_pressure_to_lattice(double p, double ref, double dt, double dx,
                               double rho) {
  util::UnitConverter u(dt, dx, {0.0, 0.0, 0.0}, rho, ref);
  return u.ConvertPressureToLatticeUnits(p);
}

double ref_pressure_diff_to_lattice(double dp, double dt, double dx,
                                    double rho) {
  util::UnitConverter u(dt, dx, {0.0, 0.0, 0.0}, rho, 0.0);
  return u.ConvertPressureDifferenceToLatticeUnits(dp);
}

double ref_pressure_gradient_to_lattice(double pg, double dt, double dx,
                                        double rho) {
  util::UnitConverter u(dt, dx, {0.0, 0.0, 0.0}, rho, 0.0);
  return u.ConvertPressureGradientToLatticeUnits(pg);
}
}
