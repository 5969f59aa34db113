/*
 * CPU oracle for HemeLB's lattice-Boltzmann timestep -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this file's library.  It is the checker, never the thing measured as the
 * product and never a fallback for it.
 *
 * Build: gcc -O2 -ffp-contract=off (see oracle/build_oracle.py).  All
 * arithmetic is IEEE binary64 evaluated in the order the reference's default
 * x86_64 build (-msse3, HEMELB_USE_SSE3 on, no FMA) evaluates it.
 *
 * Citations are to the HemeLB source tree (Code/...).
 */
#include "hlb_oracle.h"

#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846264338327950288 /* constants.h:18 */
#define ORC_CS2 (1.0 / 3.0)                         /* constants.h:41 */
#define ORC_MMHG_TO_PA 133.3223874                  /* constants.h:20 */

/* ------------------------------------------------------------------ */
/* Lattices: lb/lattices/D3Q15.h:15-56, lb/lattices/D3Q19.h:14-49      */
/* ------------------------------------------------------------------ */
static const int Q15_C[15][3] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
    {0, 0, 1},   {0, 0, -1},  {1, 1, 1},  {-1, -1, -1}, {1, 1, -1},
    {-1, -1, 1}, {1, -1, 1},  {-1, 1, -1}, {1, -1, -1}, {-1, 1, 1}};
static const double Q15_W[15] = {
    2.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
    1.0 / 9.0,  1.0 / 9.0,  1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0,
    1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0};
static const int Q19_C[19][3] = {
    {0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
    {0, 0, 1},  {0, 0, -1},  {1, 1, 0},  {-1, -1, 0}, {1, -1, 0},
    {-1, 1, 0}, {1, 0, 1},   {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1},
    {0, 1, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
static const double Q19_W[19] = {
    1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
    1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

typedef struct {
  int Q;
  int c[19][3];
  double cd[3][19]; /* CXD, CYD, CZD (Lattice.h:88-90) */
  double w[19];
  int inv[19];
} lattice_t;

static int lattice_init(int Q, lattice_t* L) {
  const int(*C)[3];
  const double* W;
  if (Q == 15) {
    C = Q15_C;
    W = Q15_W;
  } else if (Q == 19) {
    C = Q19_C;
    W = Q19_W;
  } else {
    return -1;
  }
  L->Q = Q;
  for (int i = 0; i < Q; ++i) {
    for (int a = 0; a < 3; ++a) {
      L->c[i][a] = C[i][a];
      L->cd[a][i] = (double)C[i][a];
    }
    L->w[i] = W[i];
  }
  /* detail::compute_inverses (Lattice.h:49-69) */
  for (int i = 0; i < Q; ++i) L->inv[i] = -1;
  for (int i = 0; i < Q; ++i)
    for (int j = i; j < Q; ++j)
      if (C[i][0] == -C[j][0] && C[i][1] == -C[j][1] && C[i][2] == -C[j][2]) {
        L->inv[i] = j;
        L->inv[j] = i;
      }
  return 0;
}

void orc_lattice(int Q, int* cx, int* cy, int* cz, double* w, int* inv) {
  lattice_t L;
  if (lattice_init(Q, &L)) return;
  for (int i = 0; i < Q; ++i) {
    cx[i] = L.c[i][0];
    cy[i] = L.c[i][1];
    cz[i] = L.c[i][2];
    w[i] = L.w[i];
    inv[i] = L.inv[i];
  }
}

/* Lattice::CalculateDensityAndMomentum, SSE3 variant (Lattice.h:119-178).
 * Two accumulator lanes: lane 0 starts at 0 and takes the even directions,
 * lane 1 starts with the last (odd-count) direction and takes the odd ones;
 * the result is lane0 + lane1 (_mm_hadd_pd). */
static void dens_mom(const lattice_t* L, const double* f, double* rho,
                     double* m) {
  const int Q = L->Q;
  const int n2 = (Q >> 1) << 1;
  double r0 = 0.0, r1 = 0.0, x0 = 0.0, x1 = 0.0, y0 = 0.0, y1 = 0.0,
         z0 = 0.0, z1 = 0.0;
  if (Q != n2) {
    r1 = f[Q - 1];
    x1 = L->cd[0][Q - 1] * f[Q - 1];
    y1 = L->cd[1][Q - 1] * f[Q - 1];
    z1 = L->cd[2][Q - 1] * f[Q - 1];
  }
  for (int d = 0; d < n2; d += 2) {
    r0 = r0 + f[d];
    r1 = r1 + f[d + 1];
    x0 = x0 + L->cd[0][d] * f[d];
    x1 = x1 + L->cd[0][d + 1] * f[d + 1];
    y0 = y0 + L->cd[1][d] * f[d];
    y1 = y1 + L->cd[1][d + 1] * f[d + 1];
    z0 = z0 + L->cd[2][d] * f[d];
    z1 = z1 + L->cd[2][d + 1] * f[d + 1];
  }
  *rho = r0 + r1;
  m[0] = x0 + x1;
  m[1] = y0 + y1;
  m[2] = z0 + z1;
}

/* Lattice::CalculateFeq, SSE3 variant, COMPRESSIBLE (Lattice.h:236-321).
 * The paired loop and the scalar tail element evaluate the same expression
 * tree: w_i * ((t1 + (4.5/rho)*(e*e)) + 3*e), e = (cx*mx + cy*my) + cz*mz. */
static void feq_all(const lattice_t* L, double rho, const double* m,
                    double* feq) {
  const double mx = m[0], my = m[1], mz = m[2];
  const double three_halves = (3. / 2.) * ((mx * mx + my * my) + mz * mz);
  const double t1 = rho - three_halves / rho;
  const double rho_1 = 1. / rho;
  double nine = (9. / 2.);
  nine *= rho_1;
  for (int i = 0; i < L->Q; ++i) {
    const double e =
        (L->cd[0][i] * mx + L->cd[1][i] * my) + L->cd[2][i] * mz;
    const double t2 = nine * (e * e);
    const double t3 = 3. * e;
    feq[i] = L->w[i] * ((t1 + t2) + t3);
  }
}

void orc_density_momentum(int Q, const double* f, double* rho, double* m) {
  lattice_t L;
  if (lattice_init(Q, &L)) return;
  dens_mom(&L, f, rho, m);
}

void orc_feq(int Q, double rho, const double* m, double* feq) {
  lattice_t L;
  if (lattice_init(Q, &L)) return;
  feq_all(&L, rho, m, feq);
}

/* Wall shear stress magnitude of one site (Lattice.h:652-689), from
 * f_neq = f - f_eq (LBGK.h:140-150; MRT.h:55-66) and the stress parameter
 * (1 - 1/(2 tau))/sqrt(2) (LbmParameters.h:34-39).  The second moment
 * Pi_ij = sum_l f_neq_l c_li c_lj is accumulated over l in order
 * (Lattice.h:700-727); the traction t_i = sum_j (Pi_ij n_j) temp with
 * temp = stress_parameter * -sqrt(2); result sqrt(|t|^2 - (t.n)^2). */
double orc_wall_shear_stress(int Q, double tau, const double* f,
                             const float* normal) {
  lattice_t L;
  if (lattice_init(Q, &L)) return -1.0;
  double rho, m[3], feq[19], fneq[19];
  dens_mom(&L, f, &rho, m);
  feq_all(&L, rho, m, feq);
  for (int i = 0; i < Q; ++i) fneq[i] = f[i] - feq[i];
  const double stress_parameter = (1.0 - 1.0 / (2.0 * tau)) / sqrt(2.0);
  const double temp = stress_parameter * (-sqrt(2.0));
  double pi[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b <= a; ++b) {
      double acc = 0.0;
      for (int l = 0; l < Q; ++l)
        acc += fneq[l] * L.cd[a][l] * L.cd[b][l];
      pi[a][b] = acc;
      pi[b][a] = acc;
    }
  const double n[3] = {(double)normal[0], (double)normal[1], (double)normal[2]};
  double t[3] = {0.0, 0.0, 0.0}, sq = 0.0, nn = 0.0;
  for (int a = 0; a < 3; ++a) {
    for (int b = 0; b < 3; ++b) t[a] += pi[a][b] * n[b] * temp;
    sq += t[a] * t[a];
    nn += t[a] * n[a];
  }
  return sqrt(sq - nn * nn);
}

/* Stress outputs of one site from f (fOld), following Lattice.h:
 * stress tensor (CalculateStressTensor, 630-645: Pi(f_neq) * (1 - 1/(2 tau))
 * plus (rho - 1) Cs2 on the diagonal; upper triangle xx, xy, xz, yy, yz, zz),
 * traction sigma n and its tangential projection t - n (t . n)
 * (CalculateTractionOnAPoint / CalculateTangentialProjectionTraction,
 * 560-602; computed when normal != NULL), von Mises stress (508-552) and
 * shear rate (728-753 with CalculateStrainRateTensorComponent, 890-907). */
void orc_stress_fields(int Q, double tau, const double* f, const float* normal,
                       double* tensor6, double* traction, double* tangential,
                       double* von_mises, double* shear_rate) {
  lattice_t L;
  if (lattice_init(Q, &L)) return;
  double rho, m[3], feq[19], fneq[19];
  dens_mom(&L, f, &rho, m);
  feq_all(&L, rho, m, feq);
  for (int i = 0; i < Q; ++i) fneq[i] = f[i] - feq[i];
  double pi[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b <= a; ++b) {
      double acc = 0.0;
      for (int l = 0; l < Q; ++l) acc += fneq[l] * L.cd[a][l] * L.cd[b][l];
      pi[a][b] = acc;
      pi[b][a] = acc;
    }
  double sig[3][3];
  const double fac = 1 - 1 / (2 * tau);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) sig[a][b] = pi[a][b] * fac;
  const double pressure = (rho - 1) * (1.0 / 3.0);
  for (int a = 0; a < 3; ++a) sig[a][a] += pressure;
  if (tensor6) {
    tensor6[0] = sig[0][0]; tensor6[1] = sig[0][1]; tensor6[2] = sig[0][2];
    tensor6[3] = sig[1][1]; tensor6[4] = sig[1][2]; tensor6[5] = sig[2][2];
  }
  if (normal) {
    const double n[3] = {(double)normal[0], (double)normal[1], (double)normal[2]};
    double t[3];
    for (int a = 0; a < 3; ++a) {
      t[a] = 0.0;
      for (int b = 0; b < 3; ++b) t[a] += sig[a][b] * n[b];
    }
    if (traction)
      for (int a = 0; a < 3; ++a) traction[a] = t[a];
    double tn = 0.0;
    for (int a = 0; a < 3; ++a) tn = tn + t[a] * n[a];
    if (tangential)
      for (int a = 0; a < 3; ++a) tangential[a] = t[a] - n[a] * tn;
  }
  if (von_mises) {
    const double sp = (1.0 - 1.0 / (2.0 * tau)) / sqrt(2.0);
    double xx_yy = 0.0, yy_zz = 0.0, xx_zz = 0.0, xy = 0.0, xz = 0.0, yz = 0.0;
    for (int l = 0; l < Q; ++l) {
      const int cx = L.c[l][0], cy = L.c[l][1], cz = L.c[l][2];
      xx_yy += fneq[l] * (double)(cx * cx - cy * cy);
      yy_zz += fneq[l] * (double)(cy * cy - cz * cz);
      xx_zz += fneq[l] * (double)(cx * cx - cz * cz);
      xy += fneq[l] * (double)cx * (double)cy;
      xz += fneq[l] * (double)cx * (double)cz;
      yz += fneq[l] * (double)cy * (double)cz;
    }
    const double a = xx_yy * xx_yy + yy_zz * yy_zz + xx_zz * xx_zz;
    const double b = xy * xy + xz * xz + yz * yz;
    *von_mises = sp * sqrt(a + 6.0 * b);
  }
  if (shear_rate) {
    const double k = -1.0 / (2.0 * tau * rho * (1.0 / 3.0));
    double sr = 0.0;
    for (int r = 0; r < 3; ++r)
      for (int c = r; c < 3; ++c) {
        double e = 0.0;
        for (int l = 0; l < Q; ++l) e += fneq[l] * (double)(L.c[l][r] * L.c[l][c]);
        e *= k;
        sr += (r == c) ? e * e : 2 * e * e;
      }
    *shear_rate = sqrt(2 * sr);
  }
}

/* ------------------------------------------------------------------ */
/* Collision kernels                                                   */
/* ------------------------------------------------------------------ */
typedef struct {
  double rho, m[3], u[3];
  double feq[19], fneq[19], fpc[19];
} hydro_t; /* lb/HydroVars.h:29-112 */

/* LBGK::CalculateDensityMomentumFeq + Collide (lb/kernels/LBGK.h:140-174) */
static void lbgk(const lattice_t* L, double omega, const double* f,
                 hydro_t* h) {
  dens_mom(L, f, &h->rho, h->m);
  /* Lattice::CalculateDensityMomentumFEq (Lattice.h:471-483): u = m / rho */
  for (int a = 0; a < 3; ++a) h->u[a] = h->m[a] / h->rho;
  feq_all(L, h->rho, h->m, h->feq);
  for (int i = 0; i < L->Q; ++i) h->fneq[i] = f[i] - h->feq[i];
  for (int i = 0; i < L->Q; ++i) h->fpc[i] = f[i] + h->fneq[i] * omega;
}

/* d'Humieres reduced moment bases (DHumieresD3Q15MRTBasis.h:41-53,
 * DHumieresD3Q19MRTBasis.h:38-53). */
static const int MRT15_M[11][15] = {
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
static const int MRT19_M[15][19] = {
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

typedef struct {
  int nmom;
  double M[15][19];   /* REDUCED_MOMENT_BASIS as double */
  double diag[15];    /* BASIS_TIMES_BASIS_TRANSPOSED (basis_helpers.h:14-26) */
  double N[15][19];   /* normalisedReducedMomentBasis (MRT.h:139-152) */
  double s[15];       /* collisionMatrixDiagonals */
} mrt_t;

static void mrt_rates(int Q, double tau, double* s) {
  if (Q == 15) { /* DHumieresD3Q15MRTBasis.cc:10-26 */
    const double v[11] = {1.6, 1.2, 1.6, 1.6, 1.6, 1.0 / tau,
                          1.0 / tau, 1.0 / tau, 1.0 / tau, 1.0 / tau, 1.2};
    memcpy(s, v, sizeof v);
  } else { /* DHumieresD3Q19MRTBasis.cc:11-36 */
    const double v[15] = {1.19, 1.4, 1.2, 1.2, 1.2, 1.0 / tau, 1.4, 1.0 / tau,
                          1.4, 1.0 / tau, 1.0 / tau, 1.0 / tau, 1.98, 1.98,
                          1.98};
    memcpy(s, v, sizeof v);
  }
}

static int mrt_init(int Q, double tau, const double* s_opt, mrt_t* B) {
  if (Q != 15 && Q != 19) return -1;
  B->nmom = (Q == 15) ? 11 : 15;
  for (int k = 0; k < B->nmom; ++k) {
    double acc = 0.0; /* std::inner_product(row, row, T{0}) */
    for (int i = 0; i < Q; ++i) {
      B->M[k][i] = (double)((Q == 15) ? MRT15_M[k][i] : MRT19_M[k][i]);
      acc = acc + B->M[k][i] * B->M[k][i];
    }
    B->diag[k] = acc;
  }
  for (int i = 0; i < Q; ++i)
    for (int k = 0; k < B->nmom; ++k) B->N[k][i] = B->M[k][i] / B->diag[k];
  if (s_opt)
    memcpy(B->s, s_opt, sizeof(double) * B->nmom);
  else
    mrt_rates(Q, tau, B->s);
  return 0;
}

int orc_mrt_basis(int Q, double* M, double* diag, double* s_for_tau,
                  double tau) {
  mrt_t B;
  if (mrt_init(Q, tau, NULL, &B)) return -1;
  for (int k = 0; k < B.nmom; ++k) {
    for (int i = 0; i < Q; ++i) M[k * Q + i] = B.M[k][i];
    diag[k] = B.diag[k];
    s_for_tau[k] = B.s[k];
  }
  return B.nmom;
}

/* MRT::CalculateDensityMomentumFeq + Collide (lb/kernels/MRT.h:56-105,
 * ProjectVelsIntoMomentSpace :123-135). */
static void mrt(const lattice_t* L, const mrt_t* B, const double* f,
                hydro_t* h) {
  const int Q = L->Q;
  double mneq[15];
  dens_mom(L, f, &h->rho, h->m);
  for (int a = 0; a < 3; ++a) h->u[a] = h->m[a] / h->rho;
  feq_all(L, h->rho, h->m, h->feq);
  for (int i = 0; i < Q; ++i) h->fneq[i] = f[i] - h->feq[i];
  for (int k = 0; k < B->nmom; ++k) {
    mneq[k] = 0.;
    for (int i = 0; i < Q; ++i) mneq[k] += B->M[k][i] * h->fneq[i];
  }
  for (int d = 0; d < Q; ++d) {
    double coll = 0.;
    for (int k = 0; k < B->nmom; ++k) coll += B->s[k] * B->N[k][d] * mneq[k];
    h->fpc[d] = f[d] - coll;
  }
}

void orc_collide_lbgk(int Q, double tau, const double* f, double* fpc,
                      double* rho, double* m, double* feq) {
  lattice_t L;
  hydro_t h;
  if (lattice_init(Q, &L)) return;
  const double omega = -1.0 / tau; /* LbmParameters.h:158-163 */
  lbgk(&L, omega, f, &h);
  memcpy(fpc, h.fpc, sizeof(double) * Q);
  if (rho) *rho = h.rho;
  if (m) memcpy(m, h.m, sizeof h.m);
  if (feq) memcpy(feq, h.feq, sizeof(double) * Q);
}

void orc_collide_mrt(int Q, double tau, const double* s, const double* f,
                     double* fpc) {
  lattice_t L;
  mrt_t B;
  hydro_t h;
  if (lattice_init(Q, &L) || mrt_init(Q, tau, s, &B)) return;
  mrt(&L, &B, f, &h);
  memcpy(fpc, h.fpc, sizeof(double) * Q);
}

/* ------------------------------------------------------------------ */
/* Units: LbmParameters.h:158-163, util/UnitConverter.cc:14-48         */
/* ------------------------------------------------------------------ */
double orc_tau(double dt, double dx, double rho_phys, double eta) {
  return 0.5 + (dt * eta / rho_phys) / (ORC_CS2 * dx * dx);
}

static double lattice_pressure_unit(double dt, double dx, double rho_phys) {
  const double mass = rho_phys * dx * dx * dx;
  return mass / (dx * dt * dt);
}

double orc_pressure_diff_to_lattice(double dp_mmhg, double dt, double dx,
                                    double rho_phys) {
  return dp_mmhg * ORC_MMHG_TO_PA / lattice_pressure_unit(dt, dx, rho_phys);
}

double orc_pressure_to_lattice(double p_mmhg, double ref_mmhg, double dt,
                               double dx, double rho_phys) {
  return ORC_CS2 +
         orc_pressure_diff_to_lattice(p_mmhg - ref_mmhg, dt, dx, rho_phys);
}

double orc_pressure_gradient_to_lattice(double pg, double dt, double dx,
                                        double rho_phys) {
  return pg * ORC_MMHG_TO_PA * dx / lattice_pressure_unit(dt, dx, rho_phys);
}

/* ------------------------------------------------------------------ */
/* Iolets                                                              */
/* ------------------------------------------------------------------ */
/* util::BesselJ0ComplexArgument (util/Bessel.cc:14-35) */
static double complex bessel_j0(double complex z, double tolsq) {
  double complex sum = 1.0;
  double fact = 1;
  double complex zpow = 1.0;
  double complex term = 1.0;
  for (unsigned i = 1;
       creal(term) * creal(term) + cimag(term) * cimag(term) > tolsq; ++i) {
    fact *= i;
    zpow *= (-0.25 * z) * z;
    term = zpow / (fact * fact);
    sum += term;
  }
  return sum;
}

void orc_bessel_j0(double re, double im, double tolsq, double* out) {
  double complex r = bessel_j0(CMPLX(re, im), tolsq);
  out[0] = creal(r);
  out[1] = cimag(r);
}

/* InOutLetWomersleyVelocity::iPowThreeHalves = std::pow(i, 1.5)
 * (InOutLetWomersleyVelocity.cc:12-13).  libstdc++ evaluates
 * pow(complex, real) for a non-positive-real base as
 * polar(exp(y*log|x|), y*arg x) with log from clog. */
static double complex ipow32(void) {
  double complex t = clog(CMPLX(0.0, 1.0));
  double rho = exp(1.5 * creal(t));
  double th = 1.5 * cimag(t);
  return CMPLX(rho * cos(th), rho * sin(th));
}

void orc_ipow_three_halves(double* out) {
  double complex r = ipow32();
  out[0] = creal(r);
  out[1] = cimag(r);
}

/* util::Dot: std::inner_product from RES{0} (util/Vector3D.h:583-588) */
static double dot3(const double* a, const double* b) {
  double acc = 0.0;
  acc = acc + a[0] * b[0];
  acc = acc + a[1] * b[1];
  acc = acc + a[2] * b[2];
  return acc;
}

/* InOutLetCosine::GetDensity (lb/iolets/InOutLetCosine.cc:26-43) */
double orc_cosine_density(const orc_iolet* io, uint64_t t0) {
  const double w = 2.0 * ORC_PI / io->period;
  const double target =
      io->density_mean + io->density_amp * cos(w * (double)t0 + io->phase);
  if ((int64_t)t0 >= io->warmup) return target;
  const double f = ((double)t0) / ((double)io->warmup);
  return f * target + (1. - f) * io->min_sim_density;
}

/* InOutLetParabolicVelocity::GetVelocity (InOutLetParabolicVelocity.cc:23-43)
 * and InOutLetWomersleyVelocity::GetVelocity (InOutLetWomersleyVelocity.cc:21-38) */
void orc_iolet_velocity(const orc_iolet* io, const double* x, uint64_t t,
                        double* u) {
  double displ[3] = {x[0] - io->position[0], x[1] - io->position[1],
                     x[2] - io->position[2]};
  const double z = dot3(displ, io->normal);
  if (io->kind == ORC_PARABOLIC) {
    const double rsq = (dot3(displ, displ) - z * z) / (io->radius * io->radius);
    double mx = io->max_speed;
    if ((int64_t)t < io->warmup) mx *= (double)t / (double)io->warmup;
    const double s = mx * (1. - rsq);
    for (int a = 0; a < 3; ++a) u[a] = io->normal[a] * s;
    return;
  }
  if (io->kind == ORC_WOMERSLEY) {
    const double r = sqrt(dot3(displ, displ) - z * z);
    const double omega = 2.0 * ORC_PI / io->period;
    const double density = 1.0;
    const double complex ip = ipow32();
    const double complex numer =
        bessel_j0(ip * io->womersley_number * r / io->radius, 1e-12);
    const double complex denom = bessel_j0(ip * io->womersley_number, 1e-12);
    const double complex a =
        io->pressure_gradient_amplitude / (density * omega) *
        (1.0 - numer / denom);
    const double complex e = cexp(CMPLX(0.0 * omega * (double)t,
                                        1.0 * omega * (double)t));
    const double complex p = a * e;
    const double vm = creal(p);
    for (int a2 = 0; a2 < 3; ++a2) u[a2] = io->normal[a2] * -vm;
    return;
  }
  u[0] = u[1] = u[2] = 0.0;
}

/* InOutLet density for the Nash streamer: cosine iolets use GetDensity,
 * velocity iolets return 1 (InOutLetVelocity.cc:24-35). */
static double iolet_density(const orc_iolet* io, uint64_t t0) {
  if (io->kind == ORC_COSINE) return orc_cosine_density(io, t0);
  return 1.0;
}

static double iolet_density_min(const orc_iolet* io) {
  if (io->kind == ORC_COSINE) return io->density_mean - io->density_amp;
  return 1.0;
}

/* ------------------------------------------------------------------ */
/* Whole-domain solver (push scheme)                                   */
/* ------------------------------------------------------------------ */
struct orc_sim {
  lattice_t L;
  mrt_t B;
  int kernel, wall, bc[2]; /* bc[0] inlet, bc[1] outlet */
  double tau, omega;
  int64_t n;
  int32_t* coords;
  uint32_t *type, *wmask, *imask;
  int32_t* iolet_id;
  float* cut;
  int64_t* nbr; /* push index: j*Q+d, or the rubbish slot n*Q (Domain.cc:454-471) */
  double *f_old, *f_new;
  orc_iolet* iolets[2];
  int n_iolets[2];
  uint64_t t; /* SimulationState::timeStep, starts at 1 (SimulationState.cc:15-19) */
};

static int has_bit(uint32_t mask, int d) {
  /* SiteData::HasWall / HasIolet (SiteDataBare.cc:147-163) */
  return d ? ((mask >> (d - 1)) & 1u) != 0 : 0;
}

typedef struct {
  uint64_t key;
  int64_t idx;
} site_key_t;

/* Coordinates are offset by 2^20 so that neighbours of sites at 0 pack too. */
static uint64_t pack_coords(int64_t x, int64_t y, int64_t z) {
  return ((uint64_t)(x + (1 << 20)) << 42) | ((uint64_t)(y + (1 << 20)) << 21) |
         (uint64_t)(z + (1 << 20));
}

static int cmp_key(const void* a, const void* b) {
  const uint64_t ka = ((const site_key_t*)a)->key, kb = ((const site_key_t*)b)->key;
  return ka < kb ? -1 : (ka > kb ? 1 : 0);
}

static int64_t find_key(const site_key_t* keys, int64_t n, uint64_t k) {
  int64_t lo = 0, hi = n - 1;
  while (lo <= hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (keys[mid].key == k) return keys[mid].idx;
    if (keys[mid].key < k) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

orc_sim* orc_create(int Q, int kernel, int wall, int inlet_bc, int outlet_bc,
                    double tau, int64_t n, const int32_t* coords,
                    const uint32_t* site_type, const uint32_t* wall_mask,
                    const uint32_t* iolet_mask, const int32_t* iolet_id,
                    const float* cut, const orc_iolet* inlets, int n_inlets,
                    const orc_iolet* outlets, int n_outlets) {
  orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
  if (!s || lattice_init(Q, &s->L)) {
    free(s);
    return NULL;
  }
  s->kernel = kernel;
  s->wall = wall;
  s->bc[0] = inlet_bc;
  s->bc[1] = outlet_bc;
  s->tau = tau;
  s->omega = -1.0 / tau;
  if (kernel == ORC_MRT) mrt_init(Q, tau, NULL, &s->B);
  s->n = n;
  s->t = 1;
  s->coords = (int32_t*)malloc(sizeof(int32_t) * 3 * (n ? n : 1));
  s->type = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  s->wmask = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  s->imask = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  s->iolet_id = (int32_t*)malloc(sizeof(int32_t) * (n ? n : 1));
  s->cut = (float*)malloc(sizeof(float) * (Q - 1) * (n ? n : 1));
  s->nbr = (int64_t*)malloc(sizeof(int64_t) * Q * (n ? n : 1));
  s->f_old = (double*)calloc(Q * n + 1, sizeof(double));
  s->f_new = (double*)calloc(Q * n + 1, sizeof(double));
  memcpy(s->coords, coords, sizeof(int32_t) * 3 * n);
  memcpy(s->type, site_type, sizeof(uint32_t) * n);
  memcpy(s->wmask, wall_mask, sizeof(uint32_t) * n);
  memcpy(s->imask, iolet_mask, sizeof(uint32_t) * n);
  memcpy(s->iolet_id, iolet_id, sizeof(int32_t) * n);
  if (cut)
    memcpy(s->cut, cut, sizeof(float) * (Q - 1) * n);
  else
    for (int64_t i = 0; i < (Q - 1) * n; ++i) s->cut[i] = -1.0f;

  const orc_iolet* src[2] = {inlets, outlets};
  const int cnt[2] = {n_inlets, n_outlets};
  double min_density = 1.7976931348623157e308; /* lb.hpp:127-152 */
  for (int g = 0; g < 2; ++g) {
    s->n_iolets[g] = cnt[g];
    s->iolets[g] = (orc_iolet*)calloc(cnt[g] ? cnt[g] : 1, sizeof(orc_iolet));
    for (int k = 0; k < cnt[g]; ++k) {
      orc_iolet* io = &s->iolets[g][k];
      *io = src[g][k];
      /* InOutLet::SetNormal stores normal.GetNormalised() (InOutLet.h:160-163) */
      const double mag = sqrt(dot3(src[g][k].normal, src[g][k].normal));
      for (int a = 0; a < 3; ++a) io->normal[a] = src[g][k].normal[a] / mag;
      double dm = iolet_density_min(io);
      min_density = dm < min_density ? dm : min_density;
    }
  }
  for (int g = 0; g < 2; ++g)
    for (int k = 0; k < cnt[g]; ++k) s->iolets[g][k].min_sim_density = min_density;

  /* Domain::InitialiseNeighbourLookup (Domain.cc:430-500): the push index
   * of (site, d) is the slot (neighbour, d), or the rubbish slot.  Sites are
   * found by binary search over their sorted packed coordinates. */
  site_key_t* keys = (site_key_t*)malloc(sizeof(site_key_t) * (n ? n : 1));
  for (int64_t i = 0; i < n; ++i) {
    keys[i].key = pack_coords(coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]);
    keys[i].idx = i;
  }
  qsort(keys, (size_t)n, sizeof(site_key_t), cmp_key);
  for (int64_t i = 0; i < n; ++i) {
    s->nbr[i * Q] = i * Q;
    for (int d = 1; d < Q; ++d) {
      const uint64_t k = pack_coords(coords[3 * i] + s->L.c[d][0],
                                     coords[3 * i + 1] + s->L.c[d][1],
                                     coords[3 * i + 2] + s->L.c[d][2]);
      const int64_t j = find_key(keys, n, k);
      s->nbr[i * Q + d] = (j >= 0) ? j * Q + d : n * Q;
    }
  }
  free(keys);
  return s;
}

void orc_destroy(orc_sim* s) {
  if (!s) return;
  free(s->coords);
  free(s->type);
  free(s->wmask);
  free(s->imask);
  free(s->iolet_id);
  free(s->cut);
  free(s->nbr);
  free(s->f_old);
  free(s->f_new);
  free(s->iolets[0]);
  free(s->iolets[1]);
  free(s);
}

/* EquilibriumInitialCondition::SetFs (lb/InitialCondition.hpp:39-52) */
void orc_set_equilibrium(orc_sim* s, double rho, const double* m) {
  double feq[19];
  feq_all(&s->L, rho, m, feq);
  const int Q = s->L.Q;
  for (int64_t i = 0; i < s->n; ++i)
    for (int d = 0; d < Q; ++d) s->f_old[i * Q + d] = s->f_new[i * Q + d] = feq[d];
}

void orc_set_time(orc_sim* s, uint64_t t) { s->t = t; }
uint64_t orc_get_time(const orc_sim* s) { return s->t; }

static int group_of(uint32_t type) { return type == ORC_SITE_OUTLET ? 1 : 0; }

/* One site of StreamerTypeFactory::StreamAndCollide (StreamerTypeFactory.h:48-86)
 * or BulkStreamer::StreamAndCollide (BulkStreamer.h:67-91). */
static void stream_and_collide_site(orc_sim* s, int64_t i) {
  const lattice_t* L = &s->L;
  const int Q = L->Q;
  const double* f = &s->f_old[i * Q];
  hydro_t h;
  if (s->kernel == ORC_MRT)
    mrt(L, &s->B, f, &h);
  else
    lbgk(L, s->omega, f, &h);

  const uint32_t type = s->type[i];
  const uint32_t wm = s->wmask[i], im = s->imask[i];
  /* Collision type -> which delegates exist (Domain.cc:187-228, lb.hpp:74-114) */
  const int can_wall = wm != 0;
  const int can_iolet = (type == ORC_SITE_INLET || type == ORC_SITE_OUTLET);
  const int g = group_of(type);

  for (int d = 0; d < Q; ++d) {
    const int inv = L->inv[d];
    if (can_iolet && has_bit(im, d)) {
      const orc_iolet* io = &s->iolets[g][s->iolet_id[i]];
      if (s->bc[g] == ORC_IOLET_NASH) {
        /* NashZerothOrderPressureLink::StreamLink (NashZerothOrderPressure.h:27-60) */
        const double ghost_rho = iolet_density(io, s->t - 1);
        const double nf[3] = {(double)(float)io->normal[0],
                              (double)(float)io->normal[1],
                              (double)(float)io->normal[2]};
        const double comp = dot3(h.m, nf) / h.rho;
        double gm[3], gfeq[19];
        for (int a = 0; a < 3; ++a) gm[a] = nf[a] * comp * ghost_rho;
        feq_all(L, ghost_rho, gm, gfeq);
        s->f_new[i * Q + inv] = gfeq[inv];
      } else {
        /* LaddIoletLink::StreamLink (LaddIolet.h:29-66) */
        double half[3];
        for (int a = 0; a < 3; ++a)
          half[a] = (double)s->coords[3 * i + a] + (double)L->c[d][a] * 0.5;
        double wmom[3];
        orc_iolet_velocity(io, half, (uint64_t)(unsigned int)s->t, wmom);
        for (int a = 0; a < 3; ++a) wmom[a] *= h.rho;
        const double cdv[3] = {(double)L->c[d][0], (double)L->c[d][1],
                               (double)L->c[d][2]};
        const double corr = 2. * L->w[d] * dot3(wmom, cdv) / ORC_CS2;
        s->f_new[i * Q + inv] = h.fpc[d] - corr;
      }
    } else if (can_wall && has_bit(wm, d)) {
      if (s->wall == ORC_WALL_SBB) {
        /* BounceBackLink::StreamLink (SimpleBounceBack.h:23-42) */
        s->f_new[i * Q + inv] = h.fpc[d];
      } else {
        /* BouzidiFirdaousLallemandLink::StreamLink (BouzidiFirdaousLallemand.h:41-66) */
        const double q = (double)s->cut[i * (Q - 1) + d - 1];
        if (has_bit(wm, inv) || q < 0.5)
          s->f_new[i * Q + inv] = h.fpc[d];
        else
          s->f_new[i * Q + inv] = (h.fpc[d] + (2.0 * q - 1) * h.fpc[inv]) / (2.0 * q);
      }
    } else {
      /* BulkLink::StreamLink (BulkStreamer.h:31-39) */
      s->f_new[s->nbr[i * Q + d]] = h.fpc[d];
    }
  }
}

/* StreamerTypeFactory::PostStep (StreamerTypeFactory.h:88-108) with
 * BouzidiFirdaousLallemandLink::PostStepLink (BouzidiFirdaousLallemand.h:68-91) */
static void post_step_site(orc_sim* s, int64_t i) {
  if (s->wall != ORC_WALL_BFL) return;
  const lattice_t* L = &s->L;
  const int Q = L->Q;
  const uint32_t wm = s->wmask[i];
  if (!wm) return;
  double* fn = &s->f_new[i * Q];
  for (int d = 0; d < Q; ++d) {
    if (!has_bit(wm, d)) continue;
    const int inv = L->inv[d];
    const double q = (double)s->cut[i * (Q - 1) + d - 1];
    if (!has_bit(wm, inv) && q < 0.5)
      fn[inv] = 2.0 * q * fn[inv] + (1.0 - 2.0 * q) * fn[d];
  }
}

/* Host threads the oracle's step uses (OpenMP; 1 when built without it). */
int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* SimulationMaster::DoTimeStep (SimulationMaster.impl.h:168-220) reduced to
 * the LB actor: PreSend/PreReceive stream-and-collide every site, PostReceive
 * runs PostStep over every site, then SwapOldAndNew and Increment.  The
 * StabilityTester scan (StabilityTester.h:91-118) reads fNew before the swap. */
int orc_step(orc_sim* s, int64_t nsteps) {
  int unstable = 0;
  const int Q = s->L.Q;
  const int64_t n = s->n;
  for (int64_t k = 0; k < nsteps; ++k) {
    /* Sites are independent within each phase: every fNew slot has exactly
     * one writer in the push scheme and PostStep touches only the site's own
     * slots, so the host threads split the sites without changing any
     * site's arithmetic. */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) stream_and_collide_site(s, i);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) post_step_site(s, i);
    unstable = 0;
#pragma omp parallel for schedule(static) reduction(| : unstable)
    for (int64_t i = 0; i < n * Q; ++i)
      if (!(s->f_new[i] > 0.0)) unstable |= 1;
    double* tmp = s->f_old;
    s->f_old = s->f_new;
    s->f_new = tmp;
    s->t += 1;
  }
  return unstable;
}

void orc_get_f(const orc_sim* s, double* out) {
  memcpy(out, s->f_old, sizeof(double) * s->L.Q * s->n);
}

void orc_set_f(orc_sim* s, const double* in) {
  memcpy(s->f_old, in, sizeof(double) * s->L.Q * s->n);
}

/* Density and velocity of fOld (LbDataSourceIterator.cc:34-43 semantics). */
void orc_macro(const orc_sim* s, double* rho, double* u) {
  const int Q = s->L.Q;
  for (int64_t i = 0; i < s->n; ++i) {
    double r, m[3];
    dens_mom(&s->L, &s->f_old[i * Q], &r, m);
    if (rho) rho[i] = r;
    if (u)
      for (int a = 0; a < 3; ++a) u[3 * i + a] = m[a] / r;
  }
}

void orc_wall_shear_stress_all(const orc_sim* s, const float* normal,
                               double* out) {
  const int Q = s->L.Q;
  for (int64_t i = 0; i < s->n; ++i)
    out[i] = s->wmask[i] ? orc_wall_shear_stress(Q, s->tau, &s->f_old[i * Q],
                                                     &normal[3 * i])
                             : DBL_MAX;
}
