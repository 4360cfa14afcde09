/*
 * wake_oracle.c -- CPU restatement of the wake LBM time step.
 *
 * TEST INFRASTRUCTURE ONLY (see wake_oracle.h): the checker that the HIP
 * path is compared against, never part of the product.
 *
 * Each kernel restates the ISPC export named in its comment, statement for
 * statement in the same arithmetic order, compiled with -ffp-contract=off so
 * the FP32 build is an IEEE version of the reference's gang-of-one semantics.
 * (The reference itself is compiled with ISPC --opt=fast-math, which may
 * reassociate; the parity tolerances absorb that last-ulp difference.)
 *
 * Parity status per function is recorded at the top of each section:
 *   "pinned: <test>"      a reference known-answer test, a closed form or an
 *                          invariant in tests/ checks it;
 *   "parity unpinned"     nothing beyond this restatement checks it.
 */
#include "wake_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#if defined(ORACLE_DOUBLE)
#define RSQRT(x) sqrt(x)
#define RCOS(x) cos(x)
#define RSIN(x) sin(x)
#define RABS(x) fabs(x)
#define RFLOOR(x) floor(x)
#define RCEIL(x) ceil(x)
#else
#define RSQRT(x) sqrtf(x)
#define RCOS(x) cosf(x)
#define RSIN(x) sinf(x)
#define RABS(x) fabsf(x)
#define RFLOOR(x) floorf(x)
#define RCEIL(x) ceilf(x)
#endif

#define R(x) ((real)(x))

enum { CELL_FLUID = 0, CELL_SOLID = 1, CELL_INFLOW = 2, CELL_OUTFLOW = 3, CELL_MOVING_WALL = 6 };
enum { COLL_BGK = 0, COLL_TRT = 1, COLL_MRT = 2 };

int oracle_real_bytes(void) { return (int)sizeof(real); }

#ifdef _OPENMP
#include <omp.h>
int oracle_threads(void) { return omp_get_max_threads(); }
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
#else
int oracle_threads(void) { return 1; }
void oracle_set_threads(int n) { (void)n; }
#endif

/* ------------------------------------------------------------------------
 * Lattice tables.  include/wake/core/lattice_models.hpp:14-110 and the ISPC
 * copies include/wake/ispc/lbm_d2q9.isph:8-17, lbm_d3q19.isph:8-28,
 * lbm_d3q27.isph:8-42.  The direction order is the population-plane order.
 * pinned: test_oracle_tables (the velocity sets, opposite involution with
 * c_opp = -c, sum w = 1, sum w c = 0, sum w c c = 1/3 delta).
 * ---------------------------------------------------------------------- */
static const int c9x[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int c9y[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int opp9[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
static const real w9[9] = {R(4) / R(9),  R(1) / R(9),  R(1) / R(9),  R(1) / R(9), R(1) / R(9),
                           R(1) / R(36), R(1) / R(36), R(1) / R(36), R(1) / R(36)};

static const int c19x[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int c19y[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int c19z[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
static const int opp19[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
static const real w19[19] = {R(1) / R(3),  R(1) / R(18), R(1) / R(18), R(1) / R(18), R(1) / R(18),
                             R(1) / R(18), R(1) / R(18), R(1) / R(36), R(1) / R(36), R(1) / R(36),
                             R(1) / R(36), R(1) / R(36), R(1) / R(36), R(1) / R(36), R(1) / R(36),
                             R(1) / R(36), R(1) / R(36), R(1) / R(36), R(1) / R(36)};

static const int c27x[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
static const int c27y[27] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, -1};
static const int c27z[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
/* The reference's D3Q27 opposite table (lattice_models.hpp:104-109,
 * lbm_d3q27.isph:125-130) pairs the corners as 19<->22, 20<->21, 23<->26,
 * 24<->25, which are mirror images in x and y only ((1,1,1) with (-1,-1,1)).
 * It is kept for the restatement of the d3q27_collide export; every grid-level
 * D3Q27 path uses the geometric opposite c_opp = -c (opp27_geo). */
static const int opp27[27] = {0,  2,  1,  4,  3,  6,  5,  10, 9,  8,  7,  14, 13, 12,
                              11, 18, 17, 16, 15, 22, 21, 20, 19, 26, 25, 24, 23};
static const int opp27_geo[27] = {0,  2,  1,  4,  3,  6,  5,  10, 9,  8,  7,  14, 13, 12,
                                  11, 18, 17, 16, 15, 26, 25, 24, 23, 22, 21, 20, 19};
static const real w27[27] = {R(8) / R(27),  R(2) / R(27),  R(2) / R(27),  R(2) / R(27),  R(2) / R(27),
                             R(2) / R(27),  R(2) / R(27),  R(1) / R(54),  R(1) / R(54),  R(1) / R(54),
                             R(1) / R(54),  R(1) / R(54),  R(1) / R(54),  R(1) / R(54),  R(1) / R(54),
                             R(1) / R(54),  R(1) / R(54),  R(1) / R(54),  R(1) / R(54),  R(1) / R(216),
                             R(1) / R(216), R(1) / R(216), R(1) / R(216), R(1) / R(216), R(1) / R(216),
                             R(1) / R(216), R(1) / R(216)};

typedef struct {
  int Q;
  const int *cx, *cy, *cz, *opp;
  const real* w;
} lat_t;
static const int zero27[27] = {0};
static lat_t lattice_of(int lattice) {
  lat_t L;
  if (lattice == 9) {
    L.Q = 9; L.cx = c9x; L.cy = c9y; L.cz = zero27; L.opp = opp9; L.w = w9;
  } else if (lattice == 27) {
    L.Q = 27; L.cx = c27x; L.cy = c27y; L.cz = c27z; L.opp = opp27_geo; L.w = w27;
  } else {
    L.Q = 19; L.cx = c19x; L.cy = c19y; L.cz = c19z; L.opp = opp19; L.w = w19;
  }
  return L;
}

/* Counter-based RNG for seeded noise and inlet turbulence (no reference
 * counterpart: the reference has no RNG).  splitmix64 finaliser of
 * seed ^ counter, top 24 bits to [0,1).  The product uses the same published
 * splitmix64 constants. pinned: test_oracle_rng_* (known splitmix64 outputs). */
uint64_t oracle_rng_splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
double oracle_rng_u01(uint64_t seed, uint64_t counter) {
  return (double)(oracle_rng_splitmix64(seed ^ counter) >> 40) * (1.0 / 16777216.0);
}

/* ------------------------------------------------------------------------
 * D3Q19 MRT moment basis.  The forward transform restates
 * src/ispc/d3q19_collision_mrt.ispc:66-105 (explicit sums).  The inverse is
 * the reference's table d3q19_Minv (d3q19_collision_mrt.ispc:4-24), which is
 * inv(M) rounded to 7 decimals; it is regenerated here from M rather than
 * copied.  pinned: test_oracle_mrt_forward_matches_dhumieres_basis,
 * test_oracle_mrt_inverse_is_rounded_inverse,
 * test_oracle_mrt_inverse_equals_reference_literals (round(inv(M),7) == the
 * reference's table),
 * test_oracle_mrt_mass_drift_signature.
 * ---------------------------------------------------------------------- */
static void d3q19_forward(const real* f, real* m) {
  m[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] +
         f[14] + f[15] + f[16] + f[17] + f[18];
  m[1] = R(-30) * f[0] - R(11) * (f[1] + f[2] + f[3] + f[4] + f[5] + f[6]) +
         R(8) * (f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18]);
  m[2] = R(12) * f[0] - R(4) * (f[1] + f[2] + f[3] + f[4] + f[5] + f[6]) +
         (f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18]);
  m[3] = (f[1] - f[2]) + (f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14]);
  m[4] = R(-4) * (f[1] - f[2]) + (f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14]);
  m[5] = (f[3] - f[4]) + (f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18]);
  m[6] = R(-4) * (f[3] - f[4]) + (f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18]);
  m[7] = (f[5] - f[6]) + (f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18]);
  m[8] = R(-4) * (f[5] - f[6]) + (f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18]);
  m[9] = R(2) * (f[1] + f[2]) - (f[3] + f[4] + f[5] + f[6]) +
         (f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] + f[14]) - R(2) * (f[15] + f[16] + f[17] + f[18]);
  m[10] = R(-4) * (f[1] + f[2]) + R(2) * (f[3] + f[4] + f[5] + f[6]) +
          (f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] + f[14]) - R(2) * (f[15] + f[16] + f[17] + f[18]);
  m[11] = (f[3] + f[4]) - (f[5] + f[6]) + (f[7] + f[8] + f[9] + f[10]) - (f[11] + f[12] + f[13] + f[14]);
  m[12] = R(-2) * (f[3] + f[4]) + R(2) * (f[5] + f[6]) + (f[7] + f[8] + f[9] + f[10]) -
          (f[11] + f[12] + f[13] + f[14]);
  m[13] = (f[7] - f[8] - f[9] + f[10]);
  m[14] = (f[15] - f[16] - f[17] + f[18]);
  m[15] = (f[11] - f[12] - f[13] + f[14]);
  m[16] = (f[7] - f[8] + f[9] - f[10] - f[11] + f[12] - f[13] + f[14]);
  m[17] = (-f[7] - f[8] + f[9] + f[10] + f[15] - f[16] + f[17] - f[18]);
  m[18] = (-f[11] - f[12] + f[13] + f[14] + f[15] + f[16] - f[17] - f[18]);
}

static real g_minv[19][19];
static int g_minv_ready = 0;

static void build_minv(void) {
  double M[19][19], A[19][38];
  for (int q = 0; q < 19; ++q) {
    real e[19] = {0}, m[19];
    e[q] = R(1);
    d3q19_forward(e, m);
    for (int k = 0; k < 19; ++k) M[k][q] = (double)m[k];
  }
  for (int i = 0; i < 19; ++i)
    for (int j = 0; j < 38; ++j) A[i][j] = (j < 19) ? M[i][j] : (j - 19 == i ? 1.0 : 0.0);
  for (int col = 0; col < 19; ++col) { /* Gauss-Jordan with partial pivoting */
    int piv = col;
    for (int r = col + 1; r < 19; ++r)
      if (fabs(A[r][col]) > fabs(A[piv][col])) piv = r;
    if (piv != col)
      for (int j = 0; j < 38; ++j) { double t = A[col][j]; A[col][j] = A[piv][j]; A[piv][j] = t; }
    double d = A[col][col];
    for (int j = 0; j < 38; ++j) A[col][j] /= d;
    for (int r = 0; r < 19; ++r) {
      if (r == col) continue;
      double fct = A[r][col];
      if (fct != 0.0)
        for (int j = 0; j < 38; ++j) A[r][j] -= fct * A[col][j];
    }
  }
  for (int q = 0; q < 19; ++q)
    for (int k = 0; k < 19; ++k) {
      double v = A[q][19 + k];
      double r7 = nearbyint(v * 1e7) / 1e7; /* 7-decimal rounding of the table */
      if (r7 == 0.0) r7 = 0.0;              /* no negative zeros */
      g_minv[q][k] = (real)r7;
    }
  g_minv_ready = 1;
}

void oracle_d3q19_minv(real out[19 * 19]) {
  if (!g_minv_ready) build_minv();
  for (int q = 0; q < 19; ++q)
    for (int k = 0; k < 19; ++k) out[q * 19 + k] = g_minv[q][k];
}
void oracle_d3q19_m(real out[19 * 19]) {
  for (int q = 0; q < 19; ++q) {
    real e[19] = {0}, m[19];
    e[q] = R(1);
    d3q19_forward(e, m);
    for (int k = 0; k < 19; ++k) out[k * 19 + q] = m[k];
  }
}

/* ========================================================================
 * D2Q9
 * ===================================================================== */

/* src/ispc/d2q9_moments.ispc:4-32.  pinned: test_oracle_moments_* (analytic
 * feq round trip from tests/unit/test_d2q9_moments.cpp:28-61). */
void oracle_d2q9_compute_macroscopic_moments(const real* f, real* rho, real* ux, real* uy, int count) {
  for (int i = 0; i < count; ++i) {
    real f0 = f[0 * (size_t)count + i], f1 = f[1 * (size_t)count + i], f2 = f[2 * (size_t)count + i];
    real f3 = f[3 * (size_t)count + i], f4 = f[4 * (size_t)count + i], f5 = f[5 * (size_t)count + i];
    real f6 = f[6 * (size_t)count + i], f7 = f[7 * (size_t)count + i], f8 = f[8 * (size_t)count + i];
    real local_rho = f0 + f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8;
    real inv_rho = R(1) / local_rho;
    real jx = (f1 - f3) + (f5 - f6 - f7 + f8);
    real jy = (f2 - f4) + (f5 + f6 - f7 - f8);
    rho[i] = local_rho;
    ux[i] = jx * inv_rho;
    uy[i] = jy * inv_rho;
  }
}

/* src/ispc/d2q9_moments.ispc:35-65.  pinned: test_oracle_moments_* (g shift). */
void oracle_d2q9_compute_macroscopic_moments_forced(const real* f, real* rho, real* ux, real* uy, real gx,
                                                    real gy, int count) {
  for (int i = 0; i < count; ++i) {
    real f0 = f[0 * (size_t)count + i], f1 = f[1 * (size_t)count + i], f2 = f[2 * (size_t)count + i];
    real f3 = f[3 * (size_t)count + i], f4 = f[4 * (size_t)count + i], f5 = f[5 * (size_t)count + i];
    real f6 = f[6 * (size_t)count + i], f7 = f[7 * (size_t)count + i], f8 = f[8 * (size_t)count + i];
    real local_rho = f0 + f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8;
    real inv_rho = R(1) / local_rho;
    real jx = (f1 - f3) + (f5 - f6 - f7 + f8);
    real jy = (f2 - f4) + (f5 + f6 - f7 - f8);
    rho[i] = local_rho;
    ux[i] = (jx + R(0.5) * local_rho * gx) * inv_rho;
    uy[i] = (jy + R(0.5) * local_rho * gy) * inv_rho;
  }
}

/* src/ispc/d2q9_moments.ispc:68-79. pinned: test_oracle_moments_*. */
void oracle_d2q9_compute_velocity_magnitude(const real* ux, const real* uy, real* speed, int count) {
  for (int i = 0; i < count; ++i) speed[i] = RSQRT(ux[i] * ux[i] + uy[i] * uy[i]);
}

/* Shared BGK/TRT/MRT per-node bodies.  tau_node != NULL selects the
 * LES-coupled EXTENSION (per-node shear tau); otherwise the scalar tau of the
 * ISPC signature is used. */

/* src/ispc/d2q9_collision_bgk.ispc:5-86. pinned: test_oracle_collision_*
 * (conservation, solid copy from tests/unit/test_d2q9_collision.cpp;
 * (1-omega) contraction of f-feq; Poiseuille). */
static void d2q9_bgk_node(const real* f_in, real* f_out, size_t n, size_t i, real* rho, real* ux, real* uy,
                          real tau, real gx, real gy) {
  real omega = R(1) / tau;
  real force_prefactor = (R(1) - R(0.5) * omega);
  real f[9];
  for (int q = 0; q < 9; ++q) f[q] = f_in[q * n + i];
  real local_rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
  real inv_rho = R(1) / local_rho;
  real jx = (f[1] - f[3]) + (f[5] - f[6] - f[7] + f[8]);
  real jy = (f[2] - f[4]) + (f[5] + f[6] - f[7] - f[8]);
  real lux = (jx + R(0.5) * local_rho * gx) * inv_rho;
  real luy = (jy + R(0.5) * local_rho * gy) * inv_rho;
  if (rho) { rho[i] = local_rho; ux[i] = lux; uy[i] = luy; }
  real u_dot_g = lux * gx + luy * gy;
  real u_sq = lux * lux + luy * luy;
  for (int q = 0; q < 9; ++q) {
    real ci_x = (real)c9x[q], ci_y = (real)c9y[q];
    real ci_u = ci_x * lux + ci_y * luy;
    real ci_g = ci_x * gx + ci_y * gy;
    real feq = w9[q] * local_rho * (R(1) + R(3) * ci_u + R(4.5) * ci_u * ci_u - R(1.5) * u_sq);
    real Fi = force_prefactor * w9[q] * local_rho * (R(3) * (ci_g - u_dot_g) + R(9) * ci_u * ci_g);
    f_out[q * n + i] = f[q] - omega * (f[q] - feq) + Fi;
  }
}

void oracle_d2q9_collide_bgk(const real* f_in, real* f_out, const uint8_t* cell_type, real* rho, real* ux,
                             real* uy, real tau, real gx, real gy, int count) {
  size_t n = (size_t)count;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    if (cell_type[i] == CELL_SOLID) {
      for (int q = 0; q < 9; ++q) f_out[q * n + i] = f_in[q * n + i];
      continue;
    }
    d2q9_bgk_node(f_in, f_out, n, (size_t)i, rho, ux, uy, tau, gx, gy);
  }
}

/* src/ispc/d2q9_collision_trt.ispc:7-92. pinned: test_oracle_collision_*
 * (conservation; TRT with tau_a == tau_s equals BGK). */
static void d2q9_trt_node(const real* f_in, real* f_out, size_t n, size_t i, real* rho, real* ux, real* uy,
                          real tau_s, real tau_a, real gx, real gy) {
  real omega_s = R(1) / tau_s, omega_a = R(1) / tau_a;
  real prefactor_sym = (R(1) - R(0.5) * omega_s);
  real prefactor_asym = (R(1) - R(0.5) * omega_a);
  real f[9], feq[9];
  for (int q = 0; q < 9; ++q) f[q] = f_in[q * n + i];
  real local_rho = R(0);
  for (int q = 0; q < 9; ++q) local_rho += f[q];
  real inv_rho = R(1) / local_rho;
  real jx = (f[1] - f[3]) + (f[5] - f[6] - f[7] + f[8]);
  real jy = (f[2] - f[4]) + (f[5] + f[6] - f[7] - f[8]);
  real lux = (jx + R(0.5) * local_rho * gx) * inv_rho;
  real luy = (jy + R(0.5) * local_rho * gy) * inv_rho;
  if (rho) { rho[i] = local_rho; ux[i] = lux; uy[i] = luy; }
  real u_dot_g = lux * gx + luy * gy;
  real u_sq = lux * lux + luy * luy;
  for (int q = 0; q < 9; ++q) {
    real ci_u = (real)c9x[q] * lux + (real)c9y[q] * luy;
    feq[q] = w9[q] * local_rho * (R(1) + R(3) * ci_u + R(4.5) * ci_u * ci_u - R(1.5) * u_sq);
  }
  for (int q = 0; q < 9; ++q) {
    int inv_q = opp9[q];
    real f_sym = R(0.5) * (f[q] + f[inv_q]);
    real f_asym = R(0.5) * (f[q] - f[inv_q]);
    real feq_sym = R(0.5) * (feq[q] + feq[inv_q]);
    real feq_asym = R(0.5) * (feq[q] - feq[inv_q]);
    real ci_x = (real)c9x[q], ci_y = (real)c9y[q];
    real ci_u = ci_x * lux + ci_y * luy;
    real ci_g = ci_x * gx + ci_y * gy;
    real F_asym = prefactor_asym * w9[q] * local_rho * (R(3) * ci_g);
    real F_sym = prefactor_sym * w9[q] * local_rho * (-R(3) * u_dot_g + R(9) * ci_u * ci_g);
    real Fi = F_sym + F_asym;
    f_out[q * n + i] = f[q] - omega_s * (f_sym - feq_sym) - omega_a * (f_asym - feq_asym) + Fi;
  }
}

void oracle_d2q9_collide_trt(const real* f_in, real* f_out, const uint8_t* cell_type, real* rho, real* ux,
                             real* uy, real tau_s, real tau_a, real gx, real gy, int count) {
  size_t n = (size_t)count;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    if (cell_type[i] == CELL_SOLID) {
      for (int q = 0; q < 9; ++q) f_out[q * n + i] = f_in[q * n + i];
      continue;
    }
    d2q9_trt_node(f_in, f_out, n, (size_t)i, rho, ux, uy, tau_s, tau_a, gx, gy);
  }
}

/* src/ispc/d2q9_collision_mrt.ispc:6-137 (Lallemand-Luo basis, exact
 * fraction inverse).  pinned: test_oracle_collision_* (conservation; all rates
 * 1 except conserved gives f' = feq up to the m2/m1 closure; TG decay). */
static void d2q9_mrt_node(const real* f_in, real* f_out, size_t n, size_t i, real* rho, real* ux, real* uy,
                          real tau_shear, real tau_bulk, real gx, real gy) {
  real s_rho = R(0), s_e = R(1) / tau_bulk, s_eps = R(1.4), s_jx = R(0), s_q = R(1.6), s_jy = R(0);
  real s_p = R(1) / tau_shear;
  real f0 = f_in[0 * n + i], f1 = f_in[1 * n + i], f2 = f_in[2 * n + i], f3 = f_in[3 * n + i];
  real f4 = f_in[4 * n + i], f5 = f_in[5 * n + i], f6 = f_in[6 * n + i], f7 = f_in[7 * n + i];
  real f8 = f_in[8 * n + i];
  real m0 = f0 + f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8;
  real m1 = R(-4) * f0 - (f1 + f2 + f3 + f4) + R(2) * (f5 + f6 + f7 + f8);
  real m2 = R(4) * f0 - R(2) * (f1 + f2 + f3 + f4) + (f5 + f6 + f7 + f8);
  real m3 = (f1 - f3) + (f5 - f6 - f7 + f8);
  real m4 = R(-2) * (f1 - f3) + (f5 - f6 - f7 + f8);
  real m5 = (f2 - f4) + (f5 + f6 - f7 - f8);
  real m6 = R(-2) * (f2 - f4) + (f5 + f6 - f7 - f8);
  real m7 = (f1 - f2 + f3 - f4);
  real m8 = (f5 - f6 + f7 - f8);
  real local_rho = m0;
  real inv_rho = R(1) / local_rho;
  real lux = (m3 + R(0.5) * local_rho * gx) * inv_rho;
  real luy = (m5 + R(0.5) * local_rho * gy) * inv_rho;
  if (rho) { rho[i] = local_rho; ux[i] = lux; uy[i] = luy; }
  real ux2 = lux * lux, uy2 = luy * luy, u2 = ux2 + uy2;
  real meq0 = local_rho;
  real meq1 = local_rho * (R(-2) + R(3) * u2);
  real meq2 = local_rho * (R(1) - R(3) * u2);
  real meq3 = local_rho * lux;
  real meq4 = -local_rho * lux;
  real meq5 = local_rho * luy;
  real meq6 = -local_rho * luy;
  real meq7 = local_rho * (ux2 - uy2);
  real meq8 = local_rho * (lux * luy);
  real F_pre = (R(1) - R(0.5) * s_p);
  real F_m3 = local_rho * gx, F_m5 = local_rho * gy;
  real F_m1 = R(6) * (lux * F_m3 + luy * F_m5);
  real F_m2 = R(-6) * (lux * F_m3 + luy * F_m5);
  real F_m4 = -F_m3, F_m6 = -F_m5;
  real F_m7 = R(2) * (lux * F_m3 - luy * F_m5);
  real F_m8 = (lux * F_m5 + luy * F_m3);
  m0 = m0 - s_rho * (m0 - meq0);
  m1 = m1 - s_e * (m1 - meq1) + F_pre * F_m1;
  m2 = m2 - s_eps * (m2 - meq2) + F_pre * F_m2;
  m3 = m3 - s_jx * (m3 - meq3) + F_m3;
  m4 = m4 - s_q * (m4 - meq4) + F_pre * F_m4;
  m5 = m5 - s_jy * (m5 - meq5) + F_m5;
  m6 = m6 - s_q * (m6 - meq6) + F_pre * F_m6;
  m7 = m7 - s_p * (m7 - meq7) + F_pre * F_m7;
  m8 = m8 - s_p * (m8 - meq8) + F_pre * F_m8;
  const real a9 = R(1) / R(9), a18 = R(1) / R(18), a36 = R(1) / R(36), a6 = R(1) / R(6), a12 = R(1) / R(12);
  f_out[0 * n + i] = a9 * m0 - a9 * m1 + a9 * m2;
  f_out[1 * n + i] = a9 * m0 - a36 * m1 - a18 * m2 + a6 * m3 - a6 * m4 + R(0.25) * m7;
  f_out[2 * n + i] = a9 * m0 - a36 * m1 - a18 * m2 + a6 * m5 - a6 * m6 - R(0.25) * m7;
  f_out[3 * n + i] = a9 * m0 - a36 * m1 - a18 * m2 - a6 * m3 + a6 * m4 + R(0.25) * m7;
  f_out[4 * n + i] = a9 * m0 - a36 * m1 - a18 * m2 - a6 * m5 + a6 * m6 - R(0.25) * m7;
  f_out[5 * n + i] = a9 * m0 + a18 * m1 + a36 * m2 + a6 * m3 + a12 * m4 + a6 * m5 + a12 * m6 + R(0.25) * m8;
  f_out[6 * n + i] = a9 * m0 + a18 * m1 + a36 * m2 - a6 * m3 - a12 * m4 + a6 * m5 + a12 * m6 - R(0.25) * m8;
  f_out[7 * n + i] = a9 * m0 + a18 * m1 + a36 * m2 - a6 * m3 - a12 * m4 - a6 * m5 - a12 * m6 + R(0.25) * m8;
  f_out[8 * n + i] = a9 * m0 + a18 * m1 + a36 * m2 + a6 * m3 + a12 * m4 - a6 * m5 - a12 * m6 - R(0.25) * m8;
}

void oracle_d2q9_collide_mrt(const real* f_in, real* f_out, const uint8_t* cell_type, real* rho, real* ux,
                             real* uy, real tau_shear, real tau_bulk, real gx, real gy, int count) {
  size_t n = (size_t)count;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    if (cell_type[i] == CELL_SOLID) {
      for (int q = 0; q < 9; ++q) f_out[q * n + i] = f_in[q * n + i];
      continue;
    }
    d2q9_mrt_node(f_in, f_out, n, (size_t)i, rho, ux, uy, tau_shear, tau_bulk, gx, gy);
  }
}

/* src/ispc/d2q9_streaming.ispc:4-29.  pinned: test_oracle_streaming_*
 * (permutation; a delta moves by c_q). */
void oracle_d2q9_stream_periodic(const real* f_src, real* f_dst, int nx, int ny) {
  size_t total = (size_t)nx * ny;
  for (int q = 0; q < 9; ++q) {
    int cx = c9x[q], cy = c9y[q];
    const real* s = f_src + q * total;
    real* d = f_dst + q * total;
    for (int y = 0; y < ny; ++y) {
      int src_y = (y - cy + ny) % ny;
      for (int x = 0; x < nx; ++x) {
        int src_x = (x - cx + nx) % nx;
        d[(size_t)y * nx + x] = s[(size_t)src_y * nx + src_x];
      }
    }
  }
}

/* src/ispc/d2q9_streaming.ispc:33-95 (x periodic, y walls, solid halfway
 * bounce-back, moving-wall momentum).  With per_y = 1 (EXTENSION) y wraps
 * instead of reflecting.  pinned: test_oracle_streaming_* (closed-box mass
 * conservation), test_oracle_poiseuille_* (analytic profile, reference
 * tests/validation/test_poiseuille_convergence.cpp). */
void oracle_d2q9_stream_general(const real* f_src, real* f_dst, const uint8_t* cell_type, int nx, int ny,
                                int per_y, real uw_x, real uw_y) {
  size_t total = (size_t)nx * ny;
  real mom[9];
  for (int q = 0; q < 9; ++q) {
    real ci_uw = (real)c9x[q] * uw_x + (real)c9y[q] * uw_y;
    mom[q] = R(2) * w9[q] * R(1) * R(3) * ci_uw;
  }
  for (int q = 0; q < 9; ++q) {
    int cx = c9x[q], cy = c9y[q], oq = opp9[q];
    const real* sq = f_src + q * total;
    const real* so = f_src + oq * total;
    real* d = f_dst + q * total;
    for (int y = 0; y < ny; ++y) {
      int src_y = y - cy;
      if (per_y) src_y = (src_y + ny) % ny;
      for (int x = 0; x < nx; ++x) {
        size_t idx = (size_t)y * nx + x;
        if (cell_type[idx] == CELL_SOLID) { d[idx] = sq[idx]; continue; }
        int src_x = (x - cx + nx) % nx;
        if (src_y >= 0 && src_y < ny) {
          size_t sidx = (size_t)src_y * nx + src_x;
          uint8_t nt = cell_type[sidx];
          if (nt == CELL_SOLID) d[idx] = so[idx];
          else if (nt == CELL_MOVING_WALL) d[idx] = so[idx] + mom[q];
          else d[idx] = sq[sidx];
        } else {
          d[idx] = so[idx];
        }
      }
    }
  }
}
void oracle_d2q9_stream_with_boundaries(const real* f_src, real* f_dst, const uint8_t* cell_type, int nx,
                                        int ny, real uw_x, real uw_y) {
  oracle_d2q9_stream_general(f_src, f_dst, cell_type, nx, ny, 0, uw_x, uw_y);
}

/* src/ispc/d2q9_boundaries.ispc:5-39 (uncalled by the reference grid).
 * pinned: test_d2q9_apply_bounce_back_closed_form (mass kept,
 * j -> -j + 2 u_w, an involution at u_w = 0). */
void oracle_d2q9_apply_bounce_back(real* f, const uint8_t* cell_type, real uw_x, real uw_y, int count) {
  size_t n = (size_t)count;
  real mom[9];
  for (int q = 0; q < 9; ++q) {
    real ci_uw = (real)c9x[q] * uw_x + (real)c9y[q] * uw_y;
    mom[q] = R(2) * w9[q] * R(1) * R(3) * ci_uw;
  }
  for (int i = 0; i < count; ++i) {
    if (cell_type[i] == CELL_SOLID || cell_type[i] == CELL_MOVING_WALL) {
      real fc[9];
      for (int q = 0; q < 9; ++q) fc[q] = f[q * n + i];
      f[0 * n + i] = fc[0];
      for (int q = 1; q < 9; ++q) f[q * n + i] = fc[opp9[q]] + mom[q];
    }
  }
}

/* src/ispc/d2q9_boundaries.ispc:43-81.  pinned: test_oracle_zou_he_* (after the
 * update sum f c_x = rho_in u_x and sum f = rho_in at every inlet node). */
static void d2q9_zh_inflow_node(real* f, size_t total, size_t idx, real u_in_x, real u_in_y, real* rho_o) {
  real f0 = f[0 * total + idx], f2 = f[2 * total + idx], f3 = f[3 * total + idx];
  real f4 = f[4 * total + idx], f6 = f[6 * total + idx], f7 = f[7 * total + idx];
  real rho_in = (f0 + f2 + f4 + R(2) * (f3 + f6 + f7)) / (R(1) - u_in_x);
  real f1 = f3 + (R(2) / R(3)) * rho_in * u_in_x;
  real f5 = f7 - R(0.5) * (f2 - f4) + (R(1) / R(6)) * rho_in * u_in_x + R(0.5) * rho_in * u_in_y;
  real f8 = f6 + R(0.5) * (f2 - f4) + (R(1) / R(6)) * rho_in * u_in_x - R(0.5) * rho_in * u_in_y;
  f[1 * total + idx] = f1;
  f[5 * total + idx] = f5;
  f[8 * total + idx] = f8;
  *rho_o = rho_in;
}
void oracle_d2q9_apply_zou_he_inflow_west(real* f, real* rho, real* ux, real* uy, real u_in_x, real u_in_y,
                                          int nx, int ny) {
  size_t total = (size_t)nx * ny;
  for (int y = 1; y < ny - 1; ++y) {
    size_t idx = (size_t)y * nx;
    real r;
    d2q9_zh_inflow_node(f, total, idx, u_in_x, u_in_y, &r);
    rho[idx] = r; ux[idx] = u_in_x; uy[idx] = u_in_y;
  }
}

/* src/ispc/d2q9_boundaries.ispc:85-121.  pinned: test_oracle_zou_he_* (sum f =
 * rho_out at every outlet node). */
void oracle_d2q9_apply_zou_he_outflow_east(real* f, real* rho, real* ux, real* uy, real rho_out, int nx,
                                           int ny) {
  size_t total = (size_t)nx * ny;
  for (int y = 1; y < ny - 1; ++y) {
    size_t idx = (size_t)y * nx + (nx - 1);
    real f0 = f[0 * total + idx], f1 = f[1 * total + idx], f2 = f[2 * total + idx];
    real f4 = f[4 * total + idx], f5 = f[5 * total + idx], f8 = f[8 * total + idx];
    real u_out_x = R(-1) + (f0 + f2 + f4 + R(2) * (f1 + f5 + f8)) / rho_out;
    real u_out_y = R(0);
    real f3 = f1 - (R(2) / R(3)) * rho_out * u_out_x;
    real f6 = f8 - R(0.5) * (f2 - f4) - (R(1) / R(6)) * rho_out * u_out_x + R(0.5) * rho_out * u_out_y;
    real f7 = f5 + R(0.5) * (f2 - f4) - (R(1) / R(6)) * rho_out * u_out_x - R(0.5) * rho_out * u_out_y;
    f[3 * total + idx] = f3;
    f[6 * total + idx] = f6;
    f[7 * total + idx] = f7;
    rho[idx] = rho_out; ux[idx] = u_out_x; uy[idx] = u_out_y;
  }
}

/* src/ispc/d2q9_boundaries.ispc:125-146 (uncalled by the reference grid).
 * pinned: test_convective_outlet_is_implicit_upwind_advection (zero residual
 * of the implicit upwind advection equation on random states, the alpha and
 * beta limits, the x = nx - 2 upstream column on a ramp, untouched corner
 * rows), test_oracle_convective_fixed_point. */
void oracle_d2q9_apply_convective_outflow_east(real* f, const real* f_prev, real u_conv, int nx, int ny) {
  size_t total = (size_t)nx * ny;
  real alpha = u_conv / (R(1) + u_conv);
  real beta = R(1) / (R(1) + u_conv);
  for (int y = 1; y < ny - 1; ++y) {
    size_t io = (size_t)y * nx + (nx - 1), ii = (size_t)y * nx + (nx - 2);
    for (int q = 0; q < 9; ++q) f[q * total + io] = beta * f_prev[q * total + io] + alpha * f[q * total + ii];
  }
}

/* src/ispc/d2q9_les.ispc:6-57.  pinned:
 * test_d2q9_smagorinsky_tau_equilibrium_and_brute_force (equilibrium gives
 * tau_0 exactly; the closed form on a brute-force einsum Pi_neq; monotone in
 * the stress). */
void oracle_d2q9_compute_smagorinsky_tau(const real* f, const real* rho, const real* ux, const real* uy,
                                         real* tau_eff, real tau_0, real smagorinsky_constant, int count) {
  size_t n = (size_t)count;
  real c_smag_sq = smagorinsky_constant * smagorinsky_constant;
  real prefactor = R(18) * R(1.4142135623730951) * c_smag_sq;
  for (int i = 0; i < count; ++i) {
    real lr = rho[i], lux = ux[i], luy = uy[i];
    real u_sq = lux * lux + luy * luy;
    real pxx = R(0), pxy = R(0), pyy = R(0);
    for (int q = 0; q < 9; ++q) {
      real fi = f[q * n + i];
      real ci_x = (real)c9x[q], ci_y = (real)c9y[q];
      real ci_u = ci_x * lux + ci_y * luy;
      real feq = w9[q] * lr * (R(1) + R(3) * ci_u + R(4.5) * ci_u * ci_u - R(1.5) * u_sq);
      real fn = fi - feq;
      pxx += ci_x * ci_x * fn;
      pxy += ci_x * ci_y * fn;
      pyy += ci_y * ci_y * fn;
    }
    real tensor_sum = pxx * pxx + R(2) * pxy * pxy + pyy * pyy;
    real Q_stress = RSQRT(R(2) * tensor_sum);
    real disc = tau_0 * tau_0 + (prefactor / lr) * Q_stress;
    tau_eff[i] = R(0.5) * (tau_0 + RSQRT(disc));
  }
}

/* src/ispc/forces_momentum_exchange.ispc:6-68.  Sums are accumulated in
 * double (the reference sums FP32 lanes with a gang-width-dependent
 * reduce_add; the order is not part of its contract).  pinned:
 * test_oracle_mem_* (rest equilibrium gives 0; +0.05 on f1 at (6,7) gives
 * F_x = 0.1, tests/unit/test_forces_momentum_exchange.cpp:33-68). */
static void d2q9_mem_d(const real* f, const uint8_t* cell_type, int nx, int ny, real center_x, real center_y,
                       real max_radius, double* out) {
  size_t total = (size_t)nx * ny;
  real r_max_sq = max_radius * max_radius;
  double sfx = 0, sfy = 0, stz = 0;
  for (int q = 1; q < 9; ++q) {
    int cx = c9x[q], cy = c9y[q];
    const real* fq = f + q * total;
    for (int y = 0; y < ny; ++y) {
      int ny_ = y + cy;
      if (ny_ < 0 || ny_ >= ny) continue;
      for (int x = 0; x < nx; ++x) {
        size_t idx = (size_t)y * nx + x;
        if (cell_type[idx] != CELL_FLUID) continue;
        int nxx = (x + cx + nx) % nx;
        uint8_t nt = cell_type[(size_t)ny_ * nx + nxx];
        if (nt != CELL_SOLID && nt != CELL_MOVING_WALL) continue;
        real rx = ((real)x + R(0.5) * (real)cx) - center_x;
        real ry = ((real)y + R(0.5) * (real)cy) - center_y;
        if (max_radius <= R(0) || (rx * rx + ry * ry <= r_max_sq)) {
          real fv = fq[idx];
          real dfx = R(2) * fv * (real)cx, dfy = R(2) * fv * (real)cy;
          sfx += (double)dfx;
          sfy += (double)dfy;
          stz += (double)(rx * dfy - ry * dfx);
        }
      }
    }
  }
  out[0] = sfx; out[1] = sfy; out[2] = stz;
}
void oracle_d2q9_compute_momentum_exchange_forces(const real* f, const uint8_t* cell_type, int nx, int ny,
                                                  real center_x, real center_y, real max_radius,
                                                  real* fx_out, real* fy_out, real* torque_z_out) {
  double o[3];
  d2q9_mem_d(f, cell_type, nx, ny, center_x, center_y, max_radius, o);
  *fx_out = (real)o[0]; *fy_out = (real)o[1]; *torque_z_out = (real)o[2];
}

/* ========================================================================
 * D3Q19 / D3Q27
 * ===================================================================== */

/* src/ispc/d3q19_moments.ispc:4-66, d3q27_moments.ispc:4-32.  pinned:
 * test_oracle_moments_* (analytic feq from tests/unit/test_d3q19_moments.cpp). */
static void moments3d(const lat_t* L, const real* f, real* rho, real* ux, real* uy, real* uz, real gx, real gy,
                      real gz, int forced, int count) {
  size_t n = (size_t)count;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    real lr = R(0), jx = R(0), jy = R(0), jz = R(0);
    for (int q = 0; q < L->Q; ++q) {
      real fi = f[q * n + i];
      lr += fi;
      jx += fi * (real)L->cx[q];
      jy += fi * (real)L->cy[q];
      jz += fi * (real)L->cz[q];
    }
    real inv_rho = R(1) / lr;
    rho[i] = lr;
    if (forced) {
      ux[i] = (jx + R(0.5) * lr * gx) * inv_rho;
      uy[i] = (jy + R(0.5) * lr * gy) * inv_rho;
      uz[i] = (jz + R(0.5) * lr * gz) * inv_rho;
    } else {
      ux[i] = jx * inv_rho;
      uy[i] = jy * inv_rho;
      uz[i] = jz * inv_rho;
    }
  }
}
void oracle_d3q19_compute_macroscopic_moments(const real* f, real* rho, real* ux, real* uy, real* uz,
                                              int count) {
  lat_t L = lattice_of(19);
  moments3d(&L, f, rho, ux, uy, uz, 0, 0, 0, 0, count);
}
void oracle_d3q19_compute_macroscopic_moments_forced(const real* f, real* rho, real* ux, real* uy, real* uz,
                                                     real gx, real gy, real gz, int count) {
  lat_t L = lattice_of(19);
  moments3d(&L, f, rho, ux, uy, uz, gx, gy, gz, 1, count);
}
void oracle_d3q27_compute_macroscopic_moments(const real* f, real* rho, real* ux, real* uy, real* uz,
                                              int count) {
  lat_t L = lattice_of(27);
  moments3d(&L, f, rho, ux, uy, uz, 0, 0, 0, 0, count);
}
/* src/ispc/d3q19_moments.ispc:69-82. */
void oracle_d3q19_compute_velocity_magnitude(const real* ux, const real* uy, const real* uz, real* speed,
                                             int count) {
  for (int i = 0; i < count; ++i) speed[i] = RSQRT(ux[i] * ux[i] + uy[i] * uy[i] + uz[i] * uz[i]);
}

/* src/ispc/d3q19_collision_bgk.ispc:4-75.  pinned: test_oracle_collision_*. */
static void bgk3d_node(const lat_t* L, const real* f_in, real* f_out, size_t n, size_t i, real* rho, real* ux,
                       real* uy, real* uz, real tau, real gx, real gy, real gz) {
  real omega = R(1) / tau;
  real force_prefactor = (R(1) - R(0.5) * omega);
  real f[27];
  real lr = R(0), jx = R(0), jy = R(0), jz = R(0);
  for (int q = 0; q < L->Q; ++q) {
    real fi = f_in[q * n + i];
    f[q] = fi;
    lr += fi;
    jx += fi * (real)L->cx[q];
    jy += fi * (real)L->cy[q];
    jz += fi * (real)L->cz[q];
  }
  real inv_rho = R(1) / lr;
  real lux = (jx + R(0.5) * lr * gx) * inv_rho;
  real luy = (jy + R(0.5) * lr * gy) * inv_rho;
  real luz = (jz + R(0.5) * lr * gz) * inv_rho;
  if (rho) { rho[i] = lr; ux[i] = lux; uy[i] = luy; uz[i] = luz; }
  real u_dot_g = lux * gx + luy * gy + luz * gz;
  real u_sq = lux * lux + luy * luy + luz * luz;
  for (int q = 0; q < L->Q; ++q) {
    real ci_x = (real)L->cx[q], ci_y = (real)L->cy[q], ci_z = (real)L->cz[q];
    real ci_u = ci_x * lux + ci_y * luy + ci_z * luz;
    real ci_g = ci_x * gx + ci_y * gy + ci_z * gz;
    real feq = L->w[q] * lr * (R(1) + R(3) * ci_u + R(4.5) * ci_u * ci_u - R(1.5) * u_sq);
    real Fi = force_prefactor * L->w[q] * lr * (R(3) * (ci_g - u_dot_g) + R(9) * ci_u * ci_g);
    f_out[q * n + i] = f[q] - omega * (f[q] - feq) + Fi;
  }
}

/* src/ispc/d3q19_collision_trt.ispc:5-93 and d3q27_collision_mrt.ispc:5-91
 * (the D3Q27 export is TRT).  pinned: test_oracle_collision_*. */
static void trt3d_node(const lat_t* L, const real* f_in, real* f_out, size_t n, size_t i, real* rho, real* ux,
                       real* uy, real* uz, real tau_s, real tau_a, real gx, real gy, real gz) {
  real omega_s = R(1) / tau_s, omega_a = R(1) / tau_a;
  real prefactor_sym = (R(1) - R(0.5) * omega_s);
  real prefactor_asym = (R(1) - R(0.5) * omega_a);
  real f[27], feq[27];
  real lr = R(0), jx = R(0), jy = R(0), jz = R(0);
  for (int q = 0; q < L->Q; ++q) {
    real fi = f_in[q * n + i];
    f[q] = fi;
    lr += fi;
    jx += fi * (real)L->cx[q];
    jy += fi * (real)L->cy[q];
    jz += fi * (real)L->cz[q];
  }
  real inv_rho = R(1) / lr;
  real lux = (jx + R(0.5) * lr * gx) * inv_rho;
  real luy = (jy + R(0.5) * lr * gy) * inv_rho;
  real luz = (jz + R(0.5) * lr * gz) * inv_rho;
  if (rho) { rho[i] = lr; ux[i] = lux; uy[i] = luy; uz[i] = luz; }
  real u_dot_g = lux * gx + luy * gy + luz * gz;
  real u_sq = lux * lux + luy * luy + luz * luz;
  for (int q = 0; q < L->Q; ++q) {
    real ci_u = (real)L->cx[q] * lux + (real)L->cy[q] * luy + (real)L->cz[q] * luz;
    feq[q] = L->w[q] * lr * (R(1) + R(3) * ci_u + R(4.5) * ci_u * ci_u - R(1.5) * u_sq);
  }
  for (int q = 0; q < L->Q; ++q) {
    int inv_q = L->opp[q];
    real f_sym = R(0.5) * (f[q] + f[inv_q]);
    real f_asym = R(0.5) * (f[q] - f[inv_q]);
    real feq_sym = R(0.5) * (feq[q] + feq[inv_q]);
    real feq_asym = R(0.5) * (feq[q] - feq[inv_q]);
    real ci_x = (real)L->cx[q], ci_y = (real)L->cy[q], ci_z = (real)L->cz[q];
    real ci_u = ci_x * lux + ci_y * luy + ci_z * luz;
    real ci_g = ci_x * gx + ci_y * gy + ci_z * gz;
    real F_asym = prefactor_asym * L->w[q] * lr * (R(3) * ci_g);
    real F_sym = prefactor_sym * L->w[q] * lr * (-R(3) * u_dot_g + R(9) * ci_u * ci_g);
    real Fi = F_sym + F_asym;
    f_out[q * n + i] = f[q] - omega_s * (f_sym - feq_sym) - omega_a * (f_asym - feq_asym) + Fi;
  }
}

/* src/ispc/d3q19_collision_mrt.ispc:26-176 with the rounded inverse table.
 * pinned: test_oracle_collision_* (rho, jx conservation as in
 * tests/unit/test_d3q19_collision.cpp:237-254), test_oracle_mrt_mass_drift_signature
 * (-1.1e-6 per step mass drift at rest), test_oracle_taylor_green_3d_*
 * (tests/validation/test_taylor_green_decay.cpp). */
static void mrt19_node(const real* f_in, real* f_out, size_t n, size_t i, real* rho, real* ux, real* uy,
                       real* uz, real tau_shear, real tau_bulk, real gx, real gy, real gz) {
  if (!g_minv_ready) build_minv();
  real s_rho = R(0), s_e = R(1) / tau_bulk, s_eps = R(1.4), s_j = R(0), s_q = R(1.2);
  real s_nu = R(1) / tau_shear, s_pi = R(1.4), s_m = R(1.98);
  real f[19], m[19], meq[19];
  for (int q = 0; q < 19; ++q) f[q] = f_in[q * n + i];
  d3q19_forward(f, m);
  real lr = m[0];
  real inv_rho = R(1) / lr;
  real lux = (m[3] + R(0.5) * lr * gx) * inv_rho;
  real luy = (m[5] + R(0.5) * lr * gy) * inv_rho;
  real luz = (m[7] + R(0.5) * lr * gz) * inv_rho;
  if (rho) { rho[i] = lr; ux[i] = lux; uy[i] = luy; uz[i] = luz; }
  real ux2 = lux * lux, uy2 = luy * luy, uz2 = luz * luz;
  real u2 = ux2 + uy2 + uz2;
  meq[0] = lr;
  meq[1] = lr * (R(-11) + R(19) * u2);
  meq[2] = lr * (R(3) - R(11) / R(2) * u2);
  meq[3] = lr * lux;
  meq[4] = -R(2) / R(3) * lr * lux;
  meq[5] = lr * luy;
  meq[6] = -R(2) / R(3) * lr * luy;
  meq[7] = lr * luz;
  meq[8] = -R(2) / R(3) * lr * luz;
  meq[9] = lr * (R(2) * ux2 - uy2 - uz2);
  meq[10] = R(-0.5) * meq[9];
  meq[11] = lr * (uy2 - uz2);
  meq[12] = R(-0.5) * meq[11];
  meq[13] = lr * (lux * luy);
  meq[14] = lr * (luy * luz);
  meq[15] = lr * (lux * luz);
  meq[16] = R(0);
  meq[17] = R(0);
  meq[18] = R(0);
  m[0] = m[0] - s_rho * (m[0] - meq[0]);
  m[1] = m[1] - s_e * (m[1] - meq[1]);
  m[2] = m[2] - s_eps * (m[2] - meq[2]);
  m[3] = m[3] - s_j * (m[3] - meq[3]) + lr * gx;
  m[4] = m[4] - s_q * (m[4] - meq[4]) - (R(2) / R(3)) * lr * gx;
  m[5] = m[5] - s_j * (m[5] - meq[5]) + lr * gy;
  m[6] = m[6] - s_q * (m[6] - meq[6]) - (R(2) / R(3)) * lr * gy;
  m[7] = m[7] - s_j * (m[7] - meq[7]) + lr * gz;
  m[8] = m[8] - s_q * (m[8] - meq[8]) - (R(2) / R(3)) * lr * gz;
  m[9] = m[9] - s_nu * (m[9] - meq[9]);
  m[10] = m[10] - s_pi * (m[10] - meq[10]);
  m[11] = m[11] - s_nu * (m[11] - meq[11]);
  m[12] = m[12] - s_pi * (m[12] - meq[12]);
  m[13] = m[13] - s_nu * (m[13] - meq[13]);
  m[14] = m[14] - s_nu * (m[14] - meq[14]);
  m[15] = m[15] - s_nu * (m[15] - meq[15]);
  m[16] = m[16] - s_m * (m[16] - meq[16]);
  m[17] = m[17] - s_m * (m[17] - meq[17]);
  m[18] = m[18] - s_m * (m[18] - meq[18]);
  for (int q = 0; q < 19; ++q) {
    real val = R(0);
    for (int k = 0; k < 19; ++k) val += g_minv[q][k] * m[k];
    f_out[q * n + i] = val;
  }
}

void oracle_d3q19_collide_bgk(const real* f_in, real* f_out, const uint8_t* cell_type, real* rho, real* ux,
                              real* uy, real* uz, real tau, real gx, real gy, real gz, int count) {
  lat_t L = lattice_of(19);
  size_t n = (size_t)count;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    if (cell_type[i] == CELL_SOLID) {
      for (int q = 0; q < 19; ++q) f_out[q * n + i] = f_in[q * n + i];
      continue;
    }
    bgk3d_node(&L, f_in, f_out, n, (size_t)i, rho, ux, uy, uz, tau, gx, gy, gz);
  }
}
void oracle_d3q19_collide_trt(const real* f_in, real* f_out, const uint8_t* cell_type, real* rho, real* ux,
                              real* uy, real* uz, real tau_s, real tau_a, real gx, real gy, real gz,
                              int count) {
  lat_t L = lattice_of(19);
  size_t n = (size_t)count;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    if (cell_type[i] == CELL_SOLID) {
      for (int q = 0; q < 19; ++q) f_out[q * n + i] = f_in[q * n + i];
      continue;
    }
    trt3d_node(&L, f_in, f_out, n, (size_t)i, rho, ux, uy, uz, tau_s, tau_a, gx, gy, gz);
  }
}
void oracle_d3q19_collide_mrt(const real* f_in, real* f_out, const uint8_t* cell_type, real* rho, real* ux,
                              real* uy, real* uz, real tau_shear, real tau_bulk, real gx, real gy, real gz,
                              int count) {
  size_t n = (size_t)count;
  if (!g_minv_ready) build_minv();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    if (cell_type[i] == CELL_SOLID) {
      for (int q = 0; q < 19; ++q) f_out[q * n + i] = f_in[q * n + i];
      continue;
    }
    mrt19_node(f_in, f_out, n, (size_t)i, rho, ux, uy, uz, tau_shear, tau_bulk, gx, gy, gz);
  }
}
void oracle_d3q27_collide(const real* f_in, real* f_out, const uint8_t* cell_type, real* rho, real* ux,
                          real* uy, real* uz, real tau_s, real tau_a, real gx, real gy, real gz, int count) {
  lat_t L = lattice_of(27);
  L.opp = opp27; /* the export pairs populations with the reference table */
  size_t n = (size_t)count;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    if (cell_type[i] == CELL_SOLID) {
      for (int q = 0; q < 27; ++q) f_out[q * n + i] = f_in[q * n + i];
      continue;
    }
    trt3d_node(&L, f_in, f_out, n, (size_t)i, rho, ux, uy, uz, tau_s, tau_a, gx, gy, gz);
  }
}

/* TRT antisymmetric tau from the grid: lattice_grid_3d.cpp:328-330 and
 * lattice_grid_2d.cpp:220-222 (Lambda = 1/4). */
static real trt_tau_a(real tau_s) {
  real diff = tau_s - R(0.5);
  if (diff < R(1e-4)) diff = R(1e-4);
  return R(0.5) + R(0.25) / diff;
}

/* EXTENSION: LES-coupled collision, per-node shear tau.  pinned:
 * test_oracle_les_* (tau_node == tau everywhere reproduces the scalar kernels
 * bit for bit). */
void oracle_collide_tau_node(int lattice, int collision, const real* f_in, real* f_out,
                             const uint8_t* cell_type, const real* tau_node, real gx, real gy, real gz,
                             int count) {
  lat_t L = lattice_of(lattice);
  size_t n = (size_t)count;
  if (!g_minv_ready) build_minv();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    if (cell_type[i] == CELL_SOLID) {
      for (int q = 0; q < L.Q; ++q) f_out[q * n + i] = f_in[q * n + i];
      continue;
    }
    real t = tau_node[i];
    if (lattice == 9) {
      if (collision == COLL_BGK) d2q9_bgk_node(f_in, f_out, n, (size_t)i, NULL, NULL, NULL, t, gx, gy);
      else if (collision == COLL_TRT)
        d2q9_trt_node(f_in, f_out, n, (size_t)i, NULL, NULL, NULL, t, trt_tau_a(t), gx, gy);
      else d2q9_mrt_node(f_in, f_out, n, (size_t)i, NULL, NULL, NULL, t, R(1.63), gx, gy);
    } else if (lattice == 19) {
      if (collision == COLL_BGK) bgk3d_node(&L, f_in, f_out, n, (size_t)i, NULL, NULL, NULL, NULL, t, gx, gy, gz);
      else if (collision == COLL_TRT)
        trt3d_node(&L, f_in, f_out, n, (size_t)i, NULL, NULL, NULL, NULL, t, trt_tau_a(t), gx, gy, gz);
      else mrt19_node(f_in, f_out, n, (size_t)i, NULL, NULL, NULL, NULL, t, R(1.19), gx, gy, gz);
    } else {
      if (collision == COLL_BGK) bgk3d_node(&L, f_in, f_out, n, (size_t)i, NULL, NULL, NULL, NULL, t, gx, gy, gz);
      else trt3d_node(&L, f_in, f_out, n, (size_t)i, NULL, NULL, NULL, NULL, t, trt_tau_a(t), gx, gy, gz);
    }
  }
}

/* src/ispc/d3q19_streaming.ispc:4-35, d3q27_streaming.ispc:4-35.  pinned:
 * test_oracle_streaming_* (permutation, plane sums, delta moves by c_q). */
static void stream3d_periodic(const lat_t* L, const real* f_src, real* f_dst, int nx, int ny, int nz) {
  size_t total = (size_t)nx * ny * nz;
  for (int q = 0; q < L->Q; ++q) {
    int cx = L->cx[q], cy = L->cy[q], cz = L->cz[q];
    const real* s = f_src + q * total;
    real* d = f_dst + q * total;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z) {
      int sz = (z - cz + nz) % nz;
      for (int y = 0; y < ny; ++y) {
        int sy = (y - cy + ny) % ny;
        size_t doff = ((size_t)z * ny + y) * nx, soff = ((size_t)sz * ny + sy) * nx;
        for (int x = 0; x < nx; ++x) d[doff + x] = s[soff + (x - cx + nx) % nx];
      }
    }
  }
}
void oracle_d3q19_stream_periodic(const real* f_src, real* f_dst, int nx, int ny, int nz) {
  lat_t L = lattice_of(19);
  stream3d_periodic(&L, f_src, f_dst, nx, ny, nz);
}
void oracle_d3q27_stream_periodic(const real* f_src, real* f_dst, int nx, int ny, int nz) {
  lat_t L = lattice_of(27);
  stream3d_periodic(&L, f_src, f_dst, nx, ny, nz);
}

/* src/ispc/d3q19_streaming.ispc:39-104 (x periodic; y/z out of range and
 * solid neighbours reflect; moving-wall neighbours add 2 w 3 c.u_w).  With
 * per_y / per_z (EXTENSION) the axis wraps instead, and lattice 27 applies the
 * same rules to D3Q27.  pinned: test_oracle_streaming_* (closed-box mass
 * conservation; all-fluid periodic equals the periodic kernel). */
void oracle_stream3d_general(int lattice, const real* f_src, real* f_dst, const uint8_t* cell_type, int nx,
                             int ny, int nz, int per_y, int per_z, real uw_x, real uw_y, real uw_z) {
  lat_t L = lattice_of(lattice);
  size_t total = (size_t)nx * ny * nz;
  real mom[27];
  for (int q = 0; q < L.Q; ++q) {
    real ci_uw = (real)L.cx[q] * uw_x + (real)L.cy[q] * uw_y + (real)L.cz[q] * uw_z;
    mom[q] = R(2) * L.w[q] * R(1) * R(3) * ci_uw;
  }
  for (int q = 0; q < L.Q; ++q) {
    int cx = L.cx[q], cy = L.cy[q], cz = L.cz[q], oq = L.opp[q];
    const real* sq = f_src + q * total;
    const real* so = f_src + oq * total;
    real* d = f_dst + q * total;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z) {
      int sz = z - cz;
      if (per_z) sz = (sz + nz) % nz;
      for (int y = 0; y < ny; ++y) {
        int sy = y - cy;
        if (per_y) sy = (sy + ny) % ny;
        size_t doff = ((size_t)z * ny + y) * nx;
        for (int x = 0; x < nx; ++x) {
          size_t idx = doff + x;
          if (cell_type[idx] == CELL_SOLID) { d[idx] = sq[idx]; continue; }
          int sx = (x - cx + nx) % nx;
          if (sy >= 0 && sy < ny && sz >= 0 && sz < nz) {
            size_t sidx = ((size_t)sz * ny + sy) * nx + sx;
            uint8_t nt = cell_type[sidx];
            if (nt == CELL_SOLID) d[idx] = so[idx];
            else if (nt == CELL_MOVING_WALL) d[idx] = so[idx] + mom[q];
            else d[idx] = sq[sidx];
          } else {
            d[idx] = so[idx];
          }
        }
      }
    }
  }
}
void oracle_d3q19_stream_with_boundaries(const real* f_src, real* f_dst, const uint8_t* cell_type, int nx,
                                         int ny, int nz, real uw_x, real uw_y, real uw_z) {
  oracle_stream3d_general(19, f_src, f_dst, cell_type, nx, ny, nz, 0, 0, uw_x, uw_y, uw_z);
}

/* src/ispc/d3q19_les.ispc:7-66.  pinned: test_oracle_les_*. */
static void les3d(const lat_t* L, const real* f, const real* rho, const real* ux, const real* uy,
                  const real* uz, real* tau_eff, real tau_0, real cs, int count) {
  size_t n = (size_t)count;
  real c_smag_sq = cs * cs;
  real prefactor = R(18) * R(1.4142135623730951) * c_smag_sq;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    real lr = rho[i], lux = ux[i], luy = uy[i], luz = uz[i];
    real u_sq = lux * lux + luy * luy + luz * luz;
    real pxx = R(0), pyy = R(0), pzz = R(0), pxy = R(0), pyz = R(0), pxz = R(0);
    for (int q = 0; q < L->Q; ++q) {
      real fi = f[q * n + i];
      real ci_x = (real)L->cx[q], ci_y = (real)L->cy[q], ci_z = (real)L->cz[q];
      real ci_u = ci_x * lux + ci_y * luy + ci_z * luz;
      real feq = L->w[q] * lr * (R(1) + R(3) * ci_u + R(4.5) * ci_u * ci_u - R(1.5) * u_sq);
      real fn = fi - feq;
      pxx += ci_x * ci_x * fn;
      pyy += ci_y * ci_y * fn;
      pzz += ci_z * ci_z * fn;
      pxy += ci_x * ci_y * fn;
      pyz += ci_y * ci_z * fn;
      pxz += ci_x * ci_z * fn;
    }
    real tensor_sum = (pxx * pxx + pyy * pyy + pzz * pzz) + R(2) * (pxy * pxy + pyz * pyz + pxz * pxz);
    real Q_stress = RSQRT(R(2) * tensor_sum);
    real disc = tau_0 * tau_0 + (prefactor / lr) * Q_stress;
    tau_eff[i] = R(0.5) * (tau_0 + RSQRT(disc));
  }
}
void oracle_d3q19_compute_smagorinsky_tau(const real* f, const real* rho, const real* ux, const real* uy,
                                          const real* uz, real* tau_eff, real tau_0, real smagorinsky_constant,
                                          int count) {
  lat_t L = lattice_of(19);
  les3d(&L, f, rho, ux, uy, uz, tau_eff, tau_0, smagorinsky_constant, count);
}

/* EXTENSION: Smagorinsky tau from the populations alone, for any lattice;
 * rho and u come from f with the Guo shift, exactly as the forced moments
 * the reference passes in.  pinned: test_oracle_les_* (agrees bit for bit
 * with moments_forced + the reference LES kernel on D3Q19). */
void oracle_smagorinsky_tau_from_f(int lattice, const real* f, real* tau_eff, real tau_0, real cs, real gx,
                                   real gy, real gz, int count) {
  lat_t L = lattice_of(lattice);
  real* r = (real*)malloc(sizeof(real) * (size_t)count * 4);
  real *ux = r + count, *uy = r + 2 * (size_t)count, *uz = r + 3 * (size_t)count;
  if (lattice == 9) {
    oracle_d2q9_compute_macroscopic_moments_forced(f, r, ux, uy, gx, gy, count);
    oracle_d2q9_compute_smagorinsky_tau(f, r, ux, uy, tau_eff, tau_0, cs, count);
  } else {
    moments3d(&L, f, r, ux, uy, uz, gx, gy, gz, 1, count);
    les3d(&L, f, r, ux, uy, uz, tau_eff, tau_0, cs, count);
  }
  free(r);
}

/* src/ispc/d3q19_curved.ispc:7-71; per_y/per_z wrap x_ff (EXTENSION) the way
 * the streaming does.  pinned:
 * test_bfl_reproduces_uniform_equilibrium_of_the_wall_velocity (Galilean
 * consistency of both wall terms, delta 0.01..0.99),
 * test_bfl_interpolates_a_linear_field_at_the_departure_point (interpolation
 * weights and the upstream neighbour in both branches),
 * test_bfl_limits_and_upstream_neighbour (delta -> 0 and 1, the y-edge
 * fallback, the x wrap), test_oracle_bfl_half_delta_is_halfway_bounce_back;
 * tests/test_oracle_mutants.py shows planted mistakes fail them. */
static void bfl_general(const lat_t* L, real* f_dst, const real* f_src, const int* fluid_indices,
                        const int* q_dirs, const real* delta_fractions, const real* uw_x, const real* uw_y,
                        const real* uw_z, int nx, int ny, int nz, int per_y, int per_z, int nlinks) {
  size_t total = (size_t)nx * ny * nz;
  for (int link = 0; link < nlinks; ++link) {
    int idx_f = fluid_indices[link], q = q_dirs[link];
    real delta = delta_fractions[link];
    int opp_q = L->opp[q], cx = L->cx[q], cy = L->cy[q], cz = L->cz[q];
    real ci_uw = (real)cx * uw_x[link] + (real)cy * uw_y[link] + (real)cz * uw_z[link];
    real f_i_xf = f_src[q * total + idx_f];
    real f_opp_xf = f_src[opp_q * total + idx_f];
    real reflected;
    if (delta < R(0.5)) {
      int z_f = idx_f / (nx * ny);
      int rem = idx_f - z_f * (nx * ny);
      int y_f = rem / nx;
      int x_f = rem - y_f * nx;
      int x_ff = (x_f - cx + nx) % nx;
      int y_ff = y_f - cy, z_ff = z_f - cz;
      if (per_y) y_ff = (y_ff + ny) % ny;
      if (per_z) z_ff = (z_ff + nz) % nz;
      real f_i_xff = f_i_xf;
      if (y_ff >= 0 && y_ff < ny && z_ff >= 0 && z_ff < nz) {
        int idx_ff = (z_ff * ny + y_ff) * nx + x_ff;
        f_i_xff = f_src[q * total + idx_ff];
      }
      real mom = R(2) * L->w[q] * R(1) * R(3) * ci_uw;
      reflected = R(2) * delta * f_i_xf + (R(1) - R(2) * delta) * f_i_xff - mom;
    } else {
      real inv_2delta = R(1) / (R(2) * delta);
      real mom = (R(1) / delta) * L->w[q] * R(1) * R(3) * ci_uw;
      reflected = inv_2delta * f_i_xf + (R(1) - inv_2delta) * f_opp_xf - mom;
    }
    f_dst[opp_q * total + idx_f] = reflected;
  }
}
void oracle_d3q19_apply_curved_bfl(real* f_dst, const real* f_src, const int* fluid_indices, const int* q_dirs,
                                   const real* delta_fractions, const real* uw_x, const real* uw_y,
                                   const real* uw_z, int nx, int ny, int nz, int num_boundary_links) {
  lat_t L = lattice_of(19);
  bfl_general(&L, f_dst, f_src, fluid_indices, q_dirs, delta_fractions, uw_x, uw_y, uw_z, nx, ny, nz, 0, 0,
              num_boundary_links);
}

/* src/ispc/forces_momentum_exchange.ispc:71-138, double accumulation (see the
 * 2D note).  per_y/per_z (EXTENSION) wrap the neighbour instead of skipping.
 * pinned: test_mem3d_single_population_impulses (2 delta c_q for five
 * directions at linked nodes, nothing for the opposite population, for a link
 * outside r_max or across the non-periodic z edge),
 * test_oracle_mem_3d_symmetric_body_at_rest. */
static void mem3d(const lat_t* L, const real* f, const uint8_t* cell_type, int nx, int ny, int nz, int per_y,
                  int per_z, real cxx, real cyy, real czz, real max_radius, double* out) {
  size_t total = (size_t)nx * ny * nz;
  real r_max_sq = max_radius * max_radius;
  double sfx = 0, sfy = 0, sfz = 0;
  for (int q = 1; q < L->Q; ++q) {
    int cx = L->cx[q], cy = L->cy[q], cz = L->cz[q];
    const real* fq = f + q * total;
    for (int z = 0; z < nz; ++z) {
      int nz_ = z + cz;
      if (per_z) nz_ = (nz_ + nz) % nz;
      if (nz_ < 0 || nz_ >= nz) continue;
      for (int y = 0; y < ny; ++y) {
        int ny_ = y + cy;
        if (per_y) ny_ = (ny_ + ny) % ny;
        if (ny_ < 0 || ny_ >= ny) continue;
        size_t doff = ((size_t)z * ny + y) * nx, noff = ((size_t)nz_ * ny + ny_) * nx;
        for (int x = 0; x < nx; ++x) {
          size_t idx = doff + x;
          if (cell_type[idx] != CELL_FLUID) continue;
          uint8_t nt = cell_type[noff + (x + cx + nx) % nx];
          if (nt != CELL_SOLID && nt != CELL_MOVING_WALL) continue;
          real rx = ((real)x + R(0.5) * (real)cx) - cxx;
          real ry = ((real)y + R(0.5) * (real)cy) - cyy;
          real rz = ((real)z + R(0.5) * (real)cz) - czz;
          if (max_radius <= R(0) || (rx * rx + ry * ry + rz * rz <= r_max_sq)) {
            real fv = fq[idx];
            sfx += (double)(R(2) * fv * (real)cx);
            sfy += (double)(R(2) * fv * (real)cy);
            sfz += (double)(R(2) * fv * (real)cz);
          }
        }
      }
    }
  }
  out[0] = sfx; out[1] = sfy; out[2] = sfz;
}
void oracle_d3q19_compute_momentum_exchange_forces(const real* f, const uint8_t* cell_type, int nx, int ny,
                                                   int nz, real center_x, real center_y, real center_z,
                                                   real max_radius, real* fx_out, real* fy_out,
                                                   real* fz_out) {
  lat_t L = lattice_of(19);
  double o[3];
  mem3d(&L, f, cell_type, nx, ny, nz, 0, 0, center_x, center_y, center_z, max_radius, o);
  *fx_out = (real)o[0]; *fy_out = (real)o[1]; *fz_out = (real)o[2];
}

/* ========================================================================
 * Grid step sequences: LatticeGrid2D (src/grid/lattice_grid_2d.cpp) and
 * LatticeGrid3D (src/grid/lattice_grid_3d.cpp).
 * ===================================================================== */
struct oracle_grid {
  int dim, lattice, Q, nx, ny, nz;
  size_t N;
  uint64_t step_count;
  real tau;
  int is_periodic;          /* lattice_grid_3d.hpp:98 (2D: EXTENSION, default off) */
  int ext_axes, per_y, per_z; /* EXTENSION: per-axis periodicity with walls */
  int has_moving_wall_top;  /* 3D: set_moving_wall_top; 2D: set_moving_wall_north */
  real uwx, uwz;
  int has_inflow, inflow_parabolic;
  real u_in_x, u_in_y, u_max, turb_amp;
  uint64_t turb_seed;
  int has_outflow;
  real rho_out;
  int has_conv;
  real u_conv;
  int les_coupled;
  real *f_src, *f_dst, *f_prev, *rho, *ux, *uy, *uz, *speed, *tau_eff;
  uint8_t* cell;
  int nlinks;
  int *l_idx, *l_q;
  real *l_delta, *l_uw;
  /* EXTENSION: D2Q5 temperature with Boussinesq coupling */
  int thermal;
  real tau_g, gbeta, t_ref;
  real *g_src, *g_dst, *t_wall, *temp;
};

static void* xcalloc(size_t n, size_t s) {
  void* p = calloc(n ? n : 1, s);
  return p;
}

oracle_grid* oracle_grid_create(int dim, int lattice, int nx, int ny, int nz, real tau) {
  oracle_grid* g = (oracle_grid*)xcalloc(1, sizeof(oracle_grid));
  g->dim = dim;
  g->lattice = (dim == 2) ? 9 : (lattice == 27 ? 27 : 19);
  g->Q = g->lattice;
  g->nx = nx; g->ny = ny; g->nz = (dim == 2) ? 1 : nz;
  g->N = (size_t)nx * ny * g->nz;
  g->tau = tau;
  g->is_periodic = (dim == 3) ? 1 : 0;
  g->rho_out = R(1);
  size_t qn = (size_t)g->Q * g->N;
  g->f_src = (real*)xcalloc(qn, sizeof(real));
  g->f_dst = (real*)xcalloc(qn, sizeof(real));
  g->f_prev = (real*)xcalloc(qn, sizeof(real));
  g->rho = (real*)xcalloc(g->N, sizeof(real));
  g->ux = (real*)xcalloc(g->N, sizeof(real));
  g->uy = (real*)xcalloc(g->N, sizeof(real));
  g->uz = (real*)xcalloc(g->N, sizeof(real));
  g->speed = (real*)xcalloc(g->N, sizeof(real));
  g->tau_eff = (real*)xcalloc(g->N, sizeof(real));
  g->cell = (uint8_t*)xcalloc(g->N, 1);
  oracle_grid_initialize(g, R(1), R(0), R(0), R(0));
  return g;
}

void oracle_grid_destroy(oracle_grid* g) {
  if (!g) return;
  free(g->f_src); free(g->f_dst); free(g->f_prev); free(g->rho); free(g->ux); free(g->uy); free(g->uz);
  free(g->speed); free(g->tau_eff); free(g->cell);
  free(g->l_idx); free(g->l_q); free(g->l_delta); free(g->l_uw);
  free(g->g_src); free(g->g_dst); free(g->t_wall); free(g->temp);
  free(g);
}

static real feq_of(const lat_t* L, int q, real rho, real ux, real uy, real uz) {
  real ci_u = (real)L->cx[q] * ux + (real)L->cy[q] * uy + (real)L->cz[q] * uz;
  real u_sq = ux * ux + uy * uy + uz * uz;
  return L->w[q] * rho * (R(1) + R(3) * ci_u + R(4.5) * ci_u * ci_u - R(1.5) * u_sq);
}

/* lattice_grid_3d.cpp:66-93 / lattice_grid_2d.cpp:64-95: feq into both
 * buffers, macro fields set, every flag reset to Fluid.  The 2D feq has no
 * c_z term; with uz = 0 the 3D expression is the same float sequence. */
void oracle_grid_initialize(oracle_grid* g, real rho, real ux, real uy, real uz) {
  lat_t L = lattice_of(g->lattice);
  g->step_count = 0;
  if (g->dim == 2) uz = R(0);
  for (int q = 0; q < g->Q; ++q) {
    real fe;
    if (g->dim == 2) {
      real ci_u = (real)L.cx[q] * ux + (real)L.cy[q] * uy;
      real u_sq = ux * ux + uy * uy;
      fe = L.w[q] * rho * (R(1) + R(3) * ci_u + R(4.5) * ci_u * ci_u - R(1.5) * u_sq);
    } else {
      fe = feq_of(&L, q, rho, ux, uy, uz);
    }
    for (size_t i = 0; i < g->N; ++i) g->f_src[q * g->N + i] = g->f_dst[q * g->N + i] = fe;
  }
  real usq = (g->dim == 2) ? ux * ux + uy * uy : ux * ux + uy * uy + uz * uz;
  for (size_t i = 0; i < g->N; ++i) {
    g->rho[i] = rho; g->ux[i] = ux; g->uy[i] = uy; g->uz[i] = uz;
    g->speed[i] = RSQRT(usq);
    g->tau_eff[i] = g->tau;
    g->cell[i] = CELL_FLUID;
  }
  memcpy(g->f_prev, g->f_src, sizeof(real) * g->Q * g->N);
}

/* lattice_grid_3d.cpp:95-131 / lattice_grid_2d.cpp:97-120 with the
 * per-node values given as arrays instead of a callback; flags untouched. */
void oracle_grid_initialize_fields(oracle_grid* g, const real* rho, const real* ux, const real* uy,
                                   const real* uz) {
  lat_t L = lattice_of(g->lattice);
  g->step_count = 0;
  for (size_t i = 0; i < g->N; ++i) {
    real r = rho[i], a = ux[i], b = uy[i], c = (g->dim == 2 || !uz) ? R(0) : uz[i];
    g->rho[i] = r; g->ux[i] = a; g->uy[i] = b; g->uz[i] = c;
    real usq = (g->dim == 2) ? a * a + b * b : a * a + b * b + c * c;
    g->speed[i] = RSQRT(usq);
    g->tau_eff[i] = g->tau;
    for (int q = 0; q < g->Q; ++q) {
      real fe;
      if (g->dim == 2) {
        real ci_u = (real)L.cx[q] * a + (real)L.cy[q] * b;
        fe = L.w[q] * r * (R(1) + R(3) * ci_u + R(4.5) * ci_u * ci_u - R(1.5) * usq);
      } else {
        fe = feq_of(&L, q, r, a, b, c);
      }
      g->f_src[q * g->N + i] = g->f_dst[q * g->N + i] = fe;
    }
  }
  memcpy(g->f_prev, g->f_src, sizeof(real) * g->Q * g->N);
}

/* EXTENSION: seeded uniform noise U(-a, a) added to every population of both
 * buffers; counter = node * Q + q. */
void oracle_grid_add_noise(oracle_grid* g, real amplitude, uint64_t seed) {
  for (size_t i = 0; i < g->N; ++i)
    for (int q = 0; q < g->Q; ++q) {
      double u = oracle_rng_u01(seed, (uint64_t)i * (uint64_t)g->Q + (uint64_t)q);
      real d = amplitude * (real)(2.0 * u - 1.0);
      g->f_src[q * g->N + i] += d;
      g->f_dst[q * g->N + i] += d;
    }
  memcpy(g->f_prev, g->f_src, sizeof(real) * g->Q * g->N);
}

/* lattice_grid_3d.cpp:133-140 (any non-Fluid type clears is_periodic);
 * lattice_grid_2d.cpp:122-126. */
void oracle_grid_set_cell_type(oracle_grid* g, int x, int y, int z, int type) {
  if (x >= 0 && x < g->nx && y >= 0 && y < g->ny && z >= 0 && z < g->nz) {
    g->cell[((size_t)z * g->ny + y) * g->nx + x] = (uint8_t)type;
    if (type != CELL_FLUID) g->is_periodic = 0;
  }
}
static int get_cell(const oracle_grid* g, int x, int y, int z) {
  if (x >= 0 && x < g->nx && y >= 0 && y < g->ny && z >= 0 && z < g->nz)
    return g->cell[((size_t)z * g->ny + y) * g->nx + x];
  return CELL_SOLID;
}
void oracle_grid_set_periodic(oracle_grid* g, int periodic) { g->is_periodic = periodic; }
void oracle_grid_set_periodic_axes(oracle_grid* g, int per_y, int per_z) {
  g->ext_axes = 1; g->per_y = per_y; g->per_z = per_z;
}

/* lattice_grid_3d.cpp:149-158 */
void oracle_grid_set_solid_box(oracle_grid* g, int x0, int y0, int z0, int x1, int y1, int z1) {
  g->is_periodic = 0;
  for (int z = (z0 > 0 ? z0 : 0); z <= (z1 < g->nz - 1 ? z1 : g->nz - 1); ++z)
    for (int y = (y0 > 0 ? y0 : 0); y <= (y1 < g->ny - 1 ? y1 : g->ny - 1); ++y)
      for (int x = (x0 > 0 ? x0 : 0); x <= (x1 < g->nx - 1 ? x1 : g->nx - 1); ++x)
        oracle_grid_set_cell_type(g, x, y, z, CELL_SOLID);
}

static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* lattice_grid_3d.cpp:160-182 (bit-exact flag rule).  pinned:
 * test_oracle_sphere_flags_integer_count (voxel count equals an exact-integer count of lattice
 * points in the ball when the centre and radius are integers). */
void oracle_grid_set_solid_sphere(oracle_grid* g, real cx, real cy, real cz, real radius) {
  g->is_periodic = 0;
  real r_sq = radius * radius;
  int min_x = imax(0, (int)(cx - radius - R(1))), max_x = imin(g->nx - 1, (int)(cx + radius + R(1)));
  int min_y = imax(0, (int)(cy - radius - R(1))), max_y = imin(g->ny - 1, (int)(cy + radius + R(1)));
  int min_z = imax(0, (int)(cz - radius - R(1))), max_z = imin(g->nz - 1, (int)(cz + radius + R(1)));
  for (int z = min_z; z <= max_z; ++z)
    for (int y = min_y; y <= max_y; ++y)
      for (int x = min_x; x <= max_x; ++x) {
        real dx = (real)x - cx, dy = (real)y - cy, dz = (real)z - cz;
        if (dx * dx + dy * dy + dz * dz <= r_sq) oracle_grid_set_cell_type(g, x, y, z, CELL_SOLID);
      }
}

/* primitives.cpp:73-100 (axis-aligned: the reference's arithmetic); with
 * axes != NULL the rows of the row-major 3x3 `axes` are the body axes in
 * world coordinates, the point's body coordinates are b_k = axes[k] . d and
 * the box is bounded by the largest semi-axis (no reference counterpart).
 * pinned: test_ellipsoid_voxelizer_axes_conventions (identity axes change
 * nothing, swapped axes equal swapped semi-axes, volume within 3%). */
void oracle_grid_set_solid_ellipsoid(oracle_grid* g, real cx, real cy, real cz, real rx, real ry, real rz,
                                     const real* axes) {
  real ex = rx, ey = ry, ez = rz;
  if (axes) {
    real m = rx > ry ? rx : ry;
    m = m > rz ? m : rz;
    ex = ey = ez = m;
  }
  int min_x = imax(0, (int)(cx - ex - R(1))), max_x = imin(g->nx - 1, (int)(cx + ex + R(1)));
  int min_y = imax(0, (int)(cy - ey - R(1))), max_y = imin(g->ny - 1, (int)(cy + ey + R(1)));
  int min_z = imax(0, (int)(cz - ez - R(1))), max_z = imin(g->nz - 1, (int)(cz + ez + R(1)));
  real irx = R(1) / (rx * rx), iry = R(1) / (ry * ry), irz = R(1) / (rz * rz);
  for (int z = min_z; z <= max_z; ++z)
    for (int y = min_y; y <= max_y; ++y)
      for (int x = min_x; x <= max_x; ++x) {
        real dx = (real)x - cx, dy = (real)y - cy, dz = (real)z - cz;
        if (axes) {
          real b0 = axes[0] * dx + axes[1] * dy + axes[2] * dz;
          real b1 = axes[3] * dx + axes[4] * dy + axes[5] * dz;
          real b2 = axes[6] * dx + axes[7] * dy + axes[8] * dz;
          dx = b0;
          dy = b1;
          dz = b2;
        }
        if (dx * dx * irx + dy * dy * iry + dz * dz * irz <= R(1)) oracle_grid_set_cell_type(g, x, y, z, CELL_SOLID);
      }
}

/* naca_airfoil.cpp:23-41 (digits) and :113-150 (contains_point). */
static int naca_contains(real camber, real camber_pos, real thickness, real chord, real aoa_deg, real x, real y,
                         real lead_x, real lead_y) {
  const real pi = R(3.14159265358979323846);
  real dx = x - lead_x;
  real dy = y - lead_y;
  real alpha_rad = -aoa_deg * (pi / R(180));
  real cos_a = RCOS(-alpha_rad);
  real sin_a = RSIN(-alpha_rad);
  real x_chord = dx * cos_a - dy * sin_a;
  real y_chord = dx * sin_a + dy * cos_a;
  if (x_chord < R(0) || x_chord > chord) return 0;
  real xc = x_chord / chord;
  real yt = R(5) * thickness *
            (R(0.2969) * RSQRT(xc) - R(0.1260) * xc - R(0.3516) * xc * xc + R(0.2843) * xc * xc * xc -
             R(0.1036) * xc * xc * xc * xc) *
            chord;
  real yc = R(0);
  if (camber_pos > R(1e-6)) {
    if (xc < camber_pos) {
      yc = (camber / (camber_pos * camber_pos)) * (R(2) * camber_pos * xc - xc * xc) * chord;
    } else {
      real one_minus_p = R(1) - camber_pos;
      yc = (camber / (one_minus_p * one_minus_p)) * ((R(1) - R(2) * camber_pos) + R(2) * camber_pos * xc - xc * xc) *
           chord;
    }
  }
  return RABS(y_chord - yc) <= yt;
}

/* naca_airfoil.cpp:152-190: rasterize (2D) / voxelize_extruded over z0..z1.
 * pinned: test_naca_voxelizer_shape (symmetry at 0 deg, section area
 * 0.685 t c^2, thickest section 0.12 c, the trailing edge below the leading
 * edge at +8 deg, extrusion equal to the 2D section). */
void oracle_grid_set_solid_naca(oracle_grid* g, int digits, real chord, real aoa_deg, real lead_x, real lead_y,
                                int z0, int z1) {
  real camber = (real)(digits / 1000) / R(100), camber_pos = (real)((digits / 100) % 10) / R(10);
  real thickness = (real)(digits % 100) / R(100);
  real rad_bound = chord * R(1.5);
  int min_x = imax(0, (int)(lead_x - rad_bound)), max_x = imin(g->nx - 1, (int)(lead_x + rad_bound));
  int min_y = imax(0, (int)(lead_y - rad_bound)), max_y = imin(g->ny - 1, (int)(lead_y + rad_bound));
  int za = g->dim == 2 ? 0 : imax(0, z0), zb = g->dim == 2 ? 0 : imin(g->nz - 1, z1);
  g->is_periodic = 0;
  for (int y = min_y; y <= max_y; ++y)
    for (int x = min_x; x <= max_x; ++x)
      if (naca_contains(camber, camber_pos, thickness, chord, aoa_deg, (real)x, (real)y, lead_x, lead_y))
        for (int z = za; z <= zb; ++z) oracle_grid_set_cell_type(g, x, y, z, CELL_SOLID);
}

/* lattice_grid_2d.cpp:143-159.  pinned: test_oracle_shape_setters_published_counts
 * (Gauss circle counts). */
void oracle_grid_set_solid_circle(oracle_grid* g, real cx, real cy, real radius) {
  real r_sq = radius * radius;
  int min_x = imax(0, (int)(cx - radius - R(1))), max_x = imin(g->nx - 1, (int)(cx + radius + R(1)));
  int min_y = imax(0, (int)(cy - radius - R(1))), max_y = imin(g->ny - 1, (int)(cy + radius + R(1)));
  for (int y = min_y; y <= max_y; ++y)
    for (int x = min_x; x <= max_x; ++x) {
      real dx = (real)x - cx, dy = (real)y - cy;
      if (dx * dx + dy * dy <= r_sq) oracle_grid_set_cell_type(g, x, y, 0, CELL_SOLID);
    }
}

/* lattice_grid_3d.cpp:184-203.  pinned: test_oracle_shape_setters_published_counts
 * (nz x the Gauss circle count). */
void oracle_grid_set_solid_cylinder_z(oracle_grid* g, real cx, real cy, real radius) {
  g->is_periodic = 0;
  real r_sq = radius * radius;
  int min_x = imax(0, (int)(cx - radius - R(1))), max_x = imin(g->nx - 1, (int)(cx + radius + R(1)));
  int min_y = imax(0, (int)(cy - radius - R(1))), max_y = imin(g->ny - 1, (int)(cy + radius + R(1)));
  for (int z = 0; z < g->nz; ++z)
    for (int y = min_y; y <= max_y; ++y)
      for (int x = min_x; x <= max_x; ++x) {
        real dx = (real)x - cx, dy = (real)y - cy;
        if (dx * dx + dy * dy <= r_sq) oracle_grid_set_cell_type(g, x, y, z, CELL_SOLID);
      }
}

/* lattice_grid_3d.cpp:205-215 and lattice_grid_2d.cpp:178-184 (north wall).
 * pinned: test_oracle_grid_setters_and_diagnostics_closed_forms (the cells
 * marked), test_oracle_moving_wall_adds_momentum (the wall term). */
void oracle_grid_set_moving_wall_top(oracle_grid* g, real uwx, real uwz) {
  g->is_periodic = 0;
  g->has_moving_wall_top = 1;
  g->uwx = uwx;
  g->uwz = (g->dim == 2) ? R(0) : uwz;
  for (int z = 0; z < g->nz; ++z)
    for (int x = 0; x < g->nx; ++x) oracle_grid_set_cell_type(g, x, g->ny - 1, z, CELL_MOVING_WALL);
}

/* lattice_grid_2d.cpp:161-168 (3D: EXTENSION, the whole x = 0 plane).
 * pinned: test_oracle_grid_setters_and_diagnostics_closed_forms,
 * test_oracle_zou_he_2d_identities, test_oracle_zou_he_3d_identities. */
void oracle_grid_set_inflow_west(oracle_grid* g, real ux, real uy) {
  g->has_inflow = 1; g->u_in_x = ux; g->u_in_y = uy; g->inflow_parabolic = 0;
  for (int z = 0; z < g->nz; ++z)
    for (int y = 0; y < g->ny; ++y) oracle_grid_set_cell_type(g, 0, y, z, CELL_INFLOW);
}
/* EXTENSION: parabolic profile u(y) = 4 u_max s (H - s) / H^2, s = y - 1/2,
 * H = ny - 2 (halfway walls at y = 1/2 and ny - 3/2).  pinned:
 * test_oracle_grid_setters_and_diagnostics_closed_forms (symmetry, the row
 * mean u (2/3 + 1/(3 H^2)) and the peak u (1 - 1/H^2)). */
void oracle_grid_set_inflow_parabolic(oracle_grid* g, real u_max) {
  oracle_grid_set_inflow_west(g, u_max, R(0));
  g->inflow_parabolic = 1; g->u_max = u_max;
}
/* EXTENSION: u_x(node, step) = u_in (1 + a (2 U - 1)), U = rng(seed, step*N + node).
 * pinned: test_oracle_rng_splitmix64_known_values (the generator),
 * test_oracle_regularized_inlet. */
void oracle_grid_set_inflow_turbulence(oracle_grid* g, real amplitude, uint64_t seed) {
  g->turb_amp = amplitude; g->turb_seed = seed;
}
/* lattice_grid_2d.cpp:170-176.  pinned: test_oracle_grid_setters_and_diagnostics_closed_forms
 * (the cells marked), test_oracle_zou_he_2d_identities (the closure). */
void oracle_grid_set_outflow_east(oracle_grid* g, real rho_out) {
  g->has_outflow = 1; g->has_conv = 0; g->rho_out = rho_out;
  for (int z = 0; z < g->nz; ++z)
    for (int y = 0; y < g->ny; ++y) oracle_grid_set_cell_type(g, g->nx - 1, y, z, CELL_OUTFLOW);
}
/* EXTENSION: convective outflow (the reference kernel
 * d2q9_boundaries.ispc:125-146 has no caller; here it runs each step with the
 * previous step's outlet populations). */
void oracle_grid_set_convective_outflow_east(oracle_grid* g, real u_conv) {
  g->has_conv = 1; g->has_outflow = 0; g->u_conv = u_conv;
  for (int z = 0; z < g->nz; ++z)
    for (int y = 0; y < g->ny; ++y) oracle_grid_set_cell_type(g, g->nx - 1, y, z, CELL_OUTFLOW);
  memcpy(g->f_prev, g->f_src, sizeof(real) * g->Q * g->N);
}
void oracle_grid_set_les_coupled(oracle_grid* g, int coupled) { g->les_coupled = coupled; }

/* lattice_grid_3d.cpp:217-290: links in z -> y -> x -> q order, Fluid nodes
 * only, neighbour x periodic and y/z in range (wrapped under EXTENSION
 * per-axis periodicity), FP32 ray-sphere quadratic, clamp [0.01, 0.99].
 * pinned: test_oracle_links_* (brute-force segment/sphere intersection on small
 * R; every link joins a Fluid node to a Solid neighbour). */
int oracle_grid_build_curved_sphere_links(oracle_grid* g, real cx, real cy, real cz, real radius) {
  lat_t L = lattice_of(g->lattice);
  g->is_periodic = 0;
  int cap = 1024, n = 0;
  int* li = (int*)malloc(sizeof(int) * cap);
  int* lq = (int*)malloc(sizeof(int) * cap);
  real* ld = (real*)malloc(sizeof(real) * cap);
  int per_y = g->ext_axes && g->per_y, per_z = g->ext_axes && g->per_z;
  for (int z = 0; z < g->nz; ++z)
    for (int y = 0; y < g->ny; ++y)
      for (int x = 0; x < g->nx; ++x) {
        if (get_cell(g, x, y, z) != CELL_FLUID) continue;
        int idx_f = (z * g->ny + y) * g->nx + x;
        for (int q = 1; q < L.Q; ++q) {
          int nxx = (x + L.cx[q] + g->nx) % g->nx;
          int nyy = y + L.cy[q], nzz = z + L.cz[q];
          if (per_y) nyy = (nyy + g->ny) % g->ny;
          if (per_z) nzz = (nzz + g->nz) % g->nz;
          if (nyy >= 0 && nyy < g->ny && nzz >= 0 && nzz < g->nz) {
            if (get_cell(g, nxx, nyy, nzz) == CELL_SOLID) {
              real p0x = (real)x - cx, p0y = (real)y - cy, p0z = (real)z - cz;
              real dx = (real)L.cx[q], dy = (real)L.cy[q], dz = (real)L.cz[q];
              real a = dx * dx + dy * dy + dz * dz;
              real b = R(2) * (p0x * dx + p0y * dy + p0z * dz);
              real c = (p0x * p0x + p0y * p0y + p0z * p0z) - radius * radius;
              real disc = b * b - R(4) * a * c;
              real delta = R(0.5);
              if (disc >= R(0)) {
                real t1 = (-b - RSQRT(disc)) / (R(2) * a);
                real t2 = (-b + RSQRT(disc)) / (R(2) * a);
                if (t1 >= R(0) && t1 <= R(1)) delta = t1;
                else if (t2 >= R(0) && t2 <= R(1)) delta = t2;
              }
              if (delta < R(0.01)) delta = R(0.01);
              if (delta > R(0.99)) delta = R(0.99);
              if (n == cap) {
                cap *= 2;
                li = (int*)realloc(li, sizeof(int) * cap);
                lq = (int*)realloc(lq, sizeof(int) * cap);
                ld = (real*)realloc(ld, sizeof(real) * cap);
              }
              li[n] = idx_f; lq[n] = q; ld[n] = delta; ++n;
            }
          }
        }
      }
  free(g->l_idx); free(g->l_q); free(g->l_delta); free(g->l_uw);
  g->l_idx = li; g->l_q = lq; g->l_delta = ld;
  g->l_uw = (real*)xcalloc((size_t)3 * (n ? n : 1), sizeof(real));
  g->nlinks = n;
  return n;
}
int oracle_grid_num_links(const oracle_grid* g) { return g->nlinks; }
void oracle_grid_get_links(const oracle_grid* g, int* idx, int* q, real* delta) {
  for (int i = 0; i < g->nlinks; ++i) { idx[i] = g->l_idx[i]; q[i] = g->l_q[i]; delta[i] = g->l_delta[i]; }
}

/* Rows an open boundary acts on: the reference 2D kernels use rows
 * 1..ny-2 (d2q9_boundaries.ispc:55, :96); a periodic axis (EXTENSION) uses
 * every row. */
static void bc_range(const oracle_grid* g, int* y0, int* y1, int* z0, int* z1) {
  int py = g->ext_axes && g->per_y, pz = g->ext_axes && g->per_z;
  *y0 = py ? 0 : 1; *y1 = py ? g->ny : g->ny - 1;
  if (g->dim == 2) { *z0 = 0; *z1 = 1; }
  else { *z0 = pz ? 0 : 1; *z1 = pz ? g->nz : g->nz - 1; }
}

/* EXTENSION: D3Q19 / D3Q27 regularized Zou-He velocity inlet on the x = 0
 * plane.  The unknown (c_x = +1) populations are first completed by
 * non-equilibrium bounce-back of their geometric opposite with a transverse
 * correction shared equally by the unknowns (Hecht & Harting 2010):
 *   rho  = (sum_{cx=0} f + 2 sum_{cx=-1} f) / (1 - u_x)
 *   N_t  = (sum_{cx=0} f c_t - (2/3) rho u_t) / sum_{unknown} c_t^2
 *   f_q  = f_opp(q) + 6 w_q rho (c_q . u) - (c_qy N_y + c_qz N_z)
 * and then the whole node is regularized (Latt & Chopard 2006): with
 * Pi = sum_q c_q c_q (f_q - feq_q),
 *   f_q = feq_q + 4.5 w_q (c_q c_q - I/3) : Pi.
 * Regularization keeps rho, rho u and Pi and removes the non-hydrodynamic
 * content that the plain closure feeds into an MRT collision (which
 * diverges at the inlet under perturbed inflow).  pinned:
 * test_oracle_zou_he_3d_identities (sum f = rho, sum f c = rho u) and
 * test_oracle_regularized_inlet (Pi preserved, third-order moments zero). */
static void zh3d_velocity_node(const lat_t* L, real* f, size_t total, size_t idx, real ux, real uy, real uz) {
  real fv[27], fe[27];
  for (int q = 0; q < L->Q; ++q) fv[q] = f[q * total + idx];
  real s0 = R(0), sm = R(0), ty = R(0), tz = R(0);
  int cnt_y = 0, cnt_z = 0;
  for (int q = 0; q < L->Q; ++q) {
    if (L->cx[q] == 0) {
      s0 += fv[q];
      ty += fv[q] * (real)L->cy[q];
      tz += fv[q] * (real)L->cz[q];
    } else if (L->cx[q] < 0) {
      sm += fv[q];
    } else {
      cnt_y += L->cy[q] * L->cy[q];
      cnt_z += L->cz[q] * L->cz[q];
    }
  }
  real rho = (s0 + R(2) * sm) / (R(1) - ux);
  real Ny = (ty - (R(2) / R(3)) * rho * uy) / (real)cnt_y;
  real Nz = (tz - (R(2) / R(3)) * rho * uz) / (real)cnt_z;
  for (int q = 0; q < L->Q; ++q) {
    if (L->cx[q] != 1) continue;
    real cu = (real)L->cx[q] * ux + (real)L->cy[q] * uy + (real)L->cz[q] * uz;
    fv[q] = fv[L->opp[q]] + R(6) * L->w[q] * rho * cu - ((real)L->cy[q] * Ny + (real)L->cz[q] * Nz);
  }
  real pxx = R(0), pyy = R(0), pzz = R(0), pxy = R(0), pyz = R(0), pxz = R(0);
  for (int q = 0; q < L->Q; ++q) {
    fe[q] = feq_of(L, q, rho, ux, uy, uz);
    real n = fv[q] - fe[q];
    real cx = (real)L->cx[q], cy = (real)L->cy[q], cz = (real)L->cz[q];
    pxx += cx * cx * n;
    pyy += cy * cy * n;
    pzz += cz * cz * n;
    pxy += cx * cy * n;
    pyz += cy * cz * n;
    pxz += cx * cz * n;
  }
  const real third = R(1) / R(3);
  for (int q = 0; q < L->Q; ++q) {
    real cx = (real)L->cx[q], cy = (real)L->cy[q], cz = (real)L->cz[q];
    real t = ((cx * cx - third) * pxx + (cy * cy - third) * pyy + (cz * cz - third) * pzz) +
             R(2) * ((cx * cy) * pxy + (cy * cz) * pyz + (cx * cz) * pxz);
    f[q * total + idx] = fe[q] + R(4.5) * L->w[q] * t;
  }
}

static real inlet_ux(const oracle_grid* g, int y, size_t node, uint64_t step) {
  real u = g->u_in_x;
  if (g->inflow_parabolic) {
    real s = (real)y - R(0.5);
    real H = (real)(g->ny - 2);
    u = R(4) * g->u_max * s * (H - s) / (H * H);
  }
  if (g->turb_amp != R(0)) {
    double r = oracle_rng_u01(g->turb_seed, step * (uint64_t)g->N + (uint64_t)node);
    u = u * (R(1) + g->turb_amp * (real)(2.0 * r - 1.0));
  }
  return u;
}

static void apply_open_boundaries(oracle_grid* g, uint64_t step) {
  int y0, y1, z0, z1;
  bc_range(g, &y0, &y1, &z0, &z1);
  lat_t L = lattice_of(g->lattice);
  size_t N = g->N;
  if (g->has_inflow) {
    if (g->dim == 2 && !g->inflow_parabolic && g->turb_amp == R(0) && !(g->ext_axes && g->per_y)) {
      /* reference path: lattice_grid_2d.cpp:244-249 */
      oracle_d2q9_apply_zou_he_inflow_west(g->f_src, g->rho, g->ux, g->uy, g->u_in_x, g->u_in_y, g->nx, g->ny);
    } else {
      for (int z = z0; z < z1; ++z)
        for (int y = y0; y < y1; ++y) {
          size_t idx = ((size_t)z * g->ny + y) * g->nx;
          real ux = inlet_ux(g, y, idx, step);
          if (g->dim == 2) {
            real r;
            d2q9_zh_inflow_node(g->f_src, N, idx, ux, g->u_in_y, &r);
          } else {
            zh3d_velocity_node(&L, g->f_src, N, idx, ux, g->u_in_y, R(0));
          }
        }
    }
  }
  if (g->has_outflow) {
    if (g->dim == 2 && !(g->ext_axes && g->per_y)) {
      oracle_d2q9_apply_zou_he_outflow_east(g->f_src, g->rho, g->ux, g->uy, g->rho_out, g->nx, g->ny);
    }
    /* 3D pressure outlet: not provided (the reference has none); the
     * convective outlet is the 3D open outflow. */
  }
  if (g->has_conv) {
    real alpha = g->u_conv / (R(1) + g->u_conv);
    real beta = R(1) / (R(1) + g->u_conv);
    for (int z = z0; z < z1; ++z)
      for (int y = y0; y < y1; ++y) {
        size_t io = ((size_t)z * g->ny + y) * g->nx + (g->nx - 1);
        size_t ii = io - 1;
        for (int q = 0; q < g->Q; ++q) {
          real v = beta * g->f_prev[q * N + io] + alpha * g->f_src[q * N + ii];
          g->f_src[q * N + io] = v;
          g->f_prev[q * N + io] = v;
        }
      }
  }
}

/* lattice_grid_2d.cpp:195-265 and lattice_grid_3d.cpp:301-380.  pinned:
 * test_oracle_poiseuille_reference_validation and
 * test_oracle_taylor_green_3d_reference_validation (the reference's own
 * validation tests), test_oracle_taylor_green_2d_periodic_decay,
 * test_oracle_streaming_walls_conserve_mass_in_closed_box,
 * test_oracle_fp32_fp64_agree_on_sphere_flow. */
/* ------------------------------------------------------------------------
 * EXTENSION (no reference counterpart; the reference's ThermalSolver2D,
 * src/thermal/thermal_solver.cpp:83-123, is a D2Q9 passive scalar whose
 * buoyancy reaches the flow only through the mean temperature).  Parity
 * pinned by test_thermal_* (conduction fixed point, adiabatic energy
 * conservation, de Vahl Davis cavity Nusselt number).
 *
 * D2Q5 temperature g_q (c = 0, +x, +y, -x, -y; w = 1 - 4/6, 1/6 x 4), BGK with
 * geq_q = w_q T (1 + 3 c_q . u), alpha = (tau_g - 1/2) / 3, coupled to the
 * D2Q9 flow by a per-node Boussinesq acceleration a_y = gbeta (T - t_ref)
 * added to the body force of the flow collision; u in geq is the Guo
 * velocity (j + rho a / 2) / rho of the post-stream state.  A pull source
 * that is solid (or outside a non-periodic y range) reflects g: adiabatic
 * bounce-back g_q = g*_opp, or, where the solid node carries a wall
 * temperature T_w, anti-bounce-back g_q = -g*_opp + 2 w_q T_w. */
static const int c5x[5] = {0, 1, 0, -1, 0};
static const int c5y[5] = {0, 0, 1, 0, -1};
static const int opp5[5] = {0, 3, 4, 1, 2};
/* w_0 = 1 - 4 w_1 in the working precision, so the weights sum to one and
 * the equilibrium carries T without a per-step drift */
static const real w5[5] = {R(1) - R(4) * (R(1) / R(6)), R(1) / R(6), R(1) / R(6), R(1) / R(6), R(1) / R(6)};

void oracle_grid_enable_thermal(oracle_grid* g, real tau_g, real gbeta, real t_ref, real t_init) {
  if (g->dim != 2) return;
  if (!g->thermal) {
    g->g_src = (real*)xcalloc(5 * g->N, sizeof(real));
    g->g_dst = (real*)xcalloc(5 * g->N, sizeof(real));
    g->t_wall = (real*)xcalloc(g->N, sizeof(real));
    g->temp = (real*)xcalloc(g->N, sizeof(real));
    for (size_t i = 0; i < g->N; ++i) g->t_wall[i] = (real)NAN;
  }
  g->thermal = 1;
  g->tau_g = tau_g;
  g->gbeta = gbeta;
  g->t_ref = t_ref;
  for (size_t i = 0; i < g->N; ++i) {
    g->temp[i] = t_init;
    for (int q = 0; q < 5; ++q) g->g_src[q * g->N + i] = g->g_dst[q * g->N + i] = w5[q] * t_init;
  }
}

/* Temperature field as equilibrium populations at rest (every node). */
void oracle_grid_set_temperature_field(oracle_grid* g, const real* t) {
  if (!g->thermal) return;
  for (size_t i = 0; i < g->N; ++i) {
    g->temp[i] = t[i];
    for (int q = 0; q < 5; ++q) g->g_src[q * g->N + i] = g->g_dst[q * g->N + i] = w5[q] * t[i];
  }
}

void oracle_grid_set_wall_temperature(oracle_grid* g, int x0, int y0, int x1, int y1, real t_wall) {
  if (!g->thermal) return;
  for (int y = imax(0, y0); y <= imin(g->ny - 1, y1); ++y)
    for (int x = imax(0, x0); x <= imin(g->nx - 1, x1); ++x) g->t_wall[(size_t)y * g->nx + x] = t_wall;
}

const real* oracle_grid_temperature(const oracle_grid* g) { return g->temp; }

static void thermal_step_2d(oracle_grid* g, int collision, real gx, real gy, uint64_t step) {
  const size_t n = g->N;
  const int nx = g->nx, ny = g->ny, count = (int)n;
  const int per_y = g->ext_axes && g->per_y;
  const int periodic = g->is_periodic && !g->has_moving_wall_top;
  const real omega_g = R(1) / g->tau_g;
  /* 1. collision of f (per-node Boussinesq force) and g */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    if (g->cell[i] == CELL_SOLID) {
      for (int q = 0; q < 9; ++q) g->f_dst[q * n + i] = g->f_src[q * n + i];
      for (int q = 0; q < 5; ++q) g->g_dst[q * n + i] = g->g_src[q * n + i];
      continue;
    }
    real T = R(0);
    for (int q = 0; q < 5; ++q) T += g->g_src[q * n + i];
    const real gyn = gy + g->gbeta * (T - g->t_ref);
    real rho = R(0), jx = R(0), jy = R(0);
    for (int q = 0; q < 9; ++q) {
      const real v = g->f_src[q * n + i];
      rho += v;
      jx += (real)c9x[q] * v;
      jy += (real)c9y[q] * v;
    }
    const real inv_rho = R(1) / rho;
    const real ux = (jx + R(0.5) * rho * gx) * inv_rho;
    const real uy = (jy + R(0.5) * rho * gyn) * inv_rho;
    if (collision == COLL_BGK)
      d2q9_bgk_node(g->f_src, g->f_dst, n, (size_t)i, NULL, NULL, NULL, g->tau, gx, gyn);
    else if (collision == COLL_TRT)
      d2q9_trt_node(g->f_src, g->f_dst, n, (size_t)i, NULL, NULL, NULL, g->tau, trt_tau_a(g->tau), gx, gyn);
    else
      d2q9_mrt_node(g->f_src, g->f_dst, n, (size_t)i, NULL, NULL, NULL, g->tau, R(1.63), gx, gyn);
    for (int q = 0; q < 5; ++q) {
      const real cu = (real)c5x[q] * ux + (real)c5y[q] * uy;
      const real geq = w5[q] * T * (R(1) + R(3) * cu);
      const real v = g->g_src[q * n + i];
      g->g_dst[q * n + i] = v - omega_g * (v - geq);
    }
  }
  /* 2. streaming of f, open boundaries */
  if (periodic)
    oracle_d2q9_stream_periodic(g->f_dst, g->f_src, nx, ny);
  else
    oracle_d2q9_stream_general(g->f_dst, g->f_src, g->cell, nx, ny, per_y, g->has_moving_wall_top ? g->uwx : R(0),
                               R(0));
  apply_open_boundaries(g, step);
  /* 3. streaming of g with (anti-)bounce-back */
#pragma omp parallel for schedule(static)
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const size_t i = (size_t)y * nx + x;
      if (!periodic && g->cell[i] == CELL_SOLID) {
        for (int q = 0; q < 5; ++q) g->g_src[q * n + i] = g->g_dst[q * n + i];
        continue;
      }
      for (int q = 0; q < 5; ++q) {
        const int sx = (x - c5x[q] + nx) % nx;
        int sy = y - c5y[q];
        int in = 1;
        if (periodic || per_y) sy = (sy + ny) % ny;
        else in = sy >= 0 && sy < ny;
        const size_t s = in ? (size_t)sy * nx + sx : i;
        const uint8_t t = (in && !periodic) ? g->cell[s] : CELL_FLUID;
        if (!in || t == CELL_SOLID || t == CELL_MOVING_WALL) {
          const real tw = in ? g->t_wall[s] : (real)NAN;
          const real v = g->g_dst[opp5[q] * n + i];
          g->g_src[q * n + i] = isnan(tw) ? v : -v + R(2) * w5[q] * tw;
        } else {
          g->g_src[q * n + i] = g->g_dst[q * n + s];
        }
      }
    }
  /* 4. temperature and moments with the per-node force */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < count; ++i) {
    real T = R(0);
    for (int q = 0; q < 5; ++q) T += g->g_src[q * n + i];
    g->temp[i] = T;
    const real gyn = gy + g->gbeta * (T - g->t_ref);
    real rho = R(0), jx = R(0), jy = R(0);
    for (int q = 0; q < 9; ++q) {
      const real v = g->f_src[q * n + i];
      rho += v;
      jx += (real)c9x[q] * v;
      jy += (real)c9y[q] * v;
    }
    const real inv_rho = R(1) / rho;
    g->rho[i] = rho;
    g->ux[i] = (jx + R(0.5) * rho * gx) * inv_rho;
    g->uy[i] = (jy + R(0.5) * rho * gyn) * inv_rho;
    g->speed[i] = RSQRT(g->ux[i] * g->ux[i] + g->uy[i] * g->uy[i]);
  }
  g->step_count++;
}

void oracle_grid_step(oracle_grid* g, int collision, real gx, real gy, real gz, int use_les, real smag) {
  if (g->thermal && g->dim == 2) {
    thermal_step_2d(g, collision, gx, gy, g->step_count + 1);
    return;
  }
  int count = (int)g->N;
  real et = g->tau;
  uint64_t step = g->step_count + 1;
  if (g->dim == 2) gz = R(0);
  /* 1. LES (reference: computed, result unused; EXTENSION coupled mode feeds
   *    tau_eff into the collision). */
  if (use_les) {
    if (g->les_coupled) {
      oracle_smagorinsky_tau_from_f(g->lattice, g->f_src, g->tau_eff, g->tau, smag, gx, gy, gz, count);
    } else if (g->lattice == 9) {
      oracle_d2q9_compute_smagorinsky_tau(g->f_src, g->rho, g->ux, g->uy, g->tau_eff, g->tau, smag, count);
    } else if (g->lattice == 19) {
      oracle_d3q19_compute_smagorinsky_tau(g->f_src, g->rho, g->ux, g->uy, g->uz, g->tau_eff, g->tau, smag,
                                           count);
    }
  }
  /* 2. collision f_src -> f_dst */
  if (use_les && g->les_coupled) {
    int coll = collision;
    if (g->lattice == 27 && coll == COLL_MRT) coll = COLL_TRT;
    oracle_collide_tau_node(g->lattice, coll, g->f_src, g->f_dst, g->cell, g->tau_eff, gx, gy, gz, count);
  } else if (g->lattice == 9) {
    if (collision == COLL_BGK)
      oracle_d2q9_collide_bgk(g->f_src, g->f_dst, g->cell, g->rho, g->ux, g->uy, et, gx, gy, count);
    else if (collision == COLL_TRT)
      oracle_d2q9_collide_trt(g->f_src, g->f_dst, g->cell, g->rho, g->ux, g->uy, et, trt_tau_a(et), gx, gy, count);
    else
      oracle_d2q9_collide_mrt(g->f_src, g->f_dst, g->cell, g->rho, g->ux, g->uy, et, R(1.63), gx, gy, count);
  } else if (g->lattice == 19) {
    if (collision == COLL_BGK)
      oracle_d3q19_collide_bgk(g->f_src, g->f_dst, g->cell, g->rho, g->ux, g->uy, g->uz, et, gx, gy, gz, count);
    else if (collision == COLL_TRT)
      oracle_d3q19_collide_trt(g->f_src, g->f_dst, g->cell, g->rho, g->ux, g->uy, g->uz, et, trt_tau_a(et), gx,
                               gy, gz, count);
    else
      oracle_d3q19_collide_mrt(g->f_src, g->f_dst, g->cell, g->rho, g->ux, g->uy, g->uz, et, R(1.19), gx, gy,
                               gz, count);
  } else {
    /* The reference's D3Q27 kernel is TRT only (d3q27_collision_mrt.ispc:5);
     * the grid maps MRT to TRT and keeps BGK as BGK. */
    real ta = trt_tau_a(et);
    lat_t L = lattice_of(27);
    size_t n = g->N;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < count; ++i) {
      if (g->cell[i] == CELL_SOLID) {
        for (int q = 0; q < 27; ++q) g->f_dst[q * n + i] = g->f_src[q * n + i];
        continue;
      }
      if (collision == COLL_BGK)
        bgk3d_node(&L, g->f_src, g->f_dst, n, (size_t)i, g->rho, g->ux, g->uy, g->uz, et, gx, gy, gz);
      else
        trt3d_node(&L, g->f_src, g->f_dst, n, (size_t)i, g->rho, g->ux, g->uy, g->uz, et, ta, gx, gy, gz);
    }
  }
  /* 3. streaming f_dst -> f_src */
  if (g->dim == 2) {
    if (g->is_periodic && !g->has_moving_wall_top)
      oracle_d2q9_stream_periodic(g->f_dst, g->f_src, g->nx, g->ny);
    else
      oracle_d2q9_stream_general(g->f_dst, g->f_src, g->cell, g->nx, g->ny, g->ext_axes && g->per_y,
                                 g->has_moving_wall_top ? g->uwx : R(0), R(0));
  } else {
    lat_t L = lattice_of(g->lattice);
    if (g->is_periodic && !g->has_moving_wall_top) {
      stream3d_periodic(&L, g->f_dst, g->f_src, g->nx, g->ny, g->nz);
    } else {
      oracle_stream3d_general(g->lattice, g->f_dst, g->f_src, g->cell, g->nx, g->ny, g->nz,
                              g->ext_axes && g->per_y, g->ext_axes && g->per_z,
                              g->has_moving_wall_top ? g->uwx : R(0), R(0),
                              g->has_moving_wall_top ? g->uwz : R(0));
    }
    /* 4. BFL from the post-collision buffer into the streamed one */
    if (g->nlinks > 0) {
      real* zero = g->l_uw;
      bfl_general(&L, g->f_src, g->f_dst, g->l_idx, g->l_q, g->l_delta, zero, zero + g->nlinks,
                  zero + 2 * g->nlinks, g->nx, g->ny, g->nz, g->ext_axes && g->per_y, g->ext_axes && g->per_z,
                  g->nlinks);
    }
  }
  /* open boundaries (2D reference Zou-He; 3D and convective EXTENSIONS) */
  apply_open_boundaries(g, step);
  /* 5. moments with the Guo shift, then |u| */
  if (g->dim == 2) {
    oracle_d2q9_compute_macroscopic_moments_forced(g->f_src, g->rho, g->ux, g->uy, gx, gy, count);
    oracle_d2q9_compute_velocity_magnitude(g->ux, g->uy, g->speed, count);
  } else {
    lat_t L = lattice_of(g->lattice);
    moments3d(&L, g->f_src, g->rho, g->ux, g->uy, g->uz, gx, gy, gz, 1, count);
    oracle_d3q19_compute_velocity_magnitude(g->ux, g->uy, g->uz, g->speed, count);
  }
  g->step_count++;
}

void oracle_grid_run(oracle_grid* g, int steps, int collision, real gx, real gy, real gz, int use_les,
                     real smag) {
  for (int s = 0; s < steps; ++s) oracle_grid_step(g, collision, gx, gy, gz, use_les, smag);
}

const real* oracle_grid_f_src(const oracle_grid* g) { return g->f_src; }
const real* oracle_grid_f_dst(const oracle_grid* g) { return g->f_dst; }
const real* oracle_grid_rho(const oracle_grid* g) { return g->rho; }
const real* oracle_grid_ux(const oracle_grid* g) { return g->ux; }
const real* oracle_grid_uy(const oracle_grid* g) { return g->uy; }
const real* oracle_grid_uz(const oracle_grid* g) { return g->uz; }
const real* oracle_grid_speed(const oracle_grid* g) { return g->speed; }
const uint8_t* oracle_grid_cell_type(const oracle_grid* g) { return g->cell; }

/* lattice_grid_3d.cpp:382-410 (Fluid nodes only), accumulated in double.
 * pinned: test_oracle_grid_setters_and_diagnostics_closed_forms (uniform
 * state with a solid block), test_oracle_taylor_green_3d_reference_validation. */
double oracle_grid_total_mass(const oracle_grid* g) {
  double t = 0;
  for (size_t i = 0; i < g->N; ++i)
    if (g->cell[i] == CELL_FLUID) t += (double)g->rho[i];
  return t;
}
double oracle_grid_total_kinetic_energy(const oracle_grid* g) {
  double t = 0;
  for (size_t i = 0; i < g->N; ++i)
    if (g->cell[i] == CELL_FLUID) t += (double)(R(0.5) * g->rho[i] * g->speed[i] * g->speed[i]);
  return t;
}
double oracle_grid_max_speed(const oracle_grid* g) {
  real m = R(0);
  for (size_t i = 0; i < g->N; ++i)
    if (g->cell[i] == CELL_FLUID && g->speed[i] > m) m = g->speed[i];
  return (double)m;
}

/* Momentum exchange on the post-stream populations, as the scenarios call
 * it (sphere_3d_wake.cpp:98-105, karman_vortex.cpp:107-114). */
void oracle_grid_momentum_exchange(const oracle_grid* g, real cx, real cy, real cz, real r_max, double* out3) {
  if (g->dim == 2) {
    d2q9_mem_d(g->f_src, g->cell, g->nx, g->ny, cx, cy, r_max, out3);
  } else {
    lat_t L = lattice_of(g->lattice);
    mem3d(&L, g->f_src, g->cell, g->nx, g->ny, g->nz, g->ext_axes && g->per_y, g->ext_axes && g->per_z, cx, cy,
          cz, r_max, out3);
  }
}

/* analytics_kernels.ispc:4-29: omega_z = d uy/dx - d ux/dy, central
 * differences, periodic wrap.  pinned:
 * test_vorticity_and_q_criterion_closed_forms (this and the 3D kernel). */
void oracle_compute_vorticity_2d(const real* ux, const real* uy, real* w, int nx, int ny) {
  for (int y = 0; y < ny; ++y) {
    int ym = (y - 1 + ny) % ny, yp = (y + 1) % ny;
    for (int x = 0; x < nx; ++x) {
      int xm = (x - 1 + nx) % nx, xp = (x + 1) % nx;
      real duy_dx = R(0.5) * (uy[y * nx + xp] - uy[y * nx + xm]);
      real dux_dy = R(0.5) * (ux[yp * nx + x] - ux[ym * nx + x]);
      w[y * nx + x] = duy_dx - dux_dy;
    }
  }
}

/* analytics_kernels.ispc:33-110: curl and Q = (||Omega||^2 - ||S||^2) / 2. */
void oracle_compute_vorticity_and_q_criterion_3d(const real* ux, const real* uy, const real* uz, real* ox, real* oy,
                                                 real* oz, real* qc, int nx, int ny, int nz) {
  for (int z = 0; z < nz; ++z) {
    int zm = (z - 1 + nz) % nz, zp = (z + 1) % nz;
    for (int y = 0; y < ny; ++y) {
      int ym = (y - 1 + ny) % ny, yp = (y + 1) % ny;
      size_t c = ((size_t)z * ny + y) * nx, cyp = ((size_t)z * ny + yp) * nx, cym = ((size_t)z * ny + ym) * nx;
      size_t czp = ((size_t)zp * ny + y) * nx, czm = ((size_t)zm * ny + y) * nx;
      for (int x = 0; x < nx; ++x) {
        int xm = (x - 1 + nx) % nx, xp = (x + 1) % nx;
        real dux_dx = R(0.5) * (ux[c + xp] - ux[c + xm]);
        real dux_dy = R(0.5) * (ux[cyp + x] - ux[cym + x]);
        real dux_dz = R(0.5) * (ux[czp + x] - ux[czm + x]);
        real duy_dx = R(0.5) * (uy[c + xp] - uy[c + xm]);
        real duy_dy = R(0.5) * (uy[cyp + x] - uy[cym + x]);
        real duy_dz = R(0.5) * (uy[czp + x] - uy[czm + x]);
        real duz_dx = R(0.5) * (uz[c + xp] - uz[c + xm]);
        real duz_dy = R(0.5) * (uz[cyp + x] - uz[cym + x]);
        real duz_dz = R(0.5) * (uz[czp + x] - uz[czm + x]);
        ox[c + x] = duz_dy - duy_dz;
        oy[c + x] = dux_dz - duz_dx;
        oz[c + x] = duy_dx - dux_dy;
        real s_xy = R(0.5) * (dux_dy + duy_dx), s_yz = R(0.5) * (duy_dz + duz_dy), s_xz = R(0.5) * (dux_dz + duz_dx);
        real norm_s = (dux_dx * dux_dx + duy_dy * duy_dy + duz_dz * duz_dz) +
                      R(2) * (s_xy * s_xy + s_yz * s_yz + s_xz * s_xz);
        real o_xy = R(0.5) * (dux_dy - duy_dx), o_yz = R(0.5) * (duy_dz - duz_dy), o_xz = R(0.5) * (dux_dz - duz_dx);
        real norm_o = R(2) * (o_xy * o_xy + o_yz * o_yz + o_xz * o_xz);
        qc[c + x] = R(0.5) * (norm_o - norm_s);
      }
    }
  }
}

/* ------------------------------------------------------------------------
 * Triangle-mesh voxelizer.  src/geometry/obj_loader.cpp:208-400.
 * ---------------------------------------------------------------------- */

static real rmin3(real a, real b, real c) {
  real m = a;
  if (b < m) m = b;
  if (c < m) m = c;
  return m;
}
static real rmax3(real a, real b, real c) {
  real m = a;
  if (b > m) m = b;
  if (c > m) m = c;
  return m;
}

/* obj_loader.cpp:262-271: projections a*p - b*q of the three vertices on one
 * candidate axis against the box radius fa*h1 + fb*h2. */
static int sat_axis(real a, real b, real fa, real fb, real p0, real q0, real p1, real q1, real p2, real q2, real h1,
                    real h2) {
  real s0 = a * p0 - b * q0;
  real s1 = a * p1 - b * q1;
  real s2 = a * p2 - b * q2;
  real rad = fa * h1 + fb * h2;
  return rmin3(s0, s1, s2) <= rad && rmax3(s0, s1, s2) >= -rad;
}

/* obj_loader.cpp:209-306 (test_triangle_box_overlap).  The nine edge axes
 * are called with the reference's arguments (:277-301).  pinned:
 * test_overlap_against_exact_clipping (never misses an overlap; the only
 * extra overlaps are the ones those arguments predict). */
int oracle_tri_box_overlap(const real* tri, const real* c, const real* h) {
  real v0x = tri[0] - c[0], v0y = tri[1] - c[1], v0z = tri[2] - c[2];
  real v1x = tri[3] - c[0], v1y = tri[4] - c[1], v1z = tri[5] - c[2];
  real v2x = tri[6] - c[0], v2y = tri[7] - c[1], v2z = tri[8] - c[2];
  real e[3][3] = {{v1x - v0x, v1y - v0y, v1z - v0z}, {v2x - v1x, v2y - v1y, v2z - v1z}, {v0x - v2x, v0y - v2y, v0z - v2z}};
  if (rmin3(v0x, v1x, v2x) > h[0] || rmax3(v0x, v1x, v2x) < -h[0]) return 0;
  if (rmin3(v0y, v1y, v2y) > h[1] || rmax3(v0y, v1y, v2y) < -h[1]) return 0;
  if (rmin3(v0z, v1z, v2z) > h[2] || rmax3(v0z, v1z, v2z) < -h[2]) return 0;
  /* plane of the triangle: normal e0 x e1 */
  real nx = e[0][1] * e[1][2] - e[0][2] * e[1][1];
  real ny = e[0][2] * e[1][0] - e[0][0] * e[1][2];
  real nz = e[0][0] * e[1][1] - e[0][1] * e[1][0];
  real d = nx * v0x + ny * v0y + nz * v0z;
  real mnx = nx > R(0) ? -h[0] : h[0], mxx = nx > R(0) ? h[0] : -h[0];
  real mny = ny > R(0) ? -h[1] : h[1], mxy = ny > R(0) ? h[1] : -h[1];
  real mnz = nz > R(0) ? -h[2] : h[2], mxz = nz > R(0) ? h[2] : -h[2];
  if (nx * mnx + ny * mny + nz * mnz > d || nx * mxx + ny * mxy + nz * mxz < d) return 0;
  for (int k = 0; k < 3; ++k) {
    real ex = e[k][0], ey = e[k][1], ez = e[k][2];
    real fx = RABS(ex), fy = RABS(ey), fz = RABS(ez);
    if (!sat_axis(ez, ey, fz, fy, v0y, v0z, v1y, v1z, v2y, v2z, h[1], h[2])) return 0;
    if (!sat_axis(-ez, ex, fz, fx, v0x, v0z, v1x, v1z, v2x, v2z, h[0], h[2])) return 0;
    if (!sat_axis(ey, -ex, fy, fx, v0x, v0y, v1x, v1y, v2x, v2y, h[0], h[1])) return 0;
  }
  return 1;
}

/* obj_loader.cpp:308-400 (voxelize_into_grid): surface cells by the overlap
 * test over each triangle's floor/ceil cell box, then a breadth-first walk
 * from the non-Solid cells of the six outer faces; unreached cells become
 * Solid.  Returns how many Solid writes were made (set_cell_type clears
 * the periodic mode on any of them).  pinned: tests/test_geometry_oracle.py
 * (the reference's OBJ cube test and its closed form, closed vs open shells,
 * sphere between its inscribed and circumscribed balls, torus hole). */
long long oracle_voxelize_mesh(uint8_t* flags, int nx, int ny, int nz, const real* tris, long long ntri,
                               int fill_interior) {
  const real half[3] = {R(0.5), R(0.5), R(0.5)};
  long long writes = 0;
  for (long long t = 0; t < ntri; ++t) {
    const real* v = tris + 9 * t;
    int i_min = imax(0, (int)RFLOOR(rmin3(v[0], v[3], v[6]))), i_max = imin(nx - 1, (int)RCEIL(rmax3(v[0], v[3], v[6])));
    int j_min = imax(0, (int)RFLOOR(rmin3(v[1], v[4], v[7]))), j_max = imin(ny - 1, (int)RCEIL(rmax3(v[1], v[4], v[7])));
    int k_min = imax(0, (int)RFLOOR(rmin3(v[2], v[5], v[8]))), k_max = imin(nz - 1, (int)RCEIL(rmax3(v[2], v[5], v[8])));
    for (int k = k_min; k <= k_max; ++k)
      for (int j = j_min; j <= j_max; ++j)
        for (int i = i_min; i <= i_max; ++i) {
          real c[3] = {(real)i + R(0.5), (real)j + R(0.5), (real)k + R(0.5)};
          if (oracle_tri_box_overlap(v, c, half)) {
            flags[((size_t)k * ny + j) * nx + i] = CELL_SOLID;
            ++writes;
          }
        }
  }
  if (!fill_interior) return writes;
  size_t n = (size_t)nx * ny * nz;
  uint8_t* seen = (uint8_t*)calloc(n ? n : 1, 1);
  size_t* queue = (size_t*)malloc((n ? n : 1) * sizeof(size_t));
  size_t head = 0, tail = 0;
#define PUSH(x, y, z)                                            \
  do {                                                           \
    size_t id_ = ((size_t)(z) * ny + (y)) * nx + (x);            \
    if (!seen[id_] && flags[id_] != CELL_SOLID) {                \
      seen[id_] = 1;                                             \
      queue[tail++] = id_;                                       \
    }                                                            \
  } while (0)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      PUSH(0, j, k);
      PUSH(nx - 1, j, k);
    }
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < nx; ++i) {
      PUSH(i, 0, k);
      PUSH(i, ny - 1, k);
    }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      PUSH(i, j, 0);
      PUSH(i, j, nz - 1);
    }
  while (head < tail) {
    size_t cur = queue[head++];
    int x = (int)(cur % (size_t)nx), y = (int)((cur / (size_t)nx) % (size_t)ny), z = (int)(cur / ((size_t)nx * ny));
    if (x + 1 < nx) PUSH(x + 1, y, z);
    if (x > 0) PUSH(x - 1, y, z);
    if (y + 1 < ny) PUSH(x, y + 1, z);
    if (y > 0) PUSH(x, y - 1, z);
    if (z + 1 < nz) PUSH(x, y, z + 1);
    if (z > 0) PUSH(x, y, z - 1);
  }
#undef PUSH
  for (size_t i = 0; i < n; ++i)
    if (!seen[i]) {
      flags[i] = CELL_SOLID;
      ++writes;
    }
  free(queue);
  free(seen);
  return writes;
}

/* ------------------------------------------------------------------------
 * pinned: tests/test_thermal_ibm_oracle.py (temperature conservation and the
 * equilibrium fixed point of the collision, the in-place wall sequence in
 * closed form, the buoyancy term, the interpolation's partition of unity and
 * x periodicity, the spread total in closed form).
 * The reference's D2Q9 temperature solver kernels and the IBM kernels.
 * ---------------------------------------------------------------------- */

/* thermal_kernels.ispc:4-44 */
void oracle_thermal_collide_d2q9(const real* g_in, real* g_out, const uint8_t* cell_type, const real* ux,
                                 const real* uy, real* temperature, real tau_T, int count) {
  real omega_T = R(1) / tau_T;
  for (int i = 0; i < count; ++i) {
    if (cell_type[i] == CELL_SOLID) {
      for (int q = 0; q < 9; ++q) g_out[(size_t)q * count + i] = g_in[(size_t)q * count + i];
      continue;
    }
    real t = R(0), g[9];
    for (int q = 0; q < 9; ++q) {
      g[q] = g_in[(size_t)q * count + i];
      t += g[q];
    }
    temperature[i] = t;
    for (int q = 0; q < 9; ++q) {
      real cu = (real)c9x[q] * ux[i] + (real)c9y[q] * uy[i];
      real geq = w9[q] * t * (R(1) + R(3) * cu);
      g_out[(size_t)q * count + i] = g[q] - omega_T * (g[q] - geq);
    }
  }
}

/* thermal_kernels.ispc:48-60 */
void oracle_thermal_compute_buoyancy_force_2d(const real* temperature, real* gy_body, real t_ref, real beta,
                                              real g_gravity, int count) {
  for (int i = 0; i < count; ++i) gy_body[i] += beta * (temperature[i] - t_ref) * g_gravity;
}

/* thermal_kernels.ispc:63-81: in place, q = 0..8 in turn. */
void oracle_thermal_apply_isothermal_wall(real* g, int nx, int ny, int y_wall, real t_wall) {
  size_t n = (size_t)nx * ny;
  for (int x = 0; x < nx; ++x) {
    size_t idx = (size_t)y_wall * nx + x;
    for (int q = 0; q < 9; ++q) {
      real v = g[q * n + idx];
      g[opp9[q] * n + idx] = -v + R(2) * w9[q] * t_wall;
    }
  }
}

/* ibm_kernels.ispc:4-8 */
static real peskin_1d(real r) {
  real a = RABS(r);
  if (a >= R(2)) return R(0);
  return R(0.25) * (R(1) + RCOS(R(0.5) * R(3.14159265358979323846) * a));
}

/* ibm_kernels.ispc:12-50 */
void oracle_ibm_interpolate_velocity_2d(const real* ux, const real* uy, const real* marker_x, const real* marker_y,
                                        real* marker_ux, real* marker_uy, int nx, int ny, int num_markers) {
  for (int m = 0; m < num_markers; ++m) {
    real mx = marker_x[m], my = marker_y[m], su = R(0), sv = R(0);
    for (int iy = -2; iy <= 2; ++iy) {
      int py = (int)RFLOOR(my) + iy;
      if (py < 0 || py >= ny) continue;
      real wy = peskin_1d((real)py - my);
      for (int ix = -2; ix <= 2; ++ix) {
        int px = (int)RFLOOR(mx) + ix;
        int pw = (px % nx + nx) % nx;
        real w = peskin_1d((real)px - mx) * wy;
        su += w * ux[(size_t)py * nx + pw];
        sv += w * uy[(size_t)py * nx + pw];
      }
    }
    marker_ux[m] = su;
    marker_uy[m] = sv;
  }
}

/* ibm_kernels.ispc:54-96: the row weight has no |r| < 2 cutoff (:80). */
void oracle_ibm_spread_force_2d(const real* marker_x, const real* marker_y, const real* marker_fx,
                                const real* marker_fy, const real* marker_ds, real* euler_gx, real* euler_gy, int nx,
                                int ny, int num_markers) {
  for (int m = 0; m < num_markers; ++m) {
    real mx = marker_x[m], my = marker_y[m];
    real fx = marker_fx[m] * marker_ds[m], fy = marker_fy[m] * marker_ds[m];
    int min_x = (int)RFLOOR(mx) - 2, max_x = (int)RFLOOR(mx) + 2;
    int min_y = (int)RFLOOR(my) - 2, max_y = (int)RFLOOR(my) + 2;
    for (int py = min_y; py <= max_y; ++py) {
      if (py < 0 || py >= ny) continue;
      real wy = R(0.25) * (R(1) + RCOS(R(0.5) * R(3.14159265358979323846) * RABS((real)py - my)));
      for (int px = min_x; px <= max_x; ++px) {
        size_t idx = (size_t)py * nx + (size_t)((px % nx + nx) % nx);
        real w = peskin_1d((real)px - mx) * wy;
        euler_gx[idx] += w * fx;
        euler_gy[idx] += w * fy;
      }
    }
  }
}
