// lattice.hpp -- compile-time lattice tables for the HIP kernels.
//
// Velocity sets, weights and the population-plane order follow
// include/wake/core/lattice_models.hpp:14-110 of the reference (the plane index
// q is part of the ABI: population q of node i lives at q * stride + i).
// D3Q27 keeps two opposite tables: `opp` is the geometric one (c_opp = -c),
// used by every grid-level path; `opp_ref` is the reference's table
// (lattice_models.hpp:104-109), whose corner pairs are mirrored in x and y
// only, and is used solely by the d3q27_collide export that restates it.
//
// The D3Q19 MRT inverse is the reference's d3q19_Minv
// (src/ispc/d3q19_collision_mrt.ispc:4-24), i.e. inv(M) rounded to 7 decimals.
// It is regenerated here at compile time from the d'Humieres basis M by
// Gauss-Jordan elimination in double precision and 7-decimal rounding, so the
// kernels see every entry (and every zero) as a literal.
#pragma once

#include <cstdint>

namespace wake_hip {

enum Cell : uint8_t { kFluid = 0, kSolid = 1, kInflow = 2, kOutflow = 3, kPeriodic = 4, kSymmetry = 5,
                      kMovingWall = 6 };
enum Collision : int { kBGK = 0, kTRT = 1, kMRT = 2 };

// Floats per (y, z) row of the convective outlet store: Q rounded up to a
// multiple of four, so each record is 16-byte aligned.
constexpr int prev_stride(int q) { return (q + 3) / 4 * 4; }

struct D2Q9 {
  static constexpr int Q = 9;
  static constexpr int D = 2;
  static constexpr int cx[Q] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
  static constexpr int cy[Q] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
  static constexpr int cz[Q] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int opp[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
  static constexpr int opp_ref[Q] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
  static constexpr float w[Q] = {4.0f / 9.0f,  1.0f / 9.0f,  1.0f / 9.0f,  1.0f / 9.0f, 1.0f / 9.0f,
                                 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f};
};

// D2Q5 temperature lattice (no reference counterpart): c = 0, +x, +y, -x, -y.
// w_0 = 1 - 4 w_1 in FP32, so the weights sum to exactly one.
struct D2Q5 {
  static constexpr int Q = 5;
  static constexpr int cx[Q] = {0, 1, 0, -1, 0};
  static constexpr int cy[Q] = {0, 0, 1, 0, -1};
  static constexpr int opp[Q] = {0, 3, 4, 1, 2};
  static constexpr float w[Q] = {1.0f - 4.0f * (1.0f / 6.0f), 1.0f / 6.0f, 1.0f / 6.0f, 1.0f / 6.0f, 1.0f / 6.0f};
};

struct D3Q19 {
  static constexpr int Q = 19;
  static constexpr int D = 3;
  static constexpr int cx[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  static constexpr int cy[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
  static constexpr int cz[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
  static constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
  static constexpr int opp_ref[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
  static constexpr float w[Q] = {1.0f / 3.0f,  1.0f / 18.0f, 1.0f / 18.0f, 1.0f / 18.0f, 1.0f / 18.0f,
                                 1.0f / 18.0f, 1.0f / 18.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f,
                                 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f,
                                 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f};
};

struct D3Q27 {
  static constexpr int Q = 27;
  static constexpr int D = 3;
  static constexpr int cx[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0,
                                1, -1, 1, -1, 1, -1, 1, -1};
  static constexpr int cy[Q] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1,
                                1, 1, -1, -1, 1, 1, -1, -1};
  static constexpr int cz[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1,
                                1, 1, 1, 1, -1, -1, -1, -1};
  static constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15,
                                 26, 25, 24, 23, 22, 21, 20, 19};
  static constexpr int opp_ref[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15,
                                     22, 21, 20, 19, 26, 25, 24, 23};
  static constexpr float w[Q] = {8.0f / 27.0f,  2.0f / 27.0f,  2.0f / 27.0f,  2.0f / 27.0f,  2.0f / 27.0f,
                                 2.0f / 27.0f,  2.0f / 27.0f,  1.0f / 54.0f,  1.0f / 54.0f,  1.0f / 54.0f,
                                 1.0f / 54.0f,  1.0f / 54.0f,  1.0f / 54.0f,  1.0f / 54.0f,  1.0f / 54.0f,
                                 1.0f / 54.0f,  1.0f / 54.0f,  1.0f / 54.0f,  1.0f / 54.0f,  1.0f / 216.0f,
                                 1.0f / 216.0f, 1.0f / 216.0f, 1.0f / 216.0f, 1.0f / 216.0f, 1.0f / 216.0f,
                                 1.0f / 216.0f, 1.0f / 216.0f};
};

// ---------------------------------------------------------------------------
// D3Q19 d'Humieres basis and its 7-decimal rounded inverse, computed at
// compile time.  M[k][q] is the k-th moment polynomial evaluated on c_q.
struct MrtTable {
  float minv[19][19];
};

constexpr double kabs(double v) { return v < 0 ? -v : v; }

constexpr double round7(double v) {
  double s = v * 1e7;
  // round half away from zero; no table entry lies on a tie
  long long r = static_cast<long long>(s < 0 ? s - 0.5 : s + 0.5);
  double out = static_cast<double>(r) / 1e7;
  return out == 0.0 ? 0.0 : out;
}

// Entry (k, q) of the d'Humieres D3Q19 moment matrix: basis polynomial k at
// velocity q.
constexpr double d3q19_m(int k, int q) {
  const double cx = D3Q19::cx[q], cy = D3Q19::cy[q], cz = D3Q19::cz[q];
  const double c2 = cx * cx + cy * cy + cz * cz;
  const double row[19] = {1.0,
                          19.0 * c2 - 30.0,
                          (21.0 * c2 * c2 - 53.0 * c2 + 24.0) / 2.0,
                          cx,
                          (5.0 * c2 - 9.0) * cx,
                          cy,
                          (5.0 * c2 - 9.0) * cy,
                          cz,
                          (5.0 * c2 - 9.0) * cz,
                          3.0 * cx * cx - c2,
                          (3.0 * c2 - 5.0) * (3.0 * cx * cx - c2),
                          cy * cy - cz * cz,
                          (3.0 * c2 - 5.0) * (cy * cy - cz * cz),
                          cx * cy,
                          cy * cz,
                          cx * cz,
                          cx * (cy * cy - cz * cz),
                          cy * (cz * cz - cx * cx),
                          cz * (cy * cy - cx * cx)};
  return row[k];
}

constexpr MrtTable make_d3q19_minv() {
  double M[19][19] = {};
  for (int q = 0; q < 19; ++q)
    for (int k = 0; k < 19; ++k) M[k][q] = d3q19_m(k, q);
  double A[19][38] = {};
  for (int i = 0; i < 19; ++i)
    for (int j = 0; j < 38; ++j) A[i][j] = (j < 19) ? M[i][j] : ((j - 19) == i ? 1.0 : 0.0);
  for (int col = 0; col < 19; ++col) {
    int piv = col;
    for (int r = col + 1; r < 19; ++r)
      if (kabs(A[r][col]) > kabs(A[piv][col])) piv = r;
    if (piv != col)
      for (int j = 0; j < 38; ++j) {
        double t = A[col][j];
        A[col][j] = A[piv][j];
        A[piv][j] = t;
      }
    const double d = A[col][col];
    for (int j = 0; j < 38; ++j) A[col][j] /= d;
    for (int r = 0; r < 19; ++r) {
      if (r == col) continue;
      const double f = A[r][col];
      if (f != 0.0)
        for (int j = 0; j < 38; ++j) A[r][j] -= f * A[col][j];
    }
  }
  MrtTable t = {};
  for (int q = 0; q < 19; ++q)
    for (int k = 0; k < 19; ++k) t.minv[q][k] = static_cast<float>(round7(A[q][19 + k]));
  return t;
}

inline constexpr MrtTable kD3Q19Minv = make_d3q19_minv();

// Relaxation rates fixed by the reference grids and kernels.
inline constexpr float kTauBulk2D = 1.63f;  // lattice_grid_2d.cpp:229
inline constexpr float kTauBulk3D = 1.19f;  // lattice_grid_3d.cpp:338

}  // namespace wake_hip
