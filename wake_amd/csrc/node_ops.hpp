// node_ops.hpp -- per-node LBM operators on register arrays (device code).
//
// Every operator here works on the Q populations of one node held in
// registers; the callers decide where they come from (a pull gather, a plain
// load) and where they go.  The arithmetic follows the reference's ISPC
// kernels term for term (cited per function) except that products with the
// lattice constants 0 and +-1 are resolved at compile time, so a D3Q19 node
// costs only its non-trivial FLOPs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

#include "lattice.hpp"

namespace wake_hip {

template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  [&]<int... I>(std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
  }(std::make_integer_sequence<int, N>{});
}

// c * v for a lattice component c in {-1, 0, 1}, folded at compile time.
template <int C>
__device__ __forceinline__ float cmul(float v) {
  if constexpr (C == 1) return v;
  else if constexpr (C == -1) return -v;
  else return 0.0f;
}

// c_q . (ux, uy, uz) summing only the non-zero components, in x, y, z order.
// Reciprocal and square root: correctly rounded, or the hardware's 1-ulp
// instructions (FAST) where the populations are stored in 16 bits and that
// storage rounding dwarfs them.
template <bool FAST>
__device__ __forceinline__ float recip(float x) {
  if constexpr (FAST) return __builtin_amdgcn_rcpf(x);
  else return 1.0f / x;
}
template <bool FAST>
__device__ __forceinline__ float sqrt_of(float x) {
  if constexpr (FAST) return __builtin_amdgcn_sqrtf(x);
  else return sqrtf(x);
}

template <class L, int q>
__device__ __forceinline__ float cdot(float ux, float uy, float uz) {
  constexpr int a = L::cx[q], b = L::cy[q], c = L::cz[q];
  if constexpr (a == 0 && b == 0 && c == 0) return 0.0f;
  else if constexpr (b == 0 && c == 0) return cmul<a>(ux);
  else if constexpr (a == 0 && c == 0) return cmul<b>(uy);
  else if constexpr (a == 0 && b == 0) return cmul<c>(uz);
  else if constexpr (c == 0) return cmul<a>(ux) + cmul<b>(uy);
  else if constexpr (b == 0) return cmul<a>(ux) + cmul<c>(uz);
  else if constexpr (a == 0) return cmul<b>(uy) + cmul<c>(uz);
  else return cmul<a>(ux) + cmul<b>(uy) + cmul<c>(uz);
}

// Density and momentum sums (d3q19_moments.ispc:12-24 order: rho then j).
template <class L>
__device__ __forceinline__ void density_momentum(const float (&f)[L::Q], float& rho, float& jx, float& jy,
                                                 float& jz) {
  rho = 0.0f;
  jx = 0.0f;
  jy = 0.0f;
  jz = 0.0f;
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    rho += f[q];
    if constexpr (L::cx[q] == 1) jx += f[q];
    if constexpr (L::cx[q] == -1) jx -= f[q];
    if constexpr (L::cy[q] == 1) jy += f[q];
    if constexpr (L::cy[q] == -1) jy -= f[q];
    if constexpr (L::cz[q] == 1) jz += f[q];
    if constexpr (L::cz[q] == -1) jz -= f[q];
  });
}

// Velocity with the Guo half-force shift (d3q19_moments.ispc:60-64).
template <bool FAST = false>
__device__ __forceinline__ void guo_velocity(float rho, float jx, float jy, float jz, float gx, float gy, float gz,
                                             float& ux, float& uy, float& uz) {
  const float inv_rho = recip<FAST>(rho);
  ux = (jx + 0.5f * rho * gx) * inv_rho;
  uy = (jy + 0.5f * rho * gy) * inv_rho;
  uz = (jz + 0.5f * rho * gz) * inv_rho;
}

// Second-order equilibrium (lbm_d3q19.isph:78-88).
template <class L, int q>
__device__ __forceinline__ float feq(float rho, float ux, float uy, float uz, float usq) {
  const float cu = cdot<L, q>(ux, uy, uz);
  return L::w[q] * rho * (1.0f + 3.0f * cu + 4.5f * cu * cu - 1.5f * usq);
}

// Smagorinsky relaxation time from the non-equilibrium stress
// (d3q19_les.ispc:7-66, d2q9_les.ispc:6-57):
//   Pi = sum c c (f - feq), Q = sqrt(2 Pi:Pi),
//   tau = (tau0 + sqrt(tau0^2 + 18 sqrt(2) Cs^2 Q / rho)) / 2.
template <class L>
__device__ __forceinline__ float smagorinsky_tau(const float (&f)[L::Q], float rho, float ux, float uy, float uz,
                                                 float tau0, float cs) {
  const float usq = ux * ux + uy * uy + uz * uz;
  float pxx = 0.0f, pyy = 0.0f, pzz = 0.0f, pxy = 0.0f, pyz = 0.0f, pxz = 0.0f;
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int a = L::cx[q], b = L::cy[q], c = L::cz[q];
    const float fn = f[q] - feq<L, q>(rho, ux, uy, uz, usq);
    if constexpr (a != 0) pxx += fn;
    if constexpr (b != 0) pyy += fn;
    if constexpr (c != 0) pzz += fn;
    if constexpr (a * b == 1) pxy += fn;
    if constexpr (a * b == -1) pxy -= fn;
    if constexpr (b * c == 1) pyz += fn;
    if constexpr (b * c == -1) pyz -= fn;
    if constexpr (a * c == 1) pxz += fn;
    if constexpr (a * c == -1) pxz -= fn;
  });
  float tensor_sum;
  if constexpr (L::D == 2)
    tensor_sum = pxx * pxx + 2.0f * pxy * pxy + pyy * pyy;
  else
    tensor_sum = (pxx * pxx + pyy * pyy + pzz * pzz) + 2.0f * (pxy * pxy + pyz * pyz + pxz * pxz);
  const float q_stress = sqrtf(2.0f * tensor_sum);
  const float prefactor = 18.0f * 1.4142135623730951f * (cs * cs);
  const float disc = tau0 * tau0 + (prefactor / rho) * q_stress;
  return 0.5f * (tau0 + sqrtf(disc));
}

// The same relaxation time with the non-equilibrium stress taken from the
// second moments, Pi_ab = sum c_a c_b f - (rho/3 delta_ab + rho u_a u_b), which
// equals sum c_a c_b (f - feq) for the second-order equilibrium of D2Q9,
// D3Q19 and D3Q27 (their weights are isotropic to fourth order).  Used by the
// LES-coupled BGK/TRT step: no equilibrium is formed, so none is kept live
// beside the collision's own.
template <class L, bool FAST = false>
__device__ __forceinline__ float smagorinsky_tau_moments(const float (&f)[L::Q], float rho, float ux, float uy,
                                                         float uz, float tau0, float cs) {
  float sxx = 0.0f, syy = 0.0f, szz = 0.0f, sxy = 0.0f, syz = 0.0f, sxz = 0.0f;
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int a = L::cx[q], b = L::cy[q], c = L::cz[q];
    const float v = f[q];
    if constexpr (a != 0) sxx += v;
    if constexpr (b != 0) syy += v;
    if constexpr (c != 0) szz += v;
    if constexpr (a * b == 1) sxy += v;
    if constexpr (a * b == -1) sxy -= v;
    if constexpr (b * c == 1) syz += v;
    if constexpr (b * c == -1) syz -= v;
    if constexpr (a * c == 1) sxz += v;
    if constexpr (a * c == -1) sxz -= v;
  });
  const float r3 = rho * (1.0f / 3.0f);
  const float pxx = sxx - (r3 + rho * ux * ux);
  const float pyy = syy - (r3 + rho * uy * uy);
  const float pxy = sxy - rho * ux * uy;
  float tensor_sum;
  if constexpr (L::D == 2) {
    tensor_sum = pxx * pxx + 2.0f * pxy * pxy + pyy * pyy;
  } else {
    const float pzz = szz - (r3 + rho * uz * uz);
    const float pyz = syz - rho * uy * uz;
    const float pxz = sxz - rho * ux * uz;
    tensor_sum = (pxx * pxx + pyy * pyy + pzz * pzz) + 2.0f * (pxy * pxy + pyz * pyz + pxz * pxz);
  }
  const float q_stress = sqrt_of<FAST>(2.0f * tensor_sum);
  const float prefactor = 18.0f * 1.4142135623730951f * (cs * cs);
  const float p_rho = FAST ? prefactor * recip<FAST>(rho) : prefactor / rho;
  const float disc = tau0 * tau0 + p_rho * q_stress;
  return 0.5f * (tau0 + sqrt_of<FAST>(disc));
}

// TRT antisymmetric tau from the grid (lattice_grid_3d.cpp:328-330), Lambda = 1/4.
__device__ __host__ __forceinline__ float trt_tau_a(float tau_s) {
  float diff = tau_s - 0.5f;
  diff = diff < 1e-4f ? 1e-4f : diff;
  return 0.5f + 0.25f / diff;
}

struct Macro {
  float rho, ux, uy, uz;
};

// BGK with Guo forcing, in place (d3q19_collision_bgk.ispc:4-75,
// d2q9_collision_bgk.ispc:5-86).
template <class L, bool FAST = false>
__device__ __forceinline__ Macro collide_bgk(float (&f)[L::Q], float tau, float gx, float gy, float gz) {
  const float omega = recip<FAST>(tau);
  const float pre = 1.0f - 0.5f * omega;
  float rho, jx, jy, jz, ux, uy, uz;
  density_momentum<L>(f, rho, jx, jy, jz);
  guo_velocity<FAST>(rho, jx, jy, jz, gx, gy, gz, ux, uy, uz);
  const float u_dot_g = ux * gx + uy * gy + uz * gz;
  const float usq = ux * ux + uy * uy + uz * uz;
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const float cu = cdot<L, q>(ux, uy, uz);
    const float cg = cdot<L, q>(gx, gy, gz);
    const float fe = L::w[q] * rho * (1.0f + 3.0f * cu + 4.5f * cu * cu - 1.5f * usq);
    const float Fi = pre * L::w[q] * rho * (3.0f * (cg - u_dot_g) + 9.0f * cu * cg);
    f[q] = f[q] - omega * (f[q] - fe) + Fi;
  });
  return {rho, ux, uy, uz};
}

// TRT with split Guo forcing, in place (d3q19_collision_trt.ispc:5-93,
// d3q27_collision_mrt.ispc:5-91).  OPP selects the pairing table.  Each
// opposite pair is relaxed together from its two populations and their
// equilibria formed on the fly, so no equilibrium or output array is held
// (the D3Q27 kernel stays within 5 waves per SIMD); every output is the
// same expression of the same operands as a per-population loop.
template <class L, bool REF_PAIRS = false, bool FAST = false>
__device__ __forceinline__ Macro collide_trt(float (&f)[L::Q], float tau_s, float tau_a, float gx, float gy,
                                             float gz) {
  const float omega_s = recip<FAST>(tau_s), omega_a = recip<FAST>(tau_a);
  const float pre_s = 1.0f - 0.5f * omega_s, pre_a = 1.0f - 0.5f * omega_a;
  float rho, jx, jy, jz, ux, uy, uz;
  density_momentum<L>(f, rho, jx, jy, jz);
  guo_velocity<FAST>(rho, jx, jy, jz, gx, gy, gz, ux, uy, uz);
  const float u_dot_g = ux * gx + uy * gy + uz * gz;
  const float usq = ux * ux + uy * uy + uz * uz;
  auto relax = [&](auto qc, float fq, float fo, float feq_q, float feo) {
    constexpr int q = decltype(qc)::value;
    const float f_sym = 0.5f * (fq + fo);
    const float f_asym = 0.5f * (fq - fo);
    const float fe_sym = 0.5f * (feq_q + feo);
    const float fe_asym = 0.5f * (feq_q - feo);
    const float cu = cdot<L, q>(ux, uy, uz);
    const float cg = cdot<L, q>(gx, gy, gz);
    const float F_asym = pre_a * L::w[q] * rho * (3.0f * cg);
    const float F_sym = pre_s * L::w[q] * rho * (-3.0f * u_dot_g + 9.0f * cu * cg);
    return fq - omega_s * (f_sym - fe_sym) - omega_a * (f_asym - fe_asym) + (F_sym + F_asym);
  };
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int o = REF_PAIRS ? L::opp_ref[q] : L::opp[q];
    if constexpr (q <= o) {
      const float fq = f[q], fo = f[o];
      const float feq_q = feq<L, q>(rho, ux, uy, uz, usq);
      const float feo = feq<L, o>(rho, ux, uy, uz, usq);
      f[q] = relax(qc, fq, fo, feq_q, feo);
      if constexpr (q != o) f[o] = relax(std::integral_constant<int, o>{}, fo, fq, feo, feq_q);
    }
  });
  return {rho, ux, uy, uz};
}

// D2Q9 MRT (Lallemand-Luo), in place (d2q9_collision_mrt.ispc:6-137).
// With LES = true the shear rate comes from the Smagorinsky tau, the stress
// taken from the moment deviations (sum c^2 (f - feq) = dm1 / 3,
// Pi_xx - Pi_yy = dm7, Pi_xy = dm8) as in collide_mrt_d3q19.
template <bool LES = false, bool FAST = false>
__device__ __forceinline__ Macro collide_mrt_d2q9(float (&f)[9], float tau_shear, float tau_bulk, float gx,
                                                  float gy, float cs = 0.0f) {
  const float s_e = 1.0f / tau_bulk, s_eps = 1.4f, s_q = 1.6f;
  const float f0 = f[0], f1 = f[1], f2 = f[2], f3 = f[3], f4 = f[4], f5 = f[5], f6 = f[6], f7 = f[7], f8 = f[8];
  float m0 = f0 + f1 + f2 + f3 + f4 + f5 + f6 + f7 + f8;
  float m1 = -4.0f * f0 - (f1 + f2 + f3 + f4) + 2.0f * (f5 + f6 + f7 + f8);
  float m2 = 4.0f * f0 - 2.0f * (f1 + f2 + f3 + f4) + (f5 + f6 + f7 + f8);
  float m3 = (f1 - f3) + (f5 - f6 - f7 + f8);
  float m4 = -2.0f * (f1 - f3) + (f5 - f6 - f7 + f8);
  float m5 = (f2 - f4) + (f5 + f6 - f7 - f8);
  float m6 = -2.0f * (f2 - f4) + (f5 + f6 - f7 - f8);
  float m7 = (f1 - f2 + f3 - f4);
  float m8 = (f5 - f6 + f7 - f8);
  const float rho = m0;
  const float inv_rho = recip<FAST>(rho);
  const float ux = (m3 + 0.5f * rho * gx) * inv_rho;
  const float uy = (m5 + 0.5f * rho * gy) * inv_rho;
  const float ux2 = ux * ux, uy2 = uy * uy, u2 = ux2 + uy2;
  const float meq1 = rho * (-2.0f + 3.0f * u2);
  const float meq2 = rho * (1.0f - 3.0f * u2);
  const float meq4 = -rho * ux;
  const float meq6 = -rho * uy;
  const float meq7 = rho * (ux2 - uy2);
  const float meq8 = rho * (ux * uy);
  float s_p;
  if constexpr (LES) {
    const float dc2 = (m1 - meq1) * (1.0f / 3.0f);
    const float dm7 = m7 - meq7;
    const float pxx = 0.5f * (dc2 + dm7), pyy = 0.5f * (dc2 - dm7), pxy = m8 - meq8;
    const float tensor_sum = pxx * pxx + 2.0f * pxy * pxy + pyy * pyy;
    const float q_stress = sqrt_of<FAST>(2.0f * tensor_sum);
    const float prefactor = 18.0f * 1.4142135623730951f * (cs * cs);
    const float p_rho = FAST ? prefactor * inv_rho : prefactor / rho;
    const float disc = tau_shear * tau_shear + p_rho * q_stress;
    s_p = recip<FAST>(0.5f * (tau_shear + sqrt_of<FAST>(disc)));
  } else {
    s_p = recip<FAST>(tau_shear);
  }
  const float F_pre = 1.0f - 0.5f * s_p;
  const float F_m3 = rho * gx, F_m5 = rho * gy;
  const float F_m1 = 6.0f * (ux * F_m3 + uy * F_m5);
  const float F_m2 = -6.0f * (ux * F_m3 + uy * F_m5);
  const float F_m7 = 2.0f * (ux * F_m3 - uy * F_m5);
  const float F_m8 = (ux * F_m5 + uy * F_m3);
  m1 = m1 - s_e * (m1 - meq1) + F_pre * F_m1;
  m2 = m2 - s_eps * (m2 - meq2) + F_pre * F_m2;
  m3 = m3 + F_m3;
  m4 = m4 - s_q * (m4 - meq4) + F_pre * (-F_m3);
  m5 = m5 + F_m5;
  m6 = m6 - s_q * (m6 - meq6) + F_pre * (-F_m5);
  m7 = m7 - s_p * (m7 - meq7) + F_pre * F_m7;
  m8 = m8 - s_p * (m8 - meq8) + F_pre * F_m8;
  const float a9 = 1.0f / 9.0f, a18 = 1.0f / 18.0f, a36 = 1.0f / 36.0f, a6 = 1.0f / 6.0f, a12 = 1.0f / 12.0f;
  f[0] = a9 * m0 - a9 * m1 + a9 * m2;
  f[1] = a9 * m0 - a36 * m1 - a18 * m2 + a6 * m3 - a6 * m4 + 0.25f * m7;
  f[2] = a9 * m0 - a36 * m1 - a18 * m2 + a6 * m5 - a6 * m6 - 0.25f * m7;
  f[3] = a9 * m0 - a36 * m1 - a18 * m2 - a6 * m3 + a6 * m4 + 0.25f * m7;
  f[4] = a9 * m0 - a36 * m1 - a18 * m2 - a6 * m5 + a6 * m6 - 0.25f * m7;
  f[5] = a9 * m0 + a18 * m1 + a36 * m2 + a6 * m3 + a12 * m4 + a6 * m5 + a12 * m6 + 0.25f * m8;
  f[6] = a9 * m0 + a18 * m1 + a36 * m2 - a6 * m3 - a12 * m4 + a6 * m5 + a12 * m6 - 0.25f * m8;
  f[7] = a9 * m0 + a18 * m1 + a36 * m2 - a6 * m3 - a12 * m4 - a6 * m5 - a12 * m6 + 0.25f * m8;
  f[8] = a9 * m0 + a18 * m1 + a36 * m2 + a6 * m3 + a12 * m4 - a6 * m5 - a12 * m6 - 0.25f * m8;
  return {rho, ux, uy, 0.0f};
}

// A D3Q19 moment is "odd" when its basis polynomial changes sign under
// c -> -c (the momenta, the energy fluxes and the third-order m16..m18).
struct D3Q19OddMoment {
  static constexpr bool v[19] = {false, false, false, true,  true,  true,  true,  true,  true, false,
                                 false, false, false, false, false, false, true,  true,  true};
};

// D3Q19 MRT (d'Humieres), in place (d3q19_collision_mrt.ispc:26-176).  The
// back transform applies the rounded inverse table: the odd-moment part of
// rows q and opp(q) differs only in sign, so each opposite pair shares one
// even and one odd partial sum.
//
// With LES = true the shear rate comes from the Smagorinsky tau of this node:
// the non-equilibrium stress is read off the moment deviations the collision
// already has (meq = M feq exactly for these moments):
//   sum c^2 (f - feq) = dm1 / 19,   Pi_xx = (dm9 + dm1 / 19) / 3,
//   Pi_yy, Pi_zz = ((dm1 / 19 - Pi_xx) +- dm11) / 2,
//   Pi_xy, Pi_yz, Pi_xz = dm13, dm14, dm15,
// which is sum_q c c (f - feq) of d3q19_les.ispc:139-159 without forming feq.
__device__ __forceinline__ void mrt_forward_d3q19(const float (&f)[19], float (&m)[19]) {
  m[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] +
         f[14] + f[15] + f[16] + f[17] + f[18];
  const float faces = f[1] + f[2] + f[3] + f[4] + f[5] + f[6];
  const float edges = f[7] + f[8] + f[9] + f[10] + f[11] + f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
  m[1] = -30.0f * f[0] - 11.0f * faces + 8.0f * edges;
  m[2] = 12.0f * f[0] - 4.0f * faces + edges;
  const float ex = f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14];
  const float ey = f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18];
  const float ez = f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18];
  m[3] = (f[1] - f[2]) + ex;
  m[4] = -4.0f * (f[1] - f[2]) + ex;
  m[5] = (f[3] - f[4]) + ey;
  m[6] = -4.0f * (f[3] - f[4]) + ey;
  m[7] = (f[5] - f[6]) + ez;
  m[8] = -4.0f * (f[5] - f[6]) + ez;
  const float exy = f[7] + f[8] + f[9] + f[10], exz = f[11] + f[12] + f[13] + f[14];
  const float eyz = f[15] + f[16] + f[17] + f[18];
  m[9] = 2.0f * (f[1] + f[2]) - (f[3] + f[4] + f[5] + f[6]) + (exy + exz) - 2.0f * eyz;
  m[10] = -4.0f * (f[1] + f[2]) + 2.0f * (f[3] + f[4] + f[5] + f[6]) + (exy + exz) - 2.0f * eyz;
  m[11] = (f[3] + f[4]) - (f[5] + f[6]) + exy - exz;
  m[12] = -2.0f * (f[3] + f[4]) + 2.0f * (f[5] + f[6]) + exy - exz;
  m[13] = (f[7] - f[8] - f[9] + f[10]);
  m[14] = (f[15] - f[16] - f[17] + f[18]);
  m[15] = (f[11] - f[12] - f[13] + f[14]);
  m[16] = (f[7] - f[8] + f[9] - f[10] - f[11] + f[12] - f[13] + f[14]);
  m[17] = (-f[7] - f[8] + f[9] + f[10] + f[15] - f[16] + f[17] - f[18]);
  m[18] = (-f[11] - f[12] + f[13] + f[14] + f[15] + f[16] - f[17] - f[18]);
}

// Moment-space relaxation of one node (in place on m), returning rho and u.
template <bool LES, bool FAST>
__device__ __forceinline__ Macro mrt_relax_d3q19(float (&m)[19], float tau_shear, float tau_bulk, float gx, float gy,
                                                 float gz, float cs) {
  const float s_e = 1.0f / tau_bulk, s_eps = 1.4f, s_q = 1.2f, s_pi = 1.4f, s_m = 1.98f;
  const float rho = m[0];
  const float inv_rho = recip<FAST>(rho);
  const float ux = (m[3] + 0.5f * rho * gx) * inv_rho;
  const float uy = (m[5] + 0.5f * rho * gy) * inv_rho;
  const float uz = (m[7] + 0.5f * rho * gz) * inv_rho;
  const float ux2 = ux * ux, uy2 = uy * uy, uz2 = uz * uz, u2 = ux2 + uy2 + uz2;
  const float meq9 = rho * (2.0f * ux2 - uy2 - uz2);
  const float meq11 = rho * (uy2 - uz2);
  const float meq1 = rho * (-11.0f + 19.0f * u2);
  float s_nu;
  if constexpr (LES) {
    const float dm1 = m[1] - meq1;
    const float dc2 = dm1 * (1.0f / 19.0f);
    const float pxx = (m[9] - meq9 + dc2) * (1.0f / 3.0f);
    const float dm11 = m[11] - meq11;
    const float pyy = 0.5f * ((dc2 - pxx) + dm11);
    const float pzz = 0.5f * ((dc2 - pxx) - dm11);
    const float pxy = m[13] - rho * (ux * uy);
    const float pyz = m[14] - rho * (uy * uz);
    const float pxz = m[15] - rho * (ux * uz);
    const float tensor_sum = (pxx * pxx + pyy * pyy + pzz * pzz) + 2.0f * (pxy * pxy + pyz * pyz + pxz * pxz);
    const float q_stress = sqrt_of<FAST>(2.0f * tensor_sum);
    const float prefactor = 18.0f * 1.4142135623730951f * (cs * cs);
    const float p_rho = FAST ? prefactor * inv_rho : prefactor / rho;
    const float disc = tau_shear * tau_shear + p_rho * q_stress;
    s_nu = recip<FAST>(0.5f * (tau_shear + sqrt_of<FAST>(disc)));
  } else {
    s_nu = recip<FAST>(tau_shear);
  }
  m[1] = m[1] - s_e * (m[1] - meq1);
  m[2] = m[2] - s_eps * (m[2] - rho * (3.0f - 11.0f / 2.0f * u2));
  m[3] = m[3] + rho * gx;
  m[4] = m[4] - s_q * (m[4] - (-2.0f / 3.0f * rho * ux)) - (2.0f / 3.0f) * rho * gx;
  m[5] = m[5] + rho * gy;
  m[6] = m[6] - s_q * (m[6] - (-2.0f / 3.0f * rho * uy)) - (2.0f / 3.0f) * rho * gy;
  m[7] = m[7] + rho * gz;
  m[8] = m[8] - s_q * (m[8] - (-2.0f / 3.0f * rho * uz)) - (2.0f / 3.0f) * rho * gz;
  m[9] = m[9] - s_nu * (m[9] - meq9);
  m[10] = m[10] - s_pi * (m[10] - (-0.5f * meq9));
  m[11] = m[11] - s_nu * (m[11] - meq11);
  m[12] = m[12] - s_pi * (m[12] - (-0.5f * meq11));
  m[13] = m[13] - s_nu * (m[13] - rho * (ux * uy));
  m[14] = m[14] - s_nu * (m[14] - rho * (uy * uz));
  m[15] = m[15] - s_nu * (m[15] - rho * (ux * uz));
  m[16] = m[16] - s_m * m[16];
  m[17] = m[17] - s_m * m[17];
  m[18] = m[18] - s_m * m[18];
  return {rho, ux, uy, uz};
}

// Back transform with the rounded inverse table (odd moments: see
// D3Q19OddMoment).
__device__ __forceinline__ void mrt_inverse_d3q19(const float (&m)[19], float (&f)[19]) {
  static_for<19>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int o = D3Q19::opp[q];
    if constexpr (q == 0) {
      float v = 0.0f;
      static_for<19>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr float c = kD3Q19Minv.minv[0][k];
        if constexpr (c != 0.0f) v += c * m[k];
      });
      f[0] = v;
    } else if constexpr (q < o) {
      float ev = 0.0f, od = 0.0f;
      static_for<19>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr float c = kD3Q19Minv.minv[q][k];
        if constexpr (c != 0.0f) {
          if constexpr (D3Q19OddMoment::v[k]) od += c * m[k];
          else ev += c * m[k];
        }
      });
      f[q] = ev + od;
      f[o] = ev - od;
    }
  });
}

// The two transforms for 16-bit population storage.  The storage rounds
// f - w_q to 11 (FP16) or 8 (BF16) significant bits, far above the rounding of
// any order of FP32 operations, so, as in collide_trt_fast27, the sums are
// reorganised for issue count (VALU issue, not HBM, limited these kernels):
//  * forward: each moment row is even or odd under c -> -c, so the even
//    moments are sums of the nine opposite-pair sums s and the odd moments
//    sums of the pair differences d (about 85 operations, against about 130
//    row by row);
//  * inverse: the same rounded table as mrt_inverse_d3q19, each entry applied
//    once, but the rows of the faces and of the edges share their partial
//    sums (one rho-and-energy sum per velocity shell, one stress sum per
//    plane, one odd sum per axis), and each opposite pair is formed as
//    even part +- odd part (about 75 operations, against about 110).
struct D3Q19PairLayout {
  // the velocity order both functions below are written for
  static constexpr bool ok =
      D3Q19::cx[1] == 1 && D3Q19::opp[1] == 2 && D3Q19::cy[3] == 1 && D3Q19::opp[3] == 4 && D3Q19::cz[5] == 1 &&
      D3Q19::opp[5] == 6 && D3Q19::cx[7] == 1 && D3Q19::cy[7] == 1 && D3Q19::opp[7] == 10 && D3Q19::cx[8] == -1 &&
      D3Q19::cy[8] == 1 && D3Q19::opp[8] == 9 && D3Q19::cx[11] == 1 && D3Q19::cz[11] == 1 && D3Q19::opp[11] == 14 &&
      D3Q19::cx[12] == -1 && D3Q19::cz[12] == 1 && D3Q19::opp[12] == 13 && D3Q19::cy[15] == 1 &&
      D3Q19::cz[15] == 1 && D3Q19::opp[15] == 18 && D3Q19::cy[16] == -1 && D3Q19::cz[16] == 1 &&
      D3Q19::opp[16] == 17;
};
static_assert(D3Q19PairLayout::ok, "D3Q19 velocity order changed: rewrite the pair transforms");

// Both transforms are written once for floats and for LinearForm<19>, with
// which the compiler checks below that the forward transform is the basis M
// and the inverse applies the rounded table.
template <class V>
__host__ __device__ constexpr void mrt_forward_d3q19_shared(const V (&f)[19], V (&m)[19]) {
  const V s1 = f[1] + f[2], d1 = f[1] - f[2];          // (+-1, 0, 0)
  const V s3 = f[3] + f[4], d3 = f[3] - f[4];          // (0, +-1, 0)
  const V s5 = f[5] + f[6], d5 = f[5] - f[6];          // (0, 0, +-1)
  const V s7 = f[7] + f[10], d7 = f[7] - f[10];        // +-(1, 1, 0)
  const V s8 = f[8] + f[9], d8 = f[8] - f[9];          // +-(-1, 1, 0)
  const V s11 = f[11] + f[14], d11 = f[11] - f[14];    // +-(1, 0, 1)
  const V s12 = f[12] + f[13], d12 = f[12] - f[13];    // +-(-1, 0, 1)
  const V s15 = f[15] + f[18], d15 = f[15] - f[18];    // +-(0, 1, 1)
  const V s16 = f[16] + f[17], d16 = f[16] - f[17];    // +-(0, -1, 1)
  const V syz = s3 + s5;
  const V faces = s1 + syz;
  const V exy = s7 + s8, exz = s11 + s12, eyz = s15 + s16;
  const V exyz = exy + exz;
  const V edges = exyz + eyz;
  m[0] = f[0] + faces + edges;
  m[1] = -30.0f * f[0] - 11.0f * faces + 8.0f * edges;
  m[2] = 12.0f * f[0] - 4.0f * faces + edges;
  // odd parts of the edges: c_x over the xy and xz edges, c_y over xy and
  // yz, c_z over xz and yz
  const V xp = d7 - d8, xq = d11 - d12;
  const V yp = d7 + d8, yq = d15 - d16;
  const V zp = d11 + d12, zq = d15 + d16;
  const V ex = xp + xq, ey = yp + yq, ez = zp + zq;
  m[3] = d1 + ex;
  m[4] = -4.0f * d1 + ex;
  m[5] = d3 + ey;
  m[6] = -4.0f * d3 + ey;
  m[7] = d5 + ez;
  m[8] = -4.0f * d5 + ez;
  const V t = exyz - 2.0f * eyz;
  m[9] = 2.0f * s1 - syz + t;
  m[10] = -4.0f * s1 + 2.0f * syz + t;
  const V u = s3 - s5, v = exy - exz;
  m[11] = u + v;
  m[12] = -2.0f * u + v;
  m[13] = s7 - s8;
  m[14] = s15 - s16;
  m[15] = s11 - s12;
  m[16] = xp - xq;
  m[17] = yq - yp;
  m[18] = zq - zp;
}

template <int q, int k>
inline constexpr float kMinv = kD3Q19Minv.minv[q][k];

template <class V>
__host__ __device__ constexpr void mrt_inverse_d3q19_shared(const V (&m)[19], V (&f)[19]) {
  auto pair = [&](int q, int o, const V& even, const V& odd) {
    f[q] = even + odd;
    f[o] = even - odd;
  };
  f[0] = kMinv<0, 0> * m[0] + kMinv<0, 1> * m[1] + kMinv<0, 2> * m[2];
  // faces: one shell sum; the x faces weigh m9, m10; the y and z faces weigh
  // m9, m10 alike and m11, m12 with opposite signs
  const V F = kMinv<1, 0> * m[0] + kMinv<1, 1> * m[1] + kMinv<1, 2> * m[2];
  const V G = F + kMinv<3, 9> * m[9] + kMinv<3, 10> * m[10];
  const V B = kMinv<3, 11> * m[11] + kMinv<3, 12> * m[12];
  pair(1, 2, F + kMinv<1, 9> * m[9] + kMinv<1, 10> * m[10], kMinv<1, 3> * m[3] + kMinv<1, 4> * m[4]);
  pair(3, 4, G + B, kMinv<3, 5> * m[5] + kMinv<3, 6> * m[6]);
  pair(5, 6, G - B, kMinv<5, 7> * m[7] + kMinv<5, 8> * m[8]);
  // edges: one shell sum; xy and xz edges share m9, m10 and take m11, m12
  // with opposite signs; one odd sum per axis (c = +1) and per third-order
  // moment
  const V E = kMinv<7, 0> * m[0] + kMinv<7, 1> * m[1] + kMinv<7, 2> * m[2];
  const V EP = E + kMinv<7, 9> * m[9] + kMinv<7, 10> * m[10];
  const V R = kMinv<7, 11> * m[11] + kMinv<7, 12> * m[12];
  const V X = kMinv<7, 3> * m[3] + kMinv<7, 4> * m[4];
  const V Y = kMinv<7, 5> * m[5] + kMinv<7, 6> * m[6];
  const V Z = kMinv<11, 7> * m[7] + kMinv<11, 8> * m[8];
  {  // xy edges: (1, 1, 0) and (-1, 1, 0)
    const V b = EP + R;
    const V t13 = kMinv<7, 13> * m[13];
    const V xa = X + kMinv<7, 16> * m[16], yb = Y + kMinv<7, 17> * m[17];
    pair(7, 10, b + t13, xa + yb);
    pair(8, 9, b - t13, yb - xa);
  }
  {  // xz edges: (1, 0, 1) and (-1, 0, 1)
    const V b = EP - R;
    const V t15 = kMinv<11, 15> * m[15];
    const V xc = X + kMinv<11, 16> * m[16], zc = Z + kMinv<11, 18> * m[18];
    pair(11, 14, b + t15, xc + zc);
    pair(12, 13, b - t15, zc - xc);
  }
  {  // yz edges: (0, 1, 1) and (0, -1, 1)
    const V b = E + kMinv<15, 9> * m[9] + kMinv<15, 10> * m[10];
    const V t14 = kMinv<15, 14> * m[14];
    const V yd = Y + kMinv<15, 17> * m[17], zd = Z + kMinv<15, 18> * m[18];
    pair(15, 18, b + t14, yd + zd);
    pair(16, 17, b - t14, zd - yd);
  }
}

// A linear form over N values, exact in double for the small coefficients
// the transforms use.
template <int N>
struct LinearForm {
  double c[N] = {};
  constexpr LinearForm operator+(const LinearForm& o) const {
    LinearForm r;
    for (int k = 0; k < N; ++k) r.c[k] = c[k] + o.c[k];
    return r;
  }
  constexpr LinearForm operator-(const LinearForm& o) const {
    LinearForm r;
    for (int k = 0; k < N; ++k) r.c[k] = c[k] - o.c[k];
    return r;
  }
};
template <int N>
constexpr LinearForm<N> operator*(float s, const LinearForm<N>& v) {
  LinearForm<N> r;
  for (int k = 0; k < N; ++k) r.c[k] = static_cast<double>(s) * v.c[k];
  return r;
}

// Moment k of the forward transform is sum_q M[k][q] f_q (small integers,
// exact in double).
constexpr bool mrt_forward_shared_matches_basis() {
  LinearForm<19> f[19], m[19];
  for (int q = 0; q < 19; ++q) f[q].c[q] = 1.0;
  mrt_forward_d3q19_shared(f, m);
  for (int k = 0; k < 19; ++k)
    for (int q = 0; q < 19; ++q)
      if (m[k].c[q] != d3q19_m(k, q)) return false;
  return true;
}
static_assert(mrt_forward_shared_matches_basis(), "pair-sum forward transform is not the basis M");

// Every output row of the inverse equals its table row, entry for entry
// (each entry enters once, so the comparison is exact).
constexpr bool mrt_inverse_shared_matches_table() {
  LinearForm<19> m[19], f[19];
  for (int k = 0; k < 19; ++k) m[k].c[k] = 1.0;
  mrt_inverse_d3q19_shared(m, f);
  for (int q = 0; q < 19; ++q)
    for (int k = 0; k < 19; ++k)
      if (f[q].c[k] != static_cast<double>(kD3Q19Minv.minv[q][k])) return false;
  return true;
}
static_assert(mrt_inverse_shared_matches_table(), "shared-sum inverse does not apply the rounded table");

__device__ __forceinline__ void mrt_forward_d3q19_pairs(const float (&f)[19], float (&m)[19]) {
  mrt_forward_d3q19_shared(f, m);
}
__device__ __forceinline__ void mrt_inverse_d3q19_pairs(const float (&m)[19], float (&f)[19]) {
  mrt_inverse_d3q19_shared(m, f);
}

template <bool LES = false, bool FAST = false>
__device__ __forceinline__ Macro collide_mrt_d3q19(float (&f)[19], float tau_shear, float tau_bulk, float gx,
                                                   float gy, float gz, float cs = 0.0f) {
  float m[19];
  if constexpr (FAST) mrt_forward_d3q19_pairs(f, m);
  else mrt_forward_d3q19(f, m);
  const Macro r = mrt_relax_d3q19<LES, FAST>(m, tau_shear, tau_bulk, gx, gy, gz, cs);
  if constexpr (FAST) mrt_inverse_d3q19_pairs(m, f);
  else mrt_inverse_d3q19(m, f);
  return r;
}

// --------------------------------------------------------------------------
// D3Q27 TRT for 16-bit population storage.  The storage rounds f - w_q to 11
// (FP16) or 8 (BF16) significant bits, far above the rounding of any order
// of FP32 operations, so this path is free to reorganise the arithmetic for
// issue count (VALU issue, not HBM, limited the 16-bit kernels):
//  * the moments rho, j and sum c_a c_b f come from summing the 27
//    populations along x, then y, then z (66 additions, against about 160
//    for the separate sums);
//  * each opposite pair (q, o) relaxes through its sum s and difference d
//    (trt_pairs), which is f - omega_s (f+ - feq+) - omega_a (f- - feq-) with
//    the symmetric and antisymmetric equilibria formed once per pair;
//  * the Guo force terms are compiled out when the body force is zero.

// Index of velocity (a, b, c) in lattice L, or -1.
template <class L>
constexpr int q_of(int a, int b, int c) {
  for (int q = 0; q < L::Q; ++q)
    if (L::cx[q] == a && L::cy[q] == b && L::cz[q] == c) return q;
  return -1;
}

template <class V>
struct Moments2T {
  V rho, jx, jy, jz;        // sum f, sum c f
  V sxx, syy, szz;          // sum c_a^2 f
  V sxy, syz, sxz;          // sum c_a c_b f
};
using Moments2 = Moments2T<float>;

// The x lines of lattice L: line i = 3 (b + 1) + (c + 1) holds the
// velocities (-1, b, c), (0, b, c), (1, b, c).
struct XLines {
  int q[9][3];
};
template <class L>
constexpr XLines make_x_lines() {
  XLines t = {};
  for (int i = 0; i < 9; ++i)
    for (int a = -1; a <= 1; ++a) t.q[i][a + 1] = q_of<L>(a, i / 3 - 1, i % 3 - 1);
  return t;
}

// Written once for floats and for LinearForm<27>, with which the compiler
// checks below that every moment is its sum over the 27 populations.
template <class L, class V>
__host__ __device__ constexpr Moments2T<V> moments_d3q27_t(const V (&f)[L::Q]) {
  static_assert(L::Q == 27, "the x-y-z cascade needs every velocity of {-1,0,1}^3");
  constexpr XLines xl = make_x_lines<L>();
  // x lines (y = b, z = c): orders 0, 1, 2 in c_x
  V k0[3][3], k1[3][3], k2[3][3];
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    const V t = f[xl.q[i][2]] + f[xl.q[i][0]];
    k2[i / 3][i % 3] = t;
    k1[i / 3][i % 3] = f[xl.q[i][2]] - f[xl.q[i][0]];
    k0[i / 3][i % 3] = t + f[xl.q[i][1]];
  }
  // y: for each z, orders (x0 y0, x0 y1, x0 y2, x1 y0, x1 y1, x2 y0)
  V a0[3], a1[3], a2[3], b0[3], b1[3], c0[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const V t = k0[2][c] + k0[0][c];
    a2[c] = t;
    a0[c] = t + k0[1][c];
    a1[c] = k0[2][c] - k0[0][c];
    b0[c] = (k1[2][c] + k1[0][c]) + k1[1][c];
    b1[c] = k1[2][c] - k1[0][c];
    c0[c] = (k2[2][c] + k2[0][c]) + k2[1][c];
  }
  // z
  Moments2T<V> m;
  const V t = a0[2] + a0[0];
  m.szz = t;
  m.rho = t + a0[1];
  m.jz = a0[2] - a0[0];
  m.jy = (a1[2] + a1[0]) + a1[1];
  m.syz = a1[2] - a1[0];
  m.syy = (a2[2] + a2[0]) + a2[1];
  m.jx = (b0[2] + b0[0]) + b0[1];
  m.sxz = b0[2] - b0[0];
  m.sxy = (b1[2] + b1[0]) + b1[1];
  m.sxx = (c0[2] + c0[0]) + c0[1];
  return m;
}

constexpr bool moments_d3q27_cascade_matches_sums() {
  for (int i = 0; i < 9; ++i)
    for (int a = 0; a < 3; ++a)
      if (make_x_lines<D3Q27>().q[i][a] < 0) return false;
  LinearForm<27> f[27];
  for (int q = 0; q < 27; ++q) f[q].c[q] = 1.0;
  const Moments2T<LinearForm<27>> m = moments_d3q27_t<D3Q27>(f);
  for (int q = 0; q < 27; ++q) {
    const double x = D3Q27::cx[q], y = D3Q27::cy[q], z = D3Q27::cz[q];
    if (m.rho.c[q] != 1.0 || m.jx.c[q] != x || m.jy.c[q] != y || m.jz.c[q] != z || m.sxx.c[q] != x * x ||
        m.syy.c[q] != y * y || m.szz.c[q] != z * z || m.sxy.c[q] != x * y || m.syz.c[q] != y * z ||
        m.sxz.c[q] != x * z)
      return false;
  }
  return true;
}
static_assert(moments_d3q27_cascade_matches_sums(), "x-y-z cascade is not the moment sums");

template <class L>
__device__ __forceinline__ Moments2 moments_d3q27(const float (&f)[L::Q]) {
  return moments_d3q27_t<L, float>(f);
}

// Smagorinsky tau from the second moments (as smagorinsky_tau_moments).
__device__ __forceinline__ float smagorinsky_tau_m2(const Moments2& m, float ux, float uy, float uz, float inv_rho,
                                                    float tau0, float cs) {
  const float rho = m.rho;
  const float r3 = rho * (1.0f / 3.0f);
  const float pxx = m.sxx - (r3 + rho * ux * ux);
  const float pyy = m.syy - (r3 + rho * uy * uy);
  const float pzz = m.szz - (r3 + rho * uz * uz);
  const float pxy = m.sxy - rho * ux * uy;
  const float pyz = m.syz - rho * uy * uz;
  const float pxz = m.sxz - rho * ux * uz;
  const float tensor_sum = (pxx * pxx + pyy * pyy + pzz * pzz) + 2.0f * (pxy * pxy + pyz * pyz + pxz * pxz);
  const float q_stress = __builtin_amdgcn_sqrtf(2.0f * tensor_sum);
  const float prefactor = 18.0f * 1.4142135623730951f * (cs * cs);
  const float disc = tau0 * tau0 + (prefactor * inv_rho) * q_stress;
  return 0.5f * (tau0 + __builtin_amdgcn_sqrtf(disc));
}

template <class L, bool FORCE>
__device__ __forceinline__ void trt_pairs(float (&f)[L::Q], float rho, float ux, float uy, float uz, float omega_s,
                                          float omega_a, float gx, float gy, float gz) {
  const float usq = ux * ux + uy * uy + uz * uz;
  const float K = 1.0f - 1.5f * usq;
  // With s = f_q + f_o and d = f_q - f_o, the pair relaxes to
  //   f_q' = S + A,  f_o' = S - A,
  //   S = (1 - omega_s) s / 2 + omega_s w rho (K + 4.5 (c u)^2),
  //   A = (1 - omega_a) d / 2 + 3 omega_a w rho (c u),
  // the per-shell factors omega_s w rho, its K and 4.5 multiples and
  // 3 omega_a w rho formed once per weight (three per node).
  const float as = 0.5f - 0.5f * omega_s, aa = 0.5f - 0.5f * omega_a;
  const float srho = omega_s * rho, arho = 3.0f * omega_a * rho;
  float pre_s = 0.0f, pre_a = 0.0f, ug = 0.0f;
  if constexpr (FORCE) {
    pre_s = 1.0f - 0.5f * omega_s;
    pre_a = 1.0f - 0.5f * omega_a;
    ug = ux * gx + uy * gy + uz * gz;
  }
  {
    const float wr = L::w[0] * rho;
    float v = f[0] - omega_s * (f[0] - wr * K);
    if constexpr (FORCE) v += pre_s * wr * (-3.0f * ug);
    f[0] = v;
  }
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int o = L::opp[q];
    if constexpr (q > 0 && q < o) {
      constexpr float w = L::w[q];
      const float ws = srho * w;
      const float beta = ws * K, gam = ws * 4.5f, del = arho * w;
      const float cu = cdot<L, q>(ux, uy, uz);
      const float s = f[q] + f[o];
      const float d = f[q] - f[o];
      float S = as * s + ((gam * cu) * cu + beta);
      float A = aa * d + del * cu;
      if constexpr (FORCE) {
        const float wr = w * rho;
        const float cg = cdot<L, q>(gx, gy, gz);
        A += pre_a * (3.0f * wr) * cg;
        S += pre_s * wr * (9.0f * cu * cg - 3.0f * ug);
      }
      f[q] = S + A;
      f[o] = S - A;
    }
  });
}

// TRT (Lambda = 1/4 pairing of tau_a) with the shear tau from the Smagorinsky
// closure (LES) or given.
template <class L, bool LES>
__device__ __forceinline__ Macro collide_trt_fast27(float (&f)[L::Q], float tau, float cs, float gx, float gy,
                                                    float gz) {
  const Moments2 m = moments_d3q27<L>(f);
  const float inv_rho = __builtin_amdgcn_rcpf(m.rho);
  const float ux = (m.jx + 0.5f * m.rho * gx) * inv_rho;
  const float uy = (m.jy + 0.5f * m.rho * gy) * inv_rho;
  const float uz = (m.jz + 0.5f * m.rho * gz) * inv_rho;
  const float tau_s = LES ? smagorinsky_tau_m2(m, ux, uy, uz, inv_rho, tau, cs) : tau;
  const float omega_s = __builtin_amdgcn_rcpf(tau_s);
  const float omega_a = __builtin_amdgcn_rcpf(trt_tau_a(tau_s));
  if (gx == 0.0f && gy == 0.0f && gz == 0.0f)
    trt_pairs<L, false>(f, m.rho, ux, uy, uz, omega_s, omega_a, gx, gy, gz);
  else
    trt_pairs<L, true>(f, m.rho, ux, uy, uz, omega_s, omega_a, gx, gy, gz);
  return {m.rho, ux, uy, uz};
}

// Dispatch one collision on a node.  tau_eff is the shear tau (the scalar tau
// of the step, or the per-node Smagorinsky value in LES-coupled mode).
template <class L, int COLL, bool REF_PAIRS = false, bool FAST = false>
__device__ __forceinline__ Macro collide(float (&f)[L::Q], float tau_eff, float gx, float gy, float gz) {
  if constexpr (L::Q == 27 && COLL != kBGK && FAST && !REF_PAIRS) {
    return collide_trt_fast27<L, false>(f, tau_eff, 0.0f, gx, gy, gz);
  } else if constexpr (COLL == kBGK) {
    return collide_bgk<L, FAST>(f, tau_eff, gx, gy, gz);
  } else if constexpr (COLL == kTRT || L::Q == 27) {
    return collide_trt<L, REF_PAIRS, FAST>(f, tau_eff, trt_tau_a(tau_eff), gx, gy, gz);
  } else if constexpr (L::Q == 9) {
    return collide_mrt_d2q9<false, FAST>(f, tau_eff, kTauBulk2D, gx, gy);
  } else {
    return collide_mrt_d3q19<false, FAST>(f, tau_eff, kTauBulk3D, gx, gy, gz);
  }
}

// Collision with LES-coupled shear tau (tau0 is the base tau).  MRT reads the
// stress from its moments; BGK and TRT form it from feq.
template <class L, int COLL, bool FAST = false>
__device__ __forceinline__ Macro collide_les(float (&f)[L::Q], float tau0, float cs, float gx, float gy, float gz) {
  if constexpr (L::Q == 27 && COLL != kBGK && FAST) {
    return collide_trt_fast27<L, true>(f, tau0, cs, gx, gy, gz);
  } else if constexpr (COLL == kMRT && L::Q == 19) {
    return collide_mrt_d3q19<true, FAST>(f, tau0, kTauBulk3D, gx, gy, gz, cs);
  } else if constexpr (COLL == kMRT && L::Q == 9) {
    return collide_mrt_d2q9<true, FAST>(f, tau0, kTauBulk2D, gx, gy, cs);
  } else {
    float rho, jx, jy, jz, ux, uy, uz;
    density_momentum<L>(f, rho, jx, jy, jz);
    guo_velocity<FAST>(rho, jx, jy, jz, gx, gy, gz, ux, uy, uz);
    return collide<L, COLL, false, FAST>(f, smagorinsky_tau_moments<L, FAST>(f, rho, ux, uy, uz, tau0, cs), gx, gy,
                                         gz);
  }
}

}  // namespace wake_hip
