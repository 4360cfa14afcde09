// thermal_impl.hpp -- D2Q5 temperature coupled to the D2Q9 flow (device).
//
// No reference counterpart: the reference's ThermalSolver2D
// (src/thermal/thermal_solver.cpp:83-123) advects a D2Q9 scalar whose
// buoyancy reaches the flow only through the mean temperature.  Here each
// node carries five temperature populations g_q (c = 0, +x, +y, -x, -y),
// relaxed by BGK to geq_q = w_q T (1 + 3 c_q . u), and the flow collision
// takes a per-node Boussinesq acceleration a_y = gbeta (T - t_ref) on top of
// the body force.  T is the zeroth moment of the post-stream g; u in geq is
// the Guo velocity (j + rho a / 2) / rho of the post-stream f.  A pull source
// that is solid, a moving wall, or outside a non-periodic y range reflects g:
// adiabatic bounce-back g_q = g*_opp, or anti-bounce-back
// g_q = -g*_opp + 2 w_q T_w where the solid node carries a wall temperature.
#pragma once

#include "step_impl.hpp"

namespace wake_hip {

// S of the temperature populations of node (x, y) from the post-collision G.
__device__ __forceinline__ void gather_g(const StepArgs& a, const float* __restrict__ G, int x, int y, int n,
                                         uint8_t c, float (&g)[D2Q5::Q]) {
  const int nx = a.nx, ny = a.ny;
  const long long S = a.stride;
  const int xm = x == 0 ? nx - 1 : x - 1;
  const int xp = x == nx - 1 ? 0 : x + 1;
  static_for<D2Q5::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int cx = D2Q5::cx[q], cy = D2Q5::cy[q];
    const int sx = cx == 0 ? x : (cx > 0 ? xm : xp);
    const int sy = cy == 0 ? y : wrap(y - cy, ny);
    g[q] = G[q * S + static_cast<long long>(sy * nx + sx)];
  });
  if (!(c & (kClsFrozen | kClsSlow))) return;
  if (c & kClsFrozen) {
    static_for<D2Q5::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      g[q] = G[q * S + n];
    });
    return;
  }
  static_for<D2Q5::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (q > 0) {
      constexpr int cx = D2Q5::cx[q], cy = D2Q5::cy[q], o = D2Q5::opp[q];
      const int sx = cx == 0 ? x : (cx > 0 ? xm : xp);
      int sy = y - cy;
      bool in = true;
      if constexpr (cy != 0) {
        if (a.per_y) sy = wrap(sy, ny);
        else in = sy >= 0 && sy < ny;
      }
      const int sn = in ? sy * nx + sx : n;
      const uint8_t t = (in && !a.periodic_mode) ? a.flags[sn] : static_cast<uint8_t>(kFluid);
      const bool wall = !in || t == kSolid || t == kMovingWall;
      float v = g[q];
      if (wall) {
        const float tw = in ? a.t_wall[sn] : __builtin_nanf("");
        const float r = G[o * S + n];
        v = (tw != tw) ? r : -r + 2.0f * D2Q5::w[q] * tw;
      }
      g[q] = v;
    }
  });
}

// Temperature and the node's total acceleration a_y.
__device__ __forceinline__ float thermal_t(const float (&g)[D2Q5::Q]) {
  float T = 0.0f;
  static_for<D2Q5::Q>([&](auto qc) { T += g[decltype(qc)::value]; });
  return T;
}

// Collide a fluid node's f (per-node Boussinesq force) and g.
template <int COLL>
__device__ __forceinline__ void thermal_collide(const StepArgs& a, float (&f)[D2Q9::Q], float (&g)[D2Q5::Q]) {
  const float T = thermal_t(g);
  const float gyn = a.gy + a.gbeta * (T - a.t_ref);
  float rho = 0.0f, jx = 0.0f, jy = 0.0f;
  static_for<D2Q9::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    rho += f[q];
    jx += static_cast<float>(D2Q9::cx[q]) * f[q];
    jy += static_cast<float>(D2Q9::cy[q]) * f[q];
  });
  flag_nonfinite(a, rho + T);
  const float inv_rho = 1.0f / rho;
  const float ux = (jx + 0.5f * rho * a.gx) * inv_rho;
  const float uy = (jy + 0.5f * rho * gyn) * inv_rho;
  collide<D2Q9, COLL>(f, a.tau, a.gx, gyn, 0.0f);
  const float omega_g = 1.0f / a.tau_g;
  static_for<D2Q5::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const float cu = static_cast<float>(D2Q5::cx[q]) * ux + static_cast<float>(D2Q5::cy[q]) * uy;
    const float geq = D2Q5::w[q] * T * (1.0f + 3.0f * cu);
    g[q] = g[q] - omega_g * (g[q] - geq);
  });
}

// One node of the coupled step: S of f and g (pulled from A, or read from a
// materialised S), collision, store.
template <class St, int COLL>
__device__ __forceinline__ void thermal_node(const StepArgs& a, const typename St::T* __restrict__ ain,
                                             typename St::T* __restrict__ aout, const float* __restrict__ gin,
                                             float* __restrict__ gout, int x, int y, int n, bool pull,
                                             unsigned long long s_step) {
  const uint8_t c = a.cls[n];
  const long long S = a.stride;
  float f[D2Q9::Q], g[D2Q5::Q];
  if (pull) {
    form_s_at<D2Q9, St>(a, ain, x, y, 0, n, c, true, s_step, f);
    gather_g(a, gin, x, y, n, c, g);
  } else {
    static_for<D2Q9::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      f[q] = St::template ld<D2Q9, q>(ain, q * S + n);
    });
    static_for<D2Q5::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      g[q] = gin[q * S + n];
    });
  }
  if (!(c & kClsSolid)) thermal_collide<COLL>(a, f, g);
  static_for<D2Q9::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    St::template st<D2Q9, q>(aout, q * S + n, f[q]);
  });
  static_for<D2Q5::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    __builtin_nontemporal_store(g[q], gout + q * S + n);
  });
}

}  // namespace wake_hip
