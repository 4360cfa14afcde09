// step_impl.hpp -- fused LBM time step for gfx950 (device-resident path).
//
// One lane per node, lanes contiguous along x, one work-group row per (y, z):
// every pull load of a wavefront is 64 consecutive elements of one
// population plane (a +-1 element shift for c_x = -+1), so the 19 (or 9, 27)
// plane streams are fully coalesced and each population element of the
// previous step is read exactly once.  y and z are work-group uniform, so the
// domain-edge logic is scalar.  Bulk nodes take a branch-free pull; nodes
// next to walls, solids, curved links or open boundaries (flagged once per
// geometry change in the class byte) take the general path that restates the
// reference's separate streaming, BFL and Zou-He passes:
//   d3q19_streaming.ispc:39-104, d2q9_streaming.ispc:33-95 (wall streaming)
//   d3q19_streaming.ispc:4-35 (periodic streaming, cell types ignored)
//   d3q19_curved.ispc:7-71 (BFL, read from the post-collision buffer)
//   d2q9_boundaries.ispc:43-121 (Zou-He inflow / outflow)
// followed by the collision of lattice_grid_3d.cpp:313-339.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine.hpp"
#include "lattice.hpp"
#include "node_ops.hpp"

// WAKE_MRT_MFMA=1 builds the D3Q19 MRT step with 16-bit storage with its
// moment transforms on the matrix cores (mrt_mfma.hpp); off by default.
#ifndef WAKE_MRT_MFMA
#define WAKE_MRT_MFMA 0
#endif
#if WAKE_MRT_MFMA
#include "mrt_mfma.hpp"
#endif

namespace wake_hip {

// ----------------------------------------------------------------- storage
// Code-generation knobs of the step kernel (same work, different schedule):
// WAKE_NT_LOADS marks the population loads non-temporal; WAKE_WAVES_PER_EU
// asks the compiler for that many waves per SIMD (register budget).
#ifndef WAKE_NT_LOADS
#define WAKE_NT_LOADS 0
#endif
#ifndef WAKE_WAVES_PER_EU
#define WAKE_WAVES_PER_EU 0
#endif
// WAKE_NT_STORES = 0 makes the population stores ordinary (cached) stores.
#ifndef WAKE_NT_STORES
#define WAKE_NT_STORES 1
#endif

template <class T>
__device__ __forceinline__ void st_pop(T* __restrict__ p, long long i, T v) {
#if WAKE_NT_STORES
  __builtin_nontemporal_store(v, p + i);
#else
  p[i] = v;
#endif
}

template <class T>
__device__ __forceinline__ T ld_pop(const T* __restrict__ p, long long i) {
#if WAKE_NT_LOADS
  return __builtin_nontemporal_load(p + i);
#else
  return p[i];
#endif
}

struct StF32 {
  using T = float;
  static constexpr bool kFast = false;  // correctly rounded reciprocals and roots
  static constexpr bool kBuf = false;   // addressed through global pointers
  __device__ __forceinline__ static float from_bits(unsigned b) { return __uint_as_float(b); }
  __device__ __forceinline__ static unsigned to_bits(float d) { return __float_as_uint(d); }
  template <class L, int q>
  __device__ __forceinline__ static float ld(const T* __restrict__ p, long long i) {
    return ld_pop(p, i);
  }
  template <class L, int q>
  __device__ __forceinline__ static void st(T* __restrict__ p, long long i, float v) { st_pop(p, i, v); }
};

// FP16 <-> FP32 with the weight offset in one mixed-precision FMA each:
// w + (float)h (exact conversion, one rounding: the same value as a
// conversion followed by an add) and (_Float16)(v - w) (one rounding of the
// exact difference).  h holds FP16 bits in its low half.
__device__ __forceinline__ float f16_bits_plus(unsigned h, float w) {
  float r;
  asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "s"(w));
  return r;
}
__device__ __forceinline__ unsigned short f16_bits_of_minus(float v, float w) {
  unsigned r;
  asm("v_fma_mixlo_f16 %0, %1, 1.0, %2" : "=v"(r) : "v"(v), "s"(-w));
  return static_cast<unsigned short>(r);
}

// FP16 storage of the deviation f_q - w_q; all arithmetic stays FP32.
struct StF16 {
  using T = _Float16;
  static constexpr bool kFast = true;  // 1-ulp reciprocals and roots (storage rounds far more)
  static constexpr bool kBuf = false;
  __device__ __forceinline__ static float from_bits(unsigned short h) {
    return static_cast<float>(__builtin_bit_cast(_Float16, h));
  }
  __device__ __forceinline__ static unsigned short to_bits(float d) {
    return __builtin_bit_cast(unsigned short, static_cast<_Float16>(d));
  }
  template <class L, int q>
  __device__ __forceinline__ static float ld(const T* __restrict__ p, long long i) {
    return f16_bits_plus(__builtin_bit_cast(unsigned short, ld_pop(p, i)), L::w[q]);
  }
  template <class L, int q>
  __device__ __forceinline__ static void st(T* __restrict__ p, long long i, float v) {
    st_pop(p, i, __builtin_bit_cast(_Float16, f16_bits_of_minus(v, L::w[q])));
  }
  template <class L, int q>
  __device__ __forceinline__ static float from_dev_bits(unsigned short h) {
    return f16_bits_plus(h, L::w[q]);
  }
  template <class L, int q>
  __device__ __forceinline__ static unsigned short dev_bits(float v) {
    return f16_bits_of_minus(v, L::w[q]);
  }
};

// BF16 storage of the deviation f_q - w_q (round to nearest even).
struct StBF16 {
  using T = uint16_t;
  static constexpr bool kFast = true;
  static constexpr bool kBuf = false;
  __device__ __forceinline__ static float from_bits(unsigned short h) {
    return __uint_as_float(static_cast<uint32_t>(h) << 16);
  }
  // round to nearest even: the hardware conversion (v_cvt_pk_bf16_f32)
  __device__ __forceinline__ static unsigned short to_bits(float d) {
    return __builtin_bit_cast(unsigned short, static_cast<__bf16>(d));
  }
  template <class L, int q>
  __device__ __forceinline__ static float ld(const T* __restrict__ p, long long i) {
    return L::w[q] + __uint_as_float(static_cast<uint32_t>(ld_pop(p, i)) << 16);
  }
  template <class L, int q>
  __device__ __forceinline__ static void st(T* __restrict__ p, long long i, float v) {
    st_pop(p, i, static_cast<uint16_t>(to_bits(v - L::w[q])));
  }
  template <class L, int q>
  __device__ __forceinline__ static float from_dev_bits(unsigned short h) {
    return L::w[q] + from_bits(h);
  }
  template <class L, int q>
  __device__ __forceinline__ static unsigned short dev_bits(float v) {
    return to_bits(v - L::w[q]);
  }
};

// The 16-bit storages addressed through a buffer descriptor (population
// buffers under 4 GiB): a 32-bit lane offset and a scalar offset per access,
// so the bulk pull and the store need no per-lane 64-bit address arithmetic.
// The descriptor's extent bounds every access to the buffer.
template <class Base>
struct StBuf : Base {
  static constexpr bool kBuf = true;
  // raw descriptor word 3 of the gfx9 family (32-bit data format, no swizzle)
  static constexpr int kRsrcWord3 = 0x00020000;
  static constexpr int kNonTemporal = WAKE_NT_STORES ? 2 : 0;  // cache-policy bit "nt"
  __device__ __forceinline__ static __amdgpu_buffer_rsrc_t rsrc(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), kRsrcWord3);
  }
  template <class L, int q>
  __device__ __forceinline__ static float ldb(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    if constexpr (sizeof(typename Base::T) == 4) {
      return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, static_cast<int>(voff), static_cast<int>(soff), 0));
    } else {
      const unsigned short h =
          __builtin_amdgcn_raw_buffer_load_b16(r, static_cast<int>(voff), static_cast<int>(soff), 0);
      return Base::template from_dev_bits<L, q>(h);
    }
  }
  template <class L, int q>
  __device__ __forceinline__ static void stb(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float v) {
    if constexpr (sizeof(typename Base::T) == 4) {
      __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, static_cast<int>(voff), static_cast<int>(soff),
                                            kNonTemporal);
    } else {
      __builtin_amdgcn_raw_buffer_store_b16(Base::template dev_bits<L, q>(v), r, static_cast<int>(voff),
                                            static_cast<int>(soff), kNonTemporal);
    }
  }
};
using StF32Buf = StBuf<StF32>;
using StF16Buf = StBuf<StF16>;
using StBF16Buf = StBuf<StBF16>;

// -------------------------------------------------------------- divergence
// A non-finite density sets the domain's divergence flag (an idempotent
// store; the branch is taken by no lane of a healthy flow).
__device__ __forceinline__ void flag_nonfinite(const StepArgs& a, float rho) {
  if (a.nonfinite && !__builtin_isfinite(rho)) *a.nonfinite = 1;
}

// ------------------------------------------------------------- RNG (inlet)
// splitmix64 finaliser of seed ^ counter; top 24 bits as a uniform in [0,1).
__device__ __forceinline__ float rng_sym(unsigned long long seed, unsigned long long counter) {
  unsigned long long x = seed ^ counter;
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x = x ^ (x >> 31);
  // 2 U - 1, exact in FP32
  return static_cast<float>(static_cast<uint32_t>(x >> 40)) * (1.0f / 8388608.0f) - 1.0f;
}

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// A wave-uniform value (row indices, sizes, the plane stride) as an opaque
// scalar operand.  The bulk pull forms its offsets from such copies,
// so the compiler cannot share them with the slow paths' address arithmetic:
// shared, they stay live across the whole node update and spill to VGPR
// lanes (a v_writelane per value on every node, VALU issue the 16-bit
// kernels are short of).  Only for wave-uniform values: gather takes this
// path only when its caller's lanes hold nodes of one row (UNIFORM_ROW).
__device__ __forceinline__ int opaque_uniform(int v) {
  v = __builtin_amdgcn_readfirstlane(v);
  asm("" : "+s"(v));
  return v;
}
__device__ __forceinline__ long long opaque_uniform(long long v) {
  const int lo = opaque_uniform(static_cast<int>(v));
  const int hi = opaque_uniform(static_cast<int>(v >> 32));
  return static_cast<long long>((static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
                                static_cast<unsigned>(lo));
}

// ------------------------------------------------------------------ gather
// S of node (x, y, z) from the post-collision buffer A: streaming with the
// reference's bounce-back rules, then the BFL overrides.
template <class L, class St, bool UNIFORM_ROW = false>
__device__ __forceinline__ void gather(const StepArgs& a, const typename St::T* __restrict__ A, int x, int y,
                                       int z, int n, uint8_t c, float (&f)[L::Q]) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const long long S = a.stride;
  const int xm = x == 0 ? nx - 1 : x - 1;
  const int xp = x == nx - 1 ? 0 : x + 1;
  // Bulk pull, issued for every node together with the class byte (wrapped
  // addresses are always in bounds); special nodes patch it below, so these
  // loads never wait on the class byte.
  if constexpr (St::kBuf) {
    // byte offsets: the column per lane, the plane and row per work-group row
    const auto r = St::rsrc(A, static_cast<unsigned>(L::Q * S * sizeof(typename St::T)));
    constexpr unsigned E = sizeof(typename St::T);
    const unsigned vx = static_cast<unsigned>(x) * E, vm = static_cast<unsigned>(xm) * E,
                   vp = static_cast<unsigned>(xp) * E;
    constexpr bool kOpaque = UNIFORM_ROW;
    const int by = kOpaque ? opaque_uniform(y) : y, bz = kOpaque ? opaque_uniform(z) : z;
    const int bnx = kOpaque ? opaque_uniform(nx) : nx, bny = kOpaque ? opaque_uniform(ny) : ny;
    const int bnz = kOpaque ? opaque_uniform(nz) : nz;
    const long long bS = kOpaque ? opaque_uniform(S) : S;
    static_for<L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int cx = L::cx[q], cy = L::cy[q], cz = L::cz[q];
      const int sy = cy == 0 ? by : wrap(by - cy, bny);
      const int sz = cz == 0 ? bz : wrap(bz - cz, bnz);
      const unsigned so = static_cast<unsigned>(q * bS + static_cast<long long>((sz * bny + sy) * bnx)) * E;
      f[q] = St::template ldb<L, q>(r, cx == 0 ? vx : (cx > 0 ? vm : vp), so);
    });
  } else {
    // (the D3Q27 pointer-addressed kernel keeps the shared offsets: opaque
    // copies raise it from 90 to 132 VGPRs, 5 to 3 waves per SIMD)
    constexpr bool kOpaque = UNIFORM_ROW && L::Q <= 19;
    const int by = kOpaque ? opaque_uniform(y) : y, bz = kOpaque ? opaque_uniform(z) : z;
    const int bnx = kOpaque ? opaque_uniform(nx) : nx, bny = kOpaque ? opaque_uniform(ny) : ny;
    const int bnz = kOpaque ? opaque_uniform(nz) : nz;
    const long long bS = kOpaque ? opaque_uniform(S) : S;
    static_for<L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int cx = L::cx[q], cy = L::cy[q], cz = L::cz[q];
      const int sx = cx == 0 ? x : (cx > 0 ? xm : xp);
      const int sy = cy == 0 ? by : wrap(by - cy, bny);
      const int sz = cz == 0 ? bz : wrap(bz - cz, bnz);
      f[q] = St::template ld<L, q>(A, q * bS + static_cast<long long>((sz * bny + sy) * bnx + sx));
    });
  }
  if (!(c & (kClsFrozen | kClsSlow | kClsBfl))) return;
  if (c & kClsFrozen) {
    static_for<L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      f[q] = St::template ld<L, q>(A, q * S + n);
    });
  } else if (c & kClsSlow) {
    // Cell type of the pull source along c_q (Fluid when out of range).
    auto source = [&](auto qc, bool& in) -> uint8_t {
      constexpr int q = decltype(qc)::value;
      constexpr int cx = L::cx[q], cy = L::cy[q], cz = L::cz[q];
      const int sx = cx == 0 ? x : (cx > 0 ? xm : xp);
      int sy = y - cy, sz = z - cz;
      in = true;
      if constexpr (cy != 0) {
        if (a.per_y) sy = wrap(sy, ny);
        else in = in && sy >= 0 && sy < ny;
      }
      if constexpr (cz != 0) {
        if (a.per_z) sz = wrap(sz, nz);
        else in = in && sz >= 0 && sz < nz;
      }
      const int sn = in ? (sz * ny + sy) * nx + sx : n;
      return (in && !a.periodic_mode) ? a.flags[sn] : static_cast<uint8_t>(kFluid);
    };
    // Halfway bounce-back (domain wall, solid or moving-wall source) replaces
    // the pulled value with the node's own opposite population; a moving wall
    // adds 2 w_q rho0 3 (c_q . u_w), whose products with c in {-1, 0, 1} are
    // exact, so it rounds as the host's left-to-right expression.
    auto moving_wall_term = [&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const float ci_uw = static_cast<float>(L::cx[q]) * a.uw[0] + static_cast<float>(L::cy[q]) * a.uw[1] +
                          static_cast<float>(L::cz[q]) * a.uw[2];
      return 2.0f * L::w[q] * 1.0f * 3.0f * ci_uw;
    };
    if (a.bmask) {
      // the bounce mask built with the class bytes: no cell-type reads
      const uint32_t bm = a.bmask[n];
      static_for<L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (q > 0) {
          if (bm & (1u << q)) {
            float v = St::template ld<L, L::opp[q]>(A, L::opp[q] * S + n);
            if (bm & kBounceMovingWall) {
              bool in;
              if (source(qc, in) == kMovingWall) v += moving_wall_term(qc);
            }
            f[q] = v;
          }
        }
      });
    } else {
      static_for<L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (q > 0) {
          bool in;
          const uint8_t t = source(qc, in);
          if (!in || t == kSolid || t == kMovingWall) {
            float v = St::template ld<L, L::opp[q]>(A, L::opp[q] * S + n);
            if (t == kMovingWall) v += moving_wall_term(qc);
            f[q] = v;
          }
        }
      });
    }
  }
  if (c & kClsBfl) {
    const BflNode* r = a.bfl + a.aux[n];
    const uint32_t mask = r->mask;
    static_for<L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      if constexpr (q > 0) {
        constexpr int cx = L::cx[q], cy = L::cy[q], cz = L::cz[q], o = L::opp[q];
        float val = f[o];
        if (mask & (1u << q)) {
          const float delta = r->delta[q];
          const float ciuw = r->ciuw[q];
          const float f_i_xf = St::template ld<L, q>(A, q * S + n);
          if (delta < 0.5f) {
            const int xf = cx == 0 ? x : (cx > 0 ? xm : xp);
            int yf = y - cy, zf = z - cz;
            if (a.bfl_per_y) yf = wrap(yf, ny);
            if (a.bfl_per_z) zf = wrap(zf, nz);
            const bool inside = yf >= 0 && yf < ny && zf >= 0 && zf < nz;
            const float f_i_xff = inside ? St::template ld<L, q>(A, q * S + static_cast<long long>(
                                                                        ((inside ? zf : z) * ny + (inside ? yf : y)) * nx + xf))
                                         : f_i_xf;
            const float mom = 2.0f * L::w[q] * 1.0f * 3.0f * ciuw;
            val = 2.0f * delta * f_i_xf + (1.0f - 2.0f * delta) * f_i_xff - mom;
          } else {
            const float f_opp_xf = St::template ld<L, o>(A, o * S + n);
            const float inv_2delta = 1.0f / (2.0f * delta);
            const float mom = (1.0f / delta) * L::w[q] * 1.0f * 3.0f * ciuw;
            val = inv_2delta * f_i_xf + (1.0f - inv_2delta) * f_opp_xf - mom;
          }
        }
        f[o] = val;
      }
    });
  }
}

// --------------------------------------------------------- open boundaries
// Multiply-adds are fused per source expression only ("contract on"): with
// cross-statement fusion the inlet's fused pairs depended on the code around
// the call, so kernels that differ only in addressing rounded it differently.
template <class L>
__device__ __forceinline__ void apply_inlet(const StepArgs& a, float (&f)[L::Q], int y, int n,
                                            unsigned long long s_step) {
#pragma clang fp contract(on)
  float ux = a.uin_x;
  if (a.inflow_parabolic) {
    const float s = static_cast<float>(y) - 0.5f;
    const float H = static_cast<float>(a.ny - 2);
    ux = 4.0f * a.umax * s * (H - s) / (H * H);
  }
  if (a.turb_amp != 0.0f) {
    const unsigned long long N = static_cast<unsigned long long>(a.nx) * a.ny * a.nz;
    ux = ux * (1.0f + a.turb_amp * rng_sym(a.turb_seed, s_step * N + static_cast<unsigned long long>(n)));
  }
  const float uy = a.uin_y;
  if constexpr (L::Q == 9) {
    // d2q9_boundaries.ispc:43-81
    const float rho_in = (f[0] + f[2] + f[4] + 2.0f * (f[3] + f[6] + f[7])) / (1.0f - ux);
    const float f1 = f[3] + (2.0f / 3.0f) * rho_in * ux;
    const float f5 = f[7] - 0.5f * (f[2] - f[4]) + (1.0f / 6.0f) * rho_in * ux + 0.5f * rho_in * uy;
    const float f8 = f[6] + 0.5f * (f[2] - f[4]) + (1.0f / 6.0f) * rho_in * ux - 0.5f * rho_in * uy;
    f[1] = f1;
    f[5] = f5;
    f[8] = f8;
  } else {
    // 3D: regularized Zou-He (non-equilibrium bounce-back completion with
    // equal transverse corrections, then projection on feq + second order)
    float s0 = 0.0f, sm = 0.0f, ty = 0.0f, tz = 0.0f;
    int cnt_y = 0, cnt_z = 0;
    static_for<L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      if constexpr (L::cx[q] == 0) {
        s0 += f[q];
        if constexpr (L::cy[q] == 1) ty += f[q];
        if constexpr (L::cy[q] == -1) ty -= f[q];
        if constexpr (L::cz[q] == 1) tz += f[q];
        if constexpr (L::cz[q] == -1) tz -= f[q];
      } else if constexpr (L::cx[q] < 0) {
        sm += f[q];
      } else {
        cnt_y += L::cy[q] * L::cy[q];
        cnt_z += L::cz[q] * L::cz[q];
      }
    });
    const float uz = 0.0f;
    const float rho = (s0 + 2.0f * sm) / (1.0f - ux);
    const float Ny = (ty - (2.0f / 3.0f) * rho * uy) / static_cast<float>(cnt_y);
    const float Nz = (tz - (2.0f / 3.0f) * rho * uz) / static_cast<float>(cnt_z);
    // Non-equilibrium bounce-back completion of the c_x = +1 populations,
    // g_p = f_opp(p) + 6 w rho (c_p . u) - (c_py N_y + c_pz N_z), then the
    // regularization f = feq + 4.5 w (c c - I/3) : Pi_neq with
    // Pi_neq = sum_q c c (g - feq).  Both are evaluated per opposite pair:
    // c c is even in c, so only the pair sums of g and feq enter Pi_neq and
    // both members share (c c - I/3) : Pi_neq; the pair's feq sum is
    // 2 w rho (1 - 1.5 u^2 + 4.5 (c.u)^2) and its difference 6 w rho (c.u).
    const float usq = ux * ux + uy * uy + uz * uz;
    const float K = 1.0f - 1.5f * usq;
    float fe_s[L::Q], fe_d[L::Q];
    float pxx = 0.0f, pyy = 0.0f, pzz = 0.0f, pxy = 0.0f, pyz = 0.0f, pxz = 0.0f;
    static_for<L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int o = L::opp[q];
      if constexpr (q > 0 && q < o) {
        constexpr int a = L::cx[q], b = L::cy[q], c = L::cz[q];
        const float wr = L::w[q] * rho;
        const float cu = cdot<L, q>(ux, uy, uz);
        fe_s[q] = (2.0f * wr) * ((4.5f * cu) * cu + K);
        fe_d[q] = (6.0f * wr) * cu;
        float sg;  // g_q + g_o
        if constexpr (a == 1) {
          sg = 2.0f * f[o] + (fe_d[q] - (cmul<b>(Ny) + cmul<c>(Nz)));
        } else if constexpr (a == -1) {
          sg = 2.0f * f[q] + (-fe_d[q] - (cmul<-b>(Ny) + cmul<-c>(Nz)));
        } else {
          sg = f[q] + f[o];
        }
        const float ns = sg - fe_s[q];
        if constexpr (a != 0) pxx += ns;
        if constexpr (b != 0) pyy += ns;
        if constexpr (c != 0) pzz += ns;
        if constexpr (a * b != 0) pxy += cmul<a * b>(ns);
        if constexpr (b * c != 0) pyz += cmul<b * c>(ns);
        if constexpr (a * c != 0) pxz += cmul<a * c>(ns);
      }
    });
    const float tr3 = (pxx + pyy + pzz) * (1.0f / 3.0f);
    {
      const float wr = L::w[0] * rho;
      f[0] = wr * K - (4.5f * L::w[0]) * tr3;
    }
    static_for<L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int o = L::opp[q];
      if constexpr (q > 0 && q < o) {
        constexpr int a = L::cx[q], b = L::cy[q], c = L::cz[q];
        float d = -tr3;
        if constexpr (a != 0) d += pxx;
        if constexpr (b != 0) d += pyy;
        if constexpr (c != 0) d += pzz;
        float od = 0.0f;
        if constexpr (a * b != 0) od += cmul<a * b>(pxy);
        if constexpr (b * c != 0) od += cmul<b * c>(pyz);
        if constexpr (a * c != 0) od += cmul<a * c>(pxz);
        const float t = d + 2.0f * od;
        const float base = 0.5f * fe_s[q] + (4.5f * L::w[q]) * t;
        f[q] = base + 0.5f * fe_d[q];
        f[o] = base - 0.5f * fe_d[q];
      }
    });
  }
}

// d2q9_boundaries.ispc:85-121 (pressure outlet, u_y = 0).
__device__ __forceinline__ void apply_zh_outflow_d2q9(const StepArgs& a, float (&f)[9]) {
  const float rho_out = a.rho_out;
  const float u_out_x = -1.0f + (f[0] + f[2] + f[4] + 2.0f * (f[1] + f[5] + f[8])) / rho_out;
  const float u_out_y = 0.0f;
  const float f3 = f[1] - (2.0f / 3.0f) * rho_out * u_out_x;
  const float f6 = f[8] - 0.5f * (f[2] - f[4]) - (1.0f / 6.0f) * rho_out * u_out_x + 0.5f * rho_out * u_out_y;
  const float f7 = f[5] + 0.5f * (f[2] - f[4]) - (1.0f / 6.0f) * rho_out * u_out_x - 0.5f * rho_out * u_out_y;
  f[3] = f3;
  f[6] = f6;
  f[7] = f7;
}

// Full post-stream state S_k of a node (k = s_step): gather + open
// boundaries.  The convective outlet blends its previous post-stream state
// (the outlet store, one 16-byte aligned record of prev_stride<Q> floats per
// (y, z) row, read and written with vector accesses) with the post-stream
// state of its upstream neighbour x - 1.  With ROW_LANES (lanes hold
// consecutive nodes of one row, as in every row- or tile-shaped kernel) that
// state is taken from lane - 1 by a cross-lane shuffle of the values the
// neighbour has just formed, and the bulk pull forms its offsets from opaque
// copies of the wave's row indices; otherwise (and for a lane 0 outlet) the
// outlet lane gathers it itself.
template <class L, class St, bool ROW_LANES = true>
__device__ __forceinline__ void form_s_at(const StepArgs& a, const typename St::T* __restrict__ A, int x, int y,
                                          int z, int n, uint8_t c, bool commit, unsigned long long s_step,
                                          float (&f)[L::Q]) {
  constexpr int QP = prev_stride(L::Q);
  float4* rec = reinterpret_cast<float4*>(a.f_prev + (static_cast<long long>(z) * a.ny + y) * QP);
  auto load_prev = [&](float (&prev)[QP]) {
    static_for<QP / 4>([&](auto vc) {
      constexpr int v = decltype(vc)::value;
      const float4 t = rec[v];
      prev[4 * v] = t.x;
      prev[4 * v + 1] = t.y;
      prev[4 * v + 2] = t.z;
      prev[4 * v + 3] = t.w;
    });
  };
  auto store_prev = [&](const float (&prev)[QP]) {
    static_for<QP / 4>([&](auto vc) {
      constexpr int v = decltype(vc)::value;
      rec[v] = make_float4(prev[4 * v], prev[4 * v + 1], prev[4 * v + 2], prev[4 * v + 3]);
    });
  };
  // With ROW_LANES the convective outlet's record is loaded together with
  // the pull: whether a wave holds an outlet lane follows from the lanes'
  // positions (the rule that sets kClsOutlet), so the load does not wait
  // for the class byte.  Every lane loads its own row's record (in bounds
  // for all lanes).
  float prev[QP];
  bool rec_loaded = false;
  if constexpr (ROW_LANES) {
    const bool out_pos = a.outflow == 2 && x == a.nx - 1 && y >= a.by0 && y < a.by1 && z >= a.bz0 && z < a.bz1;
    if (__any(out_pos)) {
      load_prev(prev);
      rec_loaded = true;
    }
  }
  gather<L, St, ROW_LANES>(a, A, x, y, z, n, c, f);
  if (c & kClsInlet) apply_inlet<L>(a, f, y, n, s_step);
  if constexpr (L::Q == 9) {
    if ((c & kClsOutlet) && a.outflow == 1) apply_zh_outflow_d2q9(a, f);
  }
  if (a.outflow != 2) return;
  const float alpha = a.conv_alpha, beta = a.conv_beta;
  const bool is_out = (c & kClsOutlet) != 0;
  bool done = false;
  if constexpr (ROW_LANES) {
    // wave-uniform: every lane takes part in the shuffles; only outlet lanes
    // keep results
    if (rec_loaded) {
      const bool use = is_out && __lane_id() != 0;
      static_for<L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const float v = beta * prev[q] + alpha * __shfl_up(f[q], 1);
        if (use) f[q] = v;
        prev[q] = v;
      });
      if (use && commit) store_prev(prev);
      done = use;
    }
  }
  if (!is_out || done) return;
  // the upstream neighbour is not in this wave: gather it
  float nb[L::Q];
  const uint8_t cn = a.cls[n - 1];
  gather<L, St, ROW_LANES>(a, A, x - 1, y, z, n - 1, cn, nb);
  if (cn & kClsInlet) apply_inlet<L>(a, nb, y, n - 1, s_step);
  load_prev(prev);  // (the shuffle pass above, if any, overwrote prev)
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const float v = beta * prev[q] + alpha * nb[q];
    f[q] = v;
    prev[q] = v;
  });
  if (commit) store_prev(prev);
}

template <class L, class St>
__device__ __forceinline__ void form_s(const StepArgs& a, const typename St::T* __restrict__ A, int x, int y, int z,
                                       int n, uint8_t c, bool commit, float (&f)[L::Q]) {
  form_s_at<L, St>(a, A, x, y, z, n, c, commit, a.s_step, f);
}

// ------------------------------------------------------ momentum exchange
// The momentum-exchange links of Fluid node (x, y, z): every q whose
// in-range neighbour x + c_q is Solid or MovingWall and whose link midpoint
// lies within r_max of the centre (r_max <= 0: all), as
// forces_momentum_exchange.ispc:71-138 walks them.  on_link(qc, rx, ry) is
// called per link with the midpoint offset from the centre.  The midpoint
// test rounds as the host's link builder (no fused multiply-add).  n is the
// node's index: with the bounce masks built, the walls come from its mask
// (bit opp(q) of an in-range neighbour), not from reading the neighbours'
// cell types.  Only for kClsSlow nodes (other nodes have no wall neighbour).
template <class L, class OnLink>
__device__ __forceinline__ void for_each_mem_link(const StepArgs& a, int x, int y, int z, int n, float ccx,
                                                  float ccy, float ccz, float r_max, OnLink&& on_link) {
  const float r_max_sq = r_max * r_max;
  const bool use_bm = a.bmask != nullptr;
  const uint32_t bm = use_bm ? a.bmask[n] : 0u;
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (q > 0) {
      constexpr int cx = L::cx[q], cy = L::cy[q], cz = L::cz[q];
      int ny_ = y + cy, nz_ = z + cz;
      bool in = true;
      if (a.per_y) ny_ = wrap(ny_, a.ny);
      else in = in && ny_ >= 0 && ny_ < a.ny;
      if (a.per_z) nz_ = wrap(nz_, a.nz);
      else in = in && nz_ >= 0 && nz_ < a.nz;
      bool wall;
      if (use_bm) {
        wall = in && ((bm >> L::opp[q]) & 1u);
      } else {
        const int nx_ = wrap(x + cx, a.nx);
        const uint8_t t = in ? a.flags[(nz_ * a.ny + ny_) * a.nx + nx_] : static_cast<uint8_t>(kFluid);
        wall = t == kSolid || t == kMovingWall;
      }
      if (wall) {
        const float rx = __fsub_rn(__fadd_rn(static_cast<float>(x), 0.5f * static_cast<float>(cx)), ccx);
        const float ry = __fsub_rn(__fadd_rn(static_cast<float>(y), 0.5f * static_cast<float>(cy)), ccy);
        const float rz = __fsub_rn(__fadd_rn(static_cast<float>(z), 0.5f * static_cast<float>(cz)), ccz);
        float d2 = __fadd_rn(__fmul_rn(rx, rx), __fmul_rn(ry, ry));
        if constexpr (L::D == 3) d2 = __fadd_rn(d2, __fmul_rn(rz, rz));
        if (r_max <= 0.0f || d2 <= r_max_sq) on_link(qc, rx, ry);
      }
    }
  });
}

// Whether node n is Fluid (from its bounce mask when built).
__device__ __forceinline__ bool is_fluid(const StepArgs& a, int n) {
  return a.bmask ? (a.bmask[n] & kBounceFluid) != 0 : a.flags[n] == kFluid;
}

// F += 2 f_q c_q over the links of one Fluid node on its S_k values f; in 2D
// the third sum is the torque about the centre at the link midpoint.
template <class L>
__device__ __forceinline__ void mem_node(const StepArgs& a, int x, int y, int z, int n, const float (&f)[L::Q],
                                         float ccx, float ccy, float ccz, float r_max, double& s0, double& s1,
                                         double& s2) {
  for_each_mem_link<L>(a, x, y, z, n, ccx, ccy, ccz, r_max, [&](auto qc, float rx, float ry) {
    constexpr int q = decltype(qc)::value;
    constexpr int cx = L::cx[q], cy = L::cy[q], cz = L::cz[q];
    const float dfx = 2.0f * f[q] * static_cast<float>(cx);
    const float dfy = 2.0f * f[q] * static_cast<float>(cy);
    s0 += dfx;
    s1 += dfy;
    if constexpr (L::D == 2) s2 += static_cast<double>(rx * dfy - ry * dfx);
    else s2 += static_cast<double>(2.0f * f[q] * static_cast<float>(cz));
  });
}

// Sum over the 64 lanes of a wave in a fixed butterfly order; lane 0's
// result is the same on every run.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// One wave's share of a force sample: lanes with want = true add their
// node's links (f = the node's S_k), lane 0 writes the wave's FP64 sums to
// the row's slot.  Every lane of the wave must take part (the sum is a
// cross-lane exchange); x is the lane's node column (lane 0: 64 w).
template <class L>
__device__ __forceinline__ void mem_wave(const StepArgs& a, bool want, int x, int y, int z, int n,
                                         const float (&f)[L::Q], int slot) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (__any(want)) {  // a wave without linked nodes writes zeros, as its sums are
    if (want) mem_node<L>(a, x, y, z, n, f, a.mem_c[0], a.mem_c[1], a.mem_c[2], a.mem_rmax, s0, s1, s2);
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
  }
  if (__lane_id() == 0 && x < a.nx) {
    double* p = a.mem_part + 3ll * (slot + x / 64);
    p[0] = s0;
    p[1] = s1;
    p[2] = s2;
  }
}

// ----------------------------------------------------------------- kernels
template <class L>
__global__ void k_build_class(const StepArgs a, uint8_t* __restrict__ cls, uint32_t* __restrict__ bmask) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= a.nx) return;
  const int n = (z * a.ny + y) * a.nx + x;
  const uint8_t t = a.flags[n];
  uint8_t c = 0;
  uint32_t mask = 0;
  if (t == kSolid) c |= a.periodic_mode ? kClsSolid : (kClsSolid | kClsFrozen);
  if (!a.periodic_mode && t != kSolid) {
    if (t == kFluid) mask |= kBounceFluid;
    static_for<L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      if constexpr (q > 0) {
        int sx = wrap(x - L::cx[q], a.nx), sy = y - L::cy[q], sz = z - L::cz[q];
        bool in = true;
        if (a.per_y) sy = wrap(sy, a.ny);
        else in = in && sy >= 0 && sy < a.ny;
        if (a.per_z) sz = wrap(sz, a.nz);
        else in = in && sz >= 0 && sz < a.nz;
        if (!in) mask |= 1u << q;
        else {
          const uint8_t nt = a.flags[(sz * a.ny + sy) * a.nx + sx];
          if (nt == kSolid || nt == kMovingWall) mask |= 1u << q;
          if (nt == kMovingWall) mask |= kBounceMovingWall;
        }
      }
    });
    if (mask & ~kBounceFluid) c |= kClsSlow;
  }
  if (bmask) bmask[n] = mask;
  const bool in_rows = y >= a.by0 && y < a.by1 && z >= a.bz0 && z < a.bz1;
  if (a.inflow && x == 0 && in_rows) c |= kClsInlet;
  if (a.outflow && x == a.nx - 1 && in_rows) c |= kClsOutlet;
  cls[n] = c;
}

// Waves per SIMD the register allocation must leave room for, per kernel
// instantiation (0: the compiler's choice).  The D3Q27 kernels with 16-bit
// storage are bound by arithmetic issue and need the occupancy: left alone,
// the BF16 one is scheduled into 97 VGPRs (4 waves per SIMD) and runs 13%
// slower than the FP16 one; with a floor of 5 it takes 71 VGPRs, no spills,
// and matches FP16 (the FP16 one is unchanged at 70).
#ifndef WAKE_Q27_HALF_WAVES
#define WAKE_Q27_HALF_WAVES 5
#endif
template <class L, class St>
struct StepWaves {
  static constexpr int value = 0;
};
// D3Q27 FP32 (82 VGPRs, 5 waves) is HBM-bound: a floor of 6 waves (80
// VGPRs, 2 spilled) measures the same, 7 (72, 10 spilled) is 9% slower.
#ifndef WAKE_Q27_F32_WAVES
#define WAKE_Q27_F32_WAVES 0
#endif
template <class St>
struct StepWaves<D3Q27, St> {
  static constexpr int value = sizeof(typename St::T) == 2 ? WAKE_Q27_HALF_WAVES : WAKE_Q27_F32_WAVES;
};

#if WAKE_WAVES_PER_EU > 0
#define WAKE_STEP_ATTR __attribute__((amdgpu_waves_per_eu(WAKE_WAVES_PER_EU, WAKE_WAVES_PER_EU)))
#else
#define WAKE_STEP_ATTR __attribute__((amdgpu_waves_per_eu(StepWaves<L, St>::value)))
#endif

// Nodes per thread along z (WAKE_ZK): the pull of node z+1 is issued before
// node z collides, so each lane keeps two nodes' loads in flight.
#ifndef WAKE_ZK
#define WAKE_ZK 1
#endif

template <class L, class St, int COLL, bool LES, bool PULL>
__global__ __launch_bounds__(256) WAKE_STEP_ATTR void k_step(const StepArgs a, const typename St::T* __restrict__ ain,
                                                             typename St::T* __restrict__ aout) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z0 = blockIdx.z * WAKE_ZK;
  // Force sample of the S_k this launch forms: the slot of this row, or -1
  // (rows without momentum-exchange links, launches without a sample).
  int mem_slot = -1;
  if constexpr (PULL && WAKE_ZK == 1) {
    if (a.mem_rows) mem_slot = a.mem_rows[z0 * a.ny + y];
  }
  const bool live = x < a.nx;
  if (!live && mem_slot < 0) return;
  const long long S = a.stride;
  auto fetch = [&](int z, float (&f)[L::Q], uint8_t& c, int& n) {
    n = (z * a.ny + y) * a.nx + x;
    c = a.cls[n];
    if constexpr (PULL) {
      form_s<L, St>(a, ain, x, y, z, n, c, true, f);
    } else {
      static_for<L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = St::template ld<L, q>(ain, q * S + n);
      });
    }
  };
  float fn[L::Q];
  uint8_t cn = 0;
  int nn = 0;
#if WAKE_MRT_MFMA
  constexpr bool kMfma = std::is_same_v<L, D3Q19> && COLL == kMRT && St::kFast;
  __shared__ float mfma_tile[kMfma ? 4 * kMfmaTile : 1];
#endif
  if (live) fetch(z0, fn, cn, nn);
  if constexpr (PULL && WAKE_ZK == 1) {
    if (mem_slot >= 0) {
      // every lane of the row's waves takes part in the sums
      mem_wave<L>(a, live && (cn & kClsSlow) && is_fluid(a, nn), x, y, z0, nn, fn, mem_slot);
      if (!live) return;
    }
  }
#pragma unroll
  for (int k = 0; k < WAKE_ZK; ++k) {
    const int z = z0 + k;
    if (k > 0 && z >= a.nz) break;
    float f[L::Q];
    static_for<L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      f[q] = fn[q];
    });
    const uint8_t c = cn;
    const int n = nn;
    if (k + 1 < WAKE_ZK && z + 1 < a.nz) fetch(z + 1, fn, cn, nn);
    bool collided = false;
#if WAKE_MRT_MFMA
    if constexpr (kMfma) {
      // the products need every lane of the wave; solid lanes compute and
      // keep their own values
      if (__builtin_amdgcn_read_exec() == ~0ull) {
        float g[L::Q];
        static_for<L::Q>([&](auto qc) { g[decltype(qc)::value] = f[decltype(qc)::value]; });
        const Macro m = collide_mrt_d3q19_mfma<LES, St::kFast>(g, mfma_tile + (threadIdx.x >> 6) * kMfmaTile, a.tau,
                                                               kTauBulk3D, a.gx, a.gy, a.gz, LES ? a.cs : 0.0f);
        if (!(c & kClsSolid)) {
          static_for<L::Q>([&](auto qc) { f[decltype(qc)::value] = g[decltype(qc)::value]; });
          flag_nonfinite(a, m.rho);
        }
        collided = true;
      }
    }
#endif
    if (!collided && !(c & kClsSolid)) {
      Macro m;
      if constexpr (LES) m = collide_les<L, COLL, St::kFast>(f, a.tau, a.cs, a.gx, a.gy, a.gz);
      else m = collide<L, COLL, false, St::kFast>(f, a.tau, a.gx, a.gy, a.gz);
      flag_nonfinite(a, m.rho);
    }
    if constexpr (St::kBuf) {
      const auto r = St::rsrc(aout, static_cast<unsigned>(L::Q * S * sizeof(typename St::T)));
      constexpr unsigned E = sizeof(typename St::T);
      const long long row = static_cast<long long>((z * a.ny + y) * a.nx);
      static_for<L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        St::template stb<L, q>(r, static_cast<unsigned>(x) * E, static_cast<unsigned>(q * S + row) * E, f[q]);
      });
    } else {
      static_for<L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        St::template st<L, q>(aout, q * S + n, f[q]);
      });
    }
  }
}

// ------------------------------------------- fully periodic D2Q9, T steps
// A fully periodic D2Q9 domain without solid, inlet or outlet nodes (every
// class byte zero) streams by a plain periodic shift, so kTbSteps steps can
// run in one launch: each work-group loads a (16 + 2 T)^2 tile of A_{k-1}
// around its 16 x 16 nodes into LDS, pull-collides T times in LDS (the
// valid region shrinks by one node per step), and stores A_{k-1+T} of its
// 16 x 16 nodes.  One thread per tile node (1,024 threads, 16 waves), so
// each SIMD holds four of the work-group's waves to overlap LDS latency.  Every node update is the same pull, collision and storage
// rounding as one k_step launch, so the result is the same as T launches;
// the launch cost and the HBM traffic are paid once per T steps.
inline constexpr int kTbTile = 16;

template <class St, class L, int q>
__device__ __forceinline__ float tb_get(typename St::T v) {
  if constexpr (sizeof(typename St::T) == 4) {
    return v;
  } else {
    return St::template from_dev_bits<L, q>(__builtin_bit_cast(unsigned short, v));
  }
}
template <class St, class L, int q>
__device__ __forceinline__ typename St::T tb_put(float v) {
  if constexpr (sizeof(typename St::T) == 4) {
    return v;
  } else {
    return __builtin_bit_cast(typename St::T, St::template dev_bits<L, q>(v));
  }
}

inline constexpr int kTbWidth = kTbTile + 2 * kTbSteps;  // 32
inline constexpr int kTbThreads = kTbWidth * kTbWidth;     // one per tile node

template <class St, int COLL, bool LES>
__global__ __launch_bounds__(kTbThreads) void k_step_tb2d(const StepArgs a, const typename St::T* __restrict__ ain,
                                                          typename St::T* __restrict__ aout) {
  using L = D2Q9;
  using T = typename St::T;
  constexpr int W = kTbWidth;
  static_assert((W & (W - 1)) == 0, "tile width");
  __shared__ T tile[2][L::Q][W * W];
  const int x0 = blockIdx.x * kTbTile - kTbSteps, y0 = blockIdx.y * kTbTile - kTbSteps;
  const long long S = a.stride;
  const int c = threadIdx.x;  // this thread's tile node
  const int i = c & (W - 1), j = c / W;
  {
    const int gx = ((x0 + i) % a.nx + a.nx) % a.nx;
    const int gy = ((y0 + j) % a.ny + a.ny) % a.ny;
    const long long n = static_cast<long long>(gy) * a.nx + gx;
    static_for<L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      tile[0][q][c] = ain[q * S + n];
    });
  }
  __syncthreads();
#pragma unroll 1
  for (int s = 1; s <= kTbSteps; ++s) {
    const int src = (s - 1) & 1, dst = s & 1;
    if (i >= s && i < W - s && j >= s && j < W - s) {
      float f[L::Q];
      static_for<L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = tb_get<St, L, q>(tile[src][q][(j - L::cy[q]) * W + (i - L::cx[q])]);
      });
      Macro m;
      if constexpr (LES) m = collide_les<L, COLL, St::kFast>(f, a.tau, a.cs, a.gx, a.gy, a.gz);
      else m = collide<L, COLL, false, St::kFast>(f, a.tau, a.gx, a.gy, a.gz);
      flag_nonfinite(a, m.rho);
      static_for<L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        tile[dst][q][c] = tb_put<St, L, q>(f[q]);
      });
    }
    __syncthreads();
  }
  const int li = i - kTbSteps, lj = j - kTbSteps;
  const int gx = blockIdx.x * kTbTile + li, gy = blockIdx.y * kTbTile + lj;
  if (li >= 0 && li < kTbTile && lj >= 0 && lj < kTbTile && gx < a.nx && gy < a.ny) {
    const long long n = static_cast<long long>(gy) * a.nx + gx;
    static_for<L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      aout[q * S + n] = tile[kTbSteps & 1][q][c];
    });
  }
}

// Convective outlet store from a materialised S state: record (y, z) = S at
// x = nx - 1 of that row.  One lane per row.
template <class L, class St>
__global__ __launch_bounds__(256) void k_prev_from_s(const StepArgs a, const typename St::T* __restrict__ s_state) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= a.ny * a.nz) return;
  constexpr int QP = prev_stride(L::Q);
  const long long n = static_cast<long long>(row) * a.nx + (a.nx - 1);
  float* rec = a.f_prev + static_cast<long long>(row) * QP;
  static_for<QP>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (q < L::Q) rec[q] = St::template ld<L, q>(s_state, q * a.stride + n);
    else rec[q] = 0.0f;
  });
}

// S_k into the state buffer (storage format), committing the convective store.
template <class L, class St>
__global__ __launch_bounds__(256) void k_materialize(const StepArgs a, const typename St::T* __restrict__ ain,
                                                     typename St::T* __restrict__ sout) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= a.nx) return;
  const int n = (z * a.ny + y) * a.nx + x;
  float f[L::Q];
  form_s<L, St>(a, ain, x, y, z, n, a.cls[n], true, f);
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    St::template st<L, q>(sout, q * a.stride + n, f[q]);
  });
}

// Moments of S_k with the Guo shift (d3q19_moments.ispc:35-82): rho, u, |u|
// (skipped when a.rho is null) and optionally the S planes in the reference
// layout (stride N, fp32).
template <class L, class St, bool FROM_S>
__global__ __launch_bounds__(256) void k_readback(const StepArgs a, const typename St::T* __restrict__ buf) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= a.nx) return;
  const int n = (z * a.ny + y) * a.nx + x;
  float f[L::Q];
  if constexpr (FROM_S) {
    static_for<L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      f[q] = St::template ld<L, q>(buf, q * a.stride + n);
    });
  } else {
    form_s<L, St>(a, buf, x, y, z, n, a.cls[n], false, f);
  }
  if (a.rho) {
    float rho, jx, jy, jz, ux, uy, uz;
    density_momentum<L>(f, rho, jx, jy, jz);
    guo_velocity(rho, jx, jy, jz, a.gx, a.gy, a.gz, ux, uy, uz);
    if constexpr (L::D == 2) uz = 0.0f;
    a.rho[n] = rho;
    a.ux[n] = ux;
    a.uy[n] = uy;
    if (a.uz) a.uz[n] = uz;
    a.speed[n] = L::D == 2 ? sqrtf(ux * ux + uy * uy) : sqrtf(ux * ux + uy * uy + uz * uz);
  }
  if (a.s_out) {
    const long long N = static_cast<long long>(a.nx) * a.ny * a.nz;
    static_for<L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      a.s_out[q * N + n] = f[q];
    });
  }
}

// Sum of three per-lane values over the work-group in a fixed tree order
// (blockDim.x a power of two, at most 256); the result is valid in lane 0.
__device__ __forceinline__ void block_sum3(double& s0, double& s1, double& s2) {
  __shared__ double red[3][256];
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  red[2][threadIdx.x] = s2;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  s0 = red[0][0];
  s1 = red[1][0];
  s2 = red[2][0];
}

// Rows holding a momentum-exchange link of (a.mem_c, a.mem_rmax): one lane
// per node, row_has[row] = 1 (an idempotent store; cleared by the caller).
// Whether Fluid node (x, y, z) holds a link of (a.mem_c, a.mem_rmax).
template <class L>
__device__ __forceinline__ bool has_mem_link(const StepArgs& a, int x, int y, int z, int n) {
  if (!(a.cls[n] & kClsSlow) || a.flags[n] != kFluid) return false;
  bool any = false;
  for_each_mem_link<L>(a, x, y, z, n, a.mem_c[0], a.mem_c[1], a.mem_c[2], a.mem_rmax,
                       [&](auto, float, float) { any = true; });
  return any;
}

template <class L>
__global__ __launch_bounds__(256) void k_mem_mark_rows(const StepArgs a, int* __restrict__ row_has) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= a.nx) return;
  const int n = (z * a.ny + y) * a.nx + x;
  if (has_mem_link<L>(a, x, y, z, n)) row_has[z * a.ny + y] = 1;
}

// The nodes holding links, row by row in node order: one wave per linked row
// (row_list) walks the row in 64-node segments; counts[r] = the row's nodes
// (EMIT = false), or the nodes written from nodes[offsets[r]] on (EMIT = true)
// at the place their ballot prefix gives them.
template <class L, bool EMIT>
__global__ __launch_bounds__(64) void k_mem_row_nodes(const StepArgs a, const int* __restrict__ row_list, int nrows,
                                                      int* __restrict__ counts, const int* __restrict__ offsets,
                                                      int* __restrict__ nodes) {
  const int lane = __lane_id();
  for (int r = blockIdx.x; r < nrows; r += gridDim.x) {  // uniform over the wave
    const int row = row_list[r];
    const int y = row % a.ny, z = row / a.ny;
    int base = EMIT ? offsets[r] : 0;
    for (int x0 = 0; x0 < a.nx; x0 += 64) {
      const int x = x0 + lane;
      const int n = row * a.nx + x;
      const bool want = x < a.nx && has_mem_link<L>(a, x, y, z, n);
      const unsigned long long ballot = __ballot(want);
      if constexpr (EMIT) {
        if (want) nodes[base + __popcll(ballot & ((1ull << lane) - 1ull))] = n;
      }
      base += __popcll(ballot);
    }
    if (!EMIT && lane == 0) counts[r] = base;
  }
}

// The partials of one force sample without a step: one work-group row per
// linked row (row_list), S_k read from a materialised state or formed from
// the post-collision buffer without committing the outlet store.  The S_k
// values, slots and sums are those a pull launch of k_step produces.
template <class L, class St, bool FROM_S>
__global__ __launch_bounds__(256) void k_mem_rows(const StepArgs a, const typename St::T* __restrict__ buf,
                                                  const int* __restrict__ row_list, int nrows) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = x < a.nx;
  for (int r = blockIdx.y; r < nrows; r += gridDim.y) {  // uniform over the work-group
    const int row = row_list[r];
    const int y = row % a.ny, z = row / a.ny;
    const int n = (z * a.ny + y) * a.nx + (live ? x : 0);
    const uint8_t c = live ? a.cls[n] : static_cast<uint8_t>(0);
    const bool want = live && (c & kClsSlow) && is_fluid(a, n);
    float f[L::Q];
    static_for<L::Q>([&](auto qc) { f[decltype(qc)::value] = 0.0f; });
    if (want) {
      if constexpr (FROM_S) {
        static_for<L::Q>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          f[q] = St::template ld<L, q>(buf, q * a.stride + n);
        });
      } else {
        form_s_at<L, St, false>(a, buf, x, y, z, n, c, false, a.s_step, f);
      }
    }
    mem_wave<L>(a, want, x, y, z, n, f, a.mem_rows[row]);
  }
}

}  // namespace wake_hip
