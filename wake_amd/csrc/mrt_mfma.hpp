// mrt_mfma.hpp -- the D3Q19 MRT moment transforms on the matrix cores: a
// compile-time variant of the 16-bit-storage step (WAKE_MRT_MFMA=1), kept
// out of the default build.
//
// One wave's 64 nodes form the 19 x 64 matrix F (populations x nodes), and
// m = M F and f' = Minv m' (d3q19_collision_mrt.ispc:66-174) become FP32
// matrix products: moments padded to 32 rows (2 blocks of 16), nodes in 4
// blocks of 16, populations in 5 blocks of 4 (padded to 20), that is
// 2 x 4 x 5 = 40 v_mfma_f32_16x16x4_f32 per transform.  Each lane holds one
// node, the products want a population (or moment) index that varies across
// lanes, so both transforms go through a per-wave LDS tile of 64 rows of 20
// floats: a lane writes its node's row, the products read columns, write the
// result back as rows, and each lane reads its own row again.
//
// The rounding differs from the VALU transforms only in the order of the
// FP32 sums (the hardware's FP32 products are exact, their accumulation
// rounds once per step of four).
#pragma once

#include <hip/hip_runtime.h>

#include "node_ops.hpp"

namespace wake_hip {

typedef float mfma_f4 __attribute__((ext_vector_type(4)));

inline constexpr int kMfmaRow = 20;  // floats per node row in the LDS tile (19 + a zero)
inline constexpr int kMfmaTile = 64 * kMfmaRow;

// The A operand of product (rb, kb) of lane l is entry
// [16 rb + l % 16][4 kb + l / 16] of M (t = 0) or the rounded Minv (t = 1),
// zero outside 19 x 19; ten per lane, padded to twelve for 16-byte loads.
struct MfmaOperands {
  float v[2][64][12];
};

constexpr MfmaOperands make_mfma_operands() {
  MfmaOperands t = {};
  for (int lane = 0; lane < 64; ++lane)
    for (int rb = 0; rb < 2; ++rb)
      for (int kb = 0; kb < 5; ++kb) {
        const int i = 16 * rb + lane % 16, k = 4 * kb + lane / 16;
        if (i < 19 && k < 19) {
          t.v[0][lane][rb * 5 + kb] = static_cast<float>(d3q19_m(i, k));
          t.v[1][lane][rb * 5 + kb] = kD3Q19Minv.minv[i][k];
        }
      }
  return t;
}

static __constant__ MfmaOperands kMfmaOps = make_mfma_operands();

// LDS writes of this wave's lanes become visible to its other lanes.
__device__ __forceinline__ void lds_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// out = T in over the 64 nodes of this wave (T = M for t = 0, Minv for
// t = 1).  Every lane of the wave must be active; tile is this wave's
// kMfmaTile floats of LDS.
template <int T>
__device__ __forceinline__ void mfma_transform_d3q19(const float (&in)[19], float (&out)[19], float* tile) {
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, kr = lane >> 4;
  mfma_f4* own = reinterpret_cast<mfma_f4*>(tile + lane * kMfmaRow);
  own[0] = mfma_f4{in[0], in[1], in[2], in[3]};
  own[1] = mfma_f4{in[4], in[5], in[6], in[7]};
  own[2] = mfma_f4{in[8], in[9], in[10], in[11]};
  own[3] = mfma_f4{in[12], in[13], in[14], in[15]};
  own[4] = mfma_f4{in[16], in[17], in[18], 0.0f};
  const mfma_f4* op = reinterpret_cast<const mfma_f4*>(kMfmaOps.v[T][lane]);
  const mfma_f4 o0 = op[0], o1 = op[1], o2 = op[2];
  const float av[2][5] = {{o0.x, o0.y, o0.z, o0.w, o1.x}, {o1.y, o1.z, o1.w, o2.x, o2.y}};
  lds_wave_sync();
  mfma_f4 acc[2][4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    acc[0][cb] = mfma_f4{0.0f, 0.0f, 0.0f, 0.0f};
    acc[1][cb] = mfma_f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 5; ++kb) {
      const float b = tile[(16 * cb + col) * kMfmaRow + 4 * kb + kr];
      acc[0][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[0][kb], b, acc[0][cb], 0, 0, 0);
      acc[1][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[1][kb], b, acc[1][cb], 0, 0, 0);
    }
  }
  lds_wave_sync();
  // D of product (rb, cb): lane l holds rows 16 rb + 4 (l / 16) + r, r < 4,
  // of node 16 cb + l % 16; rows 16..19 come from the lanes with l < 16.
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) {
    float* row = tile + (16 * cb + col) * kMfmaRow;
    *reinterpret_cast<mfma_f4*>(row + 4 * kr) = acc[0][cb];
    if (kr == 0) *reinterpret_cast<mfma_f4*>(row + 16) = acc[1][cb];
  }
  lds_wave_sync();
  const mfma_f4 r0 = own[0], r1 = own[1], r2 = own[2], r3 = own[3], r4 = own[4];
  out[0] = r0.x, out[1] = r0.y, out[2] = r0.z, out[3] = r0.w;
  out[4] = r1.x, out[5] = r1.y, out[6] = r1.z, out[7] = r1.w;
  out[8] = r2.x, out[9] = r2.y, out[10] = r2.z, out[11] = r2.w;
  out[12] = r3.x, out[13] = r3.y, out[14] = r3.z, out[15] = r3.w;
  out[16] = r4.x, out[17] = r4.y, out[18] = r4.z;
  lds_wave_sync();
}

template <bool LES, bool FAST>
__device__ __forceinline__ Macro collide_mrt_d3q19_mfma(float (&f)[19], float* tile, float tau_shear,
                                                        float tau_bulk, float gx, float gy, float gz, float cs) {
  float m[19];
  mfma_transform_d3q19<0>(f, m, tile);
  const Macro r = mrt_relax_d3q19<LES, FAST>(m, tau_shear, tau_bulk, gx, gy, gz, cs);
  mfma_transform_d3q19<1>(m, f, tile);
  return r;
}

}  // namespace wake_hip
