// hbm_probe.hip -- the practical HBM ceilings the step kernel is measured
// against, timed on the GPU it runs on.
//
// * contiguous copy: four floats per thread, the work-group's 1,024 floats
//   contiguous and each wave instruction one 256-byte run -- the fastest copy
//   shape found on MI355X (about 1 KiB per wave; wider or longer per-thread
//   copies run slower);
// * D3Q19 pull copy: 19 SoA planes of nx*ny*nz floats, each node taking
//   population q from x - c_q (periodic wrap) into its own slot, no
//   arithmetic: the step kernel's memory access with nothing else;
// * the same pull with 16-bit values: the FP16 / BF16 step's access.
//
// The first two move 19 * 4 bytes in and out per node, the FP32 step's
// 152 B/node; the third 19 * 2, the 16-bit step's 76 B/node.  Shapes 3 and 4
// are the D3Q27 pull over 27 planes (216 and 108 B/node).
#include <hip/hip_runtime.h>

#include <string>

#include "capi_internal.hpp"
#include "lattice.hpp"
#include "wake_hip.h"

namespace wake_hip {
namespace {

constexpr int kCopyPerThread = 4;
constexpr int kCopyBlock = 256;

__global__ __launch_bounds__(kCopyBlock) void k_probe_stream(const float* __restrict__ a, float* __restrict__ b,
                                                             long long n) {
  const long long base = static_cast<long long>(blockIdx.x) * (kCopyBlock * kCopyPerThread) + threadIdx.x;
  float v[kCopyPerThread];
#pragma unroll
  for (int u = 0; u < kCopyPerThread; ++u) {
    const long long i = base + u * kCopyBlock;
    v[u] = i < n ? a[i] : 0.0f;
  }
#pragma unroll
  for (int u = 0; u < kCopyPerThread; ++u) {
    const long long i = base + u * kCopyBlock;
    if (i < n) __builtin_nontemporal_store(v[u], b + i);
  }
}

// One work-group per (row y, z), one thread per x; nx <= 1024.  T: float
// (the FP32 step's access) or a 16-bit value (the FP16 / BF16 step's).
template <class L, class T>
__global__ __launch_bounds__(1024) void k_probe_pull(const T* __restrict__ a, T* __restrict__ b, int nx, int ny,
                                                     int nz) {
  const int x = threadIdx.x, y = blockIdx.x, z = blockIdx.y;
  if (x >= nx) return;
  const long long plane = static_cast<long long>(nx) * ny * nz;
  const long long n = (static_cast<long long>(z) * ny + y) * nx + x;
  T v[L::Q];
#pragma unroll
  for (int q = 0; q < L::Q; ++q) {
    int sx = x - L::cx[q], sy = y - L::cy[q], sz = z - L::cz[q];
    sx += sx < 0 ? nx : (sx >= nx ? -nx : 0);
    sy += sy < 0 ? ny : (sy >= ny ? -ny : 0);
    sz += sz < 0 ? nz : (sz >= nz ? -nz : 0);
    v[q] = a[q * plane + (static_cast<long long>(sz) * ny + sy) * nx + sx];
  }
#pragma unroll
  for (int q = 0; q < L::Q; ++q) __builtin_nontemporal_store(v[q], b + q * plane + n);
}

hipError_t run_probe(int shape, int nx, int ny, int nz, int reps, float* ms_out) {
  const long long count = (shape == 5 ? 9LL : (shape >= 3 ? 27LL : 19LL)) * nx * ny * nz;
  float* a = nullptr;  // count floats; the 16-bit shapes use the first count halves
  float* b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t err = hipMalloc(&a, count * sizeof(float));
  if (err == hipSuccess) err = hipMalloc(&b, count * sizeof(float));
  if (err == hipSuccess) err = hipMemset(a, 0, count * sizeof(float));
  if (err == hipSuccess) err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  auto launch = [&] {
    if (shape == 0) {
      const long long per = kCopyBlock * kCopyPerThread;
      k_probe_stream<<<static_cast<unsigned>((count + per - 1) / per), kCopyBlock>>>(a, b, count);
    } else if (shape == 1) {
      k_probe_pull<D3Q19, float><<<dim3(ny, nz), nx>>>(a, b, nx, ny, nz);
    } else if (shape == 2) {
      k_probe_pull<D3Q19, unsigned short><<<dim3(ny, nz), nx>>>(reinterpret_cast<const unsigned short*>(a),
                                                                reinterpret_cast<unsigned short*>(b), nx, ny, nz);
    } else if (shape == 3) {
      k_probe_pull<D3Q27, float><<<dim3(ny, nz), nx>>>(a, b, nx, ny, nz);
    } else if (shape == 4) {
      k_probe_pull<D3Q27, unsigned short><<<dim3(ny, nz), nx>>>(reinterpret_cast<const unsigned short*>(a),
                                                                reinterpret_cast<unsigned short*>(b), nx, ny, nz);
    } else {
      k_probe_pull<D2Q9, float><<<dim3(ny, nz), nx>>>(a, b, nx, ny, nz);
    }
  };
  if (err == hipSuccess) {
    for (int w = 0; w < 3; ++w) launch();  // warm-up
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipEventRecord(e0, nullptr);
  if (err == hipSuccess) {
    for (int r = 0; r < reps; ++r) launch();
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipEventRecord(e1, nullptr);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  if (err == hipSuccess) *ms_out = ms / reps;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return err;
}

}  // namespace
}  // namespace wake_hip

extern "C" int wake_hip_hbm_probe(int shape, int nx, int ny, int nz, int reps, float* ms_per_launch) {
  if (!ms_per_launch) return wake_hip::set_last_error("wake_hip_hbm_probe: null output");
  if (shape < 0 || shape > 5) return wake_hip::set_last_error("wake_hip_hbm_probe: shape must be 0 to 5");
  if (nx < 1 || ny < 1 || nz < 1 || nx > 1024 || reps < 1)
    return wake_hip::set_last_error("wake_hip_hbm_probe: need 1 <= nx <= 1024, ny, nz >= 1 and reps >= 1");
  if (nz > 65535) return wake_hip::set_last_error("wake_hip_hbm_probe: nz above the 65,535 grid limit");
  const hipError_t err = wake_hip::run_probe(shape, nx, ny, nz, reps, ms_per_launch);
  (void)hipGetLastError();  // a failure here is reported once, not by the next launch check
  if (err != hipSuccess)
    return wake_hip::set_last_error(std::string("wake_hip_hbm_probe: ") + hipGetErrorString(err));
  return 0;
}
