// thermal.hip -- kernels of the coupled D2Q9 flow + D2Q5 temperature step.
#include "thermal_impl.hpp"

namespace wake_hip {

template <class St, int COLL, bool PULL>
__global__ __launch_bounds__(256) void k_step_thermal(const StepArgs a, const typename St::T* __restrict__ ain,
                                                      typename St::T* __restrict__ aout,
                                                      const float* __restrict__ gin, float* __restrict__ gout) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= a.nx) return;
  thermal_node<St, COLL>(a, ain, aout, gin, gout, x, y, y * a.nx + x, PULL, a.s_step);
}

__global__ __launch_bounds__(256) void k_materialize_thermal(const StepArgs a, const float* __restrict__ gin,
                                                             float* __restrict__ gs) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= a.nx) return;
  const int n = y * a.nx + x;
  float g[D2Q5::Q];
  gather_g(a, gin, x, y, n, a.cls[n], g);
  static_for<D2Q5::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    gs[q * a.stride + n] = g[q];
  });
}

template <class St, bool FROM_S>
__global__ __launch_bounds__(256) void k_readback_thermal(const StepArgs a, const typename St::T* __restrict__ buf,
                                                          const float* __restrict__ gbuf) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= a.nx) return;
  const int n = y * a.nx + x;
  const uint8_t c = a.cls[n];
  float f[D2Q9::Q], g[D2Q5::Q];
  if constexpr (FROM_S) {
    static_for<D2Q9::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      f[q] = St::template ld<D2Q9, q>(buf, q * a.stride + n);
    });
    static_for<D2Q5::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      g[q] = gbuf[q * a.stride + n];
    });
  } else {
    form_s<D2Q9, St>(a, buf, x, y, 0, n, c, false, f);
    gather_g(a, gbuf, x, y, n, c, g);
  }
  const float T = thermal_t(g);
  const float gyn = a.gy + a.gbeta * (T - a.t_ref);
  float rho = 0.0f, jx = 0.0f, jy = 0.0f;
  static_for<D2Q9::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    rho += f[q];
    jx += static_cast<float>(D2Q9::cx[q]) * f[q];
    jy += static_cast<float>(D2Q9::cy[q]) * f[q];
  });
  const float inv_rho = 1.0f / rho;
  const float ux = (jx + 0.5f * rho * a.gx) * inv_rho;
  const float uy = (jy + 0.5f * rho * gyn) * inv_rho;
  a.temp[n] = T;
  if (a.rho) {
    a.rho[n] = rho;
    a.ux[n] = ux;
    a.uy[n] = uy;
    if (a.uz) a.uz[n] = 0.0f;
    a.speed[n] = sqrtf(ux * ux + uy * uy);
  }
}

namespace {
dim3 grid2d(const StepArgs& a) { return dim3((a.nx + 255) / 256, a.ny, 1); }

template <class F>
hipError_t with_storage_t(int storage, F&& f) {
  switch (storage) {
    case kStoreF32: return f(StF32{});
    case kStoreF16: return f(StF16{});
    case kStoreBF16: return f(StBF16{});
    default: return hipErrorInvalidValue;
  }
}
}  // namespace

hipError_t launch_step_thermal(int storage, int collision, bool pull, const StepArgs& a, const void* a_in,
                               void* a_out, const float* g_in, float* g_out, hipStream_t s) {
  return with_storage_t(storage, [&](auto st) {
    using St = decltype(st);
    using T = typename St::T;
    const T* i = static_cast<const T*>(a_in);
    T* o = static_cast<T*>(a_out);
    auto go = [&](auto coll) {
      constexpr int C = decltype(coll)::value;
      if (pull) k_step_thermal<St, C, true><<<grid2d(a), 256, 0, s>>>(a, i, o, g_in, g_out);
      else k_step_thermal<St, C, false><<<grid2d(a), 256, 0, s>>>(a, i, o, g_in, g_out);
      return hipGetLastError();
    };
    if (collision == kBGK) return go(std::integral_constant<int, kBGK>{});
    if (collision == kTRT) return go(std::integral_constant<int, kTRT>{});
    return go(std::integral_constant<int, kMRT>{});
  });
}

hipError_t launch_materialize_thermal(const StepArgs& a, const float* g_in, float* g_s, hipStream_t s) {
  k_materialize_thermal<<<grid2d(a), 256, 0, s>>>(a, g_in, g_s);
  return hipGetLastError();
}

hipError_t launch_readback_thermal(int storage, bool from_s, const StepArgs& a, const void* buf, const float* gbuf,
                                   hipStream_t s) {
  return with_storage_t(storage, [&](auto st) {
    using St = decltype(st);
    using T = typename St::T;
    if (from_s) k_readback_thermal<St, true><<<grid2d(a), 256, 0, s>>>(a, static_cast<const T*>(buf), gbuf);
    else k_readback_thermal<St, false><<<grid2d(a), 256, 0, s>>>(a, static_cast<const T*>(buf), gbuf);
    return hipGetLastError();
  });
}

}  // namespace wake_hip
