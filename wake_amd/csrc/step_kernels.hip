// step_kernels.hip -- lattice dispatch of the step launchers, BFL marking and
// the FP64 diagnostics reductions.
#include <algorithm>

#include "step_launch.hpp"

namespace wake_hip {

__global__ void k_mark_bfl(const int* __restrict__ nodes, int count, uint8_t* __restrict__ cls,
                           int* __restrict__ aux) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int n = nodes[i];
  cls[n] |= kClsBfl;
  aux[n] = i;
}

// Diagnostics: per-block FP64 partials then one fixed-order final pass.
__global__ __launch_bounds__(256) void k_diag_partial(const uint8_t* __restrict__ flags, const float* __restrict__ rho,
                                                      const float* __restrict__ speed, long long n,
                                                      double* __restrict__ partial) {
  __shared__ double sm[256], se[256], sx[256];
  double m = 0.0, e = 0.0, mx = 0.0;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += 256ll * gridDim.x) {
    if (flags[i] == kFluid) {
      const float r = rho[i], s = speed[i];
      m += r;
      e += static_cast<double>(0.5f * r * s * s);
      mx = fmax(mx, static_cast<double>(s));
    }
  }
  sm[threadIdx.x] = m;
  se[threadIdx.x] = e;
  sx[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sm[threadIdx.x] += sm[threadIdx.x + w];
      se[threadIdx.x] += se[threadIdx.x + w];
      sx[threadIdx.x] = fmax(sx[threadIdx.x], sx[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[3 * blockIdx.x + 0] = sm[0];
    partial[3 * blockIdx.x + 1] = se[0];
    partial[3 * blockIdx.x + 2] = sx[0];
  }
}

__global__ void k_diag_final(const double* __restrict__ partial, int nblocks, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double m = 0.0, e = 0.0, mx = 0.0;
  for (int b = 0; b < nblocks; ++b) {
    m += partial[3 * b];
    e += partial[3 * b + 1];
    mx = fmax(mx, partial[3 * b + 2]);
  }
  out[0] = m;
  out[1] = e;
  out[2] = mx;
}



hipError_t launch_build_class(int lattice, const StepArgs& a, uint8_t* cls, uint32_t* bmask, hipStream_t s) {
  switch (lattice) {
    case 9: return detail::build_class_d2q9(a, cls, bmask, s);
    case 19: return detail::build_class_d3q19(a, cls, bmask, s);
    case 27: return detail::build_class_d3q27(a, cls, bmask, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_step_tb(int lattice, int storage, int collision, bool les, const StepArgs& a, const void* in,
                          void* out, hipStream_t s) {
  if (lattice != 9) return hipErrorInvalidValue;
  return detail::step_tb_d2q9(storage, collision, les, a, in, out, s);
}

hipError_t launch_step(int lattice, int storage, int collision, bool les, bool pull, const StepArgs& a,
                       const void* a_in, void* a_out, int block_x, hipStream_t s) {
  switch (lattice) {
    case 9: return detail::step_d2q9(storage, collision, les, pull, a, a_in, a_out, block_x, s);
    case 19: return detail::step_d3q19(storage, collision, les, pull, a, a_in, a_out, block_x, s);
    case 27: return detail::step_d3q27(storage, collision, les, pull, a, a_in, a_out, block_x, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_prev_from_s(int lattice, int storage, const StepArgs& a, const void* s_state, hipStream_t s) {
  switch (lattice) {
    case 9: return detail::prev_from_s_d2q9(storage, a, s_state, s);
    case 19: return detail::prev_from_s_d3q19(storage, a, s_state, s);
    case 27: return detail::prev_from_s_d3q27(storage, a, s_state, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_materialize(int lattice, int storage, const StepArgs& a, const void* a_in, void* s_state,
                              hipStream_t s) {
  switch (lattice) {
    case 9: return detail::materialize_d2q9(storage, a, a_in, s_state, s);
    case 19: return detail::materialize_d3q19(storage, a, a_in, s_state, s);
    case 27: return detail::materialize_d3q27(storage, a, a_in, s_state, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_readback(int lattice, int storage, bool from_s, const StepArgs& a, const void* buf,
                           hipStream_t s) {
  switch (lattice) {
    case 9: return detail::readback_d2q9(storage, from_s, a, buf, s);
    case 19: return detail::readback_d3q19(storage, from_s, a, buf, s);
    case 27: return detail::readback_d3q27(storage, from_s, a, buf, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_mem_mark_rows(int lattice, const StepArgs& a, int* row_has, hipStream_t s) {
  switch (lattice) {
    case 9: return detail::mem_mark_rows_d2q9(a, row_has, s);
    case 19: return detail::mem_mark_rows_d3q19(a, row_has, s);
    case 27: return detail::mem_mark_rows_d3q27(a, row_has, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_mem_row_nodes(int lattice, const StepArgs& a, const int* row_list, int nrows, int* counts,
                                const int* offsets, int* nodes, hipStream_t s) {
  switch (lattice) {
    case 9: return detail::mem_row_nodes_d2q9(a, row_list, nrows, counts, offsets, nodes, s);
    case 19: return detail::mem_row_nodes_d3q19(a, row_list, nrows, counts, offsets, nodes, s);
    case 27: return detail::mem_row_nodes_d3q27(a, row_list, nrows, counts, offsets, nodes, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_mem_rows(int lattice, int storage, bool from_s, const StepArgs& a, const void* buf,
                           const int* row_list, int nrows, hipStream_t s) {
  switch (lattice) {
    case 9: return detail::mem_rows_d2q9(storage, from_s, a, buf, row_list, nrows, s);
    case 19: return detail::mem_rows_d3q19(storage, from_s, a, buf, row_list, nrows, s);
    case 27: return detail::mem_rows_d3q27(storage, from_s, a, buf, row_list, nrows, s);
    default: return hipErrorInvalidValue;
  }
}

// One work-group per force sample: the slots summed in a fixed order
// (lane-strided, then a tree over the work-group).
__global__ __launch_bounds__(256) void k_mem_fold(const double* __restrict__ part, int nslots, int ring,
                                                  long long first, double* __restrict__ out, long long out_first) {
  const long long k = first + blockIdx.x;
  const double* p = part + 3ll * nslots * (k % ring);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int i = threadIdx.x; i < nslots; i += 256) {
    s0 += p[3 * i];
    s1 += p[3 * i + 1];
    s2 += p[3 * i + 2];
  }
  block_sum3(s0, s1, s2);
  if (threadIdx.x == 0) {
    double* o = out + 3 * (out_first + blockIdx.x);
    o[0] = s0;
    o[1] = s1;
    o[2] = s2;
  }
}

hipError_t launch_mem_fold(const double* part, int nslots, int ring, long long first, int count, double* out,
                           long long out_first, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (nslots <= 0 || ring <= 0 || count > ring) return hipErrorInvalidValue;
  k_mem_fold<<<count, 256, 0, s>>>(part, nslots, ring, first, out, out_first);
  return hipGetLastError();
}

hipError_t launch_mark_bfl(const int* nodes, int count, uint8_t* cls, int* aux, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  k_mark_bfl<<<(count + 255) / 256, 256, 0, s>>>(nodes, count, cls, aux);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_count_nonfinite(const float* __restrict__ rho,
                                                         const float* __restrict__ speed, long long n,
                                                         unsigned long long* __restrict__ out) {
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += 256ll * gridDim.x) {
    if (!__builtin_isfinite(rho[i]) || !__builtin_isfinite(speed[i])) {
      atomicAdd(&out[0], 1ull);
      atomicMin(&out[1], static_cast<unsigned long long>(i));
    }
  }
}

hipError_t launch_count_nonfinite(const float* rho, const float* speed, long long n, unsigned long long* out,
                                  hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const long long blocks = std::min<long long>(1024, (n + 255) / 256);
  k_count_nonfinite<<<static_cast<int>(blocks), 256, 0, s>>>(rho, speed, n, out);
  return hipGetLastError();
}

hipError_t launch_diagnostics(const uint8_t* flags, const float* rho, const float* speed, long long n,
                              double* partial, int nblocks, double* out3, hipStream_t s) {
  k_diag_partial<<<nblocks, 256, 0, s>>>(flags, rho, speed, n, partial);
  k_diag_final<<<1, 64, 0, s>>>(partial, nblocks, out3);
  return hipGetLastError();
}

}  // namespace wake_hip
