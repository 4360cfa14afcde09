// ispc_compat.hip -- the reference's LBM kernel exports on the GPU.
//
// Each extern "C" function keeps the name, parameters and semantics of one
// ISPC export (cited per function); the caller's host arrays are staged into
// device memory, one HIP kernel with one lane per node (or per link, or per
// boundary row) does the work, and the outputs are copied back before the
// function returns, so the call is synchronous like the ISPC one.  The
// per-node arithmetic is the shared operator set of node_ops.hpp.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "wake_hip.h"
#include "domain.hpp"
#include "engine.hpp"
#include "lattice.hpp"
#include "node_ops.hpp"

namespace wake_hip {
void record_error(const std::string& msg);

namespace {

// Growable per-thread device arena; staging only, reused between calls.
class Staging {
 public:
  ~Staging() {
    if (base_) (void)hipFree(base_);
  }
  bool reserve(size_t bytes) {
    if (bytes <= cap_) return true;
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
    cap_ = 0;
    if (hipMalloc(&base_, bytes) != hipSuccess) {
      base_ = nullptr;
      return false;
    }
    cap_ = bytes;
    return true;
  }
  char* base() { return static_cast<char*>(base_); }

 private:
  void* base_ = nullptr;
  size_t cap_ = 0;
};

thread_local Staging g_stage;

// One call's staging plan: regions carved out of the arena, inputs copied in,
// outputs copied back.
class Call {
 public:
  explicit Call(const char* name) : name_(name) {}
  template <class T>
  T* in(const T* host, size_t n) {
    regions_.push_back({const_cast<T*>(host), n * sizeof(T), true, false});
    return reinterpret_cast<T*>(static_cast<uintptr_t>(regions_.size()));
  }
  template <class T>
  T* out(T* host, size_t n) {
    regions_.push_back({host, n * sizeof(T), false, true});
    return reinterpret_cast<T*>(static_cast<uintptr_t>(regions_.size()));
  }
  template <class T>
  T* inout(T* host, size_t n) {
    regions_.push_back({host, n * sizeof(T), true, true});
    return reinterpret_cast<T*>(static_cast<uintptr_t>(regions_.size()));
  }
  // Stage inputs; returns false (error recorded) on failure.
  bool begin() {
    size_t total = 0;
    for (auto& r : regions_) {
      r.offset = total;
      total += (r.bytes + 255) / 256 * 256;
    }
    if (!g_stage.reserve(total ? total : 256)) return fail("hipMalloc");
    for (auto& r : regions_)
      if (r.copy_in && r.host && r.bytes)
        if (hipMemcpy(g_stage.base() + r.offset, r.host, r.bytes, hipMemcpyHostToDevice) != hipSuccess)
          return fail("hipMemcpy H2D");
    return true;
  }
  template <class T>
  T* dev(T* token) {
    const size_t k = static_cast<size_t>(reinterpret_cast<uintptr_t>(token));
    if (k == 0) return nullptr;
    const auto& r = regions_[k - 1];
    return r.host ? reinterpret_cast<T*>(g_stage.base() + r.offset) : nullptr;
  }
  bool end() {
    if (hipGetLastError() != hipSuccess) return fail("kernel launch");
    if (hipDeviceSynchronize() != hipSuccess) return fail("kernel");
    for (auto& r : regions_)
      if (r.copy_out && r.host && r.bytes)
        if (hipMemcpy(r.host, g_stage.base() + r.offset, r.bytes, hipMemcpyDeviceToHost) != hipSuccess)
          return fail("hipMemcpy D2H");
    return true;
  }

 private:
  struct Region {
    void* host;
    size_t bytes;
    bool copy_in, copy_out;
    size_t offset = 0;
  };
  bool fail(const char* what) {
    record_error(std::string("wake: ") + name_ + ": " + what + ": " + hipGetErrorString(hipGetLastError()));
    return false;
  }
  const char* name_;
  std::vector<Region> regions_;
};

inline unsigned blocks(long long n) { return static_cast<unsigned>((n + 255) / 256); }

// Marker positions must be finite and their stencils addressable (the
// kernels take floor() of them as int); the grid must be non-empty.
bool ibm_markers_ok(const float* mx, const float* my, int nx, int ny, int nm, const char* name) {
  bool ok = mx && my && nx > 0 && ny > 0;
  for (int m = 0; ok && m < nm; ++m)
    ok = std::isfinite(mx[m]) && std::isfinite(my[m]) && std::fabs(mx[m]) < 1e8f && std::fabs(my[m]) < 1e8f;
  if (!ok) record_error(std::string("wake: ") + name + ": marker positions must be finite and the grid non-empty");
  return ok;
}

// ----------------------------------------------------------------- kernels
template <class L, bool FORCED>
__global__ void k_moments(const float* __restrict__ f, float* rho, float* ux, float* uy, float* uz, float gx,
                          float gy, float gz, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float v[L::Q];
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    v[q] = f[static_cast<long long>(q) * count + i];
  });
  float r, jx, jy, jz;
  density_momentum<L>(v, r, jx, jy, jz);
  const float inv_rho = 1.0f / r;
  rho[i] = r;
  if constexpr (FORCED) {
    ux[i] = (jx + 0.5f * r * gx) * inv_rho;
    uy[i] = (jy + 0.5f * r * gy) * inv_rho;
    if (uz) uz[i] = (jz + 0.5f * r * gz) * inv_rho;
  } else {
    ux[i] = jx * inv_rho;
    uy[i] = jy * inv_rho;
    if (uz) uz[i] = jz * inv_rho;
  }
}

__global__ void k_speed(const float* __restrict__ ux, const float* __restrict__ uy, const float* __restrict__ uz,
                        float* speed, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float a = ux[i], b = uy[i];
  if (uz) {
    const float c = uz[i];
    speed[i] = sqrtf(a * a + b * b + c * c);
  } else {
    speed[i] = sqrtf(a * a + b * b);
  }
}

// Collision kernels write the collide-time rho and u like the exports do.
template <class L, int COLL, bool REF_PAIRS>
__global__ void k_collide(const float* __restrict__ f_in, float* __restrict__ f_out, const uint8_t* __restrict__ ct,
                          float* rho, float* ux, float* uy, float* uz, float tau, float tau2, float gx, float gy,
                          float gz, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float f[L::Q];
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    f[q] = f_in[static_cast<long long>(q) * count + i];
  });
  if (ct[i] != kSolid) {
    Macro m;
    if constexpr (COLL == kBGK) m = collide_bgk<L>(f, tau, gx, gy, gz);
    else if constexpr (COLL == kTRT) m = collide_trt<L, REF_PAIRS>(f, tau, tau2, gx, gy, gz);
    else if constexpr (L::Q == 9) m = collide_mrt_d2q9(f, tau, tau2, gx, gy);
    else m = collide_mrt_d3q19(f, tau, tau2, gx, gy, gz);
    rho[i] = m.rho;
    ux[i] = m.ux;
    uy[i] = m.uy;
    if (uz) uz[i] = m.uz;
  }
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    f_out[static_cast<long long>(q) * count + i] = f[q];
  });
}

// Periodic pull streaming, one lane per node and direction loop in registers.
template <class L>
__global__ void k_stream_periodic(const float* __restrict__ src, float* __restrict__ dst, int nx, int ny, int nz) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= nx) return;
  const long long N = static_cast<long long>(nx) * ny * nz;
  const long long n = (static_cast<long long>(z) * ny + y) * nx + x;
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const int sx = (x - L::cx[q] + nx) % nx, sy = (y - L::cy[q] + ny) % ny, sz = (z - L::cz[q] + nz) % nz;
    dst[q * N + n] = src[q * N + (static_cast<long long>(sz) * ny + sy) * nx + sx];
  });
}

// Wall streaming: x periodic, y/z out of range or solid -> own opposite,
// moving wall -> own opposite + momentum term.
template <class L>
__global__ void k_stream_walls(const float* __restrict__ src, float* __restrict__ dst, const uint8_t* __restrict__ ct,
                               int nx, int ny, int nz, float uwx, float uwy, float uwz) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= nx) return;
  const long long N = static_cast<long long>(nx) * ny * nz;
  const long long n = (static_cast<long long>(z) * ny + y) * nx + x;
  const bool solid = ct[n] == kSolid;
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int o = L::opp[q];
    if (solid) {
      dst[q * N + n] = src[q * N + n];
      return;
    }
    const int sx = (x - L::cx[q] + nx) % nx, sy = y - L::cy[q], sz = z - L::cz[q];
    if (sy >= 0 && sy < ny && sz >= 0 && sz < nz) {
      const long long sn = (static_cast<long long>(sz) * ny + sy) * nx + sx;
      const uint8_t t = ct[sn];
      if (t == kSolid) {
        dst[q * N + n] = src[o * N + n];
      } else if (t == kMovingWall) {
        const float ci_uw = L::D == 2 ? static_cast<float>(L::cx[q]) * uwx + static_cast<float>(L::cy[q]) * uwy
                                      : static_cast<float>(L::cx[q]) * uwx + static_cast<float>(L::cy[q]) * uwy +
                                            static_cast<float>(L::cz[q]) * uwz;
        dst[q * N + n] = src[o * N + n] + 2.0f * L::w[q] * 1.0f * 3.0f * ci_uw;
      } else {
        dst[q * N + n] = src[q * N + sn];
      }
    } else {
      dst[q * N + n] = src[o * N + n];
    }
  });
}

__global__ void k_bounce_back_d2q9(float* f, const uint8_t* __restrict__ ct, float uwx, float uwy, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  if (ct[i] != kSolid && ct[i] != kMovingWall) return;
  float v[9];
  for (int q = 0; q < 9; ++q) v[q] = f[static_cast<long long>(q) * count + i];
  static_for<9>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if constexpr (q > 0) {
      const float ci_uw = static_cast<float>(D2Q9::cx[q]) * uwx + static_cast<float>(D2Q9::cy[q]) * uwy;
      f[static_cast<long long>(q) * count + i] = v[D2Q9::opp[q]] + 2.0f * D2Q9::w[q] * 1.0f * 3.0f * ci_uw;
    }
  });
}

__global__ void k_zh_in_d2q9(float* f, float* rho, float* ux, float* uy, float uin_x, float uin_y, int nx, int ny) {
  const int y = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= ny - 1) return;
  const long long N = static_cast<long long>(nx) * ny;
  const long long idx = static_cast<long long>(y) * nx;
  const float f0 = f[idx], f2 = f[2 * N + idx], f3 = f[3 * N + idx], f4 = f[4 * N + idx], f6 = f[6 * N + idx],
              f7 = f[7 * N + idx];
  const float rho_in = (f0 + f2 + f4 + 2.0f * (f3 + f6 + f7)) / (1.0f - uin_x);
  f[1 * N + idx] = f3 + (2.0f / 3.0f) * rho_in * uin_x;
  f[5 * N + idx] = f7 - 0.5f * (f2 - f4) + (1.0f / 6.0f) * rho_in * uin_x + 0.5f * rho_in * uin_y;
  f[8 * N + idx] = f6 + 0.5f * (f2 - f4) + (1.0f / 6.0f) * rho_in * uin_x - 0.5f * rho_in * uin_y;
  rho[idx] = rho_in;
  ux[idx] = uin_x;
  uy[idx] = uin_y;
}

__global__ void k_zh_out_d2q9(float* f, float* rho, float* ux, float* uy, float rho_out, int nx, int ny) {
  const int y = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= ny - 1) return;
  const long long N = static_cast<long long>(nx) * ny;
  const long long idx = static_cast<long long>(y) * nx + (nx - 1);
  const float f0 = f[idx], f1 = f[N + idx], f2 = f[2 * N + idx], f4 = f[4 * N + idx], f5 = f[5 * N + idx],
              f8 = f[8 * N + idx];
  const float u_out_x = -1.0f + (f0 + f2 + f4 + 2.0f * (f1 + f5 + f8)) / rho_out;
  const float u_out_y = 0.0f;
  f[3 * N + idx] = f1 - (2.0f / 3.0f) * rho_out * u_out_x;
  f[6 * N + idx] = f8 - 0.5f * (f2 - f4) - (1.0f / 6.0f) * rho_out * u_out_x + 0.5f * rho_out * u_out_y;
  f[7 * N + idx] = f5 + 0.5f * (f2 - f4) - (1.0f / 6.0f) * rho_out * u_out_x - 0.5f * rho_out * u_out_y;
  rho[idx] = rho_out;
  ux[idx] = u_out_x;
  uy[idx] = u_out_y;
}

__global__ void k_convective_d2q9(float* f, const float* __restrict__ prev, float u_conv, int nx, int ny) {
  const int y = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= ny - 1) return;
  const long long N = static_cast<long long>(nx) * ny;
  const long long io = static_cast<long long>(y) * nx + (nx - 1), ii = io - 1;
  const float alpha = u_conv / (1.0f + u_conv), beta = 1.0f / (1.0f + u_conv);
  for (int q = 0; q < 9; ++q) f[q * N + io] = beta * prev[q * N + io] + alpha * f[q * N + ii];
}

template <class L>
__global__ void k_les(const float* __restrict__ f, const float* __restrict__ rho, const float* __restrict__ ux,
                      const float* __restrict__ uy, const float* __restrict__ uz, float* tau_eff, float tau0,
                      float cs, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float v[L::Q];
  static_for<L::Q>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    v[q] = f[static_cast<long long>(q) * count + i];
  });
  tau_eff[i] = smagorinsky_tau<L>(v, rho[i], ux[i], uy[i], uz ? uz[i] : 0.0f, tau0, cs);
}

// d3q19_curved.ispc:7-71, one lane per link.
__global__ void k_bfl_d3q19(float* f_dst, const float* __restrict__ f_src, const int* __restrict__ idx,
                            const int* __restrict__ qd, const float* __restrict__ delta, const float* __restrict__ uwx,
                            const float* __restrict__ uwy, const float* __restrict__ uwz, int nx, int ny, int nz,
                            int nlinks) {
  const int link = blockIdx.x * blockDim.x + threadIdx.x;
  if (link >= nlinks) return;
  const long long N = static_cast<long long>(nx) * ny * nz;
  const int idx_f = idx[link], q = qd[link];
  const float d = delta[link];
  const int opp_q = D3Q19::opp[q], cx = D3Q19::cx[q], cy = D3Q19::cy[q], cz = D3Q19::cz[q];
  const float ci_uw = static_cast<float>(cx) * uwx[link] + static_cast<float>(cy) * uwy[link] +
                      static_cast<float>(cz) * uwz[link];
  const float f_i_xf = f_src[q * N + idx_f];
  const float f_opp_xf = f_src[opp_q * N + idx_f];
  float reflected;
  if (d < 0.5f) {
    const int z_f = idx_f / (nx * ny);
    const int rem = idx_f - z_f * (nx * ny);
    const int y_f = rem / nx;
    const int x_f = rem - y_f * nx;
    const int x_ff = (x_f - cx + nx) % nx, y_ff = y_f - cy, z_ff = z_f - cz;
    float f_i_xff = f_i_xf;
    if (y_ff >= 0 && y_ff < ny && z_ff >= 0 && z_ff < nz) f_i_xff = f_src[q * N + (z_ff * ny + y_ff) * nx + x_ff];
    const float mom = 2.0f * D3Q19::w[q] * 1.0f * 3.0f * ci_uw;
    reflected = 2.0f * d * f_i_xf + (1.0f - 2.0f * d) * f_i_xff - mom;
  } else {
    const float inv_2delta = 1.0f / (2.0f * d);
    const float mom = (1.0f / d) * D3Q19::w[q] * 1.0f * 3.0f * ci_uw;
    reflected = inv_2delta * f_i_xf + (1.0f - inv_2delta) * f_opp_xf - mom;
  }
  f_dst[opp_q * N + idx_f] = reflected;
}

// Full-grid momentum exchange (forces_momentum_exchange.ispc:6-138): FP64
// partials per work-group, one fixed-order final sum.
template <class L>
__global__ __launch_bounds__(256) void k_mem_grid(const float* __restrict__ f, const uint8_t* __restrict__ ct,
                                                  int nx, int ny, int nz, float ccx, float ccy, float ccz,
                                                  float max_radius, double* __restrict__ partial) {
  __shared__ double red[3][256];
  double s0 = 0, s1 = 0, s2 = 0;
  const long long N = static_cast<long long>(nx) * ny * nz;
  const float r_max_sq = max_radius * max_radius;
  for (long long n = blockIdx.x * 256ll + threadIdx.x; n < N; n += 256ll * gridDim.x) {
    if (ct[n] != kFluid) continue;
    const int x = static_cast<int>(n % nx);
    const int y = static_cast<int>((n / nx) % ny);
    const int z = static_cast<int>(n / (static_cast<long long>(nx) * ny));
    static_for<L::Q>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      if constexpr (q > 0) {
        constexpr int cx = L::cx[q], cy = L::cy[q], cz = L::cz[q];
        const int yy = y + cy, zz = z + cz;
        if (yy >= 0 && yy < ny && zz >= 0 && zz < nz) {
          const uint8_t t = ct[(static_cast<long long>(zz) * ny + yy) * nx + (x + cx + nx) % nx];
          if (t == kSolid || t == kMovingWall) {
            const float rx = (static_cast<float>(x) + 0.5f * static_cast<float>(cx)) - ccx;
            const float ry = (static_cast<float>(y) + 0.5f * static_cast<float>(cy)) - ccy;
            const float rz = (static_cast<float>(z) + 0.5f * static_cast<float>(cz)) - ccz;
            const float d2 = L::D == 2 ? rx * rx + ry * ry : rx * rx + ry * ry + rz * rz;
            if (max_radius <= 0.0f || d2 <= r_max_sq) {
              const float fv = f[q * N + n];
              const float dfx = 2.0f * fv * static_cast<float>(cx), dfy = 2.0f * fv * static_cast<float>(cy);
              s0 += dfx;
              s1 += dfy;
              if constexpr (L::D == 2) s2 += static_cast<double>(rx * dfy - ry * dfx);
              else s2 += static_cast<double>(2.0f * fv * static_cast<float>(cz));
            }
          }
        }
      }
    });
  }
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  red[2][threadIdx.x] = s2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) partial[3 * blockIdx.x + k] = red[k][0];
}

__global__ void k_sum_partials(const double* __restrict__ partial, int nb, float* out) {
  if (threadIdx.x != 0) return;
  double s[3] = {0, 0, 0};
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < 3; ++k) s[k] += partial[3 * b + k];
  for (int k = 0; k < 3; ++k) out[k] = static_cast<float>(s[k]);
}

constexpr int kMemBlocks = 256;

template <class L>
void mem_grid(const char* name, const float* f, const uint8_t* ct, int nx, int ny, int nz, float cx, float cy,
              float cz, float rmax, float* o0, float* o1, float* o2) {
  if (nx <= 0 || ny <= 0 || nz <= 0) {  // no node: the sums are zero
    *o0 = *o1 = *o2 = 0.0f;
    return;
  }
  const long long N = static_cast<long long>(nx) * ny * nz;
  float res[3] = {0, 0, 0};
  std::vector<double> partial(3 * kMemBlocks);
  // A grid's own f_src() mirror (the reference's scenarios pass
  // grid.f_src_data()) is already on the device, byte for byte: reduce that
  // copy instead of uploading the populations again.
  const float* resident = device_copy_of_host_populations(f, static_cast<size_t>(L::Q) * N);
  Call c(name);
  auto f_ = c.in(resident ? nullptr : f, resident ? 0 : L::Q * N);
  auto ct_ = c.in(ct, N);
  auto p_ = c.out(partial.data(), partial.size());
  auto r_ = c.out(res, 3);
  if (!c.begin()) return;
  k_mem_grid<L><<<kMemBlocks, 256>>>(resident ? resident : c.dev(f_), c.dev(ct_), nx, ny, nz, cx, cy, cz, rmax,
                                     c.dev(p_));
  k_sum_partials<<<1, 64>>>(c.dev(p_), kMemBlocks, c.dev(r_));
  if (!c.end()) return;
  *o0 = res[0];
  *o1 = res[1];
  *o2 = res[2];
}

template <class L, int COLL, bool REF_PAIRS = false>
void collide_call(const char* name, const float* f_in, float* f_out, const uint8_t* ct, float* rho, float* ux,
                  float* uy, float* uz, float tau, float tau2, float gx, float gy, float gz, int count) {
  if (count <= 0) return;
  Call c(name);
  const size_t n = static_cast<size_t>(count);
  auto fi = c.in(f_in, L::Q * n);
  auto fo = c.inout(f_out, L::Q * n);
  auto t = c.in(ct, n);
  auto r = c.inout(rho, n);
  auto a = c.inout(ux, n);
  auto b = c.inout(uy, n);
  auto z = c.inout(uz, uz ? n : 0);
  if (!c.begin()) return;
  k_collide<L, COLL, REF_PAIRS><<<blocks(count), 256>>>(c.dev(fi), c.dev(fo), c.dev(t), c.dev(r), c.dev(a), c.dev(b),
                                                        c.dev(z), tau, tau2, gx, gy, gz, count);
  c.end();
}

template <class L, bool FORCED>
void moments_call(const char* name, const float* f, float* rho, float* ux, float* uy, float* uz, float gx, float gy,
                  float gz, int count) {
  if (count <= 0) return;
  Call c(name);
  const size_t n = static_cast<size_t>(count);
  auto fi = c.in(f, L::Q * n);
  auto r = c.out(rho, n);
  auto a = c.out(ux, n);
  auto b = c.out(uy, n);
  auto z = c.out(uz, uz ? n : 0);
  if (!c.begin()) return;
  k_moments<L, FORCED><<<blocks(count), 256>>>(c.dev(fi), c.dev(r), c.dev(a), c.dev(b), c.dev(z), gx, gy, gz, count);
  c.end();
}

template <class L>
void stream_periodic_call(const char* name, const float* src, float* dst, int nx, int ny, int nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return;
  Call c(name);
  const size_t N = static_cast<size_t>(nx) * ny * nz;
  auto s = c.in(src, L::Q * N);
  auto d = c.out(dst, L::Q * N);
  if (!c.begin()) return;
  const int bx = nx >= 256 ? 256 : 64;
  k_stream_periodic<L><<<dim3((nx + bx - 1) / bx, ny, nz), bx>>>(c.dev(s), c.dev(d), nx, ny, nz);
  c.end();
}

template <class L>
void stream_walls_call(const char* name, const float* src, float* dst, const uint8_t* ct, int nx, int ny, int nz,
                       float uwx, float uwy, float uwz) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return;
  Call c(name);
  const size_t N = static_cast<size_t>(nx) * ny * nz;
  auto s = c.in(src, L::Q * N);
  auto d = c.out(dst, L::Q * N);
  auto t = c.in(ct, N);
  if (!c.begin()) return;
  const int bx = nx >= 256 ? 256 : 64;
  k_stream_walls<L><<<dim3((nx + bx - 1) / bx, ny, nz), bx>>>(c.dev(s), c.dev(d), c.dev(t), nx, ny, nz, uwx, uwy, uwz);
  c.end();
}

template <class L>
void les_call(const char* name, const float* f, const float* rho, const float* ux, const float* uy, const float* uz,
              float* tau_eff, float tau0, float cs, int count) {
  if (count <= 0) return;
  Call c(name);
  const size_t n = static_cast<size_t>(count);
  auto fi = c.in(f, L::Q * n);
  auto r = c.in(rho, n);
  auto a = c.in(ux, n);
  auto b = c.in(uy, n);
  auto z = c.in(uz, uz ? n : 0);
  auto t = c.out(tau_eff, n);
  if (!c.begin()) return;
  k_les<L><<<blocks(count), 256>>>(c.dev(fi), c.dev(r), c.dev(a), c.dev(b), c.dev(z), c.dev(t), tau0, cs, count);
  c.end();
}

// ------------------------------------------- thermal D2Q9 and IBM kernels
// thermal_kernels.ispc:4-44: BGK of the temperature populations g with the
// first-order equilibrium w_q T (1 + 3 c_q.u); T = sum g is written out.
// Solid nodes copy g.
__global__ __launch_bounds__(256) void k_thermal_collide_d2q9(const float* __restrict__ g_in, float* __restrict__ g_out,
                                                              const uint8_t* __restrict__ ct,
                                                              const float* __restrict__ ux,
                                                              const float* __restrict__ uy, float* __restrict__ temp,
                                                              float omega, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const size_t n = static_cast<size_t>(count);
  if (ct[i] == 1) {
    for (int q = 0; q < 9; ++q) g_out[q * n + i] = g_in[q * n + i];
    return;
  }
  float g[9], t = 0.0f;
  for (int q = 0; q < 9; ++q) {
    g[q] = g_in[q * n + i];
    t += g[q];
  }
  temp[i] = t;
  const float vx = ux[i], vy = uy[i];
  for (int q = 0; q < 9; ++q) {
    const float cu = static_cast<float>(D2Q9::cx[q]) * vx + static_cast<float>(D2Q9::cy[q]) * vy;
    const float geq = D2Q9::w[q] * t * (1.0f + 3.0f * cu);
    g_out[q * n + i] = g[q] - omega * (g[q] - geq);
  }
}

// thermal_kernels.ispc:48-60.  ACC: add to gy, as the export does; else
// gy = 0 + force, the same bits as clearing gy first (the solver's step).
template <bool ACC>
__global__ __launch_bounds__(256) void k_thermal_buoyancy_2d(const float* __restrict__ temp, float* __restrict__ gy,
                                                             float t_ref, float beta, float g_gravity, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) gy[i] = (ACC ? gy[i] : 0.0f) + beta * (temp[i] - t_ref) * g_gravity;
}

// thermal_kernels.ispc:63-81: on row y_wall, for q = 0..8 in turn,
// g[opp q] = -g[q] + 2 w_q T_wall in place -- later q read what earlier q
// wrote, so g_1, g_2, g_5, g_6 come back to (almost) their old values and
// only g_0 and the opposite directions change.
__global__ __launch_bounds__(256) void k_thermal_isothermal_wall(float* __restrict__ g, int nx, int ny, int y_wall,
                                                                 float t_wall) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nx) return;
  const size_t n = static_cast<size_t>(nx) * ny, idx = static_cast<size_t>(y_wall) * nx + x;
  for (int q = 0; q < 9; ++q) {
    const float v = g[q * n + idx];
    g[D2Q9::opp_ref[q] * n + idx] = -v + 2.0f * D2Q9::w[q] * t_wall;
  }
}

// ibm_kernels.ispc:4-8: the regularised 4-point cosine delta
__device__ inline float peskin_1d(float r) {
  const float a = fabsf(r);
  if (a >= 2.0f) return 0.0f;
  return 0.25f * (1.0f + cosf(0.5f * 3.14159265358979323846f * a));
}

// ibm_kernels.ispc:12-50: one lane per marker, 5 x 5 stencil from
// floor(x) - 2, x periodic, rows outside [0, ny) skipped.
__global__ __launch_bounds__(256) void k_ibm_interpolate_2d(const float* __restrict__ ux, const float* __restrict__ uy,
                                                            const float* __restrict__ mx, const float* __restrict__ my,
                                                            float* __restrict__ mux, float* __restrict__ muy, int nx,
                                                            int ny, int nm) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const float x = mx[m], y = my[m];
  const int fx = static_cast<int>(floorf(x)), fy = static_cast<int>(floorf(y));
  float su = 0.0f, sv = 0.0f;
  for (int iy = -2; iy <= 2; ++iy) {
    const int py = fy + iy;
    if (py < 0 || py >= ny) continue;
    const float wy = peskin_1d(static_cast<float>(py) - y);
    for (int ix = -2; ix <= 2; ++ix) {
      const int px = fx + ix;
      const int pw = (px % nx + nx) % nx;
      const float w = peskin_1d(static_cast<float>(px) - x) * wy;
      const size_t e = static_cast<size_t>(py) * nx + pw;
      su += w * ux[e];
      sv += w * uy[e];
    }
  }
  mux[m] = su;
  muy[m] = sv;
}

// ibm_kernels.ispc:54-96: markers in order, each adding its 5 x 5 stencil
// of weight * f * ds.  The row weight is the cosine without the |r| < 2
// cutoff (ibm_kernels.ispc:80), as there, so the row floor(y) - 2 gets a
// weight too.  One work-group: lane l < 25 owns stencil cell l; a barrier
// between markers keeps the additions to a cell in marker order.  With
// nx < 5 a stencil row wraps onto itself, and lane 0 adds it alone.
__global__ __launch_bounds__(64) void k_ibm_spread_2d(const float* __restrict__ mx, const float* __restrict__ my,
                                                      const float* __restrict__ mfx, const float* __restrict__ mfy,
                                                      const float* __restrict__ mds, float* __restrict__ gx,
                                                      float* __restrict__ gy, int nx, int ny, int nm) {
  const int l = threadIdx.x;
  // stencil cells this lane adds: [c0, c1)
  const int c0 = nx < 5 ? 0 : l;
  const int c1 = nx < 5 ? (l == 0 ? 25 : 0) : (l < 25 ? l + 1 : 0);
  for (int m = 0; m < nm; ++m) {
    const float x = mx[m], y = my[m];
    const float fx = mfx[m] * mds[m], fy = mfy[m] * mds[m];
    const int x0 = static_cast<int>(floorf(x)) - 2, y0 = static_cast<int>(floorf(y)) - 2;
    for (int c = c0; c < c1; ++c) {
      const int py = y0 + c / 5, px = x0 + c % 5;
      if (py < 0 || py >= ny) continue;
      const float wy = 0.25f * (1.0f + cosf(0.5f * 3.14159265358979323846f * fabsf(static_cast<float>(py) - y)));
      const float w = peskin_1d(static_cast<float>(px) - x) * wy;
      const size_t e = static_cast<size_t>(py) * nx + (px % nx + nx) % nx;
      atomicAdd(gx + e, w * fx);
      atomicAdd(gy + e, w * fy);
    }
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_thermal_collide_d2q9(const float* g_in, float* g_out, const uint8_t* ct, const float* ux,
                                       const float* uy, float* temp, float tau_t, int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  k_thermal_collide_d2q9<<<blocks(count), 256, 0, s>>>(g_in, g_out, ct, ux, uy, temp, 1.0f / tau_t, count);
  return hipGetLastError();
}
hipError_t launch_thermal_buoyancy_2d(const float* temp, float* gy, float t_ref, float beta, float g_gravity,
                                      int count, bool accumulate, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (accumulate) k_thermal_buoyancy_2d<true><<<blocks(count), 256, 0, s>>>(temp, gy, t_ref, beta, g_gravity, count);
  else k_thermal_buoyancy_2d<false><<<blocks(count), 256, 0, s>>>(temp, gy, t_ref, beta, g_gravity, count);
  return hipGetLastError();
}
hipError_t launch_thermal_isothermal_wall(float* g, int nx, int ny, int y_wall, float t_wall, hipStream_t s) {
  if (nx <= 0 || y_wall < 0 || y_wall >= ny) return hipErrorInvalidValue;
  k_thermal_isothermal_wall<<<blocks(nx), 256, 0, s>>>(g, nx, ny, y_wall, t_wall);
  return hipGetLastError();
}
hipError_t launch_stream_walls_d2q9(const float* src, float* dst, const uint8_t* ct, int nx, int ny, float uwx,
                                    float uwy, hipStream_t s) {
  const int bx = nx >= 256 ? 256 : 64;
  k_stream_walls<D2Q9><<<dim3((nx + bx - 1) / bx, ny, 1), bx, 0, s>>>(src, dst, ct, nx, ny, 1, uwx, uwy, 0.0f);
  return hipGetLastError();
}

}  // namespace wake_hip

using namespace wake_hip;

extern "C" {

void d2q9_compute_macroscopic_moments(const float f[], float rho[], float ux[], float uy[], int count) {
  moments_call<D2Q9, false>(__func__, f, rho, ux, uy, nullptr, 0, 0, 0, count);
}
void d2q9_compute_macroscopic_moments_forced(const float f[], float rho[], float ux[], float uy[], float gx,
                                             float gy, int count) {
  moments_call<D2Q9, true>(__func__, f, rho, ux, uy, nullptr, gx, gy, 0, count);
}
// analytics_kernels.ispc:4-29
void compute_vorticity_2d(const float ux[], const float uy[], float vorticity[], int nx, int ny) {
  if (nx <= 0 || ny <= 0) return;
  Call c(__func__);
  const size_t n = static_cast<size_t>(nx) * ny;
  auto a = c.in(ux, n);
  auto b = c.in(uy, n);
  auto w = c.out(vorticity, n);
  if (!c.begin()) return;
  (void)launch_vorticity_2d(c.dev(a), c.dev(b), c.dev(w), nx, ny, nullptr);
  c.end();
}
// analytics_kernels.ispc:33-110
void compute_vorticity_and_q_criterion_3d(const float ux[], const float uy[], const float uz[], float omega_x[],
                                          float omega_y[], float omega_z[], float q_criterion[], int nx, int ny,
                                          int nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return;
  Call c(__func__);
  const size_t n = static_cast<size_t>(nx) * ny * nz;
  auto a = c.in(ux, n);
  auto b = c.in(uy, n);
  auto d = c.in(uz, n);
  auto ox = c.out(omega_x, n);
  auto oy = c.out(omega_y, n);
  auto oz = c.out(omega_z, n);
  auto q = c.out(q_criterion, n);
  if (!c.begin()) return;
  (void)launch_vorticity_q_3d(c.dev(a), c.dev(b), c.dev(d), c.dev(ox), c.dev(oy), c.dev(oz), c.dev(q), nx, ny, nz,
                              nullptr);
  c.end();
}
void d2q9_compute_velocity_magnitude(const float ux[], const float uy[], float speed[], int count) {
  if (count <= 0) return;
  Call c(__func__);
  auto a = c.in(ux, count);
  auto b = c.in(uy, count);
  auto s = c.out(speed, count);
  if (!c.begin()) return;
  k_speed<<<blocks(count), 256>>>(c.dev(a), c.dev(b), nullptr, c.dev(s), count);
  c.end();
}
void d2q9_collide_bgk(const float f_in[], float f_out[], const uint8_t cell_type[], float rho[], float ux[],
                      float uy[], float tau, float gx, float gy, int count) {
  collide_call<D2Q9, kBGK>(__func__, f_in, f_out, cell_type, rho, ux, uy, nullptr, tau, 0, gx, gy, 0, count);
}
void d2q9_collide_trt(const float f_in[], float f_out[], const uint8_t cell_type[], float rho[], float ux[],
                      float uy[], float tau_s, float tau_a, float gx, float gy, int count) {
  collide_call<D2Q9, kTRT>(__func__, f_in, f_out, cell_type, rho, ux, uy, nullptr, tau_s, tau_a, gx, gy, 0, count);
}
void d2q9_collide_mrt(const float f_in[], float f_out[], const uint8_t cell_type[], float rho[], float ux[],
                      float uy[], float tau_shear, float tau_bulk, float gx, float gy, int count) {
  collide_call<D2Q9, kMRT>(__func__, f_in, f_out, cell_type, rho, ux, uy, nullptr, tau_shear, tau_bulk, gx, gy, 0,
                           count);
}
void d2q9_stream_periodic(const float f_src[], float f_dst[], int nx, int ny) {
  stream_periodic_call<D2Q9>(__func__, f_src, f_dst, nx, ny, 1);
}
void d2q9_stream_with_boundaries(const float f_src[], float f_dst[], const uint8_t cell_type[], int nx, int ny,
                                 float uw_x, float uw_y) {
  stream_walls_call<D2Q9>(__func__, f_src, f_dst, cell_type, nx, ny, 1, uw_x, uw_y, 0.0f);
}
void d2q9_apply_bounce_back(float f[], const uint8_t cell_type[], float uw_x, float uw_y, int count) {
  if (count <= 0) return;
  Call c(__func__);
  auto fi = c.inout(f, 9 * static_cast<size_t>(count));
  auto t = c.in(cell_type, count);
  if (!c.begin()) return;
  k_bounce_back_d2q9<<<blocks(count), 256>>>(c.dev(fi), c.dev(t), uw_x, uw_y, count);
  c.end();
}
void d2q9_apply_zou_he_inflow_west(float f[], float rho[], float ux[], float uy[], float u_in_x, float u_in_y,
                                   int nx, int ny) {
  if (nx <= 0 || ny <= 2) return;  // rows 1 .. ny-2 only
  const size_t N = static_cast<size_t>(nx) * ny;
  Call c(__func__);
  auto fi = c.inout(f, 9 * N);
  auto r = c.inout(rho, N);
  auto a = c.inout(ux, N);
  auto b = c.inout(uy, N);
  if (!c.begin()) return;
  if (ny > 2) k_zh_in_d2q9<<<blocks(ny - 2), 256>>>(c.dev(fi), c.dev(r), c.dev(a), c.dev(b), u_in_x, u_in_y, nx, ny);
  c.end();
}
void d2q9_apply_zou_he_outflow_east(float f[], float rho[], float ux[], float uy[], float rho_out, int nx,
                                    int ny) {
  if (nx <= 0 || ny <= 2) return;  // rows 1 .. ny-2 only
  const size_t N = static_cast<size_t>(nx) * ny;
  Call c(__func__);
  auto fi = c.inout(f, 9 * N);
  auto r = c.inout(rho, N);
  auto a = c.inout(ux, N);
  auto b = c.inout(uy, N);
  if (!c.begin()) return;
  if (ny > 2) k_zh_out_d2q9<<<blocks(ny - 2), 256>>>(c.dev(fi), c.dev(r), c.dev(a), c.dev(b), rho_out, nx, ny);
  c.end();
}
void d2q9_apply_convective_outflow_east(float f[], const float f_prev[], float u_conv, int nx, int ny) {
  if (nx <= 0 || ny <= 2) return;  // rows 1 .. ny-2 only
  const size_t N = static_cast<size_t>(nx) * ny;
  Call c(__func__);
  auto fi = c.inout(f, 9 * N);
  auto p = c.in(f_prev, 9 * N);
  if (!c.begin()) return;
  if (ny > 2) k_convective_d2q9<<<blocks(ny - 2), 256>>>(c.dev(fi), c.dev(p), u_conv, nx, ny);
  c.end();
}
void d2q9_compute_smagorinsky_tau(const float f[], const float rho[], const float ux[], const float uy[],
                                  float tau_eff[], float tau_0, float smagorinsky_constant, int count) {
  les_call<D2Q9>(__func__, f, rho, ux, uy, nullptr, tau_eff, tau_0, smagorinsky_constant, count);
}
void d2q9_compute_momentum_exchange_forces(const float f_post_collision[], const uint8_t cell_type[], int nx,
                                           int ny, float center_x, float center_y, float max_radius,
                                           float* fx_out, float* fy_out, float* torque_z_out) {
  mem_grid<D2Q9>(__func__, f_post_collision, cell_type, nx, ny, 1, center_x, center_y, 0.0f, max_radius, fx_out,
                 fy_out, torque_z_out);
}

void d3q19_compute_macroscopic_moments(const float f[], float rho[], float ux[], float uy[], float uz[],
                                       int count) {
  moments_call<D3Q19, false>(__func__, f, rho, ux, uy, uz, 0, 0, 0, count);
}
void d3q19_compute_macroscopic_moments_forced(const float f[], float rho[], float ux[], float uy[], float uz[],
                                              float gx, float gy, float gz, int count) {
  moments_call<D3Q19, true>(__func__, f, rho, ux, uy, uz, gx, gy, gz, count);
}
void d3q19_compute_velocity_magnitude(const float ux[], const float uy[], const float uz[], float speed[],
                                      int count) {
  if (count <= 0) return;
  Call c(__func__);
  auto a = c.in(ux, count);
  auto b = c.in(uy, count);
  auto z = c.in(uz, count);
  auto s = c.out(speed, count);
  if (!c.begin()) return;
  k_speed<<<blocks(count), 256>>>(c.dev(a), c.dev(b), c.dev(z), c.dev(s), count);
  c.end();
}
void d3q19_collide_bgk(const float f_in[], float f_out[], const uint8_t cell_type[], float rho[], float ux[],
                       float uy[], float uz[], float tau, float gx, float gy, float gz, int count) {
  collide_call<D3Q19, kBGK>(__func__, f_in, f_out, cell_type, rho, ux, uy, uz, tau, 0, gx, gy, gz, count);
}
void d3q19_collide_trt(const float f_in[], float f_out[], const uint8_t cell_type[], float rho[], float ux[],
                       float uy[], float uz[], float tau_s, float tau_a, float gx, float gy, float gz, int count) {
  collide_call<D3Q19, kTRT>(__func__, f_in, f_out, cell_type, rho, ux, uy, uz, tau_s, tau_a, gx, gy, gz, count);
}
void d3q19_collide_mrt(const float f_in[], float f_out[], const uint8_t cell_type[], float rho[], float ux[],
                       float uy[], float uz[], float tau_shear, float tau_bulk, float gx, float gy, float gz,
                       int count) {
  collide_call<D3Q19, kMRT>(__func__, f_in, f_out, cell_type, rho, ux, uy, uz, tau_shear, tau_bulk, gx, gy, gz,
                            count);
}
void d3q19_stream_periodic(const float f_src[], float f_dst[], int nx, int ny, int nz) {
  stream_periodic_call<D3Q19>(__func__, f_src, f_dst, nx, ny, nz);
}
void d3q19_stream_with_boundaries(const float f_src[], float f_dst[], const uint8_t cell_type[], int nx, int ny,
                                  int nz, float uw_x, float uw_y, float uw_z) {
  stream_walls_call<D3Q19>(__func__, f_src, f_dst, cell_type, nx, ny, nz, uw_x, uw_y, uw_z);
}
void d3q19_compute_smagorinsky_tau(const float f[], const float rho[], const float ux[], const float uy[],
                                   const float uz[], float tau_eff[], float tau_0, float smagorinsky_constant,
                                   int count) {
  les_call<D3Q19>(__func__, f, rho, ux, uy, uz, tau_eff, tau_0, smagorinsky_constant, count);
}
void d3q19_apply_curved_bfl(float f_dst[], const float f_src[], const int fluid_indices[], const int q_dirs[],
                            const float delta_fractions[], const float uw_x[], const float uw_y[],
                            const float uw_z[], int nx, int ny, int nz, int num_boundary_links) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || num_boundary_links <= 0) return;
  const size_t N = static_cast<size_t>(nx) * ny * nz;
  const size_t L = static_cast<size_t>(num_boundary_links);
  // The kernel addresses f by the caller's node indices and directions:
  // check them here (the reference reads out of its arrays on a bad link).
  if (!fluid_indices || !q_dirs || !delta_fractions || !uw_x || !uw_y || !uw_z) {
    record_error("wake: d3q19_apply_curved_bfl: null link array");
    return;
  }
  for (size_t k = 0; k < L; ++k)
    if (fluid_indices[k] < 0 || static_cast<size_t>(fluid_indices[k]) >= N || q_dirs[k] < 0 || q_dirs[k] >= 19) {
      record_error("wake: d3q19_apply_curved_bfl: link " + std::to_string(k) + " has a node index outside [0, " +
                   std::to_string(N) + ") or a direction outside [0, 19)");
      return;
    }
  Call c(__func__);
  auto d = c.inout(f_dst, 19 * N);
  auto s = c.in(f_src, 19 * N);
  auto i = c.in(fluid_indices, L);
  auto q = c.in(q_dirs, L);
  auto dl = c.in(delta_fractions, L);
  auto ax = c.in(uw_x, L);
  auto ay = c.in(uw_y, L);
  auto az = c.in(uw_z, L);
  if (!c.begin()) return;
  k_bfl_d3q19<<<blocks(static_cast<long long>(L)), 256>>>(c.dev(d), c.dev(s), c.dev(i), c.dev(q), c.dev(dl),
                                                            c.dev(ax), c.dev(ay), c.dev(az), nx, ny, nz,
                                                            num_boundary_links);
  c.end();
}
void d3q19_compute_momentum_exchange_forces(const float f_post_collision[], const uint8_t cell_type[], int nx,
                                            int ny, int nz, float center_x, float center_y, float center_z,
                                            float max_radius, float* fx_out, float* fy_out, float* fz_out) {
  mem_grid<D3Q19>(__func__, f_post_collision, cell_type, nx, ny, nz, center_x, center_y, center_z, max_radius, fx_out,
                  fy_out, fz_out);
}

void d3q27_collide(const float f_in[], float f_out[], const uint8_t cell_type[], float rho[], float ux[],
                   float uy[], float uz[], float tau_s, float tau_a, float gx, float gy, float gz, int count) {
  collide_call<D3Q27, kTRT, true>(__func__, f_in, f_out, cell_type, rho, ux, uy, uz, tau_s, tau_a, gx, gy, gz,
                                  count);
}
void d3q27_stream_periodic(const float f_src[], float f_dst[], int nx, int ny, int nz) {
  stream_periodic_call<D3Q27>(__func__, f_src, f_dst, nx, ny, nz);
}
void d3q27_compute_macroscopic_moments(const float f[], float rho[], float ux[], float uy[], float uz[],
                                       int count) {
  moments_call<D3Q27, false>(__func__, f, rho, ux, uy, uz, 0, 0, 0, count);
}

// thermal_kernels.ispc:4-44
void thermal_collide_d2q9(const float g_in[], float g_out[], const uint8_t cell_type[], const float ux[],
                          const float uy[], float temperature[], float tau_T, int count) {
  if (count <= 0) return;
  Call c(__func__);
  const size_t n = static_cast<size_t>(count);
  auto gi = c.in(g_in, 9 * n);
  auto go = c.out(g_out, 9 * n);
  auto t = c.in(cell_type, n);
  auto a = c.in(ux, n);
  auto b = c.in(uy, n);
  // solid nodes leave their temperature untouched: stage the caller's array
  auto T = c.inout(temperature, n);
  if (!c.begin()) return;
  (void)launch_thermal_collide_d2q9(c.dev(gi), c.dev(go), c.dev(t), c.dev(a), c.dev(b), c.dev(T), tau_T, count,
                                    nullptr);
  c.end();
}
// thermal_kernels.ispc:48-60
void thermal_compute_buoyancy_force_2d(const float temperature[], float gy_body[], float t_ref, float beta,
                                       float g_gravity, int count) {
  if (count <= 0) return;
  Call c(__func__);
  const size_t n = static_cast<size_t>(count);
  auto T = c.in(temperature, n);
  auto gy = c.inout(gy_body, n);
  if (!c.begin()) return;
  (void)launch_thermal_buoyancy_2d(c.dev(T), c.dev(gy), t_ref, beta, g_gravity, count, true, nullptr);
  c.end();
}
// thermal_kernels.ispc:63-81 (a wall row outside [0, ny) is an error here)
void thermal_apply_isothermal_wall(float g[], int nx, int ny, int y_wall, float t_wall) {
  if (nx <= 0 || ny <= 0 || y_wall < 0 || y_wall >= ny) {
    record_error("wake: thermal_apply_isothermal_wall: y_wall outside the grid");
    return;
  }
  Call c(__func__);
  auto gi = c.inout(g, 9 * static_cast<size_t>(nx) * ny);
  if (!c.begin()) return;
  (void)launch_thermal_isothermal_wall(c.dev(gi), nx, ny, y_wall, t_wall, nullptr);
  c.end();
}
// ibm_kernels.ispc:12-50
void ibm_interpolate_velocity_2d(const float ux[], const float uy[], const float marker_x[], const float marker_y[],
                                 float marker_ux[], float marker_uy[], int nx, int ny, int num_markers) {
  if (num_markers <= 0) return;
  if (!ibm_markers_ok(marker_x, marker_y, nx, ny, num_markers, __func__)) return;
  Call c(__func__);
  const size_t n = static_cast<size_t>(nx) * ny, m = static_cast<size_t>(num_markers);
  auto a = c.in(ux, n);
  auto b = c.in(uy, n);
  auto x = c.in(marker_x, m);
  auto y = c.in(marker_y, m);
  auto ou = c.out(marker_ux, m);
  auto ov = c.out(marker_uy, m);
  if (!c.begin()) return;
  k_ibm_interpolate_2d<<<blocks(num_markers), 256>>>(c.dev(a), c.dev(b), c.dev(x), c.dev(y), c.dev(ou), c.dev(ov), nx,
                                                     ny, num_markers);
  c.end();
}
// ibm_kernels.ispc:54-96
void ibm_spread_force_2d(const float marker_x[], const float marker_y[], const float marker_fx[],
                         const float marker_fy[], const float marker_ds[], float euler_gx[], float euler_gy[], int nx,
                         int ny, int num_markers) {
  if (num_markers <= 0) return;
  if (!ibm_markers_ok(marker_x, marker_y, nx, ny, num_markers, __func__)) return;
  Call c(__func__);
  const size_t n = static_cast<size_t>(nx) * ny, m = static_cast<size_t>(num_markers);
  auto x = c.in(marker_x, m);
  auto y = c.in(marker_y, m);
  auto fx = c.in(marker_fx, m);
  auto fy = c.in(marker_fy, m);
  auto ds = c.in(marker_ds, m);
  auto gx = c.inout(euler_gx, n);
  auto gy = c.inout(euler_gy, n);
  if (!c.begin()) return;
  k_ibm_spread_2d<<<1, 64>>>(c.dev(x), c.dev(y), c.dev(fx), c.dev(fy), c.dev(ds), c.dev(gx), c.dev(gy), nx, ny,
                             num_markers);
  c.end();
}


// ------------------------------------------------- allocation-free variant
// d3q19_compute_momentum_exchange_forces on device pointers for callers that
// cannot allocate (code under HIP graph capture, serving loops): the same
// kernels, partition and fixed-order sums, so the same bits, with the
// work-group partials in a file-scope device array instead of a staged
// buffer.  The kernels always write kMemBlocks partials, whatever the grid
// size, so kScratchPartials bounds every call.
constexpr int kScratchPartials = kMemBlocks;
static __device__ __align__(256) double scratch[3 * kScratchPartials];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on the current device (each GPU holds its
// own copy); looked up once per device with that device current.
double* scratch_on_current_device(int* device) {
  if (hipGetDevice(device) != hipSuccess || *device < 0 || *device >= kMaxScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[*device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch_ptr[*device] = static_cast<double*>(p);
  }
  return g_scratch_ptr[*device];
}
}  // namespace

int d3q19_compute_momentum_exchange_forces_noalloc(const float* f_post_collision, const uint8_t* cell_type, int nx,
                                                   int ny, int nz, float center_x, float center_y, float center_z,
                                                   float max_radius, float* forces_out, void* stream_handle) {
  static_assert(kMemBlocks <= kScratchPartials, "the partials of one call exceed the scratch array");
  const char* name = __func__;
  hipStream_t stream = static_cast<hipStream_t>(stream_handle);
  auto fail = [&](const std::string& what) {
    record_error(std::string("wake: ") + name + ": " + what);
    return 1;
  };
  if (!f_post_collision || !cell_type || !forces_out) return fail("null device pointer");
  if (nx <= 0 || ny <= 0 || nz <= 0) return fail("grid sizes must be positive");
  const long long N = static_cast<long long>(nx) * ny * nz;
  // the kernel's node coordinates are 32-bit
  if (N > 0x7fffffffll) return fail("more than 2^31 - 1 nodes");
  int device = -1;
  double* part = scratch_on_current_device(&device);
  if (!part) return fail("no scratch address on device " + std::to_string(device));
  k_mem_grid<D3Q19><<<kMemBlocks, 256, 0, stream>>>(f_post_collision, cell_type, nx, ny, nz, center_x, center_y,
                                                    center_z, max_radius, part);
  k_sum_partials<<<1, 64, 0, stream>>>(part, kMemBlocks, forces_out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(std::string("kernel launch: ") + hipGetErrorString(e));
  return 0;
}

}  // extern "C"
