// analytics.hip -- vorticity and Q-criterion of a velocity field (gfx950).
//
// src/ispc/analytics_kernels.ispc:4-110: second-order central differences
// with periodic wrap on every axis; 2D: omega_z = d uy/dx - d ux/dy; 3D: the
// curl and Q = (||Omega||^2 - ||S||^2) / 2 with S and Omega the symmetric and
// antisymmetric parts of grad u.  One lane per node, lanes along x; the
// y/z neighbours are work-group uniform rows.
#include "engine.hpp"

namespace wake_hip {

__global__ __launch_bounds__(256) void k_vorticity_2d(const float* __restrict__ ux, const float* __restrict__ uy,
                                                      float* __restrict__ w, int nx, int ny) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= nx) return;
  const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
  const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
  const int row = y * nx;
  const float duy_dx = 0.5f * (uy[row + xp] - uy[row + xm]);
  const float dux_dy = 0.5f * (ux[yp * nx + x] - ux[ym * nx + x]);
  w[row + x] = duy_dx - dux_dy;
}

__global__ __launch_bounds__(256) void k_vorticity_q_3d(const float* __restrict__ ux, const float* __restrict__ uy,
                                                        const float* __restrict__ uz, float* __restrict__ ox,
                                                        float* __restrict__ oy, float* __restrict__ oz,
                                                        float* __restrict__ qc, int nx, int ny, int nz) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= nx) return;
  const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
  const int zm = z == 0 ? nz - 1 : z - 1, zp = z == nz - 1 ? 0 : z + 1;
  const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
  const long long c = (static_cast<long long>(z) * ny + y) * nx;
  const long long cyp = (static_cast<long long>(z) * ny + yp) * nx, cym = (static_cast<long long>(z) * ny + ym) * nx;
  const long long czp = (static_cast<long long>(zp) * ny + y) * nx, czm = (static_cast<long long>(zm) * ny + y) * nx;
  const float dux_dx = 0.5f * (ux[c + xp] - ux[c + xm]);
  const float dux_dy = 0.5f * (ux[cyp + x] - ux[cym + x]);
  const float dux_dz = 0.5f * (ux[czp + x] - ux[czm + x]);
  const float duy_dx = 0.5f * (uy[c + xp] - uy[c + xm]);
  const float duy_dy = 0.5f * (uy[cyp + x] - uy[cym + x]);
  const float duy_dz = 0.5f * (uy[czp + x] - uy[czm + x]);
  const float duz_dx = 0.5f * (uz[c + xp] - uz[c + xm]);
  const float duz_dy = 0.5f * (uz[cyp + x] - uz[cym + x]);
  const float duz_dz = 0.5f * (uz[czp + x] - uz[czm + x]);
  const long long n = c + x;
  if (ox) {
    ox[n] = duz_dy - duy_dz;
    oy[n] = dux_dz - duz_dx;
    oz[n] = duy_dx - dux_dy;
  }
  if (qc) {
    const float s_xy = 0.5f * (dux_dy + duy_dx), s_yz = 0.5f * (duy_dz + duz_dy), s_xz = 0.5f * (dux_dz + duz_dx);
    const float norm_s = (dux_dx * dux_dx + duy_dy * duy_dy + duz_dz * duz_dz) +
                         2.0f * (s_xy * s_xy + s_yz * s_yz + s_xz * s_xz);
    const float o_xy = 0.5f * (dux_dy - duy_dx), o_yz = 0.5f * (duy_dz - duz_dy), o_xz = 0.5f * (dux_dz - duz_dx);
    const float norm_o = 2.0f * (o_xy * o_xy + o_yz * o_yz + o_xz * o_xz);
    qc[n] = 0.5f * (norm_o - norm_s);
  }
}

hipError_t launch_vorticity_2d(const float* ux, const float* uy, float* w, int nx, int ny, hipStream_t s) {
  k_vorticity_2d<<<dim3((nx + 255) / 256, ny, 1), 256, 0, s>>>(ux, uy, w, nx, ny);
  return hipGetLastError();
}

hipError_t launch_vorticity_q_3d(const float* ux, const float* uy, const float* uz, float* ox, float* oy, float* oz,
                                 float* q, int nx, int ny, int nz, hipStream_t s) {
  k_vorticity_q_3d<<<dim3((nx + 255) / 256, ny, nz), 256, 0, s>>>(ux, uy, uz, ox, oy, oz, q, nx, ny, nz);
  return hipGetLastError();
}

}  // namespace wake_hip
