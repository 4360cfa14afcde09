// capi.cpp -- C ABI over wake_hip::Domain (include/wake_hip.h, part 2).
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "wake_hip.h"
#include "capi_internal.hpp"
#include "lattice.hpp"

namespace {
thread_local std::string g_error;
thread_local int g_code = 0;

int fail(const char* what) {
  g_error = what;
  g_code = 1;
  return 1;
}

template <class F>
int guard(wake_domain* d, F&& f) {
  if (!d || !d->d) return fail("wake: null domain handle");
  try {
    f(*d->d);
    return 0;
  } catch (const std::exception& e) {
    return fail(e.what());
  }
}
}  // namespace

namespace wake_hip {
int set_last_error(const std::string& msg) {
  g_error = msg;
  g_code = 1;
  return 1;
}

// Shared with the kernel exports (ispc_compat.hip).
void record_error(const std::string& msg) {
  g_error = msg;
  g_code = 1;
  std::fprintf(stderr, "%s\n", msg.c_str());
}
}  // namespace wake_hip

extern "C" {

const char* wake_hip_last_error(void) { return g_error.c_str(); }

int wake_hip_clear_error(void) {
  const int c = g_code;
  g_code = 0;
  g_error.clear();
  return c;
}

int wake_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return -1;
  return n;
}

void wake_hip_d3q19_minv(float out[361]) {
  for (int q = 0; q < 19; ++q)
    for (int k = 0; k < 19; ++k) out[q * 19 + k] = wake_hip::kD3Q19Minv.minv[q][k];
}

long long wake_domain_bytes_needed(int dim, int lattice, int nx, int ny, int nz, int storage, long long plane_pad,
                                   int thermal) {
  return static_cast<long long>(
      wake_hip::Domain::bytes_needed(dim, lattice, nx, ny, nz, storage, plane_pad, thermal != 0));
}

int wake_domain_create(int dim, int lattice, int nx, int ny, int nz, float tau, int storage, long long plane_pad,
                       wake_domain** out) {
  if (!out) return fail("wake: null output handle");
  *out = nullptr;
  try {
    auto* h = new wake_domain{new wake_hip::Domain(dim, lattice, nx, ny, nz, tau, storage, plane_pad)};
    *out = h;
    return 0;
  } catch (const std::exception& e) {
    return fail(e.what());
  }
}

void wake_domain_destroy(wake_domain* d) {
  if (!d) return;
  delete d->d;
  delete d;
}

int wake_domain_initialize(wake_domain* d, float rho, float ux, float uy, float uz) {
  return guard(d, [&](wake_hip::Domain& g) { g.initialize(rho, ux, uy, uz); });
}
int wake_domain_initialize_fields(wake_domain* d, const float* rho, const float* ux, const float* uy,
                                  const float* uz) {
  return guard(d, [&](wake_hip::Domain& g) { g.initialize_fields(rho, ux, uy, uz); });
}
int wake_domain_add_noise(wake_domain* d, float amplitude, uint64_t seed) {
  return guard(d, [&](wake_hip::Domain& g) { g.add_noise(amplitude, seed); });
}
int wake_domain_set_populations(wake_domain* d, const float* f) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_populations(f); });
}
int wake_domain_set_cell_type(wake_domain* d, int x, int y, int z, int type) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_cell_type(x, y, z, static_cast<uint8_t>(type)); });
}
int wake_domain_get_cell_type(wake_domain* d, int x, int y, int z) {
  if (!d || !d->d) return -1;
  return d->d->get_cell_type(x, y, z);
}
int wake_domain_set_periodic(wake_domain* d, int periodic) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_periodic(periodic != 0); });
}
int wake_domain_set_periodic_axes(wake_domain* d, int per_y, int per_z) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_periodic_axes(per_y != 0, per_z != 0); });
}
int wake_domain_set_solid_box(wake_domain* d, int x0, int y0, int z0, int x1, int y1, int z1) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_solid_box(x0, y0, z0, x1, y1, z1); });
}
int wake_domain_set_solid_sphere(wake_domain* d, float cx, float cy, float cz, float r) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_solid_sphere(cx, cy, cz, r); });
}
int wake_domain_set_solid_ellipsoid(wake_domain* d, float cx, float cy, float cz, float rx, float ry, float rz,
                                    const float* axes) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_solid_ellipsoid(cx, cy, cz, rx, ry, rz, axes); });
}
int wake_domain_set_solid_naca(wake_domain* d, int digits, float chord, float aoa_deg, float lead_x, float lead_y,
                               int z0, int z1) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_solid_naca(digits, chord, aoa_deg, lead_x, lead_y, z0, z1); });
}
int wake_domain_voxelize_mesh(wake_domain* d, const float* tris, long long ntri, int fill_interior) {
  return guard(d, [&](wake_hip::Domain& g) { g.voxelize_mesh(tris, ntri, fill_interior != 0); });
}
int wake_domain_set_solid_circle(wake_domain* d, float cx, float cy, float r) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_solid_circle(cx, cy, r); });
}
int wake_domain_set_solid_cylinder_z(wake_domain* d, float cx, float cy, float r) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_solid_cylinder_z(cx, cy, r); });
}
int wake_domain_set_moving_wall_top(wake_domain* d, float u_wall_x, float u_wall_z) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_moving_wall_top(u_wall_x, u_wall_z); });
}
int wake_domain_set_inflow_west(wake_domain* d, float u_in_x, float u_in_y) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_inflow_west(u_in_x, u_in_y); });
}
int wake_domain_set_inflow_parabolic(wake_domain* d, float u_max) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_inflow_parabolic(u_max); });
}
int wake_domain_set_inflow_turbulence(wake_domain* d, float amplitude, uint64_t seed) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_inflow_turbulence(amplitude, seed); });
}
int wake_domain_set_outflow_east(wake_domain* d, float rho_out) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_outflow_east(rho_out); });
}
int wake_domain_set_convective_outflow_east(wake_domain* d, float u_conv) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_convective_outflow_east(u_conv); });
}
int wake_domain_set_les_coupled(wake_domain* d, int coupled) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_les_coupled(coupled != 0); });
}
int wake_domain_enable_thermal(wake_domain* d, float tau_g, float gbeta, float t_ref, float t_init) {
  return guard(d, [&](wake_hip::Domain& g) { g.enable_thermal(tau_g, gbeta, t_ref, t_init); });
}
int wake_domain_set_temperature_field(wake_domain* d, const float* t) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_temperature_field(t); });
}
int wake_domain_set_wall_temperature(wake_domain* d, int x0, int y0, int x1, int y1, float t_wall) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_wall_temperature(x0, y0, x1, y1, t_wall); });
}
int wake_domain_set_relaxation_time(wake_domain* d, float tau) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_tau(tau); });
}
int wake_domain_build_curved_sphere_links(wake_domain* d, float cx, float cy, float cz, float r, int* nlinks) {
  return guard(d, [&](wake_hip::Domain& g) {
    const int n = g.build_curved_sphere_links(cx, cy, cz, r);
    if (nlinks) *nlinks = n;
  });
}
int wake_domain_get_links(wake_domain* d, int* fluid_index, int* q_dir, float* delta) {
  return guard(d, [&](wake_hip::Domain& g) {
    const auto& k = g.curved_links();
    const size_t n = k.fluid_index.size();
    if (fluid_index) std::memcpy(fluid_index, k.fluid_index.data(), n * sizeof(int));
    if (q_dir) std::memcpy(q_dir, k.q_dir.data(), n * sizeof(int));
    if (delta) std::memcpy(delta, k.delta.data(), n * sizeof(float));
  });
}
int wake_domain_step(wake_domain* d, int collision, float gx, float gy, float gz, int use_les, float smag) {
  return guard(d, [&](wake_hip::Domain& g) { g.step(collision, gx, gy, gz, use_les != 0, smag); });
}
int wake_domain_run(wake_domain* d, int steps, int collision, float gx, float gy, float gz, int use_les,
                    float smag, float* ms) {
  return guard(d, [&](wake_hip::Domain& g) {
    const float t = g.run(steps, collision, gx, gy, gz, use_les != 0, smag, ms != nullptr);
    if (ms) *ms = t;
  });
}
int wake_domain_kernel_time(wake_domain* d, float* avg_ms) {
  return guard(d, [&](wake_hip::Domain& g) { *avg_ms = g.last_kernel_ms(); });
}
int wake_domain_synchronize(wake_domain* d) {
  return guard(d, [&](wake_hip::Domain& g) { g.synchronize(); });
}
int wake_domain_read(wake_domain* d, int field, float* out) {
  return guard(d, [&](wake_hip::Domain& g) {
    const size_t N = g.total_nodes();
    const float* src = nullptr;
    size_t n = N;
    switch (field) {
      case WAKE_FIELD_RHO: src = g.rho(); break;
      case WAKE_FIELD_UX: src = g.ux(); break;
      case WAKE_FIELD_UY: src = g.uy(); break;
      case WAKE_FIELD_UZ: src = g.uz(); break;
      case WAKE_FIELD_SPEED: src = g.speed(); break;
      case WAKE_FIELD_F_SRC: src = g.f_src(); n = N * g.lattice(); break;
      case WAKE_FIELD_F_DST: src = g.f_dst(); n = N * g.lattice(); break;
      case WAKE_FIELD_TEMPERATURE: src = g.temperature(); break;
      case WAKE_FIELD_OMEGA_X: src = g.analytics(0); break;
      case WAKE_FIELD_OMEGA_Y: src = g.analytics(1); break;
      case WAKE_FIELD_OMEGA_Z: src = g.analytics(2); break;
      case WAKE_FIELD_Q_CRITERION: src = g.analytics(3); break;
      default: throw wake_hip::HipError("unknown field");
    }
    std::memcpy(out, src, n * sizeof(float));
  });
}
int wake_domain_read_cell_types(wake_domain* d, uint8_t* out) {
  return guard(d, [&](wake_hip::Domain& g) { std::memcpy(out, g.cell_type(), g.total_nodes()); });
}
int wake_domain_update_macroscopic(wake_domain* d) {
  return guard(d, [&](wake_hip::Domain& g) { g.update_macroscopic(); });
}
int wake_domain_diagnostics(wake_domain* d, double out[3]) {
  return guard(d, [&](wake_hip::Domain& g) {
    g.diagnostics(out);
  });
}
int wake_domain_momentum_exchange(wake_domain* d, float cx, float cy, float cz, float r_max, double out[3]) {
  return guard(d, [&](wake_hip::Domain& g) { g.momentum_exchange(cx, cy, cz, r_max, out); });
}
int wake_domain_set_force_probe(wake_domain* d, float cx, float cy, float cz, float r_max) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_force_probe(cx, cy, cz, r_max); });
}
int wake_domain_clear_force_probe(wake_domain* d) {
  return guard(d, [&](wake_hip::Domain& g) { g.clear_force_probe(); });
}
int wake_domain_force_history(wake_domain* d, double* out, long long max_rows, long long* first_step,
                              long long* count) {
  return guard(d, [&](wake_hip::Domain& g) {
    if (max_rows > 0 && !out) throw wake_hip::HipError("force_history: null output");
    uint64_t first = 0;
    const std::vector<double>& h = g.force_history(&first);
    const long long rows = static_cast<long long>(h.size() / 3);
    if (first_step) *first_step = static_cast<long long>(first);
    if (count) *count = rows;
    const long long n = rows < max_rows ? rows : max_rows;
    if (n > 0) std::memcpy(out, h.data(), static_cast<size_t>(n) * 3 * sizeof(double));
  });
}
int wake_domain_nonfinite_seen(wake_domain* d, int* seen) {
  return guard(d, [&](wake_hip::Domain& g) {
    if (!seen) throw wake_hip::HipError("nonfinite_seen: null output");
    *seen = g.nonfinite_seen() ? 1 : 0;
  });
}
int wake_domain_count_nonfinite(wake_domain* d, long long* count, long long* first) {
  return guard(d, [&](wake_hip::Domain& g) {
    if (!count) throw wake_hip::HipError("count_nonfinite: null output");
    *count = g.count_nonfinite(first);
  });
}
int wake_domain_info(wake_domain* d, long long out[8]) {
  return guard(d, [&](wake_hip::Domain& g) {
    out[0] = g.nx();
    out[1] = g.ny();
    out[2] = g.nz();
    out[3] = g.lattice();
    out[4] = g.plane_stride();
    out[5] = static_cast<long long>(g.current_step());
    out[6] = g.storage();
    out[7] = static_cast<long long>(g.device_bytes());
  });
}
int wake_domain_export_vti(wake_domain* d, const char* path, int binary, int with_analytics) {
  return guard(d, [&](wake_hip::Domain& g) {
    if (!path) throw wake_hip::HipError("export_vti: null path");
    const float *v2 = nullptr, *q = nullptr;
    std::vector<float> v3;
    if (with_analytics) {
      if (g.dim() == 2) {
        v2 = g.analytics(2);
      } else {
        const size_t n = static_cast<size_t>(g.nx()) * g.ny() * g.nz();
        const float *ox = g.analytics(0), *oy = g.analytics(1), *oz = g.analytics(2);
        q = g.analytics(3);
        v3.resize(3 * n);
        for (size_t i = 0; i < n; ++i) {
          v3[3 * i] = ox[i];
          v3[3 * i + 1] = oy[i];
          v3[3 * i + 2] = oz[i];
        }
      }
    }
    if (!wake_hip::write_vti(g, path, binary != 0, v2, q, v3.empty() ? nullptr : v3.data()))
      throw wake_hip::HipError(std::string("export_vti: cannot write ") + path);
  });
}
int wake_domain_is_periodic(wake_domain* d, int* out) {
  return guard(d, [&](wake_hip::Domain& g) { *out = g.is_periodic() ? 1 : 0; });
}
int wake_domain_set_buffer_addressing(wake_domain* d, int enable) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_buffer_addressing(enable != 0); });
}
int wake_domain_set_block_x(wake_domain* d, int block_x) {
  return guard(d, [&](wake_hip::Domain& g) {
    if (block_x != 0 && block_x != 64 && block_x != 128 && block_x != 192 && block_x != 256)
      throw wake_hip::HipError("block_x must be 0, 64, 128, 192 or 256");
    g.set_block_x(block_x);
  });
}
int wake_domain_set_multistep(wake_domain* d, int on) {
  return guard(d, [&](wake_hip::Domain& g) { g.set_multistep(on != 0); });
}

}  // extern "C"
