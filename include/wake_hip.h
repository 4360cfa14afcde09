/*
 * wake_hip.h -- C ABI of the MI355X (gfx950) lattice-Boltzmann step.
 *
 * Part 1 replaces the reference's LBM kernels one for one: every function
 * has the name, parameter list and argument meaning of the ISPC export it
 * replaces (cited as file:line in the reference tree), takes host pointers
 * in the reference's SoA layout (population q of node i at q * count + i,
 * node i = (z * ny + y) * nx + x), and runs on the GPU through staged device
 * buffers.  Like the ISPC exports they return nothing; a failure of the GPU
 * runtime is reported on stderr and by wake_hip_last_error().  The
 * reference's consumers declare these names themselves in
 * `namespace ispc { extern "C" { ... } }`, so linking libwake_hip.so in place
 * of the ISPC objects is the drop-in.
 *
 * Part 2 is the device-resident path behind the reference's C++ grid API
 * (include/wake/grid/lattice_grid_3d.hpp:21-123 and lattice_grid_2d.hpp:14-113,
 * which the C++ classes in include/wake/grid/ keep): populations stay in HBM
 * between steps and one fused kernel per step pulls, bounces back, applies
 * BFL links and open boundaries, and collides.  Functions return 0 on success
 * and a non-zero code on failure, with the message in wake_hip_last_error().
 *
 * Plain pointers and sizes only.  Floating point is float (the reference's
 * real_t, include/wake/core/types.hpp:12); cell types are the reference's
 * uint8 codes (types.hpp:69-77): 0 Fluid, 1 Solid, 2 Inflow, 3 Outflow,
 * 4 Periodic, 5 Symmetry, 6 MovingWall.
 */
#ifndef WAKE_HIP_H
#define WAKE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------------
 * Library
 * -------------------------------------------------------------------- */
/* Message of the last failure on this thread ("" if none). */
const char* wake_hip_last_error(void);
/* Clear the last failure; returns the previous error code (0 = none). */
int wake_hip_clear_error(void);
/* Number of visible GPUs, or a negative value if the runtime failed. */
int wake_hip_device_count(void);
/* The D3Q19 MRT back-transform table the kernels compile in (19 x 19 row
 * major): inv(M) rounded to 7 decimals, src/ispc/d3q19_collision_mrt.ispc:4-24.
 * Host-only; needs no GPU. */
void wake_hip_d3q19_minv(float out[361]);
/* Practical HBM ceilings on the current GPU, for comparison with the step
 * kernel; no counterpart in the reference (a measurement utility).  Times
 * `reps` launches (after 3 warm-up launches) of a copy moving 19 values in
 * and out per node of an nx x ny x nz grid and returns the average
 * milliseconds per launch.
 *   shape 0: a contiguous copy of 19 * nx * ny * nz floats in the fastest
 *            copy shape (152 bytes per node);
 *   shape 1: the D3Q19 pull over 19 SoA planes of floats (population q from
 *            x - c_q, periodic), the FP32 step's memory access without
 *            arithmetic (152 bytes per node);
 *   shape 2: the same pull with 16-bit values, the FP16 / BF16 step's
 *            access (76 bytes per node);
 *   shapes 3, 4: the D3Q27 pull over 27 planes, floats (216 bytes per node)
 *            and 16-bit values (108 bytes per node);
 *   shape 5: the D2Q9 pull over 9 planes of floats (72 bytes per node), nz
 *            independent x-y layers.
 * Allocates and frees two buffers of 76 (shapes 3, 4: 108; shape 5: 36) *
 * nx * ny * nz bytes.  Needs
 * 1 <= nx <= 1024 and 1 <= nz <= 65535; returns 0 or an error code. */
int wake_hip_hbm_probe(int shape, int nx, int ny, int nz, int reps, float* ms_per_launch);

/* ----------------------------------------------------------------------
 * Part 1: the reference's LBM kernel exports (host pointers)
 * -------------------------------------------------------------------- */
/* src/ispc/d2q9_moments.ispc:4 */
void d2q9_compute_macroscopic_moments(const float f[], float rho[], float ux[], float uy[], int count);
/* src/ispc/d2q9_moments.ispc:35 */
void d2q9_compute_macroscopic_moments_forced(const float f[], float rho[], float ux[], float uy[], float gx,
                                             float gy, int count);
/* src/ispc/d2q9_moments.ispc:68 */
void d2q9_compute_velocity_magnitude(const float ux[], const float uy[], float speed[], int count);
/* src/ispc/d2q9_collision_bgk.ispc:5 */
void d2q9_collide_bgk(const float f_in[], float f_out[], const uint8_t cell_type[], float rho[], float ux[],
                      float uy[], float tau, float gx, float gy, int count);
/* src/ispc/d2q9_collision_trt.ispc:7 */
void d2q9_collide_trt(const float f_in[], float f_out[], const uint8_t cell_type[], float rho[], float ux[],
                      float uy[], float tau_s, float tau_a, float gx, float gy, int count);
/* src/ispc/d2q9_collision_mrt.ispc:6 */
void d2q9_collide_mrt(const float f_in[], float f_out[], const uint8_t cell_type[], float rho[], float ux[],
                      float uy[], float tau_shear, float tau_bulk, float gx, float gy, int count);
/* src/ispc/d2q9_streaming.ispc:4 */
void d2q9_stream_periodic(const float f_src[], float f_dst[], int nx, int ny);
/* src/ispc/d2q9_streaming.ispc:33 */
void d2q9_stream_with_boundaries(const float f_src[], float f_dst[], const uint8_t cell_type[], int nx, int ny,
                                 float uw_x, float uw_y);
/* src/ispc/d2q9_boundaries.ispc:5 */
void d2q9_apply_bounce_back(float f[], const uint8_t cell_type[], float uw_x, float uw_y, int count);
/* src/ispc/d2q9_boundaries.ispc:43 */
void d2q9_apply_zou_he_inflow_west(float f[], float rho[], float ux[], float uy[], float u_in_x, float u_in_y,
                                   int nx, int ny);
/* src/ispc/d2q9_boundaries.ispc:85 */
void d2q9_apply_zou_he_outflow_east(float f[], float rho[], float ux[], float uy[], float rho_out, int nx,
                                    int ny);
/* src/ispc/d2q9_boundaries.ispc:125 */
void d2q9_apply_convective_outflow_east(float f[], const float f_prev[], float u_conv, int nx, int ny);
/* src/ispc/d2q9_les.ispc:6 */
void d2q9_compute_smagorinsky_tau(const float f[], const float rho[], const float ux[], const float uy[],
                                  float tau_eff[], float tau_0, float smagorinsky_constant, int count);
/* src/ispc/forces_momentum_exchange.ispc:6 (the ISPC `uniform float &`
 * outputs are pointers at the ABI level) */
void d2q9_compute_momentum_exchange_forces(const float f_post_collision[], const uint8_t cell_type[], int nx,
                                           int ny, float center_x, float center_y, float max_radius,
                                           float* fx_out, float* fy_out, float* torque_z_out);

/* src/ispc/d3q19_moments.ispc:4 */
void d3q19_compute_macroscopic_moments(const float f[], float rho[], float ux[], float uy[], float uz[],
                                       int count);
/* src/ispc/d3q19_moments.ispc:35 */
void d3q19_compute_macroscopic_moments_forced(const float f[], float rho[], float ux[], float uy[], float uz[],
                                              float gx, float gy, float gz, int count);
/* src/ispc/d3q19_moments.ispc:69 */
void d3q19_compute_velocity_magnitude(const float ux[], const float uy[], const float uz[], float speed[],
                                      int count);
/* src/ispc/d3q19_collision_bgk.ispc:4 */
void d3q19_collide_bgk(const float f_in[], float f_out[], const uint8_t cell_type[], float rho[], float ux[],
                       float uy[], float uz[], float tau, float gx, float gy, float gz, int count);
/* src/ispc/d3q19_collision_trt.ispc:5 */
void d3q19_collide_trt(const float f_in[], float f_out[], const uint8_t cell_type[], float rho[], float ux[],
                       float uy[], float uz[], float tau_s, float tau_a, float gx, float gy, float gz, int count);
/* src/ispc/d3q19_collision_mrt.ispc:26 */
void d3q19_collide_mrt(const float f_in[], float f_out[], const uint8_t cell_type[], float rho[], float ux[],
                       float uy[], float uz[], float tau_shear, float tau_bulk, float gx, float gy, float gz,
                       int count);
/* src/ispc/d3q19_streaming.ispc:4 */
void d3q19_stream_periodic(const float f_src[], float f_dst[], int nx, int ny, int nz);
/* src/ispc/d3q19_streaming.ispc:39 */
void d3q19_stream_with_boundaries(const float f_src[], float f_dst[], const uint8_t cell_type[], int nx, int ny,
                                  int nz, float uw_x, float uw_y, float uw_z);
/* src/ispc/d3q19_les.ispc:7 */
void d3q19_compute_smagorinsky_tau(const float f[], const float rho[], const float ux[], const float uy[],
                                   const float uz[], float tau_eff[], float tau_0, float smagorinsky_constant,
                                   int count);
/* src/ispc/d3q19_curved.ispc:7 */
void d3q19_apply_curved_bfl(float f_dst[], const float f_src[], const int fluid_indices[], const int q_dirs[],
                            const float delta_fractions[], const float uw_x[], const float uw_y[],
                            const float uw_z[], int nx, int ny, int nz, int num_boundary_links);
/* src/ispc/forces_momentum_exchange.ispc:71 */
void d3q19_compute_momentum_exchange_forces(const float f_post_collision[], const uint8_t cell_type[], int nx,
                                            int ny, int nz, float center_x, float center_y, float center_z,
                                            float max_radius, float* fx_out, float* fy_out, float* fz_out);

/* No counterpart in the reference: d3q19_compute_momentum_exchange_forces for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  All pointers are
 * device pointers; forces_out receives fx, fy, fz (3 floats) in stream order
 * on `stream` (a hipStream_t; NULL: the default stream); the call returns 0 at once, or an error code (message in
 * wake_hip_last_error()) for null pointers, a non-positive size or more than
 * 2^31 - 1 nodes.  The same kernels and sums as the export above, so the
 * same bits.  The work-group partials live in a fixed 6 KiB device array
 * (256 partials of 3 doubles, the number every call writes, held on each
 * GPU until the process ends): calls that overlap on one GPU race.  The
 * first call on a device also looks up that array's address (no
 * allocation); make it outside any graph capture. */
int d3q19_compute_momentum_exchange_forces_noalloc(const float* f_post_collision, const uint8_t* cell_type, int nx,
                                                   int ny, int nz, float center_x, float center_y, float center_z,
                                                   float max_radius, float* forces_out, void* stream);

/* src/ispc/d3q27_collision_mrt.ispc:5 (TRT; pairs populations with the
 * reference's opposite table, include/wake/core/lattice_models.hpp:104-109) */
void d3q27_collide(const float f_in[], float f_out[], const uint8_t cell_type[], float rho[], float ux[],
                   float uy[], float uz[], float tau_s, float tau_a, float gx, float gy, float gz, int count);
/* src/ispc/d3q27_streaming.ispc:4 */
void d3q27_stream_periodic(const float f_src[], float f_dst[], int nx, int ny, int nz);
/* src/ispc/d3q27_moments.ispc:4 */
void d3q27_compute_macroscopic_moments(const float f[], float rho[], float ux[], float uy[], float uz[],
                                       int count);

/* src/ispc/analytics_kernels.ispc:4 */
void compute_vorticity_2d(const float ux[], const float uy[], float vorticity[], int nx, int ny);
/* src/ispc/analytics_kernels.ispc:33 */
void compute_vorticity_and_q_criterion_3d(const float ux[], const float uy[], const float uz[], float omega_x[],
                                          float omega_y[], float omega_z[], float q_criterion[], int nx, int ny,
                                          int nz);

/* src/ispc/thermal_kernels.ispc:4 (temperature written at non-solid nodes) */
void thermal_collide_d2q9(const float g_in[], float g_out[], const uint8_t cell_type[], const float ux[],
                          const float uy[], float temperature[], float tau_T, int count);
/* src/ispc/thermal_kernels.ispc:48 */
void thermal_compute_buoyancy_force_2d(const float temperature[], float gy_body[], float t_ref, float beta,
                                       float g_gravity, int count);
/* src/ispc/thermal_kernels.ispc:63 (y_wall outside [0, ny) is reported as an
 * error instead of writing outside the array) */
void thermal_apply_isothermal_wall(float g[], int nx, int ny, int y_wall, float t_wall);
/* src/ispc/ibm_kernels.ispc:12 */
void ibm_interpolate_velocity_2d(const float ux[], const float uy[], const float marker_x[], const float marker_y[],
                                 float marker_ux[], float marker_uy[], int nx, int ny, int num_markers);
/* src/ispc/ibm_kernels.ispc:54 (markers added in order: deterministic sums) */
void ibm_spread_force_2d(const float marker_x[], const float marker_y[], const float marker_fx[],
                         const float marker_fy[], const float marker_ds[], float euler_gx[], float euler_gy[], int nx,
                         int ny, int num_markers);

/* ----------------------------------------------------------------------
 * Part 2: device-resident domains (the LatticeGrid2D / LatticeGrid3D path)
 * -------------------------------------------------------------------- */
typedef struct wake_domain wake_domain;

enum wake_collision { WAKE_BGK = 0, WAKE_TRT = 1, WAKE_MRT = 2 };
enum wake_storage { WAKE_STORE_F32 = 0, WAKE_STORE_F16 = 1, WAKE_STORE_BF16 = 2 };
enum wake_field {
  WAKE_FIELD_RHO = 0,
  WAKE_FIELD_UX = 1,
  WAKE_FIELD_UY = 2,
  WAKE_FIELD_UZ = 3,
  WAKE_FIELD_SPEED = 4,
  WAKE_FIELD_F_SRC = 5, /* Q planes of nx*ny*nz: the post-stream populations */
  WAKE_FIELD_F_DST = 6, /* Q planes: the post-collision populations of the last step */
  WAKE_FIELD_TEMPERATURE = 7, /* thermal domains: T of the post-stream temperature populations */
  /* vorticity and Q-criterion of u (analytics_kernels.ispc:4-110, computed on
   * the device); a 2D domain has WAKE_FIELD_OMEGA_Z only */
  WAKE_FIELD_OMEGA_X = 8,
  WAKE_FIELD_OMEGA_Y = 9,
  WAKE_FIELD_OMEGA_Z = 10,
  WAKE_FIELD_Q_CRITERION = 11
};

/* LatticeGrid3D(nx, ny, nz, tau) (lattice_grid_3d.cpp:46-64) for dim = 3,
 * LatticeGrid2D(nx, ny, tau) (lattice_grid_2d.cpp:46-62) for dim = 2.
 * lattice: 9 (2D), 19 or 27 (3D), 0 for the default of the dimension (9, 19);
 * any other value is an error.  storage: WAKE_STORE_F32, or FP16 / BF16
 * storage of f_q - w_q with FP32 arithmetic.  plane_pad < 0 picks the
 * default padding of the population plane stride. */
/* Device bytes a domain of this shape allocates (with the temperature
 * populations if thermal != 0); host-only, for admission control. */
long long wake_domain_bytes_needed(int dim, int lattice, int nx, int ny, int nz, int storage, long long plane_pad,
                                   int thermal);
int wake_domain_create(int dim, int lattice, int nx, int ny, int nz, float tau, int storage, long long plane_pad,
                       wake_domain** out);
void wake_domain_destroy(wake_domain* d);
/* lattice_grid_3d.hpp:31 initialize */
int wake_domain_initialize(wake_domain* d, float rho, float ux, float uy, float uz);
/* lattice_grid_3d.hpp:32 initialize_custom, with the per-node values given as
 * arrays of nx*ny*nz (uz may be NULL) */
int wake_domain_initialize_fields(wake_domain* d, const float* rho, const float* ux, const float* uy,
                                  const float* uz);
/* Seeded uniform noise U(-a, a) on every population (counter = node*Q + q). */
int wake_domain_add_noise(wake_domain* d, float amplitude, uint64_t seed);
/* Replace the post-stream state by Q planes of nx*ny*nz (reference layout). */
int wake_domain_set_populations(wake_domain* d, const float* f);
/* lattice_grid_3d.hpp:50 set_cell_type / :51 get_cell_type */
int wake_domain_set_cell_type(wake_domain* d, int x, int y, int z, int type);
int wake_domain_get_cell_type(wake_domain* d, int x, int y, int z);
/* lattice_grid_3d.hpp:47 set_periodic */
int wake_domain_set_periodic(wake_domain* d, int periodic);
/* Wall streaming with y and/or z wrapped instead of reflected (x always wraps). */
int wake_domain_set_periodic_axes(wake_domain* d, int per_y, int per_z);
/* lattice_grid_3d.hpp:53-56, lattice_grid_2d.hpp:47-51 */
int wake_domain_set_solid_box(wake_domain* d, int x0, int y0, int z0, int x1, int y1, int z1);
int wake_domain_set_solid_sphere(wake_domain* d, float cx, float cy, float cz, float r);
/* src/geometry/primitives.cpp:73-100 (GeometryPrimitives::voxelize_ellipsoid)
 * when axes is NULL.  Otherwise axes is a row-major 3x3 whose rows are the
 * ellipsoid's unit body axes in world coordinates (a rotated ellipsoid, no
 * counterpart in the reference). */
int wake_domain_set_solid_ellipsoid(wake_domain* d, float cx, float cy, float cz, float rx, float ry, float rz,
                                    const float* axes);
/* src/geometry/naca_airfoil.cpp:113-190: NACA 4-digit airfoil (digits as an
 * integer, e.g. 12 for "0012"; camber, camber position and thickness as in
 * NacaAirfoil::configure, :23-41), chord and angle of attack in degrees,
 * leading edge at (lead_x, lead_y); rasterize (2D) or voxelize_extruded over
 * z0..z1 (3D). */
int wake_domain_set_solid_naca(wake_domain* d, int digits, float chord, float aoa_deg, float lead_x, float lead_y,
                               int z0, int z1);
/* src/geometry/obj_loader.cpp:208-400 (ObjLoader::voxelize_into_grid) on
 * the device: tris holds ntri triangles as 9 floats each (v0, v1, v2 xyz,
 * host memory, grid units).  Cells whose unit box [i, i+1]x[j, j+1]x[k, k+1]
 * overlaps a triangle become Solid (the reference's separating-axis test,
 * same FP32 expressions); with fill_interior, so does every cell that no
 * 6-neighbour path through non-Solid cells joins to the outer faces.  The
 * resulting cell types equal the CPU voxelizer's bit for bit. */
int wake_domain_voxelize_mesh(wake_domain* d, const float* tris, long long ntri, int fill_interior);
int wake_domain_set_solid_circle(wake_domain* d, float cx, float cy, float r);
int wake_domain_set_solid_cylinder_z(wake_domain* d, float cx, float cy, float r);
int wake_domain_set_moving_wall_top(wake_domain* d, float u_wall_x, float u_wall_z);
/* lattice_grid_2d.hpp:49-50 set_inflow_west / set_outflow_east (3D: x = 0 and
 * x = nx - 1 planes), plus a parabolic inlet, seeded inlet turbulence and a
 * convective outlet. */
int wake_domain_set_inflow_west(wake_domain* d, float u_in_x, float u_in_y);
int wake_domain_set_inflow_parabolic(wake_domain* d, float u_max);
int wake_domain_set_inflow_turbulence(wake_domain* d, float amplitude, uint64_t seed);
int wake_domain_set_outflow_east(wake_domain* d, float rho_out);
int wake_domain_set_convective_outflow_east(wake_domain* d, float u_conv);
/* use_les steps feed the Smagorinsky tau into the collision when coupled
 * (the reference computes it and discards it, lattice_grid_3d.cpp:312-318). */
int wake_domain_set_les_coupled(wake_domain* d, int coupled);
int wake_domain_set_relaxation_time(wake_domain* d, float tau);
/* lattice_grid_3d.hpp:59 build_curved_sphere_links; *nlinks = link count */
int wake_domain_build_curved_sphere_links(wake_domain* d, float cx, float cy, float cz, float r, int* nlinks);
/* the link list in the reference's order (fluid index, q, delta) */
int wake_domain_get_links(wake_domain* d, int* fluid_index, int* q_dir, float* delta);
/* lattice_grid_3d.hpp:35-42 step(collision, gx, gy, gz, use_les, smag) */
int wake_domain_step(wake_domain* d, int collision, float gx, float gy, float gz, int use_les, float smag);
/* `steps` steps; *ms = device time of the launches (hip events on the
 * domain's stream) or 0 when ms is NULL. */
int wake_domain_run(wake_domain* d, int steps, int collision, float gx, float gy, float gz, int use_les,
                    float smag, float* ms);
int wake_domain_synchronize(wake_domain* d);
/* Average device duration (ms) of the fused step kernel per launch over the
 * last wake_domain_run with ms != NULL: HIP events on the domain's stream
 * around the 1st, 17th, 33rd, ... launch of that run (a pair around every
 * launch would add about 8 us to each step). */
int wake_domain_kernel_time(wake_domain* d, float* avg_ms);
/* lattice_grid_3d.hpp:72-78 rho_data ... f_src_data, copied into out */
int wake_domain_read(wake_domain* d, int field, float* out);
int wake_domain_read_cell_types(wake_domain* d, uint8_t* out);
/* lattice_grid_3d.hpp:44 update_macroscopic */
int wake_domain_update_macroscopic(wake_domain* d);
/* lattice_grid_3d.hpp:86-88: total mass, total kinetic energy, max speed over
 * Fluid nodes, reduced in FP64 on the device */
int wake_domain_diagnostics(wake_domain* d, double out[3]);
/* Momentum exchange on the post-stream populations, FP64 (the call the
 * scenarios make, sphere_3d_wake.cpp:98-105): out = (Fx, Fy, Fz) in 3D and
 * (Fx, Fy, torque_z) in 2D. */
int wake_domain_momentum_exchange(wake_domain* d, float cx, float cy, float cz, float r_max, double out[3]);
/* Per-step momentum exchange (the reference's scenarios take it after every
 * step, sphere_3d_wake.cpp:94-105, karman_vortex.cpp:104-114), computed
 * inside the launch that forms each post-stream state: from the current step
 * on, every step's forces (as wake_domain_momentum_exchange would return them
 * right after that step) are kept on the device.  Setting a probe again
 * restarts the history; so does initialising the state.  r_max <= 0 takes
 * every link.  Not for temperature-coupled domains. */
int wake_domain_set_force_probe(wake_domain* d, float cx, float cy, float cz, float r_max);
int wake_domain_clear_force_probe(wake_domain* d);
/* The history so far: *first_step = step of row 0, *count = rows (steps
 * first_step .. current step); min(*count, max_rows) rows of 3 doubles are
 * copied to out (out may be NULL with max_rows 0 to ask for the count). */
int wake_domain_force_history(wake_domain* d, double* out, long long max_rows, long long* first_step,
                              long long* count);
/* Divergence (no counterpart in the reference, whose scenarios throw
 * NumericalDivergenceException, include/wake/core/error.hpp:14-18, after
 * reading every density on the host).  *seen = 1 if a step since the state
 * was last set produced a non-finite density (a flag the step kernels raise;
 * one 4-byte read). */
int wake_domain_nonfinite_seen(wake_domain* d, int* seen);
/* *count = nodes whose current density or speed is not finite, *first = the
 * first of them (-1 if none), counted on the device without downloading the
 * fields. */
int wake_domain_count_nonfinite(wake_domain* d, long long* count, long long* first);
/* out: nx, ny, nz, Q, plane stride, current step, storage, device bytes */
int wake_domain_info(wake_domain* d, long long out[8]);
/* src/export/vtk_exporter.cpp:8-191: the current fields as a VTK XML
 * ImageData file: Velocity, Speed, Density, [with_analytics: 2D Vorticity;
 * 3D Q_Criterion and Vorticity (3 components), computed on the device],
 * CellType.  binary = 1: raw little-endian appended arrays with UInt64
 * sizes; 0: the reference's ASCII layout. */
int wake_domain_export_vti(wake_domain* d, const char* path, int binary, int with_analytics);
/* lattice_grid_3d.hpp is_periodic(): 1 while the domain streams fully
 * periodic (no non-Fluid cell type was ever set, lattice_grid_3d.cpp:133-140). */
int wake_domain_is_periodic(wake_domain* d, int* out);
/* D2Q5 temperature coupled to the D2Q9 flow by a Boussinesq acceleration
 * a_y = gbeta (T - t_ref) per node (2D only; no counterpart in the reference,
 * whose src/thermal/thermal_solver.cpp:83-123 couples only the mean
 * temperature).  Temperature starts as rest equilibrium at t_init. */
int wake_domain_enable_thermal(wake_domain* d, float tau_g, float gbeta, float t_ref, float t_init);
/* Temperature field (nx * ny, reference node order) as rest equilibrium. */
int wake_domain_set_temperature_field(wake_domain* d, const float* t);
/* Anti-bounce-back wall temperature for the solid nodes in the box
 * (inclusive); NaN makes them adiabatic (bounce-back) again. */
int wake_domain_set_wall_temperature(wake_domain* d, int x0, int y0, int x1, int y1, float t_wall);
/* work-group width along x for the step kernel (64, 128, 192, 256; 0 = auto) */
int wake_domain_set_block_x(wake_domain* d, int block_x);
/* Extension (no reference counterpart): on = 1 (default) lets
 * wake_domain_run take 8 steps per launch on a fully periodic 2D domain
 * whose cells are all Fluid, with no inlet, outlet, force probe or
 * temperature, of up to 2^21 nodes (FP32) or 2^17 (16-bit storage), where
 * single launches are bound by the launch rate; the result is the same as
 * 8 single steps.  0 keeps one launch per step. */
int wake_domain_set_multistep(wake_domain* d, int on);
/* FP16 / BF16 storage: 1 (default) addresses the populations through a
 * buffer descriptor when the buffer is under 4 GiB; 0 always through 64-bit
 * pointers, as larger buffers are.  The results are the same. */
int wake_domain_set_buffer_addressing(wake_domain* d, int enable);

/* The reference's D2Q9 temperature solver (src/thermal/thermal_solver.cpp,
 * class ThermalSolver2D) with its populations in HBM.  A step
 * (thermal_solver.cpp:83-123) collides with the fluid domain's current
 * velocity, applies the isothermal south (y = 0) / north (y = ny - 1) walls,
 * streams with the fluid domain's cell types and a wall velocity of zero
 * (the reference's call leaves those two arguments undefined) and computes
 * the per-node Boussinesq acceleration beta (T - t_ref) g_gravity.  The
 * fluid domain must be 2D of the same size. */
typedef struct wake_thermal2d wake_thermal2d;
int wake_thermal2d_create(int nx, int ny, float tau_thermal, float t_ref, float beta, wake_thermal2d** out);
int wake_thermal2d_destroy(wake_thermal2d* t);
/* thermal_solver.cpp:49-60 */
int wake_thermal2d_initialize(wake_thermal2d* t, float t_init);
/* set_isothermal_wall_south / _north (a flag of 0 leaves that wall as is) */
int wake_thermal2d_set_walls(wake_thermal2d* t, int south, float t_hot, int north, float t_cold);
int wake_thermal2d_step(wake_thermal2d* t, wake_domain* fluid, float g_gravity);
/* which = 0: temperature, 1: buoyancy acceleration; nx * ny floats */
int wake_thermal2d_read(wake_thermal2d* t, int which, float* out);
/* out = {compute_mean_temperature, compute_nusselt_number} */
int wake_thermal2d_stats(wake_thermal2d* t, double out[2]);

/* ----------------------------------------------------------------------
 * Part 3: batched stepping of many domains (no counterpart in the
 * reference, which steps one grid per call of LatticeGrid*::step,
 * src/grid/lattice_grid_2d.cpp:195-265 and lattice_grid_3d.cpp:301-380).
 *
 * A batch steps every member domain with one kernel launch per step.  The
 * members share lattice, population storage, collision and LES coupling and
 * keep everything else of their own (size, cell types, boundaries, tau,
 * body force).  Each member takes exactly the steps it was added with; a
 * member's results are those of stepping it alone.  Steps run in chunks of
 * `chunk` launches captured once as a HIP graph.  A domain must not be
 * stepped, destroyed or added to another batch while it is a member; after
 * wake_batch_run returns its fields and populations read as after
 * wake_domain_step.
 * -------------------------------------------------------------------- */
typedef struct wake_batch wake_batch;

/* thermal: members are D2Q9 + D2Q5 temperature domains (lattice 9). */
int wake_batch_create(int lattice, int storage, int collision, int use_les, int thermal, int chunk,
                      wake_batch** out);
void wake_batch_destroy(wake_batch* b);
/* Add a member for `steps` more steps with body force (gx, gy, gz) and
 * Smagorinsky constant smag.  track_forces: record the momentum-exchange
 * force on S after every step over the links around (cx, cy, cz) within
 * r_max (3 doubles per step: F (3D) or Fx, Fy, torque (2D)).  *slot gets
 * the member id. */
int wake_batch_add(wake_batch* b, wake_domain* d, long long steps, float gx, float gy, float gz, float smag,
                   int track_forces, float cx, float cy, float cz, float r_max, int* slot);
int wake_batch_remove(wake_batch* b, int slot);
/* `steps` launches (each member takes min(steps, its remaining steps));
 * *ms = device time.  use_graph = 0 launches each step directly. */
int wake_batch_run(wake_batch* b, long long steps, int use_graph, float* ms);
/* Asynchronous form: enqueue one chunk (min(steps, chunk) launches) on the
 * batch's own stream and return; wake_batch_wait completes it (*ms = device
 * time).  Chunks of different batches in flight together overlap on the
 * GPU.  No other call on the batch or its members until the wait. */
int wake_batch_launch(wake_batch* b, long long steps, int use_graph);
int wake_batch_wait(wake_batch* b, float* ms);
int wake_batch_remaining(wake_batch* b, int slot, long long* remaining);
/* Force history since the member was added: *count steps, up to `capacity`
 * of them copied to out (3 doubles each). */
int wake_batch_forces(wake_batch* b, int slot, double* out, long long capacity, long long* count);
/* Autotuner: times probe_steps real steps under each candidate launch shape
 * (work-group size; one work-group per tile or a persistent grid of
 * grid_factor x compute units) and keeps the fastest.  The steps count as
 * steps of the members. */
int wake_batch_tune(wake_batch* b, int probe_steps, int* block, int* grid_factor);
/* Fix the launch shape.  Until this or wake_batch_tune is called, the
 * work-group width is the one of 256 / 128 / 64 that leaves the fewest idle
 * lanes over the members' rows, with one work-group per tile. */
int wake_batch_set_launch(wake_batch* b, int block, int grid_factor);
/* out = {members, total nodes, block, grid_factor} */
int wake_batch_info(wake_batch* b, long long out[4]);

#ifdef __cplusplus
}
#endif
#endif /* WAKE_HIP_H */
