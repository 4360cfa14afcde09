/*
 * wake_oracle.h -- CPU restatement of the wake lattice-Boltzmann kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * C++ grid, the Python package) links, loads or calls this code; only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark do, and
 * only as the checker.
 *
 * Every function restates one ISPC export of the reference, or one grid
 * method, with its argument meaning kept.  Each definition in wake_oracle.c
 * cites the reference file:line it follows.  The library is compiled twice
 * from the same source: with ORACLE_REAL=float (liboracle32.so, the ISPC-path
 * stand-in) and ORACLE_REAL=double (liboracle64.so, the FP64 oracle).
 *
 * Functions marked EXTENSION have no reference implementation; they are the
 * oracle for features the reference lacks (fully periodic 2D stepping,
 * LES-coupled collision, 3D open boundaries, D3Q27 walls, seeded noise) and
 * are pinned by physics checks in the tests instead of reference vectors.
 */
#ifndef WAKE_ORACLE_H
#define WAKE_ORACLE_H

#include <stdint.h>

#ifndef ORACLE_REAL
#define ORACLE_REAL float
#endif
typedef ORACLE_REAL real;

#ifdef __cplusplus
extern "C" {
#endif

/* ---- tables ---- */
int oracle_real_bytes(void);
int oracle_threads(void); /* OpenMP threads the node loops use */
void oracle_set_threads(int n); /* threads of the following calls (timing) */
void oracle_d3q19_minv(real out[19 * 19]);
void oracle_d3q19_m(real out[19 * 19]);
uint64_t oracle_rng_splitmix64(uint64_t x);
double oracle_rng_u01(uint64_t seed, uint64_t counter);

/* ---- D2Q9 kernels (src/ispc/d2q9_*.ispc) ---- */
void oracle_d2q9_compute_macroscopic_moments(const real* f, real* rho, real* ux, real* uy, int count);
void oracle_d2q9_compute_macroscopic_moments_forced(const real* f, real* rho, real* ux, real* uy,
                                                    real gx, real gy, int count);
void oracle_d2q9_compute_velocity_magnitude(const real* ux, const real* uy, real* speed, int count);
void oracle_d2q9_collide_bgk(const real* f_in, real* f_out, const uint8_t* cell_type, real* rho, real* ux,
                             real* uy, real tau, real gx, real gy, int count);
void oracle_d2q9_collide_trt(const real* f_in, real* f_out, const uint8_t* cell_type, real* rho, real* ux,
                             real* uy, real tau_s, real tau_a, real gx, real gy, int count);
void oracle_d2q9_collide_mrt(const real* f_in, real* f_out, const uint8_t* cell_type, real* rho, real* ux,
                             real* uy, real tau_shear, real tau_bulk, real gx, real gy, int count);
void oracle_d2q9_stream_periodic(const real* f_src, real* f_dst, int nx, int ny);
void oracle_d2q9_stream_with_boundaries(const real* f_src, real* f_dst, const uint8_t* cell_type, int nx,
                                        int ny, real uw_x, real uw_y);
void oracle_d2q9_apply_bounce_back(real* f, const uint8_t* cell_type, real uw_x, real uw_y, int count);
void oracle_d2q9_apply_zou_he_inflow_west(real* f, real* rho, real* ux, real* uy, real u_in_x, real u_in_y,
                                          int nx, int ny);
void oracle_d2q9_apply_zou_he_outflow_east(real* f, real* rho, real* ux, real* uy, real rho_out, int nx,
                                           int ny);
void oracle_d2q9_apply_convective_outflow_east(real* f, const real* f_prev, real u_conv, int nx, int ny);
void oracle_d2q9_compute_smagorinsky_tau(const real* f, const real* rho, const real* ux, const real* uy,
                                         real* tau_eff, real tau_0, real smagorinsky_constant, int count);
void oracle_d2q9_compute_momentum_exchange_forces(const real* f, const uint8_t* cell_type, int nx, int ny,
                                                  real center_x, real center_y, real max_radius,
                                                  real* fx_out, real* fy_out, real* torque_z_out);

/* ---- D3Q19 kernels (src/ispc/d3q19_*.ispc) ---- */
void oracle_d3q19_compute_macroscopic_moments(const real* f, real* rho, real* ux, real* uy, real* uz,
                                              int count);
void oracle_d3q19_compute_macroscopic_moments_forced(const real* f, real* rho, real* ux, real* uy, real* uz,
                                                     real gx, real gy, real gz, int count);
void oracle_d3q19_compute_velocity_magnitude(const real* ux, const real* uy, const real* uz, real* speed,
                                             int count);
void oracle_d3q19_collide_bgk(const real* f_in, real* f_out, const uint8_t* cell_type, real* rho, real* ux,
                              real* uy, real* uz, real tau, real gx, real gy, real gz, int count);
void oracle_d3q19_collide_trt(const real* f_in, real* f_out, const uint8_t* cell_type, real* rho, real* ux,
                              real* uy, real* uz, real tau_s, real tau_a, real gx, real gy, real gz,
                              int count);
void oracle_d3q19_collide_mrt(const real* f_in, real* f_out, const uint8_t* cell_type, real* rho, real* ux,
                              real* uy, real* uz, real tau_shear, real tau_bulk, real gx, real gy, real gz,
                              int count);
void oracle_d3q19_stream_periodic(const real* f_src, real* f_dst, int nx, int ny, int nz);
void oracle_d3q19_stream_with_boundaries(const real* f_src, real* f_dst, const uint8_t* cell_type, int nx,
                                         int ny, int nz, real uw_x, real uw_y, real uw_z);
void oracle_d3q19_compute_smagorinsky_tau(const real* f, const real* rho, const real* ux, const real* uy,
                                          const real* uz, real* tau_eff, real tau_0, real smagorinsky_constant,
                                          int count);
void oracle_d3q19_apply_curved_bfl(real* f_dst, const real* f_src, const int* fluid_indices, const int* q_dirs,
                                   const real* delta_fractions, const real* uw_x, const real* uw_y,
                                   const real* uw_z, int nx, int ny, int nz, int num_boundary_links);
void oracle_d3q19_compute_momentum_exchange_forces(const real* f, const uint8_t* cell_type, int nx, int ny,
                                                   int nz, real center_x, real center_y, real center_z,
                                                   real max_radius, real* fx_out, real* fy_out,
                                                   real* fz_out);

/* ---- D3Q27 kernels (src/ispc/d3q27_*.ispc) ---- */
void oracle_d3q27_collide(const real* f_in, real* f_out, const uint8_t* cell_type, real* rho, real* ux,
                          real* uy, real* uz, real tau_s, real tau_a, real gx, real gy, real gz, int count);
void oracle_d3q27_stream_periodic(const real* f_src, real* f_dst, int nx, int ny, int nz);
void oracle_d3q27_compute_macroscopic_moments(const real* f, real* rho, real* ux, real* uy, real* uz,
                                              int count);

/* ---- EXTENSIONS (no reference implementation) ---- */
/* Per-node relaxation (LES-coupled) collisions: tau_node[i] replaces the scalar
 * shear tau; TRT derives tau_a per node with Lambda = 1/4 as the grid does. */
void oracle_collide_tau_node(int lattice, int collision, const real* f_in, real* f_out,
                             const uint8_t* cell_type, const real* tau_node, real gx, real gy, real gz,
                             int count);
/* Smagorinsky tau for any lattice from the populations alone (rho, u taken
 * from f with the Guo shift).  lattice: 9, 19 or 27. */
void oracle_smagorinsky_tau_from_f(int lattice, const real* f, real* tau_eff, real tau_0, real cs, real gx,
                                   real gy, real gz, int count);
/* General 3D pull streaming for D3Q19/D3Q27 with per-axis periodicity and
 * halfway bounce-back on solid / moving-wall neighbours and non-periodic edges. */
void oracle_stream3d_general(int lattice, const real* f_src, real* f_dst, const uint8_t* cell_type, int nx,
                             int ny, int nz, int per_y, int per_z, real uw_x, real uw_y, real uw_z);
void oracle_d2q9_stream_general(const real* f_src, real* f_dst, const uint8_t* cell_type, int nx, int ny,
                                int per_y, real uw_x, real uw_y);

/* ---- grid-level step sequences (src/grid/lattice_grid_{2d,3d}.cpp) ---- */
typedef struct oracle_grid oracle_grid;
oracle_grid* oracle_grid_create(int dim, int lattice, int nx, int ny, int nz, real tau);
void oracle_grid_destroy(oracle_grid* g);
void oracle_grid_initialize(oracle_grid* g, real rho, real ux, real uy, real uz);
/* initialise from per-node rho, u arrays (initialize_custom with a table) */
void oracle_grid_initialize_fields(oracle_grid* g, const real* rho, const real* ux, const real* uy,
                                   const real* uz);
void oracle_grid_add_noise(oracle_grid* g, real amplitude, uint64_t seed);
void oracle_grid_set_cell_type(oracle_grid* g, int x, int y, int z, int type);
void oracle_grid_set_periodic(oracle_grid* g, int periodic);
void oracle_grid_set_periodic_axes(oracle_grid* g, int per_y, int per_z);
void oracle_grid_set_solid_box(oracle_grid* g, int x0, int y0, int z0, int x1, int y1, int z1);
void oracle_grid_set_solid_sphere(oracle_grid* g, real cx, real cy, real cz, real r);
void oracle_grid_set_solid_naca(oracle_grid* g, int digits, real chord, real aoa_deg, real lead_x, real lead_y,
                                int z0, int z1);
/* EXTENSION: D2Q5 temperature coupled to the D2Q9 flow (Boussinesq). */
void oracle_grid_enable_thermal(oracle_grid* g, real tau_g, real gbeta, real t_ref, real t_init);
void oracle_grid_set_temperature_field(oracle_grid* g, const real* t);
void oracle_grid_set_wall_temperature(oracle_grid* g, int x0, int y0, int x1, int y1, real t_wall);
const real* oracle_grid_temperature(const oracle_grid* g);
void oracle_grid_set_solid_ellipsoid(oracle_grid* g, real cx, real cy, real cz, real rx, real ry, real rz,
                                     const real* axes);
void oracle_grid_set_solid_circle(oracle_grid* g, real cx, real cy, real r);
void oracle_grid_set_solid_cylinder_z(oracle_grid* g, real cx, real cy, real r);
void oracle_grid_set_moving_wall_top(oracle_grid* g, real uwx, real uwz);
void oracle_grid_set_inflow_west(oracle_grid* g, real ux, real uy);
void oracle_grid_set_inflow_parabolic(oracle_grid* g, real u_max);
void oracle_grid_set_inflow_turbulence(oracle_grid* g, real amplitude, uint64_t seed);
void oracle_grid_set_outflow_east(oracle_grid* g, real rho_out);
void oracle_grid_set_convective_outflow_east(oracle_grid* g, real u_conv);
void oracle_grid_set_les_coupled(oracle_grid* g, int coupled);
int oracle_grid_build_curved_sphere_links(oracle_grid* g, real cx, real cy, real cz, real r);
int oracle_grid_num_links(const oracle_grid* g);
void oracle_grid_get_links(const oracle_grid* g, int* idx, int* q, real* delta);
void oracle_grid_step(oracle_grid* g, int collision, real gx, real gy, real gz, int use_les, real smag);
void oracle_grid_run(oracle_grid* g, int steps, int collision, real gx, real gy, real gz, int use_les,
                     real smag);
const real* oracle_grid_f_src(const oracle_grid* g);
const real* oracle_grid_f_dst(const oracle_grid* g);
const real* oracle_grid_rho(const oracle_grid* g);
const real* oracle_grid_ux(const oracle_grid* g);
const real* oracle_grid_uy(const oracle_grid* g);
const real* oracle_grid_uz(const oracle_grid* g);
const real* oracle_grid_speed(const oracle_grid* g);
const uint8_t* oracle_grid_cell_type(const oracle_grid* g);
double oracle_grid_total_mass(const oracle_grid* g);
double oracle_grid_total_kinetic_energy(const oracle_grid* g);
double oracle_grid_max_speed(const oracle_grid* g);
void oracle_grid_momentum_exchange(const oracle_grid* g, real cx, real cy, real cz, real r_max, double* out3);

/* obj_loader.cpp:209-306: triangle (v0, v1, v2 as 9 coordinates) against
 * the box of centre c and half extents h, separating-axis test. */
int oracle_tri_box_overlap(const real* tri, const real* c, const real* h);
/* obj_loader.cpp:308-400: voxelize ntri triangles into flags[(z*ny+y)*nx+x]
 * (surface, then the enclosed cells with fill_interior); returns the number
 * of Solid writes. */
long long oracle_voxelize_mesh(uint8_t* flags, int nx, int ny, int nz, const real* tris, long long ntri,
                               int fill_interior);

/* thermal_kernels.ispc:4-81 (the reference's D2Q9 temperature solver) */
void oracle_thermal_collide_d2q9(const real* g_in, real* g_out, const uint8_t* cell_type, const real* ux,
                                 const real* uy, real* temperature, real tau_T, int count);
void oracle_thermal_compute_buoyancy_force_2d(const real* temperature, real* gy_body, real t_ref, real beta,
                                              real g_gravity, int count);
void oracle_thermal_apply_isothermal_wall(real* g, int nx, int ny, int y_wall, real t_wall);
/* ibm_kernels.ispc:4-96 */
void oracle_ibm_interpolate_velocity_2d(const real* ux, const real* uy, const real* marker_x, const real* marker_y,
                                        real* marker_ux, real* marker_uy, int nx, int ny, int num_markers);
void oracle_ibm_spread_force_2d(const real* marker_x, const real* marker_y, const real* marker_fx,
                                const real* marker_fy, const real* marker_ds, real* euler_gx, real* euler_gy, int nx,
                                int ny, int num_markers);

/* analytics_kernels.ispc:4-110 */
void oracle_compute_vorticity_2d(const real* ux, const real* uy, real* w, int nx, int ny);
void oracle_compute_vorticity_and_q_criterion_3d(const real* ux, const real* uy, const real* uz, real* ox, real* oy,
                                                 real* oz, real* qc, int nx, int ny, int nz);

#ifdef __cplusplus
}
#endif
#endif
