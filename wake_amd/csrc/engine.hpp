// engine.hpp -- host/device contract of the device-resident LBM step.
//
// A domain's populations live in HBM as Q planes of `stride` elements
// (plane q of node n at q * stride + n, n = (z * ny + y) * nx + x, the
// reference's SoA index with a padded plane stride).  Two buffers ping-pong.
// The fused kernel pulls the post-collision populations A_{k-1} of the
// previous launch, forms the post-stream state S_k in registers (streaming,
// halfway bounce-back, moving walls, BFL curved links, open boundaries), and
// collides it into A_k.  S_k itself -- the reference's f_src after step k --
// is only materialised on demand by the read-back kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace wake_hip {

enum Storage : int { kStoreF32 = 0, kStoreF16 = 1, kStoreBF16 = 2 };

// Per-node class byte, derived on the device from the reference's cell
// types and the boundary configuration.
enum ClassBits : uint8_t {
  kClsSolid = 1,   // collision copies the node (reference: CELL_SOLID)
  kClsFrozen = 2,  // the pull copies the node's own populations (solid, wall streaming)
  kClsSlow = 4,    // some pull source is out of range, solid or a moving wall
  kClsBfl = 8,     // the node owns BFL curved-wall links (aux[n] indexes bfl[])
  kClsInlet = 16,  // Zou-He velocity inlet node
  kClsOutlet = 32  // Zou-He pressure or convective outlet node
};

// Per-node bounce mask: bit q (q >= 1) set when the pull source along c_q
// is out of range (non-periodic axis), Solid or a MovingWall, so that bit
// opp(q) of an in-range neighbour x + c_q marks a wall link along c_q;
// kBounceMovingWall set when one of them is a MovingWall; kBounceFluid set
// when the node itself is Fluid.  All zero in periodic-stream mode.
enum : uint32_t {
  kBounceFluid = 1u,
  kBounceMovingWall = 1u << 31
};

// Curved-wall (BFL) links of one fluid node: bit q of mask marks a link from
// the node along c_q into the wall, with wall fraction delta[q] and the
// wall-velocity projection c_q . u_w in ciuw[q].
struct BflNode {
  uint32_t mask;
  float delta[27];
  float ciuw[27];
};

struct StepArgs {
  int nx, ny, nz;
  long long stride;          // population plane stride in elements (>= nx*ny*nz)
  int periodic_mode;         // reference stream_periodic: the pull ignores cell types
  int per_y, per_z;          // boundary mode: wrap y / z instead of reflecting
  int bfl_per_y, bfl_per_z;  // wrap of the BFL upstream read x_f - c_q
  float tau, gx, gy, gz;     // collision: shear tau and Guo body force
  float cs;                  // Smagorinsky constant (LES-coupled collision)
  float uw[3];               // moving-wall velocity (the pull adds 2 w_q rho0 3 c_q.u_w)
  int inflow;                // 1: Zou-He velocity inlet at x = 0
  int inflow_parabolic;      // profile 4 umax s (H - s) / H^2, s = y - 1/2, H = ny - 2
  float uin_x, uin_y, umax, turb_amp;
  unsigned long long turb_seed;
  unsigned long long s_step;  // reference step index k of the S_k being formed
  int outflow;                // 1: Zou-He pressure (2D), 2: convective
  float rho_out, u_conv;
  float conv_alpha, conv_beta;  // convective outlet: u/(1+u) and 1/(1+u), formed on the host
  int by0, by1, bz0, bz1;     // rows an open boundary acts on
  int commit_prev;            // convective: store S_k(outlet) as next step's f_prev
  int global_addr;            // host side: keep 16-bit storage on global pointers (no descriptor)
  // D2Q5 temperature coupled by a Boussinesq force (2D only)
  float tau_g, gbeta, t_ref;  // temperature relaxation, a_y = gbeta (T - t_ref)
  const float* t_wall;        // per node: wall temperature of a solid node, NaN = adiabatic
  float* f_prev;              // convective outlet store: ny*nz records of prev_stride(Q) floats
  const uint8_t* flags;       // reference cell types
  const uint8_t* cls;         // ClassBits
  const int* aux;             // per node: index into bfl[] (valid where kClsBfl)
  const uint32_t* bmask;      // per node: bounce mask (valid where kClsSlow); null: from the cell types
  const BflNode* bfl;
  // read-back outputs
  float* rho;
  float* ux;
  float* uy;
  float* uz;
  float* speed;
  float* s_out;               // optional: S_k planes, fp32, stride nx*ny*nz
  float* temp;                // read-back temperature (thermal domains)
  // Momentum exchange on the S_k a launch forms (mem_rows null: none).
  // mem_rows[z * ny + y] is the first partial slot of row (y, z) or -1 for a
  // row without links; wave w of a linked row (nodes x = 64 w .. 64 w + 63)
  // writes its FP64 sum (3 doubles) to mem_part[3 (slot + w) ..].  Links run
  // from Fluid nodes to in-range Solid / MovingWall neighbours whose midpoint
  // lies within mem_rmax of mem_c (mem_rmax <= 0: every link).
  const int* mem_rows;
  double* mem_part;
  float mem_c[3];
  float mem_rmax;
  // Divergence flag (null: off): a step launch sets *nonfinite = 1 when a
  // node's density (or a thermal node's temperature) comes out of the
  // collision non-finite.  Sticky until the state is re-initialised.
  int* nonfinite;
};

// Partial slots of one force sample: rows with links x waves per row.
inline int mem_waves_per_row(int nx) { return (nx + 63) / 64; }

// Launchers (step_kernels.hip).  `lattice` is 9, 19 or 27.
hipError_t launch_build_class(int lattice, const StepArgs& a, uint8_t* cls, uint32_t* bmask, hipStream_t s);
// Steps per launch of the fully periodic D2Q9 path (launch_step_tb).
inline constexpr int kTbSteps = 8;
// kTbSteps pull + collide steps of a fully periodic D2Q9 domain whose class
// bytes are all zero (no solid, inlet or outlet node), A_{k-1} in `in` to
// A_{k-1+kTbSteps} in `out`.
hipError_t launch_step_tb(int lattice, int storage, int collision, bool les, const StepArgs& a, const void* in,
                          void* out, hipStream_t s);
hipError_t launch_mark_bfl(const int* nodes, int count, uint8_t* cls, int* aux, hipStream_t s);
// pull=false: collide-only from the materialised S state.  block_x is the
// work-group width along x (64, 128 or 256).
hipError_t launch_step(int lattice, int storage, int collision, bool les, bool pull, const StepArgs& a,
                       const void* a_in, void* a_out, int block_x, hipStream_t s);
// Thermal D2Q9 + D2Q5 domains: the same step with the temperature
// populations g (FP32, 5 planes of a.stride) pulled, coupled and collided
// beside f.  pull=false collides the materialised S state of both.
hipError_t launch_step_thermal(int storage, int collision, bool pull, const StepArgs& a, const void* a_in,
                               void* a_out, const float* g_in, float* g_out, hipStream_t s);
// S_k of the temperature populations into g_s (the flow's S is materialised
// by launch_materialize).
hipError_t launch_materialize_thermal(const StepArgs& a, const float* g_in, float* g_s, hipStream_t s);
// rho, u (with the node's buoyancy in the Guo shift), |u| and T of S_k.
hipError_t launch_readback_thermal(int storage, bool from_s, const StepArgs& a, const void* buf,
                                   const float* gbuf, hipStream_t s);
// Materialise S_k (commit=true updates the convective outlet store) into
// s_state (storage format, stride a.stride), or fields/planes for read-back.
hipError_t launch_materialize(int lattice, int storage, const StepArgs& a, const void* a_in, void* s_state,
                              hipStream_t s);
// Convective outlet store <- the x = nx - 1 plane of a materialised S state.
hipError_t launch_prev_from_s(int lattice, int storage, const StepArgs& a, const void* s_state, hipStream_t s);
hipError_t launch_readback(int lattice, int storage, bool from_s, const StepArgs& a, const void* buf,
                           hipStream_t s);
// Row marks of the momentum-exchange links of (a.mem_c, a.mem_rmax):
// row_has[z * ny + y] = 1 for rows holding a link (row_has cleared by the
// caller).
hipError_t launch_mem_mark_rows(int lattice, const StepArgs& a, int* row_has, hipStream_t s);
// The nodes holding links of (a.mem_c, a.mem_rmax) in the linked rows
// row_list[0 .. nrows): nodes == null: counts[r] = the count of row r; else
// those nodes, in node order, written from nodes[offsets[r]] on.
hipError_t launch_mem_row_nodes(int lattice, const StepArgs& a, const int* row_list, int nrows, int* counts,
                                const int* offsets, int* nodes, hipStream_t s);
// The per-wave partials of a.mem_rows / a.mem_part for one S_k without a
// step: over the `nrows` linked rows listed in row_list, from the
// materialised S state (from_s) or formed from the post-collision buffer.
// Same values and slots as a pull launch of k_step writes.
hipError_t launch_mem_rows(int lattice, int storage, bool from_s, const StepArgs& a, const void* buf,
                           const int* row_list, int nrows, hipStream_t s);
// Sums of force samples: for s < count, out[3 (out_first + s) ..] = sum over
// slots of part[3 (((first + s) % ring) nslots + slot) ..], slots in order.
hipError_t launch_mem_fold(const double* part, int nslots, int ring, long long first, int count, double* out,
                           long long out_first, hipStream_t s);
// Vorticity (2D: omega_z) and, in 3D, the curl and the Q-criterion of a
// velocity field, central differences with periodic wrap (analytics.hip).
// Null outputs are skipped.
hipError_t launch_vorticity_2d(const float* ux, const float* uy, float* w, int nx, int ny, hipStream_t s);
hipError_t launch_vorticity_q_3d(const float* ux, const float* uy, const float* uz, float* ox, float* oy, float* oz,
                                 float* q, int nx, int ny, int nz, hipStream_t s);
// Triangle mesh (ntri x 3 vertices x xyz, host memory) into the device
// cell types: cells whose unit box overlaps a triangle become Solid, and
// with fill_interior so do cells the outer faces cannot reach through
// non-Solid cells (geometry.hip).  any_set: some cell was set Solid.
hipError_t launch_voxelize_mesh(const float* tris, long long ntri, bool fill_interior, uint8_t* d_flags, int nx,
                                int ny, int nz, hipStream_t s, bool* any_set);
// BFL links to a sphere (centre, radius) from the device cell types, in
// z -> y -> x -> q order, into host vectors (geometry.hip).
hipError_t launch_build_sphere_links(int lattice, const uint8_t* d_flags, int nx, int ny, int nz, bool per_y,
                                     bool per_z, float cx, float cy, float cz, float r, hipStream_t s,
                                     std::vector<int>& node, std::vector<int>& qdir, std::vector<float>& delta);
// D2Q9 temperature populations of the reference's thermal solver
// (thermal_kernels.ispc:4-81) and the D2Q9 wall streaming
// (d2q9_streaming.ispc:33-95) on device pointers (ispc_compat.hip).
hipError_t launch_thermal_collide_d2q9(const float* g_in, float* g_out, const uint8_t* ct, const float* ux,
                                       const float* uy, float* temp, float tau_t, int count, hipStream_t s);
// accumulate: gy += force (the export); else gy = force (no separate clear,
// so a captured step records no memset).
hipError_t launch_thermal_buoyancy_2d(const float* temp, float* gy, float t_ref, float beta, float g_gravity,
                                      int count, bool accumulate, hipStream_t s);
hipError_t launch_thermal_isothermal_wall(float* g, int nx, int ny, int y_wall, float t_wall, hipStream_t s);
hipError_t launch_stream_walls_d2q9(const float* src, float* dst, const uint8_t* ct, int nx, int ny, float uwx,
                                    float uwy, hipStream_t s);
// FP64 sums over Fluid nodes of rho and rho |u|^2 / 2, and max |u|.
hipError_t launch_diagnostics(const uint8_t* flags, const float* rho, const float* speed, long long n,
                              double* partial, int nblocks, double* out3, hipStream_t s);
// Nodes (all cell types) whose density or speed is not finite: out[0] +=
// their count, out[1] = min(out[1], first node index).  The caller sets
// out = {0, ~0} first.
hipError_t launch_count_nonfinite(const float* rho, const float* speed, long long n, unsigned long long* out,
                                  hipStream_t s);

}  // namespace wake_hip
