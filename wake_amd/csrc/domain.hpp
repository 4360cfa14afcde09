// domain.hpp -- one device-resident lattice domain (host side).
//
// Domain owns the HBM state of one grid and drives the fused kernels.  It
// keeps the reference grid's observable semantics (lattice_grid_2d.cpp,
// lattice_grid_3d.cpp): after step k, f_src = S_k (post-stream), f_dst =
// A_{k-1} (post-collision), rho/u = moments of S_k with the Guo shift.  The
// device holds A_{k-1} and forms S_k inside the next launch; S_k, the macro
// fields and the populations are produced only when the host asks for them.
// Any change to the boundary configuration first materialises S_k with the
// configuration that produced it, so the deferred pull never mixes two
// configurations.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "engine.hpp"

namespace wake_hip {

class HipError : public std::runtime_error {
 public:
  explicit HipError(const std::string& m) : std::runtime_error("wake: " + m) {}
};

void check(hipError_t e, const char* what);

struct CurvedLinks {
  std::vector<int> fluid_index;
  std::vector<int> q_dir;
  std::vector<float> delta;
  std::vector<float> uw_x, uw_y, uw_z;
};

class Domain {
 public:
  Domain(int dim, int lattice, int nx, int ny, int nz, float tau, int storage = kStoreF32,
         long long plane_pad = -1);
  ~Domain();
  Domain(const Domain&) = delete;
  Domain& operator=(const Domain&) = delete;

  // --- state initialisation (lattice_grid_3d.cpp:66-131)
  void initialize(float rho, float ux, float uy, float uz);
  void initialize_fields(const float* rho, const float* ux, const float* uy, const float* uz);
  void add_noise(float amplitude, uint64_t seed);
  void set_populations(const float* f_ref);  // S state from reference-layout planes

  // --- cell types and boundary configuration
  void set_cell_type(int x, int y, int z, uint8_t type);
  uint8_t get_cell_type(int x, int y, int z) const;
  void set_periodic(bool p);
  bool is_periodic() const { return is_periodic_; }
  void set_periodic_axes(bool per_y, bool per_z);
  void set_solid_box(int x0, int y0, int z0, int x1, int y1, int z1);
  void set_solid_sphere(float cx, float cy, float cz, float r);
  void set_solid_ellipsoid(float cx, float cy, float cz, float rx, float ry, float rz, const float* axes);
  // NACA 4-digit airfoil (digits: e.g. 12 for 0012), leading edge at
  // (lead_x, lead_y), extruded over z0..z1 in 3D.
  void set_solid_naca(int digits, float chord, float aoa_deg, float lead_x, float lead_y, int z0, int z1);
  // Triangle mesh (ntri triangles of 3 xyz vertices) voxelized on the
  // device: surface cells, plus the enclosed cells with fill_interior.
  void voxelize_mesh(const float* tris, long long ntri, bool fill_interior);
  // 16-bit storage: address the populations through a buffer descriptor
  // when the buffer is under 4 GiB (default) or always through pointers.
  void set_buffer_addressing(bool enable) { buffer_addressing_ = enable; }
  void set_solid_circle(float cx, float cy, float r);
  void set_solid_cylinder_z(float cx, float cy, float r);
  void set_moving_wall_top(float uwx, float uwz);
  void set_inflow_west(float ux, float uy);
  void set_inflow_parabolic(float u_max);
  void set_inflow_turbulence(float amplitude, uint64_t seed);
  void set_outflow_east(float rho_out);
  void set_convective_outflow_east(float u_conv);
  void set_les_coupled(bool coupled) {
    les_coupled_ = coupled;
    ++version_;
  }
  bool les_coupled() const { return les_coupled_; }
  int build_curved_sphere_links(float cx, float cy, float cz, float r);
  void set_curved_links(const CurvedLinks& links);
  const CurvedLinks& curved_links() const { return links_; }

  // --- stepping
  void step(int collision, float gx, float gy, float gz, bool use_les, float smag);
  // run() takes kTbSteps steps per launch where tb_eligible() (fully periodic
  // D2Q9 without special nodes, no force probe); off: one launch per step.
  void set_multistep(bool on) { tb_enabled_ = on; }
  // `steps` steps; returns the device time in ms measured with events on the
  // domain's stream around the launches.
  float run(int steps, int collision, float gx, float gy, float gz, bool use_les, float smag, bool timed);
  void synchronize();
  // Average device duration (ms) of the main step kernel over the launches
  // of the last timed run() (events around each launch on the domain's
  // stream; the open-boundary plane kernel is excluded).
  float last_kernel_ms() const { return last_kernel_ms_; }

  // --- read-back (lazily synced host mirrors, reference layout)
  const float* rho();
  const float* ux();
  const float* uy();
  const float* uz();
  const float* speed();
  const float* f_src();
  const float* f_dst();
  const uint8_t* cell_type() const { return flags_.data(); }
  // Vorticity / Q-criterion of the current velocity field, on the device
  // (analytics_kernels.ispc:4-110): which = 0 omega_x, 1 omega_y, 2 omega_z
  // (the 2D vorticity), 3 Q-criterion (3D).
  const float* analytics(int which);
  // Device arrays of the current moments (0 rho, 1 ux, 2 uy, 3 uz, 4 |u|),
  // refreshed like the host mirrors but without the copy to the host, and
  // the device cell types; valid until the next step or configuration
  // change.  Work on them belongs on stream().
  const float* device_field(int which);
  const uint8_t* device_cell_types();
  void update_macroscopic();
  // FP64 sums over Fluid nodes of rho and rho |u|^2 / 2, and max |u|.
  void diagnostics(double out[3]);
  double total_mass();
  double total_kinetic_energy();
  double max_speed();
  // F on S_k by momentum exchange (FP64): 3D (Fx, Fy, Fz); 2D (Fx, Fy, Tz).
  void momentum_exchange(float cx, float cy, float cz, float r_max, double out[3]);
  // Per-step force history: the momentum exchange of every S_k from step
  // current_step() on (links within r_max of the centre, r_max <= 0: all),
  // computed inside the launch that forms S_k and folded per step in a
  // fixed order, so the history holds the same values as calling
  // momentum_exchange() after every step.  Setting it again restarts the
  // history; initialisation restarts it at step 0.  Not for temperature-
  // coupled domains.
  void set_force_probe(float cx, float cy, float cz, float r_max);
  void clear_force_probe();
  bool force_probe() const { return probe_; }
  // Samples of steps first .. current_step(), 3 doubles each.
  const std::vector<double>& force_history(uint64_t* first_step);
  // Divergence: whether a step since the state was last set produced a
  // non-finite density (the flag the step kernels raise; one 4-byte read).
  bool nonfinite_seen();
  // Nodes whose current density or speed is not finite, counted on the
  // device (no field download); *first = the first such node.
  long long count_nonfinite(long long* first);

  // --- accessors
  int dim() const { return dim_; }
  int lattice() const { return lattice_; }
  int nx() const { return nx_; }
  int ny() const { return ny_; }
  int nz() const { return nz_; }
  size_t total_nodes() const { return N_; }
  long long plane_stride() const { return stride_; }
  int storage() const { return storage_; }
  uint64_t current_step() const { return step_count_; }
  float tau() const { return tau_; }
  void set_tau(float t) {
    tau_ = t;
    ++version_;
  }
  // Changes with every change of state or configuration made through this
  // class (a batch re-reads a domain whose version moved).
  uint64_t version() const { return version_; }
  hipStream_t stream() const { return stream_; }
  void set_block_x(int bx) { block_x_ = bx; }
  size_t device_bytes() const { return device_bytes_; }
  // device pointer of the population buffer holding the current state
  const void* state_buffer() const;
  bool state_is_post_stream() const { return s_buf_ >= 0; }

  // --- D2Q5 temperature coupled by a Boussinesq force (2D only)
  // Temperature populations at rest equilibrium w_q t_init everywhere, wall
  // temperatures cleared (adiabatic); a_y = gbeta (T - t_ref) on the flow.
  void enable_thermal(float tau_g, float gbeta, float t_ref, float t_init);
  bool thermal() const { return thermal_; }
  // Temperature field (reference node order) as rest equilibrium.
  void set_temperature_field(const float* t);
  // Wall temperature of the solid nodes in the box (inclusive); NaN makes
  // them adiabatic again.
  void set_wall_temperature(int x0, int y0, int x1, int y1, float t_wall);
  const float* temperature();
  const float* thermal_buffer(int k) const { return gbuf_[k]; }

  // --- batched stepping (batch_host.cpp)
  // Arguments of the next step for a batch launch, after the configuration
  // has reached the device.  parity: buffer holding the current state;
  // s_state: that state is a materialised S (the next step collides only).
  void batch_prepare(float gx, float gy, float gz, float smag, StepArgs& a, void* buf[2], float* gbuf[2],
                     int& parity, bool& s_state);
  // A batch took k >= 1 steps; the state is now post-collision in buffer
  // `parity_after`.
  void batch_advance(uint64_t k, int parity_after, float gx, float gy, float gz);
  // The nodes holding momentum-exchange links around (cx, cy, cz) within
  // r_max, in node order: a device list the domain owns (valid until the
  // next call or a geometry change), *count entries.
  const int* batch_mem_nodes(float cx, float cy, float cz, float r_max, int* count);
  // Device bytes a domain of this shape allocates.
  static size_t bytes_needed(int dim, int lattice, int nx, int ny, int nz, int storage, long long plane_pad = -1,
                             bool thermal = false);

 private:
  StepArgs args(bool commit) const;
  void before_config_change();
  void materialize();
  void sync_device_config();
  void ensure_fields();
  void upload_state_from_host(const std::vector<float>& f_ref);
  void invalidate_mirrors() {
    fields_valid_ = dfields_valid_ = fsrc_valid_ = fdst_valid_ = an_valid_ = false;
    sout_is_fsrc_ = false;
  }
  void ensure_device_fields();
  void* alloc(size_t bytes);
  friend const float* device_copy_of_host_populations(const float* host, size_t count);
  void allocate_and_initialize();  // constructor body after validation
  void release() noexcept;         // wait for the stream, free everything
  void release(void* p);
  void upload_prev(const std::function<float(size_t, int)>& value);
  // Rows holding the momentum-exchange links of one (centre, r_max): per
  // (y, z) row its first partial slot or -1, and the list of linked rows.
  struct MemRows {
    int* d_rows = nullptr;
    int* d_list = nullptr;
    int nrows = 0;
    int nslots = 0;
  };
  MemRows build_mem_rows(float cx, float cy, float cz, float r_max);
  void free_mem_rows(MemRows& r);
  const MemRows& cached_mem_rows(float cx, float cy, float cz, float r_max);
  StepArgs with_mem(StepArgs a, const MemRows& r, double* part, const float c[3], float r_max) const;
  void probe_sync();                // probe rows current for the geometry
  void probe_fold(uint64_t upto);   // fold the samples of steps [fh_folded_, upto)
  void probe_sample_current();      // sample S_{step_count_} if no launch has
  double* probe_slot(uint64_t k);   // ring entry of step k (folding first if it is taken)
  void probe_restart();

  int dim_, lattice_, Q_, nx_, ny_, nz_;
  size_t N_;
  long long stride_;
  int storage_;
  size_t elem_;
  float tau_;
  uint64_t step_count_ = 0;
  uint64_t version_ = 0;
  int block_x_ = 0;
  bool time_kernels_ = false;
  bool buffer_addressing_ = true;
  size_t kev_used_ = 0;
  uint64_t timed_launches_ = 0;            // pull launches in the current timed run
  static constexpr int kTimingStride = 16;  // launches per timed sample
  std::vector<hipEvent_t> kev_;
  hipEvent_t kev_marker_ = nullptr;  // recorded just before a sample's start event
  float last_kernel_ms_ = 0.0f;

  // reference grid configuration
  bool is_periodic_;
  bool ext_axes_ = false, per_y_ = false, per_z_ = false;
  bool has_moving_wall_ = false;
  float uwx_ = 0.0f, uwz_ = 0.0f;
  bool has_inflow_ = false, inflow_parabolic_ = false;
  float uin_x_ = 0.0f, uin_y_ = 0.0f, umax_ = 0.0f, turb_amp_ = 0.0f;
  uint64_t turb_seed_ = 0;
  int outflow_ = 0;
  float rho_out_ = 1.0f, u_conv_ = 0.0f;
  bool les_coupled_ = false;
  float g_last_[3] = {0.0f, 0.0f, 0.0f};
  std::vector<uint8_t> flags_;
  CurvedLinks links_;

  // device state
  hipStream_t stream_ = nullptr;
  void* buf_[2] = {nullptr, nullptr};
  int s_buf_ = -1;  // buffer holding the materialised S_k, or -1
  int a_buf_ = -1;  // buffer holding A_{k-1}, or -1
  uint8_t* d_flags_ = nullptr;
  uint8_t* d_cls_ = nullptr;
  uint32_t* d_bmask_ = nullptr;  // per node: bounce mask of the kClsSlow nodes
  bool tb_eligible();
  void step_tb(int collision, float gx, float gy, bool use_les, float smag);
  bool all_fluid_ = false;       // every cell type Fluid (set with the class bytes)
  bool tb_enabled_ = true;       // run() may take kTbSteps steps per launch
  int* d_aux_ = nullptr;
  BflNode* d_bfl_ = nullptr;
  int* d_bfl_nodes_ = nullptr;
  int n_bfl_nodes_ = 0;
  float* d_prev_ = nullptr;
  float *d_rho_ = nullptr, *d_ux_ = nullptr, *d_uy_ = nullptr, *d_uz_ = nullptr, *d_speed_ = nullptr;
  float* d_sout_ = nullptr;
  double* d_partial_ = nullptr;
  double* d_out_ = nullptr;
  int* d_nonfinite_ = nullptr;            // divergence flag of the step kernels
  unsigned long long* d_nfcount_ = nullptr;  // count_nonfinite(): count, first node
  void clear_nonfinite();
  std::vector<void*> allocations_;
  size_t device_bytes_ = 0;
  bool flags_dirty_ = true, cls_dirty_ = true, bfl_dirty_ = true;
  uint64_t flags_version_ = 0;
  uint64_t cls_epoch_ = 0;  // counts rebuilds of the class bytes (geometry or boundary changes)
  std::map<std::tuple<float, float, float, float, uint64_t>, MemRows> mem_rows_cache_;
  MemRows batch_rows_;  // the rows of batch_mem_nodes(), with their key
  int* d_batch_nodes_ = nullptr;
  int n_batch_nodes_ = 0;
  std::tuple<float, float, float, float, uint64_t> batch_rows_key_{0.0f, 0.0f, 0.0f, 0.0f, ~0ull};
  double* d_mempart_ = nullptr;  // partials of one momentum_exchange() call
  size_t mempart_cap_ = 0;       // in doubles

  // force probe: ring of per-step partials, folded into a device history
  static constexpr int kForceRing = 64;           // steps between folds at most
  static constexpr long long kForceHistCap = 1 << 16;  // device history rows before a download
  bool probe_ = false;
  float probe_c_[3] = {0.0f, 0.0f, 0.0f};
  float probe_rmax_ = 0.0f;
  MemRows probe_rows_;
  uint64_t probe_epoch_ = ~0ull;
  double* d_ring_ = nullptr;
  size_t ring_cap_ = 0;  // in doubles
  double* d_fhist_ = nullptr;
  uint64_t fh_first_ = 0;   // step of the first sample
  uint64_t fh_next_ = 0;    // first step whose sample is not written yet
  uint64_t fh_folded_ = 0;  // samples before this step are folded into d_fhist_
  uint64_t fh_dbase_ = 0;   // step of d_fhist_ row 0
  uint64_t fh_host_ = 0;    // samples before this step are in h_fhist_
  std::vector<double> h_fhist_;

  // temperature (D2Q5), ping-pong with buf_
  bool thermal_ = false;
  float tau_g_ = 0.8f, gbeta_ = 0.0f, t_ref_ = 0.0f;
  float* gbuf_[2] = {nullptr, nullptr};
  float* d_twall_ = nullptr;
  float* d_temp_ = nullptr;
  // analytics outputs (allocated on first use)
  float* d_an_[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<float> h_an_[4];
  bool an_valid_ = false;
  std::vector<float> h_twall_, h_temp_;
  bool twall_dirty_ = false;

  // host mirrors
  std::vector<float> h_rho_, h_ux_, h_uy_, h_uz_, h_speed_, h_fsrc_, h_fdst_;
  std::vector<float> h_prev_;  // staging of the outlet store (kept until the upload completes)
  bool fields_valid_ = false, fsrc_valid_ = false, fdst_valid_ = false;
  bool sout_is_fsrc_ = false;  // d_sout_ holds the bytes of the current f_src() mirror
  bool dfields_valid_ = false;  // device moment arrays current (implied by fields_valid_)
};

// splitmix64 finaliser (same constants as the device inlet turbulence).
inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// The domain's current fields as a VTK XML ImageData file (export.cpp):
// Velocity, Speed, Density, the optional vorticity / Q-criterion arrays
// (vort2d: one value per node; qcrit: one; vort3d: three, interleaved) and
// CellType; raw appended binary or ASCII.  False if the file cannot be
// written.
bool write_vti(Domain& d, const std::string& path, bool binary, const float* vort2d, const float* qcrit,
               const float* vort3d);

// The device buffer that holds the same bytes as `host`, when `host` is the
// current f_src() mirror of a live Domain with `count` floats, else null.
// Lets the momentum-exchange export reduce a grid's own populations (as the
// reference's scenarios pass grid.f_src_data()) without uploading them again.
const float* device_copy_of_host_populations(const float* host, size_t count);

}  // namespace wake_hip
