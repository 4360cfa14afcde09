// domain.cpp -- host side of a device-resident lattice domain.
#include "domain.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>

#include "lattice.hpp"

namespace wake_hip {

// A failed runtime call also stays the thread's "last error" until read,
// and kernel launches are checked with hipGetLastError(); reset it here so
// a handled failure (e.g. an allocation refused for lack of HBM) is not
// reported again by the next launch.
void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw HipError(std::string(what) + ": " + hipGetErrorString(e));
  }
}

namespace {

struct LatTab {
  int Q;
  const int *cx, *cy, *cz, *opp;
  const float* w;
};

template <class L>
LatTab tab_of() {
  return {L::Q, L::cx, L::cy, L::cz, L::opp, L::w};
}

LatTab tab(int lattice) {
  if (lattice == 9) return tab_of<D2Q9>();
  if (lattice == 27) return tab_of<D3Q27>();
  return tab_of<D3Q19>();
}

// Round a float to the population storage format (bits of the element).
uint16_t to_half_bits(float v) {
  _Float16 h = static_cast<_Float16>(v);
  uint16_t b;
  std::memcpy(&b, &h, 2);
  return b;
}
uint16_t to_bf16_bits(float v) {
  uint32_t b;
  std::memcpy(&b, &v, 4);
  return static_cast<uint16_t>((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
float from_half_bits(uint16_t b) {
  _Float16 h;
  std::memcpy(&h, &b, 2);
  return static_cast<float>(h);
}
float from_bf16_bits(uint16_t b) {
  uint32_t u = static_cast<uint32_t>(b) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

size_t elem_size(int storage) { return storage == kStoreF32 ? 4 : 2; }

}  // namespace

// Live domains, for device_copy_of_host_populations().
namespace {
std::mutex g_live_mutex;
std::vector<const Domain*> g_live;
}  // namespace

const float* device_copy_of_host_populations(const float* host, size_t count) {
  if (!host) return nullptr;
  std::lock_guard<std::mutex> lock(g_live_mutex);
  for (const Domain* d : g_live)
    if (d->fsrc_valid_ && d->sout_is_fsrc_ && d->d_sout_ && host == d->h_fsrc_.data() &&
        count == static_cast<size_t>(d->Q_) * d->N_)
      return d->d_sout_;
  return nullptr;
}

Domain::Domain(int dim, int lattice, int nx, int ny, int nz, float tau, int storage, long long plane_pad)
    : dim_(dim), lattice_(lattice), nx_(nx), ny_(ny), nz_(dim == 2 ? 1 : nz), storage_(storage), tau_(tau) {
  if (dim != 2 && dim != 3) throw HipError("dimension must be 2 or 3");
  // 0 picks the default lattice of the dimension (D2Q9, D3Q19)
  if (dim == 2 && lattice != 0 && lattice != 9) throw HipError("a 2D domain uses D2Q9 (lattice 9)");
  if (dim == 3 && lattice != 0 && lattice != 19 && lattice != 27) throw HipError("a 3D domain uses lattice 19 or 27");
  if (dim == 2) lattice_ = 9;
  else if (lattice == 0) lattice_ = 19;
  if (storage < kStoreF32 || storage > kStoreBF16) throw HipError("unknown population storage format");
  if (nx_ < 2 || ny_ < 1 || nz_ < 1) throw HipError("grid dimensions must be positive (nx >= 2)");
  Q_ = lattice_;
  N_ = static_cast<size_t>(nx_) * ny_ * nz_;
  if (N_ >= (1ull << 31)) throw HipError("grid larger than 2^31 nodes");
  elem_ = elem_size(storage);
  const long long base = static_cast<long long>((N_ + 255) / 256 * 256);
  stride_ = base + (plane_pad >= 0 ? plane_pad : (N_ >= (1u << 20) ? 1024 : 0));
  is_periodic_ = (dim == 3);
  try {
    allocate_and_initialize();
  } catch (...) {
    release();  // the destructor does not run for a constructor that throws
    throw;
  }
  std::lock_guard<std::mutex> lock(g_live_mutex);
  g_live.push_back(this);
}

void Domain::allocate_and_initialize() {
  check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  for (int b = 0; b < 2; ++b) buf_[b] = alloc(static_cast<size_t>(Q_) * stride_ * elem_);
  d_flags_ = static_cast<uint8_t*>(alloc(N_));
  d_cls_ = static_cast<uint8_t*>(alloc(N_));
  d_aux_ = static_cast<int*>(alloc(N_ * sizeof(int)));
  d_bmask_ = static_cast<uint32_t*>(alloc(N_ * sizeof(uint32_t)));
  d_prev_ = static_cast<float*>(alloc(static_cast<size_t>(prev_stride(Q_)) * ny_ * nz_ * sizeof(float)));
  d_rho_ = static_cast<float*>(alloc(N_ * 4));
  d_ux_ = static_cast<float*>(alloc(N_ * 4));
  d_uy_ = static_cast<float*>(alloc(N_ * 4));
  d_uz_ = static_cast<float*>(alloc(N_ * 4));
  d_speed_ = static_cast<float*>(alloc(N_ * 4));
  d_partial_ = static_cast<double*>(alloc(3 * 1024 * sizeof(double)));
  d_out_ = static_cast<double*>(alloc(8 * sizeof(double)));
  d_nonfinite_ = static_cast<int*>(alloc(sizeof(int)));
  d_nfcount_ = static_cast<unsigned long long*>(alloc(2 * sizeof(unsigned long long)));
  flags_.assign(N_, kFluid);
  h_rho_.resize(N_);
  h_ux_.resize(N_);
  h_uy_.resize(N_);
  h_uz_.resize(N_);
  h_speed_.resize(N_);
  initialize(1.0f, 0.0f, 0.0f, 0.0f);
}

Domain::~Domain() { release(); }

void Domain::release() noexcept {
  {
    std::lock_guard<std::mutex> lock(g_live_mutex);
    g_live.erase(std::remove(g_live.begin(), g_live.end(), this), g_live.end());
  }
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (hipEvent_t e : kev_) (void)hipEventDestroy(e);
  kev_.clear();
  if (kev_marker_) (void)hipEventDestroy(kev_marker_);
  kev_marker_ = nullptr;
  for (void* p : allocations_) (void)hipFree(p);
  allocations_.clear();
  if (stream_) (void)hipStreamDestroy(stream_);
  stream_ = nullptr;
  (void)hipGetLastError();  // nothing of the teardown reaches the next launch check
}

void* Domain::alloc(size_t bytes) {
  void* p = nullptr;
  check(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc");
  allocations_.push_back(p);
  device_bytes_ += bytes;
  return p;
}

// The convective outlet store: one record of prev_stride(Q) floats per
// (y, z) row, value(row, q) for q < Q, zero padding.
void Domain::upload_prev(const std::function<float(size_t, int)>& value) {
  const int QP = prev_stride(Q_);
  const size_t rows = static_cast<size_t>(ny_) * nz_;
  h_prev_.assign(rows * QP, 0.0f);
  for (size_t r = 0; r < rows; ++r)
    for (int q = 0; q < Q_; ++q) h_prev_[r * QP + q] = value(r, q);
  check(hipMemcpyAsync(d_prev_, h_prev_.data(), h_prev_.size() * sizeof(float), hipMemcpyHostToDevice, stream_),
        "upload outlet store");
  check(hipStreamSynchronize(stream_), "upload outlet store");
}

void Domain::release(void* p) {
  if (!p) return;
  (void)hipFree(p);
  allocations_.erase(std::remove(allocations_.begin(), allocations_.end(), p), allocations_.end());
}

const void* Domain::state_buffer() const { return buf_[s_buf_ >= 0 ? s_buf_ : a_buf_]; }

// ------------------------------------------------------------ initialisation
// lattice_grid_3d.cpp:66-93: feq of uniform fields in both buffers, macro
// fields set, every cell type reset to Fluid.
void Domain::initialize(float rho, float ux, float uy, float uz) {
  ++version_;
  const LatTab L = tab(lattice_);
  if (dim_ == 2) uz = 0.0f;
  step_count_ = 0;
  const float u_sq = dim_ == 2 ? ux * ux + uy * uy : ux * ux + uy * uy + uz * uz;
  std::vector<float> fe_q(Q_);
  for (int q = 0; q < Q_; ++q) {
    const float ci_u = dim_ == 2 ? static_cast<float>(L.cx[q]) * ux + static_cast<float>(L.cy[q]) * uy
                                 : static_cast<float>(L.cx[q]) * ux + static_cast<float>(L.cy[q]) * uy +
                                       static_cast<float>(L.cz[q]) * uz;
    const float fe = L.w[q] * rho * (1.0f + 3.0f * ci_u + 4.5f * ci_u * ci_u - 1.5f * u_sq);
    void* plane = static_cast<char*>(buf_[0]) + static_cast<size_t>(q) * stride_ * elem_;
    if (storage_ == kStoreF32) {
      uint32_t bits;
      std::memcpy(&bits, &fe, 4);
      check(hipMemsetD32Async(static_cast<hipDeviceptr_t>(plane), bits, stride_, stream_), "fill");
    } else {
      const uint16_t bits = storage_ == kStoreF16 ? to_half_bits(fe - L.w[q]) : to_bf16_bits(fe - L.w[q]);
      check(hipMemsetD16Async(static_cast<hipDeviceptr_t>(plane), bits, stride_, stream_), "fill");
    }
    fe_q[q] = fe;
  }
  upload_prev([&](size_t, int q) { return fe_q[q]; });
  s_buf_ = 0;
  a_buf_ = -1;
  std::fill(flags_.begin(), flags_.end(), static_cast<uint8_t>(kFluid));
  flags_dirty_ = cls_dirty_ = true;
  ++flags_version_;
  const float sp = std::sqrt(u_sq);
  std::fill(h_rho_.begin(), h_rho_.end(), rho);
  std::fill(h_ux_.begin(), h_ux_.end(), ux);
  std::fill(h_uy_.begin(), h_uy_.end(), uy);
  std::fill(h_uz_.begin(), h_uz_.end(), uz);
  std::fill(h_speed_.begin(), h_speed_.end(), sp);
  const float vals[5] = {rho, ux, uy, uz, sp};
  float* dst[5] = {d_rho_, d_ux_, d_uy_, d_uz_, d_speed_};
  for (int k = 0; k < 5; ++k) {
    uint32_t b;
    std::memcpy(&b, &vals[k], 4);
    check(hipMemsetD32Async(static_cast<hipDeviceptr_t>(dst[k]), b, N_, stream_), "fill");
  }
  fields_valid_ = true;
  fsrc_valid_ = fdst_valid_ = false;
  sout_is_fsrc_ = false;
  clear_nonfinite();
  check(hipStreamSynchronize(stream_), "initialize");
  if (probe_) probe_restart();
}

void Domain::upload_state_from_host(const std::vector<float>& f_ref) {
  ++version_;
  // f_ref: Q planes of N (reference layout) -> storage format, padded stride.
  const LatTab L = tab(lattice_);
  std::vector<char> host(static_cast<size_t>(Q_) * stride_ * elem_, 0);
  for (int q = 0; q < Q_; ++q) {
    const float* src = f_ref.data() + static_cast<size_t>(q) * N_;
    char* dst = host.data() + static_cast<size_t>(q) * stride_ * elem_;
    if (storage_ == kStoreF32) {
      std::memcpy(dst, src, N_ * 4);
    } else {
      uint16_t* d16 = reinterpret_cast<uint16_t*>(dst);
      for (size_t i = 0; i < N_; ++i)
        d16[i] = storage_ == kStoreF16 ? to_half_bits(src[i] - L.w[q]) : to_bf16_bits(src[i] - L.w[q]);
    }
  }
  check(hipMemcpyAsync(buf_[0], host.data(), host.size(), hipMemcpyHostToDevice, stream_), "upload");
  clear_nonfinite();
  // convective outlet store = S_0 on the x = nx - 1 plane
  upload_prev([&](size_t row, int q) { return f_ref[static_cast<size_t>(q) * N_ + row * nx_ + (nx_ - 1)]; });
  check(hipStreamSynchronize(stream_), "upload");
  s_buf_ = 0;
  a_buf_ = -1;
  h_fsrc_ = f_ref;
  fsrc_valid_ = (storage_ == kStoreF32);
  sout_is_fsrc_ = false;
  fdst_valid_ = false;
  if (probe_) probe_restart();  // the state the samples were taken of is gone
}

// lattice_grid_3d.cpp:95-131 with the per-node values as arrays; cell types
// are left untouched.
void Domain::initialize_fields(const float* rho, const float* ux, const float* uy, const float* uz) {
  const LatTab L = tab(lattice_);
  step_count_ = 0;
  std::vector<float> f(static_cast<size_t>(Q_) * N_);
  for (size_t i = 0; i < N_; ++i) {
    const float r = rho[i], a = ux[i], b = uy[i], c = (dim_ == 2 || !uz) ? 0.0f : uz[i];
    const float u_sq = dim_ == 2 ? a * a + b * b : a * a + b * b + c * c;
    h_rho_[i] = r;
    h_ux_[i] = a;
    h_uy_[i] = b;
    h_uz_[i] = c;
    h_speed_[i] = std::sqrt(u_sq);
    for (int q = 0; q < Q_; ++q) {
      const float ci_u = dim_ == 2 ? static_cast<float>(L.cx[q]) * a + static_cast<float>(L.cy[q]) * b
                                   : static_cast<float>(L.cx[q]) * a + static_cast<float>(L.cy[q]) * b +
                                         static_cast<float>(L.cz[q]) * c;
      f[static_cast<size_t>(q) * N_ + i] = L.w[q] * r * (1.0f + 3.0f * ci_u + 4.5f * ci_u * ci_u - 1.5f * u_sq);
    }
  }
  upload_state_from_host(f);
  float* dst[5] = {d_rho_, d_ux_, d_uy_, d_uz_, d_speed_};
  const std::vector<float>* src[5] = {&h_rho_, &h_ux_, &h_uy_, &h_uz_, &h_speed_};
  for (int k = 0; k < 5; ++k)
    check(hipMemcpyAsync(dst[k], src[k]->data(), N_ * 4, hipMemcpyHostToDevice, stream_), "upload");
  check(hipStreamSynchronize(stream_), "upload");
  fields_valid_ = true;
}

void Domain::set_populations(const float* f_ref) {
  before_config_change();
  std::vector<float> f(f_ref, f_ref + static_cast<size_t>(Q_) * N_);
  upload_state_from_host(f);
  invalidate_mirrors();
  fsrc_valid_ = (storage_ == kStoreF32);
}

// Seeded uniform noise U(-a, a) on every population, counter = node * Q + q.
void Domain::add_noise(float amplitude, uint64_t seed) {
  std::vector<float> f(f_src(), f_src() + static_cast<size_t>(Q_) * N_);
  for (size_t i = 0; i < N_; ++i)
    for (int q = 0; q < Q_; ++q) {
      const uint64_t h = splitmix64(seed ^ (static_cast<uint64_t>(i) * Q_ + q));
      const double u = static_cast<double>(h >> 40) * (1.0 / 16777216.0);
      f[static_cast<size_t>(q) * N_ + i] += amplitude * static_cast<float>(2.0 * u - 1.0);
    }
  const uint64_t keep = step_count_;
  const bool fv = fields_valid_, dfv = dfields_valid_;
  upload_state_from_host(f);
  step_count_ = keep;
  if (probe_) probe_restart();
  fields_valid_ = fv;
  dfields_valid_ = dfv;
}

// ------------------------------------------------------------- cell types
void Domain::before_config_change() {
  ++version_;
  if (s_buf_ < 0 && a_buf_ >= 0) materialize();
}

void Domain::set_cell_type(int x, int y, int z, uint8_t type) {
  if (x < 0 || x >= nx_ || y < 0 || y >= ny_ || z < 0 || z >= nz_) return;
  before_config_change();
  flags_[(static_cast<size_t>(z) * ny_ + y) * nx_ + x] = type;
  if (type != kFluid) is_periodic_ = false;
  flags_dirty_ = cls_dirty_ = true;
  ++flags_version_;
}

uint8_t Domain::get_cell_type(int x, int y, int z) const {
  if (x >= 0 && x < nx_ && y >= 0 && y < ny_ && z >= 0 && z < nz_)
    return flags_[(static_cast<size_t>(z) * ny_ + y) * nx_ + x];
  return kSolid;
}

void Domain::set_periodic(bool p) {
  before_config_change();
  is_periodic_ = p;
  cls_dirty_ = true;
}

void Domain::set_periodic_axes(bool py, bool pz) {
  before_config_change();
  ext_axes_ = true;
  per_y_ = py;
  per_z_ = pz;
  cls_dirty_ = true;
}

// lattice_grid_3d.cpp:149-158
void Domain::set_solid_box(int x0, int y0, int z0, int x1, int y1, int z1) {
  before_config_change();
  is_periodic_ = false;
  for (int z = std::max(0, z0); z <= std::min(nz_ - 1, z1); ++z)
    for (int y = std::max(0, y0); y <= std::min(ny_ - 1, y1); ++y)
      for (int x = std::max(0, x0); x <= std::min(nx_ - 1, x1); ++x) set_cell_type(x, y, z, kSolid);
}

// naca_airfoil.cpp:23-41 (digits -> camber, camber position, thickness) and
// :113-150 (contains_point), statement for statement in FP32 so the flags are
// bit-exact: the query point is turned into the unrotated chord frame and
// tested against the 4-digit thickness and camber lines.
namespace {
struct Naca {
  float camber, camber_pos, thickness, chord, aoa_deg;
  bool contains(float x, float y, float lead_x, float lead_y) const {
    const float kPi = 3.14159265358979323846f;
    const float dx = x - lead_x;
    const float dy = y - lead_y;
    const float alpha_rad = -aoa_deg * (kPi / 180.0f);
    const float cos_a = std::cos(-alpha_rad);
    const float sin_a = std::sin(-alpha_rad);
    const float x_chord = dx * cos_a - dy * sin_a;
    const float y_chord = dx * sin_a + dy * cos_a;
    if (x_chord < 0.0f || x_chord > chord) return false;
    const float xc = x_chord / chord;
    const float yt = 5.0f * thickness *
                     (0.2969f * std::sqrt(xc) - 0.1260f * xc - 0.3516f * xc * xc + 0.2843f * xc * xc * xc -
                      0.1036f * xc * xc * xc * xc) *
                     chord;
    float yc = 0.0f;
    if (camber_pos > 1e-6f) {
      if (xc < camber_pos) {
        yc = (camber / (camber_pos * camber_pos)) * (2.0f * camber_pos * xc - xc * xc) * chord;
      } else {
        const float one_minus_p = 1.0f - camber_pos;
        yc = (camber / (one_minus_p * one_minus_p)) * ((1.0f - 2.0f * camber_pos) + 2.0f * camber_pos * xc - xc * xc) *
             chord;
      }
    }
    return std::abs(y_chord - yc) <= yt;
  }
};
}  // namespace

// naca_airfoil.cpp:152-190: rasterize (2D) / voxelize_extruded over z0..z1 (3D).
void Domain::set_solid_naca(int digits, float chord, float aoa_deg, float lead_x, float lead_y, int z0, int z1) {
  if (digits < 0 || digits > 9999) throw HipError("NACA designation must be 4 digits");
  before_config_change();
  is_periodic_ = false;
  const int m_digit = digits / 1000, p_digit = (digits / 100) % 10, t_digits = digits % 100;
  Naca a{static_cast<float>(m_digit) / 100.0f, static_cast<float>(p_digit) / 10.0f,
         static_cast<float>(t_digits) / 100.0f, chord, aoa_deg};
  const float rad_bound = chord * 1.5f;
  const int min_x = std::max(0, static_cast<int>(lead_x - rad_bound));
  const int max_x = std::min(nx_ - 1, static_cast<int>(lead_x + rad_bound));
  const int min_y = std::max(0, static_cast<int>(lead_y - rad_bound));
  const int max_y = std::min(ny_ - 1, static_cast<int>(lead_y + rad_bound));
  const int za = dim_ == 2 ? 0 : std::max(0, z0), zb = dim_ == 2 ? 0 : std::min(nz_ - 1, z1);
  for (int y = min_y; y <= max_y; ++y)
    for (int x = min_x; x <= max_x; ++x)
      if (a.contains(static_cast<float>(x), static_cast<float>(y), lead_x, lead_y))
        for (int z = za; z <= zb; ++z) set_cell_type(x, y, z, kSolid);
}

// primitives.cpp:73-100: FP32 ellipsoid test, bit-exact flags.  With axes
// (rows = body axes in world coordinates) the point is first projected on
// the body axes and the box is bounded by the largest semi-axis.
void Domain::set_solid_ellipsoid(float cx, float cy, float cz, float rx, float ry, float rz, const float* axes) {
  before_config_change();
  is_periodic_ = false;
  float ex = rx, ey = ry, ez = rz;
  if (axes) ex = ey = ez = std::max(rx, std::max(ry, rz));
  const int min_x = std::max(0, static_cast<int>(cx - ex - 1.0f)), max_x = std::min(nx_ - 1, static_cast<int>(cx + ex + 1.0f));
  const int min_y = std::max(0, static_cast<int>(cy - ey - 1.0f)), max_y = std::min(ny_ - 1, static_cast<int>(cy + ey + 1.0f));
  const int min_z = std::max(0, static_cast<int>(cz - ez - 1.0f)), max_z = std::min(nz_ - 1, static_cast<int>(cz + ez + 1.0f));
  const float irx = 1.0f / (rx * rx), iry = 1.0f / (ry * ry), irz = 1.0f / (rz * rz);
  for (int z = min_z; z <= max_z; ++z)
    for (int y = min_y; y <= max_y; ++y)
      for (int x = min_x; x <= max_x; ++x) {
        float dx = static_cast<float>(x) - cx, dy = static_cast<float>(y) - cy, dz = static_cast<float>(z) - cz;
        if (axes) {
          const float b0 = axes[0] * dx + axes[1] * dy + axes[2] * dz;
          const float b1 = axes[3] * dx + axes[4] * dy + axes[5] * dz;
          const float b2 = axes[6] * dx + axes[7] * dy + axes[8] * dz;
          dx = b0;
          dy = b1;
          dz = b2;
        }
        if (dx * dx * irx + dy * dy * iry + dz * dz * irz <= 1.0f) set_cell_type(x, y, z, kSolid);
      }
}

// obj_loader.cpp:208-400 on the device.  The host mirror of the cell types
// is refreshed from the device afterwards; set_cell_type's rule applies: a
// Solid write clears the all-periodic mode.
void Domain::voxelize_mesh(const float* tris, long long ntri, bool fill_interior) {
  if (ntri < 0 || (ntri > 0 && !tris)) throw HipError("voxelize_mesh: bad triangle array");
  for (long long i = 0; i < 9 * ntri; ++i)
    if (!std::isfinite(tris[i])) throw HipError("voxelize_mesh: mesh vertices must be finite");
  before_config_change();
  if (flags_dirty_) {
    check(hipMemcpyAsync(d_flags_, flags_.data(), N_, hipMemcpyHostToDevice, stream_), "upload flags");
    flags_dirty_ = false;
  }
  bool any = false;
  check(launch_voxelize_mesh(tris, ntri, fill_interior, d_flags_, nx_, ny_, nz_, stream_, &any), "voxelize mesh");
  check(hipMemcpyAsync(flags_.data(), d_flags_, N_, hipMemcpyDeviceToHost, stream_), "download flags");
  check(hipStreamSynchronize(stream_), "voxelize mesh");
  if (any) is_periodic_ = false;
  cls_dirty_ = true;
  ++flags_version_;
}

// lattice_grid_3d.cpp:160-182 (FP32 ball test, bit-exact flags)
void Domain::set_solid_sphere(float cx, float cy, float cz, float radius) {
  before_config_change();
  is_periodic_ = false;
  const float r_sq = radius * radius;
  const int min_x = std::max(0, static_cast<int>(cx - radius - 1)), max_x = std::min(nx_ - 1, static_cast<int>(cx + radius + 1));
  const int min_y = std::max(0, static_cast<int>(cy - radius - 1)), max_y = std::min(ny_ - 1, static_cast<int>(cy + radius + 1));
  const int min_z = std::max(0, static_cast<int>(cz - radius - 1)), max_z = std::min(nz_ - 1, static_cast<int>(cz + radius + 1));
  for (int z = min_z; z <= max_z; ++z)
    for (int y = min_y; y <= max_y; ++y)
      for (int x = min_x; x <= max_x; ++x) {
        const float dx = static_cast<float>(x) - cx, dy = static_cast<float>(y) - cy, dz = static_cast<float>(z) - cz;
        if (dx * dx + dy * dy + dz * dz <= r_sq) set_cell_type(x, y, z, kSolid);
      }
}

// lattice_grid_2d.cpp:143-159
void Domain::set_solid_circle(float cx, float cy, float radius) {
  before_config_change();
  const float r_sq = radius * radius;
  const int min_x = std::max(0, static_cast<int>(cx - radius - 1)), max_x = std::min(nx_ - 1, static_cast<int>(cx + radius + 1));
  const int min_y = std::max(0, static_cast<int>(cy - radius - 1)), max_y = std::min(ny_ - 1, static_cast<int>(cy + radius + 1));
  for (int y = min_y; y <= max_y; ++y)
    for (int x = min_x; x <= max_x; ++x) {
      const float dx = static_cast<float>(x) - cx, dy = static_cast<float>(y) - cy;
      if (dx * dx + dy * dy <= r_sq) set_cell_type(x, y, 0, kSolid);
    }
}

// lattice_grid_3d.cpp:184-203
void Domain::set_solid_cylinder_z(float cx, float cy, float radius) {
  before_config_change();
  is_periodic_ = false;
  const float r_sq = radius * radius;
  const int min_x = std::max(0, static_cast<int>(cx - radius - 1)), max_x = std::min(nx_ - 1, static_cast<int>(cx + radius + 1));
  const int min_y = std::max(0, static_cast<int>(cy - radius - 1)), max_y = std::min(ny_ - 1, static_cast<int>(cy + radius + 1));
  for (int z = 0; z < nz_; ++z)
    for (int y = min_y; y <= max_y; ++y)
      for (int x = min_x; x <= max_x; ++x) {
        const float dx = static_cast<float>(x) - cx, dy = static_cast<float>(y) - cy;
        if (dx * dx + dy * dy <= r_sq) set_cell_type(x, y, z, kSolid);
      }
}

// lattice_grid_3d.cpp:205-215 (3D top wall) and lattice_grid_2d.cpp:178-184
// (2D north wall).
void Domain::set_moving_wall_top(float uwx, float uwz) {
  before_config_change();
  is_periodic_ = false;
  has_moving_wall_ = true;
  uwx_ = uwx;
  uwz_ = dim_ == 2 ? 0.0f : uwz;
  for (int z = 0; z < nz_; ++z)
    for (int x = 0; x < nx_; ++x) set_cell_type(x, ny_ - 1, z, kMovingWall);
}

// lattice_grid_2d.cpp:161-168 (3D: the x = 0 plane)
void Domain::set_inflow_west(float ux, float uy) {
  before_config_change();
  has_inflow_ = true;
  inflow_parabolic_ = false;
  uin_x_ = ux;
  uin_y_ = uy;
  for (int z = 0; z < nz_; ++z)
    for (int y = 0; y < ny_; ++y) set_cell_type(0, y, z, kInflow);
  cls_dirty_ = true;
}

void Domain::set_inflow_parabolic(float u_max) {
  set_inflow_west(u_max, 0.0f);
  inflow_parabolic_ = true;
  umax_ = u_max;
}

void Domain::set_inflow_turbulence(float amplitude, uint64_t seed) {
  before_config_change();
  turb_amp_ = amplitude;
  turb_seed_ = seed;
}

// lattice_grid_2d.cpp:170-176
void Domain::set_outflow_east(float rho_out) {
  before_config_change();
  outflow_ = 1;
  rho_out_ = rho_out;
  for (int z = 0; z < nz_; ++z)
    for (int y = 0; y < ny_; ++y) set_cell_type(nx_ - 1, y, z, kOutflow);
  cls_dirty_ = true;
}

void Domain::set_convective_outflow_east(float u_conv) {
  // the outlet blends in its upstream neighbour x = nx - 2
  if (nx_ < 2) throw HipError("wake: convective outflow needs nx >= 2");
  before_config_change();
  outflow_ = 2;
  u_conv_ = u_conv;
  for (int z = 0; z < nz_; ++z)
    for (int y = 0; y < ny_; ++y) set_cell_type(nx_ - 1, y, z, kOutflow);
  // the outlet store starts from the current post-stream state (S is
  // materialised by before_config_change)
  check(launch_prev_from_s(lattice_, storage_, args(false), buf_[s_buf_], stream_), "outlet store");
  cls_dirty_ = true;
}

// lattice_grid_3d.cpp:217-290: links in z -> y -> x -> q order from Fluid
// nodes to Solid neighbours (x periodic, y/z in range or wrapped on a
// periodic axis), FP32 ray-sphere quadratic, delta clamped to [0.01, 0.99];
// built on the device from the device cell types (geometry.hip).
int Domain::build_curved_sphere_links(float cx, float cy, float cz, float radius) {
  before_config_change();
  is_periodic_ = false;
  if (flags_dirty_) {
    check(hipMemcpyAsync(d_flags_, flags_.data(), N_, hipMemcpyHostToDevice, stream_), "upload flags");
    flags_dirty_ = false;
    cls_dirty_ = true;
  }
  CurvedLinks k;
  check(launch_build_sphere_links(lattice_, d_flags_, nx_, ny_, nz_, ext_axes_ && per_y_, ext_axes_ && per_z_, cx, cy,
                                  cz, radius, stream_, k.fluid_index, k.q_dir, k.delta),
        "build curved links");
  k.uw_x.assign(k.fluid_index.size(), 0.0f);
  k.uw_y.assign(k.fluid_index.size(), 0.0f);
  k.uw_z.assign(k.fluid_index.size(), 0.0f);
  set_curved_links(k);
  return static_cast<int>(links_.fluid_index.size());
}

void Domain::set_curved_links(const CurvedLinks& links) {
  before_config_change();
  links_ = links;
  bfl_dirty_ = cls_dirty_ = true;
}

// ------------------------------------------------------------ device sync
StepArgs Domain::args(bool commit) const {
  StepArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nx = nx_;
  a.ny = ny_;
  a.nz = nz_;
  a.stride = stride_;
  a.periodic_mode = (is_periodic_ && !has_moving_wall_) ? 1 : 0;
  a.per_y = ext_axes_ && per_y_;
  a.per_z = ext_axes_ && per_z_;
  a.bfl_per_y = a.per_y;
  a.bfl_per_z = a.per_z;
  a.tau = tau_;
  a.uw[0] = has_moving_wall_ ? uwx_ : 0.0f;
  a.uw[1] = 0.0f;
  a.uw[2] = has_moving_wall_ && dim_ == 3 ? uwz_ : 0.0f;
  a.inflow = has_inflow_ ? 1 : 0;
  a.inflow_parabolic = inflow_parabolic_ ? 1 : 0;
  a.uin_x = uin_x_;
  a.uin_y = uin_y_;
  a.umax = umax_;
  a.turb_amp = turb_amp_;
  a.turb_seed = turb_seed_;
  a.s_step = step_count_;
  int outflow = outflow_;
  if (outflow == 1 && (dim_ != 2 || a.per_y)) outflow = 0;  // the pressure outlet exists in 2D only
  a.outflow = outflow;
  a.rho_out = rho_out_;
  a.u_conv = u_conv_;
  a.conv_alpha = u_conv_ / (1.0f + u_conv_);
  a.conv_beta = 1.0f / (1.0f + u_conv_);
  a.by0 = a.per_y ? 0 : 1;
  a.by1 = a.per_y ? ny_ : ny_ - 1;
  a.bz0 = dim_ == 2 ? 0 : (a.per_z ? 0 : 1);
  a.bz1 = dim_ == 2 ? 1 : (a.per_z ? nz_ : nz_ - 1);
  a.commit_prev = commit ? 1 : 0;
  a.global_addr = buffer_addressing_ ? 0 : 1;
  a.f_prev = d_prev_;
  a.flags = d_flags_;
  a.cls = d_cls_;
  a.aux = d_aux_;
  a.bmask = d_bmask_;
  a.bfl = d_bfl_;
  a.gx = g_last_[0];
  a.gy = g_last_[1];
  a.gz = g_last_[2];
  a.tau_g = tau_g_;
  a.gbeta = gbeta_;
  a.t_ref = t_ref_;
  a.t_wall = d_twall_;
  a.nonfinite = d_nonfinite_;
  return a;
}

void Domain::sync_device_config() {
  if (twall_dirty_) {
    check(hipMemcpyAsync(d_twall_, h_twall_.data(), N_ * sizeof(float), hipMemcpyHostToDevice, stream_),
          "upload wall temperatures");
    twall_dirty_ = false;
  }
  if (flags_dirty_) {
    check(hipMemcpyAsync(d_flags_, flags_.data(), N_, hipMemcpyHostToDevice, stream_), "upload flags");
    flags_dirty_ = false;
    cls_dirty_ = true;
  }
  if (bfl_dirty_) {
    // group the links (z -> y -> x -> q order) into one record per fluid node
    const LatTab L = tab(lattice_);
    std::vector<BflNode> recs;
    std::vector<int> nodes;
    for (size_t i = 0; i < links_.fluid_index.size(); ++i) {
      const int n = links_.fluid_index[i];
      if (nodes.empty() || nodes.back() != n) {
        nodes.push_back(n);
        BflNode r;
        std::memset(&r, 0, sizeof(r));
        recs.push_back(r);
      }
      BflNode& r = recs.back();
      const int q = links_.q_dir[i];
      r.mask |= 1u << q;
      r.delta[q] = links_.delta[i];
      r.ciuw[q] = static_cast<float>(L.cx[q]) * links_.uw_x[i] + static_cast<float>(L.cy[q]) * links_.uw_y[i] +
                  static_cast<float>(L.cz[q]) * links_.uw_z[i];
    }
    if (d_bfl_) {
      (void)hipStreamSynchronize(stream_);
      (void)hipFree(d_bfl_);
      (void)hipFree(d_bfl_nodes_);
      allocations_.erase(std::remove(allocations_.begin(), allocations_.end(), static_cast<void*>(d_bfl_)),
                         allocations_.end());
      allocations_.erase(std::remove(allocations_.begin(), allocations_.end(), static_cast<void*>(d_bfl_nodes_)),
                         allocations_.end());
      d_bfl_ = nullptr;
      d_bfl_nodes_ = nullptr;
    }
    n_bfl_nodes_ = static_cast<int>(nodes.size());
    if (n_bfl_nodes_ > 0) {
      d_bfl_ = static_cast<BflNode*>(alloc(recs.size() * sizeof(BflNode)));
      d_bfl_nodes_ = static_cast<int*>(alloc(nodes.size() * sizeof(int)));
      check(hipMemcpyAsync(d_bfl_, recs.data(), recs.size() * sizeof(BflNode), hipMemcpyHostToDevice, stream_), "bfl");
      check(hipMemcpyAsync(d_bfl_nodes_, nodes.data(), nodes.size() * sizeof(int), hipMemcpyHostToDevice, stream_),
            "bfl");
      check(hipStreamSynchronize(stream_), "bfl");
    }
    bfl_dirty_ = false;
    cls_dirty_ = true;
  }
  if (cls_dirty_) {
    all_fluid_ = std::all_of(flags_.begin(), flags_.end(), [](uint8_t t) { return t == kFluid; });
    const StepArgs a = args(false);
    check(launch_build_class(lattice_, a, d_cls_, d_bmask_, stream_), "build class");
    check(launch_mark_bfl(d_bfl_nodes_, n_bfl_nodes_, d_cls_, d_aux_, stream_), "mark bfl");
    cls_dirty_ = false;
    ++cls_epoch_;
  }
}

void Domain::materialize() {
  sync_device_config();
  if (probe_) {
    // the force sample of S_k is taken from the unrounded S, as a pull
    // launch would take it, before the state is stored
    probe_sync();
    probe_sample_current();
  }
  const StepArgs a = args(true);
  const int out = 1 - a_buf_;
  check(launch_materialize(lattice_, storage_, a, buf_[a_buf_], buf_[out], stream_), "materialize");
  if (thermal_) check(launch_materialize_thermal(a, gbuf_[a_buf_], gbuf_[out], stream_), "materialize temperature");
  s_buf_ = out;
}

// ------------------------------------------------------------------ step
bool Domain::tb_eligible() {
  if (dim_ != 2 || lattice_ != 9 || thermal_ || probe_ || s_buf_ >= 0 || !all_fluid_ || !tb_enabled_) return false;
  // Only where single-step launches are bound by the launch rate: measured,
  // 8 steps per launch reach about 75 G node-steps/s (the tile loads and
  // the 2.15x redundant halo updates bound it), while single steps of larger
  // domains run at the HBM rate (FP32: 58.5k vs 71.3k MLUPS at 1024^2,
  // 82.8k vs 77.5k at 4096^2; FP16: 92.3k vs 65.2k at 1024^2).
  if (N_ > (storage_ == kStoreF32 ? (1u << 21) : (1u << 17))) return false;
  const StepArgs a = args(true);
  return a.periodic_mode && !a.inflow && !a.outflow;
}

void Domain::step_tb(int collision, float gx, float gy, bool use_les, float smag) {
  version_ += kTbSteps;
  StepArgs a = args(true);
  a.tau = tau_;
  a.gx = gx;
  a.gy = gy;
  a.gz = 0.0f;
  a.cs = smag;
  const int out = 1 - a_buf_;
  check(launch_step_tb(lattice_, storage_, collision, use_les && les_coupled_, a, buf_[a_buf_], buf_[out], stream_),
        "periodic step");
  a_buf_ = out;
  g_last_[0] = gx;
  g_last_[1] = gy;
  g_last_[2] = 0.0f;
  step_count_ += kTbSteps;
  invalidate_mirrors();
}

void Domain::step(int collision, float gx, float gy, float gz, bool use_les, float smag) {
  ++version_;
  sync_device_config();
  if (probe_) probe_sync();
  if (dim_ == 2) gz = 0.0f;
  StepArgs a = args(true);
  a.tau = tau_;
  a.gx = gx;
  a.gy = gy;
  a.gz = gz;
  a.cs = smag;
  const bool les = use_les && les_coupled_;
  if (thermal_) {
    // coupled flow + temperature: one kernel, open boundaries inline
    const bool pull = s_buf_ < 0;
    const int in = pull ? a_buf_ : s_buf_;
    const int out = 1 - in;
    check(launch_step_thermal(storage_, collision, pull, a, buf_[in], buf_[out], gbuf_[in], gbuf_[out], stream_),
          "thermal step");
    a_buf_ = out;
    s_buf_ = -1;
  } else if (s_buf_ >= 0) {
    // S_k is materialised: the collide-only launch forms no S, so the force
    // sample of S_k is taken from the materialised state
    if (probe_) probe_sample_current();
    const int out = 1 - s_buf_;
    check(launch_step(lattice_, storage_, collision, les, false, a, buf_[s_buf_], buf_[out], block_x_, stream_),
          "step");
    a_buf_ = out;
    s_buf_ = -1;
  } else {
    const int out = 1 - a_buf_;
    // the pull forms S_k: its force sample is taken inside the launch
    if (probe_ && fh_next_ == step_count_) {
      a = with_mem(a, probe_rows_, probe_slot(step_count_), probe_c_, probe_rmax_);
      if (probe_rows_.nrows == 0) a.mem_rows = nullptr;
      fh_next_ = step_count_ + 1;
    }
    // A timed run brackets every kTimingStride-th pull launch with events:
    // a pair around every launch costs about 8 us (1.7% of a 256^3 FP32
    // step), which the sampled average no longer pays.
    // The start event is recorded after a plain marker: recorded straight
    // after the previous (unsampled) kernel it read the launch gap as part
    // of the kernel (462 vs 459 us per 256^3 FP32 step, rocprof 459 us).
    const bool sample = time_kernels_ && (timed_launches_++ % kTimingStride == 0);
    if (sample) {
      if (kev_used_ + 2 > kev_.size()) {
        for (int e = 0; e < 2; ++e) {
          hipEvent_t ev;
          check(hipEventCreate(&ev), "event");
          kev_.push_back(ev);
        }
      }
      if (!kev_marker_) check(hipEventCreate(&kev_marker_), "event");
      check(hipEventRecord(kev_marker_, stream_), "event");
      check(hipEventRecord(kev_[kev_used_], stream_), "event");
    }
    check(launch_step(lattice_, storage_, collision, les, true, a, buf_[a_buf_], buf_[out], block_x_, stream_),
          "step");
    if (sample) {
      check(hipEventRecord(kev_[kev_used_ + 1], stream_), "event");
      kev_used_ += 2;
    }
    a_buf_ = out;
  }
  g_last_[0] = gx;
  g_last_[1] = gy;
  g_last_[2] = gz;
  ++step_count_;
  invalidate_mirrors();
}

float Domain::run(int steps, int collision, float gx, float gy, float gz, bool use_les, float smag, bool timed) {
  sync_device_config();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (timed) {
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    check(hipEventRecord(e0, stream_), "event");
  }
  time_kernels_ = timed;
  kev_used_ = 0;
  timed_launches_ = 0;
  try {
    int s = 0;
    // fully periodic D2Q9 without special nodes: kTbSteps steps per launch
    // (an untimed run only: the kernel timer samples single-step launches)
    if (!timed && steps >= kTbSteps && tb_eligible()) {
      for (; steps - s >= kTbSteps; s += kTbSteps) step_tb(collision, gx, gy, use_les, smag);
    }
    for (; s < steps; ++s) step(collision, gx, gy, gz, use_les, smag);
  } catch (...) {
    time_kernels_ = false;
    throw;
  }
  time_kernels_ = false;
  float ms = 0.0f;
  if (timed) {
    check(hipEventRecord(e1, stream_), "event");
    check(hipEventSynchronize(e1), "event");
    check(hipEventElapsedTime(&ms, e0, e1), "event");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    double sum = 0.0;
    for (size_t k = 0; k + 1 < kev_used_; k += 2) {
      float t = 0.0f;
      check(hipEventElapsedTime(&t, kev_[k], kev_[k + 1]), "event");
      sum += t;
    }
    last_kernel_ms_ = kev_used_ ? static_cast<float>(sum / (kev_used_ / 2)) : 0.0f;
  }
  return ms;
}

void Domain::synchronize() { check(hipStreamSynchronize(stream_), "synchronize"); }

// ------------------------------------------------------------- read-back
void Domain::ensure_device_fields() {
  if (fields_valid_ || dfields_valid_) return;
  sync_device_config();
  StepArgs a = args(false);
  a.rho = d_rho_;
  a.ux = d_ux_;
  a.uy = d_uy_;
  a.uz = d_uz_;
  a.speed = d_speed_;
  const bool from_s = s_buf_ >= 0;
  const int b = from_s ? s_buf_ : a_buf_;
  if (thermal_) {
    a.temp = d_temp_;
    check(launch_readback_thermal(storage_, from_s, a, buf_[b], gbuf_[b], stream_), "readback");
  } else {
    check(launch_readback(lattice_, storage_, from_s, a, buf_[b], stream_), "readback");
  }
  dfields_valid_ = true;
}

void Domain::ensure_fields() {
  if (fields_valid_) return;
  ensure_device_fields();
  if (thermal_) check(hipMemcpyAsync(h_temp_.data(), d_temp_, N_ * 4, hipMemcpyDeviceToHost, stream_), "d2h");
  float* d[5] = {d_rho_, d_ux_, d_uy_, d_uz_, d_speed_};
  std::vector<float>* h[5] = {&h_rho_, &h_ux_, &h_uy_, &h_uz_, &h_speed_};
  for (int k = 0; k < 5; ++k) check(hipMemcpyAsync(h[k]->data(), d[k], N_ * 4, hipMemcpyDeviceToHost, stream_), "d2h");
  check(hipStreamSynchronize(stream_), "readback");
  fields_valid_ = true;
}

// lattice_grid_3d.cpp:292-299: unforced moments of f_src, then |u|.
void Domain::update_macroscopic() {
  const float keep[3] = {g_last_[0], g_last_[1], g_last_[2]};
  g_last_[0] = g_last_[1] = g_last_[2] = 0.0f;
  fields_valid_ = dfields_valid_ = false;
  ensure_fields();
  g_last_[0] = keep[0];
  g_last_[1] = keep[1];
  g_last_[2] = keep[2];
}

const float* Domain::analytics(int which) {
  if (which < 0 || which > 3) throw HipError("unknown analytics field");
  if (dim_ == 2 && which != 2) throw HipError("a 2D domain has only the z vorticity");
  if (!an_valid_) {
    ensure_fields();
    const int k0 = dim_ == 2 ? 2 : 0, k1 = dim_ == 2 ? 2 : 3;
    for (int k = k0; k <= k1; ++k)
      if (!d_an_[k]) {
        d_an_[k] = static_cast<float*>(alloc(N_ * sizeof(float)));
        h_an_[k].resize(N_);
      }
    if (dim_ == 2)
      check(launch_vorticity_2d(d_ux_, d_uy_, d_an_[2], nx_, ny_, stream_), "vorticity");
    else
      check(launch_vorticity_q_3d(d_ux_, d_uy_, d_uz_, d_an_[0], d_an_[1], d_an_[2], d_an_[3], nx_, ny_, nz_, stream_),
            "vorticity");
    for (int k = k0; k <= k1; ++k)
      check(hipMemcpyAsync(h_an_[k].data(), d_an_[k], N_ * sizeof(float), hipMemcpyDeviceToHost, stream_), "d2h");
    check(hipStreamSynchronize(stream_), "vorticity");
    an_valid_ = true;
  }
  return h_an_[which].data();
}

const float* Domain::device_field(int which) {
  if (which < 0 || which > 4) throw HipError("unknown device field");
  ensure_device_fields();
  const float* d[5] = {d_rho_, d_ux_, d_uy_, d_uz_, d_speed_};
  return d[which];
}

const uint8_t* Domain::device_cell_types() {
  sync_device_config();
  return d_flags_;
}

const float* Domain::rho() { ensure_fields(); return h_rho_.data(); }
const float* Domain::ux() { ensure_fields(); return h_ux_.data(); }
const float* Domain::uy() { ensure_fields(); return h_uy_.data(); }
const float* Domain::uz() { ensure_fields(); return h_uz_.data(); }
const float* Domain::speed() { ensure_fields(); return h_speed_.data(); }

const float* Domain::f_src() {
  if (fsrc_valid_) return h_fsrc_.data();
  sync_device_config();
  if (!d_sout_) d_sout_ = static_cast<float*>(alloc(static_cast<size_t>(Q_) * N_ * 4));
  StepArgs a = args(false);
  a.rho = d_rho_;
  a.ux = d_ux_;
  a.uy = d_uy_;
  a.uz = d_uz_;
  a.speed = d_speed_;
  a.s_out = d_sout_;
  const bool from_s = s_buf_ >= 0;
  check(launch_readback(lattice_, storage_, from_s, a, buf_[from_s ? s_buf_ : a_buf_], stream_), "readback");
  h_fsrc_.resize(static_cast<size_t>(Q_) * N_);
  check(hipMemcpyAsync(h_fsrc_.data(), d_sout_, h_fsrc_.size() * 4, hipMemcpyDeviceToHost, stream_), "d2h");
  if (!fields_valid_) {
    float* d[5] = {d_rho_, d_ux_, d_uy_, d_uz_, d_speed_};
    std::vector<float>* h[5] = {&h_rho_, &h_ux_, &h_uy_, &h_uz_, &h_speed_};
    for (int k = 0; k < 5; ++k)
      check(hipMemcpyAsync(h[k]->data(), d[k], N_ * 4, hipMemcpyDeviceToHost, stream_), "d2h");
    fields_valid_ = true;
  }
  check(hipStreamSynchronize(stream_), "readback");
  fsrc_valid_ = true;
  sout_is_fsrc_ = true;
  return h_fsrc_.data();
}

// After step k the reference's f_dst holds A_{k-1}; before any step it holds
// the initial state.
const float* Domain::f_dst() {
  if (fdst_valid_) return h_fdst_.data();
  if (a_buf_ < 0) {
    h_fdst_.assign(f_src(), f_src() + static_cast<size_t>(Q_) * N_);
    fdst_valid_ = true;
    return h_fdst_.data();
  }
  if (!d_sout_) d_sout_ = static_cast<float*>(alloc(static_cast<size_t>(Q_) * N_ * 4));
  StepArgs a = args(false);
  a.s_out = d_sout_;
  // planes only: read the buffer as-is (no moments written)
  check(launch_readback(lattice_, storage_, true, a, buf_[a_buf_], stream_), "readback");
  h_fdst_.resize(static_cast<size_t>(Q_) * N_);
  check(hipMemcpyAsync(h_fdst_.data(), d_sout_, h_fdst_.size() * 4, hipMemcpyDeviceToHost, stream_), "d2h");
  check(hipStreamSynchronize(stream_), "readback");
  fdst_valid_ = true;
  fsrc_valid_ = false;  // d_sout_ now holds f_dst
  sout_is_fsrc_ = false;
  return h_fdst_.data();
}

void Domain::diagnostics(double out[3]) {
  ensure_fields();
  sync_device_config();  // the reduction reads the device cell types
  check(launch_diagnostics(d_flags_, d_rho_, d_speed_, static_cast<long long>(N_), d_partial_, 1024, d_out_, stream_),
        "diagnostics");
  check(hipMemcpyAsync(out, d_out_, 3 * sizeof(double), hipMemcpyDeviceToHost, stream_), "d2h");
  check(hipStreamSynchronize(stream_), "diagnostics");
}

double Domain::total_mass() {
  double d[3];
  diagnostics(d);
  return d[0];
}

double Domain::total_kinetic_energy() {
  double d[3];
  diagnostics(d);
  return d[1];
}

double Domain::max_speed() {
  double d[3];
  diagnostics(d);
  return d[2];
}

// ------------------------------------------------------ momentum exchange
// Linked rows of (centre, r_max) from the device cell types: the rows are
// marked on the device, numbered here in (z, y) order, each given
// mem_waves_per_row(nx) partial slots.
Domain::MemRows Domain::build_mem_rows(float ccx, float ccy, float ccz, float r_max) {
  const size_t rows = static_cast<size_t>(ny_) * nz_;
  MemRows r;
  r.d_rows = static_cast<int*>(alloc(rows * sizeof(int)));
  StepArgs a = args(false);
  a.mem_c[0] = ccx;
  a.mem_c[1] = ccy;
  a.mem_c[2] = ccz;
  a.mem_rmax = r_max;
  check(hipMemsetAsync(r.d_rows, 0, rows * sizeof(int), stream_), "mark rows");
  check(launch_mem_mark_rows(lattice_, a, r.d_rows, stream_), "mark rows");
  std::vector<int> has(rows);
  check(hipMemcpyAsync(has.data(), r.d_rows, rows * sizeof(int), hipMemcpyDeviceToHost, stream_), "mark rows");
  check(hipStreamSynchronize(stream_), "mark rows");
  const int W = mem_waves_per_row(nx_);
  std::vector<int> list;
  for (size_t i = 0; i < rows; ++i) {
    if (has[i]) {
      has[i] = static_cast<int>(list.size()) * W;
      list.push_back(static_cast<int>(i));
    } else {
      has[i] = -1;
    }
  }
  r.nrows = static_cast<int>(list.size());
  r.nslots = r.nrows * W;
  r.d_list = static_cast<int*>(alloc(std::max<size_t>(1, list.size()) * sizeof(int)));
  check(hipMemcpyAsync(r.d_rows, has.data(), rows * sizeof(int), hipMemcpyHostToDevice, stream_), "rows");
  if (!list.empty())
    check(hipMemcpyAsync(r.d_list, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice, stream_), "rows");
  check(hipStreamSynchronize(stream_), "rows");  // the host vectors go out of scope
  return r;
}

void Domain::free_mem_rows(MemRows& r) {
  if (r.d_rows || r.d_list) check(hipStreamSynchronize(stream_), "free rows");
  release(r.d_rows);
  release(r.d_list);
  r = MemRows{};
}

const Domain::MemRows& Domain::cached_mem_rows(float ccx, float ccy, float ccz, float r_max) {
  const auto key = std::make_tuple(ccx, ccy, ccz, r_max, cls_epoch_);
  auto it = mem_rows_cache_.find(key);
  if (it != mem_rows_cache_.end()) return it->second;
  if (mem_rows_cache_.size() >= 4) {
    for (auto& kv : mem_rows_cache_) free_mem_rows(kv.second);
    mem_rows_cache_.clear();
  }
  return mem_rows_cache_.emplace(key, build_mem_rows(ccx, ccy, ccz, r_max)).first->second;
}

StepArgs Domain::with_mem(StepArgs a, const MemRows& r, double* part, const float c[3], float r_max) const {
  a.mem_rows = r.d_rows;
  a.mem_part = part;
  a.mem_c[0] = c[0];
  a.mem_c[1] = c[1];
  a.mem_c[2] = c[2];
  a.mem_rmax = r_max;
  return a;
}

// Forces on the current S_k: the partials of the linked rows (S_k formed at
// the nodes next to a wall, or read from a materialised state), summed in
// slot order.  The same kernels and order as the per-step history.
void Domain::momentum_exchange(float ccx, float ccy, float ccz, float r_max, double out[3]) {
  sync_device_config();
  const MemRows& r = cached_mem_rows(ccx, ccy, ccz, r_max);
  if (r.nrows == 0) {
    out[0] = out[1] = out[2] = 0.0;
    return;
  }
  const size_t need = 3 * static_cast<size_t>(r.nslots);
  if (need > mempart_cap_) {
    if (d_mempart_) {
      check(hipStreamSynchronize(stream_), "momentum exchange");
      release(d_mempart_);
    }
    d_mempart_ = static_cast<double*>(alloc(need * sizeof(double)));
    mempart_cap_ = need;
  }
  const float c[3] = {ccx, ccy, ccz};
  const StepArgs a = with_mem(args(false), r, d_mempart_, c, r_max);
  const bool from_s = s_buf_ >= 0;
  check(launch_mem_rows(lattice_, storage_, from_s, a, buf_[from_s ? s_buf_ : a_buf_], r.d_list, r.nrows, stream_),
        "momentum exchange");
  check(launch_mem_fold(d_mempart_, r.nslots, 1, 0, 1, d_out_, 0, stream_), "momentum exchange");
  check(hipMemcpyAsync(out, d_out_, 3 * sizeof(double), hipMemcpyDeviceToHost, stream_), "d2h");
  check(hipStreamSynchronize(stream_), "momentum exchange");
}

// -------------------------------------------------------------- divergence
void Domain::clear_nonfinite() { check(hipMemsetAsync(d_nonfinite_, 0, sizeof(int), stream_), "clear flag"); }

bool Domain::nonfinite_seen() {
  int v = 0;
  check(hipMemcpyAsync(&v, d_nonfinite_, sizeof(int), hipMemcpyDeviceToHost, stream_), "d2h");
  check(hipStreamSynchronize(stream_), "divergence flag");
  return v != 0;
}

long long Domain::count_nonfinite(long long* first) {
  ensure_device_fields();
  unsigned long long h[2] = {0ull, ~0ull};
  check(hipMemcpyAsync(d_nfcount_, h, sizeof(h), hipMemcpyHostToDevice, stream_), "non-finite count");
  check(launch_count_nonfinite(d_rho_, d_speed_, static_cast<long long>(N_), d_nfcount_, stream_),
        "non-finite count");
  check(hipMemcpyAsync(h, d_nfcount_, sizeof(h), hipMemcpyDeviceToHost, stream_), "d2h");
  check(hipStreamSynchronize(stream_), "non-finite count");
  if (first) *first = h[0] ? static_cast<long long>(h[1]) : -1;
  return static_cast<long long>(h[0]);
}

// ------------------------------------------------------------- force probe
void Domain::set_force_probe(float ccx, float ccy, float ccz, float r_max) {
  if (thermal_) throw HipError("the force history is not available on temperature-coupled domains");
  probe_ = true;
  probe_c_[0] = ccx;
  probe_c_[1] = ccy;
  probe_c_[2] = ccz;
  probe_rmax_ = r_max;
  if (probe_rows_.d_rows) free_mem_rows(probe_rows_);
  probe_epoch_ = ~0ull;
  if (!d_fhist_) d_fhist_ = static_cast<double*>(alloc(3 * static_cast<size_t>(kForceHistCap) * sizeof(double)));
  probe_restart();
}

void Domain::clear_force_probe() {
  probe_ = false;
  if (probe_rows_.d_rows) free_mem_rows(probe_rows_);
  probe_epoch_ = ~0ull;
  h_fhist_.clear();
}

void Domain::probe_restart() {
  fh_first_ = fh_next_ = fh_folded_ = fh_dbase_ = fh_host_ = step_count_;
  h_fhist_.clear();
}

// Rows for the current class bytes; samples written with the previous rows
// are folded first (their slot layout goes with those rows).
void Domain::probe_sync() {
  if (probe_epoch_ == cls_epoch_) return;
  probe_fold(fh_next_);
  if (probe_rows_.d_rows) free_mem_rows(probe_rows_);
  probe_rows_ = build_mem_rows(probe_c_[0], probe_c_[1], probe_c_[2], probe_rmax_);
  probe_epoch_ = cls_epoch_;
  // one slot at least: a geometry without links has all-zero samples
  const size_t need = 3 * static_cast<size_t>(kForceRing) * std::max(1, probe_rows_.nslots);
  if (need > ring_cap_) {
    if (d_ring_) {
      check(hipStreamSynchronize(stream_), "force ring");
      release(d_ring_);
    }
    d_ring_ = static_cast<double*>(alloc(need * sizeof(double)));
    ring_cap_ = need;
  }
  if (probe_rows_.nrows == 0) check(hipMemsetAsync(d_ring_, 0, need * sizeof(double), stream_), "force ring");
}

void Domain::probe_fold(uint64_t upto) {
  if (upto <= fh_folded_) return;
  if (upto - fh_dbase_ > static_cast<uint64_t>(kForceHistCap)) {
    // the device history is full: move what is folded to the host
    const size_t n = static_cast<size_t>(fh_folded_ - fh_host_);
    const size_t at = h_fhist_.size();
    h_fhist_.resize(at + 3 * n);
    if (n)
      check(hipMemcpyAsync(h_fhist_.data() + at, d_fhist_ + 3 * (fh_host_ - fh_dbase_), 3 * n * sizeof(double),
                           hipMemcpyDeviceToHost, stream_),
            "force history");
    check(hipStreamSynchronize(stream_), "force history");
    fh_host_ = fh_dbase_ = fh_folded_;
  }
  const int count = static_cast<int>(upto - fh_folded_);
  check(launch_mem_fold(d_ring_, std::max(1, probe_rows_.nslots), kForceRing, static_cast<long long>(fh_folded_),
                        count, d_fhist_, static_cast<long long>(fh_folded_ - fh_dbase_), stream_),
        "force fold");
  fh_folded_ = upto;
}

double* Domain::probe_slot(uint64_t k) {
  if (k - fh_folded_ >= static_cast<uint64_t>(kForceRing)) probe_fold(k);
  return d_ring_ + 3 * static_cast<size_t>(std::max(1, probe_rows_.nslots)) * (k % kForceRing);
}

// The sample of S_{step_count_} when no launch has formed it yet: read from
// a materialised state or formed from the post-collision buffer.
void Domain::probe_sample_current() {
  if (fh_next_ != step_count_) return;
  double* part = probe_slot(step_count_);
  if (probe_rows_.nrows > 0) {
    const StepArgs a = with_mem(args(false), probe_rows_, part, probe_c_, probe_rmax_);
    const bool from_s = s_buf_ >= 0;
    check(launch_mem_rows(lattice_, storage_, from_s, a, buf_[from_s ? s_buf_ : a_buf_], probe_rows_.d_list,
                          probe_rows_.nrows, stream_),
          "force sample");
  }
  fh_next_ = step_count_ + 1;
}

const std::vector<double>& Domain::force_history(uint64_t* first_step) {
  if (!probe_) throw HipError("no force probe is set on this domain");
  sync_device_config();
  probe_sync();
  probe_sample_current();
  probe_fold(fh_next_);
  const size_t n = static_cast<size_t>(fh_folded_ - fh_host_);
  if (n) {
    const size_t at = h_fhist_.size();
    h_fhist_.resize(at + 3 * n);
    check(hipMemcpyAsync(h_fhist_.data() + at, d_fhist_ + 3 * (fh_host_ - fh_dbase_), 3 * n * sizeof(double),
                         hipMemcpyDeviceToHost, stream_),
          "force history");
    check(hipStreamSynchronize(stream_), "force history");
    fh_host_ = fh_folded_;
  }
  if (first_step) *first_step = fh_first_;
  return h_fhist_;
}

// ------------------------------------------------------------- temperature
void Domain::enable_thermal(float tau_g, float gbeta, float t_ref, float t_init) {
  if (dim_ != 2) throw HipError("the temperature coupling is 2D (D2Q9 + D2Q5) only");
  if (probe_) throw HipError("clear the force probe before enabling the temperature coupling");
  if (!(tau_g > 0.5f)) throw HipError("temperature relaxation time must exceed 1/2");
  before_config_change();
  if (!thermal_) {
    for (int k = 0; k < 2; ++k) gbuf_[k] = static_cast<float*>(alloc(5 * static_cast<size_t>(stride_) * sizeof(float)));
    d_twall_ = static_cast<float*>(alloc(N_ * sizeof(float)));
    d_temp_ = static_cast<float*>(alloc(N_ * sizeof(float)));
    h_temp_.assign(N_, 0.0f);
  }
  h_twall_.assign(N_, std::numeric_limits<float>::quiet_NaN());
  twall_dirty_ = true;
  thermal_ = true;
  tau_g_ = tau_g;
  gbeta_ = gbeta;
  t_ref_ = t_ref;
  std::vector<float> t(N_, t_init);
  set_temperature_field(t.data());
}

void Domain::set_temperature_field(const float* t) {
  if (!thermal_) throw HipError("temperature is not enabled on this domain");
  before_config_change();
  std::vector<float> host(5 * static_cast<size_t>(stride_), 0.0f);
  for (int q = 0; q < 5; ++q)
    for (size_t i = 0; i < N_; ++i) host[q * static_cast<size_t>(stride_) + i] = D2Q5::w[q] * t[i];
  for (int k = 0; k < 2; ++k)
    check(hipMemcpyAsync(gbuf_[k], host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice, stream_),
          "upload temperature");
  check(hipStreamSynchronize(stream_), "upload temperature");
  fields_valid_ = dfields_valid_ = false;
}

void Domain::set_wall_temperature(int x0, int y0, int x1, int y1, float t_wall) {
  if (!thermal_) throw HipError("temperature is not enabled on this domain");
  before_config_change();
  for (int y = std::max(0, y0); y <= std::min(ny_ - 1, y1); ++y)
    for (int x = std::max(0, x0); x <= std::min(nx_ - 1, x1); ++x) h_twall_[static_cast<size_t>(y) * nx_ + x] = t_wall;
  twall_dirty_ = true;
}

const float* Domain::temperature() {
  if (!thermal_) throw HipError("temperature is not enabled on this domain");
  ensure_fields();
  return h_temp_.data();
}

// ------------------------------------------------------------ batched steps
void Domain::batch_prepare(float gx, float gy, float gz, float smag, StepArgs& a, void* buf[2], float* gbuf[2],
                           int& parity, bool& s_state) {
  sync_device_config();
  if (dim_ == 2) gz = 0.0f;
  a = args(true);
  a.gx = gx;
  a.gy = gy;
  a.gz = gz;
  a.cs = smag;
  buf[0] = buf_[0];
  buf[1] = buf_[1];
  gbuf[0] = gbuf_[0];
  gbuf[1] = gbuf_[1];
  s_state = s_buf_ >= 0;
  parity = s_state ? s_buf_ : a_buf_;
  // uploads were issued on the domain's stream; the batch runs on its own
  check(hipStreamSynchronize(stream_), "batch prepare");
}

void Domain::batch_advance(uint64_t k, int parity_after, float gx, float gy, float gz) {
  if (k == 0) return;
  a_buf_ = parity_after;
  s_buf_ = -1;
  step_count_ += k;
  g_last_[0] = gx;
  g_last_[1] = gy;
  g_last_[2] = dim_ == 2 ? 0.0f : gz;
  invalidate_mirrors();
}

const int* Domain::batch_mem_nodes(float ccx, float ccy, float ccz, float r_max, int* count) {
  sync_device_config();
  const auto key = std::make_tuple(ccx, ccy, ccz, r_max, cls_epoch_);
  if (key != batch_rows_key_ || !batch_rows_.d_list) {
    if (batch_rows_.d_rows) free_mem_rows(batch_rows_);
    if (d_batch_nodes_) {
      check(hipStreamSynchronize(stream_), "batch nodes");
      release(d_batch_nodes_);
      d_batch_nodes_ = nullptr;
    }
    batch_rows_ = build_mem_rows(ccx, ccy, ccz, r_max);
    batch_rows_key_ = key;
    n_batch_nodes_ = 0;
    const int nr = batch_rows_.nrows;
    if (nr > 0) {
      StepArgs a = args(false);
      a.mem_c[0] = ccx;
      a.mem_c[1] = ccy;
      a.mem_c[2] = ccz;
      a.mem_rmax = r_max;
      // per-row counts, their prefix on the host, then the nodes
      int* d_cnt = static_cast<int*>(alloc(2 * static_cast<size_t>(nr) * sizeof(int)));
      check(launch_mem_row_nodes(lattice_, a, batch_rows_.d_list, nr, d_cnt, nullptr, nullptr, stream_),
            "batch nodes");
      std::vector<int> cnt(nr), off(nr);
      check(hipMemcpyAsync(cnt.data(), d_cnt, nr * sizeof(int), hipMemcpyDeviceToHost, stream_), "batch nodes");
      check(hipStreamSynchronize(stream_), "batch nodes");
      for (int r = 0; r < nr; ++r) {
        off[r] = n_batch_nodes_;
        n_batch_nodes_ += cnt[r];
      }
      d_batch_nodes_ = static_cast<int*>(alloc(std::max(1, n_batch_nodes_) * sizeof(int)));
      check(hipMemcpyAsync(d_cnt + nr, off.data(), nr * sizeof(int), hipMemcpyHostToDevice, stream_), "batch nodes");
      check(launch_mem_row_nodes(lattice_, a, batch_rows_.d_list, nr, d_cnt, d_cnt + nr, d_batch_nodes_, stream_),
            "batch nodes");
      check(hipStreamSynchronize(stream_), "batch nodes");
      release(d_cnt);
    }
  }
  *count = n_batch_nodes_;
  return d_batch_nodes_;
}

size_t Domain::bytes_needed(int dim, int lattice, int nx, int ny, int nz, int storage, long long plane_pad,
                            bool thermal) {
  const int Q = dim == 2 ? 9 : (lattice == 27 ? 27 : 19);
  const size_t N = static_cast<size_t>(nx) * ny * (dim == 2 ? 1 : nz);
  const long long base = static_cast<long long>((N + 255) / 256 * 256);
  const long long stride = base + (plane_pad >= 0 ? plane_pad : (N >= (1u << 20) ? 1024 : 0));
  const size_t yz = static_cast<size_t>(ny) * (dim == 2 ? 1 : nz);
  const size_t t = thermal ? 2 * 5 * static_cast<size_t>(stride) * 4 + N * 8 : 0;
  return 2 * static_cast<size_t>(Q) * stride * elem_size(storage) + N * (1 + 1 + 4 + 4 + 5 * 4) + prev_stride(Q) * yz * 4 +
         3 * 1024 * 8 + 64 + t;
}

}  // namespace wake_hip
