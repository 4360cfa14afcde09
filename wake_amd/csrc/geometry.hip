// geometry.hip -- triangle-mesh voxelizer and curved-wall link builder on
// the device (gfx950).
//
// src/geometry/obj_loader.cpp:208-400: every cell whose unit box overlaps a
// triangle (separating-axis test) becomes Solid; optionally every cell that
// a 6-neighbour walk through non-Solid cells cannot reach from the domain's
// outer faces becomes Solid too.  The cell-type bytes must equal the CPU
// voxelizer's bit for bit, so the overlap test evaluates the same FP32
// expressions in the same order, each product and sum rounded on its own
// (this file is built without FP contraction).
//
// Work layout: the host gives every triangle its clamped cell box and an
// exclusive prefix of the box volumes; one lane per (triangle, cell) pair,
// found by binary search, so a mesh of many small triangles and one of a few
// large ones both fill the chip.  The exterior is found by line sweeps
// (along x, y and z in turn, repeated until nothing changes): the set of
// reachable cells does not depend on the order it is discovered in, so this
// equals the CPU breadth-first search.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.hpp"
#include "lattice.hpp"

namespace wake_hip {
namespace {

constexpr uint8_t kCellSolid = 1;

struct V3 {
  float x, y, z;
};

__device__ inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

__device__ inline float min3(float a, float b, float c) { return fminf(a, fminf(b, c)); }
__device__ inline float max3(float a, float b, float c) { return fmaxf(a, fmaxf(b, c)); }

// One candidate separating axis: the projections a*p - b*q of the three
// vertices against the box radius fa*h1 + fb*h2 (obj_loader.cpp:262-271).
__device__ inline bool axis_overlaps(float a, float b, float fa, float fb, const V3 (&v)[3], int ip, int iq, float h1,
                                     float h2) {
  float s[3];
  for (int m = 0; m < 3; ++m) {
    const float p = ip == 0 ? v[m].x : ip == 1 ? v[m].y : v[m].z;
    const float q = iq == 0 ? v[m].x : iq == 1 ? v[m].y : v[m].z;
    s[m] = a * p - b * q;
  }
  const float r = fa * h1 + fb * h2;
  return min3(s[0], s[1], s[2]) <= r && max3(s[0], s[1], s[2]) >= -r;
}

// obj_loader.cpp:209-306.  The edge-axis arguments are the reference's: for
// an edge e they test the directions (0, e.z, -e.y), (-e.z, 0, -e.x) and
// (e.y, e.x, 0); the last two are not e x y_hat and e x z_hat, so a few
// cells next to a triangle that the exact test would clear are marked too.
__device__ bool tri_box_overlap(const float* __restrict__ t, V3 c, V3 h) {
  const V3 v[3] = {sub({t[0], t[1], t[2]}, c), sub({t[3], t[4], t[5]}, c), sub({t[6], t[7], t[8]}, c)};
  const V3 e[3] = {sub(v[1], v[0]), sub(v[2], v[1]), sub(v[0], v[2])};
  // box face normals
  if (min3(v[0].x, v[1].x, v[2].x) > h.x || max3(v[0].x, v[1].x, v[2].x) < -h.x) return false;
  if (min3(v[0].y, v[1].y, v[2].y) > h.y || max3(v[0].y, v[1].y, v[2].y) < -h.y) return false;
  if (min3(v[0].z, v[1].z, v[2].z) > h.z || max3(v[0].z, v[1].z, v[2].z) < -h.z) return false;
  // triangle plane: n = e0 x e1, box corner most against / along n
  const V3 n = {e[0].y * e[1].z - e[0].z * e[1].y, e[0].z * e[1].x - e[0].x * e[1].z,
                e[0].x * e[1].y - e[0].y * e[1].x};
  const float d = n.x * v[0].x + n.y * v[0].y + n.z * v[0].z;
  const V3 lo = {n.x > 0.0f ? -h.x : h.x, n.y > 0.0f ? -h.y : h.y, n.z > 0.0f ? -h.z : h.z};
  const V3 hi = {-lo.x, -lo.y, -lo.z};
  if (n.x * lo.x + n.y * lo.y + n.z * lo.z > d || n.x * hi.x + n.y * hi.y + n.z * hi.z < d) return false;
  // nine edge axes
  for (int k = 0; k < 3; ++k) {
    const float ax = fabsf(e[k].x), ay = fabsf(e[k].y), az = fabsf(e[k].z);
    if (!axis_overlaps(e[k].z, e[k].y, az, ay, v, 1, 2, h.y, h.z)) return false;
    if (!axis_overlaps(-e[k].z, e[k].x, az, ax, v, 0, 2, h.x, h.z)) return false;
    if (!axis_overlaps(e[k].y, -e[k].x, ay, ax, v, 0, 1, h.x, h.y)) return false;
  }
  return true;
}

__global__ __launch_bounds__(256) void k_mesh_surface(const float* __restrict__ tris, const int* __restrict__ box,
                                                      const unsigned long long* __restrict__ start, int ntri,
                                                      unsigned long long work, uint8_t* __restrict__ flags, int nx,
                                                      int ny, int* __restrict__ any) {
  bool hit = false;
  const unsigned long long step = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
  for (unsigned long long w = static_cast<unsigned long long>(blockIdx.x) * blockDim.x + threadIdx.x; w < work;
       w += step) {
    // the triangle whose cell range holds w: the last t with start[t] <= w
    int lo = 0, hi = ntri - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (start[mid] <= w) lo = mid;
      else hi = mid - 1;
    }
    const int* b = box + 6 * lo;
    const unsigned long long wx = static_cast<unsigned long long>(b[1] - b[0] + 1);
    const unsigned long long wy = static_cast<unsigned long long>(b[3] - b[2] + 1);
    unsigned long long r = w - start[lo];
    const int i = b[0] + static_cast<int>(r % wx);
    r /= wx;
    const int j = b[2] + static_cast<int>(r % wy);
    const int k = b[4] + static_cast<int>(r / wy);
    const V3 c = {static_cast<float>(i) + 0.5f, static_cast<float>(j) + 0.5f, static_cast<float>(k) + 0.5f};
    if (tri_box_overlap(tris + 9ll * lo, c, {0.5f, 0.5f, 0.5f})) {
      flags[(static_cast<size_t>(k) * ny + j) * nx + i] = kCellSolid;
      hit = true;
    }
  }
  if (hit) atomicOr(any, 1);
}

// exterior seeds: non-Solid cells on the six outer faces (obj_loader.cpp:330-356)
__global__ __launch_bounds__(256) void k_exterior_seed(const uint8_t* __restrict__ flags, uint8_t* __restrict__ vis,
                                                       int nx, int ny, int nz) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= nx) return;
  const size_t n = (static_cast<size_t>(z) * ny + y) * nx + x;
  const bool face = x == 0 || x == nx - 1 || y == 0 || y == ny - 1 || z == 0 || z == nz - 1;
  vis[n] = (face && flags[n] != kCellSolid) ? 1 : 0;
}

// Reach along one line of `len` cells `step` apart: a non-Solid cell joins
// the exterior when the run of non-Solid cells it lies in holds an exterior
// cell (forward then backward pass).
__device__ inline bool sweep_line(const uint8_t* __restrict__ flags, uint8_t* __restrict__ vis, size_t base,
                                  size_t step, int len) {
  bool changed = false, carry = false;
  for (int m = 0; m < len; ++m) {
    const size_t n = base + m * step;
    if (flags[n] == kCellSolid) carry = false;
    else if (vis[n]) carry = true;
    else if (carry) {
      vis[n] = 1;
      changed = true;
    }
  }
  carry = false;
  for (int m = len - 1; m >= 0; --m) {
    const size_t n = base + m * step;
    if (flags[n] == kCellSolid) carry = false;
    else if (vis[n]) carry = true;
    else if (carry) {
      vis[n] = 1;
      changed = true;
    }
  }
  return changed;
}

// axis 0: one lane per (y, z) line; 1: per (x, z); 2: per (x, y)
__global__ __launch_bounds__(256) void k_exterior_sweep(const uint8_t* __restrict__ flags, uint8_t* __restrict__ vis,
                                                        int nx, int ny, int nz, int axis, int* __restrict__ changed) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t plane = static_cast<size_t>(nx) * ny;
  bool c = false;
  if (axis == 0) {
    if (t >= static_cast<long long>(ny) * nz) return;
    c = sweep_line(flags, vis, static_cast<size_t>(t) * nx, 1, nx);
  } else if (axis == 1) {
    if (t >= static_cast<long long>(nx) * nz) return;
    const long long x = t % nx, z = t / nx;
    c = sweep_line(flags, vis, static_cast<size_t>(z) * plane + x, nx, ny);
  } else {
    if (t >= static_cast<long long>(nx) * ny) return;
    c = sweep_line(flags, vis, static_cast<size_t>(t), plane, nz);
  }
  if (c) atomicOr(changed, 1);
}

// cells the exterior never reached become Solid (obj_loader.cpp:389-399)
__global__ __launch_bounds__(256) void k_fill_interior(uint8_t* __restrict__ flags, const uint8_t* __restrict__ vis,
                                                       size_t n, int* __restrict__ any) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  bool hit = false;
  if (i < n && !vis[i]) {
    flags[i] = kCellSolid;
    hit = true;
  }
  if (hit) atomicOr(any, 1);
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 1); }
};

#define WAKE_TRY(expr)                    \
  do {                                    \
    const hipError_t e_ = (expr);         \
    if (e_ != hipSuccess) return e_;      \
  } while (0)

}  // namespace

// The clamped cell box of each triangle (obj_loader.cpp:215-227, FP32 floor
// and ceil of the vertex extremes) and the exclusive prefix of the box
// volumes.  Returns the total number of (triangle, cell) pairs.
unsigned long long mesh_cell_boxes(const float* tris, long long ntri, int nx, int ny, int nz, std::vector<int>& box,
                                   std::vector<unsigned long long>& start) {
  box.assign(6 * ntri, 0);
  start.assign(ntri + 1, 0);
  auto lo_cell = [](float v) { return static_cast<int>(std::clamp(std::floor(v), -1.0e9f, 1.0e9f)); };
  auto hi_cell = [](float v) { return static_cast<int>(std::clamp(std::ceil(v), -1.0e9f, 1.0e9f)); };
  const int dims[3] = {nx, ny, nz};
  unsigned long long total = 0;
  for (long long t = 0; t < ntri; ++t) {
    const float* v = tris + 9 * t;
    unsigned long long vol = 1;
    for (int a = 0; a < 3; ++a) {
      const float mn = std::min({v[a], v[3 + a], v[6 + a]}), mx = std::max({v[a], v[3 + a], v[6 + a]});
      const int c0 = std::max(0, lo_cell(mn)), c1 = std::min(dims[a] - 1, hi_cell(mx));
      box[6 * t + 2 * a] = c0;
      box[6 * t + 2 * a + 1] = c1;
      vol = c1 < c0 ? 0 : vol * static_cast<unsigned long long>(c1 - c0 + 1);
    }
    start[t] = total;
    total += vol;
  }
  start[ntri] = total;
  return total;
}

hipError_t launch_voxelize_mesh(const float* tris, long long ntri, bool fill_interior, uint8_t* d_flags, int nx,
                                int ny, int nz, hipStream_t s, bool* any_set) {
  *any_set = false;
  std::vector<int> box;
  std::vector<unsigned long long> start;
  const unsigned long long work = ntri > 0 ? mesh_cell_boxes(tris, ntri, nx, ny, nz, box, start) : 0;
  const size_t N = static_cast<size_t>(nx) * ny * nz;
  DevBuf d_any, d_tris, d_box, d_start, d_vis, d_changed;
  WAKE_TRY(d_any.alloc(sizeof(int)));
  WAKE_TRY(hipMemsetAsync(d_any.p, 0, sizeof(int), s));
  if (work > 0) {
    WAKE_TRY(d_tris.alloc(9 * ntri * sizeof(float)));
    WAKE_TRY(d_box.alloc(box.size() * sizeof(int)));
    WAKE_TRY(d_start.alloc(start.size() * sizeof(unsigned long long)));
    WAKE_TRY(hipMemcpyAsync(d_tris.p, tris, 9 * ntri * sizeof(float), hipMemcpyHostToDevice, s));
    WAKE_TRY(hipMemcpyAsync(d_box.p, box.data(), box.size() * sizeof(int), hipMemcpyHostToDevice, s));
    WAKE_TRY(hipMemcpyAsync(d_start.p, start.data(), start.size() * sizeof(unsigned long long), hipMemcpyHostToDevice,
                            s));
    const unsigned long long blocks = std::min<unsigned long long>((work + 255) / 256, 65536ull);
    k_mesh_surface<<<static_cast<unsigned>(blocks), 256, 0, s>>>(
        static_cast<const float*>(d_tris.p), static_cast<const int*>(d_box.p),
        static_cast<const unsigned long long*>(d_start.p), static_cast<int>(ntri), work, d_flags, nx, ny,
        static_cast<int*>(d_any.p));
    WAKE_TRY(hipGetLastError());
  }
  if (fill_interior && N > 0) {
    WAKE_TRY(d_vis.alloc(N));
    WAKE_TRY(d_changed.alloc(sizeof(int)));
    k_exterior_seed<<<dim3((nx + 255) / 256, ny, nz), 256, 0, s>>>(d_flags, static_cast<uint8_t*>(d_vis.p), nx, ny,
                                                                    nz);
    WAKE_TRY(hipGetLastError());
    const long long lines[3] = {static_cast<long long>(ny) * nz, static_cast<long long>(nx) * nz,
                                static_cast<long long>(nx) * ny};
    // every round that changes something adds at least one cell, so at most
    // N + 1 rounds run (a handful for ordinary shapes)
    for (size_t round = 0; round <= N; ++round) {
      WAKE_TRY(hipMemsetAsync(d_changed.p, 0, sizeof(int), s));
      for (int axis = 0; axis < 3; ++axis) {
        k_exterior_sweep<<<static_cast<unsigned>((lines[axis] + 255) / 256), 256, 0, s>>>(
            d_flags, static_cast<uint8_t*>(d_vis.p), nx, ny, nz, axis, static_cast<int*>(d_changed.p));
        WAKE_TRY(hipGetLastError());
      }
      int changed = 0;
      WAKE_TRY(hipMemcpyAsync(&changed, d_changed.p, sizeof(int), hipMemcpyDeviceToHost, s));
      WAKE_TRY(hipStreamSynchronize(s));
      if (!changed) break;
    }
    k_fill_interior<<<static_cast<unsigned>((N + 255) / 256), 256, 0, s>>>(
        d_flags, static_cast<const uint8_t*>(d_vis.p), N, static_cast<int*>(d_any.p));
    WAKE_TRY(hipGetLastError());
  }
  int any = 0;
  WAKE_TRY(hipMemcpyAsync(&any, d_any.p, sizeof(int), hipMemcpyDeviceToHost, s));
  WAKE_TRY(hipStreamSynchronize(s));
  *any_set = any != 0;
  return hipSuccess;
}

// ------------------------------------------------------------ BFL links
// lattice_grid_3d.cpp:217-290: a link from every Fluid node x along each
// c_q (q >= 1) whose neighbour -- x periodic, y and z in range or wrapped on
// a periodic axis -- is Solid; the wall fraction delta is the first root in
// [0, 1] of |x - c + t c_q| = R (FP32, correctly rounded square root and
// division, no contraction), else 1/2, clamped to [0.01, 0.99].  One
// work-group per (y, z) row: a counting pass, a host prefix over rows, then
// an emitting pass whose in-row prefix keeps the z -> y -> x -> q order.
namespace {

struct LinkArgs {
  int nx, ny, nz, per_y, per_z;
  float cx, cy, cz, r;
};

template <class L>
__device__ inline bool links_to_solid(const uint8_t* __restrict__ flags, const LinkArgs& a, int x, int y, int z,
                                      int q) {
  const int xn = (x + L::cx[q] + a.nx) % a.nx;
  int yn = y + L::cy[q], zn = z + L::cz[q];
  if (a.per_y) yn = (yn + a.ny) % a.ny;
  if (a.per_z) zn = (zn + a.nz) % a.nz;
  if (yn < 0 || yn >= a.ny || zn < 0 || zn >= a.nz) return false;
  return flags[(static_cast<size_t>(zn) * a.ny + yn) * a.nx + xn] == kCellSolid;
}

template <class L>
__device__ inline int node_link_count(const uint8_t* __restrict__ flags, const LinkArgs& a, int x, int y, int z) {
  if (flags[(static_cast<size_t>(z) * a.ny + y) * a.nx + x] != 0) return 0;  // Fluid only
  int n = 0;
  for (int q = 1; q < L::Q; ++q) n += links_to_solid<L>(flags, a, x, y, z, q);
  return n;
}

template <class L>
__device__ inline float wall_fraction(const LinkArgs& a, int x, int y, int z, int q) {
  const float p0x = static_cast<float>(x) - a.cx, p0y = static_cast<float>(y) - a.cy,
              p0z = static_cast<float>(z) - a.cz;
  const float dx = static_cast<float>(L::cx[q]), dy = static_cast<float>(L::cy[q]), dz = static_cast<float>(L::cz[q]);
  const float qa = dx * dx + dy * dy + dz * dz;
  const float qb = 2.0f * (p0x * dx + p0y * dy + p0z * dz);
  const float qc = (p0x * p0x + p0y * p0y + p0z * p0z) - a.r * a.r;
  const float disc = qb * qb - 4.0f * qa * qc;
  float delta = 0.5f;
  if (disc >= 0.0f) {
    // sqrtf and / are correctly rounded here (-fhip-fp32-correctly-rounded-
    // divide-sqrt); __fsqrt_rn would be the native, not correctly rounded, root
    const float root = sqrtf(disc);
    const float t1 = (-qb - root) / (2.0f * qa);
    const float t2 = (-qb + root) / (2.0f * qa);
    if (t1 >= 0.0f && t1 <= 1.0f) delta = t1;
    else if (t2 >= 0.0f && t2 <= 1.0f) delta = t2;
  }
  return delta < 0.01f ? 0.01f : (delta > 0.99f ? 0.99f : delta);
}

template <class L>
__global__ __launch_bounds__(256) void k_link_rows(const uint8_t* __restrict__ flags, const LinkArgs a,
                                                   int* __restrict__ row_count) {
  __shared__ int part[256];
  const int row = blockIdx.x, y = row % a.ny, z = row / a.ny;
  int n = 0;
  for (int x = threadIdx.x; x < a.nx; x += blockDim.x) n += node_link_count<L>(flags, a, x, y, z);
  part[threadIdx.x] = n;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) row_count[row] = part[0];
}

template <class L>
__global__ __launch_bounds__(256) void k_link_emit(const uint8_t* __restrict__ flags, const LinkArgs a,
                                                   const long long* __restrict__ row_start, int* __restrict__ node,
                                                   int* __restrict__ qdir, float* __restrict__ delta) {
  __shared__ int scan[256];
  const int row = blockIdx.x, y = row % a.ny, z = row / a.ny;
  long long base = row_start[row];
  for (int x0 = 0; x0 < a.nx; x0 += blockDim.x) {
    const int x = x0 + static_cast<int>(threadIdx.x);
    const int n = x < a.nx ? node_link_count<L>(flags, a, x, y, z) : 0;
    scan[threadIdx.x] = n;
    __syncthreads();
    for (int off = 1; off < static_cast<int>(blockDim.x); off <<= 1) {
      const int v = static_cast<int>(threadIdx.x) >= off ? scan[threadIdx.x - off] : 0;
      __syncthreads();
      scan[threadIdx.x] += v;
      __syncthreads();
    }
    const int total = scan[blockDim.x - 1];
    if (n) {
      long long o = base + scan[threadIdx.x] - n;
      const int idx = (z * a.ny + y) * a.nx + x;
      for (int q = 1; q < L::Q; ++q)
        if (links_to_solid<L>(flags, a, x, y, z, q)) {
          node[o] = idx;
          qdir[o] = q;
          delta[o] = wall_fraction<L>(a, x, y, z, q);
          ++o;
        }
    }
    base += total;
    __syncthreads();
  }
}

template <class L>
hipError_t build_links(const uint8_t* d_flags, const LinkArgs& a, hipStream_t s, std::vector<int>& node,
                       std::vector<int>& qdir, std::vector<float>& delta) {
  const int rows = a.ny * a.nz;
  DevBuf d_count, d_start, d_node, d_q, d_delta;
  WAKE_TRY(d_count.alloc(rows * sizeof(int)));
  k_link_rows<L><<<rows, 256, 0, s>>>(d_flags, a, static_cast<int*>(d_count.p));
  WAKE_TRY(hipGetLastError());
  std::vector<int> count(rows);
  WAKE_TRY(hipMemcpyAsync(count.data(), d_count.p, rows * sizeof(int), hipMemcpyDeviceToHost, s));
  WAKE_TRY(hipStreamSynchronize(s));
  std::vector<long long> start(rows);
  long long total = 0;
  for (int r = 0; r < rows; ++r) {
    start[r] = total;
    total += count[r];
  }
  node.resize(total);
  qdir.resize(total);
  delta.resize(total);
  if (total == 0) return hipSuccess;
  WAKE_TRY(d_start.alloc(rows * sizeof(long long)));
  WAKE_TRY(d_node.alloc(total * sizeof(int)));
  WAKE_TRY(d_q.alloc(total * sizeof(int)));
  WAKE_TRY(d_delta.alloc(total * sizeof(float)));
  WAKE_TRY(hipMemcpyAsync(d_start.p, start.data(), rows * sizeof(long long), hipMemcpyHostToDevice, s));
  k_link_emit<L><<<rows, 256, 0, s>>>(d_flags, a, static_cast<const long long*>(d_start.p),
                                      static_cast<int*>(d_node.p), static_cast<int*>(d_q.p),
                                      static_cast<float*>(d_delta.p));
  WAKE_TRY(hipGetLastError());
  WAKE_TRY(hipMemcpyAsync(node.data(), d_node.p, total * sizeof(int), hipMemcpyDeviceToHost, s));
  WAKE_TRY(hipMemcpyAsync(qdir.data(), d_q.p, total * sizeof(int), hipMemcpyDeviceToHost, s));
  WAKE_TRY(hipMemcpyAsync(delta.data(), d_delta.p, total * sizeof(float), hipMemcpyDeviceToHost, s));
  return hipStreamSynchronize(s);
}

}  // namespace

hipError_t launch_build_sphere_links(int lattice, const uint8_t* d_flags, int nx, int ny, int nz, bool per_y,
                                     bool per_z, float cx, float cy, float cz, float r, hipStream_t s,
                                     std::vector<int>& node, std::vector<int>& qdir, std::vector<float>& delta) {
  const LinkArgs a{nx, ny, nz, per_y ? 1 : 0, per_z ? 1 : 0, cx, cy, cz, r};
  switch (lattice) {
    case 9: return build_links<D2Q9>(d_flags, a, s, node, qdir, delta);
    case 19: return build_links<D3Q19>(d_flags, a, s, node, qdir, delta);
    case 27: return build_links<D3Q27>(d_flags, a, s, node, qdir, delta);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace wake_hip
