// batch.hip -- kernels of the batched multi-domain step (gfx950).
//
// The per-node work is the device-resident step's (form_s_at + collide, the
// same device functions the single-domain kernels use), so a domain gives
// the same populations stepped alone or in a batch; only the indexing
// differs: a work-group finds its slot, row and x segment from the tile it
// runs (all uniform), and a lane its x within the segment.
#include "batch.hpp"
#include "thermal_impl.hpp"

namespace wake_hip {

// Same per-instantiation occupancy floor as the single-domain step kernel
// (StepWaves: the D3Q27 kernels with 16-bit storage).
template <class L, class St, int COLL, bool LES, bool THERMAL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(StepWaves<L, St>::value))) void k_step_batch(
    const BatchSlot* __restrict__ slots,
                                                    const int* __restrict__ tile_slot, int ntiles, int i) {
  using T = typename St::T;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const BatchSlot& sl = slots[tile_slot[t]];
    const unsigned long long cur = sl.step + static_cast<unsigned long long>(i);
    if (cur >= sl.stop_at) continue;
    const StepArgs& a = sl.a;
    const int local = t - sl.tile0;
    const int row = local / sl.tiles_per_row;
    const int x = (local - row * sl.tiles_per_row) * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
    if (x >= a.nx) continue;
    const int y = row % a.ny;
    const int z = row / a.ny;
    const int n = row * a.nx + x;
    const int par = (sl.parity + i) & 1;
    const T* __restrict__ in = static_cast<const T*>(sl.buf[par]);
    T* __restrict__ out = static_cast<T*>(sl.buf[par ^ 1]);
    if constexpr (THERMAL) {
      thermal_node<St, COLL>(a, in, out, sl.gbuf[par], sl.gbuf[par ^ 1], x, y, n, cur != sl.s_at, cur);
      continue;
    }
    const uint8_t c = a.cls[n];
    float f[L::Q];
    if (cur == sl.s_at) {
      // first step from a materialised S state: collide only
      static_for<L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = St::template ld<L, q>(in, q * a.stride + n);
      });
    } else {
      form_s_at<L, St>(a, in, x, y, z, n, c, true, cur, f);
    }
    if (!(c & kClsSolid)) {
      Macro m;
      if constexpr (LES) m = collide_les<L, COLL, St::kFast>(f, a.tau, a.cs, a.gx, a.gy, a.gz);
      else m = collide<L, COLL, false, St::kFast>(f, a.tau, a.gx, a.gy, a.gz);
      flag_nonfinite(a, m.rho);
    }
    if constexpr (St::kBuf) {
      const auto r = St::rsrc(out, static_cast<unsigned>(L::Q * a.stride * sizeof(T)));
      static_for<L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        St::template stb<L, q>(r, static_cast<unsigned>(x) * sizeof(T),
                               static_cast<unsigned>(q * a.stride + static_cast<long long>(row) * a.nx) * sizeof(T),
                               f[q]);
      });
    } else {
      static_for<L::Q>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        St::template st<L, q>(out, q * a.stride + n, f[q]);
      });
    }
  }
}

// Momentum exchange on S_{cur+1} = gather of the buffer step i just wrote,
// one work-group per member: the lanes take the member's nodes holding links
// (a device list in node order) in turn, form S there and add the links;
// lane sums in a fixed order, then a fixed tree over the work-group.  A
// member with forces but no link records zeros.
template <class L, class St>
__global__ __launch_bounds__(256) void k_mem_batch(const BatchSlot* __restrict__ slots, int i, int chunk,
                                                   double* __restrict__ hist) {
  const BatchSlot& sl = slots[blockIdx.x];
  const unsigned long long cur = sl.step + static_cast<unsigned long long>(i);
  if (!sl.mem_on || cur >= sl.stop_at) return;
  const StepArgs& a = sl.a;
  const int par = ((sl.parity + i) & 1) ^ 1;
  const auto* buf = static_cast<const typename St::T*>(sl.buf[par]);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int i = threadIdx.x; i < sl.mem_count; i += blockDim.x) {
    const int n = sl.mem_nodes[i];
    const int x = n % a.nx;
    const int y = (n / a.nx) % a.ny;
    const int z = n / (a.nx * a.ny);
    float f[L::Q];
    form_s_at<L, St, false>(a, buf, x, y, z, n, a.cls[n], false, cur + 1, f);
    mem_node<L>(a, x, y, z, n, f, a.mem_c[0], a.mem_c[1], a.mem_c[2], a.mem_rmax, s0, s1, s2);
  }
  block_sum3(s0, s1, s2);
  if (threadIdx.x == 0) {
    double* o = hist + (static_cast<long long>(blockIdx.x) * chunk + i) * 3;
    o[0] = s0;
    o[1] = s1;
    o[2] = s2;
  }
}

__global__ void k_advance_batch(BatchSlot* __restrict__ slots, int nslots, int k) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  BatchSlot& sl = slots[s];
  const unsigned long long left = sl.stop_at > sl.step ? sl.stop_at - sl.step : 0ull;
  const unsigned long long taken = left < static_cast<unsigned long long>(k) ? left : static_cast<unsigned long long>(k);
  sl.step += taken;
  sl.parity = (sl.parity + static_cast<int>(taken & 1ull)) & 1;
}

namespace {

template <class St>
struct StTag {
  using type = St;
};

// the 16-bit storages with descriptor addressing when `buf`
template <class F>
hipError_t with_step_storage_b(int storage, bool buf, F&& f) {
  switch (storage) {
    case kStoreF32: return f(StTag<StF32>{});
    case kStoreF16: return buf ? f(StTag<StF16Buf>{}) : f(StTag<StF16>{});
    case kStoreBF16: return buf ? f(StTag<StBF16Buf>{}) : f(StTag<StBF16>{});
    default: return hipErrorInvalidValue;
  }
}

template <class F>
hipError_t with_storage_b(int storage, F&& f) {
  switch (storage) {
    case kStoreF32: return f(StTag<StF32>{});
    case kStoreF16: return f(StTag<StF16>{});
    case kStoreBF16: return f(StTag<StBF16>{});
    default: return hipErrorInvalidValue;
  }
}

template <class L>
hipError_t step_batch_l(int storage, int collision, bool les, bool thermal, bool buf, const BatchSlot* slots,
                        const int* tile_slot, int ntiles, int i, int block, int grid_x, hipStream_t s) {
  return with_step_storage_b(storage, buf && !thermal, [&](auto tag) {
    using St = typename decltype(tag)::type;
    auto go = [&](auto coll) {
      constexpr int C = decltype(coll)::value;
      if constexpr (L::Q == 9) {
        if (thermal) {
          k_step_batch<L, St, C, false, true><<<grid_x, block, 0, s>>>(slots, tile_slot, ntiles, i);
          return hipGetLastError();
        }
      } else {
        if (thermal) return hipErrorInvalidValue;
      }
      if (les) k_step_batch<L, St, C, true, false><<<grid_x, block, 0, s>>>(slots, tile_slot, ntiles, i);
      else k_step_batch<L, St, C, false, false><<<grid_x, block, 0, s>>>(slots, tile_slot, ntiles, i);
      return hipGetLastError();
    };
    if (collision == kBGK) return go(std::integral_constant<int, kBGK>{});
    if constexpr (L::Q == 27) {
      return go(std::integral_constant<int, kTRT>{});
    } else {
      if (collision == kTRT) return go(std::integral_constant<int, kTRT>{});
      return go(std::integral_constant<int, kMRT>{});
    }
  });
}

template <class L>
hipError_t mem_batch_l(int storage, const BatchSlot* slots, int nslots, int i, int chunk, double* hist,
                       hipStream_t s) {
  return with_storage_b(storage, [&](auto tag) {
    using St = typename decltype(tag)::type;
    k_mem_batch<L, St><<<nslots, 256, 0, s>>>(slots, i, chunk, hist);
    return hipGetLastError();
  });
}

}  // namespace

hipError_t launch_step_batch(int lattice, int storage, int collision, bool les, bool thermal, bool buf,
                             const BatchSlot* slots, const int* tile_slot, int ntiles, int i, int block, int grid_x,
                             hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
  if (grid_x <= 0 || grid_x > ntiles) grid_x = ntiles;
  switch (lattice) {
    case 9:
      return step_batch_l<D2Q9>(storage, collision, les, thermal, buf, slots, tile_slot, ntiles, i, block, grid_x,
                                s);
    case 19:
      return step_batch_l<D3Q19>(storage, collision, les, thermal, buf, slots, tile_slot, ntiles, i, block, grid_x,
                                s);
    case 27:
      return step_batch_l<D3Q27>(storage, collision, les, thermal, buf, slots, tile_slot, ntiles, i, block, grid_x,
                                s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_mem_batch(int lattice, int storage, const BatchSlot* slots, int nslots, int i, int chunk,
                            double* hist, hipStream_t s) {
  if (nslots <= 0) return hipSuccess;
  switch (lattice) {
    case 9: return mem_batch_l<D2Q9>(storage, slots, nslots, i, chunk, hist, s);
    case 19: return mem_batch_l<D3Q19>(storage, slots, nslots, i, chunk, hist, s);
    case 27: return mem_batch_l<D3Q27>(storage, slots, nslots, i, chunk, hist, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_advance_batch(BatchSlot* slots, int nslots, int k, hipStream_t s) {
  if (nslots <= 0) return hipSuccess;
  k_advance_batch<<<(nslots + 255) / 256, 256, 0, s>>>(slots, nslots, k);
  return hipGetLastError();
}

}  // namespace wake_hip
