// step_launch.hpp -- per-lattice launchers of the step kernels.  Each lattice
// is instantiated in its own translation unit (step_d2q9.hip, step_d3q19.hip,
// step_d3q27.hip) so the kernel variants compile in parallel.
#pragma once

#include "step_impl.hpp"

namespace wake_hip {
namespace detail {

inline dim3 grid_of(const StepArgs& a, int bx) { return dim3((a.nx + bx - 1) / bx, a.ny, a.nz); }

// Work-group width along x: the one of 256 / 128 / 64 that leaves the
// fewest idle lanes per row (the larger on a tie), e.g. 128 for nx = 384.
inline int pick_block(int nx, int block_x) {
  if (block_x > 0) return block_x;
  int best = 256, best_pad = (nx + 255) / 256 * 256;
  for (int b : {128, 64}) {
    const int pad = (nx + b - 1) / b * b;
    if (pad < best_pad) {
      best = b;
      best_pad = pad;
    }
  }
  return best;
}

template <class F>
hipError_t with_storage(int storage, F&& f) {
  switch (storage) {
    case kStoreF32: return f(StF32{});
    case kStoreF16: return f(StF16{});
    case kStoreBF16: return f(StBF16{});
    default: return hipErrorInvalidValue;
  }
}

// The step kernel's storage types: the 16-bit ones with descriptor
// addressing when `buf` (the buffer fits a 32-bit byte offset).
#ifndef WAKE_BUF_F32
#define WAKE_BUF_F32 0
#endif
template <class F>
hipError_t with_step_storage(int storage, bool buf, F&& f) {
  switch (storage) {
#if WAKE_BUF_F32
    case kStoreF32: return buf ? f(StF32Buf{}) : f(StF32{});
#else
    case kStoreF32: return f(StF32{});
#endif
    case kStoreF16: return buf ? f(StF16Buf{}) : f(StF16{});
    case kStoreBF16: return buf ? f(StBF16Buf{}) : f(StBF16{});
    default: return hipErrorInvalidValue;
  }
}

template <class L, class St, int COLL>
hipError_t step_dispatch(bool les, bool pull, const StepArgs& a, const void* in, void* out, int bx, hipStream_t s) {
  using T = typename St::T;
  const T* i = static_cast<const T*>(in);
  T* o = static_cast<T*>(out);
  const dim3 g((a.nx + bx - 1) / bx, a.ny, (a.nz + WAKE_ZK - 1) / WAKE_ZK);
  if (les) {
    if (pull) k_step<L, St, COLL, true, true><<<g, bx, 0, s>>>(a, i, o);
    else k_step<L, St, COLL, true, false><<<g, bx, 0, s>>>(a, i, o);
  } else {
    if (pull) k_step<L, St, COLL, false, true><<<g, bx, 0, s>>>(a, i, o);
    else k_step<L, St, COLL, false, false><<<g, bx, 0, s>>>(a, i, o);
  }
  return hipGetLastError();
}

template <class L>
struct Launch {
  static hipError_t build_class(const StepArgs& a, uint8_t* cls, uint32_t* bmask, hipStream_t s) {
    const int bx = pick_block(a.nx, 0);
    k_build_class<L><<<grid_of(a, bx), bx, 0, s>>>(a, cls, bmask);
    return hipGetLastError();
  }

  static hipError_t step(int storage, int collision, bool les, bool pull, const StepArgs& a, const void* a_in,
                         void* a_out, int block_x, hipStream_t s) {
    const int bx = pick_block(a.nx, block_x);
    // populations in a buffer under 4 GiB: descriptor addressing
    const unsigned long long elem = storage == kStoreF32 ? 4ull : 2ull;
    const bool buf = !a.global_addr && static_cast<unsigned long long>(L::Q) * a.stride * elem < (1ull << 32);
    return with_step_storage(storage, buf, [&](auto st) {
      using St = decltype(st);
      if (collision == kBGK) return step_dispatch<L, St, kBGK>(les, pull, a, a_in, a_out, bx, s);
      if constexpr (L::Q == 27) {
        return step_dispatch<L, St, kTRT>(les, pull, a, a_in, a_out, bx, s);
      } else {
        if (collision == kTRT) return step_dispatch<L, St, kTRT>(les, pull, a, a_in, a_out, bx, s);
        return step_dispatch<L, St, kMRT>(les, pull, a, a_in, a_out, bx, s);
      }
    });
  }

  static hipError_t step_tb(int storage, int collision, bool les, const StepArgs& a, const void* a_in, void* a_out,
                            hipStream_t s) {
    if constexpr (L::Q != 9 || L::D != 2) {
      return hipErrorInvalidValue;
    } else {
      const dim3 g((a.nx + kTbTile - 1) / kTbTile, (a.ny + kTbTile - 1) / kTbTile, 1);
      return with_storage(storage, [&](auto st) {
        using St = decltype(st);
        using T = typename St::T;
        const T* i = static_cast<const T*>(a_in);
        T* o = static_cast<T*>(a_out);
        auto go = [&](auto cc) {
          constexpr int C = decltype(cc)::value;
          if (les) k_step_tb2d<St, C, true><<<g, kTbThreads, 0, s>>>(a, i, o);
          else k_step_tb2d<St, C, false><<<g, kTbThreads, 0, s>>>(a, i, o);
        };
        if (collision == kBGK) go(std::integral_constant<int, kBGK>{});
        else if (collision == kTRT) go(std::integral_constant<int, kTRT>{});
        else go(std::integral_constant<int, kMRT>{});
        return hipGetLastError();
      });
    }
  }

  static hipError_t mem_mark_rows(const StepArgs& a, int* row_has, hipStream_t s) {
    const int bx = pick_block(a.nx, 0);
    k_mem_mark_rows<L><<<grid_of(a, bx), bx, 0, s>>>(a, row_has);
    return hipGetLastError();
  }

  static hipError_t mem_row_nodes(const StepArgs& a, const int* row_list, int nrows, int* counts,
                                  const int* offsets, int* nodes, hipStream_t s) {
    if (nrows <= 0) return hipSuccess;
    const int g = nrows < 65535 ? nrows : 65535;
    if (nodes) k_mem_row_nodes<L, true><<<g, 64, 0, s>>>(a, row_list, nrows, counts, offsets, nodes);
    else k_mem_row_nodes<L, false><<<g, 64, 0, s>>>(a, row_list, nrows, counts, offsets, nodes);
    return hipGetLastError();
  }

  static hipError_t mem_rows(int storage, bool from_s, const StepArgs& a, const void* buf, const int* row_list,
                             int nrows, hipStream_t s) {
    if (nrows <= 0) return hipSuccess;
    const int bx = pick_block(a.nx, 0);
    const dim3 g((a.nx + bx - 1) / bx, nrows < 65535 ? nrows : 65535, 1);
    return with_storage(storage, [&](auto st) {
      using St = decltype(st);
      using T = typename St::T;
      if (from_s) k_mem_rows<L, St, true><<<g, bx, 0, s>>>(a, static_cast<const T*>(buf), row_list, nrows);
      else k_mem_rows<L, St, false><<<g, bx, 0, s>>>(a, static_cast<const T*>(buf), row_list, nrows);
      return hipGetLastError();
    });
  }

  static hipError_t prev_from_s(int storage, const StepArgs& a, const void* s_state, hipStream_t s) {
    const int rows = a.ny * a.nz;
    return with_storage(storage, [&](auto st) {
      using St = decltype(st);
      using T = typename St::T;
      k_prev_from_s<L, St><<<(rows + 255) / 256, 256, 0, s>>>(a, static_cast<const T*>(s_state));
      return hipGetLastError();
    });
  }

  static hipError_t materialize(int storage, const StepArgs& a, const void* a_in, void* s_state, hipStream_t s) {
    const int bx = pick_block(a.nx, 0);
    return with_storage(storage, [&](auto st) {
      using St = decltype(st);
      using T = typename St::T;
      k_materialize<L, St><<<grid_of(a, bx), bx, 0, s>>>(a, static_cast<const T*>(a_in), static_cast<T*>(s_state));
      return hipGetLastError();
    });
  }

  static hipError_t readback(int storage, bool from_s, const StepArgs& a, const void* buf, hipStream_t s) {
    const int bx = pick_block(a.nx, 0);
    return with_storage(storage, [&](auto st) {
      using St = decltype(st);
      using T = typename St::T;
      if (from_s) k_readback<L, St, true><<<grid_of(a, bx), bx, 0, s>>>(a, static_cast<const T*>(buf));
      else k_readback<L, St, false><<<grid_of(a, bx), bx, 0, s>>>(a, static_cast<const T*>(buf));
      return hipGetLastError();
    });
  }

};

// Per-lattice entry points, defined in step_d2q9.hip / step_d3q19.hip / step_d3q27.hip.
#define WAKE_DECLARE_LATTICE(NAME)                                                                             \
  hipError_t build_class_##NAME(const StepArgs& a, uint8_t* cls, uint32_t* bmask, hipStream_t s);              \
  hipError_t step_##NAME(int storage, int collision, bool les, bool pull, const StepArgs& a, const void* a_in, \
                         void* a_out, int block_x, hipStream_t s);                                             \
  hipError_t materialize_##NAME(int storage, const StepArgs& a, const void* a_in, void* s_state, hipStream_t s); \
  hipError_t prev_from_s_##NAME(int storage, const StepArgs& a, const void* s_state, hipStream_t s);             \
  hipError_t readback_##NAME(int storage, bool from_s, const StepArgs& a, const void* buf, hipStream_t s);      \
  hipError_t mem_mark_rows_##NAME(const StepArgs& a, int* row_has, hipStream_t s);                             \
  hipError_t step_tb_##NAME(int storage, int collision, bool les, const StepArgs& a, const void* a_in,          \
                            void* a_out, hipStream_t s);                                                       \
  hipError_t mem_row_nodes_##NAME(const StepArgs& a, const int* row_list, int nrows, int* counts,               \
                                  const int* offsets, int* nodes, hipStream_t s);                              \
  hipError_t mem_rows_##NAME(int storage, bool from_s, const StepArgs& a, const void* buf, const int* row_list, \
                             int nrows, hipStream_t s);

#define WAKE_DEFINE_LATTICE(NAME, LAT)                                                                          \
  hipError_t build_class_##NAME(const StepArgs& a, uint8_t* cls, uint32_t* bmask, hipStream_t s) {              \
    return Launch<LAT>::build_class(a, cls, bmask, s);                                                          \
  }                                                                                                             \
  hipError_t step_##NAME(int storage, int collision, bool les, bool pull, const StepArgs& a, const void* a_in,  \
                         void* a_out, int block_x, hipStream_t s) {                                             \
    return Launch<LAT>::step(storage, collision, les, pull, a, a_in, a_out, block_x, s);                        \
  }                                                                                                             \
  hipError_t materialize_##NAME(int storage, const StepArgs& a, const void* a_in, void* s_state, hipStream_t s) { \
    return Launch<LAT>::materialize(storage, a, a_in, s_state, s);                                              \
  }                                                                                                             \
  hipError_t prev_from_s_##NAME(int storage, const StepArgs& a, const void* s_state, hipStream_t s) {            \
    return Launch<LAT>::prev_from_s(storage, a, s_state, s);                                                    \
  }                                                                                                             \
  hipError_t readback_##NAME(int storage, bool from_s, const StepArgs& a, const void* buf, hipStream_t s) {      \
    return Launch<LAT>::readback(storage, from_s, a, buf, s);                                                   \
  }                                                                                                             \
  hipError_t mem_mark_rows_##NAME(const StepArgs& a, int* row_has, hipStream_t s) {                            \
    return Launch<LAT>::mem_mark_rows(a, row_has, s);                                                           \
  }                                                                                                             \
  hipError_t step_tb_##NAME(int storage, int collision, bool les, const StepArgs& a, const void* a_in,          \
                            void* a_out, hipStream_t s) {                                                       \
    return Launch<LAT>::step_tb(storage, collision, les, a, a_in, a_out, s);                                    \
  }                                                                                                             \
  hipError_t mem_row_nodes_##NAME(const StepArgs& a, const int* row_list, int nrows, int* counts,               \
                                  const int* offsets, int* nodes, hipStream_t s) {                             \
    return Launch<LAT>::mem_row_nodes(a, row_list, nrows, counts, offsets, nodes, s);                           \
  }                                                                                                             \
  hipError_t mem_rows_##NAME(int storage, bool from_s, const StepArgs& a, const void* buf, const int* row_list, \
                             int nrows, hipStream_t s) {                                                        \
    return Launch<LAT>::mem_rows(storage, from_s, a, buf, row_list, nrows, s);                                  \
  }

WAKE_DECLARE_LATTICE(d2q9)
WAKE_DECLARE_LATTICE(d3q19)
WAKE_DECLARE_LATTICE(d3q27)

}  // namespace detail
}  // namespace wake_hip
