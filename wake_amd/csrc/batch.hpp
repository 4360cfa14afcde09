// batch.hpp -- many independent domains stepped by one launch per step.
//
// Every domain keeps its own buffers, cell types, classes and boundary
// configuration (a Domain); a batch only adds a device table of slots and a
// tile map.  A tile is a segment of `block` consecutive x of one (y, z) row
// of one domain, so a work-group never straddles two domains or two rows:
// the per-domain arguments and the row coordinates are work-group uniform
// (scalar loads and scalar arithmetic), as in the single-domain step.
// The kernels derive everything that changes from step to step -- which
// buffer is read, the step index of the inlet noise, whether the domain has
// already reached its last step -- from the slot and the launch's position in
// the chunk, so a chunk of steps can be captured once as a graph and replayed.
#pragma once

#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace wake_hip {

constexpr unsigned long long kNoStep = ~0ull;

struct BatchSlot {
  StepArgs a;                 // the domain's step arguments (its own device arrays)
  void* buf[2];               // population ping-pong buffers
  float* gbuf[2];             // temperature populations (thermal domains), same ping-pong
  unsigned long long step;    // steps the domain has done at the start of the chunk
  unsigned long long stop_at; // the domain takes no step that would pass this count
  unsigned long long s_at;    // step count at which the state is a materialised S, or kNoStep
  int parity;                 // buffer holding the state at `step`
  int nodes;                  // nx * ny * nz
  int tile0, ntiles;          // tiles [tile0, tile0 + ntiles) belong to this slot
  int tiles_per_row;          // ceil(nx / block)
  // momentum exchange (a.mem_c, a.mem_rmax): the member's nodes holding links
  const int* mem_nodes;
  int mem_count;
  int mem_on;                 // forces recorded for this member
};

// One launch = step i of the chunk for every slot.  grid_x work-groups of
// `block` lanes walk the tiles with a stride of grid_x (grid_x = ntiles: one
// work-group per tile).
// buf: every member's population buffer is under 4 GiB, so the 16-bit
// storages may use descriptor addressing.
hipError_t launch_step_batch(int lattice, int storage, int collision, bool les, bool thermal, bool buf,
                             const BatchSlot* slots, const int* tile_slot, int ntiles, int i, int block, int grid_x,
                             hipStream_t s);
// Forces on S after step i of the chunk, one work-group per slot:
// hist[(slot * chunk + i) * 3 + k].  Slots past their stop step are skipped.
hipError_t launch_mem_batch(int lattice, int storage, const BatchSlot* slots, int nslots, int i, int chunk,
                            double* hist, hipStream_t s);
// End of a chunk of `k` launches: step += steps taken, parity updated.
hipError_t launch_advance_batch(BatchSlot* slots, int nslots, int k, hipStream_t s);

}  // namespace wake_hip
