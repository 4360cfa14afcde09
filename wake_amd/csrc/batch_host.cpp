// batch_host.cpp -- host side of the batched multi-domain step, and its C ABI.
//
// A Batch steps many Domains of one lattice / storage / collision kind with
// one kernel launch per step.  Steps run in chunks of `chunk` launches; a
// full chunk is captured once as a HIP graph (step launches, per-step force
// reductions, the slot-advance kernel) and replayed while the membership and
// launch shape stay the same.  Each slot carries its own stop step, so every
// domain takes exactly the steps it asked for even when a chunk would pass
// them.  The autotuner times real steps under each candidate launch shape;
// results do not depend on the shape (a node's arithmetic is the same in any
// tile), so tuning costs no accuracy, only the steps it spends.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "batch.hpp"
#include "capi_internal.hpp"
#include "wake_hip.h"

namespace wake_hip {

struct LaunchShape {
  int block = 256;
  int grid_factor = 0;  // 0: one work-group per tile; else factor x compute units, persistent
};

class Batch {
 public:
  Batch(int lattice, int storage, int collision, bool les, bool thermal, int chunk)
      : lattice_(lattice), storage_(storage), collision_(collision), les_(les), thermal_(thermal), chunk_(chunk) {
    if (lattice != 9 && lattice != 19 && lattice != 27) throw HipError("batch lattice must be 9, 19 or 27");
    if (thermal && lattice != 9) throw HipError("thermal batches are D2Q9 + D2Q5");
    if (chunk < 1 || chunk > 4096) throw HipError("batch chunk must be in [1, 4096]");
    try {
      check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
      int dev = 0;
      check(hipGetDevice(&dev), "hipGetDevice");
      check(hipDeviceGetAttribute(&cus_, hipDeviceAttributeMultiprocessorCount, dev), "device attribute");
      check(hipEventCreate(&e0_), "event");
      check(hipEventCreate(&e1_), "event");
    } catch (...) {
      release();  // the destructor does not run for a constructor that throws
      throw;
    }
  }

  ~Batch() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (pending_) {
      // complete the bookkeeping so member domains read as stepped
      try {
        wait();
      } catch (...) {
      }
    }
    release();
  }

  int add(Domain* d, uint64_t steps, float gx, float gy, float gz, float smag, bool forces, float cx, float cy,
          float cz, float r_max) {
    if (pending_) throw HipError("a chunk of this batch is in flight: wait first");
    if (!d) throw HipError("null domain");
    if (d->lattice() != lattice_ || d->storage() != storage_)
      throw HipError("domain lattice or storage differs from the batch's");
    if (d->les_coupled() != les_)
      throw HipError("domain LES coupling differs from the batch's");
    if (d->thermal() != thermal_)
      throw HipError("domain temperature coupling differs from the batch's");
    for (const auto& m : members_)
      if (m.d == d) throw HipError("domain is already in this batch");
    Member m;
    m.id = next_id_++;
    m.d = d;
    m.target = d->current_step() + steps;
    m.g[0] = gx;
    m.g[1] = gy;
    m.g[2] = gz;
    m.smag = smag;
    m.forces = forces;
    m.c[0] = cx;
    m.c[1] = cy;
    m.c[2] = cz;
    m.r_max = r_max;
    members_.push_back(m);
    layout_dirty_ = true;
    return m.id;
  }

  void remove(int id) {
    if (pending_) throw HipError("a chunk of this batch is in flight: wait first");
    auto it = find(id);
    members_.erase(it);
    layout_dirty_ = true;
  }

  uint64_t remaining(int id) {
    const Member& m = *find(id);
    const uint64_t done = m.d->current_step();
    return m.target > done ? m.target - done : 0;
  }

  const std::vector<double>& forces(int id) { return find(id)->hist; }

  void set_shape(LaunchShape s) {
    if (s.block != 64 && s.block != 128 && s.block != 256) throw HipError("block must be 64, 128 or 256");
    if (s.grid_factor < 0) throw HipError("grid factor must be >= 0");
    shape_ = s;
    shape_explicit_ = true;
    layout_dirty_ = true;
  }
  // The shape the next launch uses: until one is set or tuned, the
  // row-length default (resolved at launch time from the current members).
  LaunchShape shape() const {
    if (shape_explicit_) return shape_;
    return LaunchShape{members_.empty() ? shape_.block : default_block(), shape_.grid_factor};
  }
  int slots() const { return static_cast<int>(members_.size()); }
  int tiles() const { return ntiles_; }
  uint64_t nodes() const {
    uint64_t n = 0;
    for (const auto& m : members_) n += m.d->total_nodes();
    return n;
  }

  // `steps` launches (each domain takes min(steps, its remaining steps));
  // returns the device time in ms between the first and the last launch.
  float run(uint64_t steps, bool use_graph = true) {
    if (pending_) throw HipError("a chunk of this batch is in flight: wait first");
    prepare();
    if (members_.empty() || steps == 0) return 0.0f;
    check(hipEventRecord(e0_, stream_), "event");
    uint64_t left = steps;
    while (left > 0 && any_active()) {
      const int k = static_cast<int>(std::min<uint64_t>(left, static_cast<uint64_t>(chunk_)));
      if (use_graph && k == chunk_) {
        if (!graph_exec_) capture();
        check(hipGraphLaunch(graph_exec_, stream_), "hipGraphLaunch");
      } else {
        enqueue_chunk(k);
      }
      collect(k);
      left -= static_cast<uint64_t>(k);
    }
    check(hipEventRecord(e1_, stream_), "event");
    check(hipEventSynchronize(e1_), "event");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, e0_, e1_), "event");
    finish_run();
    return ms;
  }

  // One chunk (min(steps, chunk) launches) enqueued on the batch's stream;
  // wait() completes it.  Batches on their own streams overlap on the GPU.
  void launch(uint64_t steps, bool use_graph = true) {
    if (pending_) throw HipError("a chunk of this batch is already in flight");
    prepare();
    pending_k_ = 0;
    if (members_.empty() || steps == 0 || !any_active()) return;
    const int k = static_cast<int>(std::min<uint64_t>(steps, static_cast<uint64_t>(chunk_)));
    check(hipEventRecord(e0_, stream_), "event");
    if (use_graph && k == chunk_) {
      if (!graph_exec_) capture();
      check(hipGraphLaunch(graph_exec_, stream_), "hipGraphLaunch");
    } else {
      enqueue_chunk(k);
    }
    enqueue_hist_copy();
    check(hipEventRecord(e1_, stream_), "event");
    pending_ = true;
    pending_k_ = k;
  }

  float wait() {
    if (!pending_) return 0.0f;
    pending_ = false;
    check(hipEventSynchronize(e1_), "event");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, e0_, e1_), "event");
    host_collect(pending_k_);
    finish_run();
    return ms;
  }

  bool pending() const { return pending_; }

  // Times `probe` steps under each candidate shape and keeps the fastest;
  // the choice is cached per (kind, size class) for later batches.
  LaunchShape tune(int probe) {
    if (pending_) throw HipError("a chunk of this batch is in flight: wait first");
    prepare();
    // the row-length class (the width that wastes the fewest lanes on the
    // members' rows) is part of the key: a shape tuned for long rows is not
    // reused for a batch of short ones of similar total size
    const auto key = std::make_tuple(lattice_ + (thermal_ ? 100 : 0), storage_, collision_, les_,
                                     static_cast<int>(std::log2(static_cast<double>(std::max<uint64_t>(nodes(), 1)))),
                                     static_cast<int>(std::log2(static_cast<double>(std::max(slots(), 1)))),
                                     default_block());
    {
      std::lock_guard<std::mutex> lk(cache_mutex());
      auto it = cache().find(key);
      if (it != cache().end()) {
        set_shape(it->second);
        return shape_;
      }
    }
    static const LaunchShape cands[] = {{256, 0}, {128, 0}, {64, 0}, {256, 4}, {256, 8}, {128, 8}, {128, 16}, {64, 32}};
    LaunchShape best = shape_;
    float best_ms = 0.0f;
    bool have = false;
    for (const LaunchShape& c : cands) {
      if (!any_active()) break;
      set_shape(c);
      prepare();
      run(1, false);  // first launch of a shape: code object load, caches
      if (!any_active()) break;
      const float ms = run(static_cast<uint64_t>(std::max(probe, 1)), false);
      if (!have || ms < best_ms) {
        best = c;
        best_ms = ms;
        have = true;
      }
    }
    set_shape(best);
    if (have) {
      std::lock_guard<std::mutex> lk(cache_mutex());
      cache()[key] = best;
    }
    return best;
  }

 private:
  struct Member {
    int id;
    Domain* d;
    uint64_t target;
    float g[3];
    float smag;
    bool forces;
    float c[3];
    float r_max;
    uint64_t seen_version = ~0ull;
    // host copy of the device slot progress
    uint64_t step = 0;
    int parity = 0;
    uint64_t run_start = 0;
    std::vector<double> hist;
  };

  std::vector<Member>::iterator find(int id) {
    auto it = std::find_if(members_.begin(), members_.end(), [&](const Member& m) { return m.id == id; });
    if (it == members_.end()) throw HipError("unknown batch slot");
    return it;
  }

  bool any_active() const {
    for (const auto& m : members_)
      if (m.step < m.target) return true;
    return false;
  }

  static std::map<std::tuple<int, int, int, bool, int, int, int>, LaunchShape>& cache() {
    static std::map<std::tuple<int, int, int, bool, int, int, int>, LaunchShape> c;
    return c;
  }
  static std::mutex& cache_mutex() {
    static std::mutex m;
    return m;
  }

  // Until a shape is set or tuned: the work-group width of 256 / 128 / 64
  // that leaves the fewest idle lanes over every member's rows (the larger
  // on a tie), as the single-domain launch picks it.  A row of 128 nodes in
  // 256-lane tiles idles half the lanes: a batch of eight 128 x 64 x 64
  // D3Q27 domains steps in 274 us with 256-lane tiles, 228 us with the 128
  // this picks, and 200 us with the 64 that tune() finds.
  int default_block() const {
    int best = 256;
    unsigned long long best_lanes = ~0ull;
    for (int b : {256, 128, 64}) {
      unsigned long long lanes = 0;
      for (const auto& m : members_)
        lanes += static_cast<unsigned long long>(m.d->ny()) * m.d->nz() * ((m.d->nx() + b - 1) / b) * b;
      if (lanes < best_lanes) {
        best = b;
        best_lanes = lanes;
      }
    }
    return best;
  }

  // Free every device and host resource (the stream has drained).
  void release() noexcept {
    drop_graph();
    if (d_slots_) (void)hipFree(d_slots_);
    if (d_tiles_) (void)hipFree(d_tiles_);
    if (d_hist_) (void)hipFree(d_hist_);
    if (h_hist_) (void)hipHostFree(h_hist_);
    if (e0_) (void)hipEventDestroy(e0_);
    if (e1_) (void)hipEventDestroy(e1_);
    if (stream_) (void)hipStreamDestroy(stream_);
    d_slots_ = nullptr;
    d_tiles_ = nullptr;
    d_hist_ = nullptr;
    h_hist_ = nullptr;
    e0_ = nullptr;
    e1_ = nullptr;
    stream_ = nullptr;
    (void)hipGetLastError();  // nothing of the teardown reaches the next launch check
  }

  void drop_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
  }

  // Slot table and tile map for the current members; re-read a domain whose
  // state or configuration changed outside the batch.
  void prepare() {
    bool table_dirty = layout_dirty_;
    for (auto& m : members_)
      if (m.d->version() != m.seen_version) table_dirty = true;
    if (!table_dirty) return;
    check(hipStreamSynchronize(stream_), "batch");
    drop_graph();
    const int n = static_cast<int>(members_.size());
    if (!shape_explicit_) shape_.block = default_block();
    std::vector<BatchSlot> slots(n);
    std::vector<int> tile_slot;
    any_forces_ = false;
    buf_ok_ = true;
    for (int s = 0; s < n; ++s) {
      Member& m = members_[s];
      BatchSlot& b = slots[s];
      std::memset(&b, 0, sizeof(b));
      int parity = 0;
      bool s_state = false;
      m.d->batch_prepare(m.g[0], m.g[1], m.g[2], m.smag, b.a, b.buf, b.gbuf, parity, s_state);
      b.a.tau = m.d->tau();
      // descriptor addressing needs every member's buffer under 4 GiB
      const unsigned long long elem = storage_ == kStoreF32 ? 4ull : 2ull;
      if (static_cast<unsigned long long>(m.d->lattice()) * b.a.stride * elem >= (1ull << 32)) buf_ok_ = false;
      b.step = m.d->current_step();
      b.stop_at = m.target;
      b.s_at = s_state ? b.step : kNoStep;
      b.parity = parity;
      b.nodes = static_cast<int>(m.d->total_nodes());
      b.tile0 = static_cast<int>(tile_slot.size());
      b.tiles_per_row = (m.d->nx() + shape_.block - 1) / shape_.block;
      b.ntiles = m.d->ny() * m.d->nz() * b.tiles_per_row;
      tile_slot.insert(tile_slot.end(), b.ntiles, s);
      if (m.forces) {
        b.mem_nodes = m.d->batch_mem_nodes(m.c[0], m.c[1], m.c[2], m.r_max, &b.mem_count);
        b.mem_on = 1;
        b.a.mem_c[0] = m.c[0];
        b.a.mem_c[1] = m.c[1];
        b.a.mem_c[2] = m.c[2];
        b.a.mem_rmax = m.r_max;
        any_forces_ = true;
      }
      m.step = b.step;
      m.parity = parity;
      m.seen_version = m.d->version();
    }
    ntiles_ = static_cast<int>(tile_slot.size());
    if (n > slot_cap_) {
      if (d_slots_) (void)hipFree(d_slots_);
      check(hipMalloc(reinterpret_cast<void**>(&d_slots_), n * sizeof(BatchSlot)), "hipMalloc");
      slot_cap_ = n;
    }
    if (ntiles_ > tile_cap_) {
      if (d_tiles_) (void)hipFree(d_tiles_);
      check(hipMalloc(reinterpret_cast<void**>(&d_tiles_), ntiles_ * sizeof(int)), "hipMalloc");
      tile_cap_ = ntiles_;
    }
    const size_t hist_n = static_cast<size_t>(n) * chunk_ * 3;
    if (any_forces_ && hist_n > hist_cap_) {
      if (d_hist_) (void)hipFree(d_hist_);
      if (h_hist_) (void)hipHostFree(h_hist_);
      check(hipMalloc(reinterpret_cast<void**>(&d_hist_), hist_n * sizeof(double)), "hipMalloc");
      check(hipHostMalloc(reinterpret_cast<void**>(&h_hist_), hist_n * sizeof(double), hipHostMallocDefault),
            "hipHostMalloc");
      hist_cap_ = hist_n;
    }
    if (n > 0) {
      check(hipMemcpy(d_slots_, slots.data(), n * sizeof(BatchSlot), hipMemcpyHostToDevice), "slots");
      check(hipMemcpy(d_tiles_, tile_slot.data(), ntiles_ * sizeof(int), hipMemcpyHostToDevice), "tiles");
    }
    layout_dirty_ = false;
  }

  int grid_x() const {
    if (shape_.grid_factor == 0) return ntiles_;
    return std::min(ntiles_, shape_.grid_factor * cus_);
  }

  void enqueue_chunk(int k) {
    const int n = static_cast<int>(members_.size());
    for (int i = 0; i < k; ++i) {
      check(launch_step_batch(lattice_, storage_, collision_, les_, thermal_, buf_ok_, d_slots_, d_tiles_, ntiles_, i,
                              shape_.block, grid_x(), stream_),
            "batch step");
      if (any_forces_) check(launch_mem_batch(lattice_, storage_, d_slots_, n, i, chunk_, d_hist_, stream_), "batch forces");
    }
    check(launch_advance_batch(d_slots_, n, k, stream_), "batch advance");
  }

  void capture() {
    drop_graph();
    check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
    try {
      enqueue_chunk(chunk_);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    check(hipStreamEndCapture(stream_, &graph_), "end capture");
    check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "graph instantiate");
  }

  // Host bookkeeping of a chunk of k launches (mirrors k_advance_batch) and
  // the force history.
  void enqueue_hist_copy() {
    if (!any_forces_) return;
    check(hipMemcpyAsync(h_hist_, d_hist_, members_.size() * chunk_ * 3 * sizeof(double), hipMemcpyDeviceToHost,
                         stream_),
          "forces");
  }

  void collect(int k) {
    enqueue_hist_copy();
    if (any_forces_) check(hipStreamSynchronize(stream_), "forces");
    host_collect(k);
  }

  // after the chunk (and its force copy) completed
  void host_collect(int k) {
    const int n = static_cast<int>(members_.size());
    for (int s = 0; s < n; ++s) {
      Member& m = members_[s];
      const uint64_t left = m.target > m.step ? m.target - m.step : 0;
      const uint64_t taken = std::min<uint64_t>(left, static_cast<uint64_t>(k));
      if (m.forces)
        for (uint64_t i = 0; i < taken; ++i) {
          const double* o = h_hist_ + (static_cast<size_t>(s) * chunk_ + i) * 3;
          m.hist.insert(m.hist.end(), o, o + 3);
        }
      m.step += taken;
      m.parity = (m.parity + static_cast<int>(taken & 1)) & 1;
    }
  }

  void finish_run() {
    for (auto& m : members_) {
      const uint64_t k = m.step - m.d->current_step();
      if (k > 0) {
        m.d->batch_advance(k, m.parity, m.g[0], m.g[1], m.g[2]);
        m.seen_version = m.d->version();
      }
    }
  }

  int lattice_, storage_, collision_;
  bool les_, thermal_;
  int chunk_;
  int cus_ = 256;
  hipStream_t stream_ = nullptr;
  hipEvent_t e0_ = nullptr, e1_ = nullptr;
  std::vector<Member> members_;
  int next_id_ = 0;
  LaunchShape shape_;
  bool shape_explicit_ = false;  // set_shape() or tune() chose it
  bool layout_dirty_ = true;
  bool any_forces_ = false;
  bool buf_ok_ = true;
  bool pending_ = false;
  int pending_k_ = 0;
  BatchSlot* d_slots_ = nullptr;
  int slot_cap_ = 0;
  int* d_tiles_ = nullptr;
  int tile_cap_ = 0;
  int ntiles_ = 0;
  double* d_hist_ = nullptr;
  double* h_hist_ = nullptr;
  size_t hist_cap_ = 0;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
};

}  // namespace wake_hip

struct wake_batch {
  wake_hip::Batch* b;
};

namespace {
template <class F>
int bguard(wake_batch* b, F&& f) {
  if (!b || !b->b) return wake_hip::set_last_error("wake: null batch handle");
  try {
    f(*b->b);
    return 0;
  } catch (const std::exception& e) {
    return wake_hip::set_last_error(e.what());
  }
}
}  // namespace

extern "C" {

int wake_batch_create(int lattice, int storage, int collision, int use_les, int thermal, int chunk,
                      wake_batch** out) {
  if (!out) return wake_hip::set_last_error("wake: null output handle");
  *out = nullptr;
  try {
    *out = new wake_batch{new wake_hip::Batch(lattice, storage, collision, use_les != 0, thermal != 0, chunk)};
    return 0;
  } catch (const std::exception& e) {
    return wake_hip::set_last_error(e.what());
  }
}

void wake_batch_destroy(wake_batch* b) {
  if (!b) return;
  delete b->b;
  delete b;
}

int wake_batch_add(wake_batch* b, wake_domain* d, long long steps, float gx, float gy, float gz, float smag,
                   int track_forces, float cx, float cy, float cz, float r_max, int* slot) {
  return bguard(b, [&](wake_hip::Batch& B) {
    if (!d || !d->d) throw wake_hip::HipError("null domain handle");
    if (steps < 0) throw wake_hip::HipError("negative step count");
    const int id = B.add(d->d, static_cast<uint64_t>(steps), gx, gy, gz, smag, track_forces != 0, cx, cy, cz, r_max);
    if (slot) *slot = id;
  });
}

int wake_batch_remove(wake_batch* b, int slot) {
  return bguard(b, [&](wake_hip::Batch& B) { B.remove(slot); });
}

int wake_batch_run(wake_batch* b, long long steps, int use_graph, float* ms) {
  return bguard(b, [&](wake_hip::Batch& B) {
    if (steps < 0) throw wake_hip::HipError("negative step count");
    const float t = B.run(static_cast<uint64_t>(steps), use_graph != 0);
    if (ms) *ms = t;
  });
}

int wake_batch_launch(wake_batch* b, long long steps, int use_graph) {
  return bguard(b, [&](wake_hip::Batch& B) {
    if (steps < 0) throw wake_hip::HipError("negative step count");
    B.launch(static_cast<uint64_t>(steps), use_graph != 0);
  });
}

int wake_batch_wait(wake_batch* b, float* ms) {
  return bguard(b, [&](wake_hip::Batch& B) {
    const float t = B.wait();
    if (ms) *ms = t;
  });
}

int wake_batch_remaining(wake_batch* b, int slot, long long* remaining) {
  return bguard(b, [&](wake_hip::Batch& B) { *remaining = static_cast<long long>(B.remaining(slot)); });
}

int wake_batch_forces(wake_batch* b, int slot, double* out, long long capacity, long long* count) {
  return bguard(b, [&](wake_hip::Batch& B) {
    const auto& h = B.forces(slot);
    const long long n = static_cast<long long>(h.size() / 3);
    if (count) *count = n;
    if (out) std::memcpy(out, h.data(), static_cast<size_t>(std::min(n, capacity)) * 3 * sizeof(double));
  });
}

int wake_batch_tune(wake_batch* b, int probe_steps, int* block, int* grid_factor) {
  return bguard(b, [&](wake_hip::Batch& B) {
    const wake_hip::LaunchShape s = B.tune(probe_steps);
    if (block) *block = s.block;
    if (grid_factor) *grid_factor = s.grid_factor;
  });
}

int wake_batch_set_launch(wake_batch* b, int block, int grid_factor) {
  return bguard(b, [&](wake_hip::Batch& B) { B.set_shape({block, grid_factor}); });
}

int wake_batch_info(wake_batch* b, long long out[4]) {
  return bguard(b, [&](wake_hip::Batch& B) {
    out[0] = B.slots();
    out[1] = static_cast<long long>(B.nodes());
    out[2] = B.shape().block;
    out[3] = B.shape().grid_factor;
  });
}

}  // extern "C"
