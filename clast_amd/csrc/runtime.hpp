// Per-device runtime of libclast_hip: a compute stream and a copy stream
// (both default priority), a growable named workspace for stage scratch,
// pinned staging for scalar read-backs, and the error type that the C ABI
// turns into status codes.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace clast {

struct ClastError : std::runtime_error {
  int code;
  ClastError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  const int code = (e == hipErrorOutOfMemory) ? 3 : 2;
  throw ClastError(code, std::string(what) + ": " + hipGetErrorString(e));
}

// Device blocks of released handles, kept per device for the next handle:
// hipFree waits for the whole device, so a query set or result buffer freed
// by one host thread would stall the searches other threads have in flight.
// Blocks come in size classes (powers of two from 64 KiB, quarter steps
// above 64 MiB); at most kCacheBytes per device are kept, the rest is freed.
// A block is only released once no queued work uses it (every entry point
// that frees synchronises its stream first).
class DeviceCache {
 public:
  static DeviceCache& get() {
    static DeviceCache* c = new DeviceCache();  // never destroyed: blocks may outlive static teardown
    return *c;
  }
  static uint64_t size_class(uint64_t bytes) {
    uint64_t c = 64ull << 10;
    while (c < bytes && c < (64ull << 20)) c <<= 1;
    if (c >= bytes) return c;
    const uint64_t q = 16ull << 20;  // quarter of 64 MiB steps, then of each power of two
    uint64_t p = 64ull << 20;
    while (p * 2 < bytes) p <<= 1;
    const uint64_t step = std::max(q, p / 4);
    return (bytes + step - 1) / step * step;
  }
  void* take(int dev, uint64_t bytes) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto& fl = free_[dev];
      auto it = fl.lower_bound(bytes);
      if (it != fl.end() && it->first <= bytes + bytes / 4) {
        void* p = it->second;
        cached_[dev] -= it->first;
        fl.erase(it);
        return p;
      }
    }
    void* p = nullptr;
    hip_check(hipMalloc(&p, bytes), "hipMalloc");
    return p;
  }
  void give(int dev, void* p, uint64_t bytes) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (cached_[dev] + bytes <= kCacheBytes) {
        free_[dev].emplace(bytes, p);
        cached_[dev] += bytes;
        return;
      }
    }
    (void)hipFree(p);
  }

 private:
  static constexpr uint64_t kCacheBytes = 16ull << 30;
  std::mutex mu_;
  std::map<int, std::multimap<uint64_t, void*>> free_;
  std::map<int, uint64_t> cached_;
};

// A device allocation owned by a handle (through DeviceCache).
template <class T>
struct DevArray {
  T* ptr = nullptr;
  uint64_t n = 0;
  uint64_t bytes = 0;
  int dev = 0;
  DevArray() = default;
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  ~DevArray() { reset(); }
  void reset() {
    if (ptr) DeviceCache::get().give(dev, ptr, bytes);
    ptr = nullptr;
    n = 0;
    bytes = 0;
  }
  void alloc(uint64_t count) {
    reset();
    if (count == 0) count = 1;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    bytes = DeviceCache::size_class(sizeof(T) * count);
    ptr = static_cast<T*>(DeviceCache::get().take(dev, bytes));
    n = count;
  }
  void upload(const T* host, uint64_t count, hipStream_t s) {
    alloc(count);
    if (count) hip_check(hipMemcpyAsync(ptr, host, sizeof(T) * count, hipMemcpyHostToDevice, s), "upload");
  }
};

// Named scratch buffers that only grow.  A slot keeps its pointer until a
// larger request for the same name; growth waits for the stream first so no
// queued kernel still uses the old buffer.
class Workspace {
 public:
  Workspace(hipStream_t s, hipStream_t copy, hipStream_t side) : stream_(s), copy_(copy), side_(side) {}
  ~Workspace() {
    for (auto& kv : slots_) (void)hipFree(kv.second.ptr);
  }
  template <class T>
  T* get(const std::string& name, uint64_t count) {
    const uint64_t bytes = (count == 0 ? 1 : count) * sizeof(T);
    if (arena_) return static_cast<T*>(arena_get(name, bytes));
    Slot& s = slots_[name];
    if (s.bytes < bytes) {
      if (s.ptr) {
        hip_check(hipStreamSynchronize(stream_), "workspace sync");
        hip_check(hipStreamSynchronize(copy_), "workspace sync");
        hip_check(hipStreamSynchronize(side_), "workspace sync");
        hip_check(hipFree(s.ptr), "hipFree");
        s.ptr = nullptr;
        s.bytes = 0;
      }
      const uint64_t want = bytes + bytes / 4;
      hip_check(hipMalloc(&s.ptr, want), "workspace hipMalloc");
      s.bytes = want;
    }
    return static_cast<T*>(s.ptr);
  }
  uint64_t bytes() const {
    uint64_t b = 0;
    for (auto& kv : slots_) b += kv.second.bytes;
    return b;
  }

  // Arena mode: until end_arena, every slot is carved from the device block
  // [base, base + cap) by bump allocation (256-byte aligned), nothing is
  // allocated, freed or synchronised, and a request the block cannot hold
  // throws CLAST_ERR_LIMIT.  A name asked for again gets its region back if
  // it fits, else a new one (content is not kept, as on growth).  Regions
  // are laid out front to back for the call; begin_arena starts over.  The slot
  // table is fixed-size and its names are short (no heap use).
  void begin_arena(void* base, uint64_t cap) {
    arena_ = static_cast<uint8_t*>(base);
    arena_cap_ = cap;
    arena_used_ = 0;
    arena_n_ = 0;
  }
  void end_arena() { arena_ = nullptr; }
  uint64_t arena_used() const { return arena_used_; }

 private:
  struct Slot {
    void* ptr = nullptr;
    uint64_t bytes = 0;
  };
  struct ArenaSlot {
    char name[24];
    uint64_t off, bytes;
  };
  static constexpr int kArenaSlots = 160;
  void* arena_get(const std::string& name, uint64_t bytes) {
    if (name.size() >= sizeof(ArenaSlot::name)) throw ClastError(4, "workspace slot name too long: " + name);
    int k = 0;
    for (; k < arena_n_; ++k)
      if (name == arena_slots_[k].name) break;
    if (k < arena_n_ && arena_slots_[k].bytes >= bytes) return arena_ + arena_slots_[k].off;
    // the same 25% slack as a grown slot, so both modes lay buffers out alike
    bytes += bytes / 4;
    const uint64_t off = (arena_used_ + 255) & ~255ull;
    if (off > arena_cap_ || bytes > arena_cap_ - off)
      throw ClastError(4, "the search needs more than the " + std::to_string(arena_cap_ >> 20) +
                              " MiB scratch of the allocation-free entry point (slot " + name + ")");
    if (k == arena_n_) {
      if (arena_n_ == kArenaSlots) throw ClastError(4, "workspace slot table full");
      std::memcpy(arena_slots_[k].name, name.c_str(), name.size() + 1);
      ++arena_n_;
    }
    arena_slots_[k].off = off;
    arena_slots_[k].bytes = bytes;
    arena_used_ = off + bytes;
    return arena_ + off;
  }
  hipStream_t stream_, copy_, side_;
  std::map<std::string, Slot> slots_;
  uint8_t* arena_ = nullptr;
  uint64_t arena_cap_ = 0, arena_used_ = 0;
  int arena_n_ = 0;
  ArenaSlot arena_slots_[kArenaSlots];
};

class Runtime {
 public:
  explicit Runtime(int device) : device_(device) {
    hip_check(hipSetDevice(device), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamCreateWithFlags(&copy_stream_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipStreamCreateWithFlags(&side_stream_, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : out_done_) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    for (auto& e : side_ev_) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipHostMalloc(&pinned_, kPinned, hipHostMallocDefault), "hipHostMalloc");
    ws_ = new Workspace(stream_, copy_stream_, side_stream_);
    hip_check(hipDeviceGetAttribute(&ncu_, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
  }
  ~Runtime() {
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
    delete ws_;
    (void)hipHostFree(pinned_);
    for (auto& e : out_done_) (void)hipEventDestroy(e);
    for (auto& e : side_ev_) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(side_stream_);
    (void)hipStreamDestroy(copy_stream_);
    (void)hipStreamDestroy(stream_);
  }
  hipStream_t stream() const { return stream_; }
  // Device-to-host copies of result rows run here, so one search's copy
  // overlaps the next search's kernels.
  hipStream_t copy_stream() const { return copy_stream_; }
  // Kernels that overlap the compute stream's work inside one search (the
  // long DP extensions); fork(k) / join(k) order it against the compute
  // stream.
  hipStream_t side_stream() const { return side_stream_; }
  void fork(int k) {
    hip_check(hipEventRecord(side_ev_[k], stream_), "hipEventRecord");
    hip_check(hipStreamWaitEvent(side_stream_, side_ev_[k], 0), "hipStreamWaitEvent");
  }
  void join(int k) {
    hip_check(hipEventRecord(side_ev_[k], side_stream_), "hipEventRecord");
    hip_check(hipStreamWaitEvent(stream_, side_ev_[k], 0), "hipStreamWaitEvent");
  }
  int compute_units() const { return ncu_; }
  // Result rows are staged in one of two device buffers, alternating; the
  // event marks the end of the last copy out of buffer k.
  int next_out_slot() { return out_slot_ ^= 1; }
  hipEvent_t out_done(int k) const { return out_done_[k]; }
  bool out_used(int k) const { return out_used_[k]; }
  void set_out_used(int k) { out_used_[k] = true; }
  int device() const { return device_; }
  Workspace& ws() { return *ws_; }
  // Executable graphs of captured launch chains, by a key of everything the
  // launches were captured with (sizes, flags and every device address
  // passed by value); at most kGraphs kept, the oldest dropped first.
  hipGraphExec_t graph(const std::vector<uint64_t>& key) const {
    auto it = graphs_.find(key);
    return it == graphs_.end() ? nullptr : it->second;
  }
  void add_graph(const std::vector<uint64_t>& key, hipGraphExec_t g) {
    if (graphs_.size() >= kGraphs) {
      hip_check(hipStreamSynchronize(stream_), "graph cache sync");
      (void)hipGraphExecDestroy(graphs_.at(graph_order_.front()));
      graphs_.erase(graph_order_.front());
      graph_order_.erase(graph_order_.begin());
    }
    graphs_[key] = g;
    graph_order_.push_back(key);
  }
  size_t graph_count() const { return graphs_.size(); }
  std::mutex& mutex() { return mu_; }

  void check_launch(const char* what) { hip_check(hipGetLastError(), what); }
  void sync() {
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  }

  // Reads `count` device scalars of type T (count * sizeof(T) <= 4 KiB).
  template <class T>
  void read_scalars(const T* dev, T* host, int count) {
    hip_check(hipMemcpyAsync(pinned_, dev, sizeof(T) * count, hipMemcpyDeviceToHost, stream_), "scalar D2H");
    sync();
    for (int i = 0; i < count; ++i) host[i] = static_cast<const T*>(pinned_)[i];
  }
  // Several device regions in one wait: {device pointer, host pointer,
  // bytes} each (at most 4 KiB in all).
  struct Region {
    const void* dev;
    void* host;
    size_t bytes;
  };
  void read_regions(std::initializer_list<Region> regions) {
    size_t off = 0;
    for (const Region& r : regions) {
      if (off + r.bytes > kPinned) throw ClastError(4, "read-back over the staging block");
      hip_check(hipMemcpyAsync(static_cast<uint8_t*>(pinned_) + off, r.dev, r.bytes, hipMemcpyDeviceToHost, stream_),
                "scalar D2H");
      off = (off + r.bytes + 15) & ~size_t{15};
    }
    sync();
    off = 0;
    for (const Region& r : regions) {
      std::memcpy(r.host, static_cast<const uint8_t*>(pinned_) + off, r.bytes);
      off = (off + r.bytes + 15) & ~size_t{15};
    }
  }

 private:
  static constexpr size_t kPinned = 4096;
  static constexpr size_t kGraphs = 64;
  std::map<std::vector<uint64_t>, hipGraphExec_t> graphs_;
  std::vector<std::vector<uint64_t>> graph_order_;
  int device_;
  hipStream_t stream_ = nullptr;
  hipStream_t copy_stream_ = nullptr;
  hipStream_t side_stream_ = nullptr;
  hipEvent_t side_ev_[2] = {nullptr, nullptr};
  int ncu_ = 0;
  hipEvent_t out_done_[2] = {nullptr, nullptr};
  bool out_used_[2] = {false, false};
  int out_slot_ = 0;
  void* pinned_ = nullptr;
  Workspace* ws_ = nullptr;
  std::mutex mu_;
};

// Page-locked host blocks reused across calls: result rows are copied into
// them at full link speed (a pageable destination goes through a driver
// staging copy).  A block goes back to the pool when its holder releases it.
class PinnedPool {
 public:
  struct Block {
    void* ptr = nullptr;
    uint64_t bytes = 0;
  };
  Block acquire(uint64_t bytes) {
    std::lock_guard<std::mutex> lk(mu_);
    size_t best = free_.size();
    for (size_t i = 0; i < free_.size(); ++i)
      if (free_[i].bytes >= bytes && (best == free_.size() || free_[i].bytes < free_[best].bytes)) best = i;
    if (best < free_.size()) {
      const Block b = free_[best];
      free_.erase(free_.begin() + static_cast<std::ptrdiff_t>(best));
      return b;
    }
    Block b;
    b.bytes = std::max<uint64_t>(bytes + bytes / 4, 1ull << 16);
    hip_check(hipHostMalloc(&b.ptr, b.bytes, hipHostMallocDefault), "hipHostMalloc");
    return b;
  }
  void release(Block b) {
    if (!b.ptr) return;
    std::lock_guard<std::mutex> lk(mu_);
    free_.push_back(b);
    if (free_.size() > kKeep) {  // drop the smallest
      size_t k = 0;
      for (size_t i = 1; i < free_.size(); ++i)
        if (free_[i].bytes < free_[k].bytes) k = i;
      (void)hipHostFree(free_[k].ptr);
      free_.erase(free_.begin() + static_cast<std::ptrdiff_t>(k));
    }
  }

 private:
  static constexpr size_t kKeep = 8;
  std::vector<Block> free_;
  std::mutex mu_;
};

// Process-wide pool (never destroyed: blocks may outlive static teardown).
PinnedPool& pinned_pool();

// The runtime of the calling thread's current device (created on first use).
Runtime& runtime();
// The same, or null when the calling thread has none on its current device
// yet (nothing is created).
Runtime* runtime_if_exists();

}  // namespace clast
