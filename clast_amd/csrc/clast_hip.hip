// libclast_hip: C ABI over the gfx950 search pipeline (see include/clast_hip.h).
#include "../../include/clast_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "score_threshold.hpp"
#include "prims.hpp"
#include "runtime.hpp"
#include "tuner.hpp"

namespace clast {

// One runtime per (device, host thread): searches issued from different
// threads run on their own streams and workspaces and overlap on the GPU.
// When a thread ends, its runtimes go to a per-device pool that the next new
// thread draws from, so a host that keeps starting threads holds at most as
// many runtimes as it ever ran threads at once.  Returning one makes no HIP
// call (queued work, e.g. a result copy not yet waited for, stays ordered on
// the runtime's streams for its next owner).  Runtimes live until the
// process ends.
static std::mutex g_rt_mu;
static std::map<std::pair<int, std::thread::id>, Runtime*> g_rt;
static std::map<int, std::vector<Runtime*>> g_rt_free;

struct ThreadRuntimes {
  std::map<int, Runtime*> mine;  // this thread's, without the lock
  ~ThreadRuntimes() {
    std::lock_guard<std::mutex> lk(g_rt_mu);
    for (auto& kv : mine) {
      g_rt.erase(std::make_pair(kv.first, std::this_thread::get_id()));
      g_rt_free[kv.first].push_back(kv.second);
    }
  }
};
static thread_local ThreadRuntimes t_rt;

PinnedPool& pinned_pool() {
  static PinnedPool* pool = new PinnedPool();
  return *pool;
}

Runtime& runtime() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  auto hit = t_rt.mine.find(dev);
  if (hit != t_rt.mine.end()) return *hit->second;
  std::lock_guard<std::mutex> lk(g_rt_mu);
  Runtime* rt = nullptr;
  auto& pool = g_rt_free[dev];
  if (!pool.empty()) {
    rt = pool.back();
    pool.pop_back();
  } else {
    rt = new Runtime(dev);
  }
  g_rt[std::make_pair(dev, std::this_thread::get_id())] = rt;
  t_rt.mine[dev] = rt;
  return *rt;
}

Runtime* runtime_if_exists() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  auto hit = t_rt.mine.find(dev);
  return hit != t_rt.mine.end() ? hit->second : nullptr;
}

// Scratch of clast_search_batch_noalloc.  Each device holds its own copy of
// the array until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns on that device (set up once per device
// by clast_init), never through the array's name, which in host code is not
// a device address.  Calls on one device share it, so the entry point holds
// that device's lock for the whole call.
static __device__ __align__(256) uint8_t scratch[CLAST_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
struct ScratchSlot {
  void* ptr = nullptr;
  std::mutex call;  // one allocation-free call at a time on the device
};
static ScratchSlot g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

static void scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw ClastError(CLAST_ERR_LIMIT, "device index over 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[dev].ptr) return;
  void* ptr = nullptr;
  hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (!ptr) throw ClastError(CLAST_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
  g_scratch[dev].ptr = ptr;
}

// Byte arrays the DP streams read as 16-byte blocks: whole blocks, at least one.
static uint64_t padded_blocks(uint64_t bytes) {
  const uint64_t b = (bytes + 15) / 16;
  if (b >= (1ull << 31)) throw ClastError(CLAST_ERR_LIMIT, "sequence set over 32 GiB");
  return b ? b : 1;
}

static int bits_for(uint64_t count) {  // bits to hold values 0 .. count-1
  int b = 0;
  while (b < 64 && (1ull << b) < count) ++b;
  return b;
}

static void check_params(const clast_params* p) {
  if (!p) throw ClastError(CLAST_ERR_ARG, "null params");
  if (p->lmer < 1 || p->lmer > 31) throw ClastError(CLAST_ERR_ARG, "lmer must be in 1..31");
  if (p->stride < 1) throw ClastError(CLAST_ERR_ARG, "stride must be >= 1");
  if (p->gap < 2 || p->gap > 16) throw ClastError(CLAST_ERR_ARG, "gap must be in 2..16");
}

static void check_window(const clast_window& w) {
  if (w.n_seq < 0 || w.n_slots < 0) throw ClastError(CLAST_ERR_ARG, "negative window size");
  if (w.n_slots > 0 && !w.bases) throw ClastError(CLAST_ERR_ARG, "null window bases");
  if (w.n_seq > 0 && (!w.lengths || !w.gateway)) throw ClastError(CLAST_ERR_ARG, "null window arrays");
  // gateway is rebased (first sequence at slot 0) and ascending, as
  // CHostSeqList::getGateway returns it; every sequence lies inside the window
  if (w.n_seq > 0 && w.gateway[0] != 0) throw ClastError(CLAST_ERR_ARG, "window gateway must start at 0");
  for (int k = 0; k < w.n_seq; ++k) {
    if (w.lengths[k] < 0 || w.gateway[k] < 0 || w.gateway[k] > w.n_slots ||
        static_cast<int64_t>(w.gateway[k]) + w.lengths[k] > w.n_slots)
      throw ClastError(CLAST_ERR_ARG, "window gateway/length out of range");
    if (k > 0 && w.gateway[k] < static_cast<int64_t>(w.gateway[k - 1]) + w.lengths[k - 1])
      throw ClastError(CLAST_ERR_ARG, "window gateway not ascending");
  }
  if (w.n_seq == 0 && w.n_slots != 0) throw ClastError(CLAST_ERR_ARG, "window slots without sequences");
}

// Score threshold of one query length (score_threshold.hpp).
static int32_t score_threshold(const clast_params& p, int qlen) {
  return score_threshold(p.K, p.total_db_size, p.cut_off, p.lambda, qlen);
}

}  // namespace clast

using namespace clast;

struct clast_index_s {
  int device = 0;
  int lmer = 0, stride = 0, cut_repeat = 0;
  int nwin = 0, nseq = 0, max_tlen = 0;
  bool key64 = false;
  int pbits = 0, shift = 0;
  int32_t bmax = 0;  // last 16-byte block of tbase
  bool direct = false;  // cell lookup form (see use_direct)
  uint32_t nentries = 0;
  DevArray<uint8_t> tbase;
  DevArray<uint64_t> win_slot;
  DevArray<uint64_t> seq_slot;
  DevArray<int32_t> seq_len, seq_win, seq_ref;
  DevArray<int2> seq_info;  // (length, window) per target sequence
  DevArray<uint8_t> seq_clean;  // 1: no star / X class in the sequence (16-bit DP), 3: no class 4 either
  DevArray<uint32_t> tnib;      // 4-bit copy of tbase and its mirror (16-bit DP)
  uint32_t ncell = 0;
  DevArray<uint32_t> ekey32;
  DevArray<uint64_t> ekey64;
  DevArray<int2> epos;
  DevArray<unsigned long long> bitmap;
  DevArray<uint32_t> rank;
  DevArray<uint4> cells;
  std::vector<int32_t> h_seq_len, h_seq_win, h_seq_ref, h_id_begin;
  std::vector<uint64_t> h_win_slot;
};

struct clast_query_s {
  int device = 0;
  int lmer = 0;
  int nwin = 0;
  uint32_t nseq = 0, npos = 0;
  int max_qlen = 0, min_qlen = 0;
  int32_t bmax = 0;  // last 16-byte block of qbase
  DevArray<uint8_t> qbase;
  DevArray<uint32_t> qslot, pos_off, pos_strand;
  DevArray<int32_t> qlen, qref, qwin, smin;
  DevArray<uint8_t> qclean;  // no star / X class in the strand (16-bit DP)
  DevArray<uint32_t> qnib;   // 4-bit copy of qbase, other -> 7, and its mirror (16-bit DP)
  std::vector<int32_t> h_qlen, h_qref, h_qwin, h_id_begin;
  bool smin_valid = false;
  double smin_key[4] = {0, 0, 0, 0};
  std::mutex smin_mu;  // searches from several threads share the thresholds
};

struct clast_hits_s {
  // n rows in page-locked memory: 10 int32 planes (tID tIdx qID qIdx tHL
  // qHL match score twin qwin), then n E-values
  clast::PinnedPool::Block rows;
  int64_t n = 0;
  hipEvent_t ready = nullptr;  // the rows' copy into `rows` has finished
  void wait() {
    if (ready) {
      clast::hip_check(hipEventSynchronize(ready), "hipEventSynchronize");
      (void)hipEventDestroy(ready);
      ready = nullptr;
    }
  }
  const int32_t* col(int c) const { return static_cast<const int32_t*>(rows.ptr) + static_cast<int64_t>(c) * n; }
  const double* ev() const {
    return reinterpret_cast<const double*>(static_cast<const int32_t*>(rows.ptr) + 10 * n);
  }
  ~clast_hits_s() {
    if (ready) {
      (void)hipEventSynchronize(ready);
      (void)hipEventDestroy(ready);
    }
    clast::pinned_pool().release(rows);
  }
  std::vector<int32_t> stage[5][8];
  int64_t stage_n[5] = {0, 0, 0, 0, 0};
  float times[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  float ktimes[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // clast_hits_kernel_times
  clast::DPConfig dpcfg;  // DP configuration the search ran with
  // raw seeds, DP hits, DP cells, rows, 16-bit tasks, their cells, long tasks, their cells
  int64_t counts[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // clast_hits_dp_kernels: DP cells of k_dp16, of its past-the-end launch, of
  // k_dp, of k_dp_band, the four kernels' device times (ms), then the second
  // pass's tasks, cells and time (ms), and 1 for a two-pass DP stage
  double dpk[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
};

namespace {

thread_local std::string g_err;

template <class F>
int guard(F&& f) {
  try {
    f();
    return CLAST_OK;
  } catch (const ClastError& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return CLAST_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return CLAST_ERR_ARG;
  }
}

void set_device(int dev) { hip_check(hipSetDevice(dev), "hipSetDevice"); }

// ---------------------------------------------------------------- index

// Form of a new index's cell lookup: CLAST_INDEX_FORM=direct|bitmap forces
// one, otherwise the direct table is used when it is at most 16 GiB and the
// index has at least one entry per 16 prefixes.
static bool use_direct(int B, uint32_t nent) {
  const uint64_t prefixes = 1ull << B;
  if (const char* f = std::getenv("CLAST_INDEX_FORM")) {
    if (!std::strcmp(f, "direct")) return true;
    if (!std::strcmp(f, "bitmap")) return false;
    if (std::strcmp(f, "auto")) throw ClastError(CLAST_ERR_ARG, "CLAST_INDEX_FORM must be direct, bitmap or auto");
  }
  return prefixes * sizeof(uint4) <= (16ull << 30) && static_cast<uint64_t>(nent) * 16 >= prefixes;
}

// CLAST_DP16=0 sends every DP task to the int32 kernel (k_dp); by default
// the tasks that fit run in packed 16-bit lanes (k_dp16).  Both give the
// same results; the switch exists to compare them.
static bool dp16_disabled() {
  const char* f = std::getenv("CLAST_DP16");
  return f && !std::strcmp(f, "0");
}
// CLAST_GRAPH=0 launches the DP stage's task-ordering chain kernel by
// kernel; by default it is captured once per shape into a hipGraph and
// replayed.  Both give the same rows.
static bool graphs_disabled() {
  const char* f = std::getenv("CLAST_GRAPH");
  return f && !std::strcmp(f, "0");
}

// Cell lookup over the top B key bits of the sorted entries (keys in ek,
// u64): presence bitmap + rank directory + a record per present prefix, or a
// direct record table; keeps the keys as u32 or u64.
void build_cells(Runtime& rt, clast_index_s& ix, const uint64_t* ek, uint32_t nent, int keybits) {
  hipStream_t st = rt.stream();
  Workspace& ws = rt.ws();
  const int B = std::min(keybits, 30);
  ix.pbits = B;
  ix.shift = keybits - B;
  ix.direct = use_direct(B, nent);
  const uint64_t nwords = std::max<uint64_t>(1, (1ull << B) / 64);
  const uint64_t nblk = (nwords + 7) / 8;
  ix.bitmap.alloc(nwords);
  k_fill<unsigned long long><<<grid_for(nwords, kBlock), kBlock, 0, st>>>(ix.bitmap.ptr, nwords, 0ull);
  if (ix.direct) {
    // absent prefixes are never read (the bitmap answers first): no fill
    ix.rank.alloc(1);
    ix.cells.alloc(1ull << B);
  }
  uint32_t ncell = 0;
  if (nent > 0) {
    uint32_t* head = ws.get<uint32_t>("ix.chead", nent);
    uint32_t* cell = ws.get<uint32_t>("ix.cell", nent);
    uint32_t* scal = ws.get<uint32_t>("ix.cscal", 2);
    k_cell_heads<<<grid_for(nent, kBlock), kBlock, 0, st>>>(ek, nent, ix.shift, head);
    device_scan<OpAdd, uint32_t, uint32_t, uint32_t>(rt, head, nent, cell, scal, false);
    rt.read_scalars(scal, &ncell, 1);
    uint32_t* cell_off = ws.get<uint32_t>("ix.coff", static_cast<uint64_t>(ncell) + 1);
    k_cell_build<<<grid_for(nent, kBlock), kBlock, 0, st>>>(ek, nent, ix.shift, head, cell, cell_off,
                                                             ix.bitmap.ptr);
    if (!ix.direct) ix.cells.alloc(ncell);
    k_cell_records<<<grid_for(ncell, kBlock), kBlock, 0, st>>>(cell_off, ncell, ix.epos.ptr, ek, ix.shift,
                                                                ix.direct ? 1 : 0, ix.cells.ptr);
    rt.check_launch("k_cell_build");
    if (ix.key64) {
      ix.ekey64.alloc(nent);
      hip_check(hipMemcpyAsync(ix.ekey64.ptr, ek, sizeof(uint64_t) * nent, hipMemcpyDeviceToDevice, st), "keys");
    } else {
      ix.ekey32.alloc(nent);
      k_narrow_keys<<<grid_for(nent, kBlock), kBlock, 0, st>>>(ek, nent, ix.ekey32.ptr);
    }
  } else {
    if (!ix.direct) {
      ix.cells.alloc(1);
      k_fill<uint4><<<1, 64, 0, st>>>(ix.cells.ptr, 1, make_uint4(0, 0, 0, 0));
    }
    ix.ekey32.alloc(1);
    ix.ekey64.alloc(1);
  }
  ix.ncell = ncell;
  if (!ix.direct) {
    uint32_t* bc = ws.get<uint32_t>("ix.bcnt", nblk);
    ix.rank.alloc(nblk);
    k_rank_count<<<grid_for(nblk, kBlock), kBlock, 0, st>>>(ix.bitmap.ptr, nwords, nblk, bc);
    device_scan<OpAdd, uint32_t, uint32_t, uint32_t>(rt, bc, nblk, ix.rank.ptr, static_cast<uint32_t*>(nullptr),
                                                     false);
  }
  rt.check_launch("rank");
}

// Device tables of a target set derived from its host tables and class
// bytes (ix.tbase, filled): window and sequence tables, the 4-bit copy and
// the clean flags of the 16-bit DP.
static void target_tables(Runtime& rt, clast_index_s& ix, const std::vector<uint64_t>& seq_slot) {
  hipStream_t st = rt.stream();
  const std::vector<uint64_t>& win_slot = ix.h_win_slot;
  ix.win_slot.upload(win_slot.data(), win_slot.size(), st);
  ix.seq_slot.upload(seq_slot.data(), seq_slot.size(), st);
  ix.seq_len.upload(ix.h_seq_len.data(), ix.h_seq_len.size(), st);
  ix.seq_win.upload(ix.h_seq_win.data(), ix.h_seq_win.size(), st);
  {
    std::vector<int2> info(ix.h_seq_len.size());
    for (size_t i = 0; i < info.size(); ++i) info[i] = make_int2(ix.h_seq_len[i], ix.h_seq_win[i]);
    ix.seq_info.upload(info.data(), info.size(), st);
  }
  ix.seq_ref.upload(ix.h_seq_ref.data(), ix.h_seq_ref.size(), st);
  {
    const uint64_t ndw = 2 * (static_cast<uint64_t>(ix.bmax) + 1);
    ix.tnib.alloc(2 * ndw);  // with its mirror
    k_pack_nibbles<<<grid_for(ndw, kBlock), kBlock, 0, st>>>(ix.tbase.ptr, ndw, 0, ix.tnib.ptr);
  }
  ix.seq_clean.alloc(std::max(ix.nseq, 1));
  if (ix.nseq)
    k_clean_flags<uint64_t><<<grid_for(ix.nseq, kBlock / 64), kBlock, 0, st>>>(ix.tbase.ptr, ix.seq_slot.ptr,
                                                                               ix.seq_len.ptr, ix.nseq, 1,
                                                                               ix.seq_clean.ptr);
}

void build_index(const clast_params& p, const clast_window* wins, int nwin, clast_index_s& ix) {
  Runtime& rt = runtime();
  std::lock_guard<std::mutex> lk(rt.mutex());
  hipStream_t st = rt.stream();
  ix.device = rt.device();
  ix.lmer = p.lmer;
  ix.stride = p.stride;
  ix.cut_repeat = p.cut_repeat;
  ix.nwin = nwin;
  ix.key64 = p.lmer > 16;

  // host layout: window slot offsets, per target sequence tables
  std::vector<uint64_t> win_slot(nwin + 1, 0);
  std::vector<uint32_t> win_stride(nwin + 1, 0), win_seq(nwin + 1, 0);
  std::vector<uint64_t> seq_slot;
  uint64_t ns = 0;
  for (int w = 0; w < nwin; ++w) {
    check_window(wins[w]);
    win_slot[w + 1] = win_slot[w] + static_cast<uint64_t>(wins[w].n_slots);
    ns += static_cast<uint64_t>(wins[w].n_slots) / static_cast<uint64_t>(p.stride);
    if (ns >= (1ull << 32) - 1) throw ClastError(CLAST_ERR_LIMIT, "more than 2^32 strided target k-mers");
    win_stride[w + 1] = static_cast<uint32_t>(ns);
    for (int k = 0; k < wins[w].n_seq; ++k) {
      seq_slot.push_back(win_slot[w] + static_cast<uint64_t>(wins[w].gateway[k]));
      ix.h_seq_len.push_back(wins[w].lengths[k]);
      ix.h_seq_win.push_back(w);
      ix.h_seq_ref.push_back(wins[w].id_begin + k);
      ix.max_tlen = std::max(ix.max_tlen, wins[w].lengths[k]);
    }
    win_seq[w + 1] = static_cast<uint32_t>(seq_slot.size());
    ix.h_id_begin.push_back(wins[w].id_begin);
  }
  ix.nseq = static_cast<int>(seq_slot.size());
  ix.h_win_slot = win_slot;
  // a sequence's slot is never past the next sequence's; windows are in order
  ix.tbase.alloc(padded_blocks(win_slot[nwin]) * 16);
  ix.bmax = static_cast<int32_t>(padded_blocks(win_slot[nwin]) - 1);
  for (int w = 0; w < nwin; ++w)
    if (wins[w].n_slots)
      hip_check(hipMemcpyAsync(ix.tbase.ptr + win_slot[w], wins[w].bases, wins[w].n_slots, hipMemcpyHostToDevice, st),
                "upload target bases");
  if (win_slot[nwin])
    k_classify<<<grid_for(win_slot[nwin], kBlock), kBlock, 0, st>>>(ix.tbase.ptr, win_slot[nwin]);
  target_tables(rt, ix, seq_slot);
  DevArray<uint32_t> d_win_stride, d_win_seq;
  d_win_stride.upload(win_stride.data(), win_stride.size(), st);
  d_win_seq.upload(win_seq.data(), win_seq.size(), st);

  const uint32_t n = static_cast<uint32_t>(ns);
  const int keybits = 2 * p.lmer;
  const uint64_t invalid = 1ull << keybits;
  uint32_t nvalid = 0, nent = 0;
  if (n > 0) {
    Workspace& ws = rt.ws();
    uint64_t* keys = ws.get<uint64_t>("ix.keys", n);
    uint32_t* vals = ws.get<uint32_t>("ix.vals", n);
    k_target_keys<<<grid_for(n, kBlock), kBlock, 0, st>>>(ix.tbase.ptr, ix.win_slot.ptr, d_win_stride.ptr, nwin,
                                                          p.lmer, p.stride, n, invalid, keys, vals);
    rt.check_launch("k_target_keys");
    radix_sort_pairs(rt, keys, vals, n, keybits + 1);
    uint32_t* head = ws.get<uint32_t>("ix.head", n);
    uint32_t* scal = ws.get<uint32_t>("ix.scalars", 4);
    k_fill<uint32_t><<<1, 64, 0, st>>>(scal, 4, 0u);
    k_index_heads<<<grid_for(n, kBlock), kBlock, 0, st>>>(keys, vals, d_win_stride.ptr, nwin, n, invalid, head, scal);
    rt.check_launch("k_index_heads");
    rt.read_scalars(scal, &nvalid, 1);
    if (nvalid > 0) {
      uint32_t* run_ex = ws.get<uint32_t>("ix.runex", nvalid);
      uint32_t* run_start = ws.get<uint32_t>("ix.runstart", nvalid + 1);
      device_scan<OpAdd, uint32_t, uint32_t, uint32_t>(rt, head, nvalid, run_ex, scal + 1, false);
      k_index_runstart<<<grid_for(nvalid, kBlock), kBlock, 0, st>>>(head, run_ex, nvalid, run_start);
      uint32_t* keep = ws.get<uint32_t>("ix.keep", nvalid);
      int32_t* tgs = ws.get<int32_t>("ix.tgs", nvalid);
      int32_t* tti = ws.get<int32_t>("ix.tti", nvalid);
      k_index_keep<<<grid_for(nvalid, kBlock), kBlock, 0, st>>>(
          vals, head, run_ex, run_start, nvalid, ix.win_slot.ptr, d_win_stride.ptr, d_win_seq.ptr, nwin,
          ix.seq_slot.ptr, ix.seq_len.ptr, p.lmer, p.stride, p.cut_repeat, keep, tgs, tti);
      rt.check_launch("k_index_keep");
      uint32_t* eoff = ws.get<uint32_t>("ix.eoff", nvalid);
      device_scan<OpAdd, uint32_t, uint32_t, uint32_t>(rt, keep, nvalid, eoff, scal + 2, false);
      rt.read_scalars(scal + 2, &nent, 1);
      ix.epos.alloc(nent);
      uint64_t* ek = ws.get<uint64_t>("ix.ekey", nent);
      k_index_scatter<<<grid_for(nvalid, kBlock), kBlock, 0, st>>>(keys, keep, eoff, tgs, tti, nvalid, ek,
                                                                    ix.epos.ptr);
      rt.check_launch("k_index_scatter");
      build_cells(rt, ix, ek, nent, keybits);
    }
  }
  ix.nentries = nent;
  if (nent == 0) build_cells(rt, ix, nullptr, 0, keybits);
  rt.sync();
}

// ---------------------------------------------------------------- query

// Per-strand tables of a query set as they are built (both builders).
struct QueryLayout {
  std::vector<uint64_t> win_slot{0};  // window w's slots: [win_slot[w], win_slot[w + 1])
  std::vector<uint32_t> qslot, pos_off{0};
  uint64_t npos = 0;
};

static void query_add_strand(const clast_params& p, clast_query_s& q, QueryLayout& L, uint64_t slot, int len,
                             int ref, int w) {
  if (slot >= (1ull << 32)) throw ClastError(CLAST_ERR_LIMIT, "query set over 2^32 slots");
  L.qslot.push_back(static_cast<uint32_t>(slot));
  q.h_qlen.push_back(len);
  q.h_qref.push_back(ref);
  q.h_qwin.push_back(w);
  q.max_qlen = std::max(q.max_qlen, len);
  L.npos += len >= p.lmer ? static_cast<uint64_t>(len - p.lmer + 1) : 0ull;
  if (L.npos >= (1ull << 32) - 1) throw ClastError(CLAST_ERR_LIMIT, "query set over 2^32 k-mer positions");
  L.pos_off.push_back(static_cast<uint32_t>(L.npos));
}

static void query_alloc_bases(clast_query_s& q, const QueryLayout& L) {
  const uint64_t slots = L.win_slot.back();
  if (slots >= (1ull << 32)) throw ClastError(CLAST_ERR_LIMIT, "query set over 2^32 slots");
  // 48 bytes of slack: k-mer lookups read the 9 aligned words that hold
  // each k-mer (kmer_words)
  q.qbase.alloc((padded_blocks(slots) + 3) * 16);
  q.bmax = static_cast<int32_t>(padded_blocks(slots) - 1);
}

// The tables after q.qbase holds class codes: strands, k-mer positions,
// 4-bit copy, clean flags.
static void query_finish(Runtime& rt, clast_query_s& q, const QueryLayout& L) {
  hipStream_t st = rt.stream();
  q.nseq = static_cast<uint32_t>(L.qslot.size());
  q.npos = static_cast<uint32_t>(L.npos);
  if (!q.qslot.ptr)  // (build_query_reads sets it and min_qlen first)
    q.min_qlen = q.h_qlen.empty() ? 0 : *std::min_element(q.h_qlen.begin(), q.h_qlen.end());
  if (!q.qslot.ptr) q.qslot.upload(L.qslot.data(), L.qslot.size(), st);  // (build_query_reads uploads it first)
  q.pos_off.upload(L.pos_off.data(), L.pos_off.size(), st);
  q.pos_strand.alloc(std::max<uint64_t>(L.npos, 1));
  if (q.nseq) k_pos_strand<<<grid_for(q.nseq, kBlock / 64), kBlock, 0, st>>>(q.pos_off.ptr, q.nseq, q.pos_strand.ptr);
  q.qlen.upload(q.h_qlen.data(), q.h_qlen.size(), st);
  q.qref.upload(q.h_qref.data(), q.h_qref.size(), st);
  q.qwin.upload(q.h_qwin.data(), q.h_qwin.size(), st);
  {
    const uint64_t ndw = 2 * (static_cast<uint64_t>(q.bmax) + 1);
    q.qnib.alloc(2 * ndw);  // with its mirror
    k_pack_nibbles<<<grid_for(ndw, kBlock), kBlock, 0, st>>>(q.qbase.ptr, ndw, 1, q.qnib.ptr);
  }
  q.qclean.alloc(std::max<uint64_t>(q.nseq, 1));
  if (q.nseq)
    k_clean_flags<uint32_t><<<grid_for(q.nseq, kBlock / 64), kBlock, 0, st>>>(q.qbase.ptr, q.qslot.ptr, q.qlen.ptr,
                                                                              q.nseq, 0, q.qclean.ptr);
  q.smin.alloc(std::max<uint64_t>(q.nseq, 1));
  rt.check_launch("query tables");
  rt.sync();
}

void build_query(const clast_params& p, const clast_window* wins, int nwin, clast_query_s& q) {
  Runtime& rt = runtime();
  std::lock_guard<std::mutex> lk(rt.mutex());
  hipStream_t st = rt.stream();
  q.device = rt.device();
  q.lmer = p.lmer;
  q.nwin = nwin;
  QueryLayout L;
  for (int w = 0; w < nwin; ++w) {
    check_window(wins[w]);
    if (wins[w].n_seq % 2) throw ClastError(CLAST_ERR_ARG, "a query window holds '+'/'-' pairs (even n_seq)");
    const uint64_t w0 = L.win_slot.back();
    L.win_slot.push_back(w0 + static_cast<uint64_t>(wins[w].n_slots));
    for (int k = 0; k < wins[w].n_seq; ++k)
      query_add_strand(p, q, L, w0 + static_cast<uint64_t>(wins[w].gateway[k]), wins[w].lengths[k],
                       wins[w].id_begin + k, w);
    q.h_id_begin.push_back(wins[w].id_begin);
  }
  query_alloc_bases(q, L);
  for (int w = 0; w < nwin; ++w)
    if (wins[w].n_slots)
      hip_check(hipMemcpyAsync(q.qbase.ptr + L.win_slot[w], wins[w].bases, wins[w].n_slots, hipMemcpyHostToDevice,
                               st),
                "upload query bases");
  if (L.win_slot.back())
    k_classify<<<grid_for(L.win_slot.back(), kBlock), kBlock, 0, st>>>(q.qbase.ptr, L.win_slot.back());
  query_finish(rt, q, L);
}

// Query windows from the reads as given: window w holds reads
// [win[2w], win[2w + 1]) (consecutive windows), each read as its '+' and
// '-' strands (list IDs id_begin[w] + 2 (k - win[2w]) + {0, 1}); the strands
// and joints are built on the device (k_encode_reads) from one upload of the
// reads' bytes.
void build_query_reads(const clast_params& p, const uint8_t* bases, const int64_t* offsets, const int32_t* win,
                       const int32_t* id_begin, int nwin, clast_query_s& q) {
  Runtime& rt = runtime();
  std::lock_guard<std::mutex> lk(rt.mutex());
  hipStream_t st = rt.stream();
  q.device = rt.device();
  q.lmer = p.lmer;
  q.nwin = nwin;
  const int32_t r0 = win[0], r1 = win[2 * nwin - 1];
  if (r0 < 0 || r1 < r0) throw ClastError(CLAST_ERR_ARG, "bad read windows");
  for (int w = 0; w < nwin; ++w) {
    if (win[2 * w] > win[2 * w + 1] || (w > 0 && win[2 * w] != win[2 * w - 1]))
      throw ClastError(CLAST_ERR_ARG, "read windows must be consecutive ranges in order");
  }
  const int64_t joint = std::max(p.lmer - 1, 0);
  const uint32_t nreads = static_cast<uint32_t>(r1 - r0);
  const size_t nstr = 2 * static_cast<size_t>(nreads);
  // the reads' bytes go up (a copy from pageable memory blocks its thread)
  // while this thread lays out the strand tables; `up` is joined before
  // dreads goes (declared after it), also when the layout throws
  DevArray<uint8_t> dreads;
  std::future<void> up;
  if (nreads) {
    const int dev = rt.device();
    up = std::async(std::launch::async, [&dreads, bases, offsets, r0, r1, dev, st] {
      hip_check(hipSetDevice(dev), "hipSetDevice");
      dreads.upload(bases + offsets[r0], static_cast<uint64_t>(offsets[r1] - offsets[r0]), st);
    });
  }
  // one pass over the reads into arrays sized up front: read k's '+' strand
  // is strand 2 (k - r0), its '-' strand the next
  QueryLayout L;
  L.qslot.resize(nstr);
  L.pos_off.resize(nstr + 1);
  L.win_slot.reserve(static_cast<size_t>(nwin) + 1);
  q.h_qlen.resize(nstr);
  q.h_qref.resize(nstr);
  q.h_qwin.resize(nstr);
  q.h_id_begin.assign(id_begin, id_begin + nwin);
  uint64_t slot = 0, npos = 0;
  int max_len = 0, min_len = INT32_MAX;
  size_t si = 0;
  for (int w = 0; w < nwin; ++w) {
    for (int32_t k = win[2 * w]; k < win[2 * w + 1]; ++k) {
      const int64_t len = offsets[k + 1] - offsets[k];
      if (len < 0 || len > INT32_MAX - joint) throw ClastError(CLAST_ERR_ARG, "bad read offsets");
      const uint64_t span = static_cast<uint64_t>(len + joint);
      if (slot + span >= (1ull << 32)) throw ClastError(CLAST_ERR_LIMIT, "query set over 2^32 slots");
      const uint64_t kpos = len >= p.lmer ? static_cast<uint64_t>(len - p.lmer + 1) : 0ull;
      if (npos + 2 * kpos >= (1ull << 32) - 1) throw ClastError(CLAST_ERR_LIMIT, "query set over 2^32 k-mer positions");
      const int32_t ref = id_begin[w] + 2 * (k - win[2 * w]);
      for (int h = 0; h < 2; ++h, ++si) {
        L.qslot[si] = static_cast<uint32_t>(slot + h * span);
        q.h_qlen[si] = static_cast<int32_t>(len);
        q.h_qref[si] = ref + h;
        q.h_qwin[si] = w;
        npos += kpos;
        L.pos_off[si + 1] = static_cast<uint32_t>(npos);
      }
      max_len = std::max(max_len, static_cast<int>(len));
      min_len = std::min(min_len, static_cast<int>(len));
      slot += 2 * span;
    }
    L.win_slot.push_back(slot);
  }
  L.npos = npos;
  q.max_qlen = max_len;
  q.min_qlen = nreads ? min_len : 0;
  query_alloc_bases(q, L);
  DevArray<int64_t> doff;
  if (nreads) {
    const int64_t b0 = offsets[r0];
    up.get();  // the bytes are on the stream before the kernel
    doff.upload(offsets + r0, static_cast<uint64_t>(nreads) + 1, st);
    // the '+' strand slots (every other entry of the strand slots)
    q.qslot.upload(L.qslot.data(), L.qslot.size(), st);
    k_encode_reads<<<grid_for(nreads, kBlock / 64), kBlock, 0, st>>>(dreads.ptr, doff.ptr, b0, nreads, q.qslot.ptr,
                                                                     p.lmer, q.qbase.ptr);
  }
  query_finish(rt, q, L);  // ends with a sync: the uploads are done before they are freed
}

void ensure_smin(Runtime& rt, const clast_params& p, clast_query_s& q) {
  const double key[4] = {p.cut_off, p.total_db_size, p.K, p.lambda};
  std::lock_guard<std::mutex> lk(q.smin_mu);
  if (q.smin_valid && std::memcmp(key, q.smin_key, sizeof key) == 0) return;
  std::vector<int32_t> sm(q.nseq);
  if (q.max_qlen <= (1 << 20)) {
    // one threshold per length up to the longest strand, each found once
    std::vector<int32_t> table(static_cast<size_t>(q.max_qlen) + 1);
    std::vector<uint8_t> have(table.size(), 0);
    for (uint32_t s = 0; s < q.nseq; ++s) {
      const int len = std::max(q.h_qlen[s], 0);
      if (!have[len]) {
        table[len] = score_threshold(p, len);
        have[len] = 1;
      }
      sm[s] = table[len];
    }
  } else {
    std::unordered_map<int, int32_t> memo;
    for (uint32_t s = 0; s < q.nseq; ++s) {
      const int len = q.h_qlen[s];
      auto it = memo.find(len);
      if (it == memo.end()) it = memo.emplace(len, score_threshold(p, len)).first;
      sm[s] = it->second;
    }
  }
  if (q.nseq)
    hip_check(hipMemcpyAsync(q.smin.ptr, sm.data(), sizeof(int32_t) * q.nseq, hipMemcpyHostToDevice, rt.stream()),
              "upload smin");
  rt.sync();
  std::memcpy(q.smin_key, key, sizeof key);
  q.smin_valid = true;
}

template <class T>
void d2h(std::vector<T>& dst, const T* src, uint64_t n, hipStream_t st) {
  dst.resize(n);
  if (n) hip_check(hipMemcpyAsync(dst.data(), src, sizeof(T) * n, hipMemcpyDeviceToHost, st), "D2H");
}

// Event slots: kernel pairs have nothing but the kernel between them, so
// their elapsed time is the kernel's duration on the library stream.
enum {
  EV_START, EV_COUNT0, EV_COUNT1, EV_WRITE0, EV_WRITE1, EV_SORT0, EV_SORT1, EV_FILT0, EV_FILT1, EV_DP0, EV_DP1,
  EV_RES0, EV_RES1,
  // the three DP kernels, each bracketed on the stream it runs on
  EV_D16_0, EV_D16_1, EV_D32_0, EV_D32_1, EV_DBAND0, EV_DBAND1, EV_E16_0, EV_E16_1,
  // the second pass of a two-pass DP stage (side stream)
  EV_P2_0, EV_P2_1, kEvents
};

struct StageEvents {
  hipEvent_t e[kEvents];
  StageEvents() {
    for (auto& x : e) hip_check(hipEventCreate(&x), "hipEventCreate");
  }
  ~StageEvents() {
    for (auto& x : e) (void)hipEventDestroy(x);
  }
};

// The CDeviceHitList chain (CDeviceHitList.cu:228-423) for every unit of
// idx x q, rows in fixResult order (CHostResultHolder.cu:75-128): query
// pair, score descending, target window, then each unit's own order.  For a
// single unit that is the unit's CDeviceHitList order.
//
// Seeds of a query strand are contiguous after the count scan, so the sort
// by (qID, tID, diagonal, qIdx) and the isolate / duplicate filters run per
// strand in LDS (k_seed_segments, one wavefront per strand).  Strands with
// more seeds than fit take the global path: radix sort of (strand, target
// sequence, diagonal) keys and scan-based filters.  CLAST_KEEP_STAGES sends
// every strand through the global path so its intermediate arrays can be
// read back.
// Caller-owned host arrays that an allocation-free search writes its rows
// into (clast_search_batch_noalloc): no stage events, no pinned block, no
// copy stream.
struct CallerRows {
  int64_t capacity;
  int32_t* col[10];  // tID tIdx qID qIdx tHL qHL match score twin qwin (null: skipped)
  double* evalue;
  int64_t needed = 0;
};

void run_search(const clast_params& p, clast_index_s& ix, clast_query_s& q, int flags,
                clast_hits_s& out, CallerRows* rows_out = nullptr) {
  if (ix.lmer != p.lmer || q.lmer != p.lmer || ix.stride != p.stride || ix.cut_repeat != p.cut_repeat)
    throw ClastError(CLAST_ERR_ARG, "params differ from those the index/query were built with");
  if (ix.device != q.device) throw ClastError(CLAST_ERR_ARG, "index and query on different devices");
  const DPForced forced = DPForced::from_env();  // read (and checked) before anything is queued
  set_device(ix.device);
  Runtime& rt = runtime();
  std::lock_guard<std::mutex> lk(rt.mutex());
  hipStream_t st = rt.stream();
  Workspace& ws = rt.ws();
  const bool keep_stages = (flags & CLAST_KEEP_STAGES) != 0;
  const bool use_cut = p.cut_off != -1;
  if (use_cut) ensure_smin(rt, p, q);
  std::unique_ptr<StageEvents> ev(rows_out ? nullptr : new StageEvents());
  bool timed[kEvents] = {};
  auto mark_on = [&](int k, hipStream_t on) {
    if (!ev) return;
    hip_check(hipEventRecord(ev->e[k], on), "event");
    timed[k] = true;
  };
  auto mark = [&](int k) { mark_on(k, st); };
  mark(EV_START);

  const uint32_t np = q.npos;
  const int sbits = bits_for(q.nseq), gbits = bits_for(static_cast<uint64_t>(ix.nseq));
  // diagonal + d_off in [0, max_tlen + max_qlen]; one code above stays unused
  const int dbits = bits_for(static_cast<uint64_t>(ix.max_tlen) + q.max_qlen + 2);
  const int d_off = q.max_qlen;
  const uint64_t nunits = static_cast<uint64_t>(ix.nwin) * q.nwin;
  // [2, 10): DP cell counters, [10, 18): rows of 16-bit DP tasks, [18]: their
  // count, [19]: long task count, [20, 28): their rows, [28, 36): all DP rows,
  // [36, 44): rows of the second DP pass's tasks, [44, 52): the same rows
  // summed before they are listed (the second pass's long-task threshold)
  uint64_t* tot64 = ws.get<uint64_t>("s.tot64", 52);
  k_fill<uint64_t><<<1, 64, 0, st>>>(tot64, 52, 0ull);
  // [8, 11): task counts of the second DP pass's three lists
  uint32_t* tot32 = ws.get<uint32_t>("s.tot32", 16);
  uint32_t* unit_first = ws.get<uint32_t>("s.unit", nunits);
  k_fill<uint32_t><<<grid_for(nunits, kBlock), kBlock, 0, st>>>(unit_first, nunits, 0xFFFFFFFFu);

  SeedArgs a{};
  a.qbase = q.qbase.ptr;
  a.qslot = q.qslot.ptr;
  a.qlen = q.qlen.ptr;
  a.qwin = q.qwin.ptr;
  a.pos_off = q.pos_off.ptr;
  a.pos_strand = q.pos_strand.ptr;
  a.nseq = q.nseq;
  a.npos = np;
  a.lk = Lookup{ix.bitmap.ptr, ix.rank.ptr, ix.cells.ptr,
                ix.key64 ? static_cast<const void*>(ix.ekey64.ptr) : static_cast<const void*>(ix.ekey32.ptr),
                ix.epos.ptr, ix.shift, ix.key64 ? 1 : 0, ix.direct ? 1 : 0};
  a.epos = ix.epos.ptr;
  a.tinfo = ix.seq_info.ptr;
  a.nqwin = q.nwin;
  a.lmer = p.lmer;
  a.gap = p.gap;
  a.width = p.width;
  a.local = p.local;
  a.unit_first = unit_first;

  // lookup + count: raw seed count per strand, strands' seed ranges (the
  // seed total sizes the seed and hit arrays: the search's first read-back)
  uint32_t nseed = 0;
  uint4* cellp = nullptr;
  uint32_t* scnt = ws.get<uint32_t>("s.scnt", q.nseq);
  uint32_t* soff = ws.get<uint32_t>("s.soff", q.nseq);
  if (np > 0 && ix.nentries > 0) {
    cellp = ws.get<uint4>("s.cell", np);
    k_fill<uint32_t><<<grid_for(q.nseq, kBlock), kBlock, 0, st>>>(scnt, q.nseq, 0u);
    mark(EV_COUNT0);
    k_seed_count<<<grid_for(np, kBlock), kBlock, 0, st>>>(a, cellp, scnt);
    mark(EV_COUNT1);
    rt.check_launch("k_seed_count");
    device_scan<OpAdd, uint32_t, uint64_t, uint32_t>(rt, scnt, q.nseq, soff, tot64, false);
    uint64_t total = 0;
    rt.read_scalars(tot64, &total, 1);
    if (total >= (1ull << 32) - 1) throw ClastError(CLAST_ERR_LIMIT, "over 2^32 seeds: split the query set");
    nseed = static_cast<uint32_t>(total);
  }

  // per-strand sort + isolate/duplicate in LDS, the global path for strands
  // over kSegCap seeds; hits are first written into their strand's seed
  // range, then compacted in strand order
  uint32_t nhit = 0;
  const uint64_t cap = std::max<uint64_t>(nseed, 1);
  int32_t* hs = ws.get<int32_t>("h.s", cap);
  int32_t* hgs = ws.get<int32_t>("h.gs", cap);
  int32_t* ht = ws.get<int32_t>("h.t", cap);
  int32_t* hq = ws.get<int32_t>("h.q", cap);
  if (nseed > 0) {
    int32_t* ss = ws.get<int32_t>("h.ss", cap);
    int32_t* sgs = ws.get<int32_t>("h.sgs", cap);
    int32_t* stt = ws.get<int32_t>("h.st", cap);
    int32_t* sq = ws.get<int32_t>("h.sq", cap);
    uint32_t* hcnt = ws.get<uint32_t>("h.cnt", q.nseq);
    uint32_t* hbase = ws.get<uint32_t>("h.base", q.nseq);
    uint32_t* ovf_list = ws.get<uint32_t>("s.ovf", q.nseq);
    uint8_t* ovf_flag = ws.get<uint8_t>("s.ovfflag", q.nseq);
    uint32_t* cnt2 = ws.get<uint32_t>("g.cnt", q.nseq);
    uint32_t* off2 = ws.get<uint32_t>("g.off", q.nseq);
    uint32_t* fmark = ws.get<uint32_t>("s.fmark", q.nseq);
    uint32_t* flist = ws.get<uint32_t>("s.flist", nunits);
    uint32_t novf = 0;
    k_fill<uint32_t><<<1, 64, 0, st>>>(tot32, 16, 0u);
    const int qbits = bits_for(static_cast<uint64_t>(q.max_qlen));
    const int wbits = bits_for(static_cast<uint64_t>(ix.nwin));
    const bool packed = wbits + gbits + dbits + qbits <= 63;
    const SegPack sp{d_off, qbits, dbits, gbits};
    uint32_t* big_list = ws.get<uint32_t>("s.big", q.nseq);
    const SegOut so{ss, sgs, stt, sq, hcnt, big_list, tot32 + 6, ovf_list, tot32 + 1, ovf_flag};
    // nwaves: waves to launch (at most one per listed strand)
    auto segments = [&](bool small, uint32_t nwaves, const uint32_t* list, const uint32_t* nlist, int mode) {
      const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(grid_for(nwaves, kSegWaves), 1u << 20));
      if (small) {
        if (packed) k_seed_segments<true, kSegSmall><<<grid, kBlock, 0, st>>>(a, cellp, scnt, soff, sp, so, list, nlist, mode);
        else k_seed_segments<false, kSegSmall><<<grid, kBlock, 0, st>>>(a, cellp, scnt, soff, sp, so, list, nlist, mode);
      } else {
        if (packed) k_seed_segments<true, kSegBig><<<grid, kBlock, 0, st>>>(a, cellp, scnt, soff, sp, so, list, nlist, mode);
        else k_seed_segments<false, kSegBig><<<grid, kBlock, 0, st>>>(a, cellp, scnt, soff, sp, so, list, nlist, mode);
      }
      rt.check_launch("k_seed_segments");
    };
    // unit_first is final: run each unit's first strand again with its
    // element 0 (gate: nothing while *gate, i.e. before the global path)
    auto rerun = [&](const uint32_t* gate) {
      k_fill<uint32_t><<<grid_for(q.nseq, kBlock), kBlock, 0, st>>>(fmark, q.nseq, 0u);
      k_fill<uint32_t><<<1, 64, 0, st>>>(tot32 + 5, 1, 0u);
      k_fix_list<<<grid_for(nunits, kBlock), kBlock, 0, st>>>(unit_first, static_cast<uint32_t>(nunits), ovf_flag,
                                                              fmark, flist, tot32 + 5, gate);
      segments(false, static_cast<uint32_t>(nunits), flist, tot32 + 5, kSegRerun);
    };
    uint64_t t2 = 0;
    bool hits_known = false;
    mark(EV_WRITE0);
    if (!keep_stages) {
      k_fill<uint8_t><<<grid_for(q.nseq, kBlock), kBlock, 0, st>>>(ovf_flag, q.nseq, 0);
      // unit candidates as per-strand window masks when the windows fit 64 bits
      if (ix.nwin <= 64) {
        a.wmask = ws.get<uint64_t>("s.wmask", q.nseq);
        k_fill<uint64_t><<<grid_for(q.nseq, kBlock), kBlock, 0, st>>>(a.wmask, q.nseq, 0ull);
      }
      segments(true, q.nseq, nullptr, nullptr, kSegFirst);
      // the big list's length stays on the device: a bounded grid strides over it
      segments(false, std::min<uint32_t>(q.nseq, 4096u * kSegWaves), big_list, tot32 + 6, kSegBigPass);
      if (a.wmask) {
        k_unit_first<<<grid_for(q.nseq, kBlock), kBlock, 0, st>>>(a.wmask, q.qwin.ptr, q.nseq, ix.nwin, q.nwin,
                                                                 unit_first);
        rt.check_launch("k_unit_first");
        a.wmask = nullptr;  // the later passes (global path, rerun) read unit_first
      }
      // with no strand on the global path (the usual case) unit_first is
      // final now: the re-run and the hit count go ahead of the read-back,
      // and a nonzero novf (tot32[1]) stops the re-run's list
      rerun(tot32 + 1);
      device_scan<OpAdd, uint32_t, uint32_t, uint32_t>(rt, hcnt, q.nseq, hbase, tot32 + 3, false);
      mark(EV_WRITE1);
      // the global path's seeds per strand (zero without such strands)
      k_ovf_strand_counts<<<grid_for(q.nseq, kBlock), kBlock, 0, st>>>(ovf_flag, scnt, q.nseq, cnt2);
      device_scan<OpAdd, uint32_t, uint64_t, uint32_t>(rt, cnt2, q.nseq, off2, tot64 + 1, false);
      // one read-back: the global path's seeds and strands, and the hits
      uint32_t r32[4];
      rt.read_regions({{tot64 + 1, &t2, sizeof t2}, {tot32, r32, sizeof r32}});
      novf = r32[1];
      if (novf == 0) {
        nhit = r32[3];
        hits_known = true;
      }
    } else {
      // every strand on the global path: list them all
      k_iota_u32<<<grid_for(q.nseq, kBlock), kBlock, 0, st>>>(ovf_list, q.nseq);
      k_fill<uint8_t><<<grid_for(q.nseq, kBlock), kBlock, 0, st>>>(ovf_flag, q.nseq, 1);
      k_fill<uint32_t><<<grid_for(q.nseq, kBlock), kBlock, 0, st>>>(hcnt, q.nseq, 0u);
      novf = q.nseq;
      k_ovf_strand_counts<<<grid_for(q.nseq, kBlock), kBlock, 0, st>>>(ovf_flag, scnt, q.nseq, cnt2);
      device_scan<OpAdd, uint32_t, uint64_t, uint32_t>(rt, cnt2, q.nseq, off2, tot64 + 1, false);
      rt.read_scalars(tot64 + 1, &t2, 1);
    }
    SeedPack pk{gbits, dbits, d_off};
    SeedDecode dc{gbits, dbits, d_off};
    uint32_t ng = 0;
    uint64_t* skey = nullptr;
    uint32_t* sval = nullptr;
    if (novf > 0) {
      if (sbits + gbits + dbits > 64) throw ClastError(CLAST_ERR_LIMIT, "seed key over 64 bits: split the query set");
      // the listed strands' seeds in strand order (off2[s]: strand s's first
      // seed, before and after the sort, which keys on the strand first)
      if (t2 >= (1ull << 32)) throw ClastError(CLAST_ERR_LIMIT, "over 2^32 seeds: split the query set");
      ng = static_cast<uint32_t>(t2);
      skey = ws.get<uint64_t>("g.key", std::max<uint32_t>(ng, 1));
      sval = ws.get<uint32_t>("g.val", std::max<uint32_t>(ng, 1));
      // at most 2^16 workgroups, striding over the listed strands
      k_seed_write<<<std::min<uint32_t>(novf, 1u << 16), kWriteBlock, 0, st>>>(a, pk, ovf_list, novf, cellp, off2,
                                                                               ix.nwin, skey, sval);
      rt.check_launch("k_seed_write");
      if (!keep_stages) rerun(nullptr);
      mark(EV_WRITE1);
      auto dump_seeds = [&](int slot) {
        std::vector<uint64_t> hk;
        std::vector<uint32_t> hv;
        d2h(hk, skey, ng, st);
        d2h(hv, sval, ng, st);
        rt.sync();
        auto& S = out.stage[slot];
        for (int c = 0; c < 4; ++c) S[c].clear();
        const uint64_t low = (1ull << (gbits + dbits)) - 1ull;
        for (uint32_t i = 0; i < ng; ++i) {
          const uint64_t k = hk[i];
          if ((k & low) == low) continue;  // removed by the corner filter
          const uint32_t s = static_cast<uint32_t>(k >> (gbits + dbits));
          const uint32_t gs = static_cast<uint32_t>((k >> dbits) & ((1ull << gbits) - 1ull));
          const int d = static_cast<int>(k & ((1ull << dbits) - 1ull)) - d_off;
          S[0].push_back(ix.h_seq_ref[gs]);
          S[1].push_back(d + static_cast<int>(hv[i]));
          S[2].push_back(q.h_qref[s]);
          S[3].push_back(static_cast<int>(hv[i]));
        }
        out.stage_n[slot] = static_cast<int64_t>(S[0].size());
      };
      if (keep_stages) dump_seeds(CLAST_STAGE_SEEDS);
      if (ng > 1) {
        // the sort's slots grow before timing so no sync lands inside
        (void)ws.get<uint64_t>("sort.k2", ng);
        (void)ws.get<uint32_t>("sort.v2", ng);
        (void)ws.get<uint32_t>("sort.counts", static_cast<uint64_t>(grid_for(ng, kSortTile)) * kRadixBins);
        mark(EV_SORT0);
        radix_sort_pairs(rt, skey, sval, ng, sbits + gbits + dbits);
        mark(EV_SORT1);
      }
      if (keep_stages) dump_seeds(CLAST_STAGE_SORTED);
      k_fill<uint32_t><<<1, 64, 0, st>>>(tot32 + 2, 1, 0u);
      if (ng > 0) {
        mark(EV_FILT0);
        uint32_t* kidx = ws.get<uint32_t>("g.kidx", ng);
        uint32_t* gidx = ws.get<uint32_t>("g.gidx", ng);
        uint32_t* lastk = ws.get<uint32_t>("g.lastk", ng);
        uint32_t* gstart = ws.get<uint32_t>("g.gstart", ng);
        uint32_t* dflag = ws.get<uint32_t>("g.dflag", ng);
        uint32_t* hoff = ws.get<uint32_t>("g.hoff", ng);
        const uint32_t g = grid_for(ng, kBlock);
        k_seed_marks<<<g, kBlock, 0, st>>>(skey, sval, ng, dc, p.width, p.gap, ix.seq_win.ptr, q.qwin.ptr, q.nwin,
                                           unit_first, kidx, gidx);
        rt.check_launch("k_seed_marks");
        device_scan<OpMax, uint32_t, uint32_t, uint32_t>(rt, kidx, ng, lastk, static_cast<uint32_t*>(nullptr), true);
        device_scan<OpMax, uint32_t, uint32_t, uint32_t>(rt, gidx, ng, gstart, static_cast<uint32_t*>(nullptr), true);
        k_seed_dup<<<g, kBlock, 0, st>>>(skey, sval, ng, dc, p.width, p.gap, kidx, lastk, gstart, dflag);
        rt.check_launch("k_seed_dup");
        device_scan<OpAdd, uint32_t, uint32_t, uint32_t>(rt, dflag, ng, hoff, tot32 + 2, false);
        k_hits_scatter<<<g, kBlock, 0, st>>>(skey, sval, ng, dc, dflag, hoff, soff, off2, ss, sgs, stt, sq);
        rt.check_launch("k_hits_scatter");
        // per-strand hit counts of the listed strands (their total read on device)
        k_ovf_hcnt<<<grid_for(novf, kBlock), kBlock, 0, st>>>(ovf_list, novf, q.nseq, off2, ng, hoff, tot32 + 2, hcnt);
        mark(EV_FILT1);
      } else {
        k_ovf_hcnt<<<grid_for(novf, kBlock), kBlock, 0, st>>>(ovf_list, novf, q.nseq, off2, 0, hcnt, tot32 + 2, hcnt);
      }
    }
    if (!hits_known) {
      device_scan<OpAdd, uint32_t, uint32_t, uint32_t>(rt, hcnt, q.nseq, hbase, tot32 + 3, false);
      rt.read_scalars(tot32 + 3, &nhit, 1);
    }
    if (nhit)
      k_hits_compact<<<grid_for(nhit, kBlock), kBlock, 0, st>>>(nhit, q.nseq, soff, hbase, ss, sgs, stt, sq, hs, hgs,
                                                                 ht, hq);
    rt.check_launch("k_hits_compact");
  }
  if (keep_stages) {
    std::vector<int32_t> va, vb, vc, vd;
    d2h(va, hs, nhit, st); d2h(vb, hgs, nhit, st); d2h(vc, ht, nhit, st); d2h(vd, hq, nhit, st);
    rt.sync();
    auto& S = out.stage[CLAST_STAGE_HITS];
    for (int k = 0; k < 4; ++k) S[k].resize(nhit);
    for (uint32_t i = 0; i < nhit; ++i) {
      S[0][i] = ix.h_seq_ref[vb[i]]; S[1][i] = vc[i]; S[2][i] = q.h_qref[va[i]]; S[3][i] = vd[i];
    }
    out.stage_n[CLAST_STAGE_HITS] = nhit;
  }

  // banded DP
  int32_t *ot = nullptr, *oq = nullptr, *thl = nullptr, *qhl = nullptr, *om = nullptr, *osc = nullptr;
  int4* dres = nullptr;
  // two-pass DP stage (tuner.hpp): the first pass's arguments and task
  // order, which the second pass reads
  struct SecondPass {
    bool on = false;
    DPArgs da{};
    uint32_t nslot = 0;
    DPTuner::Key key{0, 0, 0, 0};
  } sp;
  if (nhit) {
    ot = ws.get<int32_t>("d.t", nhit); oq = ws.get<int32_t>("d.q", nhit);
    thl = ws.get<int32_t>("d.thl", nhit); qhl = ws.get<int32_t>("d.qhl", nhit);
    om = ws.get<int32_t>("d.m", nhit); osc = ws.get<int32_t>("d.s", nhit);
    // one task per (hit, direction), longest first; the ordering chain runs
    // over `nslot` task slots, the task count rounded up to a quarter of its
    // power of two, so that a captured chain serves every search of that size
    const uint32_t ntask = 2 * nhit;
    const uint32_t qstep = std::max<uint32_t>(256u, (1u << (31 - __builtin_clz(ntask))) / 4u);
    const uint32_t nslot = (ntask + qstep - 1) / qstep * qstep;
    uint64_t* dk = ws.get<uint64_t>("d.key", nslot);
    uint32_t* dporder = ws.get<uint32_t>("d.order", nslot);
    dres = ws.get<int4>("d.res", ntask);
    const int maxrows = std::max(q.max_qlen - p.lmer, 0);
    const int rowbits = bits_for(static_cast<uint64_t>(maxrows) + 1);
    const bool packed = !p.local && !dp16_disabled();
    const DPTuner::Key tkey{p.gap, p.local ? 1 : 0, dp_bucket(maxrows), dp_size_class(ntask)};
    // (the allocation-free entry point and the stage read-back run one pass)
    sp.on = packed && use_cut && !keep_stages && !rows_out && DPTuner::get().two_pass(tkey, forced.two_pass);
    auto* const launch16 = sp.on ? launch_dp16_score : launch_dp16;
    // a large score pass with long extensions holds its slots for long:
    // it leaves a quarter of them to the kernels beside it and to other
    // searches (searches of short reads keep the full occupancy)
    const bool reserve16 = dp_size_class(ntask) == 1 && maxrows >= 256;
    uint32_t* split = ws.get<uint32_t>("d.split", 3);
    int* band_rows = ws.get<int>("d.bandrows", 1);
    uint32_t* chunk_next = ws.get<uint32_t>("d.chunk", 1);
    TaskKeyArgs* kargs = ws.get<TaskKeyArgs>("d.kargs", 1);
    const SortSlots dsl{"dsort.k2", "dsort.v2", "dsort.counts", "dscan.tiles"};
    const bool use_graph = !graphs_disabled() && !rows_out;
    const uint32_t cu = static_cast<uint32_t>(std::max(rt.compute_units(), 1));
    hipStream_t ss = rt.side_stream();
    // The DP stage under configuration c (tuner.hpp); `final` marks the run
    // whose kernels are timed and whose counters are kept.
    auto dp_chain = [&](const DPConfig& c, bool final) -> DPArgs {
      auto fmark = [&](int k) {
        if (final) mark(k);
      };
      auto fmark_on = [&](int k, hipStream_t on) {
        if (final) mark_on(k, on);
      };
      const int bmode = maxrows >= kBandRows ? c.band : 0;
      const bool band = bmode != 0;
      // the class bits on top of the rows, inside the radix passes the rows
      // need (the row key drops its low bits when they do not fit)
      const int d = c.radix;
      const int sortbits = std::max(d, (rowbits + d - 1) / d * d);
      const int classbit = sortbits - 2;  // four classes (k_dp_task_keys)
      const int rowshift = std::max(0, rowbits - classbit);
      const TaskKeyArgs ka{nhit, hs, hgs, ht, hq, q.qlen.ptr, q.qclean.ptr, ix.seq_len.ptr, ix.seq_win.ptr,
                           ix.seq_slot.ptr, ix.win_slot.ptr, ix.seq_clean.ptr, p.lmer, p.gap, packed ? 1 : 0,
                           bmode, band_rows, maxrows, rowshift, classbit, dk, dporder,
                           reinterpret_cast<unsigned long long*>(tot64 + 10),
                           reinterpret_cast<unsigned long long*>(tot64 + 20)};
      k_store<TaskKeyArgs><<<1, 64, 0, st>>>(kargs, ka);
      auto* rsum = reinterpret_cast<unsigned long long*>(tot64 + 28);
      // the task-ordering chain: kernels only, every per-search value read
      // from kargs, sized by nslot
      auto order_chain = [&] {
        k_fill<uint64_t><<<1, 64, 0, st>>>(tot64 + 10, 26, 0ull);  // this stage's counters
        if (bmode == 2) {
          k_dp_rows_sum<<<grid_for(nslot, kBlock), kBlock, 0, st>>>(kargs, rsum);
          k_band_threshold<<<1, 64, 0, st>>>(rsum, 4ull * cu * 128ull, band_rows);
        } else if (bmode == 1) {
          k_fill<int><<<1, 64, 0, st>>>(band_rows, 1, kBandRows);
        }
        k_dp_task_keys<<<grid_for(nslot, kBlock), kBlock, 0, st>>>(kargs, nslot);
        radix_sort_pairs(rt, dk, dporder, nslot, sortbits, d, dsl);
        for (int k = 0; k < 3; ++k)
          k_count_below<<<1, 64, 0, st>>>(dk, nslot, static_cast<uint64_t>(k + 1) << classbit, split + k, nullptr);
        k_split_counts<<<1, 64, 0, st>>>(split, reinterpret_cast<unsigned long long*>(tot64 + 18));
        if (c.sched == 1) k_fill<uint32_t><<<1, 64, 0, st>>>(chunk_next, 1, 0u);
        rt.check_launch("DP task order");
      };
      if (!use_graph) {
        order_chain();
      } else {
        // the sort's slots at full size first: no workspace growth (which
        // synchronises and allocates) may happen while capturing
        const uint64_t ntiles = grid_for(nslot, kSortTile);
        uint64_t* k2 = ws.get<uint64_t>(dsl.k2, nslot);
        uint32_t* v2 = ws.get<uint32_t>(dsl.v2, nslot);
        uint32_t* cnts = ws.get<uint32_t>(dsl.counts, ntiles * kRadixBins);
        uint32_t* tiles = ws.get<uint32_t>(dsl.tiles, (ntiles * kRadixBins + kScanTile - 1) / kScanTile + 1);
        const std::vector<uint64_t> key{
            static_cast<uint64_t>(bmode), static_cast<uint64_t>(sortbits), static_cast<uint64_t>(d),
            static_cast<uint64_t>(c.sched), nslot, static_cast<uint64_t>(cu),
            reinterpret_cast<uint64_t>(kargs), reinterpret_cast<uint64_t>(dk), reinterpret_cast<uint64_t>(dporder),
            reinterpret_cast<uint64_t>(split), reinterpret_cast<uint64_t>(band_rows),
            reinterpret_cast<uint64_t>(chunk_next), reinterpret_cast<uint64_t>(tot64), reinterpret_cast<uint64_t>(k2),
            reinterpret_cast<uint64_t>(v2), reinterpret_cast<uint64_t>(cnts), reinterpret_cast<uint64_t>(tiles)};
        hipGraphExec_t ex = rt.graph(key);
        if (!ex) {
          hipGraph_t g = nullptr;
          hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
          try {
            order_chain();
          } catch (...) {
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
          }
          hip_check(hipStreamEndCapture(st, &g), "hipStreamEndCapture");
          const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
          (void)hipGraphDestroy(g);
          hip_check(e, "hipGraphInstantiate");
          rt.add_graph(key, ex);
        }
        hip_check(hipGraphLaunch(ex, st), "hipGraphLaunch");
      }
      DPArgs da{ntask, hs, hgs, ht, hq, q.qbase.ptr, q.qslot.ptr, q.qlen.ptr, ix.tbase.ptr, ix.seq_slot.ptr,
                ix.seq_len.ptr, ix.seq_win.ptr, ix.win_slot.ptr, dporder, split, q.qnib.ptr, ix.tnib.ptr, dres,
                q.bmax, ix.bmax, p.lmer, 4 * cu, chunk_next, ot, oq, thl, qhl, om, osc, q.qclean.ptr,
                ix.seq_clean.ptr};
      fmark(EV_DP0);
      // the long tasks (raised wave priority), the 16-bit tasks past a
      // sequence end and the int32 tasks run on the side stream beside the
      // 16-bit ones, so the kernels' tails overlap
      if (packed) {
        rt.fork(0);
        if (band) {
          fmark_on(EV_DBAND0, ss);
          launch_dp_band(p.gap, false, da, ss);
          fmark_on(EV_DBAND1, ss);
        }
        fmark_on(EV_E16_0, ss);
        launch16(p.gap, true, false, 0, false, da, ss);
        fmark_on(EV_E16_1, ss);
        fmark_on(EV_D32_0, ss);
        launch_dp_global(p.gap, da, ss);
        fmark_on(EV_D32_1, ss);
        fmark(EV_D16_0);
        launch16(p.gap, false, c.sched == 1, 4 * cu, reserve16, da, st);
        fmark(EV_D16_1);
        rt.join(0);
      } else {
        if (band) {
          rt.fork(0);
          fmark_on(EV_DBAND0, ss);
          launch_dp_band(p.gap, p.local != 0, da, ss);
          fmark_on(EV_DBAND1, ss);
        }
        fmark(EV_D32_0);
        if (p.local) launch_dp_local(p.gap, da, st);
        else launch_dp_global(p.gap, da, st);
        fmark(EV_D32_1);
        if (band) rt.join(0);
      }
      fmark(EV_DP1);
      rt.check_launch("k_dp");
      return da;
    };
    // configuration: the tuner's for this kind of search, found now by timed
    // trials (coordinate descent over sched, band, radix; every trial writes
    // every task's result, so the trials change no row) if it has none yet
    DPConfig cfg;
    const bool tuned = DPTuner::get().lookup(tkey, &cfg);
    cfg = forced.apply(tuned ? cfg : DPConfig{});
    // the persistent schedule runs gaps up to kPersistMaxGap only
    if (cfg.sched == 1 && p.gap > kPersistMaxGap)
      throw ClastError(CLAST_ERR_ARG, "CLAST_DP_SCHED=persist takes gaps up to " + std::to_string(kPersistMaxGap));
    // (the allocation-free entry point runs no trials: they create events)
    if (!tuned && forced.tune && !rows_out) {
      hipEvent_t e0, e1;
      hip_check(hipEventCreate(&e0), "hipEventCreate");
      hip_check(hipEventCreate(&e1), "hipEventCreate");
      std::vector<std::pair<DPConfig, float>> trials;
      // a trial's time is the least of kTrialRuns runs: searches of other
      // threads share the GPU and stretch single runs
      constexpr int kTrialRuns = 3;
      auto trial = [&](const DPConfig& c) {
        float best_ms = 0;
        for (int k = 0; k < kTrialRuns; ++k) {
          hip_check(hipEventRecord(e0, st), "hipEventRecord");
          (void)dp_chain(c, false);
          hip_check(hipEventRecord(e1, st), "hipEventRecord");
          hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
          float ms = 0;
          hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
          if (k == 0 || ms < best_ms) best_ms = ms;
        }
        trials.emplace_back(c, best_ms);
        return best_ms;
      };
      (void)trial(cfg);  // first launches of the kernels, not compared
      trials.clear();
      float best = trial(cfg);
      auto consider = [&](DPConfig c) {
        c = forced.apply(c);
        for (auto& t : trials)
          if (t.first == c) return;
        const float ms = trial(c);
        if (ms < best) {
          best = ms;
          cfg = c;
        }
      };
      if (packed && forced.sched < 0 && p.gap <= kPersistMaxGap) consider(DPConfig{1 - cfg.sched, cfg.band, cfg.radix});
      if (maxrows >= kBandRows && forced.band < 0)
        for (int b = 0; b < 3; ++b) consider(DPConfig{cfg.sched, b, cfg.radix});
      if (forced.radix < 0)
        for (int r : {8, 6, 5}) consider(DPConfig{cfg.sched, cfg.band, r});
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      DPTuner::get().store(tkey, cfg, trials);
    }
    const DPArgs da = dp_chain(cfg, true);
    out.dpcfg = cfg;
    if (sp.on) {
      sp.da = da;
      sp.nslot = nslot;
      sp.key = tkey;
    }
    k_dp_combine<<<grid_for(nhit, kBlock), kBlock, 0, st>>>(da, nhit);
    rt.check_launch("k_dp_combine");
  }
  if (keep_stages) {
    std::vector<int32_t> c[6], hsv, hgv;
    d2h(c[0], ot, nhit, st); d2h(c[1], oq, nhit, st); d2h(c[2], thl, nhit, st);
    d2h(c[3], qhl, nhit, st); d2h(c[4], om, nhit, st); d2h(c[5], osc, nhit, st);
    d2h(hsv, hs, nhit, st); d2h(hgv, hgs, nhit, st);
    rt.sync();
    auto& S = out.stage[CLAST_STAGE_ALIGNED];
    for (int k = 0; k < 8; ++k) S[k].resize(nhit);
    for (uint32_t i = 0; i < nhit; ++i) {
      S[0][i] = ix.h_seq_ref[hgv[i]]; S[1][i] = c[0][i]; S[2][i] = q.h_qref[hsv[i]]; S[3][i] = c[1][i];
      S[4][i] = c[2][i]; S[5][i] = c[3][i]; S[6][i] = c[4][i]; S[7][i] = c[5][i];
    }
    out.stage_n[CLAST_STAGE_ALIGNED] = nhit;
  }

  // E-value, filters, dedup and row order, pair by pair (hits are dense in
  // strand order, so a pair's hits are contiguous)
  mark(EV_RES0);
  uint32_t nfinal = 0;
  uint32_t p2n[3] = {0, 0, 0};  // second DP pass: lane-per-band, k_dp16, past-the-end tasks
  bool p2_pending = false;
  uint32_t* order = nullptr;
  double* hev = nullptr;
  if (nhit) {
    hev = ws.get<double>("r.ev", nhit);
    uint32_t* kflag = ws.get<uint32_t>("r.kflag", nhit);
    RowKey* rk = ws.get<RowKey>("r.key", nhit);
    uint32_t* dflag = ws.get<uint32_t>("r.dflag", nhit);
    uint32_t* gflag = ws.get<uint32_t>("r.gflag", nhit);
    uint32_t* doff = ws.get<uint32_t>("r.doff", nhit);
    uint32_t* hbase = ws.get<uint32_t>("h.base", q.nseq);
    order = ws.get<uint32_t>("r.order", nhit);
    const uint32_t g = grid_for(nhit, kBlock);
    k_hit_filter<<<g, kBlock, 0, st>>>(nhit, hs, hgs, ht, hq, osc, qhl, q.qlen.ptr, ix.seq_win.ptr, q.smin.ptr,
                                       use_cut ? 1 : 0, p.lmer + p.stride, p.K * p.total_db_size, p.lambda, hev,
                                       kflag, rk, p.lmer, 2 * p.gap + 1,
                                       reinterpret_cast<unsigned long long*>(tot64 + 2));
    rt.check_launch("k_hit_filter");
    uint32_t* order2 = nullptr;
    uint32_t* split2 = nullptr;
    if (sp.on) {
      // the second pass's three task lists (k_redo_flags), counts to tot32[8, 11)
      const uint64_t n3 = 3ull * sp.nslot;
      uint32_t* rfl = ws.get<uint32_t>("p2.flags", n3);
      uint32_t* rof = ws.get<uint32_t>("p2.offs", n3);
      order2 = ws.get<uint32_t>("p2.order", sp.nslot);
      split2 = ws.get<uint32_t>("p2.split", 3);
      const uint32_t gs2 = grid_for(sp.nslot, kBlock);
      // long tasks to k_dp_band only while they would set the pass's end by
      // themselves (over twice one SIMD's share of the pass's rows): many
      // long survivors run faster in k_dp16
      auto* rsum2 = reinterpret_cast<unsigned long long*>(tot64 + 44);
      int* band_rows2 = ws.get<int>("p2.bandrows", 1);
      k_redo_rows<<<gs2, kBlock, 0, st>>>(sp.da.order, sp.da.split, kflag, hs, hq, q.qlen.ptr, p.lmer, sp.nslot,
                                          rsum2);
      k_band_threshold<<<1, 64, 0, st>>>(rsum2, 4ull * static_cast<uint64_t>(std::max(rt.compute_units(), 1)) * 128ull,
                                         band_rows2);
      k_redo_flags<<<gs2, kBlock, 0, st>>>(sp.da.order, sp.da.split, kflag, hs, hq, q.qlen.ptr, p.lmer, sp.nslot,
                                           band_rows2, rfl, reinterpret_cast<unsigned long long*>(tot64 + 36));
      for (int c = 0; c < 3; ++c)
        device_scan<OpAdd, uint32_t, uint32_t, uint32_t>(rt, rfl + c * static_cast<uint64_t>(sp.nslot), sp.nslot,
                                                         rof + c * static_cast<uint64_t>(sp.nslot), tot32 + 8 + c,
                                                         false, "p2.tiles");
      k_redo_scatter<<<gs2, kBlock, 0, st>>>(sp.da.order, rfl, rof, sp.nslot, tot32 + 8, order2, split2);
      rt.check_launch("k_redo_flags/scatter");
    }
    PairArgs pa{nhit, q.nseq, hbase, hs, hgs, ot, oq, kflag, rk};
    k_fill<uint32_t><<<1, 64, 0, st>>>(tot32 + 4, 1, 0u);
    k_pair_dedup<<<g, kBlock, 0, st>>>(pa, dflag, gflag, tot32 + 4);
    rt.check_launch("k_pair_dedup");
    // tot32[4]: rows of pairs over kPairSmall hits, and the second pass's
    // list sizes, tot32[8, 11)
    uint32_t rb[7];
    rt.read_scalars(tot32 + 4, rb, 7);
    const uint32_t nbig = rb[0];
    if (sp.on) {
      p2n[0] = rb[4];
      p2n[1] = rb[5];
      p2n[2] = rb[6];
    }
    if (p2n[0] + p2n[1] + p2n[2] > 0) {
      // the second pass on the side stream, beside the row ordering below;
      // joined before the rows are gathered
      DPArgs d2 = sp.da;
      d2.order = order2;
      d2.split = split2;
      hipStream_t ss = rt.side_stream();
      rt.fork(1);
      mark_on(EV_P2_0, ss);
      if (p2n[0]) {
        d2.n = p2n[0];
        launch_dp_band(p.gap, false, d2, ss);
      }
      if (p2n[1]) {
        d2.n = p2n[1];
        launch_dp16(p.gap, false, false, 0, false, d2, ss);
      }
      if (p2n[2]) {
        d2.n = p2n[2];
        launch_dp16(p.gap, true, false, 0, false, d2, ss);
      }
      k_match_patch<<<g, kBlock, 0, ss>>>(nhit, kflag, dres, hs, hq, q.qlen.ptr, p.lmer, om);
      mark_on(EV_P2_1, ss);
      rt.check_launch("second DP pass");
      p2_pending = true;
    }
    uint32_t* grow = nullptr;
    uint32_t* gfirst = nullptr;
    uint32_t ngrow = 0;
    if (nbig) {
      // pairs over kPairSmall hits: complete order keys, radix sorts, dedup
      uint32_t* goff = ws.get<uint32_t>("r.goff", nhit);
      device_scan<OpAdd, uint32_t, uint32_t, uint32_t>(rt, gflag, nhit, goff, tot32 + 5, false);
      int32_t* srange = ws.get<int32_t>("r.srange", 2);
      k_fill<int32_t><<<1, 64, 0, st>>>(srange, 1, INT32_MAX);
      k_fill<int32_t><<<1, 64, 0, st>>>(srange + 1, 1, INT32_MIN);
      k_score_range<<<g, kBlock, 0, st>>>(nhit, gflag, osc, srange);
      hip_check(hipMemcpyAsync(tot32 + 6, srange, 8, hipMemcpyDeviceToDevice, st), "copy");
      uint32_t sc3[3];
      rt.read_scalars(tot32 + 5, sc3, 3);
      const uint32_t nk = sc3[0];
      // every row of the large pairs may have failed the bad-hit filters:
      // then nothing is left to order (and the score range is empty)
      if (nk > 0) {
      const int32_t smin_v = static_cast<int32_t>(sc3[1]), smax_v = static_cast<int32_t>(sc3[2]);
      const int scb = bits_for(static_cast<uint64_t>(static_cast<int64_t>(smax_v) - smin_v + 1));
      const int pbits = bits_for((q.nseq + 1) / 2);
      const int wbits = bits_for(static_cast<uint64_t>(ix.nwin));
      const int qb = bits_for(static_cast<uint64_t>(q.max_qlen) + 1);
      if (wbits + pbits + scb > 64 || gbits + dbits > 64) throw ClastError(CLAST_ERR_LIMIT, "order key over 64 bits");
      uint64_t* w0 = ws.get<uint64_t>("r.w0", nk);
      uint64_t* w1 = ws.get<uint64_t>("r.w1", nk);
      uint64_t* w2 = ws.get<uint64_t>("r.w2", nk);
      uint32_t* rec = ws.get<uint32_t>("r.rec", nk);
      uint32_t* perm = ws.get<uint32_t>("r.perm", nk);
      k_order_keys<<<g, kBlock, 0, st>>>(nhit, gflag, goff, hs, hgs, ht, hq, osc, ix.seq_win.ptr, dbits, d_off,
                                         q.max_qlen, pbits, scb, smax_v, w0, w1, w2, rec);
      rt.check_launch("k_order_keys");
      const uint64_t* words[3] = {w0, w1, w2};
      const int wb[3] = {gbits + dbits, qb + 1, wbits + pbits + scb};
      sort_perm(rt, nk, 3, words, wb, perm);
      uint32_t* bflag = ws.get<uint32_t>("r.bflag", nk);
      uint32_t* boff = ws.get<uint32_t>("r.boff", nk);
      const uint32_t gk = grid_for(nk, kBlock);
      k_dedup<<<gk, kBlock, 0, st>>>(nk, perm, rec, hs, hgs, ot, oq, ix.seq_win.ptr, q.qwin.ptr, bflag);
      device_scan<OpAdd, uint32_t, uint32_t, uint32_t>(rt, bflag, nk, boff, tot32 + 7, false);
      rt.read_scalars(tot32 + 7, &ngrow, 1);
      grow = ws.get<uint32_t>("r.grow", ngrow);
      k_compact_perm<<<gk, kBlock, 0, st>>>(nk, perm, rec, bflag, boff, grow);
      rt.check_launch("k_dedup/compact");
      if (ix.nwin > 1 && ngrow > 1) {
        // unit order is window-major: rows of a pair to (score desc, window)
        uint64_t* wf = ws.get<uint64_t>("r.wf", ngrow);
        uint32_t* pf = ws.get<uint32_t>("r.permB", ngrow);
        uint32_t* g2 = ws.get<uint32_t>("r.grow2", ngrow);
        k_final_keys<<<grid_for(ngrow, kBlock), kBlock, 0, st>>>(ngrow, grow, hs, hgs, osc, ix.seq_win.ptr, scb,
                                                                 wbits, smax_v, wf);
        const uint64_t* fw[1] = {wf};
        const int fb[1] = {pbits + scb + wbits};
        sort_perm(rt, ngrow, 1, fw, fb, pf);
        k_gather<uint32_t><<<grid_for(ngrow, kBlock), kBlock, 0, st>>>(grow, pf, g2, ngrow);
        grow = g2;
      }
      if (ngrow) {
        gfirst = ws.get<uint32_t>("r.gfirst", nhit);
        k_big_counts<<<grid_for(ngrow, kBlock), kBlock, 0, st>>>(ngrow, grow, hs, hbase, dflag, gfirst);
        rt.check_launch("k_big_counts");
      }
      }
    }
    // the row count stays on the device (tot32[3]); order has room for every hit
    device_scan<OpAdd, uint32_t, uint32_t, uint32_t>(rt, dflag, nhit, doff, tot32 + 3, false);
    k_pair_place<<<g, kBlock, 0, st>>>(pa, dflag, doff, order);
    if (ngrow)
      k_big_place<<<grid_for(ngrow, kBlock), kBlock, 0, st>>>(ngrow, grow, hs, hbase, gfirst, doff, order);
    rt.check_launch("k_pair_place");
  }
  if (p2_pending) rt.join(1);
  // the handle's earlier rows (if any) are released
  out.wait();
  clast::pinned_pool().release(out.rows);
  out.rows = {};
  out.n = 0;
  if (rows_out) {
    if (nhit) rt.read_scalars(tot32 + 3, &nfinal, 1);
    rows_out->needed = nfinal;
    if (nfinal > rows_out->capacity)
      throw ClastError(CLAST_ERR_LIMIT, "result rows exceed the caller's capacity (" + std::to_string(nfinal) +
                                            " needed)");
    if (nfinal) {
      uint8_t* ob = ws.get<uint8_t>("o0", static_cast<uint64_t>(nfinal) * 48);
      int32_t* o32 = reinterpret_cast<int32_t*>(ob);
      double* oev = reinterpret_cast<double*>(ob + static_cast<uint64_t>(nfinal) * 40);
      OutArgs oa{nfinal, order, nullptr, hs, hgs, ot, oq, thl, qhl, om, osc, hev,
                 ix.seq_ref.ptr, q.qref.ptr, ix.seq_win.ptr, q.qwin.ptr, o32, oev, nullptr, nullptr};
      k_gather_out<<<grid_for(nfinal, kBlock), kBlock, 0, st>>>(oa);
      rt.check_launch("k_gather_out");
      for (int c = 0; c < 10; ++c)
        if (rows_out->col[c])
          hip_check(hipMemcpyAsync(rows_out->col[c], o32 + static_cast<uint64_t>(c) * nfinal, 4ull * nfinal,
                                   hipMemcpyDeviceToHost, st),
                    "D2H rows");
      if (rows_out->evalue)
        hip_check(hipMemcpyAsync(rows_out->evalue, oev, 8ull * nfinal, hipMemcpyDeviceToHost, st), "D2H rows");
    }
  }
  uint8_t* ob = nullptr;  // staged rows (planes laid out for the row count the gather reads)
  int slot = 0;
  if (!rows_out && nhit) {
    // the gather reads the row count on the device, so it runs before the
    // read-back below: the search's kernels are done when it returns
    slot = rt.next_out_slot();
    const std::string tag = slot ? "o1" : "o0";
    // staging buffer `slot` is free once the copy that last read it is done
    if (rt.out_used(slot)) hip_check(hipStreamWaitEvent(st, rt.out_done(slot), 0), "hipStreamWaitEvent");
    ob = ws.get<uint8_t>(tag, static_cast<uint64_t>(nhit) * 48);
    OutArgs oa{0, order, nullptr, hs, hgs, ot, oq, thl, qhl, om, osc, hev,
               ix.seq_ref.ptr, q.qref.ptr, ix.seq_win.ptr, q.qwin.ptr, nullptr, nullptr, tot32 + 3, ob};
    k_gather_out<<<grid_for(nhit, kBlock), kBlock, 0, st>>>(oa);
    rt.check_launch("k_gather_out");
  }
  mark(EV_RES1);
  // the search's last wait: row count and work counters in one read-back
  uint64_t cnt[42];
  uint32_t r32[16] = {};
  rt.read_regions({{tot64 + 2, cnt, sizeof cnt}, {tot32, r32, sizeof r32}});
  if (!rows_out && nhit) {
    nfinal = r32[3];
    if (nfinal) {
      // the copy to pinned host memory on the copy stream (the caller's
      // reads wait for it; the next search's kernels do not)
      out.rows = clast::pinned_pool().acquire(static_cast<uint64_t>(nfinal) * 48);
      out.n = nfinal;
      hipStream_t cs = rt.copy_stream();
      hip_check(hipMemcpyAsync(out.rows.ptr, ob, static_cast<uint64_t>(nfinal) * 48, hipMemcpyDeviceToHost, cs),
                "D2H rows");
      hip_check(hipEventRecord(rt.out_done(slot), cs), "hipEventRecord");
      rt.set_out_used(slot);
      hip_check(hipEventCreateWithFlags(&out.ready, hipEventDisableTiming), "hipEventCreate");
      hip_check(hipEventRecord(out.ready, cs), "hipEventRecord");
    }
  }
  const int pairs[7][2] = {{EV_COUNT0, EV_COUNT1}, {EV_WRITE0, EV_WRITE1}, {EV_SORT0, EV_SORT1}, {EV_FILT0, EV_FILT1},
                           {EV_DP0, EV_DP1},       {EV_RES0, EV_RES1},     {EV_START, EV_RES1}};
  for (int k = 0; k < 7; ++k) {
    float t = 0;
    if (timed[pairs[k][0]] && timed[pairs[k][1]])
      hip_check(hipEventElapsedTime(&t, ev->e[pairs[k][0]], ev->e[pairs[k][1]]), "elapsed");
    out.times[k] = t;
  }
  out.times[7] = static_cast<float>(nseed);
  const int kpairs[8][2] = {{EV_COUNT0, EV_COUNT1}, {EV_D16_0, EV_D16_1}, {EV_D32_0, EV_D32_1},
                            {EV_DBAND0, EV_DBAND1}, {EV_DP0, EV_DP1},    {EV_WRITE0, EV_WRITE1},
                            {EV_RES0, EV_RES1},     {EV_START, EV_RES1}};
  for (int k = 0; k < 8; ++k) {
    float t = 0;
    if (timed[kpairs[k][0]] && timed[kpairs[k][1]])
      hip_check(hipEventElapsedTime(&t, ev->e[kpairs[k][0]], ev->e[kpairs[k][1]]), "elapsed");
    out.ktimes[k] = t;
  }
  uint64_t cells = 0, rows16 = 0, rowsband = 0;
  uint64_t rows_main = 0;
  for (int k = 0; k < 8; ++k) {
    cells += cnt[k];
    rows16 += cnt[8 + k];
    rowsband += cnt[18 + k];
    if (k < 4) rows_main += cnt[8 + k];
  }
  {
    const double w = 2 * p.gap + 1;
    out.dpk[0] = w * static_cast<double>(rows_main);
    out.dpk[1] = w * static_cast<double>(rows16 - rows_main);
    out.dpk[3] = w * static_cast<double>(rowsband);
    out.dpk[2] = static_cast<double>(cells) - out.dpk[0] - out.dpk[1] - out.dpk[3];
    const int dk[5][2] = {{EV_D16_0, EV_D16_1}, {EV_E16_0, EV_E16_1}, {EV_D32_0, EV_D32_1}, {EV_DBAND0, EV_DBAND1},
                          {EV_P2_0, EV_P2_1}};
    for (int k = 0; k < 5; ++k) {
      float t = 0;
      if (timed[dk[k][0]] && timed[dk[k][1]])
        hip_check(hipEventElapsedTime(&t, ev->e[dk[k][0]], ev->e[dk[k][1]]), "elapsed");
      out.dpk[k < 4 ? 4 + k : 10] = t;
    }
    uint64_t rows2 = 0;
    for (int k = 34; k < 42; ++k) rows2 += cnt[k];
    out.dpk[8] = static_cast<double>(p2n[0]) + p2n[1] + p2n[2];
    out.dpk[9] = sp.on ? w * static_cast<double>(rows2) : 0.0;
    out.dpk[11] = sp.on ? 1.0 : 0.0;
    if (sp.on) DPTuner::get().survived(sp.key, static_cast<uint64_t>(out.dpk[8]), cnt[16]);
  }
  out.counts[0] = nseed;
  out.counts[1] = nhit;
  out.counts[2] = static_cast<int64_t>(cells);
  out.counts[3] = nfinal;
  out.counts[4] = static_cast<int64_t>(cnt[16]);
  out.counts[5] = static_cast<int64_t>(rows16) * (2 * p.gap + 1);
  out.counts[6] = static_cast<int64_t>(cnt[17]);
  out.counts[7] = static_cast<int64_t>(rowsband) * (2 * p.gap + 1);
}
}  // namespace

// ---------------------------------------------------------------- C ABI

// Persisted index file helpers (clast_index_save / clast_index_load).
namespace {

constexpr char kIndexMagic[8] = {'C', 'L', 'A', 'S', 'T', 'I', 'X', '1'};
constexpr uint32_t kIndexVersion = 1;

struct IndexFileHead {
  char magic[8];
  uint32_t version;
  int32_t lmer, stride, cut_repeat, nwin, nseq;
  uint32_t nentries;
  uint64_t nslots;
};

struct FileCloser {
  void operator()(std::FILE* f) const { std::fclose(f); }
};
using File = std::unique_ptr<std::FILE, FileCloser>;

void file_write(std::FILE* f, const void* p, size_t n) {
  if (n && std::fwrite(p, 1, n, f) != n) throw ClastError(CLAST_ERR_IO, "index file: write failed");
}
void file_read(std::FILE* f, void* p, size_t n) {
  if (n && std::fread(p, 1, n, f) != n) throw ClastError(CLAST_ERR_IO, "index file: truncated or unreadable");
}
template <class T>
void file_write_vec(std::FILE* f, const std::vector<T>& v) { file_write(f, v.data(), sizeof(T) * v.size()); }
template <class T>
std::vector<T> file_read_vec(std::FILE* f, uint64_t n) {
  std::vector<T> v(n);
  file_read(f, v.data(), sizeof(T) * n);
  return v;
}
void bad_file(const char* what) { throw ClastError(CLAST_ERR_IO, std::string("index file: ") + what); }

}  // namespace

extern "C" {

int clast_init(int32_t device) {
  return guard([&] {
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) throw ClastError(CLAST_ERR_ARG, "no such device");
    set_device(device);
    (void)runtime();
    scratch_setup(device);
  });
}

const char* clast_last_error(void) { return g_err.c_str(); }

const char* clast_version(void) { return "libclast_hip 0.1 (gfx950)"; }

int clast_index_create(const clast_params* p, const clast_window* t, clast_index* out) {
  return clast_index_create_multi(p, t, 1, out);
}

int clast_index_create_multi(const clast_params* p, const clast_window* t, int32_t n, clast_index* out) {
  return guard([&] {
    check_params(p);
    if (!t || n < 1 || !out) throw ClastError(CLAST_ERR_ARG, "bad target windows");
    auto ix = std::make_unique<clast_index_s>();
    build_index(*p, t, n, *ix);
    *out = ix.release();
  });
}

// Inverse of index_code: the k-mer code of an index code.
static uint64_t kmer_of_index_code(uint64_t f, int lmer) {
  if (!(lmer & 1)) return f;
  const uint64_t hb = 1ull << lmer;
  const uint64_t c0 = f >> 1;
  const uint64_t c = (c0 & (hb - 1ull)) | ((c0 & ~(hb - 1ull)) << 1);
  if (!(f & 1ull)) return c;
  uint64_t x = c, rc = 0;
  for (int j = 0; j < lmer; ++j, x >>= 2) rc = (rc << 2) | (3ull - (x & 3ull));
  return rc;
}

int64_t clast_index_entry_count(clast_index idx) { return idx ? static_cast<int64_t>(idx->nentries) : -1; }

int clast_index_entries(clast_index ix, int64_t* key, int32_t* window, int32_t* tID, int32_t* tIdx) {
  return guard([&] {
    if (!ix) throw ClastError(CLAST_ERR_ARG, "null index");
    set_device(ix->device);
    Runtime& rt = runtime();
    std::lock_guard<std::mutex> lk(rt.mutex());
    const uint64_t n = ix->nentries;
    std::vector<int2> ep;
    d2h(ep, ix->epos.ptr, n, rt.stream());
    std::vector<int64_t> k(n);
    if (ix->key64) {
      std::vector<uint64_t> kk;
      d2h(kk, ix->ekey64.ptr, n, rt.stream());
      rt.sync();
      for (uint64_t i = 0; i < n; ++i) k[i] = static_cast<int64_t>(kk[i]);
    } else {
      std::vector<uint32_t> kk;
      d2h(kk, ix->ekey32.ptr, n, rt.stream());
      rt.sync();
      for (uint64_t i = 0; i < n; ++i) k[i] = static_cast<int64_t>(kk[i]);
    }
    // entries are stored in index-code order; report k-mer codes in k-mer
    // order (a stable sort keeps each k-mer's run as it is)
    for (uint64_t i = 0; i < n; ++i) k[i] = static_cast<int64_t>(kmer_of_index_code(static_cast<uint64_t>(k[i]), ix->lmer));
    std::vector<uint64_t> ord(n);
    for (uint64_t i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](uint64_t x, uint64_t y) { return k[x] < k[y]; });
    for (uint64_t r = 0; r < n; ++r) {
      const uint64_t i = ord[r];
      if (key) key[r] = k[i];
      if (window) window[r] = ix->h_seq_win[ep[i].x];
      if (tID) tID[r] = ix->h_seq_ref[ep[i].x];
      if (tIdx) tIdx[r] = ep[i].y;
    }
  });
}

// ---------------------------------------------------------------- persisted index
//
// File layout (host byte order): IndexFileHead, then
//   window slot offsets  u64 x (nwin + 1)    window list IDs  i32 x nwin
//   sequence first slots u64 x nseq          lengths, windows, list IDs  i32 x nseq each
//   class bytes          u8  x nslots        entries (seq, tIdx) i32 x 2 x nentries
//   entry index codes    u64 x nentries      (ascending)


int clast_index_save(clast_index ix, const char* path) {
  return guard([&] {
    if (!ix || !path) throw ClastError(CLAST_ERR_ARG, "null index or path");
    set_device(ix->device);
    Runtime& rt = runtime();
    std::lock_guard<std::mutex> lk(rt.mutex());
    hipStream_t st = rt.stream();
    const uint64_t nslots = ix->h_win_slot.back();
    const uint32_t n = ix->nentries;
    std::vector<uint8_t> tb;
    std::vector<uint64_t> seq_slot, keys(n);
    std::vector<int2> ep;
    std::vector<uint32_t> k32;
    d2h(tb, ix->tbase.ptr, nslots, st);
    d2h(seq_slot, ix->seq_slot.ptr, static_cast<uint64_t>(ix->nseq), st);
    d2h(ep, ix->epos.ptr, n, st);
    if (ix->key64) d2h(keys, ix->ekey64.ptr, n, st);
    else d2h(k32, ix->ekey32.ptr, n, st);
    rt.sync();
    if (!ix->key64)
      for (uint32_t i = 0; i < n; ++i) keys[i] = k32[i];
    File f(std::fopen(path, "wb"));
    if (!f) throw ClastError(CLAST_ERR_IO, std::string("index file: cannot create ") + path);
    IndexFileHead h{};
    std::memcpy(h.magic, kIndexMagic, sizeof(h.magic));
    h.version = kIndexVersion;
    h.lmer = ix->lmer;
    h.stride = ix->stride;
    h.cut_repeat = ix->cut_repeat;
    h.nwin = ix->nwin;
    h.nseq = ix->nseq;
    h.nentries = n;
    h.nslots = nslots;
    file_write(f.get(), &h, sizeof(h));
    file_write_vec(f.get(), ix->h_win_slot);
    file_write_vec(f.get(), ix->h_id_begin);
    file_write_vec(f.get(), seq_slot);
    file_write_vec(f.get(), ix->h_seq_len);
    file_write_vec(f.get(), ix->h_seq_win);
    file_write_vec(f.get(), ix->h_seq_ref);
    file_write_vec(f.get(), tb);
    file_write_vec(f.get(), ep);
    file_write_vec(f.get(), keys);
    if (std::fflush(f.get()) != 0) throw ClastError(CLAST_ERR_IO, "index file: write failed");
  });
}

int clast_index_load(const char* path, clast_index* out) {
  return guard([&] {
    if (!path || !out) throw ClastError(CLAST_ERR_ARG, "null path or output");
    // everything is read and checked on the host before any device work
    File f(std::fopen(path, "rb"));
    if (!f) throw ClastError(CLAST_ERR_IO, std::string("index file: cannot open ") + path);
    IndexFileHead h{};
    file_read(f.get(), &h, sizeof(h));
    if (std::memcmp(h.magic, kIndexMagic, sizeof(h.magic)) != 0) bad_file("not an index file");
    if (h.version != kIndexVersion) bad_file("unsupported version");
    if (h.lmer < 1 || h.lmer > 31 || h.stride < 1 || h.nwin < 1 || h.nseq < 0 || h.nslots >= (1ull << 35))
      bad_file("header out of range");
    const uint64_t nw = static_cast<uint64_t>(h.nwin), nq = static_cast<uint64_t>(h.nseq), n = h.nentries;
    const uint64_t expect = sizeof(h) + 8 * (nw + 1) + 4 * nw + 8 * nq + 12 * nq + h.nslots + 16 * n;
    if (std::fseek(f.get(), 0, SEEK_END) != 0) bad_file("not seekable");
    const long size = std::ftell(f.get());
    if (size < 0 || static_cast<uint64_t>(size) != expect) bad_file("size does not match its header");
    if (std::fseek(f.get(), static_cast<long>(sizeof(h)), SEEK_SET) != 0) bad_file("not seekable");
    auto ix = std::make_unique<clast_index_s>();
    ix->lmer = h.lmer;
    ix->stride = h.stride;
    ix->cut_repeat = h.cut_repeat;
    ix->nwin = h.nwin;
    ix->nseq = h.nseq;
    ix->key64 = h.lmer > 16;
    ix->h_win_slot = file_read_vec<uint64_t>(f.get(), nw + 1);
    ix->h_id_begin = file_read_vec<int32_t>(f.get(), nw);
    const std::vector<uint64_t> seq_slot = file_read_vec<uint64_t>(f.get(), nq);
    ix->h_seq_len = file_read_vec<int32_t>(f.get(), nq);
    ix->h_seq_win = file_read_vec<int32_t>(f.get(), nq);
    ix->h_seq_ref = file_read_vec<int32_t>(f.get(), nq);
    const std::vector<uint8_t> tb = file_read_vec<uint8_t>(f.get(), h.nslots);
    const std::vector<int2> ep = file_read_vec<int2>(f.get(), n);
    const std::vector<uint64_t> keys = file_read_vec<uint64_t>(f.get(), n);
    // the tables must describe a layout the kernels can index
    if (ix->h_win_slot[0] != 0 || ix->h_win_slot[nw] != h.nslots) bad_file("window table");
    for (uint64_t w = 0; w < nw; ++w)
      if (ix->h_win_slot[w + 1] < ix->h_win_slot[w]) bad_file("window table");
    for (uint64_t q = 0; q < nq; ++q) {
      const int32_t w = ix->h_seq_win[q], len = ix->h_seq_len[q];
      if (w < 0 || w >= h.nwin || (q > 0 && w < ix->h_seq_win[q - 1]) || len < 0 ||
          seq_slot[q] < ix->h_win_slot[w] || seq_slot[q] + static_cast<uint64_t>(len) > ix->h_win_slot[w + 1])
        bad_file("sequence table");
      ix->max_tlen = std::max(ix->max_tlen, len);
    }
    for (uint8_t c : tb)
      if (c > kClsX) bad_file("class bytes");
    const uint64_t kmax = 1ull << (2 * h.lmer);
    for (uint64_t i = 0; i < n; ++i) {
      if (keys[i] >= kmax || (i > 0 && keys[i] < keys[i - 1])) bad_file("entry codes");
      if (ep[i].x < 0 || ep[i].x >= h.nseq || ep[i].y < 0 ||
          static_cast<int64_t>(ep[i].y) + h.lmer > ix->h_seq_len[static_cast<size_t>(ep[i].x)])
        bad_file("entries");
    }
    f.reset();

    Runtime& rt = runtime();
    std::lock_guard<std::mutex> lk(rt.mutex());
    hipStream_t st = rt.stream();
    ix->device = rt.device();
    ix->tbase.alloc(padded_blocks(h.nslots) * 16);
    ix->bmax = static_cast<int32_t>(padded_blocks(h.nslots) - 1);
    if (h.nslots) hip_check(hipMemcpyAsync(ix->tbase.ptr, tb.data(), h.nslots, hipMemcpyHostToDevice, st), "upload");
    target_tables(rt, *ix, seq_slot);
    ix->epos.upload(ep.data(), n, st);
    uint64_t* ek = n ? rt.ws().get<uint64_t>("ix.ekey", n) : nullptr;
    if (n) hip_check(hipMemcpyAsync(ek, keys.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice, st), "upload");
    build_cells(rt, *ix, ek, static_cast<uint32_t>(n), 2 * h.lmer);
    ix->nentries = static_cast<uint32_t>(n);
    rt.check_launch("clast_index_load");
    rt.sync();
    *out = ix.release();
  });
}

int clast_index_destroy(clast_index idx) {
  return guard([&] {
    if (!idx) return;
    set_device(idx->device);
    delete idx;
  });
}

int clast_query_create(const clast_params* p, const clast_window* q, clast_query* out) {
  return clast_query_create_multi(p, q, 1, out);
}

int clast_query_create_multi(const clast_params* p, const clast_window* q, int32_t n, clast_query* out) {
  return guard([&] {
    check_params(p);
    if (!q || n < 1 || !out) throw ClastError(CLAST_ERR_ARG, "bad query windows");
    auto qq = std::make_unique<clast_query_s>();
    build_query(*p, q, n, *qq);
    *out = qq.release();
  });
}

int clast_query_create_reads(const clast_params* p, const uint8_t* bases, const int64_t* offsets, const int32_t* win,
                             const int32_t* id_begin, int32_t n_win, clast_query* out) {
  return guard([&] {
    check_params(p);
    if (!bases || !offsets || !win || !id_begin || n_win < 1 || !out)
      throw ClastError(CLAST_ERR_ARG, "bad read windows");
    auto qq = std::make_unique<clast_query_s>();
    build_query_reads(*p, bases, offsets, win, id_begin, n_win, *qq);
    *out = qq.release();
  });
}

int clast_query_destroy(clast_query q) {
  return guard([&] {
    if (!q) return;
    set_device(q->device);
    delete q;
  });
}

int clast_search(const clast_params* p, clast_index idx, clast_query q, int32_t t_begin, int32_t q_begin,
                 clast_hits* out) {
  return guard([&] {
    check_params(p);
    if (!idx || !q || !out) throw ClastError(CLAST_ERR_ARG, "null handle");
    if (idx->nwin != 1 || q->nwin != 1) throw ClastError(CLAST_ERR_ARG, "clast_search takes one window each");
    if (idx->h_id_begin[0] != t_begin || q->h_id_begin[0] != q_begin)
      throw ClastError(CLAST_ERR_ARG, "t_begin/q_begin differ from the windows' id_begin");
    auto h = std::make_unique<clast_hits_s>();
    run_search(*p, *idx, *q, 0, *h);
    *out = h.release();
  });
}

int clast_search_batch(const clast_params* p, clast_index idx, clast_query q, int32_t flags, clast_hits* out) {
  return guard([&] {
    check_params(p);
    if (!idx || !q || !out) throw ClastError(CLAST_ERR_ARG, "null handle");
    auto h = std::make_unique<clast_hits_s>();
    run_search(*p, *idx, *q, flags, *h);
    *out = h.release();
  });
}

int clast_search_batch_noalloc(const clast_params* p, clast_index idx, clast_query q, int64_t capacity,
                               int32_t* tID, int32_t* tIdx, int32_t* qID, int32_t* qIdx, int32_t* tHL,
                               int32_t* qHL, int32_t* match, int32_t* score, double* evalue, int32_t* twin,
                               int32_t* qwin, int64_t* n_rows) {
  return guard([&] {
    if (n_rows) *n_rows = 0;
    check_params(p);
    if (!idx || !q) throw ClastError(CLAST_ERR_ARG, "null handle");
    if (capacity < 0) throw ClastError(CLAST_ERR_ARG, "negative row capacity");
    if (idx->device < 0 || idx->device >= kMaxDevices) throw ClastError(CLAST_ERR_ARG, "bad device");
    set_device(idx->device);
    Runtime* rt = runtime_if_exists();
    ScratchSlot& sc = g_scratch[idx->device];
    if (!rt || !sc.ptr)
      throw ClastError(CLAST_ERR_ARG, "clast_search_batch_noalloc: call clast_init(device) on this thread first");
    // what the scratch cannot hold is refused before any device work: the
    // search keeps a 16-byte cell record per query k-mer position and about
    // 64 bytes per strand before the first seed
    const uint64_t floor_bytes = 16ull * q->npos + 64ull * q->nseq;
    if (floor_bytes > CLAST_NOALLOC_SCRATCH_BYTES)
      throw ClastError(CLAST_ERR_LIMIT, "query set too large for the allocation-free scratch (" +
                                            std::to_string(q->npos) + " k-mer positions): split it");
    std::lock_guard<std::mutex> lk(sc.call);
    // on every exit the streams drain before the scratch is released
    struct ArenaScope {
      Runtime* rt;
      ~ArenaScope() {
        (void)hipStreamSynchronize(rt->side_stream());
        (void)hipStreamSynchronize(rt->stream());
        rt->ws().end_arena();
      }
    } scope{rt};
    rt->ws().begin_arena(sc.ptr, CLAST_NOALLOC_SCRATCH_BYTES);
    CallerRows rows{capacity, {tID, tIdx, qID, qIdx, tHL, qHL, match, score, twin, qwin}, evalue};
    clast_hits_s h;
    try {
      run_search(*p, *idx, *q, 0, h, &rows);
    } catch (...) {
      if (n_rows) *n_rows = rows.needed;
      throw;
    }
    if (n_rows) *n_rows = rows.needed;
  });
}

int64_t clast_hits_count(clast_hits h) { return h ? h->n : -1; }

int clast_hits_wait(clast_hits h) {
  return guard([&] {
    if (!h) throw ClastError(CLAST_ERR_ARG, "null hits");
    h->wait();
  });
}

int clast_hits_copy(clast_hits h, int32_t* tID, int32_t* tIdx, int32_t* qID, int32_t* qIdx, int32_t* tHL,
                    int32_t* qHL, int32_t* match, int32_t* score, double* evalue) {
  return guard([&] {
    if (!h) throw ClastError(CLAST_ERR_ARG, "null hits");
    int32_t* dst[8] = {tID, tIdx, qID, qIdx, tHL, qHL, match, score};
    if (!h->n) return;
    h->wait();
    for (int c = 0; c < 8; ++c)
      if (dst[c]) std::memcpy(dst[c], h->col(c), static_cast<size_t>(h->n) * sizeof(int32_t));
    if (evalue) std::memcpy(evalue, h->ev(), static_cast<size_t>(h->n) * sizeof(double));
  });
}

int clast_hits_copy_windows(clast_hits h, int32_t* twin, int32_t* qwin) {
  return guard([&] {
    if (!h) throw ClastError(CLAST_ERR_ARG, "null hits");
    if (!h->n) return;
    h->wait();
    if (twin) std::memcpy(twin, h->col(8), static_cast<size_t>(h->n) * sizeof(int32_t));
    if (qwin) std::memcpy(qwin, h->col(9), static_cast<size_t>(h->n) * sizeof(int32_t));
  });
}

int64_t clast_hits_stage_count(clast_hits h, int32_t stage) {
  if (!h || stage < 1 || stage > 4) return -1;
  return h->stage_n[stage];
}

int clast_hits_stage_copy(clast_hits h, int32_t stage, int32_t* tID, int32_t* tIdx, int32_t* qID, int32_t* qIdx,
                          int32_t* tHL, int32_t* qHL, int32_t* match, int32_t* score) {
  return guard([&] {
    if (!h || stage < 1 || stage > 4) throw ClastError(CLAST_ERR_ARG, "bad stage");
    int32_t* dst[8] = {tID, tIdx, qID, qIdx, tHL, qHL, match, score};
    for (int c = 0; c < 8; ++c)
      if (dst[c] && !h->stage[stage][c].empty())
        std::memcpy(dst[c], h->stage[stage][c].data(), h->stage[stage][c].size() * sizeof(int32_t));
  });
}

int clast_hits_stage_times(clast_hits h, float* out8) {
  return guard([&] {
    if (!h || !out8) throw ClastError(CLAST_ERR_ARG, "null argument");
    std::memcpy(out8, h->times, sizeof h->times);
  });
}

int clast_hits_kernel_times(clast_hits h, float* out8) {
  return guard([&] {
    if (!h || !out8) throw ClastError(CLAST_ERR_ARG, "null argument");
    std::memcpy(out8, h->ktimes, sizeof h->ktimes);
  });
}

int clast_hits_dp_kernels(clast_hits h, double* out12) {
  return guard([&] {
    if (!h || !out12) throw ClastError(CLAST_ERR_ARG, "null argument");
    std::memcpy(out12, h->dpk, sizeof h->dpk);
  });
}

int clast_hits_dp_config(clast_hits h, int32_t* out3) {
  return guard([&] {
    if (!h || !out3) throw ClastError(CLAST_ERR_ARG, "null argument");
    out3[0] = h->dpcfg.sched;
    out3[1] = h->dpcfg.band;
    out3[2] = h->dpcfg.radix;
  });
}

int64_t clast_tuner_report(char* buf, int64_t cap) {
  const std::string r = DPTuner::get().report();
  if (buf && cap > 0) {
    const size_t n = std::min<size_t>(r.size(), static_cast<size_t>(cap - 1));
    std::memcpy(buf, r.data(), n);
    buf[n] = 0;
  }
  return static_cast<int64_t>(r.size());
}

int clast_mem_info(int64_t* free_bytes, int64_t* total_bytes) {
  return guard([&] {
    size_t f = 0, t = 0;
    hip_check(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    if (free_bytes) *free_bytes = static_cast<int64_t>(f);
    if (total_bytes) *total_bytes = static_cast<int64_t>(t);
  });
}

int clast_workspace_bytes(int64_t* bytes) {
  return guard([&] {
    if (!bytes) throw ClastError(CLAST_ERR_ARG, "null argument");
    Runtime* rt = runtime_if_exists();
    *bytes = rt ? static_cast<int64_t>(rt->ws().bytes()) : 0;
  });
}

int clast_hits_counts(clast_hits h, int64_t* out8) {
  return guard([&] {
    if (!h || !out8) throw ClastError(CLAST_ERR_ARG, "null argument");
    std::memcpy(out8, h->counts, sizeof h->counts);
  });
}

int clast_hits_destroy(clast_hits h) {
  delete h;
  return CLAST_OK;
}

int clast_result_sort(int64_t n, int32_t* tID, int32_t* tIdx, int32_t* qID, int32_t* qIdx, int32_t* tHL,
                      int32_t* qHL, int32_t* match, int32_t* score, double* evalue) {
  return guard([&] {
    if (n < 0 || n >= (1ll << 32) - 1) throw ClastError(CLAST_ERR_LIMIT, "bad row count");
    if (n <= 1) return;
    int32_t* cols[8] = {tID, tIdx, qID, qIdx, tHL, qHL, match, score};
    for (auto* c : cols)
      if (!c) throw ClastError(CLAST_ERR_ARG, "null column");
    if (!evalue) throw ClastError(CLAST_ERR_ARG, "null column");
    Runtime& rt = runtime();
    std::lock_guard<std::mutex> lk(rt.mutex());
    hipStream_t st = rt.stream();
    const uint32_t m = static_cast<uint32_t>(n);
    DevArray<int32_t> dq, ds;
    dq.upload(qID, m, st);
    ds.upload(score, m, st);
    uint64_t* w = rt.ws().get<uint64_t>("rs.w", m);
    uint32_t* perm = rt.ws().get<uint32_t>("rs.perm", m);
    k_result_keys<<<grid_for(m, kBlock), kBlock, 0, st>>>(m, dq.ptr, ds.ptr, w);
    const uint64_t* words[1] = {w};
    const int b[1] = {64};
    sort_perm(rt, m, 1, words, b, perm);
    std::vector<uint32_t> hp;
    d2h(hp, perm, m, st);
    rt.sync();
    std::vector<int32_t> tmp(m);
    for (auto* c : cols) {
      for (uint32_t i = 0; i < m; ++i) tmp[i] = c[hp[i]];
      std::memcpy(c, tmp.data(), sizeof(int32_t) * m);
    }
    std::vector<double> te(m);
    for (uint32_t i = 0; i < m; ++i) te[i] = evalue[hp[i]];
    std::memcpy(evalue, te.data(), sizeof(double) * m);
  });
}

int clast_result_finalize(int64_t n, const int32_t* tIdx, const int32_t* qID, const int32_t* qIdx,
                          const int32_t* tHL, const int32_t* qHL, const int32_t* match, const int32_t* score,
                          const double* evalue, const int32_t* qlabel, const int32_t* tlabel, const int32_t* tstart,
                          int32_t num_out, int32_t* order, uint8_t* printed, int64_t* n_printed) {
  return guard([&] {
    if (n_printed) *n_printed = 0;
    if (n < 0 || n >= (1ll << 32) - 1) throw ClastError(CLAST_ERR_LIMIT, "bad row count");
    if (num_out < -1) throw ClastError(CLAST_ERR_ARG, "num_out must be -1 or >= 0");
    if (n == 0) return;
    const void* in[] = {tIdx, qID, qIdx, tHL, qHL, match, score, evalue, qlabel, tlabel, tstart, order, printed};
    for (auto* c : in)
      if (!c) throw ClastError(CLAST_ERR_ARG, "null column");
    Runtime& rt = runtime();
    std::lock_guard<std::mutex> lk(rt.mutex());
    hipStream_t st = rt.stream();
    Workspace& ws = rt.ws();
    const uint32_t m = static_cast<uint32_t>(n);
    DevArray<int32_t> d[10];
    const int32_t* cols[10] = {tIdx, qID, qIdx, tHL, qHL, match, score, qlabel, tlabel, tstart};
    for (int k = 0; k < 10; ++k) d[k].upload(cols[k], m, st);
    DevArray<double> dev;
    dev.upload(evalue, m, st);
    // fixResult's stable order: (qID / 2 ascending, score descending)
    uint64_t* w = ws.get<uint64_t>("fin.w", m);
    uint32_t* perm = ws.get<uint32_t>("fin.perm", m);
    k_result_keys<<<grid_for(m, kBlock), kBlock, 0, st>>>(m, d[1].ptr, d[6].ptr, w);
    const uint64_t* words[1] = {w};
    const int bits[1] = {64};
    sort_perm(rt, m, 1, words, bits, perm);
    // printResult's row selection in that order
    uint32_t* dflag = ws.get<uint32_t>("fin.d", m);
    uint8_t* nflag = ws.get<uint8_t>("fin.n", m);
    uint32_t* head = ws.get<uint32_t>("fin.h", m);
    uint32_t* dpre = ws.get<uint32_t>("fin.dp", m);
    uint32_t* run = ws.get<uint32_t>("fin.r", m);
    uint8_t* pr = ws.get<uint8_t>("fin.p", m);
    uint32_t* pflag = ws.get<uint32_t>("fin.pf", m);
    uint32_t* tot = ws.get<uint32_t>("fin.tot", 1);
    const FinalCols fc{perm, d[7].ptr, d[1].ptr, d[2].ptr, d[4].ptr, d[8].ptr, d[0].ptr, d[9].ptr, d[3].ptr,
                       d[5].ptr, d[6].ptr, dev.ptr};
    const uint32_t g = grid_for(m, kBlock);
    k_final_flags<<<g, kBlock, 0, st>>>(fc, m, dflag, nflag, head);
    device_scan<OpAdd, uint32_t, uint32_t, uint32_t>(rt, dflag, m, dpre, static_cast<uint32_t*>(nullptr), false);
    device_scan<OpMax, uint32_t, uint32_t, uint32_t>(rt, head, m, run, static_cast<uint32_t*>(nullptr), true);
    k_final_print<<<g, kBlock, 0, st>>>(m, num_out, dflag, nflag, dpre, run, pr, pflag);
    device_scan<OpAdd, uint32_t, uint32_t, uint32_t>(rt, pflag, m, dflag, tot, false);
    rt.check_launch("clast_result_finalize");
    hip_check(hipMemcpyAsync(order, perm, 4ull * m, hipMemcpyDeviceToHost, st), "D2H order");
    hip_check(hipMemcpyAsync(printed, pr, m, hipMemcpyDeviceToHost, st), "D2H flags");
    uint32_t np = 0;
    rt.read_scalars(tot, &np, 1);  // waits for the copies too (same stream)
    if (n_printed) *n_printed = np;
  });
}

}  // extern "C"
