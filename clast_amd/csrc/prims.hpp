// Device-wide primitives for gfx950 (wave64): reduce-then-scan prefix sums
// and max-scans, and a stable LSD radix sort of (u64 key, u32 value) pairs
// whose per-tile ranking uses wavefront ballots instead of a block-wide
// split per bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "runtime.hpp"

namespace clast {

constexpr int kBlock = 256;       // threads per workgroup (4 wavefronts)
constexpr int kScanItems = 8;     // items per thread in scans
constexpr int kScanTile = kBlock * kScanItems;
constexpr int kSortItems = 16;    // items per thread in radix passes
constexpr int kSortTile = kBlock * kSortItems;
constexpr int kRadixBits = 8;
constexpr int kRadixBins = 1 << kRadixBits;

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

struct OpAdd {
  template <class T> __device__ __forceinline__ static T apply(T a, T b) { return a + b; }
  template <class T> __host__ __device__ __forceinline__ static T identity() { return T(0); }
};
struct OpMax {
  template <class T> __device__ __forceinline__ static T apply(T a, T b) { return a > b ? a : b; }
  template <class T> __host__ __device__ __forceinline__ static T identity() { return T(0); }  // operands are >= 0
};

// Inclusive scan across one wavefront.
template <class Op, class T>
__device__ __forceinline__ T wave_inclusive(T v) {
  const int l = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T o = __shfl_up(v, d, 64);
    if (l >= d) v = Op::apply(v, o);
  }
  return v;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total.
template <class Op, class T>
__device__ __forceinline__ T block_exclusive(T v, T* total) {
  __shared__ T wsum[kBlock / 64];
  const T inc = wave_inclusive<Op>(v);
  if (lane_id() == 63) wsum[wave_id()] = inc;
  __syncthreads();
  T before = Op::template identity<T>();
  T all = Op::template identity<T>();
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    if (w < wave_id()) before = Op::apply(before, wsum[w]);
    all = Op::apply(all, wsum[w]);
  }
  __syncthreads();
  *total = all;
  // exclusive within the wave: shift the inclusive value up by one lane
  T ex = __shfl_up(inc, 1, 64);
  if (lane_id() == 0) ex = Op::template identity<T>();
  return Op::apply(before, ex);
}

// Pass 1: per-tile reduction.
template <class Op, class Tin, class Tacc>
static __global__ __launch_bounds__(kBlock) void k_scan_reduce(const Tin* __restrict__ in, uint64_t n,
                                                        Tacc* __restrict__ tile_sum) {
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  Tacc acc = Op::template identity<Tacc>();
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) acc = Op::apply(acc, static_cast<Tacc>(in[base + k]));
  Tacc total;
  block_exclusive<Op>(acc, &total);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// Pass 2: one workgroup scans the tile sums (exclusive) and stores the total.
template <class Op, class Tacc>
static __global__ __launch_bounds__(kBlock) void k_scan_tiles(Tacc* __restrict__ tile_sum, uint64_t ntiles,
                                                       Tacc* __restrict__ total_out) {
  Tacc carry = Op::template identity<Tacc>();
  for (uint64_t b = 0; b < ntiles; b += kBlock) {
    const uint64_t i = b + threadIdx.x;
    const Tacc v = i < ntiles ? tile_sum[i] : Op::template identity<Tacc>();
    Tacc total;
    const Tacc ex = block_exclusive<Op>(v, &total);
    if (i < ntiles) tile_sum[i] = Op::apply(carry, ex);
    carry = Op::apply(carry, total);
    __syncthreads();
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// Pass 3: scan each tile with its offset.  inclusive selects the variant.
template <class Op, class Tin, class Tacc, class Tout>
static __global__ __launch_bounds__(kBlock) void k_scan_apply(const Tin* __restrict__ in, uint64_t n,
                                                       const Tacc* __restrict__ tile_off,
                                                       Tout* __restrict__ out, int inclusive) {
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  Tacc v[kScanItems];
  Tacc acc = Op::template identity<Tacc>();
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? static_cast<Tacc>(in[base + k]) : Op::template identity<Tacc>();
    acc = Op::apply(acc, v[k]);
  }
  Tacc total;
  Tacc run = Op::apply(tile_off[blockIdx.x], block_exclusive<Op>(acc, &total));
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (inclusive) run = Op::apply(run, v[k]);
    if (base + k < n) out[base + k] = static_cast<Tout>(run);
    if (!inclusive) run = Op::apply(run, v[k]);
  }
}

// ---------------------------------------------------------------- radix sort

// Digit histogram of one tile: counts[digit * ntiles + tile].
static __global__ __launch_bounds__(kBlock) void k_radix_hist(const uint64_t* __restrict__ keys, uint32_t n, int shift,
                                                       uint32_t mask, uint32_t* __restrict__ counts,
                                                       uint32_t ntiles) {
  __shared__ uint32_t h[kRadixBins];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile + wave_id() * (64 * kSortItems) + lane_id();
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const uint32_t i = base + k * 64;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & mask], 1u);
  }
  __syncthreads();
  counts[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one tile.  Item k of lane l of wave w is tile element
// w*64*kSortItems + k*64 + l; within a wave, items are ranked in (k, lane)
// order from a 8-ballot peer mask, so the sort is stable.  The tile is
// first put in digit order in LDS, then written out with consecutive
// threads on consecutive slots of each digit's run, so that the global
// writes of a digit coalesce.
static __global__ __launch_bounds__(kBlock) void k_radix_scatter(const uint64_t* __restrict__ kin,
                                                          const uint32_t* __restrict__ vin,
                                                          uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                          uint32_t n, int shift, int nbits, uint32_t mask,
                                                          const uint32_t* __restrict__ offsets, uint32_t ntiles) {
  __shared__ uint32_t wcnt[kBlock / 64][kRadixBins];
  __shared__ uint32_t gbase[kRadixBins];
  __shared__ uint32_t tstart[kRadixBins];
  __shared__ uint64_t lkey[kSortTile];
  __shared__ uint32_t lval[kSortTile];
  volatile uint32_t* vw = &wcnt[0][0];
  for (int i = threadIdx.x; i < (kBlock / 64) * kRadixBins; i += kBlock) vw[i] = 0;
  gbase[threadIdx.x] = offsets[threadIdx.x * ntiles + blockIdx.x];
  __syncthreads();
  const int l = lane_id(), w = wave_id();
  const uint64_t lt = (1ull << l) - 1ull;
  const uint32_t base = blockIdx.x * kSortTile + w * (64 * kSortItems) + l;
  uint64_t key[kSortItems];
  uint32_t val[kSortItems], rank[kSortItems], dig[kSortItems];
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const uint32_t i = base + k * 64;
    const bool valid = i < n;
    key[k] = valid ? kin[i] : 0;
    val[k] = valid ? vin[i] : 0;
    const uint32_t d = static_cast<uint32_t>(key[k] >> shift) & mask;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < nbits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t before = __popcll(peers & lt);
    uint32_t prev = 0;
    if (valid) prev = vw[w * kRadixBins + d];
    __builtin_amdgcn_wave_barrier();
    if (valid && before == 0) vw[w * kRadixBins + d] = prev + static_cast<uint32_t>(__popcll(peers));
    __builtin_amdgcn_wave_barrier();
    rank[k] = prev + before;
    dig[k] = d;
  }
  __syncthreads();
  {  // per-digit offsets of each wave inside the digit's run, and where each
     // digit's run starts in the tile
    uint32_t run = 0;
#pragma unroll
    for (int ww = 0; ww < kBlock / 64; ++ww) {
      const uint32_t c = vw[ww * kRadixBins + threadIdx.x];
      vw[ww * kRadixBins + threadIdx.x] = run;
      run += c;
    }
    uint32_t total;
    tstart[threadIdx.x] = block_exclusive<OpAdd>(run, &total);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    if (base + k * 64 < n) {
      const uint32_t at = tstart[dig[k]] + vw[w * kRadixBins + dig[k]] + rank[k];
      lkey[at] = key[k];
      lval[at] = val[k];
    }
  }
  __syncthreads();
  const uint32_t t0 = blockIdx.x * kSortTile;
  const uint32_t tn = n - t0 < static_cast<uint32_t>(kSortTile) ? n - t0 : static_cast<uint32_t>(kSortTile);
  for (uint32_t i = threadIdx.x; i < tn; i += kBlock) {
    const uint64_t k = lkey[i];
    const uint32_t d = static_cast<uint32_t>(k >> shift) & mask;
    const uint32_t pos = gbase[d] + (i - tstart[d]);
    kout[pos] = k;
    vout[pos] = lval[i];
  }
}

// ---------------------------------------------------------------- helpers

static __global__ void k_iota_u32(uint32_t* __restrict__ out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = i;
}

template <class T>
static __global__ void k_fill(T* __restrict__ out, uint64_t n, T v) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = v;
}

static __global__ void k_copy_pairs(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    kout[i] = kin[i];
    vout[i] = vin[i];
  }
}

template <class T>
static __global__ void k_gather(const T* __restrict__ in, const uint32_t* __restrict__ perm, T* __restrict__ out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[perm[i]];
}

inline uint32_t grid_for(uint64_t n, uint32_t per_block) {
  return static_cast<uint32_t>((n + per_block - 1) / per_block);
}

// ---------------------------------------------------------------- host side

// Exclusive (or inclusive) scan of n items into out; total (may be null) is
// a device scalar of type Tacc.  Uses workspace slot `tiles_slot`.
template <class Op, class Tin, class Tacc, class Tout>
void device_scan(Runtime& rt, const Tin* in, uint64_t n, Tout* out, Tacc* total, bool inclusive,
                 const char* tiles_slot = "scan.tiles") {
  const uint64_t ntiles = n == 0 ? 1 : (n + kScanTile - 1) / kScanTile;
  Tacc* tiles = rt.ws().get<Tacc>(tiles_slot, ntiles);
  if (n > 0) k_scan_reduce<Op, Tin, Tacc><<<static_cast<uint32_t>(ntiles), kBlock, 0, rt.stream()>>>(in, n, tiles);
  else k_fill<Tacc><<<1, 64, 0, rt.stream()>>>(tiles, 1, Op::template identity<Tacc>());
  k_scan_tiles<Op, Tacc><<<1, kBlock, 0, rt.stream()>>>(tiles, ntiles, total);
  if (n > 0)
    k_scan_apply<Op, Tin, Tacc, Tout><<<static_cast<uint32_t>(ntiles), kBlock, 0, rt.stream()>>>(in, n, tiles, out,
                                                                                             inclusive ? 1 : 0);
  rt.check_launch("device_scan");
}

// Workspace slots of a radix sort: the sorts of one stage can keep their
// own (the DP stage's task sort, replayed from a captured graph, needs
// addresses no other stage moves).
struct SortSlots {
  const char* k2 = "sort.k2";
  const char* v2 = "sort.v2";
  const char* counts = "sort.counts";
  const char* tiles = "scan.tiles";
};

// Stable sort of (keys, vals) by key bits [0, nbits), `digit` bits (1..8)
// per pass; results land back in keys/vals.  Launches kernels only (no
// copy or fill nodes when captured).
inline void radix_sort_pairs(Runtime& rt, uint64_t* keys, uint32_t* vals, uint32_t n, int nbits,
                             int digit = kRadixBits, const SortSlots& sl = SortSlots{}) {
  if (n <= 1 || nbits <= 0) return;
  if (digit < 1 || digit > kRadixBits) throw ClastError(1, "radix digit out of range");
  const uint32_t ntiles = grid_for(n, kSortTile);
  uint64_t* k2 = rt.ws().get<uint64_t>(sl.k2, n);
  uint32_t* v2 = rt.ws().get<uint32_t>(sl.v2, n);
  uint32_t* counts = rt.ws().get<uint32_t>(sl.counts, static_cast<uint64_t>(ntiles) * kRadixBins);
  uint64_t* kin = keys;
  uint32_t* vin = vals;
  uint64_t* kout = k2;
  uint32_t* vout = v2;
  int passes = 0;
  for (int shift = 0; shift < nbits; shift += digit, ++passes) {
    const int db = nbits - shift < digit ? nbits - shift : digit;
    const uint32_t mask = (1u << db) - 1u;
    k_radix_hist<<<ntiles, kBlock, 0, rt.stream()>>>(kin, n, shift, mask, counts, ntiles);
    // counts are digit-major: the first 2^db rows hold every used digit
    // (the rest stay zero and are never read for a key)
    device_scan<OpAdd, uint32_t, uint32_t, uint32_t>(rt, counts, static_cast<uint64_t>(ntiles) << db, counts,
                                                     static_cast<uint32_t*>(nullptr), false, sl.tiles);
    k_radix_scatter<<<ntiles, kBlock, 0, rt.stream()>>>(kin, vin, kout, vout, n, shift, db, mask, counts, ntiles);
    rt.check_launch("radix pass");
    uint64_t* tk = kin; kin = kout; kout = tk;
    uint32_t* tv = vin; vin = vout; vout = tv;
  }
  if (passes & 1) {
    k_copy_pairs<<<grid_for(n, kBlock), kBlock, 0, rt.stream()>>>(kin, vin, keys, vals, n);
    rt.check_launch("sort copy");
  }
}

// Stable permutation sort by a multi-word key: words[0] least significant.
// perm (n) receives the order.  Each word is gathered through the running
// permutation and radix-sorted with the permutation as the value.
inline void sort_perm(Runtime& rt, uint32_t n, int nwords, const uint64_t* const* words, const int* bits,
                      uint32_t* perm) {
  if (n == 0) return;
  k_iota_u32<<<grid_for(n, kBlock), kBlock, 0, rt.stream()>>>(perm, n);
  uint64_t* tmp = rt.ws().get<uint64_t>("sortperm.key", n);
  for (int w = 0; w < nwords; ++w) {
    if (bits[w] <= 0) continue;
    k_gather<uint64_t><<<grid_for(n, kBlock), kBlock, 0, rt.stream()>>>(words[w], perm, tmp, n);
    radix_sort_pairs(rt, tmp, perm, n, bits[w]);
  }
  rt.check_launch("sort_perm");
}

}  // namespace clast
