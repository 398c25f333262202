// Search-path kernels of libclast_hip for gfx950.
//
// Data layout in HBM (all resident per handle):
//   target index  : tbase (window slot arrays, concatenated, 1 B/slot),
//                   per target sequence gs: slot start, length, window, list ID;
//                   entries sorted by (k-mer, window, position): ekey
//                   (u32 for lMer <= 16 else u64), egs, etidx (4 B each), and
//                   bucket[2^P + 1] = first entry of each P-bit key prefix.
//   query set     : qbase (window slot arrays), per strand s: slot start,
//                   length, list ID, window, first valid k-mer position.
//   seeds         : u64 sort key (s | gs | diagonal + offset) + u32 qIdx.
//   hits          : SoA int32 (s, gs, tIdx, qIdx) -> DP -> 6 int32 outputs.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <type_traits>

#include "prims.hpp"

namespace clast {

// ---------------------------------------------------------------- lanes

// Cross-lane moves as DPP operand modifiers (no LDS round trip).  gfx9
// controls: quad_perm selects within groups of 4 lanes; row_shl:n / row_shr:n
// read lane i + n / i - n of the same row of 16; wave_shr:1 reads lane i - 1;
// row_bcast:15 / :31 broadcast lane 15 of each row to the next row / lane 31
// to rows 2 and 3 (rows picked by the row mask).  A lane whose source lies
// outside its row, or whose row is masked off, keeps `old` (bound_ctrl off).
constexpr int kDppQuadXor1 = 0xB1;  // quad_perm [1, 0, 3, 2]: lane i ^ 1
constexpr int kDppQuadXor2 = 0x4E;  // quad_perm [2, 3, 0, 1]: lane i ^ 2
constexpr int kDppRowShl = 0x100;   // + n
constexpr int kDppRowShr = 0x110;   // + n
constexpr int kDppWaveShr1 = 0x138;
constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;

template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t v) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(static_cast<int>(old), static_cast<int>(v), CTRL, ROWS, 0xF, false));
}

// Value of lane i ^ J (J a power of two; every lane active): DPP for
// J <= 8; for 16 and 32 the gfx950 row / half swaps, which with both
// operands v return v of lane i ^ J in the second register of the pair
// where bit J of i is clear and in the first where it is set (16: odd rows
// of the first operand trade with even rows of the second; 32: its upper 32
// lanes with the lower 32 of the second).
template <int J>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v) {
  static_assert(J == 1 || J == 2 || J == 4 || J == 8 || J == 16 || J == 32, "lane distance");
  if constexpr (J == 1) {
    return dpp_u32<kDppQuadXor1>(v, v);
  } else if constexpr (J == 2) {
    return dpp_u32<kDppQuadXor2>(v, v);
  } else if constexpr (J == 4 || J == 8) {
    const uint32_t up = dpp_u32<kDppRowShl + J>(v, v);  // lanes with bit J clear read i + J
    const uint32_t dn = dpp_u32<kDppRowShr + J>(v, v);  // lanes with bit J set read i - J
    return (lane_id() & J) ? dn : up;
  } else if constexpr (J == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane_id() & J) ? r[0] : r[1];
  } else {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane_id() & J) ? r[0] : r[1];
  }
}

template <int J>
__device__ __forceinline__ uint64_t xor_lane64(uint64_t v) {
  return (static_cast<uint64_t>(xor_lane<J>(static_cast<uint32_t>(v >> 32))) << 32) |
         xor_lane<J>(static_cast<uint32_t>(v));
}

// Value of lane i - 1 (lane 0: `old`).
__device__ __forceinline__ uint32_t lane_up1(uint32_t old, uint32_t v) { return dpp_u32<kDppWaveShr1>(old, v); }
__device__ __forceinline__ uint64_t lane_up1_64(uint64_t v) {
  const uint32_t hi = static_cast<uint32_t>(v >> 32), lo = static_cast<uint32_t>(v);
  return (static_cast<uint64_t>(lane_up1(hi, hi)) << 32) | lane_up1(lo, lo);
}

// Inclusive max scan over the wavefront: within rows of 16, then across
// rows by broadcasts.
__device__ __forceinline__ int wave_max_scan_incl(int v) {
  auto mx = [](int a, uint32_t b) { return static_cast<int>(b) > a ? static_cast<int>(b) : a; };
  const uint32_t lo = static_cast<uint32_t>(INT_MIN);
  v = mx(v, dpp_u32<kDppRowShr + 1>(lo, static_cast<uint32_t>(v)));
  v = mx(v, dpp_u32<kDppRowShr + 2>(lo, static_cast<uint32_t>(v)));
  v = mx(v, dpp_u32<kDppRowShr + 4>(lo, static_cast<uint32_t>(v)));
  v = mx(v, dpp_u32<kDppRowShr + 8>(lo, static_cast<uint32_t>(v)));
  v = mx(v, dpp_u32<kDppRowBcast15, 0xA>(lo, static_cast<uint32_t>(v)));
  v = mx(v, dpp_u32<kDppRowBcast31, 0xC>(lo, static_cast<uint32_t>(v)));
  return v;
}

// OR over the wavefront (the same value in every lane).
__device__ __forceinline__ uint32_t wave_or32(uint32_t v) {
  v |= dpp_u32<kDppRowShr + 1>(0u, v);
  v |= dpp_u32<kDppRowShr + 2>(0u, v);
  v |= dpp_u32<kDppRowShr + 4>(0u, v);
  v |= dpp_u32<kDppRowShr + 8>(0u, v);
  v |= dpp_u32<kDppRowBcast15, 0xA>(0u, v);
  v |= dpp_u32<kDppRowBcast31, 0xC>(0u, v);
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}
__device__ __forceinline__ uint64_t wave_or(uint64_t x) {
  return (static_cast<uint64_t>(wave_or32(static_cast<uint32_t>(x >> 32))) << 32) | wave_or32(static_cast<uint32_t>(x));
}

// ---------------------------------------------------------------- bases

// Base classes.  Every base byte is stored on the device as a class index:
// A 0, C 1, G 2, T 3 (the 2-bit k-mer codes of SRead2bit.cuh:12-18), any
// other letter 4 (never matches, N vs N included), E F I J L O P Q Z 5
// (scores -1), X 6 (bad area) — the 26x26 POINT_MATRIX / MATCH_COUNTER of
// krnlMatrix.cuh:11-67 as a rule.
constexpr uint32_t kClsOther = 4u, kClsStar = 5u, kClsX = 6u;

__device__ __forceinline__ uint32_t dp_class(uint32_t c) {
  const uint32_t idx = c - 'A';
  constexpr uint32_t star = (1u << 4) | (1u << 5) | (1u << 8) | (1u << 9) | (1u << 11) | (1u << 14) | (1u << 15) |
                            (1u << 16) | (1u << 25);  // E F I J L O P Q Z
  uint32_t v = (idx < 26u && ((star >> idx) & 1u)) ? kClsStar : kClsOther;
  v = c == 'X' ? kClsX : v;
  v = c == 'A' ? 0u : v;
  v = c == 'C' ? 1u : v;
  v = c == 'G' ? 2u : v;
  v = c == 'T' ? 3u : v;
  return v;
}

// 2-bit code of a class (A0 C1 G2 T3, SRead2bit.cuh:12-18); 4 or more for
// anything that is not A/C/G/T (SRepelHasOddBase.cuh:6-14 makes the k-mer -1).
__device__ __forceinline__ uint32_t base_code(uint32_t cls) { return cls; }

// Bytes -> class codes, in place (once per upload).
static __global__ __launch_bounds__(kBlock) void k_classify(uint8_t* __restrict__ p, uint64_t n) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = static_cast<uint8_t>(dp_class(p[i]));
}

// Query strands built on the device from the reads as given
// (addSequence, utilAddSequence.cu:6-47, as CHostSeqList_query uses it,
// CHostSeqList_query.cu:11-54): read k becomes its '+' strand at slot
// slot[k] and its reverse complement ('-': A<->T, C<->G, any other byte N)
// right after, each followed by lMer-1 joint bases (the strand's own first
// bases, N past a short strand's end); written as class codes.  One
// wavefront per read; `reads` holds the bytes from offset b0 on, off the
// reads' offsets (nreads + 1), qslot the strand slots (read k's '+' strand
// at qslot[2 k]).
static __global__ __launch_bounds__(kBlock) void k_encode_reads(const uint8_t* __restrict__ reads,
                                                                const int64_t* __restrict__ off, int64_t b0,
                                                                uint32_t nreads, const uint32_t* __restrict__ qslot,
                                                                int lmer, uint8_t* __restrict__ qbase) {
  const uint32_t k = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (k >= nreads) return;
  const uint32_t lane = threadIdx.x & 63u;
  const uint8_t* r = reads + (off[k] - b0);
  const int64_t len = off[k + 1] - off[k];
  const int64_t span = len + max(lmer - 1, 0);  // slots per strand
  uint8_t* plus = qbase + qslot[2 * static_cast<uint64_t>(k)];
  uint8_t* minus = plus + span;
  // kU slots per lane and round, their loads issued together (a long read's
  // wavefront waits for one load latency per round, not per slot)
  constexpr int kU = 8;
  for (int64_t i0 = lane; i0 < span; i0 += 64 * kU) {
    uint8_t bp[kU], bm[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t i = i0 + 64 * u;
      // the strand position whose base slot i holds (joint slots repeat the
      // first bases); none past a short strand's end or past the span
      const int64_t src = i < len ? i : i - len;
      const bool has = i < span && src < len;
      bp[u] = has ? r[src] : 'N';
      // '-' base at src is the complement of read base len - 1 - src
      bm[u] = has ? r[len - 1 - src] : 'N';
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t i = i0 + 64 * u;
      if (i >= span) break;
      const uint32_t cp = dp_class(bp[u]);
      // the complement keeps A C G T (as T G C A) and turns every other byte into N
      const uint32_t cr = dp_class(bm[u]);
      const uint32_t cm = cr < 4u ? 3u - cr : kClsOther;
      plus[i] = static_cast<uint8_t>(cp);
      minus[i] = static_cast<uint8_t>(cm);
    }
  }
}

// clean[k] = 1 when the bytes [slot[k], slot[k] + len[k]) of a class array
// hold no star or X class (scores and match counts of such a sequence's
// cells stay within 16 bits; see k_dp16), 3 when with other_bit they also
// hold no class 4 (k_dp16's query profiles), else 0.  One wavefront per
// sequence, four aligned 16-byte blocks per lane and step; bytes outside the
// range are masked off.
template <class SlotT>
static __global__ __launch_bounds__(kBlock) void k_clean_flags(const uint8_t* __restrict__ base,
                                                               const SlotT* __restrict__ slot,
                                                               const int32_t* __restrict__ len, uint32_t n,
                                                               int other_bit, uint8_t* __restrict__ clean) {
  const uint32_t k = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (k >= n) return;
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t a = static_cast<uint64_t>(slot[k]);
  const uint64_t e = a + static_cast<uint64_t>(max(len[k], 0));
  const uint4* blk = reinterpret_cast<const uint4*>(base);
  uint32_t bad = 0, other = 0;
  for (uint64_t b0 = (a >> 4) + lane; b0 * 16 < e; b0 += 256) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint64_t b = b0 + 64u * u;
      if (b * 16 >= e) break;
      const uint4 v = blk[b];
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const uint64_t p = b * 16 + 4u * d;  // first byte of the dword
        uint32_t in = 0xFFFFFFFFu;
#pragma unroll
        for (int y = 0; y < 4; ++y)
          if (p + y < a || p + y >= e) in &= ~(0xFFu << (8 * y));
        // bytes of classes 5 and 6 (the only ones above 4) get bit 3 set by
        // +3; bit 2 is set in classes 4, 5 and 6
        bad |= (w[d] + 0x03030303u) & 0x08080808u & in;
        other |= w[d] & 0x04040404u & in;
      }
    }
  }
  const bool any = __builtin_amdgcn_ballot_w64(bad != 0u) != 0ull;
  const bool any4 = __builtin_amdgcn_ballot_w64(other != 0u) != 0ull;
  if (lane == 0) clean[k] = any ? 0 : other_bit && !any4 ? 3 : 1;
}

template <class T>
__device__ __forceinline__ uint32_t upper_bound_u(const T* a, uint32_t lo, uint32_t hi, T v) {
  // first index in [lo, hi) with a[i] > v
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
template <class T>
__device__ __forceinline__ uint32_t lower_bound_u(const T* a, uint32_t lo, uint32_t hi, T v) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------- index build

// Index code of a valid k-mer x (2-bit MSB-first code, rc its reverse
// complement).  For odd lMer a bijection of the 2 lMer-bit codes that puts
// x and rc side by side: of the two, take c, the one whose middle base is A
// or C (exactly one is, the complement of A/C being T/G), drop that base's
// high bit (bit lMer, always 0 in c) and append which of the two x is.  A
// query strand and its reverse complement then look up the same bitmap word
// and the same cell line at mirrored positions.  Even lMer keeps the plain
// code (a k-mer can be its own reverse complement).  Cells, entry order and
// the repeat cut only need some bijection: entries of one k-mer still form
// one run, in position order.
__host__ __device__ __forceinline__ uint64_t index_code(uint64_t x, uint64_t rc, int lmer) {
  if (!(lmer & 1)) return x;
  const uint64_t hb = 1ull << lmer;
  const uint64_t o = (x >> lmer) & 1ull;
  const uint64_t c = o ? rc : x;
  return (((c & (hb - 1ull)) | ((c >> 1) & ~(hb - 1ull))) << 1) | o;
}

// 2-bit codes of the 4 class bytes of a little-endian word, first byte in
// the top bits (the MSB-first order of SRead2bit.cuh:12-18).
__device__ __forceinline__ uint32_t pack_codes4(uint32_t x) {
  const uint32_t y = x & 0x03030303u;
  return ((y << 6) & 0xC0u) | ((y >> 4) & 0x30u) | ((y >> 14) & 0x0Cu) | (y >> 24);
}
// Bit k set when byte k of the word is not A, C, G or T (classes 4..6 have
// bit 2 set, 0..3 do not).
__device__ __forceinline__ uint32_t bad_bytes4(uint32_t x) {
  const uint32_t b = (x >> 2) & 0x01010101u;
  return (b | (b >> 7) | (b >> 14) | (b >> 21)) & 0xFu;
}

// K-mer of lmer (1..31) class bytes at byte offset a of base, read as the 9
// aligned words that hold it (the buffer extends 36 bytes past the last
// k-mer): its code and reverse-complement code; false if a byte is not
// A, C, G or T (SRepelHasOddBase.cuh:7-13).
__device__ __forceinline__ bool kmer_words(const uint8_t* __restrict__ base, uint64_t a, int lmer, uint64_t& key,
                                           uint64_t& rc) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(base) + (a >> 2);
  uint32_t x[9];
#pragma unroll
  for (int d = 0; d < 9; ++d) x[d] = w[d];
  uint64_t h = 0, bad = 0;
#pragma unroll
  for (int d = 0; d < 8; ++d) {
    h = (h << 8) | pack_codes4(x[d]);
    bad |= static_cast<uint64_t>(bad_bytes4(x[d])) << (4 * d);
  }
  bad |= static_cast<uint64_t>(bad_bytes4(x[8])) << 32;
  const int o = static_cast<int>(a & 3u);
  h = (h << (2 * o)) | (pack_codes4(x[8]) >> (8 - 2 * o));  // bases o .. o + 31 of the 36 read
  key = h >> (64 - 2 * lmer);
  // reverse the 2-bit groups, then complement (A<->T, C<->G)
  uint64_t r = __builtin_bitreverse64(key);
  r = ((r >> 1) & 0x5555555555555555ull) | ((r & 0x5555555555555555ull) << 1);
  const uint64_t mask = (1ull << (2 * lmer)) - 1ull;
  rc = (r >> (64 - 2 * lmer)) ^ mask;
  return ((bad >> o) & ((1ull << lmer) - 1ull)) == 0;
}

// Strided target k-mers (CDeviceHashTable.cu:20-85): element g of window w
// is position p = (g - win_stride_off[w]) * stride of that window; a valid
// k-mer gets its index code, an invalid one (non-ACGT or running past the
// window end) `invalid`, which sorts after every valid code.
static __global__ __launch_bounds__(kBlock) void k_target_keys(const uint8_t* __restrict__ tbase,
                                                        const uint64_t* __restrict__ win_slot_off,
                                                        const uint32_t* __restrict__ win_stride_off, int nwin,
                                                        int lmer, int stride, uint32_t ns, uint64_t invalid,
                                                        uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= ns) return;
  const uint32_t w = upper_bound_u(win_stride_off, 0u, static_cast<uint32_t>(nwin + 1), g) - 1;
  const uint64_t p = static_cast<uint64_t>(g - win_stride_off[w]) * stride;
  const uint64_t nw = win_slot_off[w + 1] - win_slot_off[w];
  const uint8_t* b = tbase + win_slot_off[w] + p;
  const int top = 2 * (lmer - 1);
  uint64_t key = 0, rc = 0;
  bool ok = true;
  for (int j = 0; j < lmer; ++j) {
    const uint32_t c = (p + j < nw) ? base_code(b[j]) : 4u;
    ok = ok && c < 4u;
    key = (key << 2) | (c & 3u);
    rc = (rc >> 2) | (static_cast<uint64_t>(3u - (c & 3u)) << top);
  }
  keys[g] = ok ? index_code(key, rc, lmer) : invalid;
  vals[g] = g;
}

__device__ __forceinline__ uint32_t window_of(const uint32_t* win_stride_off, int nwin, uint32_t g) {
  return upper_bound_u(win_stride_off, 0u, static_cast<uint32_t>(nwin + 1), g) - 1;
}

// Heads of (key, window) runs over the sorted strided k-mers, and the count
// of valid ones (*nvalid; invalid codes sort last).
static __global__ __launch_bounds__(kBlock) void k_index_heads(const uint64_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ vals,
                                                        const uint32_t* __restrict__ win_stride_off, int nwin,
                                                        uint32_t ns, uint64_t invalid, uint32_t* __restrict__ head,
                                                        uint32_t* __restrict__ nvalid) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= ns) return;
  const uint64_t k = keys[j];
  const bool valid = k < invalid;
  uint32_t h = 0;
  if (valid) {
    if (j == 0) {
      h = 1;
    } else {
      const uint64_t pk = keys[j - 1];
      h = (pk != k || window_of(win_stride_off, nwin, vals[j - 1]) != window_of(win_stride_off, nwin, vals[j])) ? 1u : 0u;
    }
    if (j + 1 == ns || keys[j + 1] >= invalid) *nvalid = j + 1;
  }
  head[j] = h;
}

static __global__ __launch_bounds__(kBlock) void k_index_runstart(const uint32_t* __restrict__ head,
                                                           const uint32_t* __restrict__ run_ex, uint32_t nvalid,
                                                           uint32_t* __restrict__ run_start) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= nvalid) return;
  if (head[j]) run_start[run_ex[j]] = j;
  if (j + 1 == nvalid) run_start[run_ex[j] + head[j]] = nvalid;
}

// Repeat cut per (key, window) cell (CDeviceHashTable.cu:122-137: counts
// taken over every strided position of the cell) and the target-boundary
// seed filter (…deleteSeedsOnSequenceBoundary.cu:74-127: tIdx + lMer > tLen
// never seeds) applied once here instead of per query.
static __global__ __launch_bounds__(kBlock) void k_index_keep(const uint32_t* __restrict__ vals,
                                                       const uint32_t* __restrict__ head,
                                                       const uint32_t* __restrict__ run_ex,
                                                       const uint32_t* __restrict__ run_start, uint32_t nvalid,
                                                       const uint64_t* __restrict__ win_slot_off,
                                                       const uint32_t* __restrict__ win_stride_off,
                                                       const uint32_t* __restrict__ win_seq_off, int nwin,
                                                       const uint64_t* __restrict__ seq_slot,
                                                       const int32_t* __restrict__ seq_len, int lmer, int stride,
                                                       int cut_repeat, uint32_t* __restrict__ keep,
                                                       int32_t* __restrict__ tmp_gs, int32_t* __restrict__ tmp_tidx) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= nvalid) return;
  const uint32_t r = run_ex[j] - (head[j] ? 0u : 1u);
  const uint32_t len = run_start[r + 1] - run_start[r];
  bool k = cut_repeat == -1 || static_cast<int64_t>(len) < cut_repeat;
  const uint32_t g = vals[j];
  const uint32_t w = window_of(win_stride_off, nwin, g);
  const uint64_t slot = win_slot_off[w] + static_cast<uint64_t>(g - win_stride_off[w]) * stride;
  const uint32_t gs = upper_bound_u(seq_slot, win_seq_off[w], win_seq_off[w + 1], slot) - 1;
  const int64_t tidx = static_cast<int64_t>(slot - seq_slot[gs]);
  k = k && tidx + lmer <= seq_len[gs];
  keep[j] = k ? 1u : 0u;
  tmp_gs[j] = static_cast<int32_t>(gs);
  tmp_tidx[j] = static_cast<int32_t>(tidx);
}

// Kept entries in (key, window, position) order: full key (u64, for the
// cell build; narrowed to u32 for lMer <= 16 afterwards) and (seq, tIdx).
static __global__ __launch_bounds__(kBlock) void k_index_scatter(const uint64_t* __restrict__ keys,
                                                                 const uint32_t* __restrict__ keep,
                                                                 const uint32_t* __restrict__ off,
                                                                 const int32_t* __restrict__ tmp_gs,
                                                                 const int32_t* __restrict__ tmp_tidx, uint32_t nvalid,
                                                                 uint64_t* __restrict__ ekey, int2* __restrict__ epos) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j >= nvalid || !keep[j]) return;
  const uint32_t o = off[j];
  ekey[o] = keys[j];
  epos[o] = make_int2(tmp_gs[j], tmp_tidx[j]);
}

static __global__ __launch_bounds__(kBlock) void k_narrow_keys(const uint64_t* __restrict__ in, uint32_t n,
                                                               uint32_t* __restrict__ out) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  if (j < n) out[j] = static_cast<uint32_t>(in[j]);
}

// Presence bitmap over the top B key bits (B = min(2 lMer, 30)): bit h is set
// when some entry's key prefix is h; rank[b] is the number of set bits before
// 512-bit block b.  The r-th present prefix has a 16-byte cell record
// {first entry, entry count, target sequence and tIdx of the first entry}:
// a query k-mer whose bit is clear costs one bitmap line (the bitmap stays in
// the MI355X Infinity Cache: 128 MiB at B = 30), a present one a single
// record line more when its cell holds one entry, the common case.
static __global__ __launch_bounds__(kBlock) void k_cell_heads(const uint64_t* __restrict__ keys, uint32_t n,
                                                              int shift, uint32_t* __restrict__ head) {
  const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= n) return;
  head[e] = (e == 0 || (keys[e] >> shift) != (keys[e - 1] >> shift)) ? 1u : 0u;
}

static __global__ __launch_bounds__(kBlock) void k_cell_build(const uint64_t* __restrict__ keys, uint32_t n,
                                                              int shift, const uint32_t* __restrict__ head,
                                                              const uint32_t* __restrict__ cell,
                                                              uint32_t* __restrict__ cell_off,
                                                              unsigned long long* __restrict__ bitmap) {
  const uint32_t e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= n) return;
  if (head[e]) {
    const uint64_t h = keys[e] >> shift;
    cell_off[cell[e]] = e;
    atomicOr(&bitmap[h >> 6], 1ull << (h & 63));
  }
  if (e + 1 == n) cell_off[cell[e] + head[e]] = n;
}

// Record of the r-th present prefix, stored at r (bitmap form) or at the
// prefix value itself (direct form: a table over every prefix, so a present
// k-mer costs one record read and no rank directory).
static __global__ __launch_bounds__(kBlock) void k_cell_records(const uint32_t* __restrict__ cell_off, uint32_t ncell,
                                                                const int2* __restrict__ epos,
                                                                const uint64_t* __restrict__ keys, int shift,
                                                                int direct, uint4* __restrict__ rec) {
  const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= ncell) return;
  const uint32_t lo = cell_off[r];
  const int2 e = epos[lo];
  const uint64_t at = direct ? keys[lo] >> shift : r;
  rec[at] = make_uint4(lo, cell_off[r + 1] - lo, static_cast<uint32_t>(e.x), static_cast<uint32_t>(e.y));
}

static __global__ __launch_bounds__(kBlock) void k_rank_count(const unsigned long long* __restrict__ bitmap,
                                                              uint64_t nwords, uint64_t nblk,
                                                              uint32_t* __restrict__ cnt) {
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (b >= nblk) return;
  uint32_t c = 0;
  for (uint64_t w = b * 8; w < b * 8 + 8 && w < nwords; ++w) c += __popcll(bitmap[w]);
  cnt[b] = c;
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

struct Lookup {
  const unsigned long long* bitmap;
  const uint32_t* rank;
  const uint4* cells;
  const void* ekey;   // u32 or u64 keys; read only when the prefix is not the whole key
  const int2* epos;
  int shift;          // 2 lMer - B
  int key64;
  int direct;         // cells is indexed by the prefix itself (no rank directory)
};

// Entries of a valid query k-mer (krnlBinarySearch.cu:3-33 + the cell
// gathers of …createRawSeedList.cu:63-135): returns their count n; when
// n > 0, [lo, lo + n) is their range and first the first entry.
__device__ __forceinline__ uint32_t lookup(const Lookup& L, uint64_t key, uint32_t& lo, int2& first) {
  const uint64_t h = key >> L.shift;
  lo = 0;
  first = make_int2(0, 0);
  // presence first: the bitmap stays in the Infinity Cache, so an absent
  // k-mer never reaches HBM
  const uint64_t w = h >> 6;
  const unsigned long long word = L.bitmap[w];
  const int bit = static_cast<int>(h & 63);
  if (!((word >> bit) & 1ull)) return 0;
  uint4 c;
  if (L.direct) {
    // streamed past the caches: the bitmap is what should stay resident
    const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(&L.cells[h]));
    c = make_uint4(v.x, v.y, v.z, v.w);
  } else {
    uint32_t r = L.rank[w >> 3];
    for (uint64_t j = w & ~7ull; j < w; ++j) r += static_cast<uint32_t>(__popcll(L.bitmap[j]));
    r += static_cast<uint32_t>(__popcll(word & ((1ull << bit) - 1ull)));
    c = L.cells[r];
  }
  lo = c.x;
  uint32_t n = c.y;
  first = make_int2(static_cast<int>(c.z), static_cast<int>(c.w));
  if (L.shift) {  // the prefix cell holds several keys: narrow to this one
    uint32_t hi = lo + n;
    if (L.key64) {
      const uint64_t* k = static_cast<const uint64_t*>(L.ekey);
      lo = lower_bound_u(k, lo, hi, key);
      hi = upper_bound_u(k, lo, hi, key);
    } else {
      const uint32_t* k = static_cast<const uint32_t*>(L.ekey);
      lo = lower_bound_u(k, lo, hi, static_cast<uint32_t>(key));
      hi = upper_bound_u(k, lo, hi, static_cast<uint32_t>(key));
    }
    n = hi - lo;
    if (n) first = L.epos[lo];
  }
  return n;
}

// ---------------------------------------------------------------- seeding

struct SeedArgs {
  const uint8_t* qbase;
  const uint32_t* qslot;
  const int32_t* qlen;
  const int32_t* qwin;        // query window of each strand
  const uint32_t* pos_off;    // nseq + 1
  const uint32_t* pos_strand; // strand of each k-mer position
  uint32_t nseq, npos;
  Lookup lk;
  const int2* epos;           // (target sequence, intra-sequence index) per entry
  const int2* tinfo;          // per target sequence: (length, window)
  int nqwin;
  int lmer, gap, width, local;
  uint32_t* unit_first;       // per (target window, query window): first strand with a seed
  uint64_t* wmask;            // per strand: target windows it keeps seeds in (at most 64 windows), or null
};

// Strand of every k-mer position of a query set, written once when the set
// is uploaded: one wavefront per strand.
static __global__ __launch_bounds__(kBlock) void k_pos_strand(const uint32_t* __restrict__ pos_off, uint32_t nseq,
                                                              uint32_t* __restrict__ pos_strand) {
  const uint32_t s = blockIdx.x * (kBlock / 64) + wave_id();
  if (s >= nseq) return;
  for (uint32_t p = pos_off[s] + lane_id(); p < pos_off[s + 1]; p += 64) pos_strand[p] = s;
}

// corner filter (…deleteSeedsOnSequenceBoundary.cuh:31-49), global mode only
__device__ __forceinline__ bool corner_ok(const SeedArgs& a, int d, int qlen, int tlen) {
  return a.local || !(d + a.gap < 0 || d + qlen - a.gap > tlen);
}

// Records strand s as a candidate first strand of unit (target window tw,
// query window qw) — the strand of the unit's element 0, which the isolate
// filter keeps (…deleteIsolateSeeds.cu:28-76).  Values only decrease, so a
// stale read only costs an atomic.
__device__ __forceinline__ void note_unit_first(const SeedArgs& a, int tw, int qw, uint32_t s) {
  uint32_t* uf = &a.unit_first[tw * a.nqwin + qw];
  if (s < *uf) atomicMin(uf, s);
}

// Each unit's first strand from the strands' window masks (k_seed_segments
// with SeedArgs::wmask): thread per strand; a wavefront whose strands share
// one query window takes, per window bit, its first lane with the bit and
// makes one candidate of it, so candidates come per wavefront, not per
// strand (a read of unit_first before each atomic per strand costs a
// dependent load on the strand's path, and atomics on these few addresses
// contend).  Strands ascend with the lanes.
static __global__ __launch_bounds__(kBlock) void k_unit_first(const uint64_t* __restrict__ wmask,
                                                              const int32_t* __restrict__ qwin, uint32_t nseq,
                                                              int nwin, int nqwin, uint32_t* __restrict__ unit_first) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = s < nseq;
  const uint64_t m = valid ? wmask[s] : 0ull;
  const int qw = valid ? qwin[s] : -1;
  const int lane = lane_id();
  const int qw0 = __shfl(qw, 0, 64);
  if (__all(!valid || qw == qw0)) {
    uint32_t cand = 0xFFFFFFFFu;
    for (int b = 0; b < nwin; ++b) {
      const unsigned long long bm = __ballot(((m >> b) & 1ull) != 0ull);
      if (bm && lane == b) cand = s - static_cast<uint32_t>(lane) + static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(bm)) - 1);
    }
    if (cand != 0xFFFFFFFFu) {
      uint32_t* uf = &unit_first[lane * nqwin + qw0];
      if (cand < *uf) atomicMin(uf, cand);
    }
  } else {
    for (uint64_t r = m; r; r &= r - 1ull) {
      uint32_t* uf = &unit_first[(__ffsll(static_cast<unsigned long long>(r)) - 1) * nqwin + qw];
      if (s < *uf) atomicMin(uf, s);
    }
  }
}

// Query k-mer lookup per position (krnlBinarySearch.cu:3-33 + the cell
// gathers of …createRawSeedList.cu:63-135): the cell record {first entry,
// entry count, first entry's (seq, tIdx)} of its k-mer, and per strand the
// number of matching entries (its raw seed count, before the corner
// filter).  Only positions with qIdx + lMer <= qLen exist here: the
// query-boundary filter (…deleteSeedsOnSequenceBoundary.cu:19-72) removes
// every other one.
//
// Thread t takes the t-th position of its strand pair (strands 2k, 2k + 1).
// A pair of equal lengths is walked interleaved: '+' position j, then the
// '-' position m - 1 - j, which reads the reverse complement of the same
// bases, so with index codes (index_code) the two lanes look up one bitmap
// word and one cell line and the wavefront issues a single request for
// both.  Any other pair is walked in order.
static __global__ __launch_bounds__(kBlock) void k_seed_count(SeedArgs a, uint4* __restrict__ cellp,
                                                              uint32_t* __restrict__ scnt) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = t < a.npos;
  uint32_t s = 0xFFFFFFFFu, n = 0;
  if (valid) {
    s = __builtin_nontemporal_load(&a.pos_strand[t]);
    uint32_t qIdx = t - a.pos_off[s];
    const uint32_t s0 = s & ~1u;
    if (s0 + 1u < a.nseq) {
      const uint32_t p0 = a.pos_off[s0], p1 = a.pos_off[s0 + 1], p2 = a.pos_off[s0 + 2];
      if (p1 - p0 == p2 - p1) {
        const uint32_t u = t - p0, j = u >> 1;
        s = s0 + (u & 1u);
        qIdx = (u & 1u) ? p2 - p1 - 1u - j : j;
      }
    }
    uint64_t key, rc;
    const bool ok = kmer_words(a.qbase, static_cast<uint64_t>(a.qslot[s]) + qIdx, a.lmer, key, rc);
    uint32_t lo = 0;
    int2 first = make_int2(0, 0);
    if (ok) n = lookup(a.lk, index_code(key, rc, a.lmer), lo, first);
    u32x4 v;
    v.x = lo;
    v.y = n;
    v.z = static_cast<uint32_t>(first.x);
    v.w = static_cast<uint32_t>(first.y);
    // read again only by the next kernel
    __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(&cellp[a.pos_off[s] + qIdx]));
  }
  // strand totals: among the even lanes, and among the odd ones, each
  // strand's lanes are contiguous (pairs alternate, other strands run in
  // order); segmented inclusive scan over each half, whose last lane of a
  // strand adds the sum
  const int lane = lane_id();
  uint32_t v = n;
#pragma unroll
  for (int d = 2; d < 64; d <<= 1) {
    const uint32_t up = __shfl_up(v, d, 64);
    const uint32_t su = __shfl_up(s, d, 64);
    if (lane >= d && su == s) v += up;
  }
  const uint32_t sn = __shfl_down(s, 2, 64);
  if (valid && v && (lane >= 62 || sn != s)) atomicAdd(&scnt[s], v);
}

// hasNearPair core (…deleteDuplicateSeeds.cuh:29-38) for seeds of one
// (qID, tID) group; (pt, pd) is the predecessor.
__device__ __forceinline__ bool near_pair(int t, int d, int pt, int pd, int width, int gap) {
  return width >= t - pt && width >= pt - t && gap >= pd - d;
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t wave_sum_incl(uint32_t v) {
  const int l = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(v, d, 64);
    if (l >= d) v += o;
  }
  return v;
}

// Seeds of one strand sorted in LDS by one wavefront: the pass over every
// strand reserves room for kSegSmall (occupancy is bounded by LDS, and most
// strands hold far fewer); strands over it run in a pass with room for
// kSegBig; strands over that take the global path.
constexpr int kSegSmall = 256;
constexpr int kSegBig = 2048;
constexpr int kSegWaves = kBlock / 64;

struct SegOut {
  int32_t* hs;            // hits of strand s go to [soff[s], ...), the
  int32_t* hgs;           // strand's own seed range (a strand has no more
  int32_t* ht;            // hits than seeds)
  int32_t* hq;
  uint32_t* hcnt;         // hits per strand
  uint32_t* big;          // strands over kSegSmall seeds (first pass)
  uint32_t* nbig;
  uint32_t* ovf;          // strands over kSegBig seeds (global path)
  uint32_t* novf;
  uint8_t* ovf_flag;      // per strand
};

// LDS image of one strand's seeds.  Packed: one u64 (target window, target
// sequence, diagonal + d_off, qIdx) when the four fit in 63 bits — the
// window is monotone in the sequence index, so it does not change the order;
// otherwise (target sequence, diagonal + d_off) and qIdx apart.  Either sorts
// as (seq, diagonal, qIdx); ~0 pads a list to a power of two and sorts last.
template <bool PACKED, int CAP>
struct SegLds;
template <int CAP>
struct SegLds<true, CAP> {
  static constexpr bool kPacked = true;
  uint64_t k[kSegWaves][CAP];
};
template <int CAP>
struct SegLds<false, CAP> {
  static constexpr bool kPacked = false;
  uint64_t k[kSegWaves][CAP];
  int32_t q[kSegWaves][CAP];
};

struct SegPack {
  int d_off;
  int qbits, dbits, gbits;  // packed layout field widths
};

// One seed: target window and sequence, diagonal, qIdx.
struct SeedV {
  int tw, gs, d, q;
};

template <class Lds>
__device__ __forceinline__ void seg_put(Lds& L, int w, uint32_t at, const SegPack& pk, const SeedV& v) {
  if constexpr (Lds::kPacked) {
    L.k[w][at] = (static_cast<uint64_t>(v.tw) << (pk.gbits + pk.dbits + pk.qbits)) |
                 (static_cast<uint64_t>(v.gs) << (pk.dbits + pk.qbits)) |
                 (static_cast<uint64_t>(v.d + pk.d_off) << pk.qbits) | static_cast<uint64_t>(v.q);
  } else {
    L.k[w][at] = (static_cast<uint64_t>(v.gs) << 32) | static_cast<uint32_t>(v.d + pk.d_off);
    L.q[w][at] = v.q;
  }
}

template <class Lds>
__device__ __forceinline__ void seg_none(Lds& L, int w, uint32_t at) {
  L.k[w][at] = ~0ull;
  if constexpr (!Lds::kPacked) L.q[w][at] = INT_MAX;
}

__device__ __forceinline__ SeedV seg_decode(uint64_t k, const SegPack& pk) {
  SeedV v;
  v.tw = static_cast<int>(k >> (pk.gbits + pk.dbits + pk.qbits));
  v.gs = static_cast<int>((k >> (pk.dbits + pk.qbits)) & ((1ull << pk.gbits) - 1ull));
  v.d = static_cast<int>((k >> pk.qbits) & ((1ull << pk.dbits) - 1ull)) - pk.d_off;
  v.q = static_cast<int>(k & ((1ull << pk.qbits) - 1ull));
  return v;
}

// Seed i of a sorted LDS list (the unpacked form reads the window).
template <class Lds>
__device__ __forceinline__ SeedV seg_get(const Lds& L, int w, uint32_t i, const SegPack& pk, const int2* tinfo) {
  const uint64_t k = L.k[w][i];
  if constexpr (Lds::kPacked) {
    return seg_decode(k, pk);
  } else {
    SeedV v;
    v.gs = static_cast<int>(k >> 32);
    v.d = static_cast<int>(static_cast<uint32_t>(k)) - pk.d_off;
    v.q = L.q[w][i];
    v.tw = tinfo[v.gs].y;
    return v;
  }
}

enum SegMode { kSegFirst = 0, kSegBigPass = 1, kSegRerun = 2 };

// Bitonic sort of one u64 per lane, lanes [0, N) ascending.  Lanes from N
// on must hold ~0 (with N < 64 they never exchange with the lower lanes).
// A lane takes its partner's key o when (o < k) equals its keep-the-min
// flag (equal keys give the same result either way): one compare and one
// select per stage.
template <int KK, int J>
__device__ __forceinline__ uint64_t bitonic_merge(uint64_t k, int lane) {
  const uint64_t o = xor_lane64<J>(k);
  const bool keep_min = ((lane & KK) == 0) == ((lane & J) == 0);
  k = (o < k) == keep_min ? o : k;
  if constexpr (J > 1) return bitonic_merge<KK, J / 2>(k, lane);
  else return k;
}
template <int N, int KK = 2>
__device__ __forceinline__ uint64_t wave_sort(uint64_t k) {
  k = bitonic_merge<KK, KK / 2>(k, lane_id());
  if constexpr (KK < N) return wave_sort<N, KK * 2>(k);
  else return k;
}
// Stages J .. 1 of a bitonic merge over the wavefront's 64 lanes (a bitonic
// sequence in, sorted out), ascending when asc.
template <int J>
__device__ __forceinline__ uint64_t wave_merge(uint64_t k, int lane, bool asc) {
  const uint64_t o = xor_lane64<J>(k);
  const bool keep_min = ((lane & J) == 0) == asc;
  k = (o < k) == keep_min ? o : k;
  if constexpr (J > 1) return wave_merge<J / 2>(k, lane, asc);
  else return k;
}

// Isolate-filter flags of one sorted seed v with predecessor pv (has_pred:
// v is not the list's first): group head (new target sequence), window head
// (new target window) and near(v) (…deleteIsolateSeeds.cuh:28-37).
struct SeedMarks {
  bool ghead, twhead, nr;
};
__device__ __forceinline__ SeedMarks seed_marks(const SeedArgs& a, const SeedV& v, const SeedV& pv, bool has_pred) {
  SeedMarks m{true, true, false};
  if (has_pred) {
    if (pv.gs == v.gs) {
      m.ghead = false;
      m.twhead = false;
      m.nr = near_pair(v.d + v.q, v.d, pv.d + pv.q, pv.d, a.width, a.gap);
    } else {
      m.twhead = pv.tw != v.tw;
    }
  }
  return m;
}

// One wavefront per query strand: its seeds (positions in qIdx order, cells
// in index order; the corner filter applied) go to LDS, are sorted by
// (target sequence, diagonal, qIdx) — the …sortSeeds.cuh:14-55 order
// restricted to one qID; in registers for up to 64 seeds, bitonic in LDS
// above — and the isolate and duplicate filters run on the sorted list with
// wave scans: K(i) = near(i) or i is its unit's element 0; D(i) = K(i) and
// not near(i, previous K of the group).  Surviving seeds are written as hits
// at the start of the strand's own seed range (their order is fixed later by
// complete sort keys).
//
// Element 0 of a unit is the first seed of the unit's first strand with a
// seed in that target window.  The first pass (every strand; strands over
// CAP go to the big list) and the big pass (the big list; strands over CAP
// go to the global path) take no seed as element 0 and record the strand as
// a candidate: its window mask (SeedArgs::wmask, reduced by k_unit_first)
// or, with more than 64 target windows, unit_first itself (read first so
// the atomic is rare); the strands that end up first in some unit are then
// run again (kSegRerun) with unit_first final.
template <bool PACKED, int CAP>
static __global__ __launch_bounds__(kBlock) void k_seed_segments(SeedArgs a, const uint4* __restrict__ cellp,
                                                                 const uint32_t* __restrict__ scnt,
                                                                 const uint32_t* __restrict__ soff, SegPack pk,
                                                                 SegOut o, const uint32_t* __restrict__ list,
                                                                 const uint32_t* __restrict__ nlist, int mode) {
  __shared__ SegLds<PACKED, CAP> L;
  // wave-uniform strand index: the per-strand header comes through scalar loads
  const int w = __builtin_amdgcn_readfirstlane(wave_id()), lane = lane_id();
  const bool rerun = mode == kSegRerun;
  const uint32_t count = list ? *nlist : a.nseq;
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint32_t* __restrict__ pos_off = a.pos_off;
  const int32_t* __restrict__ qlen = a.qlen;
  const int32_t* __restrict__ qwin = a.qwin;
  for (uint32_t k = blockIdx.x * kSegWaves + w; k < count; k += gridDim.x * kSegWaves) {
    const uint32_t s = __builtin_amdgcn_readfirstlane(list ? list[k] : k);
    const uint32_t n = scnt[s];
    if (n == 0) {
      if (lane == 0) o.hcnt[s] = 0;
      continue;
    }
    if (n > static_cast<uint32_t>(CAP)) {
      if (lane == 0) {
        if (mode == kSegFirst) {
          o.big[atomicAdd(o.nbig, 1u)] = s;
        } else {
          o.ovf[atomicAdd(o.novf, 1u)] = s;
          o.ovf_flag[s] = 1;
        }
      }
      continue;
    }
    const uint32_t p0 = pos_off[s], p1 = pos_off[s + 1];
    const int qLen = qlen[s];
    const int qw = qwin[s];
    const uint32_t hbase = soff[s];
    // generate the strand's seeds into LDS in any order (they are sorted
    // next): entry j of every lane's cell per round, appended by ballot;
    // seeds the corner filter removes are not written.  The cell records of
    // up to 256 positions, and the target sequence of each first entry, are
    // requested before any is used (addresses clamped, so no load sits in a
    // branch).
    uint32_t m = 0;
    for (uint32_t b0 = p0; b0 < p1; b0 += 4 * 64) {
      uint4 cps[4];
      int2 t0s[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t p = b0 + 64u * u + lane;
        cps[u] = cellp[p < p1 ? p : p1 - 1u];  // p1 > p0: the strand has seeds
        if (p >= p1) cps[u].y = 0u;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) t0s[u] = a.tinfo[cps[u].y ? cps[u].z : 0u];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (b0 + 64u * u >= p1) break;
        const uint4 cp = cps[u];
        const int qIdx = static_cast<int>(b0 + 64u * u + lane - p0);
        for (uint32_t j = 0; __any(j < cp.y); ++j) {
          bool ok = false;
          SeedV v{0, 0, 0, qIdx};
          if (j < cp.y) {
            const int2 e = j == 0 ? make_int2(static_cast<int>(cp.z), static_cast<int>(cp.w)) : a.epos[cp.x + j];
            const int2 ti = j == 0 ? t0s[u] : a.tinfo[e.x];
            v.gs = e.x;
            v.tw = ti.y;
            v.d = e.y - qIdx;
            ok = corner_ok(a, v.d, qLen, ti.x);
          }
          const unsigned long long bm = __ballot(ok);
          if (ok) seg_put(L, w, m + static_cast<uint32_t>(__popcll(bm & below)), pk, v);
          m += static_cast<uint32_t>(__popcll(bm));
        }
      }
    }
    if (m == 0) {
      if (lane == 0) o.hcnt[s] = 0;
      continue;
    }
    if constexpr (PACKED) {
      if (m <= 64) {
        // one seed per lane: sort and filter in registers
        wave_lds_sync();
        const uint64_t k0 = lane < static_cast<int>(m) ? L.k[w][lane] : ~0ull;
        const uint64_t key = m <= 8 ? wave_sort<8>(k0) : m <= 16 ? wave_sort<16>(k0)
                             : m <= 32 ? wave_sort<32>(k0) : wave_sort<64>(k0);
        const bool valid = lane < static_cast<int>(m);
        const SeedV v = seg_decode(key, pk);
        const SeedV pv = seg_decode(lane_up1_64(key), pk);
        SeedMarks mk{false, false, false};
        bool kk = false;
        if (valid) {
          mk = seed_marks(a, v, pv, lane > 0);
          bool elem0 = false;
          if (mk.twhead) {
            if (rerun) elem0 = a.unit_first[v.tw * a.nqwin + qw] == s;
            else if (!a.wmask) note_unit_first(a, v.tw, qw, s);
          }
          kk = mk.nr || elem0;
        }
        if (!rerun && a.wmask) {
          const uint64_t wm = wave_or(valid && mk.twhead ? 1ull << (v.tw & 63) : 0ull);
          if (lane == 0 && wm) a.wmask[s] = wm;
        }
        const int kin = wave_max_scan_incl(kk ? lane : -1);
        const int gin = wave_max_scan_incl(valid && mk.ghead ? lane : -1);
        const int prevk = static_cast<int>(lane_up1(0xFFFFFFFFu, static_cast<uint32_t>(kin)));
        const SeedV kv = seg_decode(__shfl(key, prevk < 0 ? 0 : prevk, 64), pk);
        bool dd = false;
        if (valid && kk) {
          dd = !(prevk >= 0 && prevk >= gin && near_pair(v.d + v.q, v.d, kv.d + kv.q, kv.d, a.width, a.gap));
        }
        const unsigned long long hm = __ballot(dd);
        if (dd) {
          const uint32_t at = hbase + static_cast<uint32_t>(__popcll(hm & below));
          o.hs[at] = static_cast<int32_t>(s);
          o.hgs[at] = v.gs;
          o.ht[at] = v.d + v.q;
          o.hq[at] = v.q;
        }
        if (lane == 0) o.hcnt[s] = static_cast<uint32_t>(__popcll(hm));
        wave_lds_sync();
        continue;
      }
    }
    uint32_t np = 1;
    while (np < m) np <<= 1;
    for (uint32_t i = m + lane; i < np; i += 64) seg_none(L, w, i);
    wave_lds_sync();
    // Bitonic sort ascending.  One compare-exchange of stage (kk, j) per
    // lane and step: pair p is elements i = p with a zero inserted at bit j
    // and i | j, ascending when bit kk of i is clear.
    auto lds_stage = [&](uint32_t kk, uint32_t j) {
      for (uint32_t p = lane; p < np / 2; p += 64) {
        const uint32_t i = ((p & ~(j - 1u)) << 1) | (p & (j - 1u)), x = i | j;
        const uint64_t ka = L.k[w][i], kb = L.k[w][x];
        bool gt;
        if constexpr (PACKED) {
          gt = ka > kb;
        } else {
          gt = ka > kb || (ka == kb && L.q[w][i] > L.q[w][x]);
        }
        if (gt == ((i & kk) == 0)) {
          L.k[w][i] = kb;
          L.k[w][x] = ka;
          if constexpr (!PACKED) {
            const int qa = L.q[w][i];
            L.q[w][i] = L.q[w][x];
            L.q[w][x] = qa;
          }
        }
      }
      wave_lds_sync();
    };
    if constexpr (PACKED) {
      // Stages with j < 64 pair lanes of one 64-seed chunk: they run in
      // registers.  Levels up to 64 sort each chunk, ascending for even
      // chunks and descending for odd ones (stored lane-reversed) as bit 64
      // of the index asks; each later level kk runs its stages j >= 64 in
      // LDS, then j = 32 .. 1 per chunk in registers.
      for (uint32_t c = 0; c < np; c += 64) {
        const uint64_t k = wave_sort<64>(L.k[w][c + lane]);
        L.k[w][c + ((c & 64u) ? 63u - lane : lane)] = k;
      }
      wave_lds_sync();
      for (uint32_t kk = 128; kk <= np; kk <<= 1) {
        for (uint32_t j = kk >> 1; j >= 64; j >>= 1) lds_stage(kk, j);
        for (uint32_t c = 0; c < np; c += 64) {
          const uint64_t k = wave_merge<32>(L.k[w][c + lane], lane, (c & kk) == 0);
          L.k[w][c + lane] = k;
        }
        wave_lds_sync();
      }
    } else {
      for (uint32_t kk = 2; kk <= np; kk <<= 1)
        for (uint32_t j = kk >> 1; j > 0; j >>= 1) lds_stage(kk, j);
    }
    // filters over the sorted seeds, 64 at a time
    int carry_k = -1, carry_g = -1;
    uint32_t nh = 0;
    uint64_t wm = 0;  // target windows of the strand's seeds (with SeedArgs::wmask)
    for (uint32_t c0 = 0; c0 < m; c0 += 64) {
      const uint32_t i = c0 + lane;
      const bool valid = i < m;
      SeedMarks mk{false, false, false};
      SeedV v{0, 0, 0, 0};
      bool kk = false;
      if (valid) {
        v = seg_get(L, w, i, pk, a.tinfo);
        const SeedV pv = i > 0 ? seg_get(L, w, i - 1, pk, a.tinfo) : v;
        mk = seed_marks(a, v, pv, i > 0);
        bool elem0 = false;
        if (mk.twhead) {
          if (rerun) elem0 = a.unit_first[v.tw * a.nqwin + qw] == s;
          else if (!a.wmask) note_unit_first(a, v.tw, qw, s);
        }
        kk = mk.nr || elem0;
      }
      if (!rerun && a.wmask) wm |= wave_or(valid && mk.twhead ? 1ull << (v.tw & 63) : 0ull);
      const int kin = wave_max_scan_incl(kk ? static_cast<int>(i) : -1);
      const int gin = wave_max_scan_incl(valid && mk.ghead ? static_cast<int>(i) : -1);
      const int kex = static_cast<int>(lane_up1(0xFFFFFFFFu, static_cast<uint32_t>(kin)));
      const int prevk = kex > carry_k ? kex : carry_k;
      const int gstart = gin > carry_g ? gin : carry_g;
      bool dd = false;
      if (valid && kk) {
        dd = true;
        if (prevk >= 0 && prevk >= gstart) {
          const SeedV kv = seg_get(L, w, static_cast<uint32_t>(prevk), pk, a.tinfo);
          dd = !near_pair(v.d + v.q, v.d, kv.d + kv.q, kv.d, a.width, a.gap);
        }
      }
      const unsigned long long hm = __ballot(dd);
      if (dd) {
        const uint32_t at = hbase + nh + static_cast<uint32_t>(__popcll(hm & below));
        o.hs[at] = static_cast<int32_t>(s);
        o.hgs[at] = v.gs;
        o.ht[at] = v.d + v.q;
        o.hq[at] = v.q;
      }
      nh += static_cast<uint32_t>(__popcll(hm));
      const int lk = __builtin_amdgcn_readlane(kin, 63), lg = __builtin_amdgcn_readlane(gin, 63);
      carry_k = lk > carry_k ? lk : carry_k;
      carry_g = lg > carry_g ? lg : carry_g;
    }
    if (lane == 0) {
      o.hcnt[s] = nh;
      if (wm) a.wmask[s] = wm;
    }
    wave_lds_sync();
  }
}

// Strands to run again once unit_first is final: each unit's first strand,
// unless it took the global path (whose filters read unit_first directly).
// With `gate`, nothing is listed while *gate is nonzero (strands on the
// global path: unit_first is final only after their seeds are written).
static __global__ __launch_bounds__(kBlock) void k_fix_list(const uint32_t* __restrict__ unit_first, uint32_t nunits,
                                                            const uint8_t* __restrict__ ovf_flag,
                                                            uint32_t* __restrict__ mark, uint32_t* __restrict__ list,
                                                            uint32_t* __restrict__ nlist,
                                                            const uint32_t* __restrict__ gate) {
  const uint32_t u = blockIdx.x * kBlock + threadIdx.x;
  if (u >= nunits || (gate && *gate)) return;
  const uint32_t s = unit_first[u];
  if (s == 0xFFFFFFFFu || ovf_flag[s]) return;
  if (atomicExch(&mark[s], 1u) == 0u) list[atomicAdd(nlist, 1u)] = s;
}

// ---------------------------------------------------------------- global path
// (strands with more than kSegBig seeds, and the stage-retaining mode)
//
// Seed key (strand, target sequence, diagonal + d_off); a seed the corner
// filter removes keeps its slot with every bit below the strand set, which
// no seed reaches (d_off leaves the top diagonal code unused), so it sorts
// after its strand's seeds and is skipped by the filters.

struct SeedPack {
  int gs_bits, d_bits;
  int d_off;
};

struct SeedDecode {
  int gs_bits, d_bits, d_off;
  __device__ __forceinline__ uint32_t s(uint64_t k) const { return static_cast<uint32_t>(k >> (gs_bits + d_bits)); }
  __device__ __forceinline__ uint32_t gs(uint64_t k) const {
    return static_cast<uint32_t>((k >> d_bits) & ((1ull << gs_bits) - 1ull));
  }
  __device__ __forceinline__ int d(uint64_t k) const {
    return static_cast<int>(k & ((1ull << d_bits) - 1ull)) - d_off;
  }
  __device__ __forceinline__ bool none(uint64_t k) const {
    const uint64_t low = (1ull << (gs_bits + d_bits)) - 1ull;
    return (k & low) == low;
  }
};

// Raw seed counts of the strands of the global path (ovf_flag), 0 for the
// others: scanned, strand s's seeds start at off[s] in strand order.
static __global__ __launch_bounds__(kBlock) void k_ovf_strand_counts(const uint8_t* __restrict__ ovf_flag,
                                                                     const uint32_t* __restrict__ scnt, uint32_t nseq,
                                                                     uint32_t* __restrict__ cnt) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s < nseq) cnt[s] = ovf_flag[s] ? scnt[s] : 0u;
}

// One workgroup per listed strand (any order) writes its seeds from off[s]
// on, in query-position then cell order (krnlWriteSeedList.cu:3-37 order),
// and records the strand as a unit candidate like k_seed_segments.  The
// strands here hold thousands of positions: kWriteBlock of them per round,
// each position's first slot from a workgroup-wide prefix of the cell
// counts (wave scans, then the wave totals through LDS).  The round's seeds
// are then spread evenly over the threads (seed o of the round to thread o
// mod kWriteBlock, its position found by a binary search of the prefix in
// LDS), so that a repeat k-mer with hundreds of entries does not hold its
// wavefront.  With at most 64 target windows the unit candidates go as one
// window mask per strand (one note per window, not per seed).
constexpr int kWriteBlock = 1024;
static __global__ __launch_bounds__(kWriteBlock) void k_seed_write(SeedArgs a, SeedPack pk,
                                                                   const uint32_t* __restrict__ list, uint32_t nlist,
                                                                   const uint4* __restrict__ cellp,
                                                                   const uint32_t* __restrict__ off, int nwin,
                                                                   uint64_t* __restrict__ skey,
                                                                   uint32_t* __restrict__ sval) {
  constexpr int kW = kWriteBlock / 64;
  __shared__ uint32_t wsum[kW];
  __shared__ uint32_t pend[kWriteBlock];   // inclusive prefix of the round's cell counts
  __shared__ uint32_t plo[kWriteBlock];    // each position's first entry
  __shared__ int2 pfirst[kWriteBlock];     // and that entry's (sequence, tIdx)
  __shared__ unsigned long long wave_wins[kW];
  const int lane = lane_id();
  const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const uint64_t none = (1ull << (pk.gs_bits + pk.d_bits)) - 1ull;
  const bool masks = nwin <= 64;
  // workgroups stride over the listed strands (the grid is capped)
  for (uint32_t k = blockIdx.x; k < nlist; k += gridDim.x) {
    const uint32_t s = list[k];
    const uint32_t p0 = a.pos_off[s], p1 = a.pos_off[s + 1];
    const int qLen = a.qlen[s];
    const int qw = a.qwin[s];
    const uint64_t sk = static_cast<uint64_t>(s) << (pk.gs_bits + pk.d_bits);
    uint32_t base = off[s];
    uint64_t wm = 0;  // target windows of this thread's kept seeds
    for (uint32_t c0 = p0; c0 < p1; c0 += kWriteBlock) {
      const uint32_t p = c0 + threadIdx.x;
      const uint4 cp = p < p1 ? cellp[p] : make_uint4(0, 0, 0, 0);
      const uint32_t incl = wave_sum_incl(cp.y);
      if (lane == 63) wsum[w] = incl;
      __syncthreads();
      uint32_t before = 0, total = 0;
#pragma unroll
      for (int v = 0; v < kW; ++v) {
        const uint32_t t = wsum[v];
        before += v < w ? t : 0u;
        total += t;
      }
      pend[threadIdx.x] = before + incl;
      plo[threadIdx.x] = cp.x;
      pfirst[threadIdx.x] = make_int2(static_cast<int>(cp.z), static_cast<int>(cp.w));
      __syncthreads();
      for (uint32_t o = threadIdx.x; o < total; o += kWriteBlock) {
        // the round's position of seed o: the first whose prefix passes o
        uint32_t lo = 0, hi = kWriteBlock - 1;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (pend[mid] > o) hi = mid;
          else lo = mid + 1;
        }
        const uint32_t j = o - (lo ? pend[lo - 1] : 0u);
        const int qIdx = static_cast<int>(c0 + lo - p0);
        const int2 e = j == 0 ? pfirst[lo] : a.epos[plo[lo] + j];
        const int d = e.y - qIdx;
        const int2 ti = a.tinfo[e.x];
        const uint32_t at = base + o;
        if (corner_ok(a, d, qLen, ti.x)) {
          skey[at] = sk | (static_cast<uint64_t>(e.x) << pk.d_bits) | static_cast<uint64_t>(d + pk.d_off);
          if (masks) wm |= 1ull << ti.y;
          else note_unit_first(a, ti.y, qw, s);
        } else {
          skey[at] = sk | none;
        }
        sval[at] = static_cast<uint32_t>(qIdx);
      }
      base += total;
      __syncthreads();  // the round's LDS arrays are rewritten by the next
    }
    if (masks) {
      const uint64_t m = wave_or(wm);
      if (lane == 0) wave_wins[w] = m;
      __syncthreads();
      if (threadIdx.x < 64) {
        uint64_t all = 0;
#pragma unroll
        for (int v = 0; v < kW; ++v) all |= wave_wins[v];
        if ((all >> threadIdx.x) & 1ull) note_unit_first(a, static_cast<int>(threadIdx.x), qw, s);
      }
      __syncthreads();  // wave_wins is rewritten by the next strand
    }
  }
}

// Per sorted seed: near-predecessor flag, group head index (for the
// group-start max-scan) and the isolate keep flag K(i) = near(i) or i is its
// unit's element 0 (the first seed of the unit's first strand in that target
// window), stored as i+1 for the last-kept max-scan.
static __global__ __launch_bounds__(kBlock) void k_seed_marks(const uint64_t* __restrict__ skey,
                                                              const uint32_t* __restrict__ sval, uint32_t ns,
                                                              SeedDecode dc, int width, int gap,
                                                              const int32_t* __restrict__ seq_win_t,
                                                              const int32_t* __restrict__ seq_win_q, int nqwin,
                                                              const uint32_t* __restrict__ unit_first,
                                                              uint32_t* __restrict__ kidx, uint32_t* __restrict__ gidx) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= ns) return;
  const uint64_t k = skey[i];
  if (dc.none(k)) {
    kidx[i] = 0u;
    gidx[i] = i;
    return;
  }
  const uint64_t grp = k >> dc.d_bits;
  const uint32_t s = dc.s(k), gs = dc.gs(k);
  const int d = dc.d(k), q = static_cast<int>(sval[i]), t = d + q;
  bool nr = false, ghead = true, uhead = true;
  const int tw = seq_win_t[gs];
  if (i > 0) {
    // a removed seed is never the predecessor of one of its strand's seeds
    const uint64_t pk = skey[i - 1];
    if ((pk >> dc.d_bits) == grp) {
      ghead = false;
      const int pd = dc.d(pk), pq = static_cast<int>(sval[i - 1]);
      nr = near_pair(t, d, pd + pq, pd, width, gap);
    }
    uhead = dc.s(pk) != s || seq_win_t[dc.gs(pk)] != tw;
  }
  const bool elem0 = uhead && unit_first[tw * nqwin + seq_win_q[s]] == s;
  kidx[i] = (nr || elem0) ? i + 1 : 0u;
  gidx[i] = ghead ? i : 0u;
}

// Duplicate filter (…deleteDuplicateSeeds.cu:15-71) on the isolate output:
// a kept seed survives unless its previous kept seed in the same group is
// near.  lastk = inclusive max-scan of kidx, gstart = of gidx.
static __global__ __launch_bounds__(kBlock) void k_seed_dup(const uint64_t* __restrict__ skey,
                                                            const uint32_t* __restrict__ sval, uint32_t ns,
                                                            SeedDecode dc, int width, int gap,
                                                            const uint32_t* __restrict__ kidx,
                                                            const uint32_t* __restrict__ lastk,
                                                            const uint32_t* __restrict__ gstart,
                                                            uint32_t* __restrict__ dflag) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= ns) return;
  uint32_t f = 0;
  if (kidx[i]) {
    f = 1;
    const uint32_t pk = i > 0 ? lastk[i - 1] : 0u;
    if (pk >= 1 && pk - 1 >= gstart[i]) {
      const uint32_t j = pk - 1;
      const uint64_t k = skey[i], kj = skey[j];
      const int d = dc.d(k), t = d + static_cast<int>(sval[i]);
      const int dj = dc.d(kj), tj = dj + static_cast<int>(sval[j]);
      f = near_pair(t, d, tj, dj, width, gap) ? 0u : 1u;
    }
  }
  dflag[i] = f;
}

// Surviving seeds of the global path go to their strand's own seed range
// (like k_seed_segments): slot = soff[s] + rank among the strand's
// surviving seeds.  goff: the global path's offsets per position; hoff:
// exclusive scan of dflag.
static __global__ __launch_bounds__(kBlock) void k_hits_scatter(const uint64_t* __restrict__ skey,
                                                                const uint32_t* __restrict__ sval, uint32_t ns,
                                                                SeedDecode dc, const uint32_t* __restrict__ dflag,
                                                                const uint32_t* __restrict__ hoff,
                                                                const uint32_t* __restrict__ soff,
                                                                const uint32_t* __restrict__ goff,
                                                                int32_t* __restrict__ hs, int32_t* __restrict__ hgs,
                                                                int32_t* __restrict__ ht, int32_t* __restrict__ hq) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= ns || !dflag[i]) return;
  const uint64_t k = skey[i];
  const uint32_t s = dc.s(k);
  // goff[s]: strand s's first seed, in strand order before and after the sort
  const uint32_t o = soff[s] + (hoff[i] - hoff[goff[s]]);
  const int q = static_cast<int>(sval[i]);
  hs[o] = static_cast<int32_t>(s);
  hgs[o] = static_cast<int32_t>(dc.gs(k));
  ht[o] = dc.d(k) + q;
  hq[o] = q;
}

// Hit counts of the global-path strands: surviving seeds between the
// strand's first global-path seed (goff[s], strand order) and the next
// strand's.
static __global__ __launch_bounds__(kBlock) void k_ovf_hcnt(const uint32_t* __restrict__ ovf, uint32_t novf,
                                                            uint32_t nseq, const uint32_t* __restrict__ goff,
                                                            uint32_t ng, const uint32_t* __restrict__ hoff,
                                                            const uint32_t* __restrict__ nh_dev,
                                                            uint32_t* __restrict__ hcnt) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= novf) return;
  const uint32_t nh = *nh_dev;
  const uint32_t s = ovf[k];
  const uint32_t g0 = goff[s];
  const uint32_t g1 = s + 1u < nseq ? goff[s + 1u] : ng;
  const uint32_t h0 = g0 < ng ? hoff[g0] : nh;
  const uint32_t h1 = g1 < ng ? hoff[g1] : nh;
  hcnt[s] = h1 - h0;
}

// Dense hits in strand order: strand s's hits move from the start of its
// seed range (soff[s]) to hbase[s] (exclusive scan of the per-strand hit
// counts).  One thread per dense hit; its strand is the last one whose base
// is at most the hit's index (strands without hits share their
// successor's base, so the last such strand is the one that has the hit).
static __global__ __launch_bounds__(kBlock) void k_hits_compact(uint32_t nhit, uint32_t nseq,
                                                                const uint32_t* __restrict__ soff,
                                                                const uint32_t* __restrict__ hbase,
                                                                const int32_t* __restrict__ hs,
                                                                const int32_t* __restrict__ hgs,
                                                                const int32_t* __restrict__ ht,
                                                                const int32_t* __restrict__ hq,
                                                                int32_t* __restrict__ ds, int32_t* __restrict__ dgs,
                                                                int32_t* __restrict__ dt, int32_t* __restrict__ dq) {
  const uint32_t d = blockIdx.x * kBlock + threadIdx.x;
  if (d >= nhit) return;
  const uint32_t s = upper_bound_u(hbase, 0u, nseq, d) - 1u;
  const uint32_t src = soff[s] + (d - hbase[s]);
  ds[d] = hs[src];
  dgs[d] = hgs[src];
  dt[d] = ht[src];
  dq[d] = hq[src];
}

// ---------------------------------------------------------------- banded DP

struct DPArgs {
  uint32_t n;
  const int32_t* hs;
  const int32_t* hgs;
  const int32_t* ht;
  const int32_t* hq;
  const uint8_t* qbase;
  const uint32_t* qslot;
  const int32_t* qlen;
  const uint8_t* tbase;
  const uint64_t* tslot;   // slot of each target sequence's first base
  const int32_t* tlen;
  const int32_t* twin;     // window of each target sequence
  const uint64_t* wslot;   // window slot bounds (nwin + 1)
  const uint32_t* order;   // task of thread i: (hit << 1) | direction, longest first
  // order[0, split[0]) go to k_dp_band, [split[0], split[1]) to k_dp16,
  // [split[1], split[2]) to k_dp16 with 'X' past the sequence end, the rest
  // to k_dp
  const uint32_t* split;
  const uint32_t* qnib;    // 4-bit copies of qbase / tbase with mirrors (k_pack_nibbles), read by k_dp16
  const uint32_t* tnib;
  int4* dres;              // per task 2 * hit + direction: (move + end, move, match, score)
  int32_t q_bmax, t_bmax;  // last 16-byte block of qbase / tbase (both padded to whole blocks)
  int lmer;
  uint32_t round;          // wavefronts per dispatch round (one per SIMD): k_dp / k_dp16 snake order
  uint32_t* chunk_next;    // persistent k_dp16: next chunk to take (zero before the launch)
  int32_t* o_tidx;
  int32_t* o_qidx;
  int32_t* o_thl;
  int32_t* o_qhl;
  int32_t* o_match;
  int32_t* o_score;
  const uint8_t* qclean;   // per query strand / target sequence: no star or X class
  const uint8_t* tclean;   // (k_dp_band's substitution shortcut; bit 1: no class 4, k_dp16's query profiles)
};

constexpr int kGapOpen = -5, kGapExt = -2, kMatch = 1, kMismatch = -3, kBad = -1000;

// Sequential reader of one DP lane's bytes — indexes g, g + step, g + 2 step,
// ... of a device byte array padded to whole 16-byte blocks — in groups of
// 16 on a schedule shared by the whole wavefront (every lane is at the same
// stream position j, the row counter): entering a group builds its 16 bytes
// from two aligned blocks loaded one group earlier and requests the next
// group's blocks, so loads land directly in their registers and are not
// waited for until 16 rows later.  One pair of loads serves 16 rows: the
// scattered target and query lines of a wavefront are fetched once each.
// Blocks outside the array are clamped; their bytes are never used.
struct ByteWin {
  const uint4* blk;
  int32_t bi, bmax, step;  // next group's first block index, last block, +1 / -1
  uint32_t d, bs;          // the window's dword / byte offset inside its blocks (fixed per lane)
  uint4 pa, pb;            // blocks of the next group
  uint32_t w0, w1, w2, w3; // current group, stream order, low byte first

  __device__ __forceinline__ void request() {
    pa = blk[min(max(bi, 0), bmax)];
    pb = blk[min(max(bi + 1, 0), bmax)];
    bi += step;
  }
  __device__ __forceinline__ void init(const uint8_t* base, int32_t last_block, int64_t g0, int st) {
    blk = reinterpret_cast<const uint4*>(base);
    bmax = last_block;
    step = st;
    // first group's memory range: [g0, g0 + 15] forward, [g0 - 15, g0] backward
    const int64_t lo = st > 0 ? g0 : g0 - 15;
    bi = static_cast<int32_t>(lo >> 4);
    const uint32_t off = static_cast<uint32_t>(lo & 15);
    d = off >> 2;
    bs = off & 3u;
    request();
  }
  __device__ __forceinline__ void refill() {
    const uint32_t x0 = pa.x, x1 = pa.y, x2 = pa.z, x3 = pa.w, x4 = pb.x, x5 = pb.y, x6 = pb.z, x7 = pb.w;
    // y[j] = x[j + d]: shift by one dword on bit 0 of d, then by two on bit 1
    const bool s1 = (d & 1u) != 0u, s2 = (d & 2u) != 0u;
    const uint32_t a0 = s1 ? x1 : x0, a1 = s1 ? x2 : x1, a2 = s1 ? x3 : x2, a3 = s1 ? x4 : x3;
    const uint32_t a4 = s1 ? x5 : x4, a5 = s1 ? x6 : x5, a6 = s1 ? x7 : x6;
    const uint32_t y0 = s2 ? a2 : a0, y1 = s2 ? a3 : a1, y2 = s2 ? a4 : a2, y3 = s2 ? a5 : a3, y4 = s2 ? a6 : a4;
    // bytes off .. off+15 of the two blocks
    const uint32_t m0 = __builtin_amdgcn_alignbyte(y1, y0, bs);
    const uint32_t m1 = __builtin_amdgcn_alignbyte(y2, y1, bs);
    const uint32_t m2 = __builtin_amdgcn_alignbyte(y3, y2, bs);
    const uint32_t m3 = __builtin_amdgcn_alignbyte(y4, y3, bs);
    if (step > 0) {
      w0 = m0; w1 = m1; w2 = m2; w3 = m3;
    } else {  // backward: stream order is memory order reversed
      w0 = __builtin_amdgcn_perm(0u, m3, 0x00010203u);
      w1 = __builtin_amdgcn_perm(0u, m2, 0x00010203u);
      w2 = __builtin_amdgcn_perm(0u, m1, 0x00010203u);
      w3 = __builtin_amdgcn_perm(0u, m0, 0x00010203u);
    }
    request();
  }
  // Byte at stream position j; j must be the same in every lane and advance
  // by one per call.
  __device__ __forceinline__ uint32_t next(int j) {
    if ((j & 15) == 0) refill();
    const uint32_t v = w0 & 0xFFu;
    w0 >>= 8;
    if ((j & 3) == 3) {
      w0 = w1;
      w1 = w2;
      w2 = w3;
    }
    return v;
  }
};

// Target base class at intra-sequence index ti: [lo, hi) is the valid range,
// the window's slots relative to the sequence cut by the direction's
// one-sided guard — backward ti < tLen, forward ti >= 0 (krnlAlignment.cu:
// 121-127,160-165 / 234-240,271-276).  Anything outside is 'X'.
__device__ __forceinline__ uint32_t t_mask(uint32_t v, int ti, int lo, int hi) {
  return 4u * (ti >= lo && ti < hi ? v : kClsX);
}

// Substitution score (POINT_MATRIX) and match-count increment
// (MATCH_COUNTER) of query class qc against each target class, one signed
// 4-bit field per target class (field 4·tc); -8 stands for the X value
// -1000.  Fields for target A C G T other: -3 / 0, the match +1 / +1;
// E F I J L O P Q Z: -1 / -1; X: -1000 / -1000.
__device__ __forceinline__ void row_words(uint32_t qc, uint32_t& wpm, uint32_t& wmc) {
  constexpr uint32_t kPm = 0x8FDDDDDu, kMc = 0x8F00000u;  // query class with no match in the table
  constexpr uint32_t kStar = 0x8FFFFFFu, kAllX = 0x8888888u;
  const uint32_t sh = 4u * (qc & 3u);
  wpm = qc < 4u ? kPm ^ (0xCu << sh) : qc == kClsOther ? kPm : qc == kClsStar ? kStar : kAllX;
  wmc = qc < 4u ? kMc | (0x1u << sh) : qc == kClsOther ? kMc : qc == kClsStar ? kStar : kAllX;
}

// One query row of the band (calcutlateCell, krnlAlignment.cu:42-77, for
// bands 1..W in order; bands 0 and W+1 are the constant margins).  The
// horizontal value of band b-1 is only read by band b of the same row, so it
// is carried in a register (starting from band 0's constant) instead of an
// array.
template <int W, bool LOCAL>
__device__ __forceinline__ void dp_row(int (&S)[W + 2], int (&V)[W + 2], int (&M)[W + 2], const uint32_t (&T)[W],
                                       uint32_t wpm, uint32_t wmc, int r, int& bestS, int& bestM, int& bestR,
                                       int& bestB) {
  int hp = -30000;  // band 0's horizontal value (initTempNodeArray margin)
#pragma unroll
  for (int b = 1; b <= W; ++b) {
    const int v = max(S[b + 1] + (kGapOpen + kGapExt), V[b + 1] + kGapExt);
    const int h = max(S[b - 1] + (kGapOpen + kGapExt), hp + kGapExt);
    // T holds 4 x the target class: the field of this target in the row words
    const int pr = __builtin_amdgcn_sbfe(static_cast<int>(wpm), T[b - 1], 4);
    const int mr = __builtin_amdgcn_sbfe(static_cast<int>(wmc), T[b - 1], 4);
    const bool bad = pr == -8;
    const int pm = bad ? kBad : pr;
    const int mc = bad ? kBad : mr;
    const int pt = S[b] + pm;
    // the diagonal wins only if strictly above both gaps, then v over h
    // unless h is at least as high; the score is the maximum either way.
    // The match count is selected with masks so that no lane branches.
    const int vh = max(v, h);
    const int nm_gap = v > h ? M[b + 1] : M[b - 1];
    const int nm_diag = M[b] + mc;
    const int dmask = -static_cast<int>(pt > vh);
    const int nm = nm_gap ^ ((nm_gap ^ nm_diag) & dmask);
    const int ns = max(pt, vh);
    S[b] = ns;
    V[b] = v;
    M[b] = nm;
    hp = h;
    if (LOCAL) {
      const bool better = bestS < ns;
      bestS = better ? ns : bestS;
      bestM = better ? nm : bestM;
      bestR = better ? r : bestR;
      bestB = better ? b - 1 : bestB;
    }
  }
}

template <int G>
__device__ __forceinline__ void dp_init(int (&S)[2 * G + 3], int (&V)[2 * G + 3], int (&M)[2 * G + 3]) {
  constexpr int half = G + 1;
#pragma unroll
  for (int b = 0; b < 2 * G + 3; ++b) {
    if (b < half) {
      S[b] = V[b] = M[b] = -30000;
    } else if (b == half) {
      S[b] = 0; V[b] = kGapOpen; M[b] = 0;
    } else {
      const int sc = (b - half) * kGapExt + kGapOpen;
      S[b] = sc; V[b] = sc + kGapOpen; M[b] = 0;
    }
  }
}

// One thread per (hit, direction), the band held in VGPRs: the right
// extension ({global,local}AlignBackward, direction 0) or the left one
// ({global,local}AlignForward, direction 1).  The two are independent (the
// forward pass reads the seed's tIdx/qIdx, which the backward pass does not
// change), so they run as separate tasks, longest first: a long read's two
// extensions proceed side by side and every wavefront's lanes run near-equal
// row counts.  Query and target bytes arrive through 16-byte windows
// (ByteWin); k_dp_combine adds the two directions up.
// Wavefront w's chunk of a longest-first task order: wavefronts are placed
// one per SIMD per round, so within every other round the order runs
// backwards (a snake) and each SIMD gets long and short chunks alike.
__device__ __forceinline__ uint32_t snake_chunk(uint32_t w, uint32_t nchunk, uint32_t round) {
  const uint32_t r = w / round, pos = w - r * round;
  const uint32_t cnt = min(round, nchunk - r * round);
  return r * round + ((r & 1u) ? cnt - 1u - pos : pos);
}

template <int G, bool LOCAL>
static __global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_dp(DPArgs a) {
  constexpr int W = 2 * G + 1;
  const uint32_t first = a.split[2];
  const uint32_t n = a.n - first;
  const uint32_t wv = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const uint32_t nchunk = (n + 63) / 64;
  if (wv >= nchunk) return;
  const uint32_t t = snake_chunk(wv, nchunk, a.round) * 64 + (threadIdx.x & 63u);
  if (t >= n) return;
  const uint32_t task = a.order[first + t];
  const uint32_t i = task >> 1;
  const bool bwd = (task & 1u) == 0u;
  const int s = a.hs[i], gs = a.hgs[i], tIdx = a.ht[i], qIdx = a.hq[i];
  const int L = a.lmer;
  const int qLen = a.qlen[s];
  const int rows = bwd ? qLen - (qIdx + L) : qIdx;
  if (rows == 0) return;
  const int tLen = a.tlen[gs];
  const uint64_t tslot = a.tslot[gs];
  const int w = a.twin[gs];
  const int wlo = static_cast<int>(static_cast<int64_t>(a.wslot[w]) - static_cast<int64_t>(tslot));
  const int whi = static_cast<int>(static_cast<int64_t>(a.wslot[w + 1]) - static_cast<int64_t>(tslot));
  int S[W + 2], V[W + 2], M[W + 2];
  uint32_t T[W];
  const int64_t qg = static_cast<int64_t>(a.qslot[s]);
  dp_init<G>(S, V, M);
  // target index tb + ts*j feeds ring slot j (j < W) and then row j - W
  const int ts = bwd ? 1 : -1;
  const int tb = bwd ? tIdx + L - G : tIdx + G - 1;
  const int lo = bwd ? wlo : max(wlo, 0);
  const int hi = bwd ? min(whi, tLen) : whi;
  ByteWin tst, qst;
  tst.init(a.tbase, a.t_bmax, static_cast<int64_t>(tslot) + tb, ts);
  qst.init(a.qbase, a.q_bmax, qg + (bwd ? qIdx + L : qIdx - 1), ts);
#pragma unroll
  for (int k = 0; k < W; ++k) T[k] = t_mask(tst.next(k), tb + ts * k, lo, hi);
  int bestS = 0, bestM = 0, bestR = -1, bestB = G;
  int tn_i = tb + ts * W;
#pragma unroll 1
  for (int r = 0; r < rows; ++r) {
    const int ru = __builtin_amdgcn_readfirstlane(r);  // every active lane is at row r
    uint32_t wpm, wmc;
    row_words(qst.next(ru), wpm, wmc);
    const uint32_t tn = t_mask(tst.next(ru + W), tn_i, lo, hi);
    tn_i += ts;
    dp_row<W, LOCAL>(S, V, M, T, wpm, wmc, r, bestS, bestM, bestR, bestB);
#pragma unroll
    for (int k = 0; k < W - 1; ++k) T[k] = T[k + 1];
    T[W - 1] = tn;
  }
  int move, endp;
  if (LOCAL) {
    move = bestR + 1;
    endp = bestB - G;
  } else {
    // select end node: strict '<' from the constant band 0 (:170-186)
    bestS = S[0];
    bestM = M[0];
    int e = -1;
#pragma unroll
    for (int b = 1; b <= W; ++b)
      if (bestS < S[b]) {
        bestS = S[b];
        bestM = M[b];
        e = b - 1;
      }
    move = rows;
    endp = e - G;
  }
  a.dres[task] = make_int4(move + endp, move, bestM, bestS);
}

// Both directions of each hit (…alignmentHits.cu:122-348): the backward
// extension lengthens tHL/qHL; the forward one also moves tIdx/qIdx back.
static __global__ __launch_bounds__(kBlock) void k_dp_combine(DPArgs a, uint32_t nhit) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nhit) return;
  const int s = a.hs[i], tIdx = a.ht[i], qIdx = a.hq[i];
  const int L = a.lmer;
  const int qLen = a.qlen[s];
  const int4 z = make_int4(0, 0, 0, 0);
  const int4 b = qLen - (qIdx + L) > 0 ? a.dres[2 * i] : z;
  const int4 f = qIdx > 0 ? a.dres[2 * i + 1] : z;
  a.o_tidx[i] = tIdx - f.x;
  a.o_qidx[i] = qIdx - f.y;
  a.o_thl[i] = L + b.x + f.x;
  a.o_qhl[i] = L + b.y + f.y;
  a.o_match[i] = L + b.z + f.z;
  a.o_score[i] = L * kMatch + b.w + f.w;
}

// ------------------------------------------- banded DP in packed 16-bit lanes
//
// Global mode, two tasks per thread, one in each 16-bit half of every band
// register.  A task qualifies (k_dp_task_keys) when its query is at most
// kPacked16MaxLen bases, neither its query strand nor its target sequence
// holds a star or X class, and every target index its band touches lies in
// the target sequence and in the window.  Then every substitution is +1 / -3
// with match increment 1 / 0, and the int32 recurrence of k_dp stays inside
// 16 bits: a score is at most the row count and at least band 1's
// horizontal margin -30002 less 2 per band (band b's horizontal term is
// >= -30002 - 2 (b - 1)), so scores lie in [-30075, 30001]; a match count
// starts at 0 or the margin -30000 and grows by at most one per row.  The
// results equal k_dp's bit for bit.
//
// With EDGE, the tasks whose band crosses the end of the target sequence
// that k_dp guards (backward: index >= tLen, forward: index < 0) and that
// stay inside the sequence and the window on the other side.  k_dp scores
// such an 'X' cell's diagonal -1000, and that diagonal never wins: in any
// cell the vertical or horizontal term is at least the diagonal
// predecessor's score less 14 (band b's horizontal term reads band b-1 of
// the same row, whose vertical term reads band b of the row before; band 1's
// vertical term reads band 2, whose horizontal term reads band 1), so the
// cell takes a gap term and keeps that term's match count.  Scoring the
// diagonal S - 15 instead leads to the same choices and values and keeps
// 16 bits: the ring marks such positions with bit 3 of their class code
// (no query code has it), and the diagonal subtracts 12 more there.

constexpr int kPacked16MaxLen = 30000;

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pk_bits(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ s16x2 pk_val(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ s16x2 pk_splat(int v) { return s16x2{static_cast<short>(v), static_cast<short>(v)}; }
// all-ones in the halves where a < b (saturating difference, so no wrap)
__device__ __forceinline__ uint32_t pk_less(s16x2 a, s16x2 b) {
  return pk_bits(__builtin_elementwise_sub_sat(a, b) >> 15);
}
// a in the halves selected by mask m, b elsewhere
__device__ __forceinline__ uint32_t pk_pick(uint32_t m, uint32_t a, uint32_t b) { return b ^ ((a ^ b) & m); }
// 4 m + z per half in one instruction (the compiler splits a multiply by 4
// into a shift and an add); halves wrap independently
__device__ __forceinline__ s16x2 pk_mad4(s16x2 m, s16x2 z) {
  uint32_t r;
  asm("v_pk_mad_u16 %0, %1, 4, %2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(pk_bits(m)), "v"(pk_bits(z)));
  return pk_val(r);
}

// 4-bit class arrays read by k_dp16: nibble k of dword d is slot 8 d + k
// of the byte class array (padded to whole 16-byte blocks, so whole 8-byte
// units here).  The query copy maps class 4 (other) to 7, which matches no
// target class (classes 0..4 in the bands k_dp16 runs).
// Nibbles of a dword in reverse order.
__device__ __forceinline__ uint32_t nibble_reverse(uint32_t x) {
  x = __builtin_amdgcn_perm(0u, x, 0x00010203u);
  return ((x & 0x0F0F0F0Fu) << 4) | ((x >> 4) & 0x0F0F0F0Fu);
}

// nib holds 2 ndw dwords: the classes in slot order (dword d: slots 8 d ..
// 8 d + 7, slot 8 d + k in bits 4 k), then the same classes mirrored (slot
// 8 ndw + k holds slot 8 ndw - 1 - k), so a reader going down the slots
// reads the mirror upwards (NibWin).
static __global__ __launch_bounds__(kBlock) void k_pack_nibbles(const uint8_t* __restrict__ bytes, uint64_t ndw,
                                                                int query, uint32_t* __restrict__ nib) {
  const uint64_t d = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (d >= ndw) return;
  const uint2 v = reinterpret_cast<const uint2*>(bytes)[d];
  uint32_t lo = v.x, hi = v.y;
  if (query) {  // byte table: 0 1 2 3 4 5 6 7 -> 0 1 2 3 7 5 6 7
    lo = __builtin_amdgcn_perm(0x07060507u, 0x03020100u, lo);
    hi = __builtin_amdgcn_perm(0x07060507u, 0x03020100u, hi);
  }
  lo &= 0x0F0F0F0Fu;
  hi &= 0x0F0F0F0Fu;
  // byte k's nibble to bits 4k: pairs of bytes first, then the pairs
  lo |= lo >> 4;
  hi |= hi >> 4;
  const uint32_t w = __builtin_amdgcn_perm(hi, lo, 0x06040200u);
  nib[d] = w;
  nib[2 * ndw - 1 - d] = nibble_reverse(w);
}

// Sequential reader of one half's slots g, g + step, g + 2 step, ... of a
// 4-bit class array with its mirror (k_pack_nibbles; umax: its last 16-slot
// unit, mirror included) in groups of 16, on a schedule shared by the
// wavefront (every lane is at the same group).  A reader going down the
// slots (step -1) goes up the mirror, so every reader moves up: a group's
// window spans two aligned 16-slot units held in r0 / r1; after group k the
// one of them that no later group reads (r[k & 1]) is reloaded with the unit
// after the other, so each load is requested 16 rows before its use and
// lands in its own register.  Units outside [0, umax] are clamped; their
// slots are never used (nor are the other half's slots a reader may reach
// past its own end).
struct NibWin {
  int32_t u;       // unit requested next
  uint32_t s4;     // bit offset of the window inside its first unit
  uint2 r0, r1;

  __device__ __forceinline__ static uint2 load(const uint2* base, int32_t i, int32_t umax) {
    return base[min(max(i, 0), umax)];
  }
  __device__ __forceinline__ void init(const uint2* base, int32_t umax, int64_t g0, int step) {
    const int64_t nslot = 8 * (static_cast<int64_t>(umax) + 1);  // slots of each half
    // the first group's first slot (slot g's mirror: 2 nslot - 1 - g)
    const int64_t lo = step > 0 ? g0 : 2 * nslot - 1 - g0;
    const int32_t u0 = static_cast<int32_t>(lo >> 4);
    s4 = static_cast<uint32_t>(lo & 15) * 4u;
    r0 = load(base, u0, umax);
    r1 = load(base, u0 + 1, umax);
    u = u0 + 2;
  }
  // Group k, k & 1 == ODD: its 16 classes in stream order, rows 0-7 in w0
  // and 8-15 in w1 (row j in bits 4 (j & 7)).
  template <int ODD>
  __device__ __forceinline__ void group(const uint2* base, int32_t umax, uint32_t& w0, uint32_t& w1) {
    const uint2 lo = ODD ? r1 : r0;  // reloaded below
    const uint2 hi = ODD ? r0 : r1;
    // bits [s4, s4 + 64) of hi:lo
    const bool big = s4 >= 32u;
    const uint32_t x0 = big ? lo.y : lo.x, x1 = big ? hi.x : lo.y, x2 = big ? hi.y : hi.x;
    const uint32_t sh = s4 & 31u;
    w0 = __builtin_amdgcn_alignbit(x1, x0, sh);
    w1 = __builtin_amdgcn_alignbit(x2, x1, sh);
    const uint2 v = load(base, u, umax);
    u += 1;
    if (ODD) r1 = v;
    else r0 = v;
  }
};

// Row codes of the two halves for 16 rows: word k holds rows 4k .. 4k + 3,
// the low half's classes in bits 0-15 and the high half's in bits 16-31
// (row 4k + i in bits 4 i of each half).
struct PairRows {
  uint32_t p0, p1, p2, p3;
  __device__ __forceinline__ void build(uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1) {
    p0 = __builtin_amdgcn_perm(b0, a0, 0x05040100u);
    p1 = __builtin_amdgcn_perm(b0, a0, 0x07060302u);
    p2 = __builtin_amdgcn_perm(b1, a1, 0x05040100u);
    p3 = __builtin_amdgcn_perm(b1, a1, 0x07060302u);
  }
  // codes of the next row (both halves); j is the row inside the group,
  // the same in every lane
  __device__ __forceinline__ uint32_t next(int j) {
    const uint32_t v = p0 & 0x000F000Fu;
    p0 >>= 4;
    if ((j & 3) == 3) {
      p0 = p1;
      p1 = p2;
      p2 = p3;
    }
    return v;
  }
  // past the four rows of p0 without reading them
  __device__ __forceinline__ void skip4() {
    p0 = p1;
    p1 = p2;
    p2 = p3;
  }
};

// One query row (k_dp's dp_row, global mode) for two tasks.  Z holds S - 7:
// both gap terms then read it unchanged, and the diagonal term is
// S + (4 mc - 3) = Z + 4 + 4 mc.  T[b-1] and Q hold the target and query
// codes of both halves; mc = 1 - min(t ^ q, 1).  SCORE: scores only (the
// match counts M are left as they are; they do not feed the scores).
template <int W, int O, bool EDGE, bool SCORE>
__device__ __forceinline__ void dp_row16(s16x2 (&Z)[W + 2], s16x2 (&V)[W + 2], s16x2 (&M)[W + 2],
                                         const uint32_t (&T)[W + 4], uint32_t Q) {
  s16x2 hp = pk_splat(-30000);
#pragma unroll
  for (int b = 1; b <= W; ++b) {
    const s16x2 v = __builtin_elementwise_max(Z[b + 1], V[b + 1] + pk_splat(kGapExt));
    const s16x2 h = __builtin_elementwise_max(Z[b - 1], hp + pk_splat(kGapExt));
    const u16x2 x = __builtin_bit_cast(u16x2, T[O + b - 1] ^ Q);
    const s16x2 mc = __builtin_bit_cast(s16x2, __builtin_elementwise_sub_sat(u16x2{1, 1}, x));
    s16x2 pt = pk_mad4(mc, Z[b]) + pk_splat(kMatch - kMismatch);
    if constexpr (EDGE) {
      // an 'X' target position (bit 3 of x): the diagonal S - 15
      const s16x2 xb = __builtin_bit_cast(s16x2, x >> u16x2{3, 3});
      pt = pt - xb * pk_splat(12);
    }
    const s16x2 vh = __builtin_elementwise_max(v, h);
    if constexpr (!SCORE) {
      const uint32_t ngap = pk_pick(pk_less(h, v), pk_bits(M[b + 1]), pk_bits(M[b - 1]));
      const uint32_t nm = pk_pick(pk_less(vh, pt), pk_bits(M[b] + mc), ngap);
      M[b] = pk_val(nm);
    }
    Z[b] = __builtin_elementwise_max(pt, vh) + pk_splat(kGapOpen + kGapExt);
    V[b] = v;
    hp = h;
  }
}

// Rows per drift period of the score pass (dp_row16_drift).
constexpr int kDriftRows = 256;

// One query row of the score pass (SCORE, not EDGE) in drifted units: every
// value of band b at row r is held as value + d(r, b), d = 2 b + 4 (r mod
// kDriftRows)
// (Z as Z + 2 + d, i.e. S - 5 + d), which the gap terms' -2 steps cancel:
// v(r, b) reads band b + 1 of row r - 1 (d two lower, Z two higher) and
// h(r, b) band b - 1 of row r (the same), so
//   v~ = max(Z~[b+1], V~[b+1]),  h~ = max(Z~[b-1], h~[b-1]),
//   pt~ = Z~[b] + 6 + 4 mc  (the diagonal: d four lower, Z two higher),
//   Z~ = max(pt~, v~, h~) - 5,
// eight instructions per cell pair.  6 + 4 mc comes from one byte permute of
// the class codes' xor (Qs: the query codes with 0x0C in bytes 1 and 3, which
// select 0 for the high bytes).  Band 1's horizontal term and band W's
// vertical term are constants of k_dp's edges (-30002 and the band W + 1
// constants): h1 and vw, drifted by the caller.  The values change only by
// d, and 0 <= d <= 2 W + 4 (kDriftRows - 1) <= 1086, inside the 16-bit
// margin above k_dp16's range proof's scores (at most 30001, and the
// diagonal adds at most 12).
template <int W, int O>
__device__ __forceinline__ void dp_row16_drift(s16x2 (&Z)[W + 2], s16x2 (&V)[W + 2], const uint32_t (&T)[W + 4],
                                               uint32_t Qs, s16x2 h1, s16x2 vw) {
  s16x2 hp = h1;
#pragma unroll
  for (int b = 1; b <= W; ++b) {
    const s16x2 v = b == W ? vw : __builtin_elementwise_max(Z[b + 1], V[b + 1]);
    const s16x2 h = b == 1 ? h1 : __builtin_elementwise_max(Z[b - 1], hp);
    const s16x2 w = pk_val(__builtin_amdgcn_perm(0x06060606u, 0x0606060Au, T[O + b - 1] ^ Qs));
    const s16x2 pt = Z[b] + w;
    Z[b] = __builtin_elementwise_max(pt, __builtin_elementwise_max(v, h)) + pk_splat(-5);
    V[b] = v;
    hp = h;
  }
}

// dp_row16_drift with the diagonal term from one byte permute per cell pair:
// the pool is the row's query profiles (px / py: byte c = 10 when the half's
// query class is c, else 6), the selector the ring's target word (half 0's
// class in byte 0, half 1's as 4 + class in byte 2, 12 in bytes 1 and 3 for
// zero high bytes).  Target classes 0..3 only (class 4 would select py's
// byte 0): the task's target sequence holds no class 4.
// FIRST: the row's vertical terms go first, before anything that reads the
// profiles (they may still be in flight from ds_bpermute).
template <int W, int O, bool FIRST = false>
__device__ __forceinline__ void dp_row16_prof(s16x2 (&Z)[W + 2], s16x2 (&V)[W + 2], const uint32_t (&T)[W + 4],
                                              uint32_t px, uint32_t py, s16x2 h1, s16x2 vw) {
  if constexpr (FIRST) {
    // in place, rising b: V[b + 1] and Z[b + 1] are still the last row's
#pragma unroll
    for (int b = 1; b < W; ++b) V[b] = __builtin_elementwise_max(Z[b + 1], V[b + 1]);
    V[W] = vw;
    __builtin_amdgcn_sched_barrier(0);
  }
  s16x2 hp = h1;
#pragma unroll
  for (int b = 1; b <= W; ++b) {
    const s16x2 v = FIRST ? V[b] : b == W ? vw : __builtin_elementwise_max(Z[b + 1], V[b + 1]);
    const s16x2 h = b == 1 ? h1 : __builtin_elementwise_max(Z[b - 1], hp);
    const s16x2 pt = Z[b] + pk_val(__builtin_amdgcn_perm(py, px, T[O + b - 1]));
    Z[b] = __builtin_elementwise_max(pt, __builtin_elementwise_max(v, h)) + pk_splat(-5);
    V[b] = v;
    hp = h;
  }
}

// The query profile of dp_row16_prof for the class in byte K of Q (0 or 2):
// byte c of the result is 10 when c equals the class, else 6 (class 7, the
// query's other letters, equals none).
template <int K>
__device__ __forceinline__ uint32_t query_profile(uint32_t Q) {
  const uint32_t sel = __builtin_amdgcn_perm(Q, Q, K * 0x01010101u) ^ 0x03020100u;
  return __builtin_amdgcn_perm(0x06060606u, 0x0606060Au, sel);
}

// The same profiles as a table across the lanes: lane l holds the profile of
// class l mod 16 (classes 4 and up: no byte at 10, as class 7 above), so
// ds_bpermute reads the profile of the class in bits 2-5 of its byte address
// (it takes the source lane from address bits 2-7, the wavefront's 64
// lanes; bits 6-7 land on a copy of the table).
__device__ __forceinline__ uint32_t query_profile_table() {
  const uint32_t q = threadIdx.x & 15u;
  return 0x06060606u + (q < 4u ? 4u << (8u * q) : 0u);
}
__device__ __forceinline__ uint32_t query_profile_lookup(uint32_t addr, uint32_t table) {
  return static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(static_cast<int>(addr), static_cast<int>(table)));
}

// End node of the last row (k_dp's global selection: strict '<' from the
// constant band 0) of half `half`, written to dres[task].  SCORE: the match
// count is written as 0.
template <int G, bool SCORE>
__device__ __forceinline__ void dp16_end(const s16x2 (&Z)[2 * G + 3], const s16x2 (&M)[2 * G + 3], int4* dres,
                                         uint32_t task, int half, int rows) {
  constexpr int W = 2 * G + 1;
  s16x2 bz = Z[0];
  uint32_t bm = pk_bits(M[0]), be = pk_bits(pk_splat(-1));
#pragma unroll
  for (int b = 1; b <= W; ++b) {
    const uint32_t m = pk_less(bz, Z[b]);
    bz = __builtin_elementwise_max(bz, Z[b]);
    if constexpr (!SCORE) bm = pk_pick(m, pk_bits(M[b]), bm);
    be = pk_pick(m, pk_bits(pk_splat(b - 1)), be);
  }
  const int sc = (half ? bz.y : bz.x) - (kGapOpen + kGapExt);
  const int mt = SCORE ? 0 : half ? pk_val(bm).y : pk_val(bm).x;
  const int e = (half ? pk_val(be).y : pk_val(be).x) - G;
  dres[task] = make_int4(rows + e, rows, mt, sc);
}

// dp16_end of the score pass in drifted units (dp_row16_drift): band b's Z
// is Z~[b] - 2 - 2 b - d4 at the row whose drift is 2 b + d4.
template <int G>
__device__ __forceinline__ void dp16_end_drift(const s16x2 (&Z)[2 * G + 3], int4* dres, uint32_t task, int half,
                                               int rows, int d4) {
  constexpr int W = 2 * G + 1;
  // compared with the common d4 left in (band 0, never drifted, gets it)
  s16x2 bz = Z[0] + pk_splat(d4);
  uint32_t be = pk_bits(pk_splat(-1));
#pragma unroll
  for (int b = 1; b <= W; ++b) {
    const s16x2 z = Z[b] - pk_splat(2 + 2 * b);
    const uint32_t m = pk_less(bz, z);
    bz = __builtin_elementwise_max(bz, z);
    be = pk_pick(m, pk_bits(pk_splat(b - 1)), be);
  }
  const int sc = (half ? bz.y : bz.x) - d4 - (kGapOpen + kGapExt);
  const int e = (half ? pk_val(be).y : pk_val(be).x) - G;
  dres[task] = make_int4(rows + e, rows, 0, sc);
}

// Tasks order[2t] and order[2t + 1] (neighbours in the longest-first order,
// so near-equal rows) in the low and high halves; a lone last task is run
// in both.  Target and query classes come from the 4-bit arrays (a.tnib,
// a.qnib) through NibWin.  One chunk of 64 task pairs: lane l runs pair
// 64 chunk + l of the kernel's task range [first, first + n).  The rows go
// in step over the wavefront up to its longest task: a lane past its own
// rows computes on (clamped reads, no stores; values stay inside the range
// proof's bounds, which hold for any 30000 rows), and each half's end node
// is taken at its own last row, when the wavefront reaches the next such
// row of any lane (ev, uniform), so the row loop has no per-lane branch.
template <int G, bool EDGE, bool SCORE, bool PROF>
__device__ __forceinline__ void dp16_chunk(const DPArgs& a, uint32_t first, uint32_t n, uint32_t chunk) {
  constexpr int W = 2 * G + 1;
  // the score pass outside EDGE runs in drifted units (dp_row16_drift; with
  // PROF, dp_row16_prof)
  constexpr bool DRIFT = SCORE && !EDGE;
  static_assert(!PROF || DRIFT, "query profiles in the drifted score pass only");
  // the target streams start kLead positions early, so that both streams
  // enter groups on rows r = 0 (mod 16) with the same group parity
  constexpr int kLead = (32 - (W & 31)) & 31;
  constexpr int kInitGroups = (W + kLead) / 16;  // even
  const uint32_t tl = chunk * 64 + (threadIdx.x & 63u);
  const bool valid = 2 * tl < n;
  // a lane past the range runs pair 0 with no rows (no end node)
  const uint32_t t = valid ? tl : 0u;
  const uint32_t task0 = a.order[first + 2 * t];
  const uint32_t task1 = 2 * t + 1 < n ? a.order[first + 2 * t + 1] : task0;
  const int L = a.lmer;
  const uint2* tn2 = reinterpret_cast<const uint2*>(a.tnib);
  const uint2* qn2 = reinterpret_cast<const uint2*>(a.qnib);
  // one 16-slot unit per 16-byte block, the mirror as many again
  const int32_t tumax = 2 * a.t_bmax + 1, qumax = 2 * a.q_bmax + 1;
  NibWin tw[2], qw[2];
  int rows[2];
  // EDGE: ring position j (target index tb + ts j) is 'X' from j = xk[h] on
  int xk[2] = {INT32_MAX, INT32_MAX};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t task = h ? task1 : task0;
    const uint32_t i = task >> 1;
    const bool bwd = (task & 1u) == 0u;
    const int s = a.hs[i], tIdx = a.ht[i], qIdx = a.hq[i];
    rows[h] = bwd ? a.qlen[s] - (qIdx + L) : qIdx;
    const int ts = bwd ? 1 : -1;
    const int tb = bwd ? tIdx + L - G : tIdx + G - 1;
    if constexpr (EDGE) xk[h] = bwd ? a.tlen[a.hgs[i]] - tb : tb + 1;
    tw[h].init(tn2, tumax, static_cast<int64_t>(a.tslot[a.hgs[i]]) + tb - ts * kLead, ts);
    qw[h].init(qn2, qumax, static_cast<int64_t>(a.qslot[s]) + (bwd ? qIdx + L : qIdx - 1), ts);
    if (!valid) rows[h] = 0;
  }
  s16x2 Z[W + 2], V[W + 2], M[W + 2];
  constexpr int half = G + 1;
#pragma unroll
  for (int b = 0; b < W + 2; ++b) {
    int sc, vv, mm;
    if (b < half) {
      sc = vv = mm = -30000;
    } else if (b == half) {
      sc = 0; vv = kGapOpen; mm = 0;
    } else {
      sc = (b - half) * kGapExt + kGapOpen; vv = sc + kGapOpen; mm = 0;
    }
    // drifted (dp_row16_drift): as the last row of a period, d = 2 b + 4 (kDriftRows - 1)
    const int dz = DRIFT && b >= 1 && b <= W ? 2 + 2 * b + 4 * (kDriftRows - 1) : 0;
    const int dv = DRIFT && b >= 1 && b <= W ? 2 * b + 4 * (kDriftRows - 1) : 0;
    Z[b] = pk_splat(sc + kGapOpen + kGapExt + dz);
    V[b] = pk_splat(vv + dv);
    M[b] = pk_splat(mm);
  }
  // k_dp's edges: band 1's horizontal term, band W's vertical term (the band
  // W + 1 constants above)
  constexpr int kH1 = -30000 + kGapExt;
  constexpr int kVW = (W + 1 - half) * kGapExt + kGapOpen + kGapOpen + kGapExt;
  // target ring: stream positions kLead .. kLead + W - 1
  PairRows tp, qp;
  uint32_t T[W + 4];
#pragma unroll
  for (int g = 0; g < kInitGroups; ++g) {
    uint32_t a0, a1, b0, b1;
    if (g & 1) {
      tw[0].group<1>(tn2, tumax, a0, a1);
      tw[1].group<1>(tn2, tumax, b0, b1);
    } else {
      tw[0].group<0>(tn2, tumax, a0, a1);
      tw[1].group<0>(tn2, tumax, b0, b1);
    }
    tp.build(a0, a1, b0, b1);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const uint32_t v = tp.next(j);
      const int k = 16 * g + j - kLead;
      if (k >= 0 && k < W) {
        T[k] = EDGE ? v | (k >= xk[0] ? 0x8u : 0u) | (k >= xk[1] ? 0x80000u : 0u) : PROF ? v | 0x0C040C00u : v;
      }
    }
  }
  // the wavefront's longest task, and its least end row after `after`
  // (INT32_MAX: none)
  const int rmax = __builtin_amdgcn_readlane(wave_max_scan_incl(max(rows[0], rows[1])), 63);
  auto next_end = [&](int after) {
    const int e = min(rows[0] > after ? rows[0] : INT32_MAX, rows[1] > after ? rows[1] : INT32_MAX);
    return -__builtin_amdgcn_readlane(wave_max_scan_incl(-e), 63);
  };
  int ev = next_end(0);
  // 32 rows per iteration: even group, 16 rows, odd group, 16 rows (each
  // stream register then has one definition in the loop, so the loads land
  // in place); every active lane is at the same row
  // Rows go U at a time, unrolled: row i of U reads the target ring from
  // T[i] and appends its new position at T[W + i], so the ring moves back by
  // U once per U rows and, with U = 4, the code words shift at fixed rows.
  // Wider bands keep U = 1 (four rows' ring would not fit the registers).
  constexpr int U = G <= 8 ? 4 : 1;
  // the end node of half h at its last row rws - 1 (drifted values back to
  // k_dp's first)
  auto end_node = [&](uint32_t task, int h, int rws) {
    if constexpr (DRIFT) {
      dp16_end_drift<G>(Z, a.dres, task, h, rws, 4 * ((rws - 1) & (kDriftRows - 1)));
    } else {
      dp16_end<G, SCORE>(Z, M, a.dres, task, h, rws);
    }
  };
  // PROF with U = 4: the four rows' query profiles from the lane table,
  // looked up together before the rows (px / py: halves 0 / 1)
  constexpr bool kTab = PROF && U == 4;
  const uint32_t qtab = kTab ? query_profile_table() : 0u;
  uint32_t px[4], py[4];
  // rows 0-3 of w (a PairRows word): row i's classes in bits 4 i (half 0)
  // and 16 + 4 i (half 1)
  auto lookup4 = [&](uint32_t w) {
    px[0] = query_profile_lookup(w << 2, qtab);
    py[0] = query_profile_lookup(w >> 14, qtab);
    px[1] = query_profile_lookup(w >> 2, qtab);
    py[1] = query_profile_lookup(w >> 18, qtab);
    px[2] = query_profile_lookup(w >> 6, qtab);
    py[2] = query_profile_lookup(w >> 22, qtab);
    px[3] = query_profile_lookup(w >> 10, qtab);
    py[3] = query_profile_lookup(w >> 26, qtab);
  };
  auto row = [&](auto oi, int j, int r) {
    constexpr int i = decltype(oi)::value;
    // target position r + W + kLead = r (mod 16) (kTab: set before the rows)
    if constexpr (!kTab) T[W + i] = tp.next(j);
    if constexpr (EDGE) T[W + i] |= (r + W >= xk[0] ? 0x8u : 0u) | (r + W >= xk[1] ? 0x80000u : 0u);
    if constexpr (kTab) {
      const int d4 = 4 * (r & (kDriftRows - 1));
      dp_row16_prof<W, i, i == 0>(Z, V, T, px[i], py[i], pk_splat(kH1 + 2 + d4), pk_splat(kVW + 2 * W + d4));
    } else {
      const uint32_t Q = qp.next(j);  // j: the row inside the group (mod 4 is enough)
      if constexpr (PROF) {
        T[W + i] |= 0x0C040C00u;
        const int d4 = 4 * (r & (kDriftRows - 1));
        dp_row16_prof<W, i>(Z, V, T, query_profile<0>(Q), query_profile<2>(Q), pk_splat(kH1 + 2 + d4),
                            pk_splat(kVW + 2 * W + d4));
      } else if constexpr (DRIFT) {
        const int d4 = 4 * (r & (kDriftRows - 1));
        dp_row16_drift<W, i>(Z, V, T, Q | 0x0C000C00u, pk_splat(kH1 + 2 + d4), pk_splat(kVW + 2 * W + d4));
      } else {
        dp_row16<W, i, EDGE, SCORE>(Z, V, M, T, Q);
      }
    }
    if (r + 1 == ev) {
      if (rows[0] == ev) end_node(task0, 0, ev);
      if (rows[1] == ev) end_node(task1, 1, ev);
      ev = next_end(ev);
    }
  };
  auto rows16 = [&](int r0) {
    if constexpr (DRIFT) {
      if ((r0 & (kDriftRows - 1)) == 0) {
        // a new period: d drops by 4 kDriftRows (row r0 - 1 becomes row -1)
#pragma unroll
        for (int b = 1; b <= W; ++b) {
          Z[b] = Z[b] - pk_splat(4 * kDriftRows);
          V[b] = V[b] - pk_splat(4 * kDriftRows);
        }
      }
    }
#pragma unroll 1
    for (int j = 0; j < 16; j += U) {
      const int ju = __builtin_amdgcn_readfirstlane(j);
      const int r = r0 + ju;
      if (r >= rmax) break;
      if constexpr (U == 1) {
        row(std::integral_constant<int, 0>{}, ju, r);
      } else {
        if constexpr (kTab) {
          // the four rows' query profiles (qp.p0 holds their classes), then
          // their target positions, while the lookups are under way
          lookup4(qp.p0);
          qp.skip4();
#pragma unroll
          for (int i = 0; i < 4; ++i) T[W + i] = tp.next(i) | 0x0C040C00u;
        }
        row(std::integral_constant<int, 0>{}, 0, r);
        if (r + 1 >= rmax) break;
        row(std::integral_constant<int, 1>{}, 1, r + 1);
        if (r + 2 >= rmax) break;
        row(std::integral_constant<int, 2>{}, 2, r + 2);
        if (r + 3 >= rmax) break;
        row(std::integral_constant<int, 3>{}, 3, r + 3);
      }
#pragma unroll
      for (int k = 0; k < W; ++k) T[k] = T[k + U];
    }
  };
#pragma unroll 1
  for (int rg = 0; rg < rmax; rg += 32) {
    {
      uint32_t a0, a1, b0, b1, c0, c1, d0, d1;
      qw[0].group<0>(qn2, qumax, a0, a1);
      qw[1].group<0>(qn2, qumax, b0, b1);
      tw[0].group<0>(tn2, tumax, c0, c1);
      tw[1].group<0>(tn2, tumax, d0, d1);
      qp.build(a0, a1, b0, b1);
      tp.build(c0, c1, d0, d1);
    }
    rows16(rg);
    if (rg + 16 >= rmax) break;
    {
      uint32_t a0, a1, b0, b1, c0, c1, d0, d1;
      qw[0].group<1>(qn2, qumax, a0, a1);
      qw[1].group<1>(qn2, qumax, b0, b1);
      tw[0].group<1>(tn2, tumax, c0, c1);
      tw[1].group<1>(tn2, tumax, d0, d1);
      qp.build(a0, a1, b0, b1);
      tp.build(c0, c1, d0, d1);
    }
    rows16(rg + 16);
  }
}

// The chunks of the range go out longest first.  Static (PERSIST false):
// one wavefront per chunk in snake order (snake_chunk), the dispatcher
// starting workgroups as others finish.  Persistent (PERSIST true): a grid
// of about one occupancy round, each wavefront taking the next chunk from a
// device counter (a.chunk_next, zero before the launch) until none is left,
// so a wavefront that finishes early takes over the remaining work instead
// of waiting for its workgroup; every wavefront exits once the counter
// passes the last chunk.  SCORE: scores and end nodes only, match counts 0
// (the score pass of a two-pass DP stage: the tasks of the hits that pass
// the bad-hit filters run again without SCORE).
// Widest band of the persistent schedule: wider bands' kernels spill many
// registers to scratch, and with gap 16 the persistent kernel on long
// extensions gave a wrong row and then a memory fault whose cause is not
// found, so those gaps run the static schedule only.
constexpr int kPersistMaxGap = 8;

template <int G, bool EDGE, bool PERSIST, bool SCORE>
static __global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_dp16(DPArgs a) {
  const uint32_t first = a.split[EDGE ? 1 : 0];
  const uint32_t n = a.split[EDGE ? 2 : 1] - first;
  const uint32_t nchunk = ((n + 1) / 2 + 63) / 64;
  // the drifted score pass with 9 or more bands takes query profiles
  // (dp_row16_prof) when every task of the chunk lies on a target sequence
  // without class 4 (clean flag 3)
  auto run = [&](uint32_t c) {
    if constexpr (SCORE && !EDGE && 2 * G + 1 >= 9) {
      const uint32_t t = c * 64 + (threadIdx.x & 63u);
      bool prof = true;
      if (2 * t < n) {
        const uint32_t t0 = a.order[first + 2 * t];
        const uint32_t t1 = 2 * t + 1 < n ? a.order[first + 2 * t + 1] : t0;
        prof = (a.tclean[a.hgs[t0 >> 1]] & a.tclean[a.hgs[t1 >> 1]] & 2u) != 0u;
      }
      if (__all(prof)) {
        dp16_chunk<G, EDGE, SCORE, true>(a, first, n, c);
        return;
      }
    }
    dp16_chunk<G, EDGE, SCORE, false>(a, first, n, c);
  };
  if constexpr (PERSIST) {
    for (;;) {
      uint32_t c = 0;
      if ((threadIdx.x & 63u) == 0) c = atomicAdd(a.chunk_next, 1u);
      c = __builtin_amdgcn_readfirstlane(c);
      if (c >= nchunk) return;
      run(c);
    }
  } else {
    const uint32_t wv = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    if (wv >= nchunk) return;
    run(snake_chunk(wv, nchunk, a.round));
  }
}

// ------------------------------------------- banded DP, one lane per band
//
// Long extensions (at least kBandRows rows) have a latency of rows x W cells
// in k_dp / k_dp16, one lane walking the whole band row by row.  Here a task
// pair occupies W consecutive lanes, lane b holding band b (k_dp's band
// b + 1): cell (r, b) of task A is computed at step 2r + b and of task B at
// step 2r + b + 1.  Cell (r, b) reads cell (r, b-1) (horizontal term, new
// score and match count) and cell (r-1, b+1) (old score, vertical term,
// match count), both computed one step earlier by the neighbouring lanes
// (DPP lane shifts), and its own cell (r-1, b) from two steps earlier.  The
// query class of row r moves up the lanes from lane 0, which reads it; the
// target class of position r + b moves down from lane W-1.  Each lane
// alternates A and B, so every lane computes a cell every step, and a
// task's latency is two steps per row.  Arithmetic, tie rules and margins
// are k_dp's (int32, any letters, both modes).

constexpr int kBandRows = 1024;

// value of lane - 1 / lane + 1 (DPP wave_shr:1 / wave_shl:1); the wavefront's
// end lanes get 0, which no group reads (its end lanes use constants)
__device__ __forceinline__ int from_below(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, true); }
__device__ __forceinline__ int from_above(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xF, 0xF, true); }
// a where m is all ones, b where it is zero
__device__ __forceinline__ int pick(int m, int a, int b) { return b ^ ((a ^ b) & m); }

// One task's per-lane view in k_dp_band.
struct BandTask {
  int rows;
  bool clean;     // query strand and target sequence without star or X letters
  bool inwin;     // every target position the task's lane W-1 reads lies in [lo, hi)
  int r;          // this lane's row at its next step of this task
  int ti, ts;     // target index lane W-1 reads at stream position 0, and its step
  int lo, hi;     // t_mask range
  int S, M;       // this lane's cell of the previous row
  int fS, fM, fR; // local: best cell of this band (score, match, row); global: unused
  uint32_t t0;    // target class of position b - 1 (this lane's row -1 export)
  ByteWin st;     // lane 0: query classes; lane W-1: target classes
};

// row_words without branches: every lane of k_dp_band holds its own row.
__device__ __forceinline__ void row_words_sel(uint32_t qc, uint32_t& wpm, uint32_t& wmc) {
  constexpr uint32_t kPm = 0x8FDDDDDu, kMc = 0x8F00000u, kStar = 0x8FFFFFFu, kAllX = 0x8888888u;
  const uint32_t sh = 4u * (qc & 3u);
  const uint32_t m4 = 0u - static_cast<uint32_t>(qc < 4u);
  const uint32_t other_pm = qc == kClsOther ? kPm : qc == kClsStar ? kStar : kAllX;
  const uint32_t other_mc = qc == kClsOther ? kMc : qc == kClsStar ? kStar : kAllX;
  wpm = ((kPm ^ (0xCu << sh)) & m4) | (other_pm & ~m4);
  wmc = ((kMc | (0x1u << sh)) & m4) | (other_mc & ~m4);
}

// One step of task X in lane b: computes cell (X.r, b) from the exports of
// the neighbouring lanes' previous step and updates this lane's exports.
// CLEAN (every task of the wavefront without star or X letters, so the query
// class is 0-4 and the target class 0-4 or 'X' outside the window): the
// substitution is +1 / match 1 for equal classes below 4, -1000 / -1000 for
// an 'X' target, else -3 / 0 — krnlMatrix.cuh's entries for those classes,
// without the row words.  AM false: every lane is inside its task's rows
// (the wavefront's steady steps), so no lane shows its row -1 values; WIN
// false: every target position read lies in the window (no 'X' mask).
template <int G, bool LOCAL, bool CLEAN, bool AM = true, bool WIN = true>
__device__ __forceinline__ void band_step(BandTask& X, int b, int j, int Sinit, int Vinit, int Minit, int& eS,
                                          int& eV, int& eH, int& eM, uint32_t& eT, uint32_t& eQ) {
  constexpr int W = 2 * G + 1;
  constexpr int half = G + 1;
  // band W+1's constant row (k_dp's dp_init)
  constexpr int kTopS = (W + 1 - half) * kGapExt + kGapOpen;
  constexpr int kTopV = kTopS + kGapOpen;
  const int lS = from_below(eS), lH = from_below(eH), lM = from_below(eM);
  const int rS = from_above(eS), rV = from_above(eV), rM = from_above(eM);
  const uint32_t lQ = static_cast<uint32_t>(from_below(static_cast<int>(eQ)));
  const uint32_t rT = static_cast<uint32_t>(from_above(static_cast<int>(eT)));
  const uint32_t sb = X.st.next(j);
  const bool bot = b == 0, top = b == W - 1;
  const uint32_t qc = bot ? sb : lQ;
  const int ti = X.ti + X.ts * j;  // the target index of stream position j
  const uint32_t tm = WIN ? 4u * (((ti >= X.lo) & (ti < X.hi)) ? sb : kClsX) : 4u * sb;  // t_mask
  const uint32_t T = top ? tm : rT;
  const int v0 = max(rS + (kGapOpen + kGapExt), rV + kGapExt);
  const int h0 = max(lS + (kGapOpen + kGapExt), lH + kGapExt);
  const int v = top ? max(kTopS + (kGapOpen + kGapExt), kTopV + kGapExt) : v0;
  const int h = bot ? max(-30000 + (kGapOpen + kGapExt), -30000 + kGapExt) : h0;
  const int Mv = top ? 0 : rM;
  const int Mh = bot ? -30000 : lM;
  int pm, mc;
  if constexpr (CLEAN) {
    const uint32_t tcl = T >> 2;
    const bool eq = (qc == tcl) & (qc < 4u);
    const bool bad = WIN && tcl == kClsX;  // (no 'X' when every read is in the window)
    pm = bad ? kBad : eq ? kMatch : kMismatch;
    mc = bad ? kBad : eq ? 1 : 0;
  } else {
    uint32_t wpm, wmc;
    row_words_sel(qc, wpm, wmc);
    const int pr = __builtin_amdgcn_sbfe(static_cast<int>(wpm), T, 4);
    const int mr = __builtin_amdgcn_sbfe(static_cast<int>(wmc), T, 4);
    const bool bad = pr == -8;
    pm = bad ? kBad : pr;
    mc = bad ? kBad : mr;
  }
  const int pt = X.S + pm;
  const int vh = max(v, h);
  const int nm_gap = v > h ? Mv : Mh;
  const int nm_diag = X.M + mc;
  const int dmask = -static_cast<int>(pt > vh);
  const int nm = nm_gap ^ ((nm_gap ^ nm_diag) & dmask);
  const int ns = max(pt, vh);
  if constexpr (AM) {
    const int am = -static_cast<int>(static_cast<uint32_t>(X.r) < static_cast<uint32_t>(X.rows));
    if (LOCAL) {
      const int bm = am & -static_cast<int>(X.fS < ns);
      X.fS = pick(bm, ns, X.fS);
      X.fM = pick(bm, nm, X.fM);
      X.fR = pick(bm, X.r, X.fR);
    }
    X.S = pick(am, ns, X.S);
    X.M = pick(am, nm, X.M);
    // before its first row a lane shows its row -1 (k_dp's initial band)
    eS = pick(am, ns, Sinit);
    eV = pick(am, v, Vinit);
    eM = pick(am, nm, Minit);
    eT = static_cast<uint32_t>(pick(am, static_cast<int>(T), static_cast<int>(X.t0)));
  } else {
    if (LOCAL) {
      const int bm = -static_cast<int>(X.fS < ns);
      X.fS = pick(bm, ns, X.fS);
      X.fM = pick(bm, nm, X.fM);
      X.fR = pick(bm, X.r, X.fR);
    }
    X.S = ns;
    X.M = nm;
    eS = ns;
    eV = v;
    eM = nm;
    eT = T;
  }
  eH = h;
  eQ = qc;
  X.r += 1;
}

template <int G, bool LOCAL>
static __global__ __launch_bounds__(kBlock) void k_dp_band(DPArgs a) {
  constexpr int W = 2 * G + 1;
  constexpr int kPairs = 64 / W;  // task pairs per wavefront
  constexpr int half = G + 1;
  __shared__ int red[kBlock / 64][64][6];
  // these few long tasks bound the DP's end: issue before the other DP
  // kernels' wavefronts on the same SIMD
  __builtin_amdgcn_s_setprio(3);
  const int lane = lane_id();
  const int grp = lane / W;
  const int b = lane - grp * W;
  const uint32_t nband = a.split[0];
  // workgroups stride over the task pairs (the grid is an upper bound, or
  // capped); the loop bound is the same for every wavefront of a workgroup
  constexpr uint32_t kBlockPairs = (kBlock / 64) * kPairs;
  for (uint32_t blk = blockIdx.x; blk * kBlockPairs < (nband + 1) / 2; blk += gridDim.x) {
    const uint32_t pair = (blk * (kBlock / 64) + wave_id()) * kPairs + static_cast<uint32_t>(grp);
    // every lane stays to the workgroup barrier; lanes without a pair run
    // task 0's setup and no steps
    const bool live = grp < kPairs && 2 * pair < nband;
    const uint32_t taskA = live ? a.order[2 * pair] : 0u;
    const uint32_t taskB = live && 2 * pair + 1 < nband ? a.order[2 * pair + 1] : taskA;
    const int L = a.lmer;
    const int64_t tbytes = 16 * (static_cast<int64_t>(a.t_bmax) + 1);
    // X0 is the task this lane works on at even steps: A for even b, B for odd
    BandTask X0, X1;
    auto setup = [&](BandTask& x, int k) {
      const uint32_t task = ((b & 1) ^ k) ? taskB : taskA;
      const uint32_t i = task >> 1;
      const bool bwd = (task & 1u) == 0u;
      const int s = a.hs[i], gs = a.hgs[i], tIdx = a.ht[i], qIdx = a.hq[i];
      const int qLen = a.qlen[s];
      x.rows = bwd ? qLen - (qIdx + L) : qIdx;
      x.clean = a.qclean[s] && a.tclean[gs];
      // task A's cell (r, b) at step 2r + b, B's at 2r + b + 1: the row at
      // this lane's first step of the task (step k)
      const int c = (b & 1) ^ k;  // 0: task A's timing, 1: B's
      x.r = (k - b - c) >> 1;
      const int tLen = a.tlen[gs];
      const int64_t tslot = static_cast<int64_t>(a.tslot[gs]);
      const int w = a.twin[gs];
      const int wlo = static_cast<int>(static_cast<int64_t>(a.wslot[w]) - tslot);
      const int whi = static_cast<int>(static_cast<int64_t>(a.wslot[w + 1]) - tslot);
      x.ts = bwd ? 1 : -1;
      const int tb = bwd ? tIdx + L - G : tIdx + G - 1;
      x.lo = bwd ? wlo : max(wlo, 0);
      x.hi = bwd ? min(whi, tLen) : whi;
      // lane W-1 reads position r + W - 1 of row r: its stream starts at
      // position G, so stream index j = floor(step / 2) in every lane
      x.ti = tb + x.ts * G;
      // lane W-1 reads positions ti, ti + ts, ... for at most rows + W + 1 calls
      {
        const int64_t a0 = static_cast<int64_t>(x.ti), a1 = a0 + static_cast<int64_t>(x.ts) * (x.rows + W + 1);
        x.inwin = min(a0, a1) >= x.lo && max(a0, a1) < x.hi;
      }
      const bool top = b == W - 1;
      x.st.init(top ? a.tbase : a.qbase, top ? a.t_bmax : a.q_bmax,
                top ? tslot + x.ti : static_cast<int64_t>(a.qslot[s]) + (bwd ? qIdx + L : qIdx - 1), x.ts);
      const int tp = tb + x.ts * (b - 1);
      const int64_t g = min(max(tslot + tp, int64_t{0}), tbytes - 1);
      x.t0 = t_mask(a.tbase[g], tp, x.lo, x.hi);
      x.fS = 0; x.fM = 0; x.fR = -1;
    };
    setup(X0, 0);
    setup(X1, 1);
    // row -1 of band b + 1 (k_dp's dp_init)
    const int bi = b + 1;
    const int Sinit = bi < half ? -30000 : bi == half ? 0 : (bi - half) * kGapExt + kGapOpen;
    const int Vinit = bi < half ? -30000 : bi == half ? kGapOpen : Sinit + kGapOpen;
    const int Minit = bi < half ? -30000 : 0;
    X0.S = X1.S = Sinit;
    X0.M = X1.M = Minit;
    // exports of step -1 (this lane's X1 task at row -1)
    int eS = Sinit, eV = Vinit, eH = 0, eM = Minit;
    uint32_t eT = X1.t0, eQ = 0;
    // last step: task A ends at lane W-1, step 2 (rowsA - 1) + W - 1; B one later
    const bool odd = (b & 1) != 0;
    const int rowsA = odd ? X1.rows : X0.rows, rowsB = odd ? X0.rows : X1.rows;
    const int tend = live ? max(2 * (rowsA - 1) + W - 1, 2 * (rowsB - 1) + W) + 1 : 0;
    // the wavefront's steady steps: calls [i0, i1) find every lane inside
    // both its tasks' rows (call i of task X is its row X.r + i); they run
    // four calls at a time without the row -1 picks (and without the window
    // mask when every task's reads stay in its window), the calls before and
    // after them one at a time
    // (a wavefront with no pair has none; lanes without a pair count only
    // through the wavefront's largest call count, which bounds i1)
    int i0 = 0, i1 = 0;
    {
      const int lo = live ? max(-X0.r, -X1.r) : 0;
      const int hi = live ? min(X0.rows - X0.r, X1.rows - X1.r) : INT32_MAX;
      const int mlo = __builtin_amdgcn_readlane(wave_max_scan_incl(lo), 63);
      const int mhi = -__builtin_amdgcn_readlane(wave_max_scan_incl(-hi), 63);
      const int mcalls = __builtin_amdgcn_readlane(wave_max_scan_incl((tend + 1) / 2), 63);
      const int cap = min(mhi, mcalls);
      if (__any(live)) {
        i0 = (mlo + 3) & ~3;
        i1 = cap > i0 ? i0 + ((cap - i0) & ~3) : i0;
      }
    }
    // lanes without a pair ran task 0's setup: they count as clean and in the window
    const bool clean = __all(!live || (X0.clean && X1.clean));
    const bool inwin = __all(!live || (X0.inwin && X1.inwin));
    auto run = [&](auto cleanc, auto winc) {
      constexpr bool C = decltype(cleanc)::value, WN = decltype(winc)::value;
      // every lane runs the calls before i0 (lanes without a pair compute
      // on task 0 and store nothing); then the steady calls; then each lane
      // to its own end
#pragma unroll 1
      for (int i = 0; i < i0; ++i) {
        const int j = __builtin_amdgcn_readfirstlane(i);
        band_step<G, LOCAL, C>(X0, b, j, Sinit, Vinit, Minit, eS, eV, eH, eM, eT, eQ);
        band_step<G, LOCAL, C>(X1, b, j, Sinit, Vinit, Minit, eS, eV, eH, eM, eT, eQ);
      }
#pragma unroll 1
      for (int i = i0; i < i1; i += 4) {
        const int j4 = (__builtin_amdgcn_readfirstlane(i) >> 2) << 2;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          band_step<G, LOCAL, C, false, WN>(X0, b, j4 + u, Sinit, Vinit, Minit, eS, eV, eH, eM, eT, eQ);
          band_step<G, LOCAL, C, false, WN>(X1, b, j4 + u, Sinit, Vinit, Minit, eS, eV, eH, eM, eT, eQ);
        }
      }
#pragma unroll 1
      for (int t = 2 * i1; t < tend; t += 2) {
        const int j = __builtin_amdgcn_readfirstlane(t) >> 1;
        band_step<G, LOCAL, C>(X0, b, j, Sinit, Vinit, Minit, eS, eV, eH, eM, eT, eQ);
        band_step<G, LOCAL, C>(X1, b, j, Sinit, Vinit, Minit, eS, eV, eH, eM, eT, eQ);
      }
    };
    if (clean) {
      if (inwin) run(std::true_type{}, std::false_type{});
      else run(std::true_type{}, std::true_type{});
    } else {
      run(std::false_type{}, std::true_type{});
    }
    // end node per task: the group's lanes through LDS, scanned in band order
    const int wv = wave_id();
    const int aS = LOCAL ? (odd ? X1.fS : X0.fS) : (odd ? X1.S : X0.S);
    const int aM = LOCAL ? (odd ? X1.fM : X0.fM) : (odd ? X1.M : X0.M);
    const int bS = LOCAL ? (odd ? X0.fS : X1.fS) : (odd ? X0.S : X1.S);
    const int bM = LOCAL ? (odd ? X0.fM : X1.fM) : (odd ? X0.M : X1.M);
    red[wv][lane][0] = aS;
    red[wv][lane][1] = aM;
    red[wv][lane][2] = odd ? X1.fR : X0.fR;
    red[wv][lane][3] = bS;
    red[wv][lane][4] = bM;
    red[wv][lane][5] = odd ? X0.fR : X1.fR;
    __syncthreads();
    if (live && b < 2 && !(b == 1 && taskB == taskA)) {
      const int k = b;  // lane 0 finishes task A, lane 1 task B
      const int base = grp * W;
      const int rows = k ? rowsB : rowsA;
      int move, endp, bestS, bestM;
      if (LOCAL) {
        // k_dp keeps the first cell, in row-major order, of the highest score
        // above 0: highest score, then lowest row, then lowest band, as one key
        // (0 for a band that never rose above 0)
        uint64_t best = 0;
        int bq = -1;
        for (int q = 0; q < W; ++q) {
          const int sc = red[wv][base + q][3 * k], rr = red[wv][base + q][3 * k + 2];
          const uint64_t key = rr < 0 ? 0ull
                                      : (static_cast<uint64_t>(static_cast<uint32_t>(sc) ^ 0x80000000u) << 32) |
                                            (static_cast<uint64_t>(0x3FFFFFF - rr) << 6) | static_cast<uint64_t>(63 - q);
          if (key > best) {
            best = key;
            bq = q;
          }
        }
        bestS = bq < 0 ? 0 : red[wv][base + bq][3 * k];
        bestM = bq < 0 ? 0 : red[wv][base + bq][3 * k + 1];
        move = bq < 0 ? 0 : red[wv][base + bq][3 * k + 2] + 1;
        endp = (bq < 0 ? G : bq) - G;
      } else {
        bestS = -30000;  // band 0
        bestM = -30000;
        int e = -1;
        for (int q = 0; q < W; ++q) {
          const int sc = red[wv][base + q][3 * k];
          if (bestS < sc) {
            bestS = sc;
            bestM = red[wv][base + q][3 * k + 1];
            e = q;
          }
        }
        move = rows;
        endp = e - G;
      }
      a.dres[k ? taskB : taskA] = make_int4(move + endp, move, bestM, bestS);
    }
    __syncthreads();  // red is rewritten by the next round
  }
}

// Launchers of k_dp for a runtime gap (2..16); each lives in its own
// translation unit (dp_global.hip, dp_local.hip).
void launch_dp_global(int gap, const DPArgs& a, hipStream_t st);
void launch_dp_local(int gap, const DPArgs& a, hipStream_t st);
// k_dp16 over order[split[0], split[1]) and, with edge, over
// order[split[1], split[2]) (dp16.hip; its score-only form, dp16_score.hip),
// k_dp_band over order[0, split[0]) (dp_band.hip); a.n bounds the task count.
// reserve: the score pass's workgroups reserve LDS so that at most three fit
// per CU (dp16_launch.hpp); launch_dp16 ignores it.
void launch_dp16(int gap, bool edge, bool persist, uint32_t grid_cap, bool reserve, const DPArgs& a, hipStream_t st);
void launch_dp16_score(int gap, bool edge, bool persist, uint32_t grid_cap, bool reserve, const DPArgs& a,
                       hipStream_t st);
void launch_dp_band(int gap, bool local, const DPArgs& a, hipStream_t st);

// DP task order by class (key bits from classbit up): 0 the long tasks
// (k_dp_band: with band = 1 the long tasks k_dp16 cannot take, with band = 2
// every long task), 1 those k_dp16 can take, 2 those k_dp16 can take with
// 'X' past the sequence end (EDGE), 3 the rest (k_dp); each class by
// rows descending (key low bits = (maxrows - rows) >> rowshift, the shift
// keeping the class bits inside the same radix passes), so the longest
// extensions start first and a wavefront's lanes run near-equal row counts.
// The order only affects speed: every task writes its own result.
struct TaskKeyArgs {
  uint32_t nhit;
  const int32_t *hs, *hgs, *ht, *hq, *qlen;
  const uint8_t* qclean;   // per query strand: no star / X class
  const int32_t* tlen;
  const int32_t* twin;
  const uint64_t* tslot;
  const uint64_t* wslot;
  const uint8_t* tclean;   // per target sequence: no star / X class
  int lmer, gap, packed;   // packed = 0: no task goes to k_dp16 (local mode)
  int band;                // 0: no task goes to k_dp_band; 1: long tasks k_dp16 cannot take; 2: every long task
  const int* band_rows;    // tasks with at least this many rows go to k_dp_band
  int maxrows, rowshift, classbit;
  uint64_t* key;
  uint32_t* task;
  unsigned long long* rows16;    // 8 counters: rows of the k_dp16 tasks [0, 4), past-the-end ones [4, 8)
  unsigned long long* rowsband;  // 8 counters: rows of the long tasks
};

// The arguments come from device memory (k_store writes them per search),
// so a captured launch of this kernel serves every search.  Slots
// [2 nhit, nslots) get keys above every task of class 3 and below none of
// the same key's tasks (stable sort: after them): the task count padded to
// a bucket that a captured chain was built for.
static __global__ __launch_bounds__(kBlock) void k_dp_task_keys(const TaskKeyArgs* __restrict__ ap, uint32_t nslots) {
  __shared__ uint32_t wsum[3 * (kBlock / 64)];
  const TaskKeyArgs& a = *ap;
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  bool p16 = false, edge = false, long_task = false;
  int rows = 0;
  if (t >= 2 * a.nhit && t < nslots) {
    a.key[t] = (3ull << a.classbit) | static_cast<uint64_t>(a.maxrows >> a.rowshift);
    a.task[t] = t;
  }
  if (t < 2 * a.nhit) {
    const uint32_t i = t >> 1;
    const bool bwd = (t & 1u) == 0u;
    const int s = a.hs[i], q = a.hq[i];
    const int qLen = a.qlen[s];
    rows = bwd ? qLen - (q + a.lmer) : q;
    if (a.packed && rows > 0 && qLen <= kPacked16MaxLen && a.qclean[s]) {
      const int gs = a.hgs[i], tIdx = a.ht[i];
      if (a.tclean[gs]) {
        // the band covers target indices tb + ts * [0, rows + W - 2] (k_dp's ring)
        const int64_t span = rows + 2 * a.gap - 1;
        const int64_t first = bwd ? tIdx + a.lmer - a.gap : tIdx + a.gap - 1 - span;
        const int64_t last = first + span;
        const int64_t ts = static_cast<int64_t>(a.tslot[gs]);
        const int w = a.twin[gs];
        const int64_t wlo = static_cast<int64_t>(a.wslot[w]) - ts, whi = static_cast<int64_t>(a.wslot[w + 1]) - ts;
        const int64_t tl = a.tlen[gs];
        const int64_t lo = max(wlo, int64_t{0});
        const int64_t hi = min(whi, tl);
        p16 = first >= lo && last < hi;
        // past the guarded end only: backward reads past tLen ('X'), forward
        // reads before 0 ('X'), the other end inside sequence and window
        edge = !p16 && (bwd ? first >= lo && last >= tl : first < 0 && last < whi);
      }
    }
    long_task = a.band && rows >= *a.band_rows && (a.band == 2 || !(p16 || edge));
    if (long_task) p16 = edge = false;
    const uint64_t cls = long_task ? 0u : p16 ? 1u : edge ? 2u : 3u;
    a.key[t] = (cls << a.classbit) | static_cast<uint64_t>((a.maxrows - rows) >> a.rowshift);
    a.task[t] = t;
  }
  const uint32_t r16 = wave_sum_incl(p16 ? static_cast<uint32_t>(rows) : 0u);
  const uint32_t re = wave_sum_incl(edge ? static_cast<uint32_t>(rows) : 0u);
  const uint32_t rb = wave_sum_incl(long_task ? static_cast<uint32_t>(rows) : 0u);
  if (lane_id() == 63) {
    wsum[wave_id()] = r16;
    wsum[kBlock / 64 + wave_id()] = re;
    wsum[2 * (kBlock / 64) + wave_id()] = rb;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    unsigned long long tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) tot += wsum[threadIdx.x * (kBlock / 64) + w];
    // rows16[0, 4): k_dp16 tasks, rows16[4, 8): its past-the-end tasks
    unsigned long long* dst = threadIdx.x == 2 ? a.rowsband + (blockIdx.x & 7u)
                                               : a.rows16 + 4 * threadIdx.x + (blockIdx.x & 3u);
    if (tot) atomicAdd(dst, tot);
  }
}

// Rows of every DP task, summed into 8 counters (by workgroup); arguments
// from device memory as for k_dp_task_keys.
static __global__ __launch_bounds__(kBlock) void k_dp_rows_sum(const TaskKeyArgs* __restrict__ ap,
                                                               unsigned long long* __restrict__ sums) {
  __shared__ uint32_t wsum[kBlock / 64];
  const TaskKeyArgs& a = *ap;
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  int rows = 0;
  if (t < 2 * a.nhit) {
    const uint32_t i = t >> 1;
    const int q = a.hq[i];
    rows = (t & 1u) ? q : a.qlen[a.hs[i]] - (q + a.lmer);
  }
  const uint32_t r = wave_sum_incl(static_cast<uint32_t>(max(rows, 0)));
  if (lane_id() == 63) wsum[wave_id()] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) tot += wsum[w];
    if (tot) atomicAdd(&sums[blockIdx.x & 7u], tot);
  }
}

// One value into device memory (the arguments of a captured kernel).
template <class T>
static __global__ void k_store(T* __restrict__ dst, T v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *dst = v;
}

// Row threshold of the long tasks: a task over twice one SIMD's average
// share of all rows (total / slots, slots = SIMDs x 128 tasks per wavefront
// row in k_dp16) would set the end of the per-lane kernels by itself, so it
// goes to k_dp_band (whose latency per row is 2 steps, not W cells, at
// about five times the instructions per cell); never below kBandRows.
static __global__ void k_band_threshold(const unsigned long long* __restrict__ sums, unsigned long long slots,
                                        int* __restrict__ band_rows) {
  if (threadIdx.x != 0) return;
  unsigned long long tot = 0;
  for (int k = 0; k < 8; ++k) tot += sums[k];
  const unsigned long long lim = 2ull * (tot / (slots ? slots : 1ull));
  *band_rows = static_cast<int>(max(static_cast<unsigned long long>(kBandRows), min(lim, 0x7FFFFFFFull)));
}

// The 16-bit and long task counts (split[2] - split[0], split[0]) into the
// work counters.
static __global__ void k_split_counts(const uint32_t* __restrict__ split, unsigned long long* __restrict__ out) {
  if (threadIdx.x == 0) {
    out[0] = split[2] - split[0];
    out[1] = split[0];
  }
}

// Number of keys of a sorted array below `limit` (the tasks of the classes
// under a bound): one wavefront narrows [lo, hi) 64-fold per step.
static __global__ void k_count_below(const uint64_t* __restrict__ keys, uint32_t n, uint64_t limit,
                                     uint32_t* __restrict__ out, unsigned long long* __restrict__ out64) {
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t lo = 0, hi = n;  // answer in [lo, hi]
  while (lo < hi) {
    const uint32_t step = (hi - lo + 63) / 64;
    const uint32_t p = lo + lane * step;
    const bool above = p >= hi || keys[p] >= limit;
    const uint64_t b = __builtin_amdgcn_ballot_w64(above);
    const uint32_t f = b ? static_cast<uint32_t>(__builtin_ctzll(b)) : 64u;  // first lane at or past the answer
    const uint32_t nlo = f == 0 ? lo : lo + (f - 1) * step + 1;
    const uint32_t nhi = f == 64 ? hi : min(hi, lo + f * step);
    lo = nlo;
    hi = nhi;
  }
  if (lane == 0) {
    *out = lo;
    if (out64) *out64 = lo;
  }
}

// ---------------------------------------------------------------- results

// Order key of one hit, compared lexicographically (a, b, c):
//   a = (target window << 32) | ~score order     (score descending)
//   b = ((0x7FFFFFFF - seed qIdx) << 1) | strand bit
//   c = (target sequence << 32) | diagonal order
// (a, b, c) is the CDeviceHitList order of one unit; (a.low, a.high, b, c)
// is the row order fixResult leaves among the pair's rows.
// ----------------------------------------- two-pass DP stage, second pass
//
// The first pass ran the 16-bit tasks (order[split[0], split[2])) through
// k_dp16's score pass: scores and end nodes, which is all the bad-hit
// filters read.  The second pass runs the 16-bit tasks of the hits that
// passed (kflag) again with match counts, in three lists that keep the
// task order: tasks of at least *band_rows rows (k_dp_band, for latency;
// the threshold comes from the listed tasks' rows, k_redo_rows and
// k_band_threshold, as the first pass's long tasks do), k_dp16's others,
// and its past-the-end others.  flags holds one 0/1 array of nslot entries
// per list; the rows of the listed tasks are summed into rows_sum[0..8).

// Rows of the second pass's tasks (the 16-bit tasks of the hits that passed
// the filters), summed into 8 counters (by workgroup).
static __global__ __launch_bounds__(kBlock) void k_redo_rows(const uint32_t* __restrict__ order,
                                                             const uint32_t* __restrict__ split,
                                                             const uint32_t* __restrict__ kflag,
                                                             const int32_t* __restrict__ hs,
                                                             const int32_t* __restrict__ hq,
                                                             const int32_t* __restrict__ qlen, int lmer,
                                                             uint32_t nslot, unsigned long long* __restrict__ sums) {
  __shared__ uint32_t wsum[kBlock / 64];
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  int rows = 0;
  if (j < nslot && j >= split[0] && j < split[2]) {
    const uint32_t task = order[j];
    const uint32_t i = task >> 1;
    if (kflag[i]) rows = (task & 1u) == 0u ? qlen[hs[i]] - (hq[i] + lmer) : hq[i];
  }
  const uint32_t r = wave_sum_incl(static_cast<uint32_t>(max(rows, 0)));
  if (lane_id() == 63) wsum[wave_id()] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) tot += wsum[w];
    if (tot) atomicAdd(&sums[blockIdx.x & 7u], tot);
  }
}

static __global__ __launch_bounds__(kBlock) void k_redo_flags(const uint32_t* __restrict__ order,
                                                              const uint32_t* __restrict__ split,
                                                              const uint32_t* __restrict__ kflag,
                                                              const int32_t* __restrict__ hs,
                                                              const int32_t* __restrict__ hq,
                                                              const int32_t* __restrict__ qlen, int lmer,
                                                              uint32_t nslot, const int* __restrict__ band_rows,
                                                              uint32_t* __restrict__ flags,
                                                              unsigned long long* __restrict__ rows_sum) {
  __shared__ unsigned long long wsum[kBlock / 64];
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  const int brows = *band_rows;
  uint32_t cat = 0;
  int rows = 0;
  if (j < nslot && j >= split[0] && j < split[2]) {
    const uint32_t task = order[j];
    const uint32_t i = task >> 1;
    if (kflag[i]) {
      rows = (task & 1u) == 0u ? qlen[hs[i]] - (hq[i] + lmer) : hq[i];
      cat = rows >= brows ? 1u : j < split[1] ? 2u : 3u;
    }
  }
  if (j < nslot) {
    flags[j] = cat == 1u;
    flags[nslot + j] = cat == 2u;
    flags[2ull * nslot + j] = cat == 3u;
  }
  unsigned long long c = cat ? static_cast<unsigned long long>(rows) : 0ull;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
  if (lane_id() == 0) wsum[wave_id()] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int k = 0; k < kBlock / 64; ++k) t += wsum[k];
    if (t) atomicAdd(&rows_sum[blockIdx.x & 7u], t);
  }
}

// The three lists into order2 (list k at the sum of the counts before it;
// offs: exclusive scans of the flags, cnt: their totals) and their bounds
// into split2 in k_dp16 / k_dp_band's layout: [0, split2[0]) lane per band,
// [split2[0], split2[1]) k_dp16, [split2[1], split2[2]) past the end.
static __global__ __launch_bounds__(kBlock) void k_redo_scatter(const uint32_t* __restrict__ order,
                                                                const uint32_t* __restrict__ flags,
                                                                const uint32_t* __restrict__ offs, uint32_t nslot,
                                                                const uint32_t* __restrict__ cnt,
                                                                uint32_t* __restrict__ order2,
                                                                uint32_t* __restrict__ split2) {
  const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t b1 = cnt[0], b2 = cnt[0] + cnt[1];
  if (j == 0) {
    split2[0] = b1;
    split2[1] = b2;
    split2[2] = b2 + cnt[2];
  }
  if (j >= nslot) return;
  if (flags[j]) order2[offs[j]] = order[j];
  else if (flags[nslot + j]) order2[b1 + offs[nslot + j]] = order[j];
  else if (flags[2ull * nslot + j]) order2[b2 + offs[2ull * nslot + j]] = order[j];
}

// Match counts of the hits that passed the filters from the second pass's
// results (k_dp_combine's sum; the other fields are the first pass's and
// unchanged).
static __global__ __launch_bounds__(kBlock) void k_match_patch(uint32_t nhit, const uint32_t* __restrict__ kflag,
                                                               const int4* __restrict__ dres,
                                                               const int32_t* __restrict__ hs,
                                                               const int32_t* __restrict__ hq,
                                                               const int32_t* __restrict__ qlen, int lmer,
                                                               int32_t* __restrict__ om) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nhit || !kflag[i]) return;
  const int qIdx = hq[i];
  const int b = qlen[hs[i]] - (qIdx + lmer) > 0 ? dres[2 * i].z : 0;
  const int f = qIdx > 0 ? dres[2 * i + 1].z : 0;
  om[i] = lmer + b + f;
}

struct RowKey {
  uint64_t a, b, c;
};

__device__ __forceinline__ bool unit_less(const RowKey& x, const RowKey& y) {
  return x.a < y.a || (x.a == y.a && (x.b < y.b || (x.b == y.b && x.c < y.c)));
}

__device__ __forceinline__ uint64_t swap_halves(uint64_t v) { return (v << 32) | (v >> 32); }

__device__ __forceinline__ bool row_less(const RowKey& x, const RowKey& y) {
  const uint64_t xa = swap_halves(x.a), ya = swap_halves(y.a);
  return xa < ya || (xa == ya && (x.b < y.b || (x.b == y.b && x.c < y.c)));
}

// calculateEvalue (krnlCalculateEvalue.cu:18-21) in FP64, the E-value cut
// as an integer score threshold per strand (score >= smin[s] <=> E <= cutOff,
// threshold computed on the host with the same formula), the too-short cut
// (…deleteBadHits.cuh:41-48), the hit's order key, and the number of DP
// cells the hits took (summed over cells[0..8)).
static __global__ __launch_bounds__(kBlock) void k_hit_filter(uint32_t n, const int32_t* __restrict__ hs,
                                                              const int32_t* __restrict__ hgs,
                                                              const int32_t* __restrict__ ht,
                                                              const int32_t* __restrict__ hq,
                                                              const int32_t* __restrict__ score,
                                                              const int32_t* __restrict__ qhl,
                                                              const int32_t* __restrict__ qlen,
                                                              const int32_t* __restrict__ seq_win_t,
                                                              const int32_t* __restrict__ smin, int use_cut,
                                                              int min_qhl, double kd, double lambda,
                                                              double* __restrict__ ev, uint32_t* __restrict__ flag,
                                                              RowKey* __restrict__ rk, int lmer, int band,
                                                              unsigned long long* __restrict__ cells) {
  __shared__ unsigned long long wsum[kBlock / 64];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  // DP cells of the hits (both directions: band x (qLen - lMer) rows): one
  // atomic per workgroup, spread over 8 counters
  unsigned long long c = i < n ? static_cast<unsigned long long>(band) * (qlen[hs[i]] - lmer) : 0ull;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
  if (lane_id() == 0) wsum[wave_id()] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int k = 0; k < kBlock / 64; ++k) t += wsum[k];
    if (t) atomicAdd(&cells[blockIdx.x & 7u], t);
  }
  if (i >= n) return;
  const int s = hs[i];
  const int sc = score[i];
  ev[i] = kd * static_cast<double>(qlen[s]) * exp(-lambda * static_cast<double>(sc));
  const bool keep = (!use_cut || sc >= smin[s]) && !(min_qhl > qhl[i]);
  flag[i] = keep ? 1u : 0u;
  const int gs = hgs[i], q = hq[i];
  RowKey k;
  k.a = (static_cast<uint64_t>(static_cast<uint32_t>(seq_win_t[gs])) << 32) |
        static_cast<uint64_t>(0x7FFFFFFFu - static_cast<uint32_t>(sc));
  k.b = (static_cast<uint64_t>(0x7FFFFFFFu - static_cast<uint32_t>(q)) << 1) | static_cast<uint64_t>(s & 1);
  k.c = (static_cast<uint64_t>(static_cast<uint32_t>(gs)) << 32) |
        static_cast<uint64_t>(static_cast<uint32_t>(ht[i] - q) ^ 0x80000000u);
  rk[i] = k;
}

// Order keys of the kept hits, least significant first: word0 = (target
// sequence, seed diagonal), word1 = (maxq - seed qIdx, strand), word2 =
// (target window, query pair, maxscore - score).  Sorting stably by word0,
// then word1, then word2 gives each unit's CDeviceHitList order whatever
// order the hits arrived in.
static __global__ __launch_bounds__(kBlock) void k_order_keys(uint32_t n, const uint32_t* __restrict__ kept,
                                                              const uint32_t* __restrict__ hoff,
                                                              const int32_t* __restrict__ hs,
                                                              const int32_t* __restrict__ hgs,
                                                              const int32_t* __restrict__ ht,
                                                              const int32_t* __restrict__ hq,
                                                              const int32_t* __restrict__ score,
                                                              const int32_t* __restrict__ seq_win_t, int d_bits,
                                                              int d_off, int maxq, int pair_bits, int score_bits,
                                                              int smax, uint64_t* __restrict__ w0,
                                                              uint64_t* __restrict__ w1, uint64_t* __restrict__ w2,
                                                              uint32_t* __restrict__ rec) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || !kept[i]) return;
  const uint32_t o = hoff[i];
  const uint32_t s = static_cast<uint32_t>(hs[i]);
  w0[o] = (static_cast<uint64_t>(hgs[i]) << d_bits) | static_cast<uint64_t>(ht[i] - hq[i] + d_off);
  w1[o] = (static_cast<uint64_t>(maxq - hq[i]) << 1) | (s & 1u);
  w2[o] = (static_cast<uint64_t>(seq_win_t[hgs[i]]) << (pair_bits + score_bits)) |
          (static_cast<uint64_t>(s >> 1) << score_bits) | static_cast<uint64_t>(smax - score[i]);
  rec[o] = i;
}

// deleteHits_duplicateResult (CDeviceHitList.cu:32-128): in unit order, drop
// a row equal in (tID, tIdx, qID, qIdx) to its predecessor of the same unit.
static __global__ __launch_bounds__(kBlock) void k_dedup(uint32_t n, const uint32_t* __restrict__ perm,
                                                  const uint32_t* __restrict__ rec, const int32_t* __restrict__ hs,
                                                  const int32_t* __restrict__ hgs, const int32_t* __restrict__ ot,
                                                  const int32_t* __restrict__ oq,
                                                  const int32_t* __restrict__ seq_win_t,
                                                  const int32_t* __restrict__ seq_win_q, uint32_t* __restrict__ flag) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t f = 1;
  if (i > 0) {
    const uint32_t a = rec[perm[i]], b = rec[perm[i - 1]];
    const bool same_unit = seq_win_t[hgs[a]] == seq_win_t[hgs[b]] && seq_win_q[hs[a]] == seq_win_q[hs[b]];
    if (same_unit && hgs[a] == hgs[b] && ot[a] == ot[b] && hs[a] == hs[b] && oq[a] == oq[b]) f = 0;
  }
  flag[i] = f;
}

static __global__ __launch_bounds__(kBlock) void k_compact_perm(uint32_t n, const uint32_t* __restrict__ perm,
                                                         const uint32_t* __restrict__ rec,
                                                         const uint32_t* __restrict__ flag,
                                                         const uint32_t* __restrict__ off, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || !flag[i]) return;
  out[off[i]] = rec[perm[i]];
}

// fixResult order key (query pair, maxscore - score, target window).
static __global__ __launch_bounds__(kBlock) void k_final_keys(uint32_t n, const uint32_t* __restrict__ order,
                                                       const int32_t* __restrict__ hs, const int32_t* __restrict__ hgs,
                                                       const int32_t* __restrict__ score,
                                                       const int32_t* __restrict__ seq_win_t, int score_bits,
                                                       int win_bits, int smax, uint64_t* __restrict__ w) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = order[i];
  w[i] = (static_cast<uint64_t>(static_cast<uint32_t>(hs[r]) >> 1) << (score_bits + win_bits)) |
         (static_cast<uint64_t>(smax - score[r]) << win_bits) | static_cast<uint64_t>(seq_win_t[hgs[r]]);
}

struct OutArgs {
  uint32_t n;
  const uint32_t* order;   // record index per output row (may be composed with perm)
  const uint32_t* perm;    // optional second level (null = identity)
  const int32_t* hs;
  const int32_t* hgs;
  const int32_t* ot;
  const int32_t* oq;
  const int32_t* thl;
  const int32_t* qhl;
  const int32_t* match;
  const int32_t* score;
  const double* ev;
  const int32_t* tref;
  const int32_t* qref;
  const int32_t* twin;
  const int32_t* qwin;
  int32_t* out_i32;   // 10 planes of n: tID tIdx qID qIdx tHL qHL match score twin qwin
  double* out_ev;
  // with n_dev: n = *n_dev (the grid covers an upper bound), and the planes
  // and E-values are laid out from out_base as above (10 x 4 n bytes, then
  // 8 n)
  const uint32_t* n_dev;
  uint8_t* out_base;
};

static __global__ __launch_bounds__(kBlock) void k_gather_out(OutArgs a) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t n = a.n_dev ? *a.n_dev : a.n;
  if (i >= n) return;
  const uint32_t r = a.order[a.perm ? a.perm[i] : i];
  const int gs = a.hgs[r], s = a.hs[r];
  int32_t* o = a.n_dev ? reinterpret_cast<int32_t*>(a.out_base) : a.out_i32;
  double* oev = a.n_dev ? reinterpret_cast<double*>(a.out_base + 40ull * n) : a.out_ev;
  o[0 * n + i] = a.tref[gs];
  o[1 * n + i] = a.ot[r];
  o[2 * n + i] = a.qref[s];
  o[3 * n + i] = a.oq[r];
  o[4 * n + i] = a.thl[r];
  o[5 * n + i] = a.qhl[r];
  o[6 * n + i] = a.match[r];
  o[7 * n + i] = a.score[r];
  o[8 * n + i] = a.twin[gs];
  o[9 * n + i] = a.qwin[s];
  oev[i] = a.ev[r];
}

// ---------------------------------------------------------------- per-pair results

// Pairs with up to kPairSmall hits are ordered by thread-per-hit ranking
// inside the pair; larger pairs take the sort path.
constexpr uint32_t kPairSmall = 256;

// Hits of query pair k ('+' strand 2k, '-' strand 2k+1) are [hbase[2k],
// pair_end) in the dense strand-ordered hit list.
__device__ __forceinline__ uint32_t pair_end(const uint32_t* hbase, uint32_t nseq, uint32_t nhit, uint32_t k) {
  return 2 * k + 2 < nseq ? hbase[2 * k + 2] : nhit;
}

struct PairArgs {
  uint32_t nhit, nseq;
  const uint32_t* hbase;   // per strand: first dense hit
  const int32_t* hs;
  const int32_t* hgs;
  const int32_t* ot;       // extended tIdx / qIdx
  const int32_t* oq;
  const uint32_t* kflag;   // passed the E-value and length cuts
  const RowKey* rk;
};

// deleteHits_duplicateResult (CDeviceHitList.cu:32-128) within one pair: a
// kept hit is dropped when its predecessor in unit order is in the same
// target window and equal in (tID, tIdx, qID, qIdx).  A row of another pair
// never precedes it within a unit (pairs are contiguous in unit order) and
// could not be equal.  Hits of pairs over kPairSmall go to the sort path
// (gflag) and count in nbig once per pair.
static __global__ __launch_bounds__(kBlock) void k_pair_dedup(PairArgs a, uint32_t* __restrict__ dflag,
                                                              uint32_t* __restrict__ gflag,
                                                              uint32_t* __restrict__ nbig) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nhit) return;
  const uint32_t k = static_cast<uint32_t>(a.hs[i]) >> 1;
  const uint32_t lo = a.hbase[2 * k], hi = pair_end(a.hbase, a.nseq, a.nhit, k);
  const bool big = hi - lo > kPairSmall;
  if (big && i == lo) atomicAdd(nbig, 1u);
  const bool kept = a.kflag[i] != 0;
  uint32_t f = 0;
  if (kept && !big) {
    const RowKey me = a.rk[i];
    uint32_t pred = 0xFFFFFFFFu;
    RowKey best{};
    for (uint32_t j = lo; j < hi; ++j) {
      if (!a.kflag[j]) continue;
      const RowKey o = a.rk[j];
      if (unit_less(o, me) && (pred == 0xFFFFFFFFu || unit_less(best, o))) {
        best = o;
        pred = j;
      }
    }
    f = 1;
    if (pred != 0xFFFFFFFFu && (best.a >> 32) == (me.a >> 32) && a.hgs[pred] == a.hgs[i] && a.ot[pred] == a.ot[i] &&
        a.hs[pred] == a.hs[i] && a.oq[pred] == a.oq[i])
      f = 0;
  }
  dflag[i] = f;
  gflag[i] = kept && big ? 1u : 0u;
}

// Row slot of each surviving hit of a small pair: the pair's first output
// row (exclusive scan of the per-hit row counts at the pair's first hit)
// plus the hit's rank among the pair's survivors in row order.
static __global__ __launch_bounds__(kBlock) void k_pair_place(PairArgs a, const uint32_t* __restrict__ dflag,
                                                              const uint32_t* __restrict__ doff,
                                                              uint32_t* __restrict__ order) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nhit) return;
  const uint32_t k = static_cast<uint32_t>(a.hs[i]) >> 1;
  const uint32_t lo = a.hbase[2 * k], hi = pair_end(a.hbase, a.nseq, a.nhit, k);
  if (hi - lo > kPairSmall || !dflag[i]) return;
  const RowKey me = a.rk[i];
  uint32_t rank = 0;
  for (uint32_t j = lo; j < hi; ++j)
    if (dflag[j] && row_less(a.rk[j], me)) ++rank;
  order[doff[lo] + rank] = i;
}

// Rows of the large pairs from the sort path (pair-major, row order within
// each pair): each pair's row count lands on its first hit's slot of the
// per-hit count array, and its first row index in gfirst at the same slot.
static __global__ __launch_bounds__(kBlock) void k_big_counts(uint32_t n, const uint32_t* __restrict__ grow,
                                                              const int32_t* __restrict__ hs,
                                                              const uint32_t* __restrict__ hbase,
                                                              uint32_t* __restrict__ rows, uint32_t* __restrict__ gfirst) {
  const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= n) return;
  const uint32_t k = static_cast<uint32_t>(hs[grow[r]]) >> 1;
  const uint32_t lo = hbase[2 * k];
  if (r == 0 || (static_cast<uint32_t>(hs[grow[r - 1]]) >> 1) != k) gfirst[lo] = r;
  atomicAdd(&rows[lo], 1u);
}

static __global__ __launch_bounds__(kBlock) void k_big_place(uint32_t n, const uint32_t* __restrict__ grow,
                                                             const int32_t* __restrict__ hs,
                                                             const uint32_t* __restrict__ hbase,
                                                             const uint32_t* __restrict__ gfirst,
                                                             const uint32_t* __restrict__ doff,
                                                             uint32_t* __restrict__ order) {
  const uint32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= n) return;
  const uint32_t lo = hbase[static_cast<uint32_t>(hs[grow[r]]) & ~1u];
  order[doff[lo] + (r - gfirst[lo])] = grow[r];
}

// Score range of the hits sent to the sort path (its key width).
static __global__ __launch_bounds__(kBlock) void k_score_range(uint32_t n, const uint32_t* __restrict__ flag,
                                                               const int32_t* __restrict__ score,
                                                               int32_t* __restrict__ srange) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  int lo = INT32_MAX, hi = INT32_MIN;
  if (i < n && flag[i]) lo = hi = score[i];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const int a = __shfl_xor(lo, d, 64), b = __shfl_xor(hi, d, 64);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if (lane_id() == 0 && lo <= hi) {
    atomicMin(&srange[0], lo);
    atomicMax(&srange[1], hi);
  }
}

// ---------------------------------------------------------------- printResult row selection
//
// printResult (CHostResultHolder.cu:131-243) walks the fixResult order with
// the last printed row's ten fields and a per-label count: a row equal to
// the last printed one is skipped, and after a label's first row at most
// numOut rows of that label print.  The last printed row always equals the
// row just before in every field that decides: a row not printed was either
// equal to the last printed row (then it has its fields) or past its label's
// limit (then it has its label, and so does the next row of that label, also
// past the limit).  So, in sorted order, with row i-1 for the last printed
// row: distinct(i) = i is first or differs from i-1 in a printed field;
// new(i) = its query label differs from i-1's (row 0: from the empty label
// the walk starts with); printed(i) = distinct(i) and (new(i) or numOut = -1
// or rank(i) < numOut), rank(i) = distinct rows of i's label run before i.
struct FinalCols {
  const uint32_t* perm;     // sorted position -> row
  const int32_t* qlabel;    // per row: query label id (0: the empty label)
  const int32_t* qid;
  const int32_t* qidx;
  const int32_t* qhl;
  const int32_t* tlabel;    // per row: target label id
  const int32_t* tidx;
  const int32_t* tstart;    // per row: the target's start offset ('@' line), added to tIdx
  const int32_t* thl;
  const int32_t* match;
  const int32_t* score;
  const double* evalue;
};

static __global__ __launch_bounds__(kBlock) void k_final_flags(FinalCols c, uint32_t n, uint32_t* __restrict__ dflag,
                                                               uint8_t* __restrict__ nflag,
                                                               uint32_t* __restrict__ head) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = c.perm[i];
  bool nw, dist;
  if (i == 0) {
    nw = c.qlabel[r] != 0;
    dist = true;
  } else {
    const uint32_t o = c.perm[i - 1];
    nw = c.qlabel[r] != c.qlabel[o];
    dist = nw || c.qidx[r] != c.qidx[o] || ((c.qid[r] ^ c.qid[o]) & 1) != 0 || c.tlabel[r] != c.tlabel[o] ||
           c.tidx[r] + c.tstart[r] != c.tidx[o] + c.tstart[o] || c.thl[r] != c.thl[o] || c.qhl[r] != c.qhl[o] || c.match[r] != c.match[o] ||
           c.score[r] != c.score[o] || c.evalue[r] != c.evalue[o];
  }
  dflag[i] = dist ? 1u : 0u;
  nflag[i] = nw ? 1 : 0;
  head[i] = nw ? i : 0u;  // max-scanned: the start of i's label run
}

static __global__ __launch_bounds__(kBlock) void k_final_print(uint32_t n, int num_out,
                                                               const uint32_t* __restrict__ dflag,
                                                               const uint8_t* __restrict__ nflag,
                                                               const uint32_t* __restrict__ dpre,
                                                               const uint32_t* __restrict__ run,
                                                               uint8_t* __restrict__ printed,
                                                               uint32_t* __restrict__ pflag) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = run[i];                  // start of i's label run (0 before the first new label)
  const bool nw = nflag[i] != 0;
  const uint32_t rank = dpre[i] - dpre[s];    // distinct rows of the run before i
  const bool p = dflag[i] && (nw || num_out == -1 || rank < static_cast<uint32_t>(num_out));
  printed[i] = p ? 1 : 0;
  pflag[i] = p ? 1u : 0u;
}

// resultSorting key for clast_result_sort: (qID/2, -score) biased to u64.
static __global__ void k_result_keys(uint32_t n, const int32_t* __restrict__ qid, const int32_t* __restrict__ score,
                              uint64_t* __restrict__ w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // signed -> order-preserving unsigned by flipping the sign bit; the score
  // half is complemented so that a larger score sorts first
  const uint64_t q = static_cast<uint64_t>(static_cast<uint32_t>(qid[i] / 2) ^ 0x80000000u);
  const uint64_t sc = static_cast<uint64_t>(0xFFFFFFFFu - (static_cast<uint32_t>(score[i]) ^ 0x80000000u));
  w[i] = (q << 32) | sc;
}

}  // namespace clast
