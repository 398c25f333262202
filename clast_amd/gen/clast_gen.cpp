// Seeded synthetic DNA generator for the CLAST workloads: targets (uniform
// ACGT, N bases, diverged repeat-family copies, (AT)n/(CA)n tracts) and reads
// sampled from them (substitutions, geometric indels, reverse complement,
// random no-hit reads, fixed / log-uniform / mixed lengths).
// One xoshiro256** stream per call, seeded through splitmix64, so every
// workload is reproducible from its seed alone.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (auto& x : s) {
      seed += 0x9E3779B97F4A7C15ull;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      x = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
  // number of failures before the first success, success probability p
  int64_t geometric_skip(double p) {
    if (p <= 0) return INT64_MAX;
    if (p >= 1) return 0;
    const double u = 1.0 - uniform();
    return static_cast<int64_t>(std::floor(std::log(u) / std::log1p(-p)));
  }
};

const char kACGT[4] = {'A', 'C', 'G', 'T'};

char other_base(Rng& r, char b) {
  int k = 0;
  switch (b) { case 'A': k = 0; break; case 'C': k = 1; break; case 'G': k = 2; break; case 'T': k = 3; break;
    default: return kACGT[r.below(4)]; }
  return kACGT[(k + 1 + r.below(3)) & 3];
}

char comp(char b) {
  switch (b) { case 'A': return 'T'; case 'T': return 'A'; case 'C': return 'G'; case 'G': return 'C'; default: return 'N'; }
}

void fill_uniform(Rng& r, char* out, int64_t n) {
  int64_t i = 0;
  while (i < n) {
    uint64_t x = r.next();
    for (int k = 0; k < 32 && i < n; ++k, ++i, x >>= 2) out[i] = kACGT[x & 3];
  }
}

}  // namespace

extern "C" {

struct clast_gen_target_spec {
  uint64_t seed;
  int32_t n_seq;
  int64_t seq_len;
  double n_rate;          // per-base probability of 'N'
  double repeat_frac;     // fraction of bases covered by family copies
  int32_t n_families;
  int32_t fam_min, fam_max;
  double div_min, div_max;  // per-family substitution divergence range
  double lowc_frac;       // fraction covered by (AT)n / (CA)n tracts
};

struct clast_gen_target {
  std::vector<char> bases;           // n_seq * seq_len
  std::vector<int64_t> copy_pos;     // global offset of each repeat copy
  std::vector<int32_t> copy_len;
};

void* clast_gen_target_create(const clast_gen_target_spec* sp) {
  auto* t = new clast_gen_target();
  Rng r(sp->seed);
  const int64_t total = static_cast<int64_t>(sp->n_seq) * sp->seq_len;
  t->bases.resize(total);
  fill_uniform(r, t->bases.data(), total);
  if (sp->repeat_frac > 0 && sp->n_families > 0 && sp->seq_len > sp->fam_max) {
    std::vector<std::vector<char>> fam(sp->n_families);
    std::vector<double> div(sp->n_families);
    for (int f = 0; f < sp->n_families; ++f) {
      const int len = sp->fam_min + static_cast<int>(r.below(sp->fam_max - sp->fam_min + 1));
      fam[f].resize(len);
      fill_uniform(r, fam[f].data(), len);
      div[f] = sp->div_min + (sp->div_max - sp->div_min) * r.uniform();
    }
    int64_t covered = 0;
    const int64_t want = static_cast<int64_t>(sp->repeat_frac * total);
    while (covered < want) {
      const int f = static_cast<int>(r.below(sp->n_families));
      const int len = static_cast<int>(fam[f].size());
      const int64_t seq = r.below(sp->n_seq);
      const int64_t off = r.below(sp->seq_len - len + 1);
      char* dst = t->bases.data() + seq * sp->seq_len + off;
      for (int k = 0; k < len; ++k) dst[k] = r.uniform() < div[f] ? other_base(r, fam[f][k]) : fam[f][k];
      t->copy_pos.push_back(seq * sp->seq_len + off);
      t->copy_len.push_back(len);
      covered += len;
    }
  }
  if (sp->lowc_frac > 0 && sp->seq_len > 600) {
    int64_t covered = 0;
    const int64_t want = static_cast<int64_t>(sp->lowc_frac * total);
    while (covered < want) {
      const int len = 50 + static_cast<int>(r.below(451));
      const int64_t seq = r.below(sp->n_seq);
      const int64_t off = r.below(sp->seq_len - len + 1);
      const char* unit = r.below(2) ? "AT" : "CA";
      char* dst = t->bases.data() + seq * sp->seq_len + off;
      for (int k = 0; k < len; ++k) dst[k] = unit[k & 1];
      covered += len;
    }
  }
  if (sp->n_rate > 0) {
    int64_t i = r.geometric_skip(sp->n_rate);
    while (i < total) {
      t->bases[i] = 'N';
      const int64_t s = r.geometric_skip(sp->n_rate);
      if (s >= total) break;
      i += 1 + s;
    }
  }
  return t;
}

int64_t clast_gen_target_size(void* h) { return static_cast<int64_t>(static_cast<clast_gen_target*>(h)->bases.size()); }
int64_t clast_gen_target_ncopies(void* h) { return static_cast<int64_t>(static_cast<clast_gen_target*>(h)->copy_pos.size()); }
void clast_gen_target_copy(void* h, char* out, int64_t* copy_pos, int32_t* copy_len) {
  auto* t = static_cast<clast_gen_target*>(h);
  std::memcpy(out, t->bases.data(), t->bases.size());
  if (copy_pos) std::memcpy(copy_pos, t->copy_pos.data(), t->copy_pos.size() * sizeof(int64_t));
  if (copy_len) std::memcpy(copy_len, t->copy_len.data(), t->copy_len.size() * sizeof(int32_t));
}
void clast_gen_target_free(void* h) { delete static_cast<clast_gen_target*>(h); }

enum { LEN_FIXED = 0, LEN_LOGUNIFORM = 1, LEN_MIXED = 2 };

struct clast_gen_read_spec {
  uint64_t seed;
  int64_t n_reads;
  int32_t len_mode;
  int32_t len_a, len_b;   // fixed: len_a; log-uniform: [len_a, len_b]
  int32_t len_c;          // mixed: len_c with probability mix_frac, else log-uniform [len_a, len_b]
  double mix_frac;
  double sub_rate, indel_rate;
  double indel_p;         // geometric length parameter
  int32_t indel_max;
  double rc_frac;         // reverse-complemented fraction of sampled reads
  double random_frac;     // i.i.d. random (no-hit) reads
  double repeat_read_frac;  // reads started inside a repeat copy (needs copies)
};

struct clast_gen_reads {
  std::vector<char> bases;
  std::vector<int64_t> offsets;  // n_reads + 1
  std::vector<int64_t> origin;   // global target offset of the first sampled base, -1 random
  std::vector<int8_t> strand;    // 0 '+', 1 '-'
};

// targets: concatenated sequences, seq_off: n_seq+1 offsets into it.
void* clast_gen_reads_create(const clast_gen_read_spec* sp, const char* target, const int64_t* seq_off,
                             int32_t n_seq, const int64_t* copy_pos, const int32_t* copy_len, int64_t n_copies) {
  auto* o = new clast_gen_reads();
  Rng r(sp->seed);
  const int64_t total = seq_off[n_seq];
  o->offsets.push_back(0);
  std::vector<char> read;
  for (int64_t i = 0; i < sp->n_reads; ++i) {
    int len = sp->len_a;
    if (sp->len_mode == LEN_LOGUNIFORM || (sp->len_mode == LEN_MIXED && r.uniform() >= sp->mix_frac)) {
      const double la = std::log(static_cast<double>(sp->len_a)), lb = std::log(static_cast<double>(sp->len_b));
      len = static_cast<int>(std::lround(std::exp(la + (lb - la) * r.uniform())));
    } else if (sp->len_mode == LEN_MIXED) {
      len = sp->len_c;
    }
    read.clear();
    int64_t origin = -1;
    int8_t strand = 0;
    if (r.uniform() < sp->random_frac || total == 0) {
      read.resize(len);
      fill_uniform(r, read.data(), len);
    } else {
      // choose a start: inside a repeat copy, or uniform over the target
      const int64_t span = static_cast<int64_t>(len) + len / 4 + 16;
      int64_t seq = 0, start = 0;
      for (int tries = 0; tries < 64; ++tries) {
        int64_t g;
        if (n_copies > 0 && r.uniform() < sp->repeat_read_frac) {
          const int64_t c = r.below(n_copies);
          g = copy_pos[c] + r.below(copy_len[c]);
        } else {
          g = r.below(total);
        }
        seq = std::upper_bound(seq_off, seq_off + n_seq + 1, g) - seq_off - 1;
        const int64_t slen = seq_off[seq + 1] - seq_off[seq];
        start = g - seq_off[seq];
        if (slen >= span && start + span <= slen) break;
        if (slen >= span) { start = slen - span; break; }
      }
      const int64_t slen = seq_off[seq + 1] - seq_off[seq];
      const char* src = target + seq_off[seq];
      int64_t p = start;
      origin = seq_off[seq] + start;
      while (static_cast<int>(read.size()) < len) {
        if (p >= slen) { read.push_back(kACGT[r.below(4)]); continue; }
        if (r.uniform() < sp->indel_rate) {
          int k = 1;
          while (k < sp->indel_max && r.uniform() >= sp->indel_p) ++k;
          if (r.below(2)) {
            for (int j = 0; j < k && static_cast<int>(read.size()) < len; ++j) read.push_back(kACGT[r.below(4)]);
          } else {
            p += k;
          }
          continue;
        }
        char b = src[p++];
        if (r.uniform() < sp->sub_rate) b = other_base(r, b);
        read.push_back(b);
      }
      if (r.uniform() < sp->rc_frac) {
        std::reverse(read.begin(), read.end());
        for (auto& b : read) b = comp(b);
        strand = 1;
      }
    }
    o->bases.insert(o->bases.end(), read.begin(), read.end());
    o->offsets.push_back(static_cast<int64_t>(o->bases.size()));
    o->origin.push_back(origin);
    o->strand.push_back(strand);
  }
  return o;
}

int64_t clast_gen_reads_total(void* h) { return static_cast<int64_t>(static_cast<clast_gen_reads*>(h)->bases.size()); }
void clast_gen_reads_copy(void* h, char* bases, int64_t* offsets, int64_t* origin, int8_t* strand) {
  auto* o = static_cast<clast_gen_reads*>(h);
  std::memcpy(bases, o->bases.data(), o->bases.size());
  std::memcpy(offsets, o->offsets.data(), o->offsets.size() * sizeof(int64_t));
  if (origin) std::memcpy(origin, o->origin.data(), o->origin.size() * sizeof(int64_t));
  if (strand) std::memcpy(strand, o->strand.data(), o->strand.size());
}
void clast_gen_reads_free(void* h) { delete static_cast<clast_gen_reads*>(h); }

}  // extern "C"
