// vfb_kernels.hip -- gfx950 kernels of the VFB bottom-up cached search.
//
// Reference: processOperator<op> (full-cache.cu:333-372) computes one
// candidate per thread from two AoS cache rows, fingerprints it
// (makeUnqChk, :126-134), inserts the fingerprint into a warpcore HashSet
// (64-bit atomicCAS, hash_set.cuh:193-254) and writes the row, a per-sample
// bool mask and the child indices to a temp cache that thrust compacts.
//
// Here a level is three passes over its candidate index space, which keeps
// the representative of each new fingerprint deterministic (the first
// candidate in enumeration order, where the reference takes whichever
// thread wins the CAS race):
//   insert: signature -> fingerprint -> open-addressing insert (64-bit CAS on
//           the key, atomicMin of (level << 48 | candidate index) on the
//           slot's companion word, so keys of earlier levels always win); the
//           candidate's slot is stored, and candidates equal to the target
//           record their index;
//   flag:   keep = (companion == level << 48 | own index); per-chunk counts;
//   write:  prefix over chunks + in-chunk prefix -> row position; kept
//           candidates recompute their signature and append it in order.
// Rows are stored SoA, rows[sample * cap + row], so the lanes of a wave that
// walk consecutive right rows read consecutive words of every sample.
#include <type_traits>

#include "vfb_kernels.h"

namespace mba {

// --- fingerprint (full-cache.cu:113-134) -----------------------------------
__device__ __forceinline__ uint32_t fmix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x85ebca6bu;
  x ^= x >> 13;
  x *= 0xc2b2ae35u;
  x ^= x >> 16;
  return x;
}

__device__ __forceinline__ uint64_t fmix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

// One step of makeUnqChk.  For 32-bit words the reference adds
// simpleHash(x) + 0x9e3779b9 in 32-bit unsigned arithmetic (both operands
// are unsigned int) before widening; the 64-bit extension uses the 64-bit
// MurmurHash3 finalizer and a 64-bit sum.
template <typename word>
__device__ __forceinline__ uint64_t fp_step(uint64_t seed, word x, int i) {
  uint64_t v;
  if (sizeof(word) == 4) {
    const uint32_t h = fmix32((uint32_t)x) + 0x9e3779b9u;
    v = (uint64_t)h + (seed << 6) + (seed >> 2);
  } else {
    v = fmix64((uint64_t)x) + 0x9e3779b9ull + (seed << 6) + (seed >> 2);
  }
  if (i & 1) v = (v >> 32) | (v << 32);
  return seed ^ v;
}

// --- candidate decoding ----------------------------------------------------
struct Cand {
  uint32_t l, r;
  int op;
};

__device__ __forceinline__ Cand decode(const VfbBlock *blk, int nblocks, uint64_t idx) {
  int b = 0;
  while (b + 1 < nblocks && idx >= blk[b + 1].begin) ++b;
  const VfbBlock &B = blk[b];
  uint64_t local = idx - B.begin;
  Cand c;
  c.op = B.op;
  if (is_unary(B.op)) {
    c.l = B.l0 + (uint32_t)local;
    c.r = c.l;
    return c;
  }
  int sw = 0;
  if (B.op == SUB) {
    sw = (int)(local & 1);
    local >>= 1;
  }
  const uint64_t q = local / B.nr;
  c.l = B.l0 + (uint32_t)q;
  c.r = B.r0 + (uint32_t)(local - q * B.nr);
  if (sw) {
    const uint32_t t = c.l;
    c.l = c.r;
    c.r = t;
  }
  return c;
}

// fp_step at a sample index of known parity (odd: the rotation).
template <typename word, int ODD>
__device__ __forceinline__ uint64_t fp_step_p(uint64_t seed, word x) {
  uint64_t v;
  if (sizeof(word) == 4) {
    const uint32_t h = fmix32((uint32_t)x) + 0x9e3779b9u;
    v = (uint64_t)h + (seed << 6) + (seed >> 2);
  } else {
    v = fmix64((uint64_t)x) + 0x9e3779b9ull + (seed << 6) + (seed >> 2);
  }
  if (ODD) v = (v >> 32) | (v << 32);
  return seed ^ v;
}

__device__ __forceinline__ uint64_t probe_start(uint64_t key, uint64_t mask) {
  return (key ^ (key >> 29)) & mask;
}

// Open addressing with linear probing, continued from slot s (its `step`-th
// probe); returns the slot of `key`, or ~0 when the probe sequence reaches
// `limit` (table too full: the host reports a capacity overflow), so every
// lane terminates.  Keys only ever change from EMPTY to a key, so a plain
// (cached, possibly stale) load is enough: a stale EMPTY is caught by the
// compare-and-swap, which returns the key that is there.  `inserted` is set
// when this call filled an empty slot (the key is new).
__device__ __forceinline__ uint64_t table_insert(uint64_t *keys, uint64_t mask,
                                                 uint64_t key, uint64_t s, uint64_t step,
                                                 uint64_t &probes, uint64_t limit,
                                                 bool &inserted, uint64_t &cas) {
  for (;; ++step) {
    if (step >= limit) return ~0ull;
    ++probes;
    const uint64_t k = keys[2 * s];
    if (k == key) return s;
    if (k == kVfbEmpty) {
      ++cas;
      const uint64_t old = atomicCAS((unsigned long long *)&keys[2 * s],
                                     (unsigned long long)kVfbEmpty,
                                     (unsigned long long)key);
      if (old == kVfbEmpty) {
        inserted = true;
        return s;
      }
      if (old == key) return s;
    }
    s = (s + 1) & mask;
  }
}

__device__ __forceinline__ uint64_t table_find(const uint64_t *keys, uint64_t mask,
                                               uint64_t key) {
  uint64_t s = probe_start(key, mask);
  for (uint64_t step = 0; step <= mask; ++step) {
    const uint64_t k = keys[2 * s];
    if (k == key) return s;
    if (k == kVfbEmpty) return ~0ull;
    s = (s + 1) & mask;
  }
  return ~0ull;
}

struct InsertStats {
  uint64_t probes = 0, fresh = 0, cas = 0, amin = 0;
};

// Inserts the N fingerprints of one lane's candidates (one row pair).  The
// first probes of all of them are issued together, then the compare-and-
// swaps into the slots that read EMPTY, so a lane waits on memory about
// twice per pair rather than twice per candidate; the rest of a probe
// sequence (an occupied first slot, mostly in the same cache line) follows
// per candidate.  Then the companions (first occurrence: the smallest level
// tag | index) and the candidates' slots.  An invalid fingerprint (EMPTY or
// TOMBSTONE) is dropped (hash_set.cuh:200-204).
template <int N>
__device__ __forceinline__ void insert_group(const VfbLevelArgs &a, const uint64_t (&idx)[N],
                                             const uint64_t (&fp)[N], uint64_t tfp,
                                             InsertStats &is) {
  uint64_t s[N], k[N];
#pragma unroll
  for (int c = 0; c < N; ++c) {
    s[c] = probe_start(fp[c], a.mask);
    k[c] = fp[c] < kVfbTomb ? a.keys[2 * s[c]] : 0;
  }
#pragma unroll
  for (int c = 0; c < N; ++c) {
    if (fp[c] < kVfbTomb && k[c] == kVfbEmpty) {
      ++is.cas;
      const uint64_t old = atomicCAS((unsigned long long *)&a.keys[2 * s[c]],
                                     (unsigned long long)kVfbEmpty, (unsigned long long)fp[c]);
      // k = the key now in the slot; EMPTY marks this lane's insert
      k[c] = old == kVfbEmpty ? kVfbEmpty - 1 : old;
      if (old == kVfbEmpty) ++is.fresh;
    }
  }
#pragma unroll
  for (int c = 0; c < N; ++c) {
    if (fp[c] >= kVfbTomb) {
      if (!a.last) a.slot_of[idx[c]] = ~0u;
      continue;
    }
    ++is.probes;
    uint64_t sl = s[c];
    if (k[c] != kVfbEmpty - 1 && k[c] != fp[c]) {  // occupied by another key
      bool ins = false;
      sl = table_insert(a.keys, a.mask, fp[c], (s[c] + 1) & a.mask, 1, is.probes, a.max_probes,
                        ins, is.cas);
      is.fresh += ins ? 1u : 0u;
      if (sl == ~0ull) {
        atomicOr(a.fail, 1u);
        if (!a.last) a.slot_of[idx[c]] = ~0u;
        continue;
      }
    }
    if ((!a.last || fp[c] == tfp) && a.minidx[2 * sl] > (a.tag | idx[c])) {
      ++is.amin;
      atomicMin((unsigned long long *)&a.minidx[2 * sl], (unsigned long long)(a.tag | idx[c]));
    }
    if (!a.last) a.slot_of[idx[c]] = (uint32_t)sl;
  }
}

// Whether op(l, r) equals the target on every sample (a candidate whose
// fingerprint is the target's).
template <typename word>
__device__ __noinline__ bool equals_target(const word *rows, uint64_t cap, int n, int op,
                                           uint32_t l, uint32_t r, const word *tgt) {
  for (int i = 0; i < n; ++i)
    if (apply_op<word>(op, rows[(uint64_t)i * cap + l], rows[(uint64_t)i * cap + r]) != tgt[i])
      return false;
  return true;
}

// Pair-major form of the insert pass: one lane takes one (left, right) row
// pair of a split (or one row for the unary ops) and produces every
// candidate built on it -- & | ^ + * and both orders of '-' (or ~ and neg) --
// so each pair of rows is read once for up to seven candidates.  Candidate
// indices are those of the enumeration order (opbegin[op] + local index).
// Within a split the pairs are taken in bands of kVfbBand right rows, every
// left row of a band before the next band, so a band's right rows are read
// by neighbouring lanes and workgroups close together in time (cache hits)
// rather than once per left row across the whole right level; the
// candidate indices stay those of the left-major order.  Consecutive
// chunks of pairs go to workgroups of one XCD (workgroup w runs on XCD
// w mod 8), whose L2 then holds the band.
// The last level is hash-only: no rows are stored, so only the target's
// fingerprint needs its first occurrence (the companion) and the level's new
// fingerprints are counted as the empty slots this pass fills.
template <typename word>
__global__ void __launch_bounds__(kVfbThreads)
k_vfb_insert_pairs(VfbLevelArgs a, const VfbPairBlock *pblocks, int npb, uint64_t npairs) {
  __shared__ VfbPairBlock s_pb[kVfbMaxPairBlocks];
  __shared__ word s_tgt[kMaxSamples];
  __shared__ uint64_t s_tfp;
  for (int i = threadIdx.x; i < npb; i += blockDim.x) s_pb[i] = pblocks[i];
  for (int i = threadIdx.x; i < a.n; i += blockDim.x) s_tgt[i] = ((const word *)a.target)[i];
  __syncthreads();
  if (threadIdx.x == 0) {  // the target's fingerprint
    uint64_t t = 0;
    for (int i = 0; i < a.n; ++i) t = fp_step<word>(t, s_tgt[i], i);
    s_tfp = t;
  }
  __syncthreads();
  const uint64_t tfp = s_tfp;
  const word *rows = (const word *)a.rows;
  const int n = a.n;
  InsertStats is;
  const uint32_t G = gridDim.x;
  const uint32_t wg = G % 8 == 0 ? (blockIdx.x % 8) * (G / 8) + blockIdx.x / 8 : blockIdx.x;
  const uint64_t stride = (uint64_t)G * blockDim.x;
  for (uint64_t p = (uint64_t)wg * blockDim.x + threadIdx.x; p < npairs; p += stride) {
    int b = 0;
    while (b + 1 < npb && p >= s_pb[b + 1].pbegin) ++b;
    const VfbPairBlock &B = s_pb[b];
    const uint64_t local = p - B.pbegin;
    if (B.unary) {
      const uint32_t l = B.l0 + (uint32_t)local;
      uint64_t sn = 0, sg = 0;
      int i = 0;
      for (; i + 1 < n; i += 2) {
        const word x0 = rows[(uint64_t)i * a.cap + l], x1 = rows[(uint64_t)(i + 1) * a.cap + l];
        sn = fp_step_p<word, 0>(sn, word(~x0));
        sg = fp_step_p<word, 0>(sg, word(word(0) - x0));
        sn = fp_step_p<word, 1>(sn, word(~x1));
        sg = fp_step_p<word, 1>(sg, word(word(0) - x1));
      }
      if (i < n) {
        const word x0 = rows[(uint64_t)i * a.cap + l];
        sn = fp_step_p<word, 0>(sn, word(~x0));
        sg = fp_step_p<word, 0>(sg, word(word(0) - x0));
      }
      const uint64_t idx[2] = {B.opbegin[NOT] + local, B.opbegin[NEG] + local};
      const uint64_t fp[2] = {sn, sg};
      insert_group<2>(a, idx, fp, tfp, is);
      if (sn == tfp && equals_target<word>(rows, a.cap, n, NOT, l, l, s_tgt))
        atomicMin(a.found, (unsigned long long)idx[0]);
      if (sg == tfp && equals_target<word>(rows, a.cap, n, NEG, l, l, s_tgt))
        atomicMin(a.found, (unsigned long long)idx[1]);
      continue;
    }
    // band of kVfbBand right rows (the last band may be narrower)
    const uint64_t bw = (uint64_t)B.nl * kVfbBand;
    const uint64_t band = local / bw;
    const uint64_t in_band = local - band * bw;
    const uint64_t r0 = band * kVfbBand;
    const uint64_t w = B.nr - r0 < (uint64_t)kVfbBand ? B.nr - r0 : (uint64_t)kVfbBand;
    const uint64_t q = in_band / w;
    const uint32_t l = B.l0 + (uint32_t)q, r = B.r0 + (uint32_t)(r0 + (in_band - q * w));
    const uint64_t lr = q * B.nr + r0 + (in_band - q * w);  // left-major pair index
    uint64_t sd[7] = {0, 0, 0, 0, 0, 0, 0};
    auto sample = [&](auto odd, int i) {
      constexpr int O = decltype(odd)::value;
      const word x = rows[(uint64_t)i * a.cap + l];
      const word y = rows[(uint64_t)i * a.cap + r];
      sd[0] = fp_step_p<word, O>(sd[0], word(x & y));
      sd[1] = fp_step_p<word, O>(sd[1], word(x | y));
      sd[2] = fp_step_p<word, O>(sd[2], word(x ^ y));
      sd[3] = fp_step_p<word, O>(sd[3], word(x + y));
      sd[4] = fp_step_p<word, O>(sd[4], word(x * y));
      sd[5] = fp_step_p<word, O>(sd[5], word(x - y));
      sd[6] = fp_step_p<word, O>(sd[6], word(y - x));
    };
    int i = 0;
    for (; i + 1 < n; i += 2) {
      sample(std::integral_constant<int, 0>(), i);
      sample(std::integral_constant<int, 1>(), i + 1);
    }
    if (i < n) sample(std::integral_constant<int, 0>(), i);
    const uint64_t idx[7] = {B.opbegin[AND] + lr, B.opbegin[OR] + lr, B.opbegin[XOR] + lr,
                             B.opbegin[ADD] + lr, B.opbegin[MUL] + lr, B.opbegin[SUB] + 2 * lr,
                             B.opbegin[SUB] + 2 * lr + 1};
    insert_group<7>(a, idx, sd, tfp, is);
    static constexpr int kOps[7] = {AND, OR, XOR, ADD, MUL, SUB, SUB};
#pragma unroll
    for (int c = 0; c < 7; ++c)
      if (sd[c] == tfp &&
          equals_target<word>(rows, a.cap, n, kOps[c], c == 6 ? r : l, c == 6 ? l : r, s_tgt))
        atomicMin(a.found, (unsigned long long)idx[c]);
  }
  if (a.probes) atomicAdd(a.probes, (unsigned long long)is.probes);
  if (is.fresh) atomicAdd(a.new_keys, (unsigned long long)is.fresh);
  if (a.atomics && (is.cas | is.amin)) {
    atomicAdd(a.atomics, (unsigned long long)is.cas);
    atomicAdd(a.atomics + 1, (unsigned long long)is.amin);
  }
}

// keep = first occurrence in this level of a fingerprint unseen before.
__global__ void __launch_bounds__(kVfbThreads) k_vfb_flag(VfbLevelArgs a) {
  __shared__ uint32_t s_cnt;
  const uint64_t nchunks = (a.total + kVfbChunk - 1) / kVfbChunk;
  for (uint64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    uint32_t mine = 0;
    for (int k = 0; k < kVfbChunk / kVfbThreads; ++k) {
      const uint64_t idx = ch * kVfbChunk + (uint64_t)k * kVfbThreads + threadIdx.x;
      if (idx >= a.total) break;
      const uint32_t s = a.slot_of[idx];
      const uint8_t kp = (s != ~0u) && (a.minidx[2 * (uint64_t)s] == (a.tag | idx));
      a.keep[idx] = kp;
      mine += kp;
    }
    // wave reduction, one LDS atomic per wave
    for (int off = 32; off > 0; off >>= 1) mine += __shfl_down(mine, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&s_cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) a.chunk_count[ch] = s_cnt;
    __syncthreads();
  }
}

// Exclusive prefix of the chunk counts (single workgroup).
__global__ void __launch_bounds__(1024) k_vfb_scan(const uint32_t *cnt, uint64_t *base,
                                                   uint64_t n) {
  __shared__ uint64_t s_part[1024];
  const uint64_t per = (n + 1023) / 1024;
  const uint64_t b0 = threadIdx.x * per, b1 = b0 + per < n ? b0 + per : n;
  uint64_t sum = 0;
  for (uint64_t i = b0; i < b1; ++i) sum += cnt[i];
  s_part[threadIdx.x] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t acc = 0;
    for (int i = 0; i < 1024; ++i) {
      const uint64_t v = s_part[i];
      s_part[i] = acc;
      acc += v;
    }
    base[n] = acc;
  }
  __syncthreads();
  uint64_t acc = s_part[threadIdx.x];
  for (uint64_t i = b0; i < b1; ++i) {
    base[i] = acc;
    acc += cnt[i];
  }
}

template <typename word>
__global__ void __launch_bounds__(kVfbThreads) k_vfb_write(VfbLevelArgs a) {
  __shared__ VfbBlock s_blk[kVfbMaxBlocks];
  __shared__ uint32_t s_wave[kVfbThreads / 64];
  for (int i = threadIdx.x; i < a.nblocks; i += blockDim.x) s_blk[i] = a.blocks[i];
  __syncthreads();
  const word *rows = (const word *)a.rows;
  word *out = (word *)a.out_rows;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t nchunks = (a.total + kVfbChunk - 1) / kVfbChunk;
  for (uint64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    uint64_t pos = a.out_base + a.chunk_base[ch];
    for (int k = 0; k < kVfbChunk / kVfbThreads; ++k) {
      const uint64_t idx = ch * kVfbChunk + (uint64_t)k * kVfbThreads + threadIdx.x;
      const bool kp = idx < a.total && a.keep[idx];
      // prefix of kept flags in index order across the 4 waves of this step
      const uint64_t bal = __ballot(kp);
      const uint32_t below = __popcll(bal & ((1ull << lane) - 1));
      if (lane == 0) s_wave[wave] = __popcll(bal);
      __syncthreads();
      uint32_t wbase = 0, tot = 0;
      for (int w = 0; w < kVfbThreads / 64; ++w) {
        if (w < wave) wbase += s_wave[w];
        tot += s_wave[w];
      }
      if (kp) {
        const Cand c = decode(s_blk, a.nblocks, idx);
        const uint64_t row = pos + wbase + below;
        for (int i = 0; i < a.n; ++i) {
          const word x = rows[(uint64_t)i * a.cap + c.l];
          const word y = rows[(uint64_t)i * a.cap + c.r];
          out[(uint64_t)i * a.cap + row] = apply_op<word>(c.op, x, y);
        }
        a.left[row] = (int32_t)c.l;
        a.right[row] = is_unary(c.op) ? -1 : (int32_t)c.r;
        a.row_op[row] = (uint8_t)c.op;
      }
      pos += tot;
      __syncthreads();
    }
  }
}

template <typename word>
__global__ void k_vfb_seed(VfbLevelArgs a, int L) {
  const word *rows = (const word *)a.rows;
  const int v = threadIdx.x;
  if (v >= L) return;
  uint64_t seed = 0;
  for (int i = 0; i < a.n; ++i) seed = fp_step<word>(seed, rows[(uint64_t)i * a.cap + v], i);
  if (seed >= kVfbTomb) return;
  uint64_t probes = 0, cas = 0;
  bool ins = false;
  const uint64_t s = table_insert(a.keys, a.mask, seed, probe_start(seed, a.mask), 0, probes,
                                  a.max_probes, ins, cas);
  if (s == ~0ull) {
    atomicOr(a.fail, 1u);
    return;
  }
  a.minidx[2 * s] = 0;  // present before any level >= 2
}

template <typename word>
__global__ void k_vfb_lookup_target(VfbLevelArgs a, uint64_t *out) {
  if (threadIdx.x != 0) return;
  const word *t = (const word *)a.target;
  uint64_t seed = 0;
  for (int i = 0; i < a.n; ++i) seed = fp_step<word>(seed, t[i], i);
  const uint64_t s = seed >= kVfbTomb ? ~0ull : table_find(a.keys, a.mask, seed);
  *out = s == ~0ull ? ~0ull : a.minidx[2 * s];
}

__global__ void k_vfb_rpn(const int32_t *left, const int32_t *right, const uint8_t *op,
                          int L, int32_t root_op, int32_t root_l, int32_t root_r,
                          uint8_t *out, int32_t *out_len, int cap) {
  if (threadIdx.x != 0) return;
  // iterative post-order: (row, state); the root is the unsaved candidate
  int32_t st_row[256];
  uint8_t st_state[256];
  int sp = 0, n = 0;
  auto push_child = [&](int32_t row) {
    if (sp < 256) {
      st_row[sp] = row;
      st_state[sp] = 0;
      ++sp;
    }
  };
  const bool un = is_unary(root_op);
  // children of the root, left then right, then the root op itself
  int32_t kids[2] = {root_l, root_r};
  for (int k = 0; k < (un ? 1 : 2); ++k) {
    push_child(kids[k]);
    while (sp > 0) {
      const int32_t row = st_row[sp - 1];
      if (row < L) {
        if (n < cap) out[n] = (uint8_t)row;
        ++n;
        --sp;
        continue;
      }
      const int o = op[row];
      const uint8_t s = st_state[sp - 1];
      if (s == 0) {
        st_state[sp - 1] = 1;
        push_child(left[row]);
      } else if (s == 1 && !is_unary(o)) {
        st_state[sp - 1] = 2;
        push_child(right[row]);
      } else {
        if (n < cap) out[n] = (uint8_t)(kTokOp + o);
        ++n;
        --sp;
      }
    }
  }
  if (n < cap) out[n] = (uint8_t)(kTokOp + root_op);
  ++n;
  *out_len = n;
}

// Fingerprints of `count` signatures of n words (vals [count][n]), one lane
// per signature, by the fp_step of the insert pass.
template <typename word>
__global__ void __launch_bounds__(256)
k_vfb_fingerprints(const uint64_t *__restrict__ vals, int n, uint64_t count,
                   uint64_t *__restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < count; r += stride) {
    uint64_t seed = 0;
    for (int i = 0; i < n; ++i) seed = fp_step<word>(seed, (word)vals[r * n + i], i);
    out[r] = seed;
  }
}

// --- random-access ceiling of the insert pass -------------------------------
// Keys of the measurement: keys[i] = SplitMix64(SplitMix64(i + salt) mod
// distinct), i < count, kept below TOMBSTONE (Steele, Lea and Flood, "Fast
// splittable pseudorandom number generators", OOPSLA 2014).
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

__global__ void k_vfb_ceiling_keys(uint64_t *keys, uint64_t count, uint64_t distinct,
                                   uint64_t salt) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const uint64_t k = splitmix64(splitmix64(i + salt) % distinct);
    keys[i] = k >= kVfbTomb ? k - 2 : k;
  }
}

// One probe load per key (the insert pass's first probe, no write); the
// keys that read an occupied slot are counted so every load is used.
__global__ void __launch_bounds__(kVfbThreads)
k_vfb_ceiling_probe(const uint64_t *__restrict__ keys, uint64_t count,
                    const uint64_t *__restrict__ slots, uint64_t mask,
                    unsigned long long *occupied) {
  constexpr int N = 7;
  uint64_t occ = 0;
  const uint64_t groups = (count + N - 1) / N;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += stride) {
    uint64_t k[N];
#pragma unroll
    for (int c = 0; c < N; ++c) {
      const uint64_t i = g * N + c;
      k[c] = i < count ? slots[2 * probe_start(keys[i], mask)] : kVfbEmpty;
    }
#pragma unroll
    for (int c = 0; c < N; ++c) occ += k[c] != kVfbEmpty;
  }
  if (occ) atomicAdd(occupied, (unsigned long long)occ);
}

// The insert pass's table code (insert_group, hash-only form) on ready-made
// keys, seven per lane as in the pass; no fingerprint work.
__global__ void __launch_bounds__(kVfbThreads)
k_vfb_ceiling_insert(VfbLevelArgs a, const uint64_t *__restrict__ keys, uint64_t count) {
  constexpr int N = 7;
  InsertStats is;
  const uint64_t groups = (count + N - 1) / N;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += stride) {
    uint64_t idx[N], fp[N];
#pragma unroll
    for (int c = 0; c < N; ++c) {
      idx[c] = g * N + c;
      fp[c] = idx[c] < count ? keys[idx[c]] : kVfbTomb;  // TOMBSTONE: dropped
    }
    insert_group<N>(a, idx, fp, kVfbEmpty, is);
  }
  if (is.fresh) atomicAdd(a.new_keys, (unsigned long long)is.fresh);
}

hipError_t vfb_launch_ceiling(int kind, const VfbLevelArgs &a, uint64_t *keys, uint64_t count,
                              uint64_t distinct, uint64_t salt, unsigned long long *counter,
                              int grid, hipStream_t s) {
  if (count == 0) return hipSuccess;
  if (kind == 0)
    hipLaunchKernelGGL(k_vfb_ceiling_keys, dim3(grid), dim3(256), 0, s, keys, count, distinct,
                       salt);
  else if (kind == 1)
    hipLaunchKernelGGL(k_vfb_ceiling_probe, dim3(grid), dim3(kVfbThreads), 0, s, keys, count,
                       a.keys, a.mask, counter);
  else
    hipLaunchKernelGGL(k_vfb_ceiling_insert, dim3(grid), dim3(kVfbThreads), 0, s, a, keys,
                       count);
  return hipGetLastError();
}

__global__ void k_fill_u64(uint64_t *p, uint64_t n, uint64_t v) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = v;
}

// --- launchers -------------------------------------------------------------
hipError_t vfb_launch_insert_pairs(int width, const VfbLevelArgs &a, const VfbPairBlock *d_pb,
                                   int npb, uint64_t npairs, int grid, hipStream_t s) {
  if (width == 32)
    hipLaunchKernelGGL(k_vfb_insert_pairs<uint32_t>, dim3(grid), dim3(kVfbThreads), 0, s, a,
                       d_pb, npb, npairs);
  else
    hipLaunchKernelGGL(k_vfb_insert_pairs<uint64_t>, dim3(grid), dim3(kVfbThreads), 0, s, a,
                       d_pb, npb, npairs);
  return hipGetLastError();
}

hipError_t vfb_launch_flag(int, const VfbLevelArgs &a, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_vfb_flag, dim3(grid), dim3(kVfbThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t vfb_launch_scan(const VfbLevelArgs &a, uint64_t nchunks, hipStream_t s) {
  hipLaunchKernelGGL(k_vfb_scan, dim3(1), dim3(1024), 0, s, a.chunk_count, a.chunk_base,
                     nchunks);
  return hipGetLastError();
}

hipError_t vfb_launch_write(int width, const VfbLevelArgs &a, int grid, hipStream_t s) {
  if (width == 32)
    hipLaunchKernelGGL(k_vfb_write<uint32_t>, dim3(grid), dim3(kVfbThreads), 0, s, a);
  else
    hipLaunchKernelGGL(k_vfb_write<uint64_t>, dim3(grid), dim3(kVfbThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t vfb_launch_seed(int width, const VfbLevelArgs &a, int L, hipStream_t s) {
  if (width == 32)
    hipLaunchKernelGGL(k_vfb_seed<uint32_t>, dim3(1), dim3(64), 0, s, a, L);
  else
    hipLaunchKernelGGL(k_vfb_seed<uint64_t>, dim3(1), dim3(64), 0, s, a, L);
  return hipGetLastError();
}

hipError_t vfb_launch_lookup_target(int width, const VfbLevelArgs &a, uint64_t *out,
                                    hipStream_t s) {
  if (width == 32)
    hipLaunchKernelGGL(k_vfb_lookup_target<uint32_t>, dim3(1), dim3(64), 0, s, a, out);
  else
    hipLaunchKernelGGL(k_vfb_lookup_target<uint64_t>, dim3(1), dim3(64), 0, s, a, out);
  return hipGetLastError();
}

hipError_t vfb_launch_rpn(const int32_t *left, const int32_t *right, const uint8_t *op,
                          int L, int32_t root_op, int32_t root_l, int32_t root_r,
                          uint8_t *out, int32_t *out_len, int cap, hipStream_t s) {
  hipLaunchKernelGGL(k_vfb_rpn, dim3(1), dim3(64), 0, s, left, right, op, L, root_op,
                     root_l, root_r, out, out_len, cap);
  return hipGetLastError();
}

hipError_t vfb_launch_fingerprints(int width, const uint64_t *d_vals, int n, uint64_t count,
                                   uint64_t *d_out, hipStream_t s) {
  if (count == 0) return hipSuccess;
  const uint64_t blocks = (count + 255) / 256;
  const dim3 grid((unsigned)(blocks < 4096 ? blocks : 4096));
  if (width == 32)
    hipLaunchKernelGGL(k_vfb_fingerprints<uint32_t>, grid, dim3(256), 0, s, d_vals, n, count,
                       d_out);
  else
    hipLaunchKernelGGL(k_vfb_fingerprints<uint64_t>, grid, dim3(256), 0, s, d_vals, n, count,
                       d_out);
  return hipGetLastError();
}

hipError_t vfb_fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  const int grid = (int)(blocks < 16384 ? blocks : 16384);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace mba
