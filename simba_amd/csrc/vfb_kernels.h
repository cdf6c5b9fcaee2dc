// vfb_kernels.h -- launch interface of the VFB (bottom-up, cached) kernels.
#pragma once
#include "mba_common.h"

namespace mba {

constexpr uint64_t kVfbEmpty = ~0ull;        // empty hash slot (full-cache.cu:738)
constexpr uint64_t kVfbTomb = ~0ull - 1;     // tombstone key (full-cache.cu:739)
constexpr int kVfbMaxBlocks = 96;
constexpr int kVfbChunk = 4096;              // candidates per workgroup step
constexpr int kVfbThreads = 256;
constexpr int kVfbBand = 256;               // right rows per band of a split (insert pass)

// One (op, split[, swap]) block of a level's candidate space.  Candidate
// indices of the level are the concatenation of the blocks in enumeration
// order (full-cache.cu:568-663): unary blocks over the rows of level L-1,
// binary blocks over left rows (outer) x right rows (inner); for '-' every
// pair gives (l, r) at local index 2p and (r, l) at 2p + 1.
struct VfbBlock {
  uint64_t begin;  // first candidate index of the block
  uint64_t count;  // candidates in the block
  uint32_t l0, nl; // left rows [l0, l0+nl)
  uint32_t r0, nr; // right rows (binary)
  int32_t op;
  int32_t pad;
};

struct VfbLevelArgs {
  const void *rows;       // SoA row store: rows[sample * cap + row]
  uint64_t cap;           // row capacity
  int n;                  // samples
  int nblocks;
  const VfbBlock *blocks; // device copy
  uint64_t total;         // candidates of the level
  uint64_t *keys;         // hash slots, 16 B each: key (fingerprint) at [2s]
  uint64_t *minidx;       // = keys + 1: companion of slot s at minidx[2s],
                          // (level << 48) | min candidate index; a
                          // key of an earlier level keeps a smaller value
  uint64_t tag;           // level << 48 of the level being processed
  uint64_t mask;          // table capacity - 1 (power of two)
  uint32_t *slot_of;      // per candidate: slot of its fingerprint (or ~0u: invalid)
  uint8_t *keep;          // per candidate: kept (first occurrence of a new fingerprint)
  uint32_t *chunk_count;  // per chunk of kVfbChunk candidates: kept count
  uint64_t *chunk_base;   // exclusive prefix of chunk_count
  const void *target;     // n target words
  unsigned long long *found;  // min index of a candidate equal to the target
  unsigned int *fail;     // set when a probe sequence exceeds max_probes
  uint64_t max_probes;
  void *out_rows;         // rows of this level start at row `out_base`
  uint64_t out_base;
  int32_t *left, *right;  // child rows per row
  uint8_t *row_op;        // root op per row
  unsigned long long *probes;  // total probe steps (statistics)
  unsigned long long *new_keys;  // fingerprints this pass inserted (new keys)
  unsigned long long *atomics;   // [0] key CAS attempts, [1] companion atomicMin
  int last;               // hash-only level: no slot_of, companions of the
                          // target's fingerprint only
  int pad;
};

// The row pairs of a level: the unary block (rows of level L-1, ops ~ and
// neg) and one block per split i (left rows of level i x right rows of level
// L-1-i, every binary op).  opbegin[op] = first candidate index of that op's
// block for this split (the order of full-cache.cu:568-663).
constexpr int kVfbMaxPairBlocks = 24;
struct VfbPairBlock {
  uint64_t pbegin;       // first pair index of the block
  uint64_t opbegin[8];
  uint32_t l0, nl, r0, nr;
  int32_t unary;
  int32_t pad;
};

hipError_t vfb_launch_insert_pairs(int width, const VfbLevelArgs &a, const VfbPairBlock *d_pb,
                                   int npb, uint64_t npairs, int grid, hipStream_t s);
hipError_t vfb_launch_flag(int width, const VfbLevelArgs &a, int grid, hipStream_t s);
hipError_t vfb_launch_scan(const VfbLevelArgs &a, uint64_t nchunks, hipStream_t s);
hipError_t vfb_launch_write(int width, const VfbLevelArgs &a, int grid, hipStream_t s);
// Leaves: fingerprint and insert rows [0, L) (hashSetInit, full-cache.cu:142-157).
hipError_t vfb_launch_seed(int width, const VfbLevelArgs &a, int L, hipStream_t s);
// Writes, for the fingerprint of the target, the slot's minidx into *out
// (~0 if absent).
hipError_t vfb_launch_lookup_target(int width, const VfbLevelArgs &a, uint64_t *out,
                                    hipStream_t s);
// Post-order walk of the solution tree into RPN tokens (at most `cap`).
hipError_t vfb_launch_rpn(const int32_t *left, const int32_t *right, const uint8_t *op,
                          int L, int32_t root_op, int32_t root_l, int32_t root_r,
                          uint8_t *out, int32_t *out_len, int cap, hipStream_t s);
hipError_t vfb_fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t s);
// Random-access ceiling of the insert pass: kind 0 writes `count` keys with
// `distinct` values (salted) to `keys`; kind 1 runs one probe load per key
// against a.keys (occupied slots counted in *counter); kind 2 inserts the
// keys with the insert pass's table code (hash-only, fresh keys counted in
// *a.new_keys).
hipError_t vfb_launch_ceiling(int kind, const VfbLevelArgs &a, uint64_t *keys, uint64_t count,
                              uint64_t distinct, uint64_t salt, unsigned long long *counter,
                              int grid, hipStream_t s);
// out[r] = fingerprint of the n words vals[r * n .. r * n + n) (the insert
// pass's makeUnqChk), r < count.
hipError_t vfb_launch_fingerprints(int width, const uint64_t *d_vals, int n, uint64_t count,
                                   uint64_t *d_out, hipStream_t s);

}  // namespace mba
