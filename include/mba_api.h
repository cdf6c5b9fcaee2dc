/*
 * mba_api.h -- C ABI of the MI355X (gfx950) MBA synthesis engine.
 *
 * Drop-in for the search of the reference's two tools:
 *   - SIMBA plain search   NoCache::solve(const MbaInstance&, int max_size)
 *                          simba/src/nocache.cu:197-289, CLI main :291-342
 *   - VFB cached search    string MBA(int maxLen, int numExamples,
 *                                     uint32_t* in, uint32_t* out)
 *                          vfb/src/full-cache.cu:665-774, CLI main :776-853
 * and the instance readers MbaInstance::from_json_file
 * (simba/src/instance.h:19-29, instance.cpp:7-27) and readJsonFile
 * (vfb/src/utils.h:15-44).
 *
 * Plain pointers and sizes only.  Inputs are owned by the caller; results are
 * written into caller-provided structs.  No global mutable state: every
 * context owns its device buffers, so distinct contexts may be used from
 * distinct threads.  Return codes: 0 = solution found, 1 = not found within
 * the size range, 2 = time budget or cancel flag hit, 3 = cache capacity
 * exhausted (VFB), < 0 = error (see mba_strerror).
 */
#ifndef MBA_API_H
#define MBA_API_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBA_MAX_VARS 16      /* leaves = vars + consts <= 16            */
#define MBA_MAX_SAMPLES 128  /* reference: 100 (globals.cuh:6)          */
#define MBA_MAX_SIZE 32      /* reference: 32 (globals.cuh:7)           */
#define MBA_EXPR_CAP 1024

enum {
  MBA_FOUND = 0,
  MBA_NOT_FOUND = 1,
  MBA_BUDGET = 2,
  MBA_OVERFLOW = 3,
  MBA_ERR_ARG = -1,
  MBA_ERR_LIMIT = -2,   /* sizes/vars/samples beyond the supported limits */
  MBA_ERR_COUNT = -3,   /* count table overflows uint64 at this max size  */
  MBA_ERR_HIP = -4,     /* a HIP runtime call failed                       */
  MBA_ERR_NOMEM = -5,
  MBA_ERR_IO = -6,      /* instance file missing or malformed              */
  MBA_ERR_NODEV = -7    /* no gfx950 device / HIP code object not loadable */
};

/* One synthesis instance.  Layout of `inputs` is [sample][var], as the
 * reference (instance.cpp:14-24, utils.h:27-41).  Values are taken mod 2^width.
 * Constant leaves (an extension; empty = reference grammar) follow the
 * variables in leaf order. */
typedef struct {
  uint32_t width;      /* 32 (reference semantics) or 64 */
  uint32_t n_vars;
  uint32_t n_samples;
  uint32_t n_consts;
  const uint64_t *inputs;   /* n_samples * n_vars */
  const uint64_t *outputs;  /* n_samples          */
  const uint64_t *consts;   /* n_consts (may be NULL when 0) */
} mba_instance;

typedef struct {
  int32_t min_size;         /* inclusive; SIMBA native default 1            */
  int32_t max_size;         /* inclusive                                     */
  int32_t compat;           /* 1: reference ranges: SIMBA sizes 2..max-1,    */
                            /*    VFB levels 1..max (nocache.cu:246,         */
                            /*    full-cache.cu:754)                         */
  int32_t exhaustive;       /* keep enumerating up to max after a solution   */
                            /* (the answer is still the first solution)      */
  double time_budget_s;     /* <= 0: none                                    */
  volatile int32_t *cancel; /* host-written; polled between launches; NULL ok */
  uint64_t table_bytes;     /* SIMBA: value-table budget, 0 = default        */
  uint64_t cache_bytes;     /* VFB: row-store budget, 0 = default            */
  int32_t verbose;          /* progress on stderr                            */
  int32_t wg_per_cu;        /* launch tuning: workgroups per CU (0: 32       */
                            /* single-instance, 8 batched)                   */
  int32_t lanes_b;          /* launch tuning: B values per lane, 4 or 8 (0: 8) */
  int32_t table_delta;      /* tabulate Kt - table_delta sizes (0: budget)   */
  int32_t block_log2;       /* sweep one (op, split) block per launch from   */
                            /* 2^block_log2 trees per size (0: defaults)    */
  int32_t hit_cap;          /* window-hit buffer, in hits (0: per-instance   */
                            /* 2^22, batched 2^26; both above the hits of    */
                            /* one tile); a launch group (batched: a step)   */
                            /* that overflows it is replayed in parts, the   */
                            /* batched step by halves of its jobs, then of   */
                            /* one job's tiles (a smaller buffer that one    */
                            /* tile overflows is an error)                   */
  int32_t stop_size;        /* SIMBA: end the sweep after this size (0:     */
                            /* max_size); the unary-chain windows stay those */
                            /* of max_size, so sizes 1..stop_size are swept  */
                            /* exactly as in a search to max_size that is    */
                            /* interrupted there (benchmarks of one shape)   */
  int32_t reserved[1];
} mba_opts;

typedef struct {
  int32_t status;
  int32_t size;          /* size of the solution (SIMBA) / level (VFB)      */
  uint64_t rank;         /* SIMBA: rank within its size, in the reference's */
                         /* enumeration order; VFB: enumeration index       */
                         /* within the level                                */
  int32_t op;            /* root operator 0..7 = ~ & | ^ neg + - *, -1 leaf */
  int32_t split;         /* SIMBA: left child size (0: unary or leaf)       */
  char expr[MBA_EXPR_CAP];     /* reference printing conventions            */
  char rpn[MBA_MAX_SIZE + 1];  /* SIMBA: reference RPN chars ('0'+leaf, op) */
  double seconds;        /* search-only time (device work + host driver)    */
  double kernel_seconds; /* sum of the hot kernels' HIP-event times         */
  uint64_t candidates;   /* candidates of every completed size/level        */
  int32_t size_reached;  /* last size/level completely searched             */
  uint32_t kernel_launches; /* intervals summed in kernel_seconds (SIMBA:   */
                            /* one per pair-kernel launch; VFB: one per     */
                            /* level; batch: one per graph replay)          */
  int32_t hit_replays;      /* SIMBA per-instance: launch groups replayed    */
                            /* in parts after a hit-buffer overflow         */
  int32_t ordered_sizes;    /* learned search: sizes whose blocks were       */
                            /* visited in the network's order               */
  int32_t split_reached;    /* SIMBA exhaustive sweep stopped by a budget    */
                            /* inside size size_reached + 1: its splits      */
                            /* {j, s-1-j}, j = 1..split_reached, were fully  */
                            /* checked (0: none)                             */
  int32_t reserved;
  uint64_t split_pairs;     /* and of split split_reached + 1, the (A, B)    */
                            /* value pairs of the B blocks whose launches    */
                            /* had completed when the budget stopped it      */
} mba_result;

/* VFB per-level statistics (PRINT_CACHE_SIZE mode, full-cache.cu:763-765). */
typedef struct {
  uint64_t level_new[MBA_MAX_SIZE + 1];   /* rows kept at each level         */
  uint64_t level_cand[MBA_MAX_SIZE + 1];  /* candidates enumerated per level */
  double level_seconds[MBA_MAX_SIZE + 1];
  uint64_t level_bytes[MBA_MAX_SIZE + 1]; /* algorithmic HBM bytes per level */
  uint64_t table_slots;
  uint64_t rows_capacity;
  uint64_t level_cas[MBA_MAX_SIZE + 1];     /* 64-bit key CAS attempts      */
  uint64_t level_atomics[MBA_MAX_SIZE + 1]; /* CAS + companion atomicMin    */
} mba_vfb_stats;

typedef struct mba_ctx mba_ctx;

/* Context on HIP device `device`.  Fails with MBA_ERR_NODEV when the device
 * is absent or the gfx950 code object cannot be loaded. */
int mba_ctx_create(int device, mba_ctx **out);
void mba_ctx_destroy(mba_ctx *ctx);
const char *mba_strerror(int code);
/* Text of the last error a context recorded (diagnostics). */
const char *mba_ctx_last_error(mba_ctx *ctx);
void mba_default_opts(mba_opts *o);

/* SIMBA plain search (replaces NoCache::solve, nocache.cu:197-289).  First
 * solution = minimum rank within the smallest size that has one. */
int mba_simba_solve(mba_ctx *ctx, const mba_instance *inst, const mba_opts *opts,
                    mba_result *res);

/* SIMBA plain search that allocates no device memory.  No counterpart in
 * the reference: it is for callers that cannot allocate during a call
 * (serving loops, harnesses that forbid hipMalloc).  Same answers as
 * mba_simba_solve; every device buffer of the search is a region of a static
 * 64 MiB device array (one per device, kept until the process ends, located
 * when a context is created on that device): value tables up to about
 * 55.8 MiB (opts->table_bytes: 0 or at most that; larger searches form more
 * subtree values per candidate instead) and at most 2^17 window-hit records
 * (opts->hit_cap: 0 or at most 131072; a tile with more hits is
 * MBA_ERR_LIMIT).  Larger requests fail with MBA_ERR_LIMIT.  Calls that
 * overlap on one device share the array and race: serialise them.  Host-side
 * bookkeeping uses the C++ heap. */
int mba_simba_solve_noalloc(mba_ctx *ctx, const mba_instance *inst, const mba_opts *opts,
                            mba_result *res);

/* SIMBA-RTid ablation (nocache_rtid.cu, the same driver as nocache.cu with
 * the kernel of :85-95): in each (size, op) block of N trees, thread g of
 * the ceil(N/1024)*1024 launched checks rank ((g * 2246822507) mod 2^64)
 * mod N, so some ranks are visited twice and others never when
 * gcd(2246822507, N) > 1.  The first block with a match decides; within it
 * the tree of the smallest such g (the reference keeps whichever thread
 * wins atomicCAS).  Size 1 has no op block and is not searched. */
int mba_simba_rtid_solve(mba_ctx *ctx, const mba_instance *inst, const mba_opts *opts,
                         mba_result *res);

/* VFB bottom-up cached search (replaces MBA(), full-cache.cu:665-774) with
 * the intended keep-iff-fingerprint-new dedup and the first kept match in
 * enumeration order as the solution.  `stats` may be NULL. */
int mba_vfb_solve(mba_ctx *ctx, const mba_instance *inst, const mba_opts *opts,
                  mba_result *res, mba_vfb_stats *stats);

/* VFB candidate fingerprints as the device computes them for dedup:
 * makeUnqChk (full-cache.cu:126-134, with simpleHash :113-120) of each of
 * `count` signatures values[r * n_words .. + n_words) (each word taken mod
 * 2^width); out[r] receives the 64-bit fingerprint.  Width 32 is the
 * reference's; width 64 uses the 64-bit MurmurHash3 finalizer and a 64-bit
 * sum.  The reference computes them inside processOperator and exports no
 * entry point: this one exposes the kernels' own routine for checks. */
int mba_vfb_fingerprints(mba_ctx *ctx, uint32_t width, const uint64_t *values,
                         uint32_t n_words, size_t count, uint64_t *out);

/* Random-access ceiling of the VFB insert pass on this device.  No
 * counterpart in the reference: a measurement of what bounds the insert
 * kernel (one probe of a random 16-byte slot per candidate, a 64-bit
 * compare-and-swap per new key).  A table of 2^slots_log2 slots in the
 * insert pass's layout holds `prefill` keys; then `count` keys with
 * `distinct` different values (SplitMix64, random order) are probed once
 * each (*probe_s: seconds) and, on the same starting table, inserted by the
 * insert pass's own table code with no fingerprint work (*insert_s;
 * *fresh: keys the insert added).  Allocates and frees its own buffers
 * (16 * 2^slots_log2 + 8 * (count + prefill) bytes). */
int mba_vfb_probe_ceiling(mba_ctx *ctx, uint32_t slots_log2, uint64_t count,
                          uint64_t distinct, uint64_t prefill, double *probe_s,
                          double *insert_s, uint64_t *fresh);

/* Batched SIMBA: `count` instances solved together.  Consecutive sizes of
 * all instances run as captured launch sequences: per size step, two
 * persistent kernels (tabulated B sides, then B sides formed from blocks)
 * drain a device work queue of every instance's jobs, window hits are
 * resolved on the device, and instances settled at a smaller size skip
 * their remaining jobs.  From the size at which the instances' value tables
 * no longer fit the table pool (opts->table_bytes, 0: 96 GiB at most 2/5 of
 * the device), instances are swept in cohorts whose tables do.  Each result equals
 * mba_simba_solve's on that instance.  `opts->time_budget_s` (0 = none)
 * bounds each instance's share of the device time (a launch sequence's time
 * split over the instances by the tiles of their jobs), checked between
 * launch sequences: an instance past it returns MBA_BUDGET. */
int mba_batch_solve(mba_ctx *ctx, const mba_instance *insts, size_t count,
                    const mba_opts *opts, mba_result *results);

/* ---- learned block proposal (no reference counterpart: an extension that
 * orders the plain search's blocks by a learned score).  A 2-layer MLP
 * 256 -> 512 -> 128, BF16 MFMA operands with FP32 accumulation, trained on
 * the device with Adam.  Output class =
 * op * 16 + min(left size, 15) of a block's root (op 0..7, left size 0 for
 * unary roots).  An mba_mlp belongs to the context it was created on. */
typedef struct mba_mlp mba_mlp;
#define MBA_MLP_IN 256
#define MBA_MLP_HIDDEN 512
#define MBA_MLP_OUT 128

/* 256 features of an instance from its hashed sample input/output bits. */
int mba_features(const mba_instance *inst, float *out);
/* The same features of `count` instances computed on the context's device
 * (one workgroup per instance); out [count x 256], equal to mba_features. */
int mba_features_batch(mba_ctx *ctx, const mba_instance *insts, size_t count, float *out);
int mba_mlp_create(mba_ctx *ctx, uint64_t seed, mba_mlp **out);
/* Destroy every network of a context before the context itself. */
void mba_mlp_destroy(mba_mlp *m);
/* Host copies of the FP32 master weights: w1 [256 x 512], b1 [512],
 * w2 [512 x 128], b2 [128], row-major (input-major). */
int mba_mlp_get_weights(mba_mlp *m, float *w1, float *b1, float *w2, float *b2);
int mba_mlp_set_weights(mba_mlp *m, const float *w1, const float *b1, const float *w2,
                        const float *b2);
/* logits [batch x 128] of features x [batch x 256]. */
int mba_mlp_forward(mba_mlp *m, const float *x, int batch, float *logits);
/* Mean softmax cross-entropy over rows with label >= 0 and its gradients
 * (any output pointer may be NULL). */
int mba_mlp_grads(mba_mlp *m, const float *x, const int32_t *labels, int batch, float *loss,
                  float *gw1, float *gb1, float *gw2, float *gb2);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on a batch; returns the loss. */
int mba_mlp_train_step(mba_mlp *m, const float *x, const int32_t *labels, int batch, float lr,
                       float *loss);
/* SIMBA search with the splits of each size of at least 2^26 trees
 * (opts->block_log2 lowers it) visited in descending probability the
 * network gives their blocks per pair checked; returns a verified solution
 * of the minimal size. */
int mba_learned_solve(mba_ctx *ctx, mba_mlp *m, const mba_instance *inst,
                      const mba_opts *opts, mba_result *res);
/* Batched form of mba_learned_solve (rounds of splits over all instances). */
int mba_learned_batch_solve(mba_ctx *ctx, mba_mlp *m, const mba_instance *insts,
                            size_t count, const mba_opts *opts, mba_result *results);

/* Instance reader for the reference JSON schema (sample.py:43-65):
 * {"num_variables", "num_examples", "examples":[{"inputs":[{"value","size"}],
 *  "output":[{"value","size"}]}]}.  Width = the first value's "size" field
 * (32 when absent).  Arrays are malloc'ed; free with mba_instance_free. */
int mba_instance_read_json(const char *path, mba_instance *out);
void mba_instance_free(mba_instance *inst);

/* Reference printing of a SIMBA RPN string (nocache.cu:162-195). */
int mba_rpn_to_infix(const char *rpn, uint32_t n_vars, const uint64_t *consts,
                     uint32_t n_consts, uint32_t width, char *out, size_t cap);

#ifdef __cplusplus
}
#endif
#endif
