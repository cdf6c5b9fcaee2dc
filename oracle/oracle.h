/*
 * oracle.h -- CPU restatement of the SIMBA / VFB search semantics.
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product never links it.
 */
#ifndef MBA_ORACLE_H
#define MBA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_SIZE 32
#define ORC_MAX_LEAVES 16
#define ORC_MAX_RPN 128
#define ORC_MAX_SAMPLES 128
#define ORC_TOK_OP 32

enum { ORC_NOT = 0, ORC_AND, ORC_OR, ORC_XOR, ORC_NEG, ORC_ADD, ORC_SUB, ORC_MUL };
enum { ORC_FOUND = 0, ORC_NOT_FOUND = 1, ORC_BUDGET = 2, ORC_OVERFLOW = 3 };

typedef struct {
  int32_t status;
  int32_t size;      /* size (SIMBA) or level (VFB) of the solution */
  uint64_t rank;     /* SIMBA: rank within size; VFB: enumeration index in level */
  int32_t op;        /* root operator, -1 for a leaf */
  int32_t split;     /* SIMBA: left child size (0 for unary/leaf) */
  uint8_t rpn[ORC_MAX_RPN];
  char expr[1024];
  double seconds;
  uint64_t candidates;
  uint64_t rows; /* VFB: cache rows at the end */
} orc_result;

int orc_count_table(int L, int max_size, uint64_t *T);
int orc_unrank(const uint64_t *T, int max_size, int size, uint64_t rank, uint8_t *rpn);
uint64_t orc_rank(const uint64_t *T, const uint8_t *rpn, int size);
uint64_t orc_apply(int op, uint64_t l, uint64_t r, int W);
uint64_t orc_eval(const uint8_t *rpn, int size, const uint64_t *leafvals, int W);
int orc_is_solution(const uint8_t *rpn, int size, int L, int n_samples,
                    const uint64_t *leafvals, const uint64_t *outputs, int W);
int orc_infix(const uint8_t *rpn, int size, int n_vars, const uint64_t *consts,
              int W, int vfb_style, char *out, int cap);
int orc_simba_search(int n_vars, int n_consts, const uint64_t *consts, int W,
                     int n_samples, const uint64_t *inputs, const uint64_t *outputs,
                     int min_size, int max_size, int nthreads, double budget_s,
                     int exhaustive, orc_result *res);
/* SIMBA-RTid (nocache_rtid.cu:85-95): first (size, op) block with a visited
 * solution; within it the smallest thread index (written to *thread_index). */
int orc_simba_rtid_search(int n_vars, int n_consts, const uint64_t *consts, int W,
                          int n_samples, const uint64_t *inputs, const uint64_t *outputs,
                          int min_size, int max_size, int nthreads, orc_result *res,
                          uint64_t *thread_index);
uint32_t orc_fmix32(uint32_t x);
uint64_t orc_fmix64(uint64_t x);
uint64_t orc_fingerprint(const uint64_t *cs, int n, int W);
int orc_vfb_search(int n_vars, int n_consts, const uint64_t *consts, int W,
                   int n_samples, const uint64_t *inputs, const uint64_t *outputs,
                   int max_len, int dedup_mode, int exhaustive, uint64_t max_rows,
                   orc_result *res, uint64_t *level_new, uint64_t *level_cand);
/* Multithreaded VFB with the intended dedup (the CPU baseline): the same
 * solution, per-level counts and candidates as orc_vfb_search, every level
 * swept fully; nthreads workers pinned to the process's CPUs. */
int orc_vfb_search_mt(int n_vars, int n_consts, const uint64_t *consts, int W,
                      int n_samples, const uint64_t *inputs, const uint64_t *outputs,
                      int max_len, int exhaustive, int nthreads, orc_result *res,
                      uint64_t *level_new, uint64_t *level_cand);

#ifdef __cplusplus
}
#endif
#endif
