/*
 * oracle.c -- CPU restatement of the SIMBA and VFB search semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (simba_amd/) links,
 * loads or calls this file.  It is used by tests/ as the parity checker, by
 * __graft_entry__.smoke() as the checker of one small GPU solve, and by
 * bench.py's cpu_baseline leg (multithreaded, "port" kind).
 *
 * Every function cites the reference file:line it restates.  Reference
 * semantics are uint32 (mod 2^32); this restatement is parameterised by the
 * word width W in {32, 64} and by extra constant leaves (appended after the
 * variables), which reduce to the reference exactly for W=32 and no
 * constants.
 *
 * Pinning: every function is checked by tests/ against brute force, closed
 * forms or golden vectors.  The VFB fingerprint's finalizers are checked
 * through the published MurmurHash3 vectors (fmix32: x86_32, fmix64:
 * x64_128), its combine step (makeUnqChk) by known answers of one sample,
 * the 32-bit wrap of its first sum and the rotation of odd indices.
 *
 * Token encoding of an RPN formula (one byte per node):
 *   0..31   leaf index (variables first, then constants)
 *   32+o    operator o, o in 0..7 = ~ & | ^ neg + - *   (globals.cuh:13-15)
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <pthread.h>
#include <sched.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static inline uint64_t wmask(int W) { return W == 64 ? ~0ull : 0xffffffffull; }

/* Pins the calling thread to the k-th CPU (mod the count) of the process's
 * allowed set, so worker k of a pool runs on its own core. */
static void pin_self(int k) {
  cpu_set_t allowed, one;
  if (sched_getaffinity(0, sizeof allowed, &allowed) != 0) return;
  int n = CPU_COUNT(&allowed);
  if (n <= 0) return;
  int want = k % n, seen = 0;
  for (int c = 0; c < CPU_SETSIZE; ++c) {
    if (!CPU_ISSET(c, &allowed)) continue;
    if (seen++ == want) {
      CPU_ZERO(&one);
      CPU_SET(c, &one);
      pthread_setaffinity_np(pthread_self(), sizeof one, &one);
      return;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Count table: nocache.cu:117-160 (compute_n_formulas).                     */
/* Columns 0..7 = ops in order ~ & | ^ neg + - *, column 8 = total.          */
/* Size-1 op columns are zero; T[1][8] = L.  Returns -1 on u64 overflow.     */
/* ------------------------------------------------------------------------ */
int orc_count_table(int L, int max_size, uint64_t *T) {
  if (max_size < 1 || max_size > ORC_MAX_SIZE || L < 1 || L > ORC_MAX_LEAVES)
    return -2;
  memset(T, 0, sizeof(uint64_t) * 9 * (size_t)(max_size + 1));
  T[1 * 9 + 8] = (uint64_t)L;
  for (int s = 2; s <= max_size; ++s) {
    unsigned __int128 U = T[(s - 1) * 9 + 8], C = 0, NC = 0;
    for (int j = 1; j <= (s - 1) / 2; ++j)
      C += (unsigned __int128)T[j * 9 + 8] * T[(s - 1 - j) * 9 + 8];
    for (int j = 1; j <= s - 2; ++j)
      NC += (unsigned __int128)T[j * 9 + 8] * T[(s - 1 - j) * 9 + 8];
    unsigned __int128 tot = 2 * U + 5 * C + NC;
    if (tot >> 64) return -1;
    T[s * 9 + 0] = T[s * 9 + 4] = (uint64_t)U;
    T[s * 9 + 1] = T[s * 9 + 2] = T[s * 9 + 3] = T[s * 9 + 5] =
        T[s * 9 + 7] = (uint64_t)C;
    T[s * 9 + 6] = (uint64_t)NC;
    T[s * 9 + 8] = (uint64_t)tot;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Unrank: nocache.cu:23-83 (numberToFormula).                               */
/* The rank is the index within all trees of `size` (op blocks in order).    */
/* Left child occupies rpn[0..j), right child rpn[j..size-1), op at the end. */
/* ------------------------------------------------------------------------ */
static void unrank_rec(const uint64_t *T, int size, uint64_t n, uint8_t *rpn) {
  if (size == 1) {
    rpn[0] = (uint8_t)n;
    return;
  }
  uint64_t acc = 0;
  int op = 0;
  while (n >= acc + T[size * 9 + op]) acc += T[size * 9 + op++];
  n -= acc;
  int cs = size - 1; /* total size of the children */
  rpn[cs] = (uint8_t)(ORC_TOK_OP + op);
  if (op == ORC_NOT || op == ORC_NEG) {
    unrank_rec(T, cs, n, rpn);
    return;
  }
  uint64_t fsum = 0;
  int j = 1;
  while (n >= fsum + T[j * 9 + 8] * T[(cs - j) * 9 + 8]) {
    fsum += T[j * 9 + 8] * T[(cs - j) * 9 + 8];
    ++j;
  }
  n -= fsum;
  uint64_t rs = T[(cs - j) * 9 + 8];
  unrank_rec(T, j, n / rs, rpn);
  unrank_rec(T, cs - j, n % rs, rpn + j);
}

int orc_unrank(const uint64_t *T, int max_size, int size, uint64_t rank,
               uint8_t *rpn) {
  if (size < 1 || size > max_size || rank >= T[size * 9 + 8]) return -1;
  unrank_rec(T, size, rank, rpn);
  return 0;
}

/* Inverse of orc_unrank (not in the reference; used to test the bijection). */
static int subtree_len(const uint8_t *rpn, int end) {
  /* length of the subtree whose root is rpn[end] */
  int need = 1, i = end;
  while (need > 0) {
    uint8_t t = rpn[i--];
    if (t < ORC_TOK_OP) need -= 1;
    else if (t == ORC_TOK_OP + ORC_NOT || t == ORC_TOK_OP + ORC_NEG) need += 0;
    else need += 1;
  }
  return end - i;
}

uint64_t orc_rank(const uint64_t *T, const uint8_t *rpn, int size) {
  if (size == 1) return rpn[0];
  int op = rpn[size - 1] - ORC_TOK_OP;
  uint64_t r = 0;
  for (int o = 0; o < op; ++o) r += T[size * 9 + o];
  int cs = size - 1;
  if (op == ORC_NOT || op == ORC_NEG) return r + orc_rank(T, rpn, cs);
  int rlen = subtree_len(rpn, size - 2);
  int j = cs - rlen;
  for (int jj = 1; jj < j; ++jj) r += T[jj * 9 + 8] * T[(cs - jj) * 9 + 8];
  return r + orc_rank(T, rpn, j) * T[(cs - j) * 9 + 8] + orc_rank(T, rpn + j, rlen);
}

/* ------------------------------------------------------------------------ */
/* Evaluation: mba.cuh:11-66 (is_solution_rpn), one sample.                   */
/* Binary ops pop right then left.                                           */
/* ------------------------------------------------------------------------ */
uint64_t orc_apply(int op, uint64_t l, uint64_t r, int W) {
  uint64_t m = wmask(W), v = 0;
  switch (op) {
  case ORC_NOT: v = ~l; break;
  case ORC_AND: v = l & r; break;
  case ORC_OR: v = l | r; break;
  case ORC_XOR: v = l ^ r; break;
  case ORC_NEG: v = 0 - l; break;
  case ORC_ADD: v = l + r; break;
  case ORC_SUB: v = l - r; break;
  case ORC_MUL: v = l * r; break;
  }
  return v & m;
}

uint64_t orc_eval(const uint8_t *rpn, int size, const uint64_t *leafvals,
                  int W) {
  uint64_t st[ORC_MAX_SIZE];
  int sp = -1;
  for (int i = 0; i < size; ++i) {
    uint8_t t = rpn[i];
    if (t < ORC_TOK_OP) {
      st[++sp] = leafvals[t] & wmask(W);
    } else {
      int op = t - ORC_TOK_OP;
      if (op == ORC_NOT || op == ORC_NEG) {
        st[sp] = orc_apply(op, st[sp], 0, W);
      } else {
        uint64_t r = st[sp--];
        st[sp] = orc_apply(op, st[sp], r, W);
      }
    }
  }
  return st[0];
}

/* All samples, in order, first mismatch exits (mba.cuh:13-66). */
int orc_is_solution(const uint8_t *rpn, int size, int L, int n_samples,
                    const uint64_t *leafvals, const uint64_t *outputs, int W) {
  for (int i = 0; i < n_samples; ++i)
    if (orc_eval(rpn, size, leafvals + (size_t)i * L, W) !=
        (outputs[i] & wmask(W)))
      return 0;
  return 1;
}

/* ------------------------------------------------------------------------ */
/* Printers.                                                                 */
/* SIMBA infix: nocache.cu:162-195 (MBAToString): leaf v<d>, (~X), (-X),     */
/* (L op R).  VFB infix: full-cache.cu:483-509 (toString): leaf v<i>, ~(X),  */
/* -(X), (L)op(R).  Constant leaves (an extension) print their value, signed */
/* when the top bit of the W-bit word is set.                                */
/* ------------------------------------------------------------------------ */
static const char OPCH[8] = {'~', '&', '|', '^', '-', '+', '-', '*'};

static int leaf_str(char *buf, int cap, int leaf, int n_vars,
                    const uint64_t *consts, int W) {
  if (leaf < n_vars) return snprintf(buf, cap, "v%d", leaf);
  uint64_t c = consts[leaf - n_vars] & wmask(W);
  uint64_t sign = 1ull << (W - 1);
  if (c & sign) return snprintf(buf, cap, "-%llu", (unsigned long long)((0 - c) & wmask(W)));
  return snprintf(buf, cap, "%llu", (unsigned long long)c);
}

typedef struct { char *s; int len; } sstr;

static int infix_rec(const uint8_t *rpn, int end, int n_vars,
                     const uint64_t *consts, int W, int vfb_style, char *out,
                     int cap, int *used) {
  /* returns start index of the subtree rooted at `end` */
  uint8_t t = rpn[end];
  if (t < ORC_TOK_OP) {
    char tmp[64];
    int n = leaf_str(tmp, sizeof tmp, t, n_vars, consts, W);
    if (*used + n < cap) { memcpy(out + *used, tmp, n); *used += n; }
    return end;
  }
  int op = t - ORC_TOK_OP;
#define PUT(str)                                                             \
  do {                                                                       \
    const char *_p = (str);                                                  \
    int _n = (int)strlen(_p);                                                \
    if (*used + _n < cap) { memcpy(out + *used, _p, _n); *used += _n; }      \
  } while (0)
  if (op == ORC_NOT || op == ORC_NEG) {
    const char *o = op == ORC_NOT ? "~" : "-";
    if (vfb_style) { PUT(o); PUT("("); } else { PUT("("); PUT(o); }
    int st = infix_rec(rpn, end - 1, n_vars, consts, W, vfb_style, out, cap, used);
    PUT(")");
    return st;
  }
  /* binary: right subtree ends at end-1; find its start first */
  int rlen = subtree_len(rpn, end - 1);
  int lend = end - 1 - rlen;
  char ob[4] = {0};
  if (vfb_style) {
    PUT("(");
    int st = infix_rec(rpn, lend, n_vars, consts, W, vfb_style, out, cap, used);
    PUT(")");
    ob[0] = OPCH[op];
    PUT(ob);
    PUT("(");
    infix_rec(rpn, end - 1, n_vars, consts, W, vfb_style, out, cap, used);
    PUT(")");
    return st;
  } else {
    PUT("(");
    int st = infix_rec(rpn, lend, n_vars, consts, W, vfb_style, out, cap, used);
    ob[0] = ' '; ob[1] = OPCH[op]; ob[2] = ' ';
    PUT(ob);
    infix_rec(rpn, end - 1, n_vars, consts, W, vfb_style, out, cap, used);
    PUT(")");
    return st;
  }
#undef PUT
}

int orc_infix(const uint8_t *rpn, int size, int n_vars, const uint64_t *consts,
              int W, int vfb_style, char *out, int cap) {
  int used = 0;
  if (size < 1) { if (cap) out[0] = 0; return 0; }
  infix_rec(rpn, size - 1, n_vars, consts, W, vfb_style, out, cap, &used);
  out[used < cap ? used : cap - 1] = 0;
  return used;
}

/* ------------------------------------------------------------------------ */
/* SIMBA search: nocache.cu:197-289 (NoCache::solve) with the deterministic   */
/* first solution = minimum rank inside the first size that has a solution  */
/* (the reference latches whichever thread wins atomicCAS, nocache.cu:103). */
/* Multithreaded: ranks of one size are split into chunks pulled from an    */
/* atomic counter; the minimum matching rank wins.                          */
/* ------------------------------------------------------------------------ */
typedef struct {
  const uint64_t *T;
  int size, L, n_samples, W;
  const uint64_t *leafvals, *outputs;
  uint64_t total;
  _Atomic uint64_t next_chunk;
  _Atomic uint64_t best;
  _Atomic uint64_t evaluated;
  _Atomic int stop;
  _Atomic int next_worker;
  double deadline;
  int exhaustive, pin;
} simba_job;

#define ORC_CHUNK 16384ull

static void *simba_worker(void *arg) {
  simba_job *J = (simba_job *)arg;
  if (J->pin) pin_self(atomic_fetch_add(&J->next_worker, 1));
  uint8_t rpn[ORC_MAX_SIZE];
  uint64_t done = 0;
  for (;;) {
    if (atomic_load(&J->stop)) break;
    uint64_t c = atomic_fetch_add(&J->next_chunk, 1);
    uint64_t lo = c * ORC_CHUNK;
    if (lo >= J->total || (!J->exhaustive && lo > atomic_load(&J->best))) break;
    uint64_t hi = lo + ORC_CHUNK;
    if (hi > J->total) hi = J->total;
    for (uint64_t r = lo; r < hi; ++r) {
      unrank_rec(J->T, J->size, r, rpn);
      ++done;
      if (orc_is_solution(rpn, J->size, J->L, J->n_samples, J->leafvals,
                          J->outputs, J->W)) {
        uint64_t b = atomic_load(&J->best);
        while (r < b && !atomic_compare_exchange_weak(&J->best, &b, r)) {
        }
        if (!J->exhaustive) break; /* later ranks of this chunk cannot be smaller */
      }
    }
    if (J->deadline > 0 && now_s() > J->deadline) {
      atomic_store(&J->stop, 1);
      break;
    }
  }
  atomic_fetch_add(&J->evaluated, done);
  return NULL;
}

int orc_simba_search(int n_vars, int n_consts, const uint64_t *consts, int W,
                     int n_samples, const uint64_t *inputs,
                     const uint64_t *outputs, int min_size, int max_size,
                     int nthreads, double budget_s, int exhaustive, orc_result *res) {
  memset(res, 0, sizeof *res);
  int L = n_vars + n_consts;
  if (L < 1 || L > ORC_MAX_LEAVES || max_size > ORC_MAX_SIZE || min_size < 1 ||
      n_samples < 1)
    return -2;
  uint64_t T[(ORC_MAX_SIZE + 1) * 9];
  if (orc_count_table(L, max_size, T) != 0) return -1;
  uint64_t *lv = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n_samples * L);
  for (int i = 0; i < n_samples; ++i) {
    for (int v = 0; v < n_vars; ++v) lv[i * L + v] = inputs[i * n_vars + v] & wmask(W);
    for (int c = 0; c < n_consts; ++c) lv[i * L + n_vars + c] = consts[c] & wmask(W);
  }
  if (nthreads < 1) nthreads = 1;
  double t0 = now_s();
  uint64_t evaluated = 0;
  int status = ORC_NOT_FOUND;
  for (int s = min_size; s <= max_size; ++s) {
    simba_job J;
    J.T = T; J.size = s; J.L = L; J.n_samples = n_samples; J.W = W;
    J.leafvals = lv; J.outputs = outputs; J.total = T[s * 9 + 8];
    atomic_init(&J.next_chunk, 0);
    atomic_init(&J.best, UINT64_MAX);
    atomic_init(&J.evaluated, 0);
    atomic_init(&J.stop, 0);
    atomic_init(&J.next_worker, 0);
    J.deadline = budget_s > 0 ? t0 + budget_s : 0;
    J.exhaustive = exhaustive;
    pthread_t th[256];
    int nt = nthreads > 256 ? 256 : nthreads;
    J.pin = nt > 1;
    for (int k = 0; k < nt; ++k) pthread_create(&th[k], NULL, simba_worker, &J);
    for (int k = 0; k < nt; ++k) pthread_join(th[k], NULL);
    evaluated += atomic_load(&J.evaluated);
    uint64_t b = atomic_load(&J.best);
    if (atomic_load(&J.stop) && b == UINT64_MAX) {
      status = ORC_BUDGET;
      res->size = s; /* size reached */
      break;
    }
    if (b != UINT64_MAX && status != ORC_FOUND) {
      status = ORC_FOUND;
      res->size = s;
      res->rank = b;
      unrank_rec(T, s, b, res->rpn);
      if (s > 1) {
        int op = res->rpn[s - 1] - ORC_TOK_OP;
        res->op = op;
        res->split = (op == ORC_NOT || op == ORC_NEG) ? 0 : s - 1 - subtree_len(res->rpn, s - 2);
      } else {
        res->op = -1;
      }
      orc_infix(res->rpn, s, n_vars, consts, W, 0, res->expr, sizeof res->expr);
      if (!exhaustive) break;
    }
  }
  res->status = status;
  res->seconds = now_s() - t0;
  res->candidates = evaluated;
  free(lv);
  return status;
}

/* ------------------------------------------------------------------------ */
/* SIMBA-RTid: nocache_rtid.cu:85-95.  Each (size, op) block of N trees is    */
/* launched as ceil(N/1024) x 1024 threads; thread g checks the tree of rank  */
/* offset + ((g * C) mod 2^64) mod N with C = 2246822507, so the visited      */
/* ranks are a multiset that may miss ranks when gcd(C, N) > 1.  The first    */
/* (size, op) block, in order, with a visited solution decides; the          */
/* reference keeps the thread that wins atomicCAS, restated here as the      */
/* smallest thread index g (`rank` = its rank, `candidates` = threads run).   */
/* ------------------------------------------------------------------------ */
#define ORC_RTID_C 2246822507ull

typedef struct {
  const uint64_t *T;
  int size, L, n_samples, W;
  const uint64_t *leafvals, *outputs;
  uint64_t offset, N, G;
  _Atomic uint64_t next_chunk;
  _Atomic uint64_t best_g;
  _Atomic uint64_t evaluated;
} rtid_job;

static void *rtid_worker(void *arg) {
  rtid_job *J = (rtid_job *)arg;
  uint8_t rpn[ORC_MAX_SIZE];
  uint64_t done = 0;
  for (;;) {
    uint64_t c = atomic_fetch_add(&J->next_chunk, 1);
    uint64_t lo = c * ORC_CHUNK;
    if (lo >= J->G || lo > atomic_load(&J->best_g)) break;
    uint64_t hi = lo + ORC_CHUNK;
    if (hi > J->G) hi = J->G;
    for (uint64_t g = lo; g < hi; ++g) {
      uint64_t r = J->offset + (g * ORC_RTID_C) % J->N;
      unrank_rec(J->T, J->size, r, rpn);
      ++done;
      if (orc_is_solution(rpn, J->size, J->L, J->n_samples, J->leafvals, J->outputs, J->W)) {
        uint64_t b = atomic_load(&J->best_g);
        while (g < b && !atomic_compare_exchange_weak(&J->best_g, &b, g)) {
        }
        break;
      }
    }
  }
  atomic_fetch_add(&J->evaluated, done);
  return NULL;
}

int orc_simba_rtid_search(int n_vars, int n_consts, const uint64_t *consts, int W,
                          int n_samples, const uint64_t *inputs, const uint64_t *outputs,
                          int min_size, int max_size, int nthreads, orc_result *res,
                          uint64_t *thread_index) {
  memset(res, 0, sizeof *res);
  int L = n_vars + n_consts;
  if (L < 1 || L > ORC_MAX_LEAVES || max_size > ORC_MAX_SIZE || min_size < 1 ||
      n_samples < 1)
    return -2;
  uint64_t T[(ORC_MAX_SIZE + 1) * 9];
  if (orc_count_table(L, max_size, T) != 0) return -1;
  uint64_t *lv = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n_samples * L);
  for (int i = 0; i < n_samples; ++i) {
    for (int v = 0; v < n_vars; ++v) lv[i * L + v] = inputs[i * n_vars + v] & wmask(W);
    for (int c = 0; c < n_consts; ++c) lv[i * L + n_vars + c] = consts[c] & wmask(W);
  }
  if (nthreads < 1) nthreads = 1;
  double t0 = now_s();
  uint64_t evaluated = 0;
  int status = ORC_NOT_FOUND;
  for (int s = min_size; s <= max_size && status != ORC_FOUND; ++s) {
    uint64_t offset = 0;
    for (int op = 0; op < 8; ++op) {
      uint64_t N = T[s * 9 + op];
      if (N == 0) continue;
      rtid_job J;
      J.T = T; J.size = s; J.L = L; J.n_samples = n_samples; J.W = W;
      J.leafvals = lv; J.outputs = outputs;
      J.offset = offset; J.N = N; J.G = (N + 1023) / 1024 * 1024;
      atomic_init(&J.next_chunk, 0);
      atomic_init(&J.best_g, UINT64_MAX);
      atomic_init(&J.evaluated, 0);
      pthread_t th[256];
      int nt = nthreads > 256 ? 256 : nthreads;
      for (int k = 0; k < nt; ++k) pthread_create(&th[k], NULL, rtid_worker, &J);
      for (int k = 0; k < nt; ++k) pthread_join(th[k], NULL);
      evaluated += atomic_load(&J.evaluated);
      uint64_t g = atomic_load(&J.best_g);
      if (g != UINT64_MAX) {
        uint64_t r = offset + (g * ORC_RTID_C) % N;
        status = ORC_FOUND;
        res->size = s;
        res->rank = r;
        if (thread_index) *thread_index = g;
        unrank_rec(T, s, r, res->rpn);
        res->op = op;
        res->split = (op == ORC_NOT || op == ORC_NEG) ? 0 : s - 1 - subtree_len(res->rpn, s - 2);
        orc_infix(res->rpn, s, n_vars, consts, W, 0, res->expr, sizeof res->expr);
        break;
      }
      offset += N;
    }
  }
  if (status != ORC_FOUND) strcpy(res->expr, "Not found !");
  res->status = status;
  res->seconds = now_s() - t0;
  res->candidates = evaluated;
  free(lv);
  return status;
}

/* ------------------------------------------------------------------------ */
/* VFB fingerprint: full-cache.cu:113-134 (simpleHash, makeUnqChk).           */
/* W=32 is the reference.  Note that `simpleHash(x) + 0x9e3779b9` is a sum   */
/* of two 32-bit unsigned ints, so it wraps mod 2^32 before being widened.  */
/* W=64 (extension): the 64-bit MurmurHash3 finalizer replaces the 32-bit   */
/* one and the sum is formed in 64 bits.                                    */
/* ------------------------------------------------------------------------ */
uint32_t orc_fmix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x85ebca6bu;
  x ^= x >> 13;
  x *= 0xc2b2ae35u;
  x ^= x >> 16;
  return x;
}

uint64_t orc_fmix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

uint64_t orc_fingerprint(const uint64_t *cs, int n, int W) {
  uint64_t seed = 0;
  for (int i = 0; i < n; ++i) {
    uint64_t val;
    if (W == 32) {
      uint32_t h = orc_fmix32((uint32_t)cs[i]) + 0x9e3779b9u; /* 32-bit wrap */
      val = (uint64_t)h + (seed << 6) + (seed >> 2);
    } else {
      val = orc_fmix64(cs[i]) + 0x9e3779b9ull + (seed << 6) + (seed >> 2);
    }
    if (i & 1) val = (val >> 32) | (val << 32);
    seed ^= val;
  }
  return seed;
}

/* ------------------------------------------------------------------------ */
/* VFB search: full-cache.cu:665-774 (MBA) + generateMBAs 568-663 +           */
/* processOperator 333-372 + insertInCache 293-321.                          */
/* Enumeration order at level Lv: ops ~ & | ^ neg + - * ; unary over rows of */
/* level Lv-1; binary over splits i = 1..floor((Lv-1)/2) with left rows of   */
/* level i (outer) and right rows of level Lv-1-i (inner); '-' also (r,l).   */
/* Dedup: fingerprint set.  mode 0 (intended, full-cache.cu:122-125,275-276) */
/* keeps a candidate iff its fingerprint is new; mode 1 (literal: hash_set   */
/* .cuh:214-218,243-246 vs full-cache.cu:286) keeps iff it was already       */
/* present.  Fingerprints equal to the table's EMPTY (2^64-1) or TOMBSTONE  */
/* (2^64-2) keys are invalid and dropped (hash_set.cuh:200-204).            */
/* The first kept candidate equal to the target, in enumeration order, is   */
/* the solution.                                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
  uint64_t *keys;
  uint64_t cap; /* power of two */
  uint64_t count;
} fpset;

static void fpset_init(fpset *s, uint64_t cap) {
  uint64_t c = 1024;
  while (c < cap) c <<= 1;
  s->cap = c;
  s->count = 0;
  s->keys = (uint64_t *)malloc(sizeof(uint64_t) * c);
  memset(s->keys, 0xff, sizeof(uint64_t) * c);
}

static void fpset_grow(fpset *s);

/* returns 1 if already present, 0 if inserted now */
static int fpset_insert(fpset *s, uint64_t k) {
  if (2 * (s->count + 1) > s->cap) fpset_grow(s);
  uint64_t m = s->cap - 1, i = orc_fmix64(k) & m;
  for (;;) {
    if (s->keys[i] == k) return 1;
    if (s->keys[i] == ~0ull) {
      s->keys[i] = k;
      s->count++;
      return 0;
    }
    i = (i + 1) & m;
  }
}

static void fpset_grow(fpset *s) {
  uint64_t *old = s->keys, oc = s->cap;
  fpset_init(s, oc * 2);
  for (uint64_t i = 0; i < oc; ++i)
    if (old[i] != ~0ull) fpset_insert(s, old[i]);
  free(old);
}

typedef struct {
  int n, W;
  uint64_t *vals; /* row-major [row][n] */
  int32_t *left, *right;
  uint8_t *op;
  uint64_t rows, cap;
} vcache;

static void vc_push(vcache *c, const uint64_t *cs, int32_t l, int32_t r, int op) {
  if (c->rows == c->cap) {
    c->cap = c->cap ? c->cap * 2 : 1024;
    c->vals = (uint64_t *)realloc(c->vals, sizeof(uint64_t) * c->n * c->cap);
    c->left = (int32_t *)realloc(c->left, sizeof(int32_t) * c->cap);
    c->right = (int32_t *)realloc(c->right, sizeof(int32_t) * c->cap);
    c->op = (uint8_t *)realloc(c->op, c->cap);
  }
  memcpy(c->vals + c->rows * c->n, cs, sizeof(uint64_t) * c->n);
  c->left[c->rows] = l;
  c->right[c->rows] = r;
  c->op[c->rows] = (uint8_t)op;
  c->rows++;
}

/* Rebuild an RPN from the child-index tree (full-cache.cu:422-553). */
static int vfb_rpn(const vcache *c, int L, int64_t row, uint8_t *rpn, int pos,
                   int cap) {
  if (pos >= cap) return -1;
  if (row < L) {
    rpn[pos] = (uint8_t)row;
    return pos + 1;
  }
  int op = c->op[row];
  if (op == ORC_NOT || op == ORC_NEG) {
    pos = vfb_rpn(c, L, c->left[row], rpn, pos, cap);
  } else {
    pos = vfb_rpn(c, L, c->left[row], rpn, pos, cap);
    if (pos < 0) return -1;
    pos = vfb_rpn(c, L, c->right[row], rpn, pos, cap);
  }
  if (pos < 0 || pos >= cap) return -1;
  rpn[pos] = (uint8_t)(ORC_TOK_OP + op);
  return pos + 1;
}

int orc_vfb_search(int n_vars, int n_consts, const uint64_t *consts, int W,
                   int n_samples, const uint64_t *inputs,
                   const uint64_t *outputs, int max_len, int dedup_mode,
                   int exhaustive, uint64_t max_rows, orc_result *res,
                   uint64_t *level_new /* [max_len+1] or NULL */,
                   uint64_t *level_cand /* [max_len+1] or NULL */) {
  memset(res, 0, sizeof *res);
  int L = n_vars + n_consts, n = n_samples;
  if (L < 1 || L > ORC_MAX_LEAVES || n < 1 || max_len < 1 || max_len > 100)
    return -2;
  double t0 = now_s();
  uint64_t mask = wmask(W);
  vcache C = {0};
  C.n = n;
  C.W = W;
  uint64_t *tgt = (uint64_t *)malloc(sizeof(uint64_t) * n);
  uint64_t *cs = (uint64_t *)malloc(sizeof(uint64_t) * n);
  for (int i = 0; i < n; ++i) tgt[i] = outputs[i] & mask;
  if (level_new) memset(level_new, 0, sizeof(uint64_t) * (max_len + 1));
  if (level_cand) memset(level_cand, 0, sizeof(uint64_t) * (max_len + 1));
  fpset S;
  fpset_init(&S, 1 << 16);
  /* level 1: leaves, checked in order (full-cache.cu:694-707) */
  int64_t found_row = -1;
  int found_level = 0;
  uint64_t found_index = 0;
  uint64_t evaluated = 0;
  for (int v = 0; v < L; ++v) {
    for (int i = 0; i < n; ++i)
      cs[i] = (v < n_vars ? inputs[i * n_vars + v] : consts[v - n_vars]) & mask;
    vc_push(&C, cs, -1, -1, 255);
    evaluated++;
    if (found_row < 0 && !memcmp(cs, tgt, sizeof(uint64_t) * n)) {
      found_row = v;
      found_level = 1;
      found_index = (uint64_t)v;
    }
  }
  if (level_new) level_new[1] = L;
  if (level_cand) level_cand[1] = L;
  /* hashSetInit (full-cache.cu:142-157): insert all leaf fingerprints */
  for (int v = 0; v < L; ++v) {
    uint64_t fp = orc_fingerprint(C.vals + (size_t)v * n, n, W);
    if (fp < ~1ull) fpset_insert(&S, fp);
  }
  uint64_t lvl_start[102];
  lvl_start[1] = 0;
  lvl_start[2] = C.rows;
  int overflow = 0;
  for (int lev = 2; lev <= max_len && (found_row < 0 || exhaustive) && !overflow;
       ++lev) {
    uint64_t idx = 0; /* enumeration index within the level */
    int stop_level = 0;
    for (int op = 0; op < 8 && !stop_level; ++op) {
      int unary = (op == ORC_NOT || op == ORC_NEG);
      int nsplit = unary ? 1 : (lev - 1) / 2;
      for (int sp = 1; sp <= nsplit && !stop_level; ++sp) {
        uint64_t l0, l1, r0 = 0, r1 = 1;
        if (unary) {
          l0 = lvl_start[lev - 1];
          l1 = lvl_start[lev];
        } else {
          l0 = lvl_start[sp];
          l1 = lvl_start[sp + 1];
          r0 = lvl_start[lev - 1 - sp];
          r1 = lvl_start[lev - sp];
        }
        for (uint64_t l = l0; l < l1 && !stop_level; ++l) {
          for (uint64_t r = r0; r < r1 && !stop_level; ++r) {
            int nsw = (op == ORC_SUB) ? 2 : 1;
            for (int sw = 0; sw < nsw; ++sw, ++idx) {
              uint64_t a = sw ? r : l, b = sw ? l : r;
              const uint64_t *A = C.vals + a * n;
              const uint64_t *B = unary ? NULL : C.vals + b * n;
              for (int i = 0; i < n; ++i)
                cs[i] = orc_apply(op, A[i], unary ? 0 : B[i], W);
              evaluated++;
              if (level_cand) level_cand[lev]++;
              uint64_t fp = orc_fingerprint(cs, n, W);
              if (fp >= ~1ull) continue; /* invalid key: dropped */
              int present = fpset_insert(&S, fp);
              int keep = dedup_mode == 0 ? !present : present;
              if (!keep) continue;
              if (max_rows && C.rows >= max_rows) {
                overflow = 1;
                stop_level = 1;
                break;
              }
              vc_push(&C, cs, (int32_t)a, unary ? -1 : (int32_t)b, op);
              if (level_new) level_new[lev]++;
              if (found_row < 0 && !memcmp(cs, tgt, sizeof(uint64_t) * n)) {
                found_row = (int64_t)C.rows - 1;
                found_level = lev;
                found_index = idx;
                if (!exhaustive) {
                  stop_level = 1;
                  break;
                }
              }
            }
          }
        }
      }
    }
    lvl_start[lev + 1] = C.rows;
  }
  int status;
  if (found_row >= 0) {
    status = ORC_FOUND;
    res->size = found_level;
    res->rank = found_index;
    int len = vfb_rpn(&C, L, found_row, res->rpn, 0, ORC_MAX_RPN);
    if (len < 0) {
      strcpy(res->expr, "Size of the output is too big !");
      res->op = -1;
    } else {
      res->op = found_level > 1 ? res->rpn[len - 1] - ORC_TOK_OP : -1;
      orc_infix(res->rpn, len, n_vars, consts, W, 1, res->expr, sizeof res->expr);
    }
  } else {
    status = overflow ? ORC_OVERFLOW : ORC_NOT_FOUND;
    strcpy(res->expr, "Not found !");
  }
  res->status = status;
  res->candidates = evaluated;
  res->rows = C.rows;
  res->seconds = now_s() - t0;
  free(C.vals);
  free(C.left);
  free(C.right);
  free(C.op);
  free(S.keys);
  free(tgt);
  free(cs);
  return status;
}

/* ------------------------------------------------------------------------ */
/* Multithreaded VFB (the CPU baseline's form of orc_vfb_search, intended    */
/* dedup): the same enumeration order, dedup and solution, computed level by */
/* level with every level swept fully.  Per level, three passes over the     */
/* candidate index space, split into chunks pulled from an atomic counter:   */
/*   1. signature -> fingerprint -> a shared open-addressing set (CAS on the */
/*      key; the slot's companion takes the atomic minimum of                */
/*      level << 48 | index, so keys of earlier levels win and the first     */
/*      occurrence in enumeration order is the representative);             */
/*   2. keep = companion == own tag; per-chunk counts (prefix on one thread);*/
/*   3. kept rows written in enumeration order.                              */
/* The solution is the first candidate equal to the target when its         */
/* fingerprint's companion holds that candidate (full-cache.cu:293-321).     */
/* ------------------------------------------------------------------------ */
#define MT_CHUNK 65536ull

typedef struct {
  uint64_t begin, count; /* candidate index range of the block */
  uint64_t l0, nl, r0, nr;
  int op, unary;
} mt_block;

typedef struct {
  int n, W, nblocks, nthreads;
  const uint64_t *vals; /* rows [row][n] */
  const uint64_t *tgt;
  const mt_block *blk;
  uint64_t total, tag, mask;
  uint64_t *keys, *comp; /* the set: key and companion per slot */
  uint32_t *slot_of;     /* per candidate, ~0 = invalid key */
  uint8_t *keep;
  uint64_t *chunk_cnt;   /* per chunk: kept count, then base */
  uint64_t *out_vals;    /* rows of this level */
  int32_t *out_l, *out_r;
  uint8_t *out_op;
  _Atomic uint64_t next;
  _Atomic uint64_t found;
  _Atomic int next_worker;
  int pass;
} mt_level;

static void mt_decode(const mt_level *M, uint64_t idx, int *op, uint64_t *a, uint64_t *b,
                      int *unary) {
  int lo = 0, hi = M->nblocks - 1;
  while (lo < hi) { /* last block whose begin <= idx */
    int mid = (lo + hi + 1) / 2;
    if (M->blk[mid].begin <= idx) lo = mid; else hi = mid - 1;
  }
  const mt_block *B = &M->blk[lo];
  uint64_t local = idx - B->begin;
  *op = B->op;
  *unary = B->unary;
  if (B->unary) {
    *a = B->l0 + local;
    *b = *a;
    return;
  }
  int sw = 0;
  if (B->op == ORC_SUB) {
    sw = (int)(local & 1);
    local >>= 1;
  }
  uint64_t l = B->l0 + local / B->nr, r = B->r0 + local % B->nr;
  *a = sw ? r : l;
  *b = sw ? l : r;
}

static void mt_sig(const mt_level *M, uint64_t idx, uint64_t *cs, int32_t *l, int32_t *r,
                   int *op_out) {
  int op, unary;
  uint64_t a, b;
  mt_decode(M, idx, &op, &a, &b, &unary);
  const uint64_t *A = M->vals + a * M->n, *Bv = M->vals + b * M->n;
  for (int i = 0; i < M->n; ++i) cs[i] = orc_apply(op, A[i], unary ? 0 : Bv[i], M->W);
  *l = (int32_t)a;
  *r = unary ? -1 : (int32_t)b;
  *op_out = op;
}

static uint64_t mt_insert(mt_level *M, uint64_t fp) {
  uint64_t s = orc_fmix64(fp) & M->mask;
  for (;;) {
    uint64_t k = __atomic_load_n(&M->keys[s], __ATOMIC_RELAXED);
    if (k == fp) return s;
    if (k == ~0ull) {
      uint64_t expect = ~0ull;
      if (__atomic_compare_exchange_n(&M->keys[s], &expect, fp, 0, __ATOMIC_RELAXED,
                                      __ATOMIC_RELAXED) || expect == fp)
        return s;
    }
    s = (s + 1) & M->mask;
  }
}

static uint64_t mt_find(const mt_level *M, uint64_t fp) {
  uint64_t s = orc_fmix64(fp) & M->mask;
  for (;;) {
    uint64_t k = M->keys[s];
    if (k == fp) return s;
    if (k == ~0ull) return ~0ull;
    s = (s + 1) & M->mask;
  }
}

static void *mt_worker(void *arg) {
  mt_level *M = (mt_level *)arg;
  pin_self(atomic_fetch_add(&M->next_worker, 1));
  uint64_t cs[ORC_MAX_SAMPLES];
  const uint64_t nchunks = (M->total + MT_CHUNK - 1) / MT_CHUNK;
  for (;;) {
    uint64_t c = atomic_fetch_add(&M->next, 1);
    if (c >= nchunks) break;
    uint64_t i0 = c * MT_CHUNK, i1 = i0 + MT_CHUNK < M->total ? i0 + MT_CHUNK : M->total;
    if (M->pass == 1) {
      for (uint64_t idx = i0; idx < i1; ++idx) {
        int32_t l, r;
        int op;
        mt_sig(M, idx, cs, &l, &r, &op);
        uint64_t fp = orc_fingerprint(cs, M->n, M->W);
        if (fp >= ~1ull) { /* invalid key: dropped (hash_set.cuh:200-204) */
          M->slot_of[idx] = ~0u;
          continue;
        }
        uint64_t s = mt_insert(M, fp);
        M->slot_of[idx] = (uint32_t)s;
        uint64_t want = M->tag | idx, cur = __atomic_load_n(&M->comp[s], __ATOMIC_RELAXED);
        while (want < cur && !__atomic_compare_exchange_n(&M->comp[s], &cur, want, 1,
                                                          __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
        }
        if (!memcmp(cs, M->tgt, sizeof(uint64_t) * M->n)) {
          uint64_t f = atomic_load(&M->found);
          while (idx < f && !atomic_compare_exchange_weak(&M->found, &f, idx)) {
          }
        }
      }
    } else if (M->pass == 2) {
      uint64_t cnt = 0;
      for (uint64_t idx = i0; idx < i1; ++idx) {
        uint32_t s = M->slot_of[idx];
        uint8_t k = s != ~0u && M->comp[s] == (M->tag | idx);
        M->keep[idx] = k;
        cnt += k;
      }
      M->chunk_cnt[c] = cnt;
    } else {
      uint64_t row = M->chunk_cnt[c];
      for (uint64_t idx = i0; idx < i1; ++idx) {
        if (!M->keep[idx]) continue;
        int32_t l, r;
        int op;
        mt_sig(M, idx, M->out_vals + row * M->n, &l, &r, &op);
        M->out_l[row] = l;
        M->out_r[row] = r;
        M->out_op[row] = (uint8_t)op;
        ++row;
      }
    }
  }
  return NULL;
}

static void mt_run(mt_level *M, int pass) {
  M->pass = pass;
  atomic_store(&M->next, 0);
  atomic_store(&M->next_worker, 0);
  pthread_t th[256];
  int nt = M->nthreads > 256 ? 256 : M->nthreads;
  for (int k = 0; k < nt; ++k) pthread_create(&th[k], NULL, mt_worker, M);
  for (int k = 0; k < nt; ++k) pthread_join(th[k], NULL);
}

/* Set of capacity >= want slots (power of two) holding the keys (and their
 * companions) of the old set. */
static int mt_regrow(uint64_t **keys, uint64_t **comp, uint64_t *cap, uint64_t want) {
  uint64_t c = *cap ? *cap : 1024;
  while (c < want) c <<= 1;
  if (c == *cap) return 0;
  uint64_t *nk = (uint64_t *)malloc(sizeof(uint64_t) * c);
  uint64_t *nc = (uint64_t *)malloc(sizeof(uint64_t) * c);
  if (!nk || !nc) { free(nk); free(nc); return -1; }
  memset(nk, 0xff, sizeof(uint64_t) * c);
  memset(nc, 0xff, sizeof(uint64_t) * c);
  for (uint64_t i = 0; i < *cap; ++i) {
    uint64_t k = (*keys)[i];
    if (k == ~0ull) continue;
    uint64_t s = orc_fmix64(k) & (c - 1);
    while (nk[s] != ~0ull) s = (s + 1) & (c - 1);
    nk[s] = k;
    nc[s] = (*comp)[i];
  }
  free(*keys);
  free(*comp);
  *keys = nk;
  *comp = nc;
  *cap = c;
  return 0;
}

int orc_vfb_search_mt(int n_vars, int n_consts, const uint64_t *consts, int W,
                      int n_samples, const uint64_t *inputs, const uint64_t *outputs,
                      int max_len, int exhaustive, int nthreads, orc_result *res,
                      uint64_t *level_new, uint64_t *level_cand) {
  memset(res, 0, sizeof *res);
  int L = n_vars + n_consts, n = n_samples;
  if (L < 1 || L > ORC_MAX_LEAVES || n < 1 || n > ORC_MAX_SAMPLES || max_len < 1 ||
      max_len > 40)
    return -2;
  if (nthreads < 1) nthreads = 1;
  double t0 = now_s();
  const uint64_t mask = wmask(W);
  uint64_t *tgt = (uint64_t *)malloc(sizeof(uint64_t) * n);
  for (int i = 0; i < n; ++i) tgt[i] = outputs[i] & mask;
  if (level_new) memset(level_new, 0, sizeof(uint64_t) * (max_len + 1));
  if (level_cand) memset(level_cand, 0, sizeof(uint64_t) * (max_len + 1));
  vcache C = {0};
  C.n = n;
  C.W = W;
  uint64_t cs[ORC_MAX_SAMPLES];
  int64_t found_row = -1;
  int found_level = 0;
  uint64_t found_index = 0, evaluated = 0;
  for (int v = 0; v < L; ++v) { /* level 1 (full-cache.cu:694-707) */
    for (int i = 0; i < n; ++i)
      cs[i] = (v < n_vars ? inputs[i * n_vars + v] : consts[v - n_vars]) & mask;
    vc_push(&C, cs, -1, -1, 255);
    evaluated++;
    if (found_row < 0 && !memcmp(cs, tgt, sizeof(uint64_t) * n)) {
      found_row = v;
      found_level = 1;
      found_index = (uint64_t)v;
    }
  }
  if (level_new) level_new[1] = L;
  if (level_cand) level_cand[1] = L;
  uint64_t *keys = NULL, *comp = NULL, cap = 0, nkeys = 0;
  int status = ORC_NOT_FOUND, err = 0;
  if (mt_regrow(&keys, &comp, &cap, 1 << 16)) err = 1;
  for (int v = 0; v < L && !err; ++v) { /* hashSetInit (full-cache.cu:142-157) */
    uint64_t fp = orc_fingerprint(C.vals + (size_t)v * n, n, W);
    if (fp >= ~1ull) continue;
    uint64_t s = orc_fmix64(fp) & (cap - 1);
    while (keys[s] != ~0ull && keys[s] != fp) s = (s + 1) & (cap - 1);
    if (keys[s] == ~0ull) nkeys++;
    keys[s] = fp;
    comp[s] = 0;
  }
  uint64_t lvl_start[42];
  lvl_start[1] = 0;
  lvl_start[2] = C.rows;
  mt_block blk[8 * 20 + 8];
  for (int lev = 2; lev <= max_len && (found_row < 0 || exhaustive) && !err; ++lev) {
    int nb = 0;
    uint64_t total = 0;
    for (int op = 0; op < 8; ++op) {
      int unary = op == ORC_NOT || op == ORC_NEG;
      int nsplit = unary ? 1 : (lev - 1) / 2;
      for (int sp = 1; sp <= nsplit; ++sp) {
        mt_block b;
        memset(&b, 0, sizeof b);
        b.op = op;
        b.unary = unary;
        if (unary) {
          b.l0 = lvl_start[lev - 1];
          b.nl = lvl_start[lev] - b.l0;
          b.nr = 1;
        } else {
          b.l0 = lvl_start[sp];
          b.nl = lvl_start[sp + 1] - b.l0;
          b.r0 = lvl_start[lev - 1 - sp];
          b.nr = lvl_start[lev - sp] - b.r0;
        }
        b.count = b.nl * b.nr * (op == ORC_SUB ? 2 : 1);
        if (!b.count) continue;
        b.begin = total;
        total += b.count;
        blk[nb++] = b;
      }
    }
    if (level_cand) level_cand[lev] = total;
    evaluated += total;
    if (total == 0) {
      lvl_start[lev + 1] = C.rows;
      continue;
    }
    /* room for every key this level can add, at most half full */
    if (mt_regrow(&keys, &comp, &cap, 2 * (nkeys + total))) { err = 1; break; }
    mt_level M;
    memset(&M, 0, sizeof M);
    M.n = n; M.W = W; M.nblocks = nb; M.nthreads = nthreads;
    M.vals = C.vals; M.tgt = tgt; M.blk = blk; M.total = total;
    M.tag = (uint64_t)lev << 48;
    M.keys = keys; M.comp = comp; M.mask = cap - 1;
    const uint64_t nchunks = (total + MT_CHUNK - 1) / MT_CHUNK;
    M.slot_of = (uint32_t *)malloc(sizeof(uint32_t) * total);
    M.keep = (uint8_t *)malloc(total);
    M.chunk_cnt = (uint64_t *)malloc(sizeof(uint64_t) * nchunks);
    if (!M.slot_of || !M.keep || !M.chunk_cnt || cap > 0xffffffffull) {
      free(M.slot_of); free(M.keep); free(M.chunk_cnt);
      err = 1;
      break;
    }
    atomic_init(&M.found, UINT64_MAX);
    mt_run(&M, 1);
    mt_run(&M, 2);
    uint64_t kept = 0;
    for (uint64_t c = 0; c < nchunks; ++c) {
      uint64_t k = M.chunk_cnt[c];
      M.chunk_cnt[c] = C.rows + kept;
      kept += k;
    }
    /* rows of this level appended in order */
    uint64_t need = C.rows + kept;
    if (need > C.cap) {
      uint64_t nc = C.cap ? C.cap : 1024;
      while (nc < need) nc *= 2;
      C.vals = (uint64_t *)realloc(C.vals, sizeof(uint64_t) * n * nc);
      C.left = (int32_t *)realloc(C.left, sizeof(int32_t) * nc);
      C.right = (int32_t *)realloc(C.right, sizeof(int32_t) * nc);
      C.op = (uint8_t *)realloc(C.op, nc);
      C.cap = nc;
    }
    M.vals = C.vals;
    M.out_vals = C.vals;
    M.out_l = C.left;
    M.out_r = C.right;
    M.out_op = C.op;
    mt_run(&M, 3);
    uint64_t f = atomic_load(&M.found);
    if (found_row < 0 && f != UINT64_MAX) {
      uint64_t tfp = orc_fingerprint(tgt, n, W);
      uint64_t s = tfp >= ~1ull ? ~0ull : mt_find(&M, tfp);
      if (s != ~0ull && comp[s] == (M.tag | f)) {
        uint64_t c = f / MT_CHUNK, row = M.chunk_cnt[c];
        for (uint64_t i = c * MT_CHUNK; i < f; ++i) row += M.keep[i];
        found_row = (int64_t)row;
        found_level = lev;
        found_index = f;
      }
    }
    nkeys += kept;
    C.rows += kept;
    if (level_new) level_new[lev] = kept;
    lvl_start[lev + 1] = C.rows;
    free(M.slot_of);
    free(M.keep);
    free(M.chunk_cnt);
  }
  if (err) {
    status = ORC_OVERFLOW;
    strcpy(res->expr, "Not found !");
  } else if (found_row >= 0) {
    status = ORC_FOUND;
    res->size = found_level;
    res->rank = found_index;
    int len = vfb_rpn(&C, L, found_row, res->rpn, 0, ORC_MAX_RPN);
    if (len < 0) {
      strcpy(res->expr, "Size of the output is too big !");
      res->op = -1;
    } else {
      res->op = found_level > 1 ? res->rpn[len - 1] - ORC_TOK_OP : -1;
      orc_infix(res->rpn, len, n_vars, consts, W, 1, res->expr, sizeof res->expr);
    }
  } else {
    strcpy(res->expr, "Not found !");
  }
  res->status = status;
  res->candidates = evaluated;
  res->rows = C.rows;
  res->seconds = now_s() - t0;
  free(C.vals);
  free(C.left);
  free(C.right);
  free(C.op);
  free(keys);
  free(comp);
  free(tgt);
  return status;
}
