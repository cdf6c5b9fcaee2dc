// mlp_host.cpp -- learned block proposal: instance features, the MLP object
// (device weights, Adam state, forward/backward on gfx950) and the learned
// search entry points.
//
// The network scores the top-level (root operator, left-child size) blocks
// of a size level; the search visits the binary blocks of each size in
// descending score and stops at the first verified solution of that size,
// so the solution has the minimal size (the plain search's size); sizes
// without a solution are swept completely, as in the plain search.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "context.h"
#include "mlp_kernels.h"
#include "simba_host.h"

struct mba_mlp {
  mba_ctx *owner = nullptr;
  mba::DevBuf params, grads, adam_m, adam_v, acts, labels;
  int cap = 0;   // batch capacity of `acts`
  int step = 0;
};

namespace mba {

namespace {

constexpr size_t kW1 = (size_t)kMlpIn * kMlpHidden, kB1 = kMlpHidden,
                 kW2 = (size_t)kMlpHidden * kMlpOut, kB2 = kMlpOut;
constexpr size_t kParams = kW1 + kB1 + kW2 + kB2;

struct Views {
  float *w1, *b1, *w2, *b2;
};

Views views(float *base) {
  return {base, base + kW1, base + kW1 + kB1, base + kW1 + kB1 + kW2};
}

// activations: X [cap x 256], H [cap x 512], Z [cap x 128], dZ [cap x 128],
// dH [cap x 512], loss rows [cap]
struct Acts {
  float *X, *H, *Z, *dZ, *dH, *loss;
};

size_t acts_floats(int cap) {
  return (size_t)cap * (kMlpIn + kMlpHidden + kMlpOut + kMlpOut + kMlpHidden + 1);
}

Acts acts_of(float *b, int cap) {
  Acts a;
  a.X = b;
  a.H = a.X + (size_t)cap * kMlpIn;
  a.Z = a.H + (size_t)cap * kMlpHidden;
  a.dZ = a.Z + (size_t)cap * kMlpOut;
  a.dH = a.dZ + (size_t)cap * kMlpOut;
  a.loss = a.dH + (size_t)cap * kMlpHidden;
  return a;
}

int ensure_cap(Ctx &ctx, mba_mlp *m, int rows) {
  if (rows <= m->cap) return 0;
  MBA_HIP(m->acts.ensure(acts_floats(rows) * 4));
  MBA_HIP(m->labels.ensure((size_t)rows * 4));
  m->cap = rows;
  return 0;
}

GemmArgs gemm(const float *A, const float *B, float *C, int M, int N, int K, int64_t sa_i,
              int64_t sa_k, int64_t sb_k, int64_t sb_j, int64_t ldc) {
  GemmArgs g;
  std::memset(&g, 0, sizeof g);
  g.A = A;
  g.B = B;
  g.C = C;
  g.M = M;
  g.N = N;
  g.K = K;
  g.sa_i = sa_i;
  g.sa_k = sa_k;
  g.sb_k = sb_k;
  g.sb_j = sb_j;
  g.ldc = ldc;
  return g;
}

// Forward over `rows` (multiple of 16) rows already in acts.X.
int forward_dev(Ctx &ctx, mba_mlp *m, int rows) {
  const Views P = views(m->params.as<float>());
  const Acts a = acts_of(m->acts.as<float>(), m->cap);
  GemmArgs g1 = gemm(a.X, P.w1, a.H, rows, kMlpHidden, kMlpIn, kMlpIn, 1, kMlpHidden, 1,
                     kMlpHidden);
  g1.bias = P.b1;
  g1.relu = 1;
  MBA_HIP(launch_gemm(g1, ctx.stream));
  GemmArgs g2 = gemm(a.H, P.w2, a.Z, rows, kMlpOut, kMlpHidden, kMlpHidden, 1, kMlpOut, 1,
                     kMlpOut);
  g2.bias = P.b2;
  MBA_HIP(launch_gemm(g2, ctx.stream));
  return 0;
}

// Backward of the mean cross-entropy; gradients into m->grads.
int backward_dev(Ctx &ctx, mba_mlp *m, int rows, int n_labelled) {
  const Views P = views(m->params.as<float>());
  const Views G = views(m->grads.as<float>());
  const Acts a = acts_of(m->acts.as<float>(), m->cap);
  const float inv_n = n_labelled > 0 ? 1.f / (float)n_labelled : 0.f;
  MBA_HIP(launch_softmax_ce(a.Z, m->labels.as<int32_t>(), rows, kMlpOut, inv_n, a.dZ, a.loss,
                            ctx.stream));
  // dW2 = H^T dZ ; db2 = colsum dZ
  MBA_HIP(launch_gemm(gemm(a.H, a.dZ, G.w2, kMlpHidden, kMlpOut, rows, 1, kMlpHidden, kMlpOut,
                           1, kMlpOut),
                      ctx.stream));
  MBA_HIP(launch_colsum(a.dZ, rows, kMlpOut, G.b2, ctx.stream));
  // dH = (dZ W2^T) * [H > 0]
  GemmArgs gh = gemm(a.dZ, P.w2, a.dH, rows, kMlpHidden, kMlpOut, kMlpOut, 1, 1, kMlpOut,
                     kMlpHidden);
  gh.gate = a.H;
  MBA_HIP(launch_gemm(gh, ctx.stream));
  // dW1 = X^T dH ; db1 = colsum dH
  MBA_HIP(launch_gemm(gemm(a.X, a.dH, G.w1, kMlpIn, kMlpHidden, rows, 1, kMlpIn, kMlpHidden, 1,
                           kMlpHidden),
                      ctx.stream));
  MBA_HIP(launch_colsum(a.dH, rows, kMlpHidden, G.b1, ctx.stream));
  return 0;
}

int upload_batch(Ctx &ctx, mba_mlp *m, const float *x, const int32_t *labels, int batch,
                 int &rows, int &n_labelled) {
  rows = (batch + 15) / 16 * 16;
  if (int rc = ensure_cap(ctx, m, rows)) return rc;
  const Acts a = acts_of(m->acts.as<float>(), m->cap);
  MBA_HIP(hipMemsetAsync(a.X, 0, (size_t)rows * kMlpIn * 4, ctx.stream));
  MBA_HIP(hipMemcpyAsync(a.X, x, (size_t)batch * kMlpIn * 4, hipMemcpyHostToDevice, ctx.stream));
  n_labelled = 0;
  if (labels) {
    std::vector<int32_t> lab(rows, -1);
    for (int i = 0; i < batch; ++i) {
      lab[i] = labels[i] >= 0 && labels[i] < kMlpOut ? labels[i] : -1;
      n_labelled += lab[i] >= 0;
    }
    MBA_HIP(hipMemcpyAsync(m->labels.p, lab.data(), (size_t)rows * 4, hipMemcpyHostToDevice,
                           ctx.stream));
    MBA_HIP(hipStreamSynchronize(ctx.stream));  // `lab` leaves scope
  }
  return 0;
}

uint32_t fmix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x85ebca6bu;
  x ^= x >> 13;
  x *= 0xc2b2ae35u;
  x ^= x >> 16;
  return x;
}

}  // namespace

// Features of instances [0, count) on the device into ctx.f_x (count x 256
// floats, row-major, equal to instance_features): inputs and outputs packed
// and uploaded, one k_features launch, on ctx.stream.
int features_dev(Ctx &ctx, const mba_instance *insts, size_t count, float **d_x) {
  std::vector<uint64_t> in_all, out_all;
  std::vector<FeatInst> fi(count);
  for (size_t i = 0; i < count; ++i) {
    const mba_instance &in = insts[i];
    if (!in.outputs || (in.n_vars && !in.inputs)) return MBA_ERR_ARG;
    if (in.width != 32 && in.width != 64) return MBA_ERR_LIMIT;
    fi[i] = {in_all.size(), out_all.size(), in.n_samples, in.n_vars, in.width, 0u};
    in_all.insert(in_all.end(), in.inputs, in.inputs + (size_t)in.n_samples * in.n_vars);
    out_all.insert(out_all.end(), in.outputs, in.outputs + in.n_samples);
  }
  MBA_HIP(ctx.f_in.ensure(8 * std::max<size_t>(in_all.size(), 1)));
  MBA_HIP(ctx.f_out.ensure(8 * std::max<size_t>(out_all.size(), 1)));
  MBA_HIP(ctx.f_meta.ensure(sizeof(FeatInst) * std::max<size_t>(count, 1)));
  MBA_HIP(ctx.f_x.ensure(4 * kMlpIn * std::max<size_t>(count, 1)));
  MBA_HIP(hipMemcpyAsync(ctx.f_in.p, in_all.data(), 8 * in_all.size(), hipMemcpyHostToDevice,
                         ctx.stream));
  MBA_HIP(hipMemcpyAsync(ctx.f_out.p, out_all.data(), 8 * out_all.size(), hipMemcpyHostToDevice,
                         ctx.stream));
  MBA_HIP(hipMemcpyAsync(ctx.f_meta.p, fi.data(), sizeof(FeatInst) * count,
                         hipMemcpyHostToDevice, ctx.stream));
  MBA_HIP(launch_features(ctx.f_in.as<uint64_t>(), ctx.f_out.as<uint64_t>(),
                          ctx.f_meta.as<FeatInst>(), (uint32_t)count, ctx.f_x.as<float>(),
                          ctx.stream));
  // the host vectors leave scope: finish the copies before returning
  MBA_HIP(hipStreamSynchronize(ctx.stream));
  *d_x = ctx.f_x.as<float>();
  return 0;
}

// 256 features of an instance: for every sample and bit position k, the
// pattern of (output bit, input bits of up to 5 variables) at k, hashed
// together with a bit-position group (k = 0..7 individually, then 8-15,
// 16-31, 32-63) and the variable count into one of the first 240 buckets;
// bucket counts are normalised by samples x width.  The last 16 slots are a
// one-hot of the size being ordered, left zero here (set per size by the
// search and by the training rows).
void instance_features(const mba_instance &in, float *out) {
  std::memset(out, 0, sizeof(float) * kMlpIn);
  const int W = (int)in.width, n = (int)in.n_samples, nv = (int)in.n_vars;
  const uint64_t M = W == 64 ? ~0ull : 0xffffffffull;
  for (int s = 0; s < n; ++s) {
    const uint64_t o = in.outputs[s] & M;
    for (int k = 0; k < W; ++k) {
      uint32_t p = (uint32_t)((o >> k) & 1);
      for (int v = 0; v < nv && v < 5; ++v)
        p |= (uint32_t)((in.inputs[(size_t)s * nv + v] >> k) & 1) << (v + 1);
      const uint32_t g = k < 8 ? (uint32_t)k : k < 16 ? 8u : k < 32 ? 9u : 10u;
      const uint32_t b = feat_bucket(fmix32(p * 64u + g + 1024u * (uint32_t)nv));
      out[b] += 1.f;
    }
  }
  const float scale = 8.f / (float)(n * W);
  for (int i = 0; i < kMlpIn; ++i) out[i] *= scale;
}

}  // namespace mba

using namespace mba;

extern "C" {

int mba_features(const mba_instance *inst, float *out) {
  if (!inst || !out || !inst->outputs || (inst->n_vars && !inst->inputs)) return MBA_ERR_ARG;
  if (inst->width != 32 && inst->width != 64) return MBA_ERR_LIMIT;
  instance_features(*inst, out);
  return 0;
}

int mba_mlp_create(mba_ctx *c, uint64_t seed, mba_mlp **out) {
  if (!c || !out) return MBA_ERR_ARG;
  *out = nullptr;
  mba_mlp *m = new (std::nothrow) mba_mlp;
  if (!m) return MBA_ERR_NOMEM;
  m->owner = c;
  Ctx &ctx = c->c;
  if (hipSetDevice(ctx.device) != hipSuccess || m->params.ensure(kParams * 4) != hipSuccess ||
      m->grads.ensure(kParams * 4) != hipSuccess || m->adam_m.ensure(kParams * 4) != hipSuccess ||
      m->adam_v.ensure(kParams * 4) != hipSuccess) {
    delete m;
    return MBA_ERR_HIP;
  }
  // He-uniform initialisation from SplitMix64 (host), biases zero
  std::vector<float> p(kParams, 0.f);
  uint64_t st = seed;
  auto next = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  const double l1 = std::sqrt(6.0 / kMlpIn), l2 = std::sqrt(6.0 / kMlpHidden);
  for (size_t i = 0; i < kW1; ++i) p[i] = (float)((2 * next() - 1) * l1);
  for (size_t i = 0; i < kW2; ++i) p[kW1 + kB1 + i] = (float)((2 * next() - 1) * l2);
  if (hipMemcpy(m->params.p, p.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->adam_m.p, 0, kParams * 4) != hipSuccess ||
      hipMemset(m->adam_v.p, 0, kParams * 4) != hipSuccess) {
    delete m;
    return MBA_ERR_HIP;
  }
  *out = m;
  return 0;
}

void mba_mlp_destroy(mba_mlp *m) {
  if (!m) return;
  (void)hipSetDevice(m->owner->c.device);
  (void)hipStreamSynchronize(m->owner->c.stream);
  m->params.release();
  m->grads.release();
  m->adam_m.release();
  m->adam_v.release();
  m->acts.release();
  m->labels.release();
  delete m;
}

int mba_mlp_get_weights(mba_mlp *m, float *w1, float *b1, float *w2, float *b2) {
  if (!m) return MBA_ERR_ARG;
  Ctx &ctx = m->owner->c;
  std::vector<float> p(kParams);
  MBA_HIP(hipStreamSynchronize(ctx.stream));
  MBA_HIP(hipMemcpy(p.data(), m->params.p, kParams * 4, hipMemcpyDeviceToHost));
  if (w1) std::memcpy(w1, p.data(), kW1 * 4);
  if (b1) std::memcpy(b1, p.data() + kW1, kB1 * 4);
  if (w2) std::memcpy(w2, p.data() + kW1 + kB1, kW2 * 4);
  if (b2) std::memcpy(b2, p.data() + kW1 + kB1 + kW2, kB2 * 4);
  return 0;
}

int mba_mlp_set_weights(mba_mlp *m, const float *w1, const float *b1, const float *w2,
                        const float *b2) {
  if (!m || !w1 || !b1 || !w2 || !b2) return MBA_ERR_ARG;
  Ctx &ctx = m->owner->c;
  std::vector<float> p(kParams);
  std::memcpy(p.data(), w1, kW1 * 4);
  std::memcpy(p.data() + kW1, b1, kB1 * 4);
  std::memcpy(p.data() + kW1 + kB1, w2, kW2 * 4);
  std::memcpy(p.data() + kW1 + kB1 + kW2, b2, kB2 * 4);
  MBA_HIP(hipStreamSynchronize(ctx.stream));
  MBA_HIP(hipMemcpy(m->params.p, p.data(), kParams * 4, hipMemcpyHostToDevice));
  MBA_HIP(hipMemset(m->adam_m.p, 0, kParams * 4));
  MBA_HIP(hipMemset(m->adam_v.p, 0, kParams * 4));
  m->step = 0;
  return 0;
}

int mba_mlp_forward(mba_mlp *m, const float *x, int batch, float *logits) {
  if (!m || !x || !logits || batch < 1) return MBA_ERR_ARG;
  Ctx &ctx = m->owner->c;
  MBA_HIP(hipSetDevice(ctx.device));
  int rows, nl;
  if (int rc = upload_batch(ctx, m, x, nullptr, batch, rows, nl)) return rc;
  if (int rc = forward_dev(ctx, m, rows)) return rc;
  const Acts a = acts_of(m->acts.as<float>(), m->cap);
  MBA_HIP(hipMemcpyAsync(logits, a.Z, (size_t)batch * kMlpOut * 4, hipMemcpyDeviceToHost,
                         ctx.stream));
  MBA_HIP(hipStreamSynchronize(ctx.stream));
  return 0;
}

int mba_mlp_grads(mba_mlp *m, const float *x, const int32_t *labels, int batch, float *loss,
                  float *gw1, float *gb1, float *gw2, float *gb2) {
  if (!m || !x || !labels || batch < 1) return MBA_ERR_ARG;
  Ctx &ctx = m->owner->c;
  MBA_HIP(hipSetDevice(ctx.device));
  int rows, nl;
  if (int rc = upload_batch(ctx, m, x, labels, batch, rows, nl)) return rc;
  if (int rc = forward_dev(ctx, m, rows)) return rc;
  if (int rc = backward_dev(ctx, m, rows, nl)) return rc;
  const Acts a = acts_of(m->acts.as<float>(), m->cap);
  std::vector<float> lr(rows), g(kParams);
  MBA_HIP(hipMemcpyAsync(lr.data(), a.loss, (size_t)rows * 4, hipMemcpyDeviceToHost, ctx.stream));
  MBA_HIP(hipMemcpyAsync(g.data(), m->grads.p, kParams * 4, hipMemcpyDeviceToHost, ctx.stream));
  MBA_HIP(hipStreamSynchronize(ctx.stream));
  if (loss) {
    double s = 0;
    for (int i = 0; i < rows; ++i) s += lr[i];
    *loss = nl ? (float)(s / nl) : 0.f;
  }
  if (gw1) std::memcpy(gw1, g.data(), kW1 * 4);
  if (gb1) std::memcpy(gb1, g.data() + kW1, kB1 * 4);
  if (gw2) std::memcpy(gw2, g.data() + kW1 + kB1, kW2 * 4);
  if (gb2) std::memcpy(gb2, g.data() + kW1 + kB1 + kW2, kB2 * 4);
  return 0;
}

int mba_mlp_train_step(mba_mlp *m, const float *x, const int32_t *labels, int batch, float lr,
                       float *loss) {
  if (!m || !x || !labels || batch < 1) return MBA_ERR_ARG;
  Ctx &ctx = m->owner->c;
  MBA_HIP(hipSetDevice(ctx.device));
  int rows, nl;
  if (int rc = upload_batch(ctx, m, x, labels, batch, rows, nl)) return rc;
  if (int rc = forward_dev(ctx, m, rows)) return rc;
  if (int rc = backward_dev(ctx, m, rows, nl)) return rc;
  m->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
  MBA_HIP(launch_adam(m->params.as<float>(), m->grads.as<float>(), m->adam_m.as<float>(),
                      m->adam_v.as<float>(), (int)kParams, lr, b1, b2, 1e-8f, c1, c2,
                      ctx.stream));
  if (loss) {
    const Acts a = acts_of(m->acts.as<float>(), m->cap);
    std::vector<float> lrow(rows);
    MBA_HIP(hipMemcpyAsync(lrow.data(), a.loss, (size_t)rows * 4, hipMemcpyDeviceToHost,
                           ctx.stream));
    MBA_HIP(hipStreamSynchronize(ctx.stream));
    double s = 0;
    for (int i = 0; i < rows; ++i) s += lrow[i];
    *loss = nl ? (float)(s / nl) : 0.f;
  }
  return 0;
}

// Scores of the (op, left size) blocks of one instance from its logits.
static void block_scores_from_logits(const float *logits, float *scores) {
  for (int op = 0; op < 8; ++op)
    for (int j = 0; j < 32; ++j) scores[op * 32 + j] = logits[op * 16 + (j < 15 ? j : 15)];
}

int mba_learned_solve(mba_ctx *c, mba_mlp *m, const mba_instance *inst, const mba_opts *opts,
                      mba_result *res) {
  if (!c || !m || !inst || !res) return MBA_ERR_ARG;
  mba_opts d;
  if (!opts) {
    mba_default_opts(&d);
    opts = &d;
  }
  // the scores order the blocks of sizes with at least kLearnedBlockMin
  // trees; the network runs for each such size the search reaches, with that
  // size in its input (its time is inside the search's, so time to solution
  // includes it), and not at all for a search settled below
  float scores[8 * 32];
  std::fill(scores, scores + 8 * 32, 0.f);
  float feat[kMlpIn];
  bool have_feat = false;
  const std::function<int(int)> score_size = [&](int size) -> int {
    float logits[kMlpOut];
    if (!have_feat) {
      if (int rc = mba_features(inst, feat)) return rc;
      have_feat = true;
    }
    std::fill(feat + kFeatHashed, feat + kMlpIn, 0.f);
    feat[feat_size_slot(size)] = 1.f;
    if (int rc = mba_mlp_forward(m, feat, 1, logits)) return rc;
    block_scores_from_logits(logits, scores);
    return hipSetDevice(c->c.device) == hipSuccess ? 0 : MBA_ERR_HIP;
  };
  if (hipSetDevice(c->c.device) != hipSuccess) return MBA_ERR_HIP;
  return simba_solve(c->c, *inst, *opts, *res, scores, false, &score_size);
}

int mba_learned_batch_solve(mba_ctx *c, mba_mlp *m, const mba_instance *insts, size_t count,
                            const mba_opts *opts, mba_result *results) {
  if (!c || !m || (!insts && count) || (!results && count)) return MBA_ERR_ARG;
  mba_opts d;
  if (!opts) {
    mba_default_opts(&d);
    opts = &d;
  }
  if (count == 0) return 0;
  // features once on the network's device; for each size the search orders,
  // the size's one-hot is set and the forward pass runs on every instance
  // (rows in chunks of 4096)
  Ctx &ctx = m->owner->c;
  MBA_HIP(hipSetDevice(ctx.device));
  float *d_x = nullptr;
  if (int rc = features_dev(ctx, insts, count, &d_x)) return rc;
  std::vector<float> feats(count * kMlpIn), logits(count * kMlpOut);
  MBA_HIP(hipMemcpyAsync(feats.data(), d_x, count * kMlpIn * 4, hipMemcpyDeviceToHost,
                         ctx.stream));
  MBA_HIP(hipStreamSynchronize(ctx.stream));
  const std::function<int(int, float *)> score_size = [&](int size, float *scores) -> int {
    MBA_HIP(hipSetDevice(ctx.device));
    for (size_t i = 0; i < count; ++i) {
      float *f = feats.data() + i * kMlpIn;
      std::fill(f + kFeatHashed, f + kMlpIn, 0.f);
      f[feat_size_slot(size)] = 1.f;
    }
    for (size_t b = 0; b < count; b += 4096) {
      const int nb = (int)std::min<size_t>(4096, count - b);
      const int rows = (nb + 15) / 16 * 16;
      if (int rc = ensure_cap(ctx, m, rows)) return rc;
      const Acts a = acts_of(m->acts.as<float>(), m->cap);
      MBA_HIP(hipMemsetAsync(a.X, 0, (size_t)rows * kMlpIn * 4, ctx.stream));
      MBA_HIP(hipMemcpyAsync(a.X, feats.data() + b * kMlpIn, (size_t)nb * kMlpIn * 4,
                             hipMemcpyHostToDevice, ctx.stream));
      if (int rc = forward_dev(ctx, m, rows)) return rc;
      MBA_HIP(hipMemcpyAsync(logits.data() + b * kMlpOut, a.Z, (size_t)nb * kMlpOut * 4,
                             hipMemcpyDeviceToHost, ctx.stream));
    }
    MBA_HIP(hipStreamSynchronize(ctx.stream));
    for (size_t i = 0; i < count; ++i)
      block_scores_from_logits(logits.data() + i * kMlpOut, scores + i * 256);
    return hipSetDevice(c->c.device) == hipSuccess ? 0 : MBA_ERR_HIP;
  };
  if (hipSetDevice(c->c.device) != hipSuccess) return MBA_ERR_HIP;
  return batch_solve(c->c, insts, count, *opts, results, &score_size);
}

int mba_features_batch(mba_ctx *c, const mba_instance *insts, size_t count, float *out) {
  if (!c || (!insts && count) || (!out && count)) return MBA_ERR_ARG;
  if (count == 0) return 0;
  Ctx &ctx = c->c;
  MBA_HIP(hipSetDevice(ctx.device));
  float *d_x = nullptr;
  if (int rc = features_dev(ctx, insts, count, &d_x)) return rc;
  MBA_HIP(hipMemcpyAsync(out, d_x, count * kMlpIn * 4, hipMemcpyDeviceToHost, ctx.stream));
  MBA_HIP(hipStreamSynchronize(ctx.stream));
  return 0;
}

}  // extern "C"
