// mlp.cpp -- host side of the scattering-source surrogate: parameters,
// feature normalisation and training hull, the training step (captured once
// in a hipGraph and replayed), inference, and the C ABI jur_mlp_*.
//
// Training minimises the mean squared error of the normalised target with
// Adam; the forward pass keeps activations in BF16 and pre-activations in
// FP32, the backward pass forms BF16 dz and FP32 weight gradients (split-K
// matrix-core products summed in a fixed order, so a step is deterministic).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "jur_host.h"
#include "jur_mlp.h"
#include "jur_scatter.h"
#include "jur_surrogate.h"

namespace {

constexpr long LOSS_CAP = 1 << 20;  // loss history ring (steps)

struct Mlp {
  bool init = false;
  DBuf params, wbf, grad, m, v, step, wt[MLP_NHID];
  DBuf X, Y, xb, yb, z[MLP_NHID], act[MLP_NHID], da, dz, y, dy, part, loss, nrm;
  long ntot = 0;
  int gB = 0;
  float glr = 0.f;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  MlpNorm norm;
};
Mlp& M() {
  static Mlp m;
  return m;
}

uint16_t f2bf_host(float f) { return mlp_f2bf(f); }

// Split-K factor for a K = B reduction: slabs of at least 256 rows, at most 16.
int ksplit_for(int B) {
  for (int s : {16, 8, 4, 2})
    if (B % (32 * s) == 0 && B / s >= 256) return s;
  return 1;
}

void sync_copies(hipStream_t s) {
  Mlp& m = M();
  launch_f2bf(m.params.as<float>(), m.wbf.as<uint16_t>(), MLP_NPARAM, s);
  const uint16_t* w = m.wbf.as<uint16_t>();
  launch_transpose_bf(w + MLP_W1, MLP_INP, MLP_H, m.wt[0].as<uint16_t>(), s);
  launch_transpose_bf(w + MLP_W2, MLP_H, MLP_H, m.wt[1].as<uint16_t>(), s);
  launch_transpose_bf(w + MLP_W3, MLP_H, MLP_H, m.wt[2].as<uint16_t>(), s);
  launch_transpose_bf(w + MLP_W4, MLP_H, MLP_H, m.wt[3].as<uint16_t>(), s);
}

void upload_norm(hipStream_t s) {
  Mlp& m = M();
  float* d = (float*)m.nrm.need(sizeof(MlpNorm));
  HIP_OK(hipMemcpyAsync(d, &m.norm, sizeof(MlpNorm), hipMemcpyHostToDevice, s));
}

void drop_graph();

// Sizes the step buffers for batch B; a captured training graph holds their
// addresses, so it is dropped when any of them moves.
void ensure_buffers(int B) {
  Mlp& m = M();
  DBuf* all[] = {&m.xb, &m.yb, &m.z[0], &m.z[1], &m.z[2], &m.z[3], &m.act[0], &m.act[1], &m.act[2],
                 &m.act[3], &m.da, &m.dz, &m.y, &m.dy, &m.part, &m.loss};
  void* before[16];
  for (int i = 0; i < 16; i++) before[i] = all[i]->p;
  m.xb.need((size_t)B * MLP_INP * 2);
  m.yb.need((size_t)B * 4);
  for (int l = 0; l < MLP_NHID; l++) {
    m.z[l].need((size_t)B * MLP_H * 4);
    m.act[l].need((size_t)B * MLP_H * 2);
  }
  m.da.need((size_t)B * MLP_H * 4);
  m.dz.need((size_t)B * MLP_H * 2);
  m.y.need((size_t)B * 4);
  m.dy.need((size_t)B * 4);
  const size_t slab = (size_t)MLP_H * MLP_H;
  m.part.need(std::max((size_t)16 * slab, (size_t)((B + 63) / 64) * (MLP_H + 1)) * 4);
  m.loss.need(LOSS_CAP * 4);
  for (int i = 0; i < 16; i++)
    if (all[i]->p != before[i]) {
      drop_graph();
      break;
    }
}

// Forward + backward on the batch in xb / yb, then (optionally) Adam.
void enqueue_step(int B, float lr, bool gather, bool update, hipStream_t s) {
  Mlp& m = M();
  float* P = m.params.as<float>();
  const uint16_t* W = m.wbf.as<uint16_t>();
  float* G = m.grad.as<float>();
  int* step = m.step.as<int>();
  if (gather)
    launch_gather_batch(m.X.as<uint16_t>(), m.Y.as<float>(), m.ntot, B, step, 0, m.xb.as<uint16_t>(),
                        m.yb.as<float>(), s);
  const int woff[MLP_NHID] = {MLP_W1, MLP_W2, MLP_W3, MLP_W4};
  const int boff[MLP_NHID] = {MLP_B1, MLP_B2, MLP_B3, MLP_B4};
  const int kin[MLP_NHID] = {MLP_INP, MLP_H, MLP_H, MLP_H};
  // forward
  for (int l = 0; l < MLP_NHID; l++) {
    GemmArgs g{};
    g.M = B;
    g.N = MLP_H;
    g.K = kin[l];
    g.ksplit = 1;
    g.A = l == 0 ? m.xb.as<uint16_t>() : m.act[l - 1].as<uint16_t>();
    g.sam = kin[l];
    g.sak = 1;
    g.B = W + woff[l];
    g.sbk = MLP_H;
    g.sbn = 1;
    g.C = m.z[l].as<float>();
    g.ldc = MLP_H;
    g.cslab = 0;
    g.epi = 1;
    g.bias = P + boff[l];
    g.act = m.act[l].as<uint16_t>();
    g.ldact = MLP_H;
    launch_gemm(g, s);
  }
  launch_out_fwd(m.act[3].as<uint16_t>(), W + MLP_W5, P + MLP_B5, B, m.y.as<float>(), s);
  launch_loss_grad(m.y.as<float>(), m.yb.as<float>(), B, m.dy.as<float>(), m.loss.as<float>(), step,
                   (int)(LOSS_CAP - 1), s);
  // backward
  launch_out_bwd(m.act[3].as<uint16_t>(), m.z[3].as<float>(), m.dy.as<float>(), W + MLP_W5, B, G + MLP_W5,
                 G + MLP_B5, m.dz.as<uint16_t>(), m.part.as<float>(), s);
  const int ks = ksplit_for(B);
  const long slab = (long)MLP_H * MLP_H;
  for (int l = MLP_NHID - 1; l >= 0; l--) {
    // dW_l = a_{l-1}^T dz_l  (K = B, split-K slabs summed in order)
    GemmArgs g{};
    g.M = kin[l];
    g.N = MLP_H;
    g.K = B;
    g.ksplit = ks;
    g.A = l == 0 ? m.xb.as<uint16_t>() : m.act[l - 1].as<uint16_t>();
    g.sam = 1;
    g.sak = kin[l];
    g.B = m.dz.as<uint16_t>();
    g.sbk = MLP_H;
    g.sbn = 1;
    g.C = m.part.as<float>();
    g.ldc = MLP_H;
    g.cslab = slab;
    g.epi = 0;
    launch_gemm(g, s);
    launch_splitk_sum(m.part.as<float>(), ks, slab, (long)kin[l] * MLP_H, G + woff[l], s);
    // db_l = column sums of dz_l
    launch_colsum(m.dz.as<uint16_t>(), B, MLP_H, m.part.as<float>(), s);
    launch_splitk_sum(m.part.as<float>(), (B + 63) / 64, MLP_H, MLP_H, G + boff[l], s);
    if (l > 0) {
      // da_{l-1} = dz_l W_l^T, dz_{l-1} = da silu'(z_{l-1})
      GemmArgs h{};
      h.M = B;
      h.N = MLP_H;
      h.K = MLP_H;
      h.ksplit = 1;
      h.A = m.dz.as<uint16_t>();
      h.sam = MLP_H;
      h.sak = 1;
      h.B = W + woff[l];
      h.sbk = 1;
      h.sbn = MLP_H;
      h.C = m.da.as<float>();
      h.ldc = MLP_H;
      h.cslab = 0;
      h.epi = 0;
      launch_gemm(h, s);
      launch_silu_back(m.da.as<float>(), m.z[l - 1].as<float>(), (long)B * MLP_H, m.dz.as<uint16_t>(), s);
    }
  }
  if (update) {
    launch_adam(P, G, m.m.as<float>(), m.v.as<float>(), m.wbf.as<uint16_t>(), MLP_NPARAM, step, lr, 0.9f,
                0.999f, 1e-8f, s);
    launch_step_inc(step, s);
  }
}

void drop_graph() {
  Mlp& m = M();
  if (m.exec) HIP_OK(hipGraphExecDestroy(m.exec));
  if (m.graph) HIP_OK(hipGraphDestroy(m.graph));
  m.exec = nullptr;
  m.graph = nullptr;
  m.gB = 0;
}

void ensure_init() {
  if (!M().init) jur_mlp_init(0x4A55u);
}

}  // namespace

// ---- shared with the surrogate path ----------------------------------------
const MlpNorm& mlp_norm() {
  ensure_init();
  return M().norm;
}
const float* mlp_norm_dev() {
  ensure_init();
  return M().nrm.as<float>();
}
void mlp_infer_dev(const uint16_t* x, long n, float* y, hipStream_t s) {
  ensure_init();
  Mlp& m = M();
  InferArgs a;
  a.n = (int)n;
  a.x = x;
  a.w1t = m.wt[0].as<uint16_t>();
  a.w2t = m.wt[1].as<uint16_t>();
  a.w3t = m.wt[2].as<uint16_t>();
  a.w4t = m.wt[3].as<uint16_t>();
  a.params = m.params.as<float>();
  a.wbf = m.wbf.as<uint16_t>();
  a.y = y;
  a.variant = engine_option_int("mlp_infer", 1);
  launch_mlp_infer(a, s);
}
bool mlp_trained() { return M().init && M().norm.trained; }

// ---- C ABI -------------------------------------------------------------------
extern "C" int jur_mlp_init(unsigned seed) {
  Mlp& m = M();
  hipStream_t s = engine_stream();
  drop_graph();
  std::vector<float> p(MLP_NPARAM, 0.f);
  std::mt19937 rng(seed);
  auto fill = [&](int off, int rows, int cols, int fan_in, int used_rows) {
    const float a = std::sqrt(6.f / (float)fan_in);
    std::uniform_real_distribution<float> u(-a, a);
    for (int r = 0; r < rows; r++)
      for (int c = 0; c < cols; c++) p[off + r * cols + c] = r < used_rows ? u(rng) : 0.f;
  };
  fill(MLP_W1, MLP_INP, MLP_H, MLP_IN, MLP_IN);
  fill(MLP_W2, MLP_H, MLP_H, MLP_H, MLP_H);
  fill(MLP_W3, MLP_H, MLP_H, MLP_H, MLP_H);
  fill(MLP_W4, MLP_H, MLP_H, MLP_H, MLP_H);
  fill(MLP_W5, MLP_H, 1, 2 * MLP_H, MLP_H);
  float* d = (float*)m.params.need(MLP_NPARAM * 4);
  HIP_OK(hipMemcpyAsync(d, p.data(), MLP_NPARAM * 4, hipMemcpyHostToDevice, s));
  m.wbf.need(MLP_NPARAM * 2);
  m.grad.need(MLP_NPARAM * 4);
  launch_fill_f32(m.grad.as<float>(), MLP_NPARAM, 0.f, s);
  launch_fill_f32((float*)m.m.need(MLP_NPARAM * 4), MLP_NPARAM, 0.f, s);
  launch_fill_f32((float*)m.v.need(MLP_NPARAM * 4), MLP_NPARAM, 0.f, s);
  launch_zero_int((int*)m.step.need(sizeof(int)), 1, s);
  m.wt[0].need((size_t)MLP_INP * MLP_H * 2);
  for (int l = 1; l < MLP_NHID; l++) m.wt[l].need((size_t)MLP_H * MLP_H * 2);
  m.norm = MlpNorm();
  for (int i = 0; i < MLP_IN; i++) {
    m.norm.mu[i] = 0.f;
    m.norm.sd[i] = 1.f;
    m.norm.lo[i] = -1e30f;
    m.norm.hi[i] = 1e30f;
  }
  m.norm.ymu = 0.f;
  m.norm.ysd = 1.f;
  m.norm.trained = 0;
  m.init = true;
  sync_copies(s);
  upload_norm(s);
  HIP_OK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int jur_mlp_set_data(const float* x, const float* y, long n, int normalize, unsigned seed) {
  if (n <= 0) JUR_FAIL("No training data!");
  ensure_init();
  Mlp& m = M();
  hipStream_t s = engine_stream();
  MlpNorm& N = m.norm;
  if (normalize) {
    for (int i = 0; i < MLP_IN; i++) {
      double a = 0, b = 0;
      for (long r = 0; r < n; r++) a += x[r * MLP_IN + i];
      a /= (double)n;
      for (long r = 0; r < n; r++) b += (x[r * MLP_IN + i] - a) * (x[r * MLP_IN + i] - a);
      b = std::sqrt(b / (double)n);
      N.mu[i] = (float)a;
      N.sd[i] = b > 1e-6 ? (float)b : 1.f;
    }
    double a = 0, b = 0;
    for (long r = 0; r < n; r++) a += y[r];
    a /= (double)n;
    for (long r = 0; r < n; r++) b += (y[r] - a) * (y[r] - a);
    b = std::sqrt(b / (double)n);
    N.ymu = (float)a;
    N.ysd = b > 1e-6 ? (float)b : 1.f;
  }
  std::vector<long> order(n);
  for (long r = 0; r < n; r++) order[r] = r;
  if (normalize) {
    std::mt19937_64 rng(seed);
    std::shuffle(order.begin(), order.end(), rng);
  }
  std::vector<uint16_t> xb((size_t)n * MLP_INP, 0);
  std::vector<float> yb(n);
  for (int i = 0; i < MLP_IN; i++) {
    N.lo[i] = 1e30f;
    N.hi[i] = -1e30f;
  }
  for (long k = 0; k < n; k++) {
    const long r = order[k];
    for (int i = 0; i < MLP_IN; i++) {
      const float v = (x[r * MLP_IN + i] - N.mu[i]) / N.sd[i];
      N.lo[i] = std::min(N.lo[i], v);
      N.hi[i] = std::max(N.hi[i], v);
      xb[(size_t)k * MLP_INP + i] = f2bf_host(v);
    }
    yb[k] = (y[r] - N.ymu) / N.ysd;
  }
  HIP_OK(hipMemcpyAsync(m.X.need(xb.size() * 2), xb.data(), xb.size() * 2, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(m.Y.need(yb.size() * 4), yb.data(), yb.size() * 4, hipMemcpyHostToDevice, s));
  m.ntot = n;
  upload_norm(s);
  drop_graph();
  HIP_OK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int jur_mlp_train(int steps, int batch, float lr, float* loss_hist) {
  ensure_init();
  Mlp& m = M();
  if (m.ntot <= 0) JUR_FAIL("No training data!");
  if (batch < 32 || batch % 32) JUR_FAIL("Batch size must be a multiple of 32!");
  hipStream_t s = engine_stream();
  ensure_buffers(batch);
  if (!m.exec || m.gB != batch || m.glr != lr) {
    drop_graph();
    HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    enqueue_step(batch, lr, true, true, s);
    HIP_OK(hipStreamEndCapture(s, &m.graph));
    HIP_OK(hipGraphInstantiate(&m.exec, m.graph, nullptr, nullptr, 0));
    m.gB = batch;
    m.glr = lr;
  }
  int step0 = 0;
  HIP_OK(hipMemcpyAsync(&step0, m.step.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  hipEvent_t a, b;
  HIP_OK(hipEventCreate(&a));
  HIP_OK(hipEventCreate(&b));
  HIP_OK(hipEventRecord(a, s));
  for (int i = 0; i < steps; i++) HIP_OK(hipGraphLaunch(m.exec, s));
  HIP_OK(hipEventRecord(b, s));
  HIP_OK(hipEventSynchronize(b));
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, a, b));
  HIP_OK(hipEventDestroy(a));
  HIP_OK(hipEventDestroy(b));
  engine_set_stat("train_ms", ms);
  engine_set_stat("train_steps", steps);
  sync_copies(s);
  if (loss_hist) {
    std::vector<float> all(LOSS_CAP);
    HIP_OK(hipMemcpyAsync(all.data(), m.loss.p, LOSS_CAP * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    for (int i = 0; i < steps; i++) loss_hist[i] = all[(step0 + i) % LOSS_CAP];
  }
  m.norm.trained = 1;
  upload_norm(s);
  HIP_OK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int jur_mlp_grads(const float* x, const float* y, int batch, float* grads, float* loss) {
  ensure_init();
  Mlp& m = M();
  if (batch < 32 || batch % 32) JUR_FAIL("Batch size must be a multiple of 32!");
  hipStream_t s = engine_stream();
  ensure_buffers(batch);
  const MlpNorm& N = m.norm;
  std::vector<uint16_t> xb((size_t)batch * MLP_INP, 0);
  std::vector<float> yb(batch);
  for (int r = 0; r < batch; r++) {
    for (int i = 0; i < MLP_IN; i++) xb[(size_t)r * MLP_INP + i] = f2bf_host((x[r * MLP_IN + i] - N.mu[i]) / N.sd[i]);
    yb[r] = (y[r] - N.ymu) / N.ysd;
  }
  HIP_OK(hipMemcpyAsync(m.xb.p, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(m.yb.p, yb.data(), yb.size() * 4, hipMemcpyHostToDevice, s));
  enqueue_step(batch, 0.f, false, false, s);
  HIP_OK(hipMemcpyAsync(grads, m.grad.p, MLP_NPARAM * 4, hipMemcpyDeviceToHost, s));
  int st = 0;
  HIP_OK(hipMemcpyAsync(&st, m.step.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  if (loss) HIP_OK(hipMemcpy(loss, m.loss.as<float>() + (st % LOSS_CAP), 4, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int jur_mlp_forward(const float* x, long n, float* y) {
  ensure_init();
  Mlp& m = M();
  hipStream_t s = engine_stream();
  const MlpNorm& N = m.norm;
  std::vector<uint16_t> xb((size_t)n * MLP_INP, 0);
  for (long r = 0; r < n; r++)
    for (int i = 0; i < MLP_IN; i++) xb[(size_t)r * MLP_INP + i] = f2bf_host((x[r * MLP_IN + i] - N.mu[i]) / N.sd[i]);
  static DBuf dx, dy;
  uint16_t* xd = (uint16_t*)dx.need(xb.size() * 2 + 16);
  float* yd = (float*)dy.need((size_t)n * 4 + 16);
  HIP_OK(hipMemcpyAsync(xd, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, s));
  mlp_infer_dev(xd, n, yd, s);
  HIP_OK(hipMemcpyAsync(y, yd, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  for (long r = 0; r < n; r++) y[r] = y[r] * N.ysd + N.ymu;
  return 0;
}

extern "C" int jur_mlp_get_params(float* p) {
  ensure_init();
  HIP_OK(hipMemcpy(p, M().params.p, MLP_NPARAM * 4, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int jur_mlp_set_params(const float* p) {
  ensure_init();
  hipStream_t s = engine_stream();
  HIP_OK(hipMemcpyAsync(M().params.p, p, MLP_NPARAM * 4, hipMemcpyHostToDevice, s));
  sync_copies(s);
  HIP_OK(hipStreamSynchronize(s));
  return 0;
}

extern "C" int jur_mlp_nparam(void) { return MLP_NPARAM; }

extern "C" int jur_mlp_save(const char* path) {
  ensure_init();
  std::vector<float> p(MLP_NPARAM);
  jur_mlp_get_params(p.data());
  FILE* f = fopen(path, "wb");
  if (!f) JUR_FAIL("Cannot create file!");
  const char magic[8] = {'J', 'U', 'R', 'M', 'L', 'P', '1', 0};
  bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(p.data(), 4, MLP_NPARAM, f) == (size_t)MLP_NPARAM &&
            fwrite(&M().norm, sizeof(MlpNorm), 1, f) == 1;
  fclose(f);
  if (!ok) JUR_FAIL("Error while writing!");
  return 0;
}

extern "C" int jur_mlp_load(const char* path) {
  ensure_init();
  FILE* f = fopen(path, "rb");
  if (!f) JUR_FAIL("Cannot open file!");
  char magic[8];
  std::vector<float> p(MLP_NPARAM);
  MlpNorm N;
  bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "JURMLP1", 8) == 0 &&
            fread(p.data(), 4, MLP_NPARAM, f) == (size_t)MLP_NPARAM && fread(&N, sizeof(MlpNorm), 1, f) == 1;
  fclose(f);
  if (!ok) JUR_FAIL("Not a surrogate parameter file!");
  jur_mlp_set_params(p.data());
  M().norm = N;
  upload_norm(engine_stream());
  HIP_OK(hipStreamSynchronize(engine_stream()));
  return 0;
}

extern "C" long jur_sca_collect(float* feat, float* jth, long cap) { return scatter_collected(feat, jth, cap); }
