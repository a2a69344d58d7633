// mlp.hip — the cutoff surrogate network on BF16 MFMA (gfx950).
//
// Network 6 -> 128 -> 128 -> 1 with GELU, BF16 weights and activations, FP32
// accumulation (v_mfma_f32_16x16x16_bf16).  Inference: one workgroup (4
// wavefronts) per 64 samples, wave w owning rows 16w..16w+15 of every
// activation tile.  Training: one workgroup per 16 samples, wave w owning
// hidden units 32w..32w+31, forward, loss and backward fused in one kernel
// that writes a per-workgroup gradient partial; a second kernel reduces the
// partials in a fixed order, applies Adam to the
// FP32 master weights, refreshes the BF16 copies (W2 is kept in both
// orientations so every MFMA operand is a contiguous 64-bit load) and
// advances the step count and batch offset, kept in two device-memory slots
// that consecutive steps use in turn.  Two training steps are captured in one
// hipGraph.
//
// Parameter layout: W1[128][6] | b1[128] | W2[128][128] | b2[128] | w3[128] | b3
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "amps_common.h"

using namespace amps;

namespace {

constexpr int DIN = 6, HID = 128, TB = 64;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + HID * DIN, OFF_W2 = OFF_B1 + HID, OFF_B2 = OFF_W2 + HID * HID,
              OFF_W3 = OFF_B2 + HID, OFF_B3 = OFF_W3 + HID, NPARAM = OFF_B3 + 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * __expf(-0.5f * z * z) * 0.39894228040143268f;
}
__device__ __forceinline__ bf16x4 ld4(const unsigned short* p) { return *(const bf16x4*)p; }
__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float sum16(float x) {  // over lanes sharing lane/16
  x += __shfl_xor(x, 1);
  x += __shfl_xor(x, 2);
  x += __shfl_xor(x, 4);
  x += __shfl_xor(x, 8);
  return x;
}
__device__ __forceinline__ float sum_rows(float x) {  // over the 4 row groups (lane/16)
  x += __shfl_xor(x, 16);
  x += __shfl_xor(x, 32);
  return x;
}

struct MlpView {
  const float* theta;             // FP32 master parameters
  const unsigned short* W1b;      // [128][16] bf16, k >= 6 zero
  const unsigned short* W2b;      // [128][128] bf16 (n, k)
  const unsigned short* W2Tb;     // [128][128] bf16 (k, n)
};

// Layers 1-2 for the wave's 16 rows.  z1/z2 hold pre-activations in the MFMA
// D layout (row 16w + 4(l/16) + v, column 16 nt + l%16).
__device__ __forceinline__ void forward_rows(const MlpView& mv, const unsigned short* Xb,
                                             unsigned short* H1b, unsigned short* H1Tb, int w,
                                             int l, float z1[8][4], float z2[8][4],
                                             float h2[8][4]) {
  const int r0 = 16 * w, rl = l & 15, kq = 4 * (l >> 4);
  const bf16x4 ax = ld4(Xb + (r0 + rl) * 16 + kq);
#pragma unroll
  for (int nt = 0; nt < 8; nt++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(ax, ld4(mv.W1b + (16 * nt + rl) * 16 + kq), acc);
    const float b1 = mv.theta[OFF_B1 + 16 * nt + rl];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      z1[nt][v] = acc[v] + b1;
      const unsigned short h = f2bf(gelu(z1[nt][v]));
      const int m = r0 + kq + v, n = 16 * nt + rl;
      H1b[m * HID + n] = h;
      H1Tb[n * TB + m] = h;
    }
  }
  __syncthreads();
#pragma unroll
  for (int nt = 0; nt < 8; nt++) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 8; kt++)
      acc = mfma(ld4(H1b + (r0 + rl) * HID + 16 * kt + kq), ld4(mv.W2b + (16 * nt + rl) * HID + 16 * kt + kq), acc);
    const float b2 = mv.theta[OFF_B2 + 16 * nt + rl];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      z2[nt][v] = acc[v] + b2;
      h2[nt][v] = bf2f(f2bf(gelu(z2[nt][v])));
    }
  }
}

__device__ __forceinline__ void load_x(const float* X, int64_t row0, int64_t n, unsigned short* Xb,
                                       unsigned short* XTb) {
  for (int i = threadIdx.x; i < TB * 16; i += blockDim.x) {
    const int m = i >> 4, k = i & 15;
    const int64_t r = row0 + m;
    const float val = (k < DIN && r < n) ? X[r * DIN + k] : 0.0f;
    const unsigned short h = f2bf(val);
    Xb[m * 16 + k] = h;
    XTb[k * TB + m] = h;
  }
}

__global__ __launch_bounds__(256) void k_mlp_forward(MlpView mv, const float* __restrict__ X, int64_t n,
                                                      float* __restrict__ y) {
  __shared__ unsigned short Xb[TB * 16], XTb[16 * TB], H1b[TB * HID], H1Tb[HID * TB];
  const int64_t row0 = (int64_t)blockIdx.x * TB;
  load_x(X, row0, n, Xb, XTb);
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  float z1[8][4], z2[8][4], h2[8][4];
  forward_rows(mv, Xb, H1b, H1Tb, w, l, z1, z2, h2);
  const float b3 = mv.theta[OFF_B3];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    float s = 0.f;
#pragma unroll
    for (int nt = 0; nt < 8; nt++) s += h2[nt][v] * mv.theta[OFF_W3 + 16 * nt + (l & 15)];
    s = sum16(s);
    const int64_t r = row0 + 16 * w + 4 * (l >> 4) + v;
    if ((l & 15) == 0 && r < n) y[r] = s + b3;
  }
}

// Fused forward + MSE loss + backward for TS = 16 rows of the batch starting
// at rows [*batch_off, ...) of the data (wrapping at n), spread over the
// columns: wavefront w owns hidden units 32w .. 32w + 31 of both layers (its
// two 16-wide column tiles), so a workgroup's chain of dependent products is
// a quarter of what one wavefront per 16 rows would take, and a 4,096-sample
// batch is 256 workgroups.  Writes this workgroup's gradient partial (NPARAM
// floats) and loss partial.  LDS rows are padded (HP, TP) off the bank
// period; every MFMA operand stays an aligned 8-byte read.
constexpr int TS = 16;               // samples per training workgroup
constexpr int HP = HID + 4;          // row stride of [TS][HID] tiles
constexpr int TP = TS + 4;           // row stride of [HID][TS] and [16][TS] tiles

__global__ __launch_bounds__(256) void k_mlp_train(MlpView mv, const float* __restrict__ X,
                                                    const float* __restrict__ T, int64_t n, int batch,
                                                    const int64_t* __restrict__ batch_off,
                                                    float* __restrict__ gpart,
                                                    float* __restrict__ lpart) {
  __shared__ __align__(16) unsigned short Xb[TS * 16], XTb[16 * TP];
  __shared__ __align__(16) unsigned short H1b[TS * HP], H1Tb[HID * TP];
  __shared__ __align__(16) unsigned short D2b[TS * HP], D2Tb[HID * TP], D1Tb[HID * TP];
  __shared__ float ypart[4][TS];
  const int64_t base = (batch_off ? *batch_off : 0) + (int64_t)blockIdx.x * TS;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, rl = l & 15, kq = 4 * (l >> 4);
  // everything this lane reads from the weights and targets, loaded up front
  // so that the load latency is paid once, not once per phase
  bf16x4 bw1[2], bw2[2][8], bw2t[2][8];
  float pb1[2], pb2[2], pw3[2], tv[4];
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int nn = 16 * (2 * w + t) + rl;
    bw1[t] = ld4(mv.W1b + nn * 16 + kq);
#pragma unroll
    for (int kt = 0; kt < 8; kt++) {
      bw2[t][kt] = ld4(mv.W2b + nn * HID + 16 * kt + kq);
      bw2t[t][kt] = ld4(mv.W2Tb + nn * HID + 16 * kt + kq);
    }
    pb1[t] = mv.theta[OFF_B1 + nn];
    pb2[t] = mv.theta[OFF_B2 + nn];
    pw3[t] = mv.theta[OFF_W3 + nn];
  }
  const float b3 = mv.theta[OFF_B3];
#pragma unroll
  for (int v = 0; v < 4; v++) tv[v] = T[(base + kq + v) % n];
  for (int i = threadIdx.x; i < TS * 16; i += blockDim.x) {
    const int m = i >> 4, k = i & 15;
    const int64_t r = (base + m) % n;
    const unsigned short h = f2bf((k < DIN) ? X[r * DIN + k] : 0.0f);
    Xb[m * 16 + k] = h;
    XTb[k * TP + m] = h;
  }
  __syncthreads();
  float* g = gpart + (int64_t)blockIdx.x * NPARAM;
  // layer 1 for this wave's two column tiles; the D layout: row kq + v, column 16 nt + rl
  float z1[2][4], p1[2][4];  // pre-activation and Phi(z) = (1 + erf(z / sqrt 2)) / 2
  const bf16x4 ax = ld4(Xb + rl * 16 + kq);
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int nt = 2 * w + t, nn = 16 * nt + rl;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(ax, bw1[t], acc);
    const float b1 = pb1[t];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const float z = acc[v] + b1;
      const float ph = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
      z1[t][v] = z;
      p1[t][v] = ph;
      const unsigned short h = f2bf(z * ph);
      H1b[(kq + v) * HP + nn] = h;
      H1Tb[nn * TP + kq + v] = h;
    }
  }
  __syncthreads();
  // layer 2 for the same columns, and the output layer's partial sums
  float z2[2][4], p2[2][4], h2[2][4];
  float ys[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int nt = 2 * w + t, nn = 16 * nt + rl;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 8; kt++)
      acc = mfma(ld4(H1b + rl * HP + 16 * kt + kq), bw2[t][kt], acc);
    const float b2 = pb2[t], w3 = pw3[t];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const float z = acc[v] + b2;
      const float ph = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
      z2[t][v] = z;
      p2[t][v] = ph;
      h2[t][v] = bf2f(f2bf(z * ph));
      ys[v] += h2[t][v] * w3;
    }
  }
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const float sv = sum16(ys[v]);
    if (rl == 0) ypart[w][kq + v] = sv;
  }
  __syncthreads();
  // the loss gradient of this lane's rows (every wave computes the same)
  float dy[4];
  float lsum = 0.f;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int m = kq + v;
    const float y = (((ypart[0][m] + ypart[1][m]) + ypart[2][m]) + ypart[3][m]) + b3;
    const float res = y - tv[v];
    dy[v] = 2.0f * res / (float)batch;
    lsum += res * res;
  }
  // dZ2 = dy w3 gelu'(z2) on this wave's columns; their b2 and w3 gradients
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int nn = 16 * (2 * w + t) + rl;
    const float w3 = pw3[t];
    float sb = 0.f, sw = 0.f;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const float z = z2[t][v];
      const float gg = p2[t][v] + z * __expf(-0.5f * z * z) * 0.39894228040143268f;
      const float d = dy[v] * w3 * gg;
      const unsigned short h = f2bf(d);
      D2b[(kq + v) * HP + nn] = h;
      D2Tb[nn * TP + kq + v] = h;
      sb += d;
      sw += dy[v] * h2[t][v];
    }
    sb = sum_rows(sb);
    sw = sum_rows(sw);
    if (l < 16) {
      g[OFF_B2 + nn] = sb;
      g[OFF_W3 + nn] = sw;
    }
  }
  if (w == 0) {
    float ds = dy[0] + dy[1] + dy[2] + dy[3];
    float ls = lsum;
    ds = sum_rows(ds);
    ls = sum_rows(ls);
    if (l == 0) {
      g[OFF_B3] = ds;
      lpart[blockIdx.x] = ls;
    }
  }
  __syncthreads();
  // dW2[n][k] = sum_m dZ2[m][n] H1[m][k] for this wave's rows n (one
  // 16-sample step per tile); dH1 = dZ2 W2 for this wave's columns k
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int it = 2 * w + t;
    const bf16x4 a = ld4(D2Tb + (16 * it + rl) * TP + kq);
#pragma unroll
    for (int jt = 0; jt < 8; jt++) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma(a, ld4(H1Tb + (16 * jt + rl) * TP + kq), acc);
#pragma unroll
      for (int v = 0; v < 4; v++) g[OFF_W2 + (16 * it + kq + v) * HID + 16 * jt + rl] = acc[v];
    }
  }
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int jt = 2 * w + t, kk = 16 * jt + rl;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 8; ks++)
      acc = mfma(ld4(D2b + rl * HP + 16 * ks + kq), bw2t[t][ks], acc);
    // dZ1 = dH1 gelu'(z1), the same columns as this wave's layer 1
    float sb = 0.f;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const float z = z1[t][v];
      const float gg = p1[t][v] + z * __expf(-0.5f * z * z) * 0.39894228040143268f;
      const float d = acc[v] * gg;
      D1Tb[kk * TP + kq + v] = f2bf(d);
      sb += d;
    }
    sb = sum_rows(sb);
    if (l < 16) g[OFF_B1 + kk] = sb;
  }
  // dW1[n][j] = sum_m dZ1[m][n] X[m][j]: the wave's own rows of D1Tb (the
  // LDS operations of one wavefront execute in issue order, so a compiler
  // barrier is all the ordering needed)
  asm volatile("" ::: "memory");
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int it = 2 * w + t;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(ld4(D1Tb + (16 * it + rl) * TP + kq), ld4(XTb + rl * TP + kq), acc);
    if (rl < DIN) {
#pragma unroll
      for (int v = 0; v < 4; v++) g[OFF_W1 + (16 * it + kq + v) * DIN + rl] = acc[v];
    }
  }
}

// Sum the workgroup partials in a fixed order, Adam on the FP32 master copy,
// BF16 refresh.  A workgroup takes 64 consecutive parameters: wavefront w of
// its 16 sums rows [w r, (w + 1) r) of the partials (r = nwg / 16 rounded
// up) for the 64 parameters, one coalesced 256-byte row per load, and the
// sixteen sums are added in order.  The step count and batch offset are
// read from ctr_in and the advanced ones written to ctr_out by the first
// workgroup: two slots used in turn by consecutive steps, so no workgroup
// reads what another is writing.
constexpr int ADAM_PARAMS = 64, ADAM_WAVES = 16;

__global__ __launch_bounds__(64 * ADAM_WAVES) void k_mlp_adam(const float* __restrict__ gpart, int nwg, float* theta,
                                                   float* m, float* v, float* grad_out,
                                                   const int64_t* __restrict__ ctr_in, int64_t* ctr_out,
                                                   int64_t batch, int64_t nrows, float lr, int apply,
                                                   unsigned short* W1b, unsigned short* W2b, unsigned short* W2Tb) {
  __shared__ float part[ADAM_WAVES][ADAM_PARAMS];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int j = blockIdx.x * ADAM_PARAMS + l;
  const int64_t step = ctr_in[0] + 1;
  const int rows = (nwg + ADAM_WAVES - 1) / ADAM_WAVES, k0 = w * rows, k1 = min(nwg, k0 + rows);
  float gsum = 0.f;
  if (j < NPARAM) {
#pragma unroll 8
    for (int k = k0; k < k1; k++) gsum += gpart[(int64_t)k * NPARAM + j];
  }
  part[w][l] = gsum;
  __syncthreads();
  if (w == 0 && j < NPARAM) {
    float gt = part[0][l];
#pragma unroll
    for (int q = 1; q < ADAM_WAVES; q++) gt += part[q][l];
    if (grad_out) grad_out[j] = gt;
    if (apply) {
      const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
      const float mj = b1 * m[j] + (1.f - b1) * gt;
      const float vj = b2 * v[j] + (1.f - b2) * gt * gt;
      m[j] = mj;
      v[j] = vj;
      const float mh = mj / (1.f - powf(b1, (float)step));
      const float vh = vj / (1.f - powf(b2, (float)step));
      const float th = theta[j] - lr * mh / (sqrtf(vh) + eps);
      theta[j] = th;
      const unsigned short hb = f2bf(th);
      if (j < OFF_B1) {
        W1b[(j / DIN) * 16 + (j % DIN)] = hb;
      } else if (j >= OFF_W2 && j < OFF_B2) {
        const int q = j - OFF_W2, nn = q / HID, kk = q % HID;
        W2b[nn * HID + kk] = hb;
        W2Tb[kk * HID + nn] = hb;
      }
    }
  }
  if (apply && blockIdx.x == 0 && threadIdx.x == 0) {
    ctr_out[0] = step;
    ctr_out[1] = (ctr_in[1] + batch) % nrows;
  }
}

__global__ __launch_bounds__(256) void k_mlp_refresh_bf16(const float* theta, unsigned short* W1b,
                                                           unsigned short* W2b, unsigned short* W2Tb) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < HID * 16) {
    const int nn = j / 16, kk = j % 16;
    W1b[j] = (kk < DIN) ? f2bf(theta[OFF_W1 + nn * DIN + kk]) : (unsigned short)0;
  }
  if (j < HID * HID) {
    const int nn = j / HID, kk = j % HID;
    const unsigned short h = f2bf(theta[OFF_W2 + j]);
    W2b[j] = h;
    W2Tb[kk * HID + nn] = h;
  }
}

}  // namespace

// =====================================================================
// host side
// =====================================================================
struct MlpState {
  float* theta = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  unsigned short *W1b = nullptr, *W2b = nullptr, *W2Tb = nullptr;
  float *X = nullptr, *T = nullptr;
  int64_t n = 0, cap = 0;
  float* gpart = nullptr;
  float* lpart = nullptr;
  int64_t gpart_wg = 0;
  // two slots of (Adam step count, batch offset), read and written in turn by
  // consecutive steps; slot 0 is the current one between calls
  int64_t* step_dev = nullptr;
  float* grad = nullptr;
  float* qX = nullptr;
  float* qy = nullptr;
  int64_t qcap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_batch = 0;
  float graph_lr = 0.f;
};

namespace amps {
void mlp_free(amps_ctx* c) {
  MlpState* s = (MlpState*)c->mlp_state;
  if (!s) return;
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  for (void* p : {(void*)s->theta, (void*)s->m, (void*)s->v, (void*)s->W1b, (void*)s->W2b,
                  (void*)s->W2Tb, (void*)s->X, (void*)s->T, (void*)s->gpart, (void*)s->lpart,
                  (void*)s->step_dev, (void*)s->grad, (void*)s->qX, (void*)s->qy})
    (void)hipFree(p);
  delete s;
  c->mlp_state = nullptr;
}
}  // namespace amps

static MlpView view(MlpState* s) { return MlpView{s->theta, s->W1b, s->W2b, s->W2Tb}; }

// The captured training step holds the data and partial buffers' addresses
// and the row count: it is dropped whenever one of them changes.
static void drop_graph(MlpState* s) {
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->gexec = nullptr;
  s->graph = nullptr;
}

static int ensure_gpart(amps_ctx* c, MlpState* s, int64_t nwg) {
  if (nwg <= s->gpart_wg) return 0;
  drop_graph(s);
  (void)hipFree(s->gpart);
  (void)hipFree(s->lpart);
  s->gpart = nullptr;
  s->lpart = nullptr;
  HIP_TRY(hipMalloc(&s->gpart, sizeof(float) * NPARAM * nwg));
  HIP_TRY(hipMalloc(&s->lpart, sizeof(float) * nwg));
  s->gpart_wg = nwg;
  (void)c;
  return 0;
}

extern "C" int amps_gpu_mlp_init(amps_ctx* c, const float* theta) {
  if (!c || !theta) return fail(AMPS_E_ARG, "mlp_init: null argument");
  (void)hipSetDevice(c->device);
  MlpState* s = (MlpState*)c->mlp_state;
  if (!s) {
    s = new MlpState();
    c->mlp_state = s;
    HIP_TRY(hipMalloc(&s->theta, sizeof(float) * NPARAM));
    HIP_TRY(hipMalloc(&s->m, sizeof(float) * NPARAM));
    HIP_TRY(hipMalloc(&s->v, sizeof(float) * NPARAM));
    HIP_TRY(hipMalloc(&s->grad, sizeof(float) * NPARAM));
    HIP_TRY(hipMalloc(&s->W1b, 2 * HID * 16));
    HIP_TRY(hipMalloc(&s->W2b, 2 * HID * HID));
    HIP_TRY(hipMalloc(&s->W2Tb, 2 * HID * HID));
    HIP_TRY(hipMalloc(&s->step_dev, sizeof(int64_t) * 4));
  }
  HIP_TRY(hipMemcpyAsync(s->theta, theta, sizeof(float) * NPARAM, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemsetAsync(s->m, 0, sizeof(float) * NPARAM, c->stream));
  HIP_TRY(hipMemsetAsync(s->v, 0, sizeof(float) * NPARAM, c->stream));
  HIP_TRY(hipMemsetAsync(s->step_dev, 0, sizeof(int64_t) * 4, c->stream));
  hipLaunchKernelGGL(k_mlp_refresh_bf16, dim3((HID * HID + 255) / 256), dim3(256), 0, c->stream, s->theta,
                     s->W1b, s->W2b, s->W2Tb);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AMPS_OK;
}

static MlpState* mstate(amps_ctx* c) { return c ? (MlpState*)c->mlp_state : nullptr; }

extern "C" int amps_gpu_mlp_get_params(amps_ctx* c, float* theta, int64_t* step) {
  MlpState* s = mstate(c);
  if (!s) return fail(AMPS_E_STATE, "mlp: call amps_gpu_mlp_init first");
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (theta) HIP_TRY(hipMemcpy(theta, s->theta, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  if (step) HIP_TRY(hipMemcpy(step, s->step_dev, sizeof(int64_t), hipMemcpyDeviceToHost));
  return AMPS_OK;
}

extern "C" int amps_gpu_mlp_set_data(amps_ctx* c, int64_t n, const float* X, const float* T) {
  MlpState* s = mstate(c);
  if (!s) return fail(AMPS_E_STATE, "mlp: call amps_gpu_mlp_init first");
  if (n <= 0 || !X || !T) return fail(AMPS_E_ARG, "mlp_set_data: empty data");
  (void)hipSetDevice(c->device);
  HIP_TRY(hipStreamSynchronize(c->stream));
  drop_graph(s);
  if (n > s->cap) {
    (void)hipFree(s->X);
    (void)hipFree(s->T);
    s->X = nullptr;
    s->T = nullptr;
    HIP_TRY(hipMalloc(&s->X, sizeof(float) * DIN * n));
    HIP_TRY(hipMalloc(&s->T, sizeof(float) * n));
    s->cap = n;
  }
  HIP_TRY(hipMemcpyAsync(s->X, X, sizeof(float) * DIN * n, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(s->T, T, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
  const int64_t zero = 0;
  HIP_TRY(hipMemcpyAsync(s->step_dev + 1, &zero, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  s->n = n;
  return AMPS_OK;
}

constexpr int ADAM_BLOCKS = (NPARAM + ADAM_PARAMS - 1) / ADAM_PARAMS;

// one Adam step reading the counters of slot `from` and writing the other
static void enqueue_step(amps_ctx* c, MlpState* s, int batch, float lr, int from) {
  const int nwg = batch / TS;
  const int64_t* in = s->step_dev + 2 * from;
  int64_t* out = s->step_dev + 2 * (1 - from);
  hipLaunchKernelGGL(k_mlp_train, dim3(nwg), dim3(256), 0, c->stream, view(s), s->X, s->T, s->n,
                     batch, in + 1, s->gpart, s->lpart);
  hipLaunchKernelGGL(k_mlp_adam, dim3(ADAM_BLOCKS), dim3(64 * ADAM_WAVES), 0, c->stream, s->gpart, nwg, s->theta,
                     s->m, s->v, (float*)nullptr, in, out, (int64_t)batch, s->n, lr, 1, s->W1b, s->W2b, s->W2Tb);
}

// Gradient of the mean-squared error over the first n rows of the resident
// data (n a multiple of 64), optionally followed by one Adam step.
extern "C" int amps_gpu_mlp_grad(amps_ctx* c, int64_t n, float* grad_out, float* loss_out) {
  MlpState* s = mstate(c);
  if (!s || s->n == 0) return fail(AMPS_E_STATE, "mlp_grad: init and set_data first");
  if (n <= 0 || n % TB || n > s->n || n > (1 << 24)) return fail(AMPS_E_ARG, "mlp_grad: n must be a positive multiple of 64 <= data size");
  const int nwg = (int)(n / TS);
  if (int r = ensure_gpart(c, s, nwg)) return r;
  int64_t zero = 0;
  HIP_TRY(hipMemcpyAsync(s->step_dev + 1, &zero, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_mlp_train, dim3(nwg), dim3(256), 0, c->stream, view(s), s->X, s->T, s->n, (int)n,
                     s->step_dev + 1, s->gpart, s->lpart);
  hipLaunchKernelGGL(k_mlp_adam, dim3(ADAM_BLOCKS), dim3(64 * ADAM_WAVES), 0, c->stream, s->gpart, nwg, s->theta,
                     s->m, s->v, s->grad, s->step_dev, s->step_dev + 2, (int64_t)0, s->n, 0.f, 0, s->W1b, s->W2b,
                     s->W2Tb);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (grad_out) HIP_TRY(hipMemcpy(grad_out, s->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  if (loss_out) {
    std::vector<float> lp((size_t)nwg);
    HIP_TRY(hipMemcpy(lp.data(), s->lpart, sizeof(float) * nwg, hipMemcpyDeviceToHost));
    double L = 0;
    for (float x : lp) L += x;
    *loss_out = (float)(L / (double)n);
  }
  return AMPS_OK;
}

// nsteps Adam steps on consecutive batches of the resident data (wrapping),
// replayed from one captured hipGraph.  loss_out[nsteps] (may be NULL)
// receives each step's batch loss.
extern "C" int amps_gpu_mlp_train(amps_ctx* c, int64_t nsteps, int32_t batch, double lr, float* loss_out) {
  MlpState* s = mstate(c);
  if (!s || s->n == 0) return fail(AMPS_E_STATE, "mlp_train: init and set_data first");
  if (batch <= 0 || batch % TB || batch > (1 << 20)) return fail(AMPS_E_ARG, "mlp_train: batch must be a positive multiple of 64");
  if (nsteps < 0) return fail(AMPS_E_ARG, "mlp_train: nsteps < 0");
  (void)hipSetDevice(c->device);
  const int nwg = batch / TS;
  if (int r = ensure_gpart(c, s, nwg)) return r;
  if (!s->gexec || s->graph_batch != batch || s->graph_lr != (float)lr) {
    drop_graph(s);
    // two steps, slot 0 -> 1 -> 0: a replay leaves the counters in slot 0
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(c, s, batch, (float)lr, 0);
    enqueue_step(c, s, batch, (float)lr, 1);
    HIP_TRY(hipStreamEndCapture(c->stream, &s->graph));
    HIP_TRY(hipGraphInstantiate(&s->gexec, s->graph, nullptr, nullptr, 0));
    s->graph_batch = batch;
    s->graph_lr = (float)lr;
  }
  std::vector<float> lp;
  tbegin(c);
  if (loss_out) {
    // one step at a time, each batch loss read back
    lp.resize((size_t)nwg);
    for (int64_t i = 0; i < nsteps; i++) {
      enqueue_step(c, s, batch, (float)lr, (int)(i & 1));
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(lp.data(), s->lpart, sizeof(float) * nwg, hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
      double Ls = 0;
      for (float x : lp) Ls += x;
      loss_out[i] = (float)(Ls / batch);
    }
  } else {
    for (int64_t i = 0; i + 1 < nsteps; i += 2) HIP_TRY(hipGraphLaunch(s->gexec, c->stream));
    if (nsteps & 1) {
      enqueue_step(c, s, batch, (float)lr, 0);
      HIP_TRY(hipGetLastError());
    }
  }
  if (nsteps & 1)  // the counters back in slot 0
    HIP_TRY(hipMemcpyAsync(s->step_dev, s->step_dev + 2, sizeof(int64_t) * 2, hipMemcpyDeviceToDevice, c->stream));
  tend(c, AMPS_K_MLP);
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AMPS_OK;
}

extern "C" int amps_gpu_mlp_predict(amps_ctx* c, int64_t n, const float* X, float* y) {
  MlpState* s = mstate(c);
  if (!s) return fail(AMPS_E_STATE, "mlp: call amps_gpu_mlp_init first");
  if (n < 0 || (n > 0 && (!X || !y))) return fail(AMPS_E_ARG, "mlp_predict: bad arguments");
  if (n == 0) return AMPS_OK;
  (void)hipSetDevice(c->device);
  if (n > s->qcap) {
    (void)hipFree(s->qX);
    (void)hipFree(s->qy);
    s->qX = nullptr;
    s->qy = nullptr;
    HIP_TRY(hipMalloc(&s->qX, sizeof(float) * DIN * n));
    HIP_TRY(hipMalloc(&s->qy, sizeof(float) * n));
    s->qcap = n;
  }
  HIP_TRY(hipMemcpyAsync(s->qX, X, sizeof(float) * DIN * n, hipMemcpyHostToDevice, c->stream));
  const int64_t nwg = (n + TB - 1) / TB;
  tbegin(c);
  hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)nwg), dim3(256), 0, c->stream, view(s), s->qX, n, s->qy);
  HIP_TRY(hipGetLastError());
  tend(c, AMPS_K_MLP);
  HIP_TRY(hipMemcpyAsync(y, s->qy, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AMPS_OK;
}
