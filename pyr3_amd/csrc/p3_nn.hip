// Training kernels of the learned denoiser on gfx950 (BF16 MFMA, FP32
// accumulate). A 3x3 convolution over NHWC bf16 activations is written as
//   im2col   col[p][t * Cin + c] = x(p's tap t, channel c)   (bf16, zero pad)
//   forward  y[p][o]  = sum_k col[p][k] W[o][k]              (GEMM, "NT")
//   dgrad    dcol[p][k] = sum_o dy[p][o] W[o][k]             (GEMM, A [P][O], B [O][K])
//   col2im   dx(q, c) = sum of dcol over the (p, t) that read q (gather)
//   wgrad    dW[o][k] = sum_p dy[p][o] col[p][k]             (GEMM, split over p)
// with the padding, stride 2, nearest 2x upsample and channel concat of
// the U-Net folded into im2col / col2im. One GEMM kernel serves all three
// products: a 64 x 64 output tile per block of four waves, K staged 32 at a
// time through LDS (either operand may be stored K-minor or K-major; the
// staging transposes), each wave a 32 x 32 quarter as 2 x 2 fragments of
//   v_mfma_f32_16x16x32_bf16: A[m][k] lane l: m = l % 16, k = 8 (l / 16) + 0..7
//                             B[k][n] lane l: n = l % 16, k = 8 (l / 16) + 0..7
//                             D[m][n] lane l: n = l % 16, m = 4 (l / 16) + 0..3
// (A and B take the same k of each lane slot, so the product is that of the
// tiles whatever order the slots sum in). Epilogues: f32 store, f32 atomic
// add (split-K), or bias + ReLU to bf16 (the forward activations).
// Elementwise kernels: features from the histogram, ReLU backward with the
// bf16 cast, bias gradients, the 1x1 head forward/backward with the MSE loss,
// and Adam.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pyr3nn.h"
#include "p3_nn_common.h"

namespace {

using p3nn::from_bf16;
using p3nn::to_bf16;

typedef __bf16 v8bf16 __attribute__((ext_vector_type(8)));
typedef float v4f32 __attribute__((ext_vector_type(4)));

constexpr int kBM = 64, kBN = 64, kBK = 32, kLdsK = kBK + 8;  // LDS row: 40 bf16 (bank spread)

// Stages rows [r0, r0 + 64) x k [k0, k0 + 32) of an operand into LDS as
// [row][k]. kKMajor = false: stored [rows][K] with leading dimension ld
// (K-minor); true: stored [K][rows]. Out-of-range rows / k read as zero. K
// and ld are multiples of 8.
template <bool kKMajor>
__device__ inline void stage(uint16_t (*S)[kLdsK], const uint16_t* G, int ld, int rows, int K,
                             int r0, int k0) {
  const int t = threadIdx.x;
  if (!kKMajor) {
    const int r = t >> 2, q = t & 3;            // 64 rows x 4 quads of 8 k
    const int gr = r0 + r, gk = k0 + q * 8;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (gr < rows && gk < K) v = *(const uint4*)(G + (size_t)gr * ld + gk);
    *(uint4*)&S[r][q * 8] = v;
  } else {
    const int k = t >> 3, q = t & 7;            // 32 k x 8 quads of 8 rows
    const int gk = k0 + k, gr = r0 + q * 8;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (gk < K && gr < rows) {
      if (gr + 8 <= rows) {
        v = *(const uint4*)(G + (size_t)gk * ld + gr);
      } else {
        uint16_t e[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = 0; gr + i < rows; ++i) e[i] = G[(size_t)gk * ld + gr + i];
        v = make_uint4(e[0] | (e[1] << 16), e[2] | (e[3] << 16), e[4] | (e[5] << 16),
                       e[6] | (e[7] << 16));
      }
    }
    const uint16_t* e = (const uint16_t*)&v;
#pragma unroll
    for (int i = 0; i < 8; ++i) S[q * 8 + i][k] = e[i];
  }
}

template <bool kAKMajor, bool kBKMajor, int kEpi>
__global__ void __launch_bounds__(256) gemm_bf16(P3nnGemm g) {
  __shared__ __align__(16) uint16_t As[kBM][kLdsK];
  __shared__ __align__(16) uint16_t Bs[kBN][kLdsK];
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  // split-K: block z covers k in [z * kc, (z + 1) * kc)
  const int kc = ((g.K + gridDim.z - 1) / gridDim.z + kBK - 1) / kBK * kBK;
  const int kb = blockIdx.z * kc, ke = min(g.K, kb + kc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  v4f32 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4f32{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = kb; k0 < ke; k0 += kBK) {
    stage<kAKMajor>(As, (const uint16_t*)g.A, g.lda, g.M, ke, m0, k0);
    stage<kBKMajor>(Bs, (const uint16_t*)g.B, g.ldb, g.N, ke, n0, k0);
    __syncthreads();
    v8bf16 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const uint4 va = *(const uint4*)&As[wm + 16 * i + fr][fk];
      const uint4 vb = *(const uint4*)&Bs[wn + 16 * i + fr][fk];
      a[i] = __builtin_bit_cast(v8bf16, va);
      b[i] = __builtin_bit_cast(v8bf16, vb);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn + 16 * j + fr;
      if (n >= g.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + (lane >> 4) * 4 + r;
        if (m >= g.M) continue;
        const float v = acc[i][j][r];
        if (kEpi == P3NN_EPI_STORE) {
          ((float*)g.C)[(size_t)m * g.ldc + n] = v;
        } else if (kEpi == P3NN_EPI_ATOMIC) {
          atomicAdd((float*)g.C + (size_t)m * g.ldc + n, v);
        } else if (kEpi == P3NN_EPI_BIAS_RELU_BF16) {
          ((uint16_t*)g.C)[(size_t)m * g.ldc + n] = to_bf16(fmaxf(v + g.bias[n], 0.0f));
        } else {  // P3NN_EPI_BIAS_F32
          ((float*)g.C)[(size_t)m * g.ldc + n] = v + g.bias[n];
        }
      }
    }
}

template <bool kA, bool kB>
hipError_t gemm_epi(const P3nnGemm& g, int epi, int split, hipStream_t s) {
  const dim3 grid((g.N + kBN - 1) / kBN, (g.M + kBM - 1) / kBM, split);
  switch (epi) {
    case P3NN_EPI_STORE: gemm_bf16<kA, kB, P3NN_EPI_STORE><<<grid, 256, 0, s>>>(g); break;
    case P3NN_EPI_ATOMIC: gemm_bf16<kA, kB, P3NN_EPI_ATOMIC><<<grid, 256, 0, s>>>(g); break;
    case P3NN_EPI_BIAS_RELU_BF16:
      gemm_bf16<kA, kB, P3NN_EPI_BIAS_RELU_BF16><<<grid, 256, 0, s>>>(g);
      break;
    default: gemm_bf16<kA, kB, P3NN_EPI_BIAS_F32><<<grid, 256, 0, s>>>(g); break;
  }
  return hipGetLastError();
}

// ------------------------------------------------------------- im2col --

// Source pixel of output pixel (oy, ox), tap (dy, dx) in the layer's input
// (low-resolution pixel for an upsampling layer), or false for padding.
__device__ inline bool tap_source(const P3nnConv& c, int oy, int ox, int dy, int dx, int& iy,
                                  int& ix) {
  if (c.up) {
    const int uy = oy + dy, ux = ox + dx;   // upsampled plane = output size
    if (uy < 0 || uy >= c.out_h || ux < 0 || ux >= c.out_w) return false;
    iy = uy >> 1;
    ix = ux >> 1;
    return true;
  }
  iy = oy * c.stride + dy;
  ix = ox * c.stride + dx;
  return iy >= 0 && iy < c.in_h && ix >= 0 && ix < c.in_w;
}

// One thread per (output pixel, tap, 8-channel quad); quads past 9 Cin up to
// the padded row length are zero.
__global__ void __launch_bounds__(256) im2col(P3nnConv c, const uint16_t* a, const uint16_t* b,
                                              uint16_t* col, int kp) {
  const int cin = c.ca + c.cb;
  const int qpr = kp / 8;   // quads per col row
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t P = (size_t)c.n * c.out_h * c.out_w;
  if (i >= P * qpr) return;
  const int q = (int)(i % qpr);
  const size_t p = i / qpr;
  const int ox = (int)(p % c.out_w), oy = (int)((p / c.out_w) % c.out_h);
  const int img = (int)(p / ((size_t)c.out_w * c.out_h));
  const int k = q * 8, t = k / cin, ch = k % cin;
  uint4 v = make_uint4(0u, 0u, 0u, 0u);
  int iy, ix;
  if (t < 9 && tap_source(c, oy, ox, t / 3 - 1, t % 3 - 1, iy, ix)) {
    const size_t pix = ((size_t)img * c.in_h + iy) * c.in_w + ix;
    v = ch < c.ca ? *(const uint4*)(a + pix * c.ca + ch)
                  : *(const uint4*)(b + pix * c.cb + (ch - c.ca));
  }
  *(uint4*)(col + p * kp + k) = v;
}

// Gather backward of im2col: one thread per (input pixel, channel) of the
// input it belongs to (a: ca channels at input resolution; b: cb channels
// at output resolution, the concat's second input); adds into da / db.
__global__ void __launch_bounds__(256) col2im(P3nnConv c, const float* dcol, int kp, float* da,
                                              float* db) {
  const int cin = c.ca + c.cb;
  const size_t na = (size_t)c.n * c.in_h * c.in_w * c.ca;
  const size_t nb = (size_t)c.n * c.out_h * c.out_w * c.cb;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= na + nb) return;
  const bool in_b = i >= na;
  const size_t j = in_b ? i - na : i;
  const int cc = in_b ? c.cb : c.ca;
  const int ch = (int)(j % cc);
  const size_t pix = j / cc;
  const int W = in_b ? c.out_w : c.in_w, H = in_b ? c.out_h : c.in_h;
  const int x = (int)(pix % W), y = (int)((pix / W) % H);
  const int img = (int)(pix / ((size_t)W * H));
  const int k_ch = in_b ? c.ca + ch : ch;
  float s = 0.0f;
  for (int t = 0; t < 9; ++t) {
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    // output pixels (oy, ox) whose tap t read this input pixel
    int ys[2], xs[2], ny = 0, nx = 0;
    if (c.up && !in_b) {
      for (int u = 0; u < 2; ++u) {
        const int uy = 2 * y + u, oy = uy - dy;
        if (uy < c.out_h && oy >= 0 && oy < c.out_h) ys[ny++] = oy;
        const int ux = 2 * x + u, ox = ux - dx;
        if (ux < c.out_w && ox >= 0 && ox < c.out_w) xs[nx++] = ox;
      }
    } else {
      const int st = in_b ? 1 : c.stride;
      const int ry = y - dy, rx = x - dx;
      if (ry >= 0 && ry % st == 0 && ry / st < c.out_h) ys[ny++] = ry / st;
      if (rx >= 0 && rx % st == 0 && rx / st < c.out_w) xs[nx++] = rx / st;
    }
    for (int a = 0; a < ny; ++a)
      for (int bq = 0; bq < nx; ++bq) {
        const size_t p = ((size_t)img * c.out_h + ys[a]) * c.out_w + xs[bq];
        s += dcol[p * kp + t * cin + k_ch];
      }
  }
  float* d = in_b ? db : da;
  d[j] += s;
}

// ------------------------------------------------------ elementwise ---

// dy (f32, [P][C]) times the ReLU mask (y > 0, the stored bf16 activation)
// to bf16.
__global__ void __launch_bounds__(256) relu_bwd(const float* dy, const uint16_t* y, uint16_t* out,
                                                size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  out[i] = from_bf16(y[i]) > 0.0f ? to_bf16(dy[i]) : (uint16_t)0;
}

// db[o] += sum_p dy[p][o] (dy bf16, as it enters the weight gradient).
__global__ void __launch_bounds__(256) bias_grad(const uint16_t* dy, size_t P, int C, float* db) {
  __shared__ float part[256];
  const int o = blockIdx.x;
  float s = 0.0f;
  for (size_t p = (size_t)blockIdx.y * 256 + threadIdx.x; p < P; p += (size_t)gridDim.y * 256)
    s += from_bf16(dy[p * C + o]);
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(db + o, part[0]);
}

// Network input (8 bf16 channels per pixel) of a frame: p3nn::pixel_features.
__global__ void __launch_bounds__(256) features(const uint32_t* hist, const float* img, float k2,
                                                size_t pixels, uint16_t* out) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= pixels) return;
  ((uint4*)out)[i] = p3nn::pixel_features(((const uint4*)hist)[i], ((const float4*)img)[i], k2);
}

// 1x1 head: out[p][c] = base[p][c] + b[c] + sum_i w[c][i] x[p][i] (x bf16,
// 16 channels; base: the DE-off tonemapped f32 RGBA image); with a
// target, also the loss term and dout = 2 (out - target) / count, and the
// head's input gradient dx and weight / bias gradients.
__global__ void __launch_bounds__(256) head_fwd_bwd(const uint16_t* x, const float* base,
                                                    const float* w, const float* target,
                                                    size_t P, float* out, float* dx, float* dw,
                                                    float* loss) {
  __shared__ float sdw[51];
  __shared__ float sl;
  if (threadIdx.x < 51) sdw[threadIdx.x] = 0.0f;
  if (threadIdx.x == 0) sl = 0.0f;
  __syncthreads();
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (p < P) {
    float xi[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) xi[i] = from_bf16(x[p * 16 + i]);
    float d[3];
    float l = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float s = base[p * 4 + c] + w[48 + c];
#pragma unroll
      for (int i = 0; i < 16; ++i) s += w[c * 16 + i] * xi[i];
      out[p * 3 + c] = s;
      d[c] = 0.0f;
      if (target) {
        const float e = s - target[p * 3 + c];
        l += e * e;
        d[c] = 2.0f * e / (float)(P * 3);
      }
    }
    if (target) {
      atomicAdd(&sl, l);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        dx[p * 16 + i] = d[0] * w[i] + d[1] * w[16 + i] + d[2] * w[32 + i];
#pragma unroll
        for (int c = 0; c < 3; ++c) atomicAdd(&sdw[c * 16 + i], d[c] * xi[i]);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) atomicAdd(&sdw[48 + c], d[c]);
    }
  }
  __syncthreads();
  if (target) {
    if (threadIdx.x < 51) atomicAdd(dw + threadIdx.x, sdw[threadIdx.x]);
    if (threadIdx.x == 0) atomicAdd(loss, sl / (float)(P * 3));
  }
}

// Adam (bias-corrected) on f32 parameters, then the bf16 copy the GEMMs
// read (null: none).
__global__ void __launch_bounds__(256) adam(float* w, const float* g, float* m, float* v, size_t n,
                                            float lr, float b1, float b2, float eps, float c1,
                                            float c2, uint16_t* wbf) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wbf) wbf[i] = to_bf16(wi);
}

__global__ void __launch_bounds__(256) cast_bf16(const float* x, uint16_t* y, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = to_bf16(x[i]);
}

unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" {

int p3nn_gemm(const P3nnGemm* g, int a_kmajor, int b_kmajor, int epilogue, int split,
              void* stream) {
  if (!g || g->M <= 0 || g->N <= 0 || g->K <= 0 || split < 1 || split > 65535) return -1;
  // a K-minor operand is read 8 k at a time; a K-major one 8 rows at a time
  if (((!a_kmajor || !b_kmajor) && g->K % 8) || g->lda % 8 || g->ldb % 8) return -1;
  if (epilogue < 0 || epilogue > 3) return -1;
  if ((epilogue == P3NN_EPI_BIAS_RELU_BF16 || epilogue == P3NN_EPI_BIAS_F32) && !g->bias) return -1;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (!a_kmajor && !b_kmajor) e = gemm_epi<false, false>(*g, epilogue, split, s);
  else if (!a_kmajor && b_kmajor) e = gemm_epi<false, true>(*g, epilogue, split, s);
  else if (a_kmajor && !b_kmajor) e = gemm_epi<true, false>(*g, epilogue, split, s);
  else e = gemm_epi<true, true>(*g, epilogue, split, s);
  return e == hipSuccess ? 0 : -2;
}

static int conv_ok(const P3nnConv* c) {
  if (!c || c->n < 1 || c->in_w < 1 || c->in_h < 1 || c->out_w < 1 || c->out_h < 1) return 0;
  if (c->ca < 8 || c->ca % 8 || c->cb < 0 || c->cb % 8) return 0;
  if (c->stride != 1 && c->stride != 2) return 0;
  if (c->up && (c->stride != 1 || c->cb)) return 0;
  if (c->cb && (c->in_w != c->out_w || c->in_h != c->out_h || c->stride != 1)) return 0;
  return 1;
}

int p3nn_im2col(const P3nnConv* c, const uint16_t* a, const uint16_t* b, uint16_t* col, int kp,
                void* stream) {
  if (!conv_ok(c) || !a || (c->cb && !b) || !col || kp % 8 || kp < 9 * (c->ca + c->cb)) return -1;
  const size_t n = (size_t)c->n * c->out_h * c->out_w * (kp / 8);
  im2col<<<blocks(n), 256, 0, (hipStream_t)stream>>>(*c, a, b, col, kp);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int p3nn_col2im(const P3nnConv* c, const float* dcol, int kp, float* da, float* db, void* stream) {
  if (!conv_ok(c) || !dcol || !da || (c->cb && !db) || kp < 9 * (c->ca + c->cb)) return -1;
  const size_t n = (size_t)c->n * (c->in_h * c->in_w * c->ca + c->out_h * c->out_w * c->cb);
  col2im<<<blocks(n), 256, 0, (hipStream_t)stream>>>(*c, dcol, kp, da, db);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int p3nn_relu_bwd(const float* dy, const uint16_t* y, uint16_t* out, size_t n, void* stream) {
  if (!dy || !y || !out) return -1;
  relu_bwd<<<blocks(n), 256, 0, (hipStream_t)stream>>>(dy, y, out, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int p3nn_bias_grad(const uint16_t* dy, size_t P, int C, float* db, void* stream) {
  if (!dy || !db || C < 1) return -1;
  const unsigned ys = (unsigned)((P + 65535) / 65536 < 64 ? (P + 65535) / 65536 + 1 : 64);
  bias_grad<<<dim3(C, ys), 256, 0, (hipStream_t)stream>>>(dy, P, C, db);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int p3nn_features(const uint32_t* hist, const float* img, float k2, size_t pixels, uint16_t* out,
                  void* stream) {
  if (!hist || !img || !out) return -1;
  features<<<blocks(pixels), 256, 0, (hipStream_t)stream>>>(hist, img, k2, pixels, out);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int p3nn_head(const uint16_t* x, const float* base, const float* w, const float* target,
              size_t P, float* out, float* dx, float* dw, float* loss, void* stream) {
  if (!x || !base || !w || !out || (target && (!dx || !dw || !loss))) return -1;
  head_fwd_bwd<<<blocks(P), 256, 0, (hipStream_t)stream>>>(x, base, w, target, P, out, dx, dw,
                                                           loss);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int p3nn_adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
              float eps, int step, uint16_t* wbf, void* stream) {
  if (!w || !g || !m || !v || step < 1) return -1;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  adam<<<blocks(n), 256, 0, (hipStream_t)stream>>>(w, g, m, v, n, lr, b1, b2, eps, c1, c2, wbf);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int p3nn_cast_bf16(const float* x, uint16_t* y, size_t n, void* stream) {
  if (!x || !y) return -1;
  cast_bf16<<<blocks(n), 256, 0, (hipStream_t)stream>>>(x, y, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // extern "C"
