/*
 * pyr3nn — C ABI of the learned denoiser's training kernels (gfx950, BF16
 * MFMA with FP32 accumulation). No counterpart in the reference: the
 * denoiser is an addition of this build (its inference runs through
 * p3_denoise in pyr3hip.h). Every pointer is device memory; every call is
 * queued on `stream` (a hipStream_t; null = the null stream) and returns 0,
 * -1 for an invalid argument or -2 for a launch failure.
 *
 * Tensors: activations NHWC bf16 (channel-minor); a 3x3 convolution's
 * weights W[o][k] with k = tap * Cin + c (tap = 3 (dy + 1) + (dx + 1)),
 * rows padded to kp (a multiple of 8) columns.
 */
#ifndef PYR3NN_H
#define PYR3NN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define P3NN_EPI_STORE 0           /* C f32 = A B^T */
#define P3NN_EPI_ATOMIC 1          /* C f32 += A B^T (atomic: split-K) */
#define P3NN_EPI_BIAS_RELU_BF16 2  /* C bf16 = relu(A B^T + bias) */
#define P3NN_EPI_BIAS_F32 3        /* C f32 = A B^T + bias */

/* C[m][n] (epilogue) sum_k A(m, k) B(n, k), bf16 operands, f32 accumulate.
 * A is stored [M][K] (a_kmajor = 0, leading dimension lda) or [K][M]
 * (a_kmajor = 1); B [N][K] (b_kmajor = 0) or [K][N] (b_kmajor = 1). lda and
 * ldb are multiples of 8, and so is K when an operand is stored K-minor.
 * split > 1 divides K over blocks (use the atomic epilogue). */
typedef struct {
  const void* A;
  const void* B;
  void* C;
  const float* bias;
  int M, N, K, lda, ldb, ldc;
} P3nnGemm;
int p3nn_gemm(const P3nnGemm* g, int a_kmajor, int b_kmajor, int epilogue, int split,
              void* stream);

/* A 3x3 convolution layer (padding 1): n images; input a (ca channels,
 * in_w x in_h) and optionally b (cb channels, concatenated after a, at the
 * output size); stride 1 or 2; up = 1: the input is nearest-upsampled 2x
 * and cropped to the output size first. ca, cb multiples of 8. */
typedef struct {
  int n, in_w, in_h, out_w, out_h, ca, cb, stride, up;
} P3nnConv;
/* col [n out_h out_w][kp] bf16, kp >= 9 (ca + cb); columns past 9 Cin are 0. */
int p3nn_im2col(const P3nnConv* c, const uint16_t* a, const uint16_t* b, uint16_t* col, int kp,
                void* stream);
/* Adds the gradient of im2col's inputs: da (f32, like a) and db (like b)
 * from dcol (f32 [P][kp]). */
int p3nn_col2im(const P3nnConv* c, const float* dcol, int kp, float* da, float* db, void* stream);
/* out (bf16) = dy (f32) where the bf16 activation y > 0, else 0. */
int p3nn_relu_bwd(const float* dy, const uint16_t* y, uint16_t* out, size_t n, void* stream);
/* db[o] += sum_p dy[p][o], dy bf16 [P][C]. */
int p3nn_bias_grad(const uint16_t* dy, size_t P, int C, float* db, void* stream);
/* Network input (8 bf16 channels per pixel) from the u32 RGBA histogram
 * (oversample 1) and the DE-off tonemapped f32 RGBA image of the same frame:
 * base RGB, log1p(16 count k2) / 4, mean splat colour (count units),
 * occupancy (k2: the frame's calibration constant). */
int p3nn_features(const uint32_t* hist, const float* img, float k2, size_t pixels, uint16_t* out,
                  void* stream);
/* 1x1 head on x (bf16 [P][16]): out[p][c] = base[p].rgb + w[48 + c] +
 * sum_i w[c * 16 + i] x[p][i] (base: f32 RGBA). With target (f32 [P][3]):
 * loss += mean squared error, dx (f32 [P][16]) = its gradient at x, dw += its
 * gradient at w (51 floats). */
int p3nn_head(const uint16_t* x, const float* base, const float* w, const float* target,
              size_t P, float* out, float* dx, float* dw, float* loss, void* stream);
/* Adam step `step` (1-based) on n f32 parameters w with gradient g and
 * moments m, v; also writes the bf16 copy wbf when non-null. */
int p3nn_adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
              float eps, int step, uint16_t* wbf, void* stream);
int p3nn_cast_bf16(const float* x, uint16_t* y, size_t n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PYR3NN_H */
