// Chaos kernels for any genome shape, built with the library.
//
// The per-shape kernels of p3_rtc.cpp need a hipRTC compile (1-35 s) the
// first time a shape is seen. These take the shape from the xform records at
// run time instead: the chain loops over each record's variation slots and
// enters the catalog switch per slot (src/shaders/chaos.wgsl:8226-8267,
// 7731-8107), in the same order and with the same arithmetic as the
// generated straight-line cases — pre_blur slots first, then the weighted
// sum in slot order with the direct-colour check after each slot, then the
// post-affine — so a dispatch gives the same histogram through either
// kernel. The catalog sits behind non-inlined functions, so the kernel is
// slower per turn than a specialised one; it serves shapes whose specialised
// kernel is not compiled yet and batches of one-off shapes.
#include "p3_chaos_kernel.h"
#include "p3_launch.h"

namespace p3 {
namespace {

__device__ __noinline__ float2 generic_variation(int idx, float2 p, xrec_t xf, int k, float w,
                                                 Rng& rng) {
  return apply_variation(idx, p, xf, k, w, rng);
}

__device__ __noinline__ bool generic_dc(int idx, float2 p, xrec_t xf, int k, float3& rgb) {
  return dc_color_for(idx, p, xf, k, rgb);
}

P3_INLINE int slot_count(xrec_t xf) {
  const float n = xf[7];
  return n < 8.0f ? (int)n : 8;
}

// pre_blur slots, then the weighted variation sum (direct colour after each
// slot); the post-affine is left to the caller.
P3_INLINE float2 generic_sum(xrec_t xf, float2 pam, Rng& rng, bool& dc_on, float3& dc_rgb) {
  const int nv = slot_count(xf);
  for (int k = 0; k < nv; ++k)
    if ((uint32_t)xf[20 + 4 * k] == 97u) pam = pre_blur_slot(xf, k, pam, rng);
  float2 pv = make_float2(0.0f, 0.0f);
  for (int k = 0; k < nv; ++k) {
    const uint32_t idx = (uint32_t)xf[20 + 4 * k];
    const float w = xf[20 + 4 * k + 1];
    if (idx != 97u) {
      const float2 o = generic_variation((int)idx, pam, xf, k, w, rng);
      pv.x = pv.x + o.x;
      pv.y = pv.y + o.y;
    }
    if (is_dc_index(idx) && xf[11] > 0.5f && w > 0.0f) {
      float3 dcc;
      if (generic_dc((int)idx, pam, xf, k, dcc)) {
        dc_on = true;
        dc_rgb = dcc;
      }
    }
  }
  return pv;
}

struct GenericSpec {
  static constexpr bool kPairAdds = false;
  static constexpr int kRegular = -1;  // xform count known at run time only
  static P3_INLINE bool has_final(const ChaosArgs& A) { return A.final_idx >= 0; }
  static P3_INLINE int final_slot(const ChaosArgs& A) { return A.final_idx; }
  static P3_INLINE bool final_draw_free(const ChaosArgs&) { return false; }
  static P3_INLINE Chain chain(uint32_t fn, xrec_t X, float x, float y, float z, Rng& rng,
                               bool& dc_on, float3& dc_rgb) {
    const xrec_t xf = X + (size_t)fn * kXformFloats;
    Chain c;
    c.pa = make_float2(xf[0] * x + xf[1] * y + xf[2], xf[4] * x + xf[5] * y + xf[6]);
    const float2 pv = generic_sum(xf, c.pa, rng, dc_on, dc_rgb);
    c.pv_pre = pv;
    c.pv = xf[15] != 0.0f ? post_affine(xf, pv) : pv;
    c.z = mixf(z, xf[8], xf[9]);
    return c;
  }
  static P3_INLINE float2 final_sum(xrec_t X, int slot, float2 pa, Rng& rng, bool& dc_on,
                                    float3& dc_rgb) {
    const xrec_t xf = X + (size_t)slot * kXformFloats;
    const float2 pv = generic_sum(xf, pa, rng, dc_on, dc_rgb);
    return xf[15] != 0.0f ? post_affine(xf, pv) : pv;
  }
};

template <bool kSat>
__global__ void __launch_bounds__(kChaosBlock) chaos_generic(ChaosArgs A) {
  __shared__ BlockLds<kChaosBlock> L;
  chaos_binned<kSat, kChaosBlock, GenericSpec>(A, L, blockIdx.x);
}

template <bool kSat>
__global__ void __launch_bounds__(kChaosBlock) chaos_generic_multi(MultiArgs M) {
  __shared__ BlockLds<kChaosBlock> L;
  const uint32_t k = multi_find(M, blockIdx.x);
  const ChaosArgs A = M.args[k];
  chaos_binned<kSat, kChaosBlock, GenericSpec>(A, L, blockIdx.x - M.start[k]);
}

}  // namespace

hipError_t launch_chaos_generic(const ChaosArgs& a, bool saturating, hipStream_t stream) {
  const unsigned grid = (a.walkers + kChaosBlock - 1) / kChaosBlock;
  if (grid == 0) return hipSuccess;
  if (saturating)
    hipLaunchKernelGGL(chaos_generic<true>, dim3(grid), dim3(kChaosBlock), 0, stream, a);
  else
    hipLaunchKernelGGL(chaos_generic<false>, dim3(grid), dim3(kChaosBlock), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_chaos_generic_multi(const MultiArgs& m, bool saturating, hipStream_t stream) {
  const unsigned grid = m.start[m.n];
  if (grid == 0) return hipSuccess;
  if (saturating)
    hipLaunchKernelGGL(chaos_generic_multi<true>, dim3(grid), dim3(kChaosBlock), 0, stream, m);
  else
    hipLaunchKernelGGL(chaos_generic_multi<false>, dim3(grid), dim3(kChaosBlock), 0, stream, m);
  return hipGetLastError();
}

}  // namespace p3
