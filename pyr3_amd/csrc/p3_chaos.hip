// Chaos-game kernel for gfx950: one walker per lane, ISAAC stream in LDS,
// u32 RGBA histogram scatter with native global atomics.
//
// Behaviour follows chaos_main (src/shaders/chaos.wgsl:8269-8732):
//   init draws rand_11, rand_11, rand01 -> (x, y, z)
//   per turn: pick = table[row * 16384 + (irand & 16383)], row = previous
//   accepted xform or the fallback row; chain A (affine, pre_blur, variation
//   sum, direct colour, post-affine, colour mix); optional xform blend; bad
//   check (NaN or |v| > 1e10) with rand_11 reseed and up to four retried
//   turns that do not advance i (except at i = 0); final-xform lens with an
//   opacity gate; scale-relative jitter (two rand01); splat after the fuse.
// The splat adds u32(pal * opacity * 255) and u32(opacity * 255) to the four
// channels of bin (y * W + x). The reference's adds saturate at 2^32 - 1; a
// saturating add is order-independent (the final value is min(sum, 2^32-1)),
// so it is reproduced with a native atomic add whose returned old value
// detects a wrap, followed by an exchange to 2^32 - 1: after the last wrap
// of a bin the last operation on it is always that exchange. When the host
// has proven that no bin can reach 2^32 (bound on 255 x samples since the
// last reset) the adds are issued without a return value.
#include "p3_device.h"
#include "p3_launch.h"
#include "p3_variations.h"

namespace p3 {

struct Chain {
  float2 pv, pv_pre, pa;
  float z;
};

__device__ inline int nvars_of(const float* xf) {
  const uint32_t n = f2u_sat(xf[7]);
  return (int)(n < (uint32_t)kMaxVars ? n : (uint32_t)kMaxVars);
}

// chaos.wgsl:8164-8181
__device__ inline float2 pre_blur(const float* xf, int nv, float2 c, Rng& rng) {
  for (int k = 0; k < nv; ++k) {
    const float* v = xf + 20 + k * 4;
    if (f2u_sat(v[0]) == 97u) {
      const float r0 = rng.rand01(), r1 = rng.rand01(), r2 = rng.rand01(), r3 = rng.rand01();
      const float g = v[1] * (r0 + r1 + r2 + r3 - 2.0f);
      const float r4 = rng.rand01();
      const float ang = r4 * kTau;
      c.x = c.x + cosf(ang) * g;
      c.y = c.y + sinf(ang) * g;
    }
  }
  return c;
}

// Variation sum with the "last active direct colour wins" override.
__device__ inline float2 var_sum(const float* xf, int nv, float2 pa, Rng& rng, bool& dc_on,
                                 float3& dc_rgb) {
  float2 pv = make_float2(0.0f, 0.0f);
  const bool dc_flag = xf[11] > 0.5f;
  for (int k = 0; k < nv; ++k) {
    const float* v = xf + 20 + k * 4;
    const int idx = (int)f2u_sat(v[0]);
    const float w = v[1];
    if (idx != 97) {
      const float2 o = apply_variation(idx, pa, xf, k, w, rng);
      pv.x = pv.x + o.x;
      pv.y = pv.y + o.y;
    }
    if (dc_flag && w > 0.0f) {
      float3 rgb;
      if (dc_color_for(idx, pa, xf, k, rgb)) {
        dc_on = true;
        dc_rgb = rgb;
      }
    }
  }
  return pv;
}

__device__ inline float2 post_affine(const float* xf, float2 pv) {
  if (xf[15] != 0.0f) {
    return make_float2(xf[12] * pv.x + xf[13] * pv.y + xf[14],
                       xf[16] * pv.x + xf[17] * pv.y + xf[18]);
  }
  return pv;
}

// chaos.wgsl:8226-8267
__device__ inline Chain xform_chain(const float* xf, float x, float y, float z, Rng& rng,
                                    bool& dc_on, float3& dc_rgb) {
  Chain c;
  c.pa = make_float2(xf[0] * x + xf[1] * y + xf[2], xf[4] * x + xf[5] * y + xf[6]);
  const int nv = nvars_of(xf);
  const float2 pam = pre_blur(xf, nv, c.pa, rng);
  const float2 pv = var_sum(xf, nv, pam, rng, dc_on, dc_rgb);
  c.pv_pre = pv;
  c.pv = post_affine(xf, pv);
  c.z = mixf(z, xf[8], xf[9]);
  return c;
}

__device__ inline bool is_bad(float2 v) {
  return isnan(v.x) || isnan(v.y) || fabsf(v.x) > 1e10f || fabsf(v.y) > 1e10f;
}

template <bool kSat>
__device__ inline void hist_add(uint32_t* slot, uint32_t d) {
  if (d == 0u) return;
  if (kSat) {
    const uint32_t old = atomicAdd(slot, d);
    if (old > 0xffffffffu - d) atomicExch(slot, 0xffffffffu);
  } else {
    atomicAdd(slot, d);
  }
}

// Per-walker loop state that stays with the walker between turns.
struct Walker {
  float x, y, z;
  int prev;
  uint32_t consec, iter;
};

// One turn of walker `wid` (global id) after its pick `fn` (from draw `raw`)
// is known: chain, blend, bad check and retry window, final lens, jitter,
// splat. Returns nothing; advances `wk.iter` unless the turn is retried.
template <bool kSat>
__device__ inline void run_turn(const ChaosArgs& A, Walker& wk, Rng& rng, uint32_t fn,
                                uint32_t raw, uint32_t wid) {
  const uint32_t i = wk.iter;
  const float* xf = A.xforms + (size_t)fn * kXformFloats;
  bool dc_on = false;
  float3 dc_rgb = make_float3(0.0f, 0.0f, 0.0f);
  const Chain ca = xform_chain(xf, wk.x, wk.y, wk.z, rng, dc_on, dc_rgb);
  float2 pv = ca.pv;
  float nz = ca.z;
  if (A.xform_blend > 0.0f) {
    if (rng.rand01() < A.xform_blend) {
      const uint32_t rb = rng.irand();
      const uint32_t bi = A.distrib[(size_t)A.fallback_row * kGrain + (rb & (kGrain - 1))];
      bool dc_b = false;
      float3 rgb_b;
      const Chain cb = xform_chain(A.xforms + (size_t)bi * kXformFloats, wk.x, wk.y, wk.z, rng,
                                   dc_b, rgb_b);
      const float t = rng.rand01();
      pv.x = mixf(pv.x, cb.pv.x, t);
      pv.y = mixf(pv.y, cb.pv.y, t);
      nz = mixf(nz, cb.z, t);
    }
  }
  const bool bad = is_bad(pv);
  if (bad) {
    const float rx = rng.rand11();
    const float ry = rng.rand11();
    pv = make_float2(rx, ry);
    wk.consec += 1u;
  } else {
    wk.consec = 0u;
  }
  float* tr = nullptr;
  if (A.trace != nullptr && wid < A.trace_walkers && i < A.trace_iters) {
    tr = A.trace + ((size_t)wid * A.trace_iters + i) * kTraceFloats;
    tr[0] = (float)i; tr[1] = (float)fn;
    tr[2] = ca.pa.x; tr[3] = ca.pa.y; tr[4] = ca.pv_pre.x; tr[5] = ca.pv_pre.y;
    tr[6] = pv.x; tr[7] = pv.y; tr[8] = bad ? 1.0f : 0.0f; tr[9] = nz;
    tr[11] = __uint_as_float(raw);
    tr[12] = pv.x; tr[13] = pv.y; tr[14] = 0.0f; tr[15] = 0.0f;
  }
  const float ppx = pv.x, ppy = pv.y, ppz = nz;
  if (wk.consec > 0u && wk.consec < 5u) {
    // Retried turn: the loop index only advances when it is 0.
    wk.x = ppx; wk.y = ppy; wk.z = ppz;
    if (tr) tr[10] = __uint_as_float(rng.draws);
    if (i == 0u) wk.iter = 1u;
    return;
  }
  if (wk.consec >= 5u) wk.consec = 0u;
  wk.prev = (int)fn;
  float sx = ppx, sy = ppy, sz = ppz;
  if (A.final_idx >= 0) {
    const float* fx = A.xforms + (size_t)A.final_idx * kXformFloats;
    const bool apply = fx[10] == 1.0f ? true : (rng.rand01() < fx[10]);
    if (apply) {
      const float2 fpa = make_float2(fx[0] * pv.x + fx[1] * pv.y + fx[2],
                                     fx[4] * pv.x + fx[5] * pv.y + fx[6]);
      const int fnv = nvars_of(fx);
      const float2 fpam = pre_blur(fx, fnv, fpa, rng);
      float2 fpv = var_sum(fx, fnv, fpam, rng, dc_on, dc_rgb);
      fpv = post_affine(fx, fpv);
      const float fz = mixf(nz, fx[8], fx[9]);
      if (!is_bad(fpv)) { sx = fpv.x; sy = fpv.y; sz = fz; }
    }
  }
  const float mag = fmaxf(fmaxf(fabsf(ppx), fabsf(ppy)), 1e-30f);
  const float amp = mag * A.jitter;
  const float jx = (rng.rand01() - 0.5f) * amp;
  const float jy = (rng.rand01() - 0.5f) * amp;
  wk.x = ppx + jx; wk.y = ppy + jy; wk.z = ppz;
  if (tr) {
    tr[12] = wk.x; tr[13] = wk.y; tr[14] = sx; tr[15] = sy;
    tr[10] = __uint_as_float(rng.draws);
  }
  wk.iter = i + 1u;
  if (i >= A.fuse) {
    const float opacity = xf[10];
    const float dx = sx - A.cx, dy = sy - A.cy;
    const float rx = dx * A.cos_r - dy * A.sin_r;
    const float ry = dx * A.sin_r + dy * A.cos_r;
    const float px = rx * A.scale + (float)A.width * 0.5f;
    const float py = ry * A.scale + (float)A.height * 0.5f;
    // floor(p) in [0, W) <=> p in [0, W) for finite p; NaN fails both tests.
    if (px >= 0.0f && px < (float)A.width && py >= 0.0f && py < (float)A.height) {
      const uint32_t xi = (uint32_t)floorf(px), yi = (uint32_t)floorf(py);
      const float cxf = fminf(fmaxf(sz, 0.0f), 1.0f) * 255.0f;
      float4 pal;
      if (A.palette_mode == 1u) {
        uint32_t i0 = f2u_sat(cxf);
        float frac = cxf - (float)i0;
        if (i0 >= 255u) { i0 = 254u; frac = 1.0f; }
        const float4 p0 = A.palette[i0], p1 = A.palette[i0 + 1u];
        pal = make_float4(mixf(p0.x, p1.x, frac), mixf(p0.y, p1.y, frac),
                          mixf(p0.z, p1.z, frac), mixf(p0.w, p1.w, frac));
      } else {
        const uint32_t pi = min(f2u_sat(cxf), 255u);
        pal = A.palette[pi];
      }
      if (dc_on) { pal.x = dc_rgb.x; pal.y = dc_rgb.y; pal.z = dc_rgb.z; }
      const float wgt = opacity * 255.0f;
      uint32_t* b = A.hist + ((size_t)yi * A.width + xi) * 4u;
      hist_add<kSat>(b + 0, f2u_sat(pal.x * wgt));
      hist_add<kSat>(b + 1, f2u_sat(pal.y * wgt));
      hist_add<kSat>(b + 2, f2u_sat(pal.z * wgt));
      hist_add<kSat>(b + 3, f2u_sat(wgt));
    }
  }
}

// LDS image of a block's walkers (structure of arrays, kB walkers):
// ISAAC words [32][kB] (mem 0..15, rsl 16..31) and the per-walker scalars.
template <int kB>
struct BlockLds {
  uint32_t isaac[32 * kB];
  uint32_t cnt[kB], ra[kB], rb[kB], rc[kB];
  float x[kB], y[kB], z[kB];
  int32_t prev[kB];
  uint32_t consec[kB], iter[kB], draws[kB], raw[kB];
  uint16_t fn[kB], order[kB];
  uint32_t bucket[kMaxXforms + 2], base[kMaxXforms + 2];
};

// Persistent block of kB walkers. Every turn:
//   A  lane t draws walker t's pick and counts it into an LDS bucket per xform
//   B  one wave turns the bucket counts into offsets
//   C  walkers are listed in xform order
//   D  lane t runs the turn of the t-th walker in that order
// so the lanes of a wave mostly share one xform and its variation code.
// Each walker keeps its own ISAAC column and loop state; only which lane
// computes it changes, so the per-walker draw sequence, arithmetic and
// splats are those of the one-lane-per-walker loop.
template <bool kSat, int kB>
__global__ void __launch_bounds__(kB) chaos_binned(ChaosArgs A) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  BlockLds<kB>& L = *reinterpret_cast<BlockLds<kB>*>(lds_raw);
  const int t = threadIdx.x;
  const uint32_t w0 = blockIdx.x * kB;
  const uint32_t total = A.iters + A.fuse;
  const uint32_t idle = A.fallback_row + 1u;   // bucket after the last xform

  {  // seed walker t
    const uint32_t wid = w0 + t;
    Rng rng;
    rng.m.base = L.isaac + t;
    rng.m.stride = kB;
    rng.draws = 0u;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    uint32_t iter = total;
    if (wid < A.walkers) {
      if (A.isaac_in) {
        const uint32_t* st = A.isaac_in + (size_t)wid * 36u;
        rng.s.cnt = st[0]; rng.s.a = st[1]; rng.s.b = st[2]; rng.s.c = st[3];
        for (int k = 0; k < 16; ++k) { rng.m.mem(k) = st[4 + k]; rng.m.rsl(k) = st[20 + k]; }
      } else {
        isaac_seed(rng.s, rng.m, A.seed, wid);
      }
      x = rng.rand11();
      y = rng.rand11();
      z = rng.rand01();
      iter = 0u;
    }
    L.cnt[t] = rng.s.cnt; L.ra[t] = rng.s.a; L.rb[t] = rng.s.b; L.rc[t] = rng.s.c;
    L.x[t] = x; L.y[t] = y; L.z[t] = z;
    L.prev[t] = -1; L.consec[t] = 0u; L.iter[t] = iter; L.draws[t] = rng.draws;
  }
  for (int k = t; k < kMaxXforms + 2; k += kB) L.bucket[k] = 0u;
  __syncthreads();

  for (;;) {
    // A: pick for walker t
    const bool active = L.iter[t] < total;
    uint32_t f = idle;
    if (active) {
      Rng rng;
      rng.m.base = L.isaac + t;
      rng.m.stride = kB;
      rng.s.cnt = L.cnt[t]; rng.s.a = L.ra[t]; rng.s.b = L.rb[t]; rng.s.c = L.rc[t];
      rng.draws = L.draws[t];
      const int prev = L.prev[t];
      const uint32_t row = prev >= 0 ? (uint32_t)prev : A.fallback_row;
      const uint32_t raw = rng.irand();
      f = A.distrib[(size_t)row * kGrain + (raw & (kGrain - 1))];
      L.cnt[t] = rng.s.cnt; L.ra[t] = rng.s.a; L.rb[t] = rng.s.b; L.rc[t] = rng.s.c;
      L.draws[t] = rng.draws;
      L.raw[t] = raw;
    }
    L.fn[t] = (uint16_t)f;
    const uint32_t pos = atomicAdd(&L.bucket[f], 1u);
    if (__syncthreads_count(active) == 0) break;
    // B: exclusive scan of the bucket counts (first wave, 64 at a time)
    if (t < 64) {
      uint32_t carry = 0u;
      for (uint32_t k0 = 0; k0 <= idle; k0 += 64u) {
        const uint32_t k = k0 + (uint32_t)t;
        const uint32_t v = k <= idle ? L.bucket[k] : 0u;
        uint32_t incl = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
          const uint32_t o = __shfl_up(incl, d, 64);
          if (t >= d) incl += o;
        }
        if (k <= idle) L.base[k] = carry + incl - v;
        carry += __shfl(incl, 63, 64);
      }
    }
    __syncthreads();
    // C: list walkers in xform order
    L.order[L.base[f] + pos] = (uint16_t)t;
    __syncthreads();
    // D: run the turn of the t-th walker in that order
    {
      const int wl = L.order[t];
      if (L.iter[wl] < total) {
        Rng rng;
        rng.m.base = L.isaac + wl;
        rng.m.stride = kB;
        rng.s.cnt = L.cnt[wl]; rng.s.a = L.ra[wl]; rng.s.b = L.rb[wl]; rng.s.c = L.rc[wl];
        rng.draws = L.draws[wl];
        Walker wk{L.x[wl], L.y[wl], L.z[wl], L.prev[wl], L.consec[wl], L.iter[wl]};
        run_turn<kSat>(A, wk, rng, L.fn[wl], L.raw[wl], w0 + wl);
        L.cnt[wl] = rng.s.cnt; L.ra[wl] = rng.s.a; L.rb[wl] = rng.s.b; L.rc[wl] = rng.s.c;
        L.draws[wl] = rng.draws;
        L.x[wl] = wk.x; L.y[wl] = wk.y; L.z[wl] = wk.z;
        L.prev[wl] = wk.prev; L.consec[wl] = wk.consec; L.iter[wl] = wk.iter;
      }
      for (int k = t; k < kMaxXforms + 2; k += kB) L.bucket[k] = 0u;
    }
    __syncthreads();
  }
}

constexpr int kChaosBlock = 256;

size_t chaos_lds_bytes() { return sizeof(BlockLds<kChaosBlock>); }

hipError_t launch_chaos(const ChaosArgs& a, bool saturating, hipStream_t stream) {
  const unsigned grid = (a.walkers + kChaosBlock - 1) / kChaosBlock;
  if (grid == 0) return hipSuccess;
  const size_t shmem = sizeof(BlockLds<kChaosBlock>);
  if (saturating)
    chaos_binned<true, kChaosBlock><<<grid, kChaosBlock, shmem, stream>>>(a);
  else
    chaos_binned<false, kChaosBlock><<<grid, kChaosBlock, shmem, stream>>>(a);
  return hipGetLastError();
}

}  // namespace p3
