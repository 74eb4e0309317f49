// C ABI of the gfx950 render path (declared in include/pyr3hip.h).
//
// A p3_renderer owns one HIP stream and every device buffer of the path:
// the super-res u32 RGBA histogram, the DE staging (pre-scaled RGBA, radius
// bytes, per-tile max radius), the filtered grid, the RGBA8 and RGBA f32
// images, and small uploads (xform records, pick table, palette, taps,
// kernel-norm LUT, optional ISAAC states and trace). Host blobs are
// validated before any launch so a malformed pick table or xform index is an
// error return, never an out-of-bounds device read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pyr3hip.h"
#include "p3_denoise.h"
#include "p3_launch.h"
#include "p3_host.h"


namespace {

constexpr int kTimers = 6;  // merge, chaos, density, tonemap, reset, denoise
enum { T_MERGE = 0, T_CHAOS = 1, T_DE = 2, T_TM = 3, T_RESET = 4, T_DENOISE = 5 };

std::string g_create_error = "";

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t count) {
    if (count <= n && p) return hipSuccess;
    if (p) { (void)hipFree(p); p = nullptr; n = 0; }
    if (count == 0) return hipSuccess;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e == hipSuccess) n = count;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Catalog entries the chaos kernel implements: all of 0..325 (index 97 is
// pre_blur, applied by the chain itself).
bool supported(uint32_t idx) {
  return idx < (uint32_t)p3::kCatalogSize;
}

// Every active slot of every record must name an implemented entry.
int check_records(p3_renderer* r, const float* xforms, int n_slots);

}  // namespace

namespace p3h {

bool pick_bounds(const uint32_t* table, int n_rows, std::vector<uint16_t>& out) {
  if (n_rows > p3::kPickRows) return false;
  out.assign((size_t)n_rows * p3::kPickCols, (uint16_t)P3_GRAIN);
  for (int r = 0; r < n_rows; ++r) {
    const uint32_t* row = table + (size_t)r * P3_GRAIN;
    uint16_t* b = out.data() + (size_t)r * p3::kPickCols;
    uint32_t prev = 0;
    for (uint32_t c = 0; c < (uint32_t)P3_GRAIN; ++c) {
      const uint32_t v = row[c];
      if (v < prev) return false;
      for (; prev < v; ++prev) b[prev] = (uint16_t)c;
    }
  }
  return true;
}

}  // namespace p3h

using namespace p3h;

struct p3_renderer {
  int width = 0, height = 0, os = 1, sw = 0, sh = 0;
  int n_taps = 0;
  float taps_host[17] = {0};
  bool tables_on_device = false;  // taps and kernel-norm table uploaded since allocation
  hipStream_t stream = nullptr;
  DevBuf<uint32_t> hist;
  DevBuf<float4> filtered, pre, imagef;
  DevBuf<uint8_t> rad;
  DevBuf<uint32_t> tile_rmax, image8;
  DevBuf<float> taps, lut, palette, xforms;
  DevBuf<float4> linear;        // p3_read_linear output (allocated on first use)
  DevBuf<float> curves;         // 5 x 256 channel-curve LUT (p3_set_curves)
  std::vector<float> curves_host;
  bool curves_set = false;
  DevBuf<uint32_t> distrib, isaac;
  DevBuf<uint16_t> bounds;      // step boundaries of the pick table, when it has them
  bool bounds_ok = false;
  std::vector<uint16_t> bounds_host, mbounds_host;  // staging kept alive past async copies
  std::vector<p3::ChaosArgs> margs_host;
  DevBuf<float> trace;
  DevBuf<uint32_t> idx;         // colour-index capture (one u32 per bin), once requested
  bool idx_live = false;        // idx sized to the current grid and in use
  DevBuf<float> mxforms;        // batched dispatch: per-sub-dispatch records
  DevBuf<uint32_t> mdistrib;    //   and pick tables
  DevBuf<uint16_t> mbounds;     //   and their step boundaries
  DevBuf<char> margs;           //   and kernel arguments
  bool palette_set = false;
  bool have_image = false;
  bool have_filtered = false;
  // Set while build() (re)sizes the buffers and left set when it fails: the
  // buffers may then be missing or smaller than the grid, so every call that
  // touches them refuses with P3_E_STATE until a p3_resize succeeds.
  bool broken = false;
  // Pending adds outside the u32 histogram: the packed accumulator (1 u64
  // per bin) and its carry counts (4 u64 per bin). All zero when acc_clean;
  // acc_dirty when a dispatch may have left adds that hist_merge has not yet
  // applied to the histogram.
  DevBuf<uint64_t> acc, carry;
  bool acc_clean = false, acc_dirty = false;
  int trace_walkers = 0, trace_iters = 0;
  bool timing = false;
  bool want_image_f32 = false;
  bool have_image_f32 = false;
  float last_k2 = 0.0f;        // calibration k2 of the last present (denoiser input)
  bool last_de = false;        // the last present ran DE
  std::vector<float> last_xforms;       // host copies of the last uploads
  std::vector<uint32_t> last_distrib;
  hipEvent_t ev[kTimers][2] = {};
  bool ev_used[kTimers] = {};
  std::string err;

  int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }
  int hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return P3_OK;
    return fail(P3_E_HIP, "%s: %s", what, hipGetErrorString(e));
  }
  void tic(int t) {
    if (timing) { (void)hipEventRecord(ev[t][0], stream); }
  }
  void toc(int t) {
    if (timing) { (void)hipEventRecord(ev[t][1], stream); ev_used[t] = true; }
  }
};

struct p3_denoiser {
  DevBuf<uint16_t> w;
  DevBuf<float> b, head;
  size_t w_off[p3::kDenoiseLayers] = {}, b_off[p3::kDenoiseLayers] = {};
  DevBuf<uint16_t> act[p3::kDenoiseLayers - 1];  // activations, grown on demand
};

namespace {

// f32 -> bf16 bits, round to nearest even.
uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}


}  // namespace

namespace p3h {

// Per-radius kernel sums over the integer disc, sigma = r / 3, in f64
// (src/density.ts:78-99), stored as f32: ~22 M exp() for radii up to 256,
// so computed once per process (it does not depend on the frame).
const float* kernel_norm_lut() {
  static const std::vector<float> lut = [] {
    std::vector<float> t(257);
    t[0] = 1.0f;
    for (int rr = 1; rr <= 256; ++rr) {
      const double sigma = rr / 3.0, inv = 1.0 / (2.0 * sigma * sigma), r2 = (double)rr * rr;
      double s = 0.0;
      for (int dy = -rr; dy <= rr; ++dy)
        for (int dx = -rr; dx <= rr; ++dx) {
          const double d2 = (double)dx * dx + (double)dy * dy;
          if (d2 <= r2) s += std::exp(-d2 * inv);
        }
      t[rr] = (float)s;
    }
    return t;
  }();
  return lut.data();
}

}  // namespace p3h

namespace {

int build(p3_renderer* r, int width, int height, int os, const float* taps, int n_taps) {
  if (width <= 0 || height <= 0 || os < 1 || os > 16)
    return r->fail(P3_E_ARG, "invalid size %dx%d oversample %d", width, height, os);
  if (!taps || n_taps < 3 || n_taps > 17)
    return r->fail(P3_E_ARG, "taps must hold 3..17 values (got %d)", n_taps);
  const long long sw = (long long)width * os, sh = (long long)height * os;
  if (sw * sh > (1LL << 31))
    return r->fail(P3_E_ARG, "super-res grid %lldx%lld too large", sw, sh);
  // Until every buffer is sized for the new grid the handle is unusable: a
  // failed allocation below leaves it so (calls return P3_E_STATE).
  r->broken = true;
  r->palette_set = false;
  r->have_image = r->have_filtered = r->have_image_f32 = false;
  r->acc_clean = r->acc_dirty = false;
  r->idx_live = false;
  r->width = width; r->height = height; r->os = os; r->sw = (int)sw; r->sh = (int)sh;
  // The taps upload below reads taps_host asynchronously, so a change of
  // taps waits for the stream before the copy is overwritten; unchanged taps
  // (the usual resize between batch jobs) are already on the device.
  const bool same_taps = r->tables_on_device && n_taps == r->n_taps &&
                         std::memcmp(r->taps_host, taps, sizeof(float) * n_taps) == 0;
  int rc;
  if (r->tables_on_device && !same_taps &&
      (rc = r->hip(hipStreamSynchronize(r->stream), "sync before new taps")))
    return rc;
  r->n_taps = n_taps;
  if (!same_taps) {
    r->tables_on_device = false;  // until the upload below is queued
    std::memcpy(r->taps_host, taps, sizeof(float) * n_taps);
  }
  const size_t bins = (size_t)sw * sh;
  const size_t tiles = (size_t)((sw + 15) / 16) * ((sh + 15) / 16);
  if ((rc = r->hip(r->hist.ensure(bins * 4), "hipMalloc(histogram)"))) return rc;
  if ((rc = r->hip(r->filtered.ensure(bins), "hipMalloc(filtered)"))) return rc;
  if ((rc = r->hip(r->pre.ensure(bins), "hipMalloc(de pre)"))) return rc;
  if ((rc = r->hip(r->rad.ensure(bins), "hipMalloc(de radius)"))) return rc;
  if ((rc = r->hip(r->tile_rmax.ensure(tiles), "hipMalloc(de tiles)"))) return rc;
  if ((rc = r->hip(r->image8.ensure((size_t)width * height), "hipMalloc(image)"))) return rc;
  if ((rc = r->hip(r->imagef.ensure((size_t)width * height), "hipMalloc(image f32)"))) return rc;
  if ((rc = r->hip(r->taps.ensure(17), "hipMalloc(taps)"))) return rc;
  if ((rc = r->hip(r->lut.ensure(257), "hipMalloc(lut)"))) return rc;
  if ((rc = r->hip(r->palette.ensure(1024), "hipMalloc(palette)"))) return rc;
  if ((rc = r->hip(r->xforms.ensure((size_t)(P3_MAX_XFORMS + 1) * P3_XFORM_FLOATS),
                   "hipMalloc(xforms)")))
    return rc;
  if ((rc = r->hip(r->distrib.ensure((size_t)(P3_MAX_XFORMS + 1) * P3_GRAIN),
                   "hipMalloc(distrib)")))
    return rc;
  // From the handle's copy and the process-wide table, so the host continues
  // before the uploads run: a resize between batch jobs does not wait for the
  // previous job.
  if (!r->tables_on_device) {
    if ((rc = r->hip(hipMemcpyAsync(r->taps.p, r->taps_host, sizeof(float) * n_taps,
                                    hipMemcpyHostToDevice, r->stream),
                     "upload taps")))
      return rc;
    if ((rc = r->hip(hipMemcpyAsync(r->lut.p, kernel_norm_lut(), sizeof(float) * 257,
                                    hipMemcpyHostToDevice, r->stream),
                     "upload lut")))
      return rc;
    r->tables_on_device = true;
  }
  r->last_xforms.clear();   // the capture restarts (and is re-sized) at its next use
  r->last_distrib.clear();
  r->broken = false;
  return P3_OK;
}

// Bins of the current super-sampled grid (buffers may be larger: a resize
// keeps the allocations and grows them only when needed).
size_t grid_bins(const p3_renderer* r) { return (size_t)r->sw * r->sh; }

void release(p3_renderer* r) {
  r->hist.release(); r->filtered.release(); r->pre.release(); r->rad.release();
  r->tables_on_device = false;
  r->tile_rmax.release(); r->image8.release(); r->imagef.release(); r->taps.release();
  r->lut.release(); r->palette.release(); r->xforms.release(); r->distrib.release();
  r->bounds.release(); r->mbounds.release(); r->acc.release(); r->carry.release();
  r->isaac.release(); r->trace.release(); r->idx.release();
  r->mxforms.release(); r->mdistrib.release(); r->margs.release();
  r->linear.release(); r->curves.release();
}

}  // namespace

namespace p3h {

std::string records_error(const float* xforms, int n_slots) {
  char msg[128];
  for (int s = 0; s < n_slots; ++s) {
    const float* xf = xforms + (size_t)s * P3_XFORM_FLOATS;
    const float nvf = xf[7];
    if (!(nvf >= 0.0f) || nvf > 1e9f) {
      std::snprintf(msg, sizeof msg, "xform %d: bad variation count", s);
      return msg;
    }
    const int nv = std::min((int)nvf, 8);
    for (int k = 0; k < nv; ++k) {
      const float idxf = xf[20 + 4 * k];
      if (!(idxf >= 0.0f) || idxf > 1e9f || !supported((uint32_t)idxf)) {
        std::snprintf(msg, sizeof msg, "xform %d slot %d: catalog entry %g is not implemented",
                      s, k, (double)idxf);
        return msg;
      }
    }
  }
  return std::string();
}

}  // namespace p3h

namespace {

int check_records(p3_renderer* r, const float* xforms, int n_slots) {
  const std::string e = records_error(xforms, n_slots);
  return e.empty() ? P3_OK : r->fail(P3_E_ARG, "%s", e.c_str());
}

// Checks one dispatch's blobs against the renderer (sizes, final slot, uniform
// grid size, catalog support, pick-table range); `check_tables` = false skips
// the O(table) scans when the blobs are known to equal validated ones.
int validate_dispatch(p3_renderer* r, const float* xforms, int n_slots, const uint32_t* distrib,
                      int n_rows, const uint32_t* u, bool check_records_too, bool check_table,
                      int& n_regular, int32_t& final_idx) {
  if (!xforms || !distrib || !u) return r->fail(P3_E_ARG, "null argument");
  if (n_rows < 2 || n_rows > P3_MAX_XFORMS + 1)
    return r->fail(P3_E_ARG, "n_rows must be 2..%d (got %d)", P3_MAX_XFORMS + 1, n_rows);
  n_regular = n_rows - 1;
  if (n_slots < n_regular || n_slots > n_regular + 1)
    return r->fail(P3_E_ARG, "n_slots %d inconsistent with %d regular xforms", n_slots, n_regular);
  final_idx = (int32_t)u[8];
  if (final_idx != -1 && (final_idx < 0 || final_idx >= n_slots))
    return r->fail(P3_E_ARG, "final xform slot %d outside %d slots", final_idx, n_slots);
  if (u[15] > 3u) return r->fail(P3_E_ARG, "colour mode %u is not 0..3", u[15]);
  if (u[14] > 1u) return r->fail(P3_E_ARG, "capture flag %u is not 0 or 1", u[14]);
  if ((int)u[0] != r->sw || (int)u[1] != r->sh)
    return r->fail(P3_E_ARG, "uniform size %ux%u does not match super-res grid %dx%d", u[0], u[1],
                   r->sw, r->sh);
  int rc;
  if (check_records_too && (rc = check_records(r, xforms, n_slots))) return rc;
  if (check_table) {
    const size_t table = (size_t)n_rows * P3_GRAIN;
    for (size_t i = 0; i < table; ++i)
      if (distrib[i] >= (uint32_t)n_regular)
        return r->fail(P3_E_ARG, "pick table entry %zu = %u outside %d xforms", i, distrib[i],
                       n_regular);
  }
  return P3_OK;
}

// The packed accumulator and its carry counts, sized to the grid and zeroed.
int prepare_acc(p3_renderer* r) {
  int rc;
  const size_t bins = grid_bins(r);
  if ((rc = r->hip(r->acc.ensure(bins), "hipMalloc(packed accumulator)"))) return rc;
  if ((rc = r->hip(r->carry.ensure(bins * 4), "hipMalloc(packed carries)"))) return rc;
  if (!r->acc_clean) {
    if ((rc = r->hip(p3::launch_fill(r->acc.p, bins * 8, 0u, r->stream),
                     "clear packed accumulator")) ||
        (rc = r->hip(p3::launch_fill(r->carry.p, bins * 32, 0u, r->stream),
                     "clear packed carries")))
      return rc;
    r->acc_clean = true;
  }
  return P3_OK;
}

// Applies pending adds: every reader of the histogram calls it.
int settle_hist(p3_renderer* r) {
  if (!r->acc_dirty) return P3_OK;
  r->acc_dirty = false;
  r->tic(T_MERGE);
  const int rc = r->hip(p3::launch_hist_merge(r->hist.p, r->acc.p, r->carry.p, grid_bins(r),
                                              r->stream),
                        "packed merge");
  r->toc(T_MERGE);
  return rc;
}

// Drops pending adds when the histogram is overwritten wholesale.
int drop_pending(p3_renderer* r) {
  if (!r->acc_dirty) return P3_OK;
  r->acc_dirty = false;
  int rc;
  if ((rc = r->hip(p3::launch_fill(r->acc.p, grid_bins(r) * 8, 0u, r->stream),
                   "clear packed accumulator")))
    return rc;
  return r->hip(p3::launch_fill(r->carry.p, grid_bins(r) * 32, 0u, r->stream),
                "clear packed carries");
}

// Attaches the packed accumulator to a dispatch.
int attach_adds(p3_renderer* r, p3::ChaosArgs& a) {
  const int rc = prepare_acc(r);
  if (rc) return rc;
  a.acc = r->acc.p;
  a.carry = r->carry.p;
  return P3_OK;
}

// P3_E_STATE when a failed resize left the handle without its buffers.
int usable(p3_renderer* r) {
  if (!r->broken) return P3_OK;
  return r->fail(P3_E_STATE, "a failed resize left the renderer without its buffers; "
                             "p3_resize again or destroy it");
}

}  // namespace

namespace p3h {

p3::ChaosArgs dispatch_args_at(const uint32_t* u, int n_regular, int32_t final_idx,
                               const float* dxf, const uint32_t* dds, const float* pal,
                               uint32_t* hist) {
  p3::ChaosArgs a{};
  a.isaac_in = nullptr;
  a.trace = nullptr;
  a.xforms = dxf;
  a.distrib = dds;
  a.palette = (const float4*)pal;
  a.hist = hist;
  a.fallback_row = (uint32_t)n_regular;
  a.width = u[0]; a.height = u[1]; a.iters = u[2]; a.fuse = u[3];
  a.scale = bits_f(u[4]); a.cx = bits_f(u[5]); a.cy = bits_f(u[6]);
  a.seed = u[7];
  a.final_idx = final_idx;
  const float rot = bits_f(u[9]);
  a.cos_r = std::cos(rot);
  a.sin_r = std::sin(rot);
  a.palette_mode = u[10];
  a.walkers = u[12];
  a.jitter = bits_f(u[13]);
  a.xform_blend = bits_f(u[30]);
  a.color_mode = u[15];
  a.flow_strength = bits_f(u[16]); a.flow_scale = bits_f(u[17]);
  a.trap_kind = u[18]; a.trap_mode = u[19];
  a.trap_cx = bits_f(u[20]); a.trap_cy = bits_f(u[21]); a.trap_radius = bits_f(u[22]);
  a.trap_nx = bits_f(u[23]); a.trap_ny = bits_f(u[24]); a.trap_falloff = bits_f(u[25]);
  a.trap_freq = bits_f(u[26]); a.trap_strength = bits_f(u[27]);
  a.phase_strength = bits_f(u[28]); a.phase_freq = bits_f(u[29]);
  a.idx_sum = nullptr;
  return a;
}

std::string de_params(const uint32_t* de, p3::DeArgs& a) {
  a.max_rad = bits_f(de[2]); a.min_rad = bits_f(de[3]); a.curve = bits_f(de[4]);
  a.k1 = bits_f(de[5]); a.k2 = bits_f(de[6]);
  if (!(a.max_rad >= 0.0f) || std::isinf(a.max_rad))
    return "DE max radius must be finite and >= 0";
  a.R = (int)std::ceil(std::min(std::max(a.max_rad, 0.0f), 256.0f));
  if (a.R > 254)
    return "DE search radius " + std::to_string(a.R) + " above the 254 limit";
  return std::string();
}

void tm_params(const uint32_t* viz, p3::TmArgs& t) {
  t.fwidth = viz[9];
  t.k1 = bits_f(viz[2]); t.k2 = bits_f(viz[3]); t.gamma = bits_f(viz[4]);
  t.vibrancy = bits_f(viz[5]); t.highpow = bits_f(viz[6]); t.linrange = bits_f(viz[7]);
  t.bg0 = bits_f(viz[12]); t.bg1 = bits_f(viz[13]); t.bg2 = bits_f(viz[14]);
  t.transparent = viz[21];
  t.curves_mask = viz[16];
  t.hsl_on = viz[17];
  t.hsl_hue = bits_f(viz[18]); t.hsl_sat = bits_f(viz[19]); t.hsl_light = bits_f(viz[20]);
}

}  // namespace p3h

namespace {

// Kernel arguments of one dispatch on the renderer's palette and histogram.
p3::ChaosArgs dispatch_args(p3_renderer* r, const uint32_t* u, int n_regular, int32_t final_idx,
                            const float* dxf, const uint32_t* dds) {
  return dispatch_args_at(u, n_regular, final_idx, dxf, dds, r->palette.p, r->hist.p);
}

// Capture buffer of the colour index (allocated and cleared on first use,
// cleared by every reset afterwards) when the uniforms ask for it.
int attach_capture(p3_renderer* r, p3::ChaosArgs& a, const uint32_t* u) {
  if (u[14] != 1u) return P3_OK;
  int rc;
  if (!r->idx_live) {
    if ((rc = r->hip(r->idx.ensure(grid_bins(r)), "hipMalloc(capture)"))) return rc;
    if ((rc = r->hip(p3::launch_fill(r->idx.p, grid_bins(r) * 4, 0u, r->stream), "clear capture")))
      return rc;
  }
  r->idx_live = true;
  a.idx_sum = r->idx.p;
  return P3_OK;
}

// Trace buffer (cleared) for the configured walkers x turns, or null.
int attach_trace(p3_renderer* r, p3::ChaosArgs& a) {
  if (r->trace_walkers <= 0 || r->trace_iters <= 0) return P3_OK;
  const size_t n = (size_t)r->trace_walkers * r->trace_iters * P3_TRACE_FLOATS;
  int rc;
  if ((rc = r->hip(r->trace.ensure(n), "hipMalloc(trace)"))) return rc;
  if ((rc = r->hip(p3::launch_fill(r->trace.p, n * 4, 0u, r->stream), "clear trace"))) return rc;
  a.trace = r->trace.p;
  a.trace_walkers = (uint32_t)r->trace_walkers;
  a.trace_iters = (uint32_t)r->trace_iters;
  return P3_OK;
}

int iterate_impl(p3_renderer* r, const float* xforms, int n_slots, const uint32_t* distrib,
                 int n_rows, const uint32_t* u, const uint32_t* states) {
  if (!r) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  if (!r->palette_set) return r->fail(P3_E_STATE, "iterate before reset (no palette)");
  if (!xforms || !distrib || !u) return r->fail(P3_E_ARG, "null argument");
  const size_t n_xf = (size_t)P3_XFORM_FLOATS * (n_slots > 0 ? n_slots : 0);
  const size_t table = (size_t)(n_rows > 0 ? n_rows : 0) * P3_GRAIN;
  // Blobs identical to the last upload (successive frames or chunks of one
  // genome) are neither re-validated nor re-sent.
  const bool new_xf = r->last_xforms.size() != n_xf ||
                      std::memcmp(r->last_xforms.data(), xforms, n_xf * sizeof(float)) != 0;
  const bool new_ds = r->last_distrib.size() != table ||
                      std::memcmp(r->last_distrib.data(), distrib, table * sizeof(uint32_t)) != 0;
  int n_regular = 0;
  int32_t final_idx = -1;
  int rc;
  if ((rc = validate_dispatch(r, xforms, n_slots, distrib, n_rows, u, new_xf, new_ds, n_regular,
                              final_idx)))
    return rc;
  const uint32_t walkers = u[12];
  if (walkers == 0) return P3_OK;
  if (new_xf) {
    if ((rc = r->hip(hipMemcpyAsync(r->xforms.p, xforms, n_xf * sizeof(float),
                                    hipMemcpyHostToDevice, r->stream),
                     "upload xforms")))
      return rc;
    r->last_xforms.assign(xforms, xforms + n_xf);
  }
  if (new_ds) {
    if ((rc = r->hip(hipMemcpyAsync(r->distrib.p, distrib, table * sizeof(uint32_t),
                                    hipMemcpyHostToDevice, r->stream),
                     "upload pick table")))
      return rc;
    r->bounds_ok = pick_bounds(distrib, n_rows, r->bounds_host);
    if (r->bounds_ok) {
      if ((rc = r->hip(r->bounds.ensure((size_t)p3::kPickRows * p3::kPickCols),
                       "hipMalloc(pick bounds)")))
        return rc;
      if ((rc = r->hip(hipMemcpyAsync(r->bounds.p, r->bounds_host.data(),
                                      r->bounds_host.size() * sizeof(uint16_t),
                                      hipMemcpyHostToDevice, r->stream),
                       "upload pick bounds")))
        return rc;
    }
    r->last_distrib.assign(distrib, distrib + table);
  }
  p3::ChaosArgs a = dispatch_args(r, u, n_regular, final_idx, r->xforms.p, r->distrib.p);
  a.pick_bounds = r->bounds_ok ? r->bounds.p : nullptr;
  if (states) {
    if ((rc = r->hip(r->isaac.ensure((size_t)walkers * 36), "hipMalloc(isaac)"))) return rc;
    if ((rc = r->hip(hipMemcpyAsync(r->isaac.p, states, sizeof(uint32_t) * 36 * walkers,
                                    hipMemcpyHostToDevice, r->stream),
                     "upload isaac states")))
      return rc;
    a.isaac_in = r->isaac.p;
  }
  if ((rc = attach_trace(r, a))) return rc;
  if ((rc = attach_capture(r, a, u))) return rc;
  if ((rc = attach_adds(r, a))) return rc;
  std::string rtc_err;
  {  // compile (or fetch) the kernel for this shape before the timed launch
    const std::vector<char>* code = nullptr;
    if (!p3::chaos_code(xforms, n_regular, final_idx, &code, rtc_err))
      return r->fail(P3_E_HIP, "chaos kernel specialisation: %s", rtc_err.c_str());
  }
  r->tic(T_CHAOS);
  const hipError_t le = p3::launch_chaos_rtc(a, xforms, n_regular, final_idx, r->stream, rtc_err);
  if (le != hipSuccess)
    return r->fail(P3_E_HIP, "chaos launch: %s %s", hipGetErrorString(le), rtc_err.c_str());
  r->acc_dirty = true;
  r->toc(T_CHAOS);
  r->have_image = false;
  return P3_OK;
}

// n sub-dispatches of one genome shape (motion-blur sub-frames, chunks with
// their own seeds) in one batched launch; each keeps its own records, pick
// table, uniforms (seed, walkers, camera) and walker numbering.
int iterate_multi_impl(p3_renderer* r, int n_sub, const float* const* xforms, int n_slots,
                       const uint32_t* const* distribs, int n_rows,
                       const uint32_t* const* uniforms) {
  if (!r) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  if (!r->palette_set) return r->fail(P3_E_STATE, "iterate before reset (no palette)");
  if (n_sub < 1 || n_sub > p3::kMaxSub)
    return r->fail(P3_E_ARG, "n_sub must be 1..%d (got %d)", p3::kMaxSub, n_sub);
  if (!xforms || !distribs || !uniforms) return r->fail(P3_E_ARG, "null argument");
  int rc, n_regular = 0;
  int32_t final_idx = -1;
  std::string key0;
  for (int k = 0; k < n_sub; ++k) {
    int nr;
    int32_t fi;
    if ((rc = validate_dispatch(r, xforms[k], n_slots, distribs[k], n_rows, uniforms[k], true,
                                true, nr, fi)))
      return rc;
    const std::string key = p3::chaos_shape_key(xforms[k], nr, fi);
    if (k == 0) {
      key0 = key;
      n_regular = nr;
      final_idx = fi;
    } else if (key != key0) {
      return r->fail(P3_E_ARG, "sub-dispatch %d has a different genome shape from sub-dispatch 0",
                     k);
    }
  }
  const size_t n_xf = (size_t)P3_XFORM_FLOATS * n_slots, table = (size_t)n_rows * P3_GRAIN;
  if ((rc = r->hip(r->mxforms.ensure(n_xf * n_sub), "hipMalloc(batched xforms)"))) return rc;
  if ((rc = r->hip(r->mdistrib.ensure(table * n_sub), "hipMalloc(batched pick tables)"))) return rc;
  if ((rc = r->hip(r->margs.ensure(sizeof(p3::ChaosArgs) * n_sub), "hipMalloc(batched args)")))
    return rc;
  const size_t nb = (size_t)p3::kPickRows * p3::kPickCols;
  if ((rc = r->hip(r->mbounds.ensure(nb * n_sub), "hipMalloc(batched pick bounds)"))) return rc;
  std::vector<uint16_t>& bounds = r->mbounds_host;
  bounds.assign(nb * n_sub, (uint16_t)P3_GRAIN);
  std::vector<p3::ChaosArgs>& args = r->margs_host;
  args.assign((size_t)n_sub, p3::ChaosArgs{});
  p3::MultiArgs M{};
  M.args = (const p3::ChaosArgs*)r->margs.p;
  M.n = (uint32_t)n_sub;
  for (int k = 0; k < n_sub; ++k) {
    float* dxf = r->mxforms.p + n_xf * k;
    uint32_t* dds = r->mdistrib.p + table * k;
    if ((rc = r->hip(hipMemcpyAsync(dxf, xforms[k], n_xf * sizeof(float), hipMemcpyHostToDevice,
                                    r->stream),
                     "upload batched xforms")))
      return rc;
    if ((rc = r->hip(hipMemcpyAsync(dds, distribs[k], table * sizeof(uint32_t),
                                    hipMemcpyHostToDevice, r->stream),
                     "upload batched pick table")))
      return rc;
    args[k] = dispatch_args(r, uniforms[k], n_regular, final_idx, dxf, dds);
    std::vector<uint16_t> b;
    if (pick_bounds(distribs[k], n_rows, b)) {
      std::copy(b.begin(), b.end(), bounds.begin() + nb * k);
      args[k].pick_bounds = r->mbounds.p + nb * k;
    }
    if (k == 0 && (rc = attach_trace(r, args[0]))) return rc;
    if ((rc = attach_capture(r, args[k], uniforms[k]))) return rc;
    const uint32_t blk = (uint32_t)p3::chaos_block();
    M.start[k + 1] = M.start[k] + (args[k].walkers + blk - 1) / blk;
  }
  if (M.start[n_sub] == 0) return P3_OK;
  for (int k = 0; k < n_sub; ++k)
    if ((rc = attach_adds(r, args[k]))) return rc;
  if ((rc = r->hip(hipMemcpyAsync(r->mbounds.p, bounds.data(), bounds.size() * sizeof(uint16_t),
                                  hipMemcpyHostToDevice, r->stream),
                   "upload batched pick bounds")))
    return rc;
  if ((rc = r->hip(hipMemcpyAsync(r->margs.p, args.data(), sizeof(p3::ChaosArgs) * n_sub,
                                  hipMemcpyHostToDevice, r->stream),
                   "upload batched args")))
    return rc;
  std::string rtc_err;
  {
    const std::vector<char>* code = nullptr;
    if (!p3::chaos_code(xforms[0], n_regular, final_idx, &code, rtc_err))
      return r->fail(P3_E_HIP, "chaos kernel specialisation: %s", rtc_err.c_str());
  }
  r->tic(T_CHAOS);
  const hipError_t le =
      p3::launch_chaos_rtc_multi(M, xforms[0], n_regular, final_idx, r->stream, rtc_err);
  if (le != hipSuccess)
    return r->fail(P3_E_HIP, "batched chaos launch: %s %s", hipGetErrorString(le),
                   rtc_err.c_str());
  r->acc_dirty = true;
  r->toc(T_CHAOS);
  r->have_image = false;
  return P3_OK;
}

int read_dev(p3_renderer* r, void* dst, const void* src, size_t bytes, const char* what) {
  int rc = r->hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, r->stream), what);
  if (rc) return rc;
  return r->hip(hipStreamSynchronize(r->stream), what);
}

// ------------------------------------------- allocation-free dispatch --
// Device memory of p3_iterate_noalloc, one copy per GPU that loads the
// library, held until the process ends. Laid out per call as 256-byte
// aligned parts: packed accumulator (8 B per bin), carry counts (32 B per
// bin), xform records, pick table, pick-table step boundaries, palette.
// In host code `scratch` names a host-side stand-in, not the device array:
// kernels, fills and copies take the address that hipGetSymbolAddress
// returns for the current device (scratch_base), never the name.
static __device__ __align__(256) unsigned char scratch[P3_NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_dev;                       // per device
std::vector<std::vector<uint16_t>> g_scratch_bounds;             // per device: host staging

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

int noalloc_fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_create_error = buf;
  return code;
}

// This device's scratch address, looked up once per device with the device
// current; null on failure (message in g_create_error).
unsigned char* scratch_base(int dev) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((size_t)dev >= g_scratch_dev.size()) {
    g_scratch_dev.resize((size_t)dev + 1, nullptr);
    g_scratch_bounds.resize((size_t)dev + 1);
  }
  if (!g_scratch_dev[(size_t)dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      noalloc_fail(P3_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                   hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_dev[(size_t)dev] = (unsigned char*)p;
    g_scratch_bounds[(size_t)dev].assign((size_t)p3::kPickRows * p3::kPickCols, 0);
  }
  return g_scratch_dev[(size_t)dev];
}

}  // namespace

extern "C" {

int p3_create(int width, int height, int oversample, const float* taps, int n_taps,
              p3_renderer** out) {
  if (!out) { g_create_error = "null output handle"; return P3_E_ARG; }
  *out = nullptr;
  if (width <= 0 || height <= 0 || oversample < 1 || oversample > 16 || !taps || n_taps < 3 ||
      n_taps > 17) {
    g_create_error = "invalid size, oversample or taps";
    return P3_E_ARG;
  }
  p3_renderer* r = new p3_renderer();
  hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_create_error = std::string("hipStreamCreate: ") + hipGetErrorString(e);
    delete r;
    return P3_E_HIP;
  }
  for (int t = 0; t < kTimers; ++t) {
    (void)hipEventCreate(&r->ev[t][0]);
    (void)hipEventCreate(&r->ev[t][1]);
  }
  int rc = build(r, width, height, oversample, taps, n_taps);
  if (rc) {
    g_create_error = r->err;
    p3_destroy(r);
    return rc;
  }
  *out = r;
  return P3_OK;
}

int p3_resize(p3_renderer* r, int width, int height, int oversample, const float* taps,
              int n_taps) {
  if (!r) return P3_E_ARG;
  // Keeps the allocations (build grows them when needed; freeing one waits
  // for the device), so work already queued on the stream stays valid and
  // the host does not wait for it.
  return build(r, width, height, oversample, taps, n_taps);
}

void p3_destroy(p3_renderer* r) {
  if (!r) return;
  if (r->stream) (void)hipStreamSynchronize(r->stream);
  release(r);
  for (int t = 0; t < kTimers; ++t) {
    if (r->ev[t][0]) (void)hipEventDestroy(r->ev[t][0]);
    if (r->ev[t][1]) (void)hipEventDestroy(r->ev[t][1]);
  }
  if (r->stream) (void)hipStreamDestroy(r->stream);
  delete r;
}

int p3_reset(p3_renderer* r, const float* palette_rgba) {
  if (!r) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  if (!palette_rgba) return r->fail(P3_E_ARG, "null palette");
  int rc;
  if ((rc = r->hip(hipMemcpyAsync(r->palette.p, palette_rgba, sizeof(float) * 1024,
                                  hipMemcpyHostToDevice, r->stream),
                   "upload palette")))
    return rc;
  r->tic(T_RESET);
  if ((rc = r->hip(p3::launch_fill(r->hist.p, grid_bins(r) * 16, 0u, r->stream), "clear histogram")))
    return rc;
  if ((rc = drop_pending(r))) return rc;
  if (r->idx_live &&
      (rc = r->hip(p3::launch_fill(r->idx.p, grid_bins(r) * 4, 0u, r->stream), "clear capture")))
    return rc;
  r->toc(T_RESET);
  r->palette_set = true;
  r->have_image = r->have_filtered = false;
  return P3_OK;
}

int p3_iterate(p3_renderer* r, const float* xforms, int n_slots, const uint32_t* distrib,
               int n_rows, const uint32_t* uniforms) {
  return iterate_impl(r, xforms, n_slots, distrib, n_rows, uniforms, nullptr);
}

int p3_iterate_multi(p3_renderer* r, int n_sub, const float* const* xforms, int n_slots,
                     const uint32_t* const* distribs, int n_rows,
                     const uint32_t* const* uniforms) {
  return iterate_multi_impl(r, n_sub, xforms, n_slots, distribs, n_rows, uniforms);
}

int p3_iterate_states(p3_renderer* r, const float* xforms, int n_slots, const uint32_t* distrib,
                      int n_rows, const uint32_t* uniforms, const uint32_t* isaac_states) {
  if (!isaac_states) return r ? r->fail(P3_E_ARG, "null isaac states") : P3_E_ARG;
  return iterate_impl(r, xforms, n_slots, distrib, n_rows, uniforms, isaac_states);
}

int p3_present(p3_renderer* r, const uint32_t* de, const uint32_t* viz) {
  if (!r) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  if (!viz) return r->fail(P3_E_ARG, "null tonemap uniforms");
  if ((int)viz[0] != r->width || (int)viz[1] != r->height || (int)viz[8] != r->os)
    return r->fail(P3_E_ARG, "tonemap uniforms %ux%u os %u do not match renderer %dx%d os %d",
                   viz[0], viz[1], viz[8], r->width, r->height, r->os);
  if ((int)viz[9] != r->n_taps)
    return r->fail(P3_E_ARG, "tonemap fwidth %u does not match %d taps", viz[9], r->n_taps);
  int rc;
  if ((rc = settle_hist(r))) return rc;
  if (de) {
    if ((int)de[0] != r->sw || (int)de[1] != r->sh)
      return r->fail(P3_E_ARG, "DE uniforms %ux%u do not match super-res grid %dx%d", de[0], de[1],
                     r->sw, r->sh);
    p3::DeArgs a{};
    const std::string bad = de_params(de, a);
    if (!bad.empty()) return r->fail(P3_E_ARG, "%s", bad.c_str());
    a.hist = (const uint4*)r->hist.p;
    a.pre = r->pre.p;
    a.rad = r->rad.p;
    a.tile_rmax = r->tile_rmax.p;
    a.out = r->filtered.p;
    a.lut = r->lut.p;
    a.width = (uint32_t)r->sw; a.height = (uint32_t)r->sh;
    a.tiles_x = (uint32_t)((r->sw + 15) / 16);
    a.tiles_y = (uint32_t)((r->sh + 15) / 16);
    r->tic(T_DE);
    if ((rc = r->hip(p3::launch_density(a, r->stream), "density launch"))) return rc;
    r->toc(T_DE);
    r->have_filtered = true;
  }
  p3::TmArgs t{};
  t.filtered = r->filtered.p;
  t.hist = (const uint4*)r->hist.p;
  t.taps = r->taps.p;
  t.out8 = r->image8.p;
  t.outf = r->want_image_f32 ? r->imagef.p : nullptr;
  t.width = (uint32_t)r->width; t.height = (uint32_t)r->height;
  t.os = (uint32_t)r->os;
  tm_params(viz, t);
  r->last_k2 = t.k2;
  r->last_de = de != nullptr;
  if (viz[16] > 31u) return r->fail(P3_E_ARG, "curves mask %u has bits above 4", viz[16]);
  if (viz[16] != 0u && !r->curves_set)
    return r->fail(P3_E_STATE, "curves mask %u set but no curves LUT uploaded (p3_set_curves)",
                   viz[16]);
  t.curves = r->curves.p;
  r->tic(T_TM);
  if ((rc = r->hip(p3::launch_tonemap(t, de != nullptr, r->stream), "tonemap launch"))) return rc;
  r->toc(T_TM);
  r->have_image = true;
  r->have_image_f32 = r->want_image_f32;
  return P3_OK;
}

int p3_probe_atomics(p3_renderer* r, uint32_t blocks, uint32_t per_thread,
                     double* requests_per_s, int* requests_per_splat) {
  if (!r || !requests_per_s || !requests_per_splat) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  if (blocks == 0 || blocks > 65535 || per_thread == 0 || per_thread > 4096)
    return r->fail(P3_E_ARG, "probe size out of range (blocks 1..65535, per_thread 1..4096)");
  const uint32_t cells = (uint32_t)((size_t)r->sw * r->sh);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc;
  if ((rc = drop_pending(r))) return rc;
  if ((rc = prepare_acc(r))) return rc;   // the requests land in the accumulator
  if ((rc = r->hip(hipEventCreate(&e0), "hipEventCreate")) ||
      (rc = r->hip(hipEventCreate(&e1), "hipEventCreate")))
    return rc;
  float ms = 0.0f;
  rc = r->hip(hipEventRecord(e0, r->stream), "hipEventRecord");
  if (!rc) rc = r->hip(p3::launch_atomic_probe(r->acc.p, cells, blocks, per_thread, 12345u,
                                                r->stream), "atomic probe launch");
  if (!rc) rc = r->hip(hipEventRecord(e1, r->stream), "hipEventRecord");
  if (!rc) rc = r->hip(hipEventSynchronize(e1), "hipEventSynchronize");
  if (!rc) rc = r->hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  // The accumulator now holds probe counts (no carries are counted: exact
  // below 2^16 adds per cell); they are merged into the histogram at its
  // next read, so reset before the next render.
  r->have_image = false;
  r->acc_dirty = true;
  *requests_per_splat = 1;
  *requests_per_s = 256.0 * blocks * per_thread / (ms * 1e-3);
  return P3_OK;
}

int p3_read_linear(p3_renderer* r, double k1, double k2, float* out, size_t n) {
  if (!r) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  const size_t want = (size_t)r->width * r->height * 4;
  if (!out || n != want) return r->fail(P3_E_ARG, "linear export needs %zu floats", want);
  if (!std::isfinite(k1) || !std::isfinite(k2))
    return r->fail(P3_E_ARG, "linear export calibration must be finite");
  int rc;
  if ((rc = settle_hist(r))) return rc;
  if ((rc = r->hip(r->linear.ensure((size_t)r->width * r->height), "hipMalloc(linear)"))) return rc;
  p3::LinArgs a{};
  a.hist = (const uint4*)r->hist.p;
  a.out = r->linear.p;
  a.width = (uint32_t)r->width; a.height = (uint32_t)r->height; a.os = (uint32_t)r->os;
  a.scale = (k1 * k2) / (double)((size_t)r->os * r->os);
  if ((rc = r->hip(p3::launch_linear_export(a, r->stream), "linear export launch"))) return rc;
  return read_dev(r, out, r->linear.p, want * sizeof(float), "read linear export");
}

int p3_set_curves(p3_renderer* r, const float* lut, size_t n) {
  if (!r) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  if (!lut || n == 0) {
    r->curves_set = false;
    return P3_OK;
  }
  if (n != 5 * 256) return r->fail(P3_E_ARG, "curves LUT needs 1280 floats (5 x 256), got %zu", n);
  for (size_t i = 0; i < n; ++i)
    if (!(lut[i] >= 0.0f && lut[i] <= 1.0f))
      return r->fail(P3_E_ARG, "curves LUT entry %zu is not in [0, 1]", i);
  int rc;
  if ((rc = r->hip(r->curves.ensure(n), "hipMalloc(curves)"))) return rc;
  r->curves_host.assign(lut, lut + n);   // stays alive while the async copy runs
  if ((rc = r->hip(hipMemcpyAsync(r->curves.p, r->curves_host.data(), n * sizeof(float),
                                  hipMemcpyHostToDevice, r->stream),
                   "upload curves")))
    return rc;
  r->curves_set = true;
  return P3_OK;
}

int p3_read_histogram(p3_renderer* r, uint32_t* out, size_t n) {
  if (!r) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  if (!out || n != grid_bins(r) * 4)
    return r->fail(P3_E_ARG, "histogram needs %zu words", grid_bins(r) * 4);
  const int rc = settle_hist(r);
  if (rc) return rc;
  return read_dev(r, out, r->hist.p, n * 4, "read histogram");
}

int p3_read_index(p3_renderer* r, uint32_t* out, size_t n) {
  if (!r) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  if (!r->idx_live) return r->fail(P3_E_STATE, "no dispatch has captured the colour index");
  if (!out || n != grid_bins(r))
    return r->fail(P3_E_ARG, "capture buffer needs %zu words", grid_bins(r));
  return read_dev(r, out, r->idx.p, n * 4, "read capture");
}

int p3_read_filtered(p3_renderer* r, float* out, size_t n) {
  if (!r) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  if (!out || n != grid_bins(r) * 4)
    return r->fail(P3_E_ARG, "filtered grid needs %zu floats", grid_bins(r) * 4);
  if (!r->have_filtered) return r->fail(P3_E_STATE, "no DE output yet");
  return read_dev(r, out, r->filtered.p, n * 4, "read filtered");
}

int p3_read_image(p3_renderer* r, uint8_t* out, size_t n) {
  if (!r) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  if (!out || n != (size_t)r->width * r->height * 4)
    return r->fail(P3_E_ARG, "image needs %zu bytes", (size_t)r->width * r->height * 4);
  if (!r->have_image) return r->fail(P3_E_STATE, "present has not run since the last change");
  return read_dev(r, out, r->image8.p, n, "read image");
}

int p3_read_image_f32(p3_renderer* r, float* out, size_t n) {
  if (!r) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  if (!out || n != (size_t)r->width * r->height * 4)
    return r->fail(P3_E_ARG, "image needs %zu floats", (size_t)r->width * r->height * 4);
  if (!r->have_image || !r->have_image_f32)
    return r->fail(P3_E_STATE, "no f32 image: enable it with p3_set_image_f32 before present");
  return read_dev(r, out, r->imagef.p, n * 4, "read image f32");
}

int p3_write_histogram(p3_renderer* r, const uint32_t* data, size_t n) {
  if (!r) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  if (!data || n != grid_bins(r) * 4)
    return r->fail(P3_E_ARG, "histogram needs %zu words", grid_bins(r) * 4);
  int rc = drop_pending(r);
  if (rc) return rc;
  if ((rc = r->hip(hipMemcpyAsync(r->hist.p, data, n * 4, hipMemcpyHostToDevice, r->stream),
                   "write histogram")))
    return rc;
  r->have_image = r->have_filtered = false;
  return r->hip(hipStreamSynchronize(r->stream), "write histogram");
}

int p3_set_trace(p3_renderer* r, int walkers, int iters) {
  if (!r) return P3_E_ARG;
  if (walkers < 0 || iters < 0 || (long long)walkers * iters > (1LL << 26))
    return r->fail(P3_E_ARG, "trace %d x %d out of range", walkers, iters);
  r->trace_walkers = walkers;
  r->trace_iters = iters;
  return P3_OK;
}

int p3_read_trace(p3_renderer* r, float* out, size_t n) {
  if (!r) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  const size_t need = (size_t)r->trace_walkers * r->trace_iters * P3_TRACE_FLOATS;
  if (!out || n != need || need == 0) return r->fail(P3_E_ARG, "trace needs %zu floats", need);
  if (r->trace.n < need) return r->fail(P3_E_STATE, "no trace recorded");
  return read_dev(r, out, r->trace.p, need * 4, "read trace");
}

int p3_set_image_f32(p3_renderer* r, int enabled) {
  if (!r) return P3_E_ARG;
  r->want_image_f32 = enabled != 0;
  return P3_OK;
}

int p3_set_timing(p3_renderer* r, int enabled) {
  if (!r) return P3_E_ARG;
  r->timing = enabled != 0;
  for (auto& u : r->ev_used) u = false;
  return P3_OK;
}

int p3_read_timings(p3_renderer* r, float* out6) {
  if (!r || !out6) return P3_E_ARG;
  int rc = r->hip(hipStreamSynchronize(r->stream), "sync for timings");
  if (rc) return rc;
  for (int t = 0; t < kTimers; ++t) {
    out6[t] = 0.0f;
    if (r->ev_used[t]) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, r->ev[t][0], r->ev[t][1]) == hipSuccess) out6[t] = ms;
    }
  }
  return P3_OK;
}

int p3_sync(p3_renderer* r) {
  if (!r) return P3_E_ARG;
  return r->hip(hipStreamSynchronize(r->stream), "stream sync");
}

int p3_dims(p3_renderer* r, int* out4) {
  if (!r || !out4) return P3_E_ARG;
  out4[0] = r->width; out4[1] = r->height; out4[2] = r->sw; out4[3] = r->sh;
  return P3_OK;
}

// Argument checks shared by p3_chaos_prepare and p3_chaos_ready.
int check_shape_args(const float* xforms, int n_slots, int n_rows, int final_idx) {
  if (!xforms || n_rows < 2 || n_rows > P3_MAX_XFORMS + 1) return P3_E_ARG;
  const int n_regular = n_rows - 1;
  if (n_slots < n_regular || n_slots > n_regular + 1) return P3_E_ARG;
  if (final_idx != -1 && (final_idx < 0 || final_idx >= n_slots)) return P3_E_ARG;
  return P3_OK;
}

int p3_chaos_ready(const float* xforms, int n_slots, int n_rows, int final_idx) {
  if (check_shape_args(xforms, n_slots, n_rows, final_idx)) return P3_E_ARG;
  return p3::chaos_code_ready(xforms, n_rows - 1, final_idx) ? 1 : 0;
}

int p3_chaos_prepare(const float* xforms, int n_slots, int n_rows, int final_idx) {
  if (!xforms || n_rows < 2 || n_rows > P3_MAX_XFORMS + 1) return P3_E_ARG;
  const int n_regular = n_rows - 1;
  if (n_slots < n_regular || n_slots > n_regular + 1) return P3_E_ARG;
  if (final_idx != -1 && (final_idx < 0 || final_idx >= n_slots)) return P3_E_ARG;
  for (int s = 0; s < n_slots; ++s) {
    const float* xf = xforms + (size_t)s * P3_XFORM_FLOATS;
    const int nv = xf[7] >= 0.0f && xf[7] < 8.0f ? (int)xf[7] : (xf[7] >= 8.0f ? 8 : -1);
    if (nv < 0) return P3_E_ARG;
    for (int k = 0; k < nv; ++k) {
      const float idxf = xf[20 + 4 * k];
      if (!(idxf >= 0.0f) || idxf > 1e9f || !supported((uint32_t)idxf)) return P3_E_ARG;
    }
  }
  const std::vector<char>* code = nullptr;
  if (!p3::chaos_code(xforms, n_regular, final_idx, &code, g_create_error)) return P3_E_HIP;
  return P3_OK;
}

int p3_chaos_source(const float* xforms, int n_slots, int n_rows, int final_idx, char* out,
                    size_t cap, size_t* len) {
  if (!xforms || n_rows < 2 || n_rows > P3_MAX_XFORMS + 1 || n_slots < n_rows - 1) return P3_E_ARG;
  (void)n_slots;
  const std::string src = p3::chaos_source(xforms, n_rows - 1, final_idx);
  if (len) *len = src.size();
  if (out && cap > 0) {
    const size_t n = std::min(cap - 1, src.size());
    std::memcpy(out, src.data(), n);
    out[n] = '\0';
  }
  return P3_OK;
}

int p3_variation_supported(int index) { return index >= 0 && supported((uint32_t)index) ? 1 : 0; }

// ---------------------------------------------------------------- denoiser

size_t p3_denoiser_param_count(void) {
  size_t n = p3::kDenoiseHead;
  for (int l = 0; l < p3::kDenoiseLayers; ++l)
    n += (size_t)p3::kDenoiseShape[l][0] * p3::kDenoiseShape[l][1] * 9 + p3::kDenoiseShape[l][1];
  return n;
}

int p3_denoiser_create(const float* params, size_t n, p3_denoiser** out) {
  if (!out) return P3_E_ARG;
  *out = nullptr;
  if (!params || n != p3_denoiser_param_count()) {
    g_create_error = "denoiser: parameter count does not match the network";
    return P3_E_ARG;
  }
  p3_denoiser* d = new p3_denoiser();
  // Each 3x3 layer's weight [c_out][c_in][3][3] packed as bf16 [c_out][kp]
  // with k = tap * c_in + c (zero past 9 c_in); biases f32.
  std::vector<uint16_t> w;
  std::vector<float> b;
  const float* q = params;
  for (int l = 0; l < p3::kDenoiseLayers; ++l) {
    const int ci = p3::kDenoiseShape[l][0], co = p3::kDenoiseShape[l][1];
    const int kp = p3::denoise_kp(l);
    d->w_off[l] = w.size();
    d->b_off[l] = b.size();
    w.resize(w.size() + (size_t)co * kp, 0);
    uint16_t* dst = w.data() + d->w_off[l];
    for (int o = 0; o < co; ++o)
      for (int c = 0; c < ci; ++c)
        for (int t = 0; t < 9; ++t)
          dst[(size_t)o * kp + t * ci + c] = bf16_bits(q[((size_t)o * ci + c) * 9 + t]);
    q += (size_t)co * ci * 9;
    b.insert(b.end(), q, q + co);
    q += co;
  }
  std::vector<float> head(q, q + p3::kDenoiseHead);
  hipError_t e;
  if ((e = d->w.ensure(w.size())) != hipSuccess || (e = d->b.ensure(b.size())) != hipSuccess ||
      (e = d->head.ensure(head.size())) != hipSuccess ||
      (e = hipMemcpy(d->w.p, w.data(), w.size() * 2, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d->b.p, b.data(), b.size() * 4, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d->head.p, head.data(), head.size() * 4, hipMemcpyHostToDevice)) !=
          hipSuccess) {
    g_create_error = std::string("denoiser upload: ") + hipGetErrorString(e);
    p3_denoiser_destroy(d);
    return P3_E_HIP;
  }
  *out = d;
  return P3_OK;
}

int p3_denoiser_destroy(p3_denoiser* d) {
  if (!d) return P3_OK;
  d->w.release(); d->b.release(); d->head.release();
  for (auto& a : d->act) a.release();
  delete d;
  return P3_OK;
}

int p3_denoise(p3_renderer* r, p3_denoiser* d) {
  if (!r) return P3_E_ARG;
  if (usable(r)) return P3_E_STATE;
  if (!d) return r->fail(P3_E_ARG, "null denoiser");
  if (r->os != 1) return r->fail(P3_E_STATE, "the denoiser runs at oversample 1");
  if (!r->have_image || !r->have_image_f32 || r->last_de)
    return r->fail(P3_E_STATE,
                   "denoise needs a present without DE and with the f32 image enabled");
  const int W = r->width, H = r->height;
  size_t lw[4] = {(size_t)W, 0, 0, 0}, lh[4] = {(size_t)H, 0, 0, 0};
  for (int i = 1; i < 4; ++i) { lw[i] = (lw[i - 1] + 1) / 2; lh[i] = (lh[i - 1] + 1) / 2; }
  int rc;
  p3::DenoiseArgs a{};
  for (int l = 0; l < p3::kDenoiseLayers - 1; ++l) {
    const int lev = p3::kDenoiseShape[l][2];
    if ((rc = r->hip(d->act[l].ensure(lw[lev] * lh[lev] * p3::kDenoiseShape[l][1]),
                     "hipMalloc(denoise activations)")))
      return rc;
    a.act[l] = d->act[l].p;
  }
  for (int l = 0; l < p3::kDenoiseLayers; ++l) {
    a.w[l] = d->w.p + d->w_off[l];
    a.b[l] = d->b.p + d->b_off[l];
  }
  a.head = d->head.p;
  a.img = r->imagef.p;
  a.hist = r->hist.p;
  a.out8 = r->image8.p;
  a.k2 = r->last_k2;
  a.width = W;
  a.height = H;
  r->tic(T_DENOISE);
  if ((rc = r->hip(p3::launch_denoise(a, r->stream), "denoise"))) return rc;
  r->toc(T_DENOISE);
  return P3_OK;
}

int p3_iterate_noalloc(const float* xforms, int n_slots, const uint32_t* distrib, int n_rows,
                       const uint32_t* u, const float* palette_rgba, uint32_t* hist,
                       size_t n_hist, void* stream_) {
  // every check that needs no device comes first
  if (!xforms || !distrib || !u || !palette_rgba || !hist)
    return noalloc_fail(P3_E_ARG, "null argument");
  if (n_rows < 2 || n_rows > P3_MAX_XFORMS + 1)
    return noalloc_fail(P3_E_ARG, "n_rows must be 2..%d (got %d)", P3_MAX_XFORMS + 1, n_rows);
  const int n_regular = n_rows - 1;
  if (n_slots < n_regular || n_slots > n_regular + 1)
    return noalloc_fail(P3_E_ARG, "n_slots %d inconsistent with %d regular xforms", n_slots,
                        n_regular);
  const int32_t final_idx = (int32_t)u[8];
  if (final_idx != -1 && (final_idx < 0 || final_idx >= n_slots))
    return noalloc_fail(P3_E_ARG, "final xform slot %d outside %d slots", final_idx, n_slots);
  if (u[15] > 3u) return noalloc_fail(P3_E_ARG, "colour mode %u is not 0..3", u[15]);
  if (u[14] != 0u)
    return noalloc_fail(P3_E_ARG, "the colour-index capture needs a renderer (p3_iterate)");
  const size_t bins = (size_t)u[0] * u[1];
  if (bins == 0) return noalloc_fail(P3_E_ARG, "empty grid %ux%u", u[0], u[1]);
  if (n_hist != bins * 4)
    return noalloc_fail(P3_E_ARG, "histogram holds %zu u32, the %ux%u grid needs %zu", n_hist,
                        u[0], u[1], bins * 4);
  const std::string rec = records_error(xforms, n_slots);
  if (!rec.empty()) return noalloc_fail(P3_E_ARG, "%s", rec.c_str());
  const size_t n_xf = (size_t)n_slots * P3_XFORM_FLOATS;
  const size_t table = (size_t)n_rows * P3_GRAIN;
  for (size_t i = 0; i < table; ++i)
    if (distrib[i] >= (uint32_t)n_regular)
      return noalloc_fail(P3_E_ARG, "pick table entry %zu = %u outside %d xforms", i, distrib[i],
                          n_regular);
  const size_t off_carry = align256(bins * 8);
  const size_t off_xf = off_carry + align256(bins * 32);
  const size_t off_ds = off_xf + align256(n_xf * sizeof(float));
  const size_t off_bd = off_ds + align256(table * sizeof(uint32_t));
  const size_t off_pal = off_bd + align256((size_t)p3::kPickRows * p3::kPickCols * 2);
  const size_t need = off_pal + 256 * 4 * sizeof(float);
  if (need > (size_t)P3_NOALLOC_SCRATCH_BYTES)
    return noalloc_fail(P3_E_ARG,
                        "a %ux%u grid with %d pick rows needs %zu bytes of scratch, above the "
                        "%zu-byte maximum (40 bytes per bin)",
                        u[0], u[1], n_rows, need, (size_t)P3_NOALLOC_SCRATCH_BYTES);
  if (u[12] == 0) return P3_OK;
  // the device, its scratch, and the shape's module loaded earlier
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return noalloc_fail(P3_E_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  if (!p3::chaos_module_loaded(dev, xforms, n_regular, final_idx))
    return noalloc_fail(P3_E_STATE,
                        "the chaos kernel of this genome shape is not loaded on device %d: run "
                        "one p3_iterate of the shape there first (loading code allocates)",
                        dev);
  unsigned char* base = scratch_base(dev);
  if (!base) return P3_E_HIP;
  hipStream_t stream = (hipStream_t)stream_;
  uint64_t* acc = (uint64_t*)base;
  uint64_t* carry = (uint64_t*)(base + off_carry);
  float* dxf = (float*)(base + off_xf);
  uint32_t* dds = (uint32_t*)(base + off_ds);
  uint16_t* dbd = (uint16_t*)(base + off_bd);
  float* dpal = (float*)(base + off_pal);
  std::vector<uint16_t> bounds;
  const bool have_bounds = pick_bounds(distrib, n_rows, bounds);
  p3::ChaosArgs a = dispatch_args_at(u, n_regular, final_idx, dxf, dds, dpal, hist);
  a.pick_bounds = have_bounds ? dbd : nullptr;
  a.acc = acc;
  a.carry = carry;
  // uploads (the bounds through this device's host staging, which outlives
  // the asynchronous copy), then the cleared accumulator
  if ((e = hipMemcpyAsync(dxf, xforms, n_xf * sizeof(float), hipMemcpyHostToDevice, stream)) ||
      (e = hipMemcpyAsync(dds, distrib, table * sizeof(uint32_t), hipMemcpyHostToDevice, stream)) ||
      (e = hipMemcpyAsync(dpal, palette_rgba, 256 * 4 * sizeof(float), hipMemcpyHostToDevice,
                          stream)))
    return noalloc_fail(P3_E_HIP, "upload: %s", hipGetErrorString(e));
  if (have_bounds) {
    std::vector<uint16_t>* stage;
    {
      std::lock_guard<std::mutex> lock(g_scratch_mu);
      stage = &g_scratch_bounds[(size_t)dev];
    }
    std::copy(bounds.begin(), bounds.end(), stage->begin());
    if ((e = hipMemcpyAsync(dbd, stage->data(), bounds.size() * sizeof(uint16_t),
                            hipMemcpyHostToDevice, stream)))
      return noalloc_fail(P3_E_HIP, "upload pick bounds: %s", hipGetErrorString(e));
  }
  if ((e = p3::launch_fill(acc, off_xf, 0u, stream)))
    return noalloc_fail(P3_E_HIP, "clear packed accumulator: %s", hipGetErrorString(e));
  if ((e = p3::launch_chaos_rtc_loaded(a, xforms, n_regular, final_idx, stream)))
    return noalloc_fail(P3_E_HIP, "chaos launch: %s", hipGetErrorString(e));
  if ((e = p3::launch_hist_merge(hist, acc, carry, bins, stream)))
    return noalloc_fail(P3_E_HIP, "packed merge: %s", hipGetErrorString(e));
  return P3_OK;
}

const char* p3_last_error(p3_renderer* r) {
  if (!r) return g_create_error.c_str();
  return r->err.c_str();
}

}  // extern "C"
