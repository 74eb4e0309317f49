// Batch object: many output frames (genomes of any shapes and sizes, each
// possibly several sub-dispatches such as motion-blur sub-frames) rendered
// by a few persistent chaos launches on one captured HIP graph.
//
// Frames are added on the host (p3_batch_add), then p3_batch_commit:
//   * lays every frame's buffers out in a few device arrays (histograms,
//     packed accumulators and carries, DE staging, filtered grids, images),
//     so the frames of a launch group cover one contiguous range of bins;
//   * splits the frames, in order, into launch groups of at most
//     `max_shapes` distinct genome shapes (frames with more than 16 xforms
//     form groups of their own: their picks read the pick table) and
//     compiles one union kernel per group (p3_rtc.cpp): every shape's xform
//     cases in one switch, selected per block by its sub-dispatch;
//   * uploads every record blob, pick table (only where the kernel reads it),
//     pick step boundaries, palette, filter taps and kernel argument block;
//   * captures the whole batch in a HIP graph over `lanes` streams: per
//     group, a histogram fill, ONE chaos launch over all the group's
//     sub-dispatches (each block finds its sub-dispatch by a binary search
//     over the group's block offsets), one packed-accumulator merge over the
//     group's bins, then per frame DE and tonemap; groups go round-robin to
//     the lanes, so one group's DE/tonemap and the next group's chaos run
//     side by side.
// p3_batch_launch replays the graph (asynchronous); the frames' images and
// histograms are read back per frame. Every replay renders the same frames
// from the same seeds (a histogram fill starts each group; the merge leaves
// the accumulators cleared), so replays are bit-identical.
//
// No counterpart in the reference: its batch loop renders one genome or
// sub-frame per dispatch (src/animate-render.ts:144-160).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/pyr3hip.h"
#include "p3_host.h"
#include "p3_launch.h"

namespace {

constexpr int kMaxLanes = 4;

template <typename T>
struct Dev {
  T* p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t count) {
    release();
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

struct Sub {
  std::vector<float> xf;         // n_slots x 116
  std::vector<uint32_t> table;   // pick table, kept only when the kernel may read it
  std::vector<uint16_t> bounds;  // step boundaries (empty: none)
  uint32_t u[P3_CHAOS_UNIFORM_WORDS];
  int n_regular = 0;
  int32_t final_idx = -1;
};

struct Frame {
  int w = 0, h = 0, os = 1, sw = 0, sh = 0, n_taps = 0;
  float taps[17] = {0};
  float palette[1024];
  std::vector<Sub> subs;
  bool de_on = false;
  uint32_t de[P3_DE_UNIFORM_WORDS];
  uint32_t viz[P3_VIZ_UNIFORM_WORDS];
  std::string shape;     // shape key of its subs
  size_t bin0 = 0;       // first bin in the batch's bin arrays
  size_t tile0 = 0;      // first DE tile
  size_t pix0 = 0;       // first output pixel
  size_t bins() const { return (size_t)sw * sh; }
  size_t tiles() const { return (size_t)((sw + 15) / 16) * ((sh + 15) / 16); }
  size_t pixels() const { return (size_t)w * h; }
};

struct Group {
  int f0 = 0, f1 = 0;                    // launch-order positions [f0, f1)
  std::vector<int> shape_of_frame;       // union shape index per position (pos - f0)
  std::vector<int> shape_frame;          // first frame (index) of each union shape
  uint32_t blocks = 0;
  size_t args0 = 0, start0 = 0;          // offsets in the argument / start arrays
  uint32_t n_sub = 0;
};

}  // namespace

struct p3_batch {
  std::vector<Frame> frames;             // in p3_batch_add order
  std::vector<int> order;                // launch order (frame indices)
  std::vector<Group> groups;
  hipStream_t lane[kMaxLanes] = {};
  int lanes = 0;
  hipEvent_t fork = nullptr, join[kMaxLanes] = {}, t0 = nullptr, t1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool committed = false, launched = false;
  Dev<uint32_t> hist, tiles, image;
  Dev<uint64_t> acc, carry;
  Dev<float4> pre, filtered;
  Dev<uint8_t> rad;
  Dev<float> xforms, palettes, taps, lut;
  Dev<uint32_t> tables, starts;
  Dev<uint16_t> bounds;
  Dev<p3::ChaosArgs> args;
  uint32_t total_blocks = 0, total_subs = 0, total_shapes = 0;
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
};

namespace {

void free_device(p3_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->hist.release(); b->tiles.release(); b->image.release(); b->acc.release();
  b->carry.release(); b->pre.release(); b->filtered.release(); b->rad.release();
  b->xforms.release(); b->palettes.release(); b->taps.release(); b->lut.release();
  b->tables.release(); b->starts.release(); b->bounds.release(); b->args.release();
}

// One sub-dispatch's blobs checked against its frame's grid (as p3_iterate
// checks them against a renderer's).
int check_sub(p3_batch* b, const Frame& f, const float* xf, int n_slots, const uint32_t* table,
              int n_rows, const uint32_t* u, bool scan_table) {
  if (!xf || !table || !u) return b->fail(P3_E_ARG, "null sub-dispatch blob");
  if (n_rows < 2 || n_rows > P3_MAX_XFORMS + 1)
    return b->fail(P3_E_ARG, "n_rows must be 2..%d (got %d)", P3_MAX_XFORMS + 1, n_rows);
  const int n_regular = n_rows - 1;
  if (n_slots < n_regular || n_slots > n_regular + 1)
    return b->fail(P3_E_ARG, "n_slots %d inconsistent with %d regular xforms", n_slots, n_regular);
  const int32_t fi = (int32_t)u[8];
  if (fi != -1 && (fi < 0 || fi >= n_slots))
    return b->fail(P3_E_ARG, "final xform slot %d outside %d slots", fi, n_slots);
  if (u[15] > 3u) return b->fail(P3_E_ARG, "colour mode %u is not 0..3", u[15]);
  if (u[14] != 0u) return b->fail(P3_E_ARG, "colour-index capture is not available in a batch");
  if ((int)u[0] != f.sw || (int)u[1] != f.sh)
    return b->fail(P3_E_ARG, "uniform size %ux%u does not match the frame's grid %dx%d", u[0],
                   u[1], f.sw, f.sh);
  const std::string rec = p3h::records_error(xf, n_slots);
  if (!rec.empty()) return b->fail(P3_E_ARG, "%s", rec.c_str());
  if (scan_table) {
    const size_t n = (size_t)n_rows * P3_GRAIN;
    for (size_t i = 0; i < n; ++i)
      if (table[i] >= (uint32_t)n_regular)
        return b->fail(P3_E_ARG, "pick table entry %zu = %u outside %d xforms", i, table[i],
                       n_regular);
  }
  return P3_OK;
}

template <typename T>
int upload(p3_batch* b, T* dst, const std::vector<T>& src, const char* what) {
  if (src.empty()) return P3_OK;
  return b->hip(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice), what);
}

// Launch order and groups: the frames with at most 16 regular xforms in the
// order they were added, then the others (their picks read the pick table,
// so they never share a union kernel with boundary-picked shapes); groups
// are consecutive runs of that order with at most max_shapes distinct
// shapes.
void make_groups(p3_batch* b, int max_shapes) {
  b->groups.clear();
  b->order.clear();
  const int nf = (int)b->frames.size();
  for (int wide = 0; wide < 2; ++wide)
    for (int f = 0; f < nf; ++f)
      if ((b->frames[f].subs[0].n_regular > 16) == (wide == 1)) b->order.push_back(f);
  Group g;
  std::map<std::string, int> idx;
  bool g_wide = false;
  for (int pos = 0; pos < nf; ++pos) {
    const Frame& fr = b->frames[b->order[pos]];
    const bool wide = fr.subs[0].n_regular > 16;
    const bool full = !idx.count(fr.shape) && (int)idx.size() >= max_shapes;
    if (!g.shape_of_frame.empty() && (full || wide != g_wide)) {
      b->groups.push_back(g);
      g = Group();
      idx.clear();
    }
    if (g.shape_of_frame.empty()) { g.f0 = pos; g_wide = wide; }
    auto it = idx.find(fr.shape);
    int sh;
    if (it == idx.end()) {
      sh = (int)idx.size();
      idx.emplace(fr.shape, sh);
      g.shape_frame.push_back(b->order[pos]);
    } else {
      sh = it->second;
    }
    g.shape_of_frame.push_back(sh);
    g.f1 = pos + 1;
  }
  if (!g.shape_of_frame.empty()) b->groups.push_back(g);
}

}  // namespace

extern "C" {

int p3_batch_create(p3_batch** out) {
  if (!out) return P3_E_ARG;
  *out = new p3_batch();
  return P3_OK;
}

void p3_batch_destroy(p3_batch* b) {
  if (!b) return;
  for (int i = 0; i < b->lanes; ++i)
    if (b->lane[i]) (void)hipStreamSynchronize(b->lane[i]);
  free_device(b);
  for (int i = 0; i < b->lanes; ++i) {
    if (b->join[i]) (void)hipEventDestroy(b->join[i]);
    if (b->lane[i]) (void)hipStreamDestroy(b->lane[i]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->t0) (void)hipEventDestroy(b->t0);
  if (b->t1) (void)hipEventDestroy(b->t1);
  delete b;
}

int p3_batch_add(p3_batch* b, int width, int height, int oversample, const float* taps,
                 int n_taps, const float* palette_rgba, int n_sub, const float* const* xforms,
                 int n_slots, const uint32_t* const* distribs, int n_rows,
                 const uint32_t* const* uniforms, const uint32_t* de_uniforms,
                 const uint32_t* viz_uniforms) {
  if (!b) return P3_E_ARG;
  if (b->committed) return b->fail(P3_E_STATE, "frames cannot be added after p3_batch_commit");
  if (width <= 0 || height <= 0 || oversample < 1 || oversample > 16)
    return b->fail(P3_E_ARG, "invalid size %dx%d oversample %d", width, height, oversample);
  if (!taps || n_taps < 3 || n_taps > 17)
    return b->fail(P3_E_ARG, "taps must hold 3..17 values (got %d)", n_taps);
  if (!palette_rgba || !viz_uniforms || !xforms || !distribs || !uniforms)
    return b->fail(P3_E_ARG, "null argument");
  if (n_sub < 1 || n_sub > 4096) return b->fail(P3_E_ARG, "n_sub must be 1..4096 (got %d)", n_sub);
  const long long sw = (long long)width * oversample, sh = (long long)height * oversample;
  if (sw * sh > (1LL << 31)) return b->fail(P3_E_ARG, "super-res grid %lldx%lld too large", sw, sh);
  Frame f;
  f.w = width; f.h = height; f.os = oversample; f.sw = (int)sw; f.sh = (int)sh;
  f.n_taps = n_taps;
  std::memcpy(f.taps, taps, sizeof(float) * n_taps);
  std::memcpy(f.palette, palette_rgba, sizeof f.palette);
  if ((int)viz_uniforms[0] != width || (int)viz_uniforms[1] != height ||
      (int)viz_uniforms[8] != oversample || (int)viz_uniforms[9] != n_taps)
    return b->fail(P3_E_ARG, "tonemap uniforms do not match the frame (size, oversample, taps)");
  if (viz_uniforms[16] != 0u)
    return b->fail(P3_E_ARG, "channel curves are not available in a batch");
  std::memcpy(f.viz, viz_uniforms, sizeof f.viz);
  f.de_on = de_uniforms != nullptr;
  if (f.de_on) {
    if ((int)de_uniforms[0] != f.sw || (int)de_uniforms[1] != f.sh)
      return b->fail(P3_E_ARG, "DE uniforms do not match the frame's grid");
    p3::DeArgs a{};
    const std::string bad = p3h::de_params(de_uniforms, a);
    if (!bad.empty()) return b->fail(P3_E_ARG, "%s", bad.c_str());
    std::memcpy(f.de, de_uniforms, sizeof f.de);
  }
  const size_t n_xf = (size_t)n_slots * P3_XFORM_FLOATS;
  const size_t n_tab = (size_t)(n_rows > 0 ? n_rows : 0) * P3_GRAIN;
  for (int k = 0; k < n_sub; ++k) {
    // a table equal to the previous sub-dispatch's (sub-frames of one
    // genome share their weights) was already scanned
    const bool same = k > 0 && n_rows > 0 &&
                      std::memcmp(distribs[k], distribs[k - 1], n_tab * sizeof(uint32_t)) == 0;
    int rc;
    if ((rc = check_sub(b, f, xforms[k], n_slots, distribs[k], n_rows, uniforms[k], !same)))
      return rc;
    Sub s;
    s.xf.assign(xforms[k], xforms[k] + n_xf);
    s.n_regular = n_rows - 1;
    s.final_idx = (int32_t)uniforms[k][8];
    std::memcpy(s.u, uniforms[k], sizeof s.u);
    const bool has_bounds = p3h::pick_bounds(distribs[k], n_rows, s.bounds);
    if (!has_bounds) s.bounds.clear();
    const bool blend = p3h::bits_f(s.u[30]) > 0.0f;
    if (!has_bounds || blend || s.n_regular > 16)
      s.table.assign(distribs[k], distribs[k] + n_tab);
    const std::string key = p3::chaos_shape_key(s.xf.data(), s.n_regular, s.final_idx);
    if (k == 0) f.shape = key;
    else if (key != f.shape)
      return b->fail(P3_E_ARG, "sub-dispatch %d has a different genome shape from sub-dispatch 0",
                     k);
    f.subs.push_back(std::move(s));
  }
  b->frames.push_back(std::move(f));
  return (int)b->frames.size() - 1;
}

int p3_batch_commit(p3_batch* b, int max_shapes, int lanes) {
  if (!b) return P3_E_ARG;
  if (b->committed) return b->fail(P3_E_STATE, "batch already committed");
  if (b->frames.empty()) return b->fail(P3_E_STATE, "no frames added");
  if (max_shapes < 1 || max_shapes > 256)
    return b->fail(P3_E_ARG, "max_shapes must be 1..256 (got %d)", max_shapes);
  if (lanes < 1 || lanes > kMaxLanes)
    return b->fail(P3_E_ARG, "lanes must be 1..%d (got %d)", kMaxLanes, lanes);
  make_groups(b, max_shapes);
  // layout in launch order: a group's frames cover one range of bins
  size_t bins = 0, tiles = 0, pix = 0, n_xf = 0, n_tab = 0, n_bnd = 0, n_sub = 0;
  for (int fi : b->order) {
    Frame& f = b->frames[fi];
    f.bin0 = bins; f.tile0 = tiles; f.pix0 = pix;
    bins += f.bins(); tiles += f.tiles(); pix += f.pixels();
    for (const Sub& s : f.subs) {
      n_xf += s.xf.size();
      n_tab += s.table.size();
      n_bnd += s.bounds.size();
      ++n_sub;
    }
  }
  int rc;
  if ((rc = b->hip(b->hist.alloc(bins * 4), "hipMalloc(histograms)")) ||
      (rc = b->hip(b->acc.alloc(bins), "hipMalloc(packed accumulators)")) ||
      (rc = b->hip(b->carry.alloc(bins * 4), "hipMalloc(packed carries)")) ||
      (rc = b->hip(b->pre.alloc(bins), "hipMalloc(DE staging)")) ||
      (rc = b->hip(b->filtered.alloc(bins), "hipMalloc(filtered grids)")) ||
      (rc = b->hip(b->rad.alloc(bins), "hipMalloc(DE radii)")) ||
      (rc = b->hip(b->tiles.alloc(tiles), "hipMalloc(DE tiles)")) ||
      (rc = b->hip(b->image.alloc(pix), "hipMalloc(images)")) ||
      (rc = b->hip(b->xforms.alloc(n_xf), "hipMalloc(records)")) ||
      (rc = b->hip(b->tables.alloc(n_tab), "hipMalloc(pick tables)")) ||
      (rc = b->hip(b->bounds.alloc(n_bnd), "hipMalloc(pick bounds)")) ||
      (rc = b->hip(b->palettes.alloc(b->frames.size() * 1024), "hipMalloc(palettes)")) ||
      (rc = b->hip(b->taps.alloc(b->frames.size() * 17), "hipMalloc(taps)")) ||
      (rc = b->hip(b->lut.alloc(257), "hipMalloc(kernel-norm table)")) ||
      (rc = b->hip(b->args.alloc(n_sub), "hipMalloc(chaos arguments)")) ||
      (rc = b->hip(b->starts.alloc(n_sub + b->groups.size()), "hipMalloc(block offsets)"))) {
    free_device(b);
    return rc;
  }
  // host images of the uploads
  std::vector<float> h_xf, h_pal, h_taps;
  std::vector<uint32_t> h_tab, h_start;
  std::vector<uint16_t> h_bnd;
  std::vector<p3::ChaosArgs> h_args;
  h_xf.reserve(n_xf); h_tab.reserve(n_tab); h_bnd.reserve(n_bnd);
  h_pal.reserve(b->frames.size() * 1024); h_taps.assign(b->frames.size() * 17, 0.0f);
  for (size_t fi = 0; fi < b->frames.size(); ++fi) {
    const Frame& f = b->frames[fi];
    h_pal.insert(h_pal.end(), f.palette, f.palette + 1024);
    std::memcpy(&h_taps[fi * 17], f.taps, sizeof(float) * f.n_taps);
  }
  b->total_shapes = 0;
  const uint32_t blk = (uint32_t)p3::chaos_block();
  for (Group& g : b->groups) {
    g.args0 = h_args.size();
    g.start0 = h_start.size();
    g.blocks = 0;
    b->total_shapes += (uint32_t)g.shape_frame.size();
    std::vector<int> case_base(g.shape_frame.size(), 0);
    for (size_t s = 1; s < g.shape_frame.size(); ++s)
      case_base[s] = case_base[s - 1] + b->frames[g.shape_frame[s - 1]].subs[0].n_regular;
    for (int pos = g.f0; pos < g.f1; ++pos) {
      const int fi = b->order[pos];
      const Frame& f = b->frames[fi];
      const int shape = g.shape_of_frame[pos - g.f0];
      for (const Sub& s : f.subs) {
        const float* dxf = b->xforms.p + h_xf.size();
        const uint32_t* dtab = s.table.empty() ? nullptr : b->tables.p + h_tab.size();
        p3::ChaosArgs a = p3h::dispatch_args_at(s.u, s.n_regular, s.final_idx, dxf, dtab,
                                                b->palettes.p + fi * 1024,
                                                b->hist.p + f.bin0 * 4);
        a.pick_bounds = s.bounds.empty() ? nullptr : b->bounds.p + h_bnd.size();
        a.acc = b->acc.p + f.bin0;
        a.carry = b->carry.p + f.bin0 * 4;
        a.shape = (uint32_t)shape;
        a.case_base = (uint32_t)case_base[shape];
        h_xf.insert(h_xf.end(), s.xf.begin(), s.xf.end());
        h_tab.insert(h_tab.end(), s.table.begin(), s.table.end());
        h_bnd.insert(h_bnd.end(), s.bounds.begin(), s.bounds.end());
        h_start.push_back(g.blocks);
        g.blocks += (a.walkers + blk - 1) / blk;
        h_args.push_back(a);
      }
    }
    h_start.push_back(g.blocks);
    g.n_sub = (uint32_t)(h_args.size() - g.args0);
  }
  b->total_subs = (uint32_t)h_args.size();
  b->total_blocks = 0;
  for (const Group& g : b->groups) b->total_blocks += g.blocks;
  if ((rc = upload(b, b->xforms.p, h_xf, "upload records")) ||
      (rc = upload(b, b->tables.p, h_tab, "upload pick tables")) ||
      (rc = upload(b, b->bounds.p, h_bnd, "upload pick bounds")) ||
      (rc = upload(b, b->palettes.p, h_pal, "upload palettes")) ||
      (rc = upload(b, b->taps.p, h_taps, "upload taps")) ||
      (rc = upload(b, b->args.p, h_args, "upload chaos arguments")) ||
      (rc = upload(b, b->starts.p, h_start, "upload block offsets")) ||
      (rc = b->hip(hipMemcpy(b->lut.p, p3h::kernel_norm_lut(), sizeof(float) * 257,
                             hipMemcpyHostToDevice),
                   "upload kernel-norm table"))) {
    free_device(b);
    return rc;
  }
  // union kernels: compiled (or fetched from the caches) and loaded before
  // the capture, so the capture records launches only
  std::vector<std::vector<p3::ShapeRef>> unions(b->groups.size());
  for (size_t gi = 0; gi < b->groups.size(); ++gi) {
    for (int f : b->groups[gi].shape_frame) {
      const Sub& s = b->frames[f].subs[0];
      unions[gi].push_back({s.xf.data(), s.n_regular, s.final_idx});
    }
    std::string rtc_err;
    const std::vector<char>* code = nullptr;
    if (!p3::union_code(unions[gi], &code, rtc_err)) {
      free_device(b);
      return b->fail(P3_E_HIP, "union kernel of group %zu: %s", gi, rtc_err.c_str());
    }
  }
  // streams and events (once per batch object)
  if (b->lanes == 0) {
    for (int i = 0; i < kMaxLanes; ++i) {
      if ((rc = b->hip(hipStreamCreateWithFlags(&b->lane[i], hipStreamNonBlocking),
                       "hipStreamCreate")) ||
          (rc = b->hip(hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming),
                       "hipEventCreate"))) {
        free_device(b);
        return rc;
      }
      b->lanes = i + 1;
    }
    if ((rc = b->hip(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate")) ||
        (rc = b->hip(hipEventCreate(&b->t0), "hipEventCreate")) ||
        (rc = b->hip(hipEventCreate(&b->t1), "hipEventCreate"))) {
      free_device(b);
      return rc;
    }
  }
  // accumulators start cleared (each group's merge clears its range again)
  hipStream_t s0 = b->lane[0];
  if ((rc = b->hip(p3::launch_fill(b->acc.p, bins * 8, 0u, s0), "clear accumulators")) ||
      (rc = b->hip(p3::launch_fill(b->carry.p, bins * 32, 0u, s0), "clear carries")) ||
      (rc = b->hip(hipStreamSynchronize(s0), "sync after clears"))) {
    free_device(b);
    return rc;
  }
  // Modules are loaded on first launch: a throw-away launch of zero blocks
  // per union loads them outside the capture.
  for (size_t gi = 0; gi < b->groups.size(); ++gi) {
    std::string rtc_err;
    const hipError_t e = p3::launch_union_batch(p3::BatchArgs{}, 0u, unions[gi], s0, rtc_err);
    if (e != hipSuccess) {
      free_device(b);
      return b->fail(P3_E_HIP, "loading union kernel %zu: %s %s", gi, hipGetErrorString(e),
                     rtc_err.c_str());
    }
  }
  // capture
  if ((rc = b->hip(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal), "begin capture"))) {
    free_device(b);
    return rc;
  }
  auto abort_capture = [&](int code) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s0, &g);
    if (g) (void)hipGraphDestroy(g);
    free_device(b);
    return code;
  };
  if ((rc = b->hip(hipEventRecord(b->fork, s0), "fork"))) return abort_capture(rc);
  for (int i = 1; i < lanes; ++i)
    if ((rc = b->hip(hipStreamWaitEvent(b->lane[i], b->fork, 0), "fork wait")))
      return abort_capture(rc);
  for (size_t gi = 0; gi < b->groups.size(); ++gi) {
    const Group& g = b->groups[gi];
    hipStream_t st = b->lane[gi % lanes];
    const Frame& first = b->frames[b->order[g.f0]];
    const Frame& last = b->frames[b->order[g.f1 - 1]];
    const size_t g_bins = last.bin0 + last.bins() - first.bin0;
    if ((rc = b->hip(p3::launch_fill(b->hist.p + first.bin0 * 4, g_bins * 16, 0u, st),
                     "clear histograms")))
      return abort_capture(rc);
    p3::BatchArgs ba{b->args.p + g.args0, b->starts.p + g.start0, g.n_sub};
    std::string rtc_err;
    const hipError_t e = p3::launch_union_batch(ba, g.blocks, unions[gi], st, rtc_err);
    if (e != hipSuccess)
      return abort_capture(b->fail(P3_E_HIP, "union chaos launch %zu: %s %s", gi,
                                   hipGetErrorString(e), rtc_err.c_str()));
    if ((rc = b->hip(p3::launch_hist_merge(b->hist.p + first.bin0 * 4, b->acc.p + first.bin0,
                                           b->carry.p + first.bin0 * 4, g_bins, st),
                     "packed merge")))
      return abort_capture(rc);
    for (int pos = g.f0; pos < g.f1; ++pos) {
      const int fi = b->order[pos];
      const Frame& f = b->frames[fi];
      const uint4* fh = (const uint4*)(b->hist.p + f.bin0 * 4);
      if (f.de_on) {
        p3::DeArgs a{};
        (void)p3h::de_params(f.de, a);
        a.hist = fh;
        a.pre = b->pre.p + f.bin0;
        a.rad = b->rad.p + f.bin0;
        a.tile_rmax = b->tiles.p + f.tile0;
        a.out = b->filtered.p + f.bin0;
        a.lut = b->lut.p;
        a.width = (uint32_t)f.sw; a.height = (uint32_t)f.sh;
        a.tiles_x = (uint32_t)((f.sw + 15) / 16);
        a.tiles_y = (uint32_t)((f.sh + 15) / 16);
        if ((rc = b->hip(p3::launch_density(a, st), "density launch"))) return abort_capture(rc);
      }
      p3::TmArgs t{};
      t.filtered = b->filtered.p + f.bin0;
      t.hist = fh;
      t.taps = b->taps.p + fi * 17;
      t.out8 = b->image.p + f.pix0;
      t.outf = nullptr;
      t.width = (uint32_t)f.w; t.height = (uint32_t)f.h; t.os = (uint32_t)f.os;
      p3h::tm_params(f.viz, t);
      t.curves = nullptr;
      if ((rc = b->hip(p3::launch_tonemap(t, f.de_on, st), "tonemap launch")))
        return abort_capture(rc);
    }
  }
  for (int i = 1; i < lanes; ++i) {
    if ((rc = b->hip(hipEventRecord(b->join[i], b->lane[i]), "join")) ||
        (rc = b->hip(hipStreamWaitEvent(s0, b->join[i], 0), "join wait")))
      return abort_capture(rc);
  }
  if ((rc = b->hip(hipStreamEndCapture(s0, &b->graph), "end capture"))) {
    free_device(b);
    return rc;
  }
  if ((rc = b->hip(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0),
                   "graph instantiate"))) {
    free_device(b);
    return rc;
  }
  b->committed = true;
  return P3_OK;
}

int p3_batch_launch(p3_batch* b) {
  if (!b) return P3_E_ARG;
  if (!b->committed) return b->fail(P3_E_STATE, "launch before p3_batch_commit");
  int rc;
  if ((rc = b->hip(hipEventRecord(b->t0, b->lane[0]), "event")) ||
      (rc = b->hip(hipGraphLaunch(b->exec, b->lane[0]), "graph launch")) ||
      (rc = b->hip(hipEventRecord(b->t1, b->lane[0]), "event")))
    return rc;
  b->launched = true;
  return P3_OK;
}

int p3_batch_sync(p3_batch* b) {
  if (!b) return P3_E_ARG;
  if (!b->committed) return P3_OK;
  return b->hip(hipStreamSynchronize(b->lane[0]), "batch sync");
}

int p3_batch_elapsed(p3_batch* b, float* ms) {
  if (!b || !ms) return P3_E_ARG;
  if (!b->launched) return b->fail(P3_E_STATE, "no launch yet");
  int rc;
  if ((rc = b->hip(hipEventSynchronize(b->t1), "event sync"))) return rc;
  return b->hip(hipEventElapsedTime(ms, b->t0, b->t1), "elapsed time");
}

int p3_batch_frames(p3_batch* b, int* out) {
  if (!b || !out) return P3_E_ARG;
  *out = (int)b->frames.size();
  return P3_OK;
}

int p3_batch_stats(p3_batch* b, uint32_t* out4) {
  if (!b || !out4) return P3_E_ARG;
  if (!b->committed) return b->fail(P3_E_STATE, "stats before p3_batch_commit");
  out4[0] = (uint32_t)b->groups.size();
  out4[1] = b->total_shapes;
  out4[2] = b->total_subs;
  out4[3] = b->total_blocks;
  return P3_OK;
}

int p3_batch_read_image(p3_batch* b, int frame, uint8_t* out, size_t n) {
  if (!b || !out) return P3_E_ARG;
  if (!b->launched) return b->fail(P3_E_STATE, "read before p3_batch_launch");
  if (frame < 0 || frame >= (int)b->frames.size())
    return b->fail(P3_E_ARG, "frame %d outside %zu frames", frame, b->frames.size());
  const Frame& f = b->frames[frame];
  if (n != f.pixels() * 4)
    return b->fail(P3_E_ARG, "image needs %zu bytes (got %zu)", f.pixels() * 4, n);
  int rc;
  if ((rc = p3_batch_sync(b))) return rc;
  return b->hip(hipMemcpy(out, b->image.p + f.pix0, n, hipMemcpyDeviceToHost), "read image");
}

int p3_batch_read_histogram(p3_batch* b, int frame, uint32_t* out, size_t n) {
  if (!b || !out) return P3_E_ARG;
  if (!b->launched) return b->fail(P3_E_STATE, "read before p3_batch_launch");
  if (frame < 0 || frame >= (int)b->frames.size())
    return b->fail(P3_E_ARG, "frame %d outside %zu frames", frame, b->frames.size());
  const Frame& f = b->frames[frame];
  if (n != f.bins() * 4)
    return b->fail(P3_E_ARG, "histogram needs %zu words (got %zu)", f.bins() * 4, n);
  int rc;
  if ((rc = p3_batch_sync(b))) return rc;
  return b->hip(hipMemcpy(out, b->hist.p + f.bin0 * 4, n * sizeof(uint32_t),
                          hipMemcpyDeviceToHost),
                "read histogram");
}

int p3_union_prepare(int n, const float* const* xforms, const int* n_slots, const int* n_rows,
                     const int* final_idx) {
  if (n < 1 || !xforms || !n_slots || !n_rows || !final_idx) return P3_E_ARG;
  std::vector<p3::ShapeRef> shapes;
  for (int i = 0; i < n; ++i) {
    const int nr = n_rows[i], ns = n_slots[i], fi = final_idx[i];
    if (!xforms[i] || nr < 2 || nr > P3_MAX_XFORMS + 1) return P3_E_ARG;
    if (ns < nr - 1 || ns > nr) return P3_E_ARG;
    if (fi != -1 && (fi < 0 || fi >= ns)) return P3_E_ARG;
    if (!p3h::records_error(xforms[i], ns).empty()) return P3_E_ARG;
    shapes.push_back({xforms[i], nr - 1, fi});
  }
  std::string err;
  const std::vector<char>* code = nullptr;
  return p3::union_code(shapes, &code, err) ? P3_OK : P3_E_HIP;
}

const char* p3_batch_last_error(p3_batch* b) {
  return b ? b->err.c_str() : "null batch";
}

}  // extern "C"
