// Run-time specialisation of the chaos kernel with hipRTC.
//
// A genome's "shape" is, per xform record, the variation catalog index of
// every slot and whether a post-affine is present, plus the same for the
// final xform (or its absence). For each new shape a small source is
// generated: one `case` per xform with straight-line pre_blur / variation /
// post-affine code (the catalog switch folds away because the index is a
// constant), reading weights and parameters from the record at fixed offsets.
// Weights and parameters stay in the record, so every genome of the same
// shape (animation frames, quality sweeps) reuses the compiled module. It is
// compiled for gfx950 with the same flags as the rest of the library, loaded
// with hipModuleLoadData and cached for the life of the process.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "p3_launch.h"
#include "p3_var_draws.h"

namespace {

#include "p3_rtc_source.inc"  // defines kRtcSource[] (generated by the build)

struct Module {
  hipModule_t mod = nullptr;
  hipFunction_t single = nullptr, multi = nullptr;   // one-shape kernels
  hipFunction_t batch = nullptr;                     // union kernels
};

std::mutex g_mu;
std::unordered_map<std::string, std::vector<char>> g_code;  // key: shape
std::unordered_map<std::string, Module> g_mods;             // key: device + shape

uint32_t slot_index(const float* xf, int k) { return (uint32_t)xf[20 + 4 * k]; }
int slot_count(const float* xf) {
  const float n = xf[7];
  return n < 8.0f ? (int)n : 8;
}

void shape_of(const float* xf, std::ostringstream& key) {
  const int nv = slot_count(xf);
  key << nv << (xf[15] != 0.0f ? "p" : "n");
  for (int k = 0; k < nv; ++k) key << "," << slot_index(xf, k);
  key << ";";
}

// Straight-line chain of one record at X + slot * 116.
// True when the final xform exists and none of its slots draws from the RNG.
bool final_draw_free(const float* xforms, int final_slot) {
  if (final_slot < 0) return false;
  const float* xf = xforms + (size_t)final_slot * p3::kXformFloats;
  for (int k = 0; k < slot_count(xf); ++k)
    if (p3::var_draws(slot_index(xf, k))) return false;
  return true;
}

void emit_body(std::ostringstream& s, const float* xf, int slot, bool full) {
  const int nv = slot_count(xf);
  s << "      const xrec_t xf = X + " << slot * p3::kXformFloats << ";\n";
  if (full) {
    s << "      c.pa = make_float2(xf[0] * x + xf[1] * y + xf[2], xf[4] * x + xf[5] * y + xf[6]);\n"
      << "      float2 pam = c.pa;\n";
  } else {
    s << "      float2 pam = pa;\n";
  }
  for (int k = 0; k < nv; ++k)
    if (slot_index(xf, k) == 97u) s << "      pam = pre_blur_slot(xf, " << k << ", pam, rng);\n";
  s << "      float2 pv = make_float2(0.0f, 0.0f);\n";
  for (int k = 0; k < nv; ++k) {
    const uint32_t idx = slot_index(xf, k);
    if (idx != 97u)
      s << "      { const float2 o = apply_variation(" << idx << ", pam, xf, " << k << ", xf["
        << 20 + 4 * k + 1 << "], rng); pv.x = pv.x + o.x; pv.y = pv.y + o.y; }\n";
    // Direct-colour entries override the splat colour when the record's
    // dc flag is set and the slot weight is positive (last one wins).
    if (p3::is_dc_index(idx))
      s << "      if (xf[11] > 0.5f && xf[" << 20 + 4 * k + 1 << "] > 0.0f) { float3 dcc; if (dc_color_for("
        << idx << ", pam, xf, " << k << ", dcc)) { dc_on = true; dc_rgb = dcc; } }\n";
  }
  const bool post = xf[15] != 0.0f;
  if (full) {
    s << "      c.pv_pre = pv;\n"
      << "      c.pv = " << (post ? "post_affine(xf, pv)" : "pv") << ";\n"
      << "      c.z = mixf(z, xf[8], xf[9]);\n";
  } else {
    s << "      return " << (post ? "post_affine(xf, pv)" : "pv") << ";\n";
  }
}

// The xform cases of one shape, numbered from `base`.
void emit_cases(std::ostringstream& s, const float* xforms, int n_regular, int base) {
  for (int j = 0; j < n_regular; ++j) {
    s << "    case " << base + j << ": {\n";
    emit_body(s, xforms + (size_t)j * p3::kXformFloats, j, true);
    s << "    } break;\n";
  }
}

const char* const kChainHead =
    "  static P3_INLINE Chain chain(const ChaosArgs& A, uint32_t fn, xrec_t X, float x,\n"
    "      float y, float z, Rng& rng, bool& dc_on, float3& dc_rgb) {\n"
    "    (void)A; (void)dc_on; (void)dc_rgb;\n"
    "    Chain c;\n";
const char* const kFinalHead =
    "  static P3_INLINE float2 final_sum(const ChaosArgs& A, xrec_t X, int, float2 pa,\n"
    "      Rng& rng, bool& dc_on, float3& dc_rgb) {\n"
    "    (void)A; (void)dc_on; (void)dc_rgb;\n";

std::string generate(const float* xforms, int n_regular, int final_slot) {
  std::ostringstream s;
  s << "namespace p3 {\nstruct Spec {\n"
    << "  static constexpr int kRegular = " << n_regular << ";\n"
    << "  static P3_INLINE bool has_final(const ChaosArgs&) { return "
    << (final_slot >= 0 ? "true" : "false") << "; }\n"
    << "  static P3_INLINE int final_slot(const ChaosArgs&) { return "
    << (final_slot >= 0 ? final_slot : 0) << "; }\n"
    << "  static P3_INLINE bool final_draw_free(const ChaosArgs&) { return "
    << (final_draw_free(xforms, final_slot) ? "true" : "false") << "; }\n"
    << kChainHead
    << "    switch (fn) {\n";
  emit_cases(s, xforms, n_regular, 0);
  s << "    default: __builtin_unreachable();\n"
    << "    }\n"
    << "    return c;\n"
    << "  }\n"
    << kFinalHead;
  if (final_slot >= 0) {
    s << "    {\n";
    emit_body(s, xforms + (size_t)final_slot * p3::kXformFloats, final_slot, false);
    s << "    }\n";
  } else {
    s << "    (void)X; (void)rng; return pa;\n";
  }
  s << "  }\n};\n}  // namespace p3\n\n"
    << "constexpr int kBlk = " << p3::chaos_block() << ";\n"
    << "extern \"C\" __global__ void __launch_bounds__(kBlk)\n"
    << "p3_chaos(p3::ChaosArgs A) {\n"
    << "  __shared__ p3::BlockLds<kBlk> L;\n"
    << "  p3::chaos_binned<kBlk, p3::Spec>(A, L, blockIdx.x);\n"
    << "}\n"
    // Batched launch: sub-dispatch k (same genome shape, its own records,
    // pick table, seed, camera and walker count) owns blocks
    // [start[k], start[k+1]) of one grid.
    << "extern \"C\" __global__ void __launch_bounds__(kBlk)\n"
    << "p3_chaos_multi(p3::MultiArgs M) {\n"
    << "  __shared__ p3::BlockLds<kBlk> L;\n"
    << "  const uint32_t k = p3::multi_find(M, blockIdx.x);\n"
    << "  const p3::ChaosArgs A = ((const P3_CONST p3::ChaosArgs*)M.args)[k];\n"
    << "  p3::chaos_binned<kBlk, p3::Spec>(A, L, blockIdx.x - M.start[k]);\n"
    << "}\n";
  return s.str();
}

// Union kernel of several shapes: shape u's xform j is case
// case_base(u) + j of one switch (the dispatch's ChaosArgs carry its shape
// index and case base), its final xform case u of a second switch. A block
// belongs to one sub-dispatch, so both switches are block-uniform. The pick
// path is the step-boundary one when every shape has at most 16 xforms
// (kRegular = the largest count: a table of fewer xforms has its unused
// boundaries at 16384, which no column reaches).
std::string generate_union(const std::vector<p3::ShapeRef>& shapes) {
  std::ostringstream s;
  int n_max = 1;
  for (const auto& sh : shapes) n_max = std::max(n_max, sh.n_regular);
  s << "namespace p3 {\nstruct Spec {\n"
    << "  static constexpr int kRegular = " << n_max << ";\n"
    << "  static P3_INLINE bool has_final(const ChaosArgs& A) { return A.final_idx >= 0; }\n"
    << "  static P3_INLINE int final_slot(const ChaosArgs& A) { return A.final_idx; }\n"
    << "  static P3_INLINE bool final_draw_free(const ChaosArgs& A) {\n"
    << "    switch (A.shape) {\n";
  for (size_t u = 0; u < shapes.size(); ++u)
    if (final_draw_free(shapes[u].xforms, shapes[u].final_slot))
      s << "    case " << u << ":\n";
  s << "      return true;\n"
    << "    default: return false;\n"
    << "    }\n"
    << "  }\n"
    << kChainHead
    << "    switch (A.case_base + fn) {\n";
  int base = 0;
  for (const auto& sh : shapes) {
    emit_cases(s, sh.xforms, sh.n_regular, base);
    base += sh.n_regular;
  }
  s << "    default: __builtin_unreachable();\n"
    << "    }\n"
    << "    return c;\n"
    << "  }\n"
    << kFinalHead
    << "    switch (A.shape) {\n";
  for (size_t u = 0; u < shapes.size(); ++u) {
    if (shapes[u].final_slot < 0) continue;
    s << "    case " << u << ": {\n";
    emit_body(s, shapes[u].xforms + (size_t)shapes[u].final_slot * p3::kXformFloats,
              shapes[u].final_slot, false);
    s << "    }\n";
  }
  s << "    default: (void)X; (void)rng; return pa;\n"
    << "    }\n"
    << "  }\n};\n}  // namespace p3\n\n"
    << "constexpr int kBlk = " << p3::chaos_block() << ";\n"
    // Sub-dispatch k (any shape of the union, its own records, pick table,
    // seed, camera, walker count and histogram) owns blocks
    // [start[k], start[k+1]) of one grid.
    << "extern \"C\" __global__ void __launch_bounds__(kBlk)\n"
    << "p3_chaos_batch(p3::BatchArgs M) {\n"
    << "  __shared__ p3::BlockLds<kBlk> L;\n"
    << "  const uint32_t k = p3::batch_find(M, blockIdx.x);\n"
    << "  const p3::ChaosArgs A = ((const P3_CONST p3::ChaosArgs*)M.args)[k];\n"
    << "  p3::chaos_binned<kBlk, p3::Spec>(A, L,\n"
    << "      blockIdx.x - ((const P3_CONST uint32_t*)M.start)[k]);\n"
    << "}\n";
  return s.str();
}

// Persistent code-object cache: when P3_KERNEL_CACHE names a directory,
// compiled shapes are stored there as <dir>/p3chaos-<fnv64 of source and
// options>.co and reused by later processes (and hosts of the same arch).
uint64_t fnv1a(const char* p, size_t n, uint64_t h) {
  for (size_t i = 0; i < n; ++i) { h ^= (unsigned char)p[i]; h *= 1099511628211ull; }
  return h;
}

std::string cache_path(const std::string& src, const char* const* opts, int n_opts) {
  const char* dir = std::getenv("P3_KERNEL_CACHE");
  if (!dir || !*dir) return std::string();
  uint64_t h = fnv1a(src.data(), src.size(), 1469598103934665603ull);
  for (int i = 0; i < n_opts; ++i) h = fnv1a(opts[i], std::strlen(opts[i]) + 1, h);
  char name[48];
  std::snprintf(name, sizeof name, "/p3chaos-%016llx.co", (unsigned long long)h);
  return std::string(dir) + name;
}

bool read_cached(const std::string& path, std::vector<char>& code) {
  if (path.empty()) return false;
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  bool ok = n > 0;
  if (ok) {
    code.resize((size_t)n);
    ok = std::fread(code.data(), 1, (size_t)n, f) == (size_t)n;
  }
  std::fclose(f);
  return ok;
}

void write_cached(const std::string& path, const std::vector<char>& code) {
  if (path.empty()) return;
  const std::string dir = path.substr(0, path.rfind('/'));
  ::mkdir(dir.c_str(), 0755);  // one level; an existing directory is fine
  const std::string tmp = path + ".tmp." + std::to_string((long)::getpid());
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return;
  const bool ok = std::fwrite(code.data(), 1, code.size(), f) == code.size();
  std::fclose(f);
  if (ok) std::rename(tmp.c_str(), path.c_str());
  else std::remove(tmp.c_str());
}

const char* const kOpts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
constexpr int kNumOpts = (int)(sizeof kOpts / sizeof kOpts[0]);

bool compile(const std::string& src, std::vector<char>& code, std::string& err) {
  const std::string cached = cache_path(src, kOpts, kNumOpts);
  if (read_cached(cached, code)) return true;
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "p3_chaos_spec.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS) {
    err = "hiprtcCreateProgram failed";
    return false;
  }
  const hiprtcResult rc = hiprtcCompileProgram(prog, kNumOpts, const_cast<const char**>(kOpts));
  if (rc != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    err = std::string("hiprtc compile failed: ") + hiprtcGetErrorString(rc) + "\n" + log;
    hiprtcDestroyProgram(&prog);
    return false;
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  code.resize(n);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  write_cached(cached, code);
  return true;
}

}  // namespace

namespace p3 {

// Walkers per chaos block: 640, or P3_CHAOS_BLOCK (a multiple of 64 up to
// 640), read once per process.
int chaos_block() {
  static const int b = [] {
    const char* e = std::getenv("P3_CHAOS_BLOCK");
    const int v = e ? std::atoi(e) : kChaosBlock;
    // whole waves; 640 walkers x 236 B of block state fit one CU's 160 KiB LDS
    return (v >= 64 && v <= 640 && v % 64 == 0) ? v : kChaosBlock;
  }();
  return b;
}

// When P3_SHAPE_LOG names a file, the key of every shape this process
// compiles or loads is appended to it (one line each), so a run's shapes can
// be compiled ahead of the next run (pyr3_amd/kcache.py). Called under g_mu.
void log_shape(const std::string& key) {
  const char* path = std::getenv("P3_SHAPE_LOG");
  if (!path || !*path) return;
  FILE* f = std::fopen(path, "a");
  if (!f) return;
  std::fprintf(f, "%s\n", key.c_str());
  std::fclose(f);
}

std::string chaos_shape_key(const float* xforms, int n_regular, int final_slot) {
  std::ostringstream key;
  for (int j = 0; j < n_regular; ++j) shape_of(xforms + (size_t)j * kXformFloats, key);
  key << "F";
  if (final_slot >= 0) {
    key << final_slot << ":";
    shape_of(xforms + (size_t)final_slot * kXformFloats, key);
  }
  return key.str();
}

std::string chaos_source(const float* xforms, int n_regular, int final_slot) {
  return std::string(kRtcSource) + generate(xforms, n_regular, final_slot);
}

// Code object for this shape, compiled on first use (no device needed). The
// compile runs outside the lock, so launches of other shapes are not held up;
// entries of g_code are never erased, so the returned pointer stays valid.
bool chaos_code(const float* host_xforms, int n_regular, int final_slot,
                const std::vector<char>** out, std::string& err) {
  const std::string key = chaos_shape_key(host_xforms, n_regular, final_slot);
  {
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_code.find(key);
    if (it != g_code.end()) {
      *out = &it->second;
      return true;
    }
  }
  std::vector<char> code;
  if (!compile(chaos_source(host_xforms, n_regular, final_slot), code, err)) return false;
  std::lock_guard<std::mutex> lock(g_mu);
  log_shape(key);
  *out = &g_code.emplace(key, std::move(code)).first->second;  // first insert wins
  return true;
}

bool chaos_code_ready(const float* host_xforms, int n_regular, int final_slot) {
  const std::string key = chaos_shape_key(host_xforms, n_regular, final_slot);
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_code.count(key)) return true;
  }
  std::vector<char> code;
  if (!read_cached(cache_path(chaos_source(host_xforms, n_regular, final_slot), kOpts, kNumOpts),
                   code))
    return false;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_code.emplace(key, std::move(code)).second) log_shape(key);
  return true;
}

// Loaded module of this shape on device `dev` (loaded once per device).
hipError_t module_for(int dev, const float* host_xforms, int n_regular, int final_slot,
                      const std::vector<char>& code, Module& m) {
  const std::string key =
      std::to_string(dev) + "|" + chaos_shape_key(host_xforms, n_regular, final_slot);
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_mods.find(key);
  if (it != g_mods.end()) {
    m = it->second;
    return hipSuccess;
  }
  hipError_t e;
  if ((e = hipModuleLoadData(&m.mod, code.data())) != hipSuccess) return e;
  if ((e = hipModuleGetFunction(&m.single, m.mod, "p3_chaos")) != hipSuccess) return e;
  if ((e = hipModuleGetFunction(&m.multi, m.mod, "p3_chaos_multi")) != hipSuccess) return e;
  g_mods.emplace(key, m);
  return hipSuccess;
}

// Compiles (or fetches) the kernel for this shape and launches it on `stream`.
hipError_t launch_chaos_rtc(const ChaosArgs& a, const float* host_xforms, int n_regular,
                            int final_slot, hipStream_t stream, std::string& err) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const std::vector<char>* code = nullptr;
  if (!chaos_code(host_xforms, n_regular, final_slot, &code, err)) return hipErrorInvalidImage;
  Module m;
  if ((e = module_for(dev, host_xforms, n_regular, final_slot, *code, m)) != hipSuccess) return e;
  const unsigned blk = (unsigned)chaos_block();
  const unsigned grid = (a.walkers + blk - 1) / blk;
  if (grid == 0) return hipSuccess;
  ChaosArgs args = a;
  void* params[] = {&args};
  return hipModuleLaunchKernel(m.single, grid, 1, 1, blk, 1, 1,
                               0, stream, params, nullptr);
}

// True when this shape's module is loaded on device `dev`.
bool chaos_module_loaded(int dev, const float* host_xforms, int n_regular, int final_slot) {
  const std::string key =
      std::to_string(dev) + "|" + chaos_shape_key(host_xforms, n_regular, final_slot);
  std::lock_guard<std::mutex> lock(g_mu);
  return g_mods.count(key) != 0;
}

// Launch of the kernel of this shape only when its module is already loaded
// on the current device (neither compiles nor loads: loading a code object
// allocates device memory); hipErrorNotReady when it is not.
hipError_t launch_chaos_rtc_loaded(const ChaosArgs& a, const float* host_xforms, int n_regular,
                                   int final_slot, hipStream_t stream) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const std::string key =
      std::to_string(dev) + "|" + chaos_shape_key(host_xforms, n_regular, final_slot);
  Module m;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_mods.find(key);
    if (it == g_mods.end()) return hipErrorNotReady;
    m = it->second;
  }
  const unsigned blk = (unsigned)chaos_block();
  const unsigned grid = (a.walkers + blk - 1) / blk;
  if (grid == 0) return hipSuccess;
  ChaosArgs args = a;
  void* params[] = {&args};
  return hipModuleLaunchKernel(m.single, grid, 1, 1, blk, 1, 1, 0, stream, params, nullptr);
}

// Batched launch of n sub-dispatches sharing one genome shape: `dev_args`
// holds their ChaosArgs in device memory, start[] their first blocks.
hipError_t launch_chaos_rtc_multi(const MultiArgs& margs, const float* host_xforms, int n_regular,
                                  int final_slot, hipStream_t stream, std::string& err) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const std::vector<char>* code = nullptr;
  if (!chaos_code(host_xforms, n_regular, final_slot, &code, err)) return hipErrorInvalidImage;
  Module m;
  if ((e = module_for(dev, host_xforms, n_regular, final_slot, *code, m)) != hipSuccess) return e;
  const unsigned grid = margs.start[margs.n];
  if (grid == 0) return hipSuccess;
  MultiArgs args = margs;
  void* params[] = {&args};
  return hipModuleLaunchKernel(m.multi, grid, 1, 1, chaos_block(), 1, 1, 0,
                               stream, params, nullptr);
}

std::string union_key(const std::vector<ShapeRef>& shapes) {
  std::string key = "U";
  for (const auto& sh : shapes)
    key += "|" + chaos_shape_key(sh.xforms, sh.n_regular, sh.final_slot);
  return key;
}

std::string union_source(const std::vector<ShapeRef>& shapes) {
  return std::string(kRtcSource) + generate_union(shapes);
}

// Code object of the union kernel, compiled on first use (no device
// needed) and cached like the one-shape kernels.
bool union_code(const std::vector<ShapeRef>& shapes, const std::vector<char>** out,
                std::string& err) {
  const std::string key = union_key(shapes);
  {
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_code.find(key);
    if (it != g_code.end()) {
      *out = &it->second;
      return true;
    }
  }
  std::vector<char> code;
  if (!compile(union_source(shapes), code, err)) return false;
  std::lock_guard<std::mutex> lock(g_mu);
  log_shape(key);
  *out = &g_code.emplace(key, std::move(code)).first->second;
  return true;
}

hipError_t launch_union_batch(const BatchArgs& bargs, uint32_t n_blocks,
                              const std::vector<ShapeRef>& shapes, hipStream_t stream,
                              std::string& err) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const std::vector<char>* code = nullptr;
  if (!union_code(shapes, &code, err)) return hipErrorInvalidImage;
  const std::string key = std::to_string(dev) + "|" + union_key(shapes);
  Module m;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_mods.find(key);
    if (it != g_mods.end()) {
      m = it->second;
    } else {
      if ((e = hipModuleLoadData(&m.mod, code->data())) != hipSuccess) return e;
      if ((e = hipModuleGetFunction(&m.batch, m.mod, "p3_chaos_batch")) != hipSuccess) return e;
      g_mods.emplace(key, m);
    }
  }
  if (n_blocks == 0) return hipSuccess;
  BatchArgs args = bargs;
  void* params[] = {&args};
  return hipModuleLaunchKernel(m.batch, n_blocks, 1, 1, chaos_block(), 1, 1, 0, stream, params,
                               nullptr);
}

size_t chaos_compiled_shapes() {
  std::lock_guard<std::mutex> lock(g_mu);
  return g_code.size();
}

}  // namespace p3
