// pypde_hip host library: hipRTC specialisation of the ADER-WENO kernels,
// resident device state, step loop and the C ABI of include/pypde_hip.h.
//
// Reference counterparts: src/api.cpp:5-48 (entry points),
// src/solvers/iterator.cpp:38-151 (time loop), pypde/cfuncs.py:30-75 (the JIT
// step that turned Python physics into callable code; here hipRTC).
#include "pypde_hip.h"

#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <utime.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "tables.h"

namespace {

#include "kernels_embed.inc"  // PRELUDE_SRC, COMMON_SRC, KERNELS_SRC, STEP_SRC

constexpr int MAXV = 17;
thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[2048];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      return fail(PYPDE_ERR_HIP, "%s failed: %s", #x, hipGetErrorString(e_));       \
  } while (0)

// ------------------------------------------------------------ device structs
// Must match Geom / StepState / FaceBufs in kernels/ader_kernels.h (refl sized
// for the largest V the build accepts).
struct Geom {
  int nx[2];
  int nw[2];
  int nb[2];
  int bc[2];
  long ncell;
  long nwc;
  long nf[2];
  double dX[2];
  double refl[2][MAXV];
  int ndt;
  unsigned fmul[2], fsh[2], rmul[2], rsh[2];
  int idx32;
};

// Multiplier and shift of the device's udiv for a divisor 1 <= d < 2^31:
// sh = 31 + ceil(log2 d), mul = ceil(2^sh / d) (below 2^32).
static void udiv_magic(long d, unsigned* mul, unsigned* sh) {
  int L = 0;
  while ((1L << L) < d) L++;
  *sh = (unsigned)(31 + L);
  *mul = (unsigned)(((1ULL << (31 + L)) + (unsigned long long)d - 1) / (unsigned long long)d);
}

struct StepState {
  double t, dt, tf, cfl;
  long long count, max_steps;
  int push, done, nan_step, pad;
  unsigned long long maxbits;
  unsigned int nan_flag, pad2;
  unsigned long long pred_iters, pred_cells, pred_capped, pred_fallbacks;
  unsigned int pred_maxit, pad3;
  unsigned long long qhead;
  unsigned long long pred_nfull, pred_nshort;
  unsigned long long pred_iters32;
  unsigned long long hot_count;
  unsigned int dt_ticket, step_ticket;
  // predictor statistics spread over 64 slots, one 128-byte line each, so the
  // wavefronts' atomics do not queue on one address; the corrector's last
  // workgroup folds them into the totals above and clears them ([0]
  // iterations, [1] cells, [2] capped, [3] fallbacks, [4] max iterations, [5]
  // full and [6] shortcut node corrections, [7] FP32-phase iterations; [8]
  // k_dt's partial maximum, folded by k_dt's last workgroup)
  unsigned long long pstat[64][16];
};

struct FaceBufs {
  const double* ff[2];
  const double* fb[2];
};

// Must match Member / BlkMap of the batched ensemble kernels (ader_step.h).
struct Member {
  Geom g;
  double *u, *wh, *tr, *vol, *trd;
  double *ff[2], *fb[2];
  StepState* st;
  unsigned int* hot;
};
struct BlkMap {
  unsigned int start, gx, gy, gz;
};

// ------------------------------------------------------------ compilation
uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ull) {
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ull;
  }
  return h;
}

std::string cache_dir() {
  const char* e = getenv("PYPDE_HIP_CACHE");
  if (e && *e) return e;
  Dl_info info;
  if (dladdr((void*)&cache_dir, &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    size_t k = p.rfind('/');
    if (k != std::string::npos) return p.substr(0, k) + "/_jitcache";
  }
  return "/tmp/pypde_hip_cache";
}

struct Spec {
  std::string src;
  bool useF, useB, useS, useM, stiff, guess;
  int D, N, V, b_quirk;
  int flux = 0;  // 0 Rusanov, 1 Roe, 2 Osher (fluxes.h:6-8)
  bool second = false;  // F depends on the gradient (viscous, secondOrder)
  bool mixed = false;   // FP32 predictor phase before the FP64 one
  bool batch = false;   // also the batched ensemble kernels (one launch per stage)
};

std::string full_source(const Spec& s, const pypde::Tables& T) {
  std::ostringstream o;
  o << "// generated: ndim=" << s.D << " N=" << s.N << " V=" << s.V << "\n";
  o << "#define PDE_D " << s.D << "\n#define PDE_N " << s.N << "\n#define PDE_V " << s.V << "\n";
  o << "#define PDE_USE_F " << (int)s.useF << "\n#define PDE_USE_B " << (int)s.useB << "\n";
  o << "#define PDE_USE_S " << (int)s.useS << "\n#define PDE_USE_M " << (int)(s.useS && s.useM)
    << "\n";
  o << "#define PDE_STIFF " << (int)s.stiff << "\n#define PDE_B_QUIRK " << s.b_quirk << "\n";
  o << "#define PDE_GUESS " << (int)s.guess << "\n#define PDE_FLUX " << s.flux << "\n";
  o << "#define PDE_SECOND " << (int)s.second << "\n#define PDE_MIXED " << (int)s.mixed << "\n";
  o << "#define PDE_BATCH " << (int)s.batch << "\n";
  o << "constexpr int ipow_c(int b, int e) { return e <= 0 ? 1 : b * ipow_c(b, e - 1); }\n";
  o << pypde::tables_source(T, s.D);
  o << PRELUDE_SRC << "\n";
  o << "#line 1 \"system\"\n" << s.src << "\n";
  o << COMMON_SRC << "\n";
  if (s.mixed)  // the same physics once more, in FP32, for the mixed predictor's first phase
    o << REAL32_SRC << "\nnamespace f32 {\n#define double R\n#line 1 \"system (FP32)\"\n" << s.src
      << "\n#undef double\n}  // namespace f32\n";
  o << KERNELS_SRC << "\n" << STEP_SRC << "\n";
  return o.str();
}

const std::vector<std::string>& compile_options() {
  static const std::vector<std::string> o = {
      "--offload-arch=gfx950", "-O3", "-std=c++20", "-ffp-contract=off",
      "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  return o;
}

std::mutex g_mu;
std::map<uint64_t, std::vector<char>> g_codes;

// Compiles `src` for gfx950 with hipRTC, through a code-object cache keyed by
// the source text, the compile options and the hipRTC version (in memory, then
// <cache_dir>/<key>.hsaco, written to a temporary name and renamed into place).
// PYPDE_HIP_RECOMPILE=1 compiles even when a cached code object exists (the
// compile check of `python -m pypde_amd.build`); the result still replaces the
// cache entry.
int rtc_compile(const std::string& src, const char* what, std::vector<char>* code,
                uint64_t* key_out = nullptr) {
  std::string key = src;
  for (auto& x : compile_options()) key += "\n" + x;
  int ver = 0;
  hiprtcVersion(&ver, &ver);
  key += "\nhiprtc" + std::to_string(ver);
  uint64_t h = fnv1a(key);
  if (key_out) *key_out = h;
  const char* re = getenv("PYPDE_HIP_RECOMPILE");
  const bool recompile = re && *re && *re != '0';
  if (!recompile) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_codes.find(h);
    if (it != g_codes.end()) {
      *code = it->second;
      return 0;
    }
  }
  char name[64];
  snprintf(name, sizeof name, "%016llx", (unsigned long long)h);
  std::string dir = cache_dir();
  std::string path = dir + "/" + name + ".hsaco";
  if (!recompile) {
    std::ifstream f(path, std::ios::binary);
    if (f) {
      std::vector<char> c((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
      if (!c.empty()) {
        (void)utime(path.c_str(), nullptr);  // marks the entry as in use (the build prunes the rest)
        std::lock_guard<std::mutex> lk(g_mu);
        g_codes[h] = c;
        *code = c;
        return 0;
      }
    }
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), what, 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    return fail(PYPDE_ERR_COMPILE, "hiprtcCreateProgram failed (%s)", what);
  std::vector<const char*> opts;
  for (auto& x : compile_options()) opts.push_back(x.c_str());
  hiprtcResult r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  std::string log(ls, '\0');
  if (ls) hiprtcGetProgramLog(prog, &log[0]);
  if (r != HIPRTC_SUCCESS) {
    hiprtcDestroyProgram(&prog);
    return fail(PYPDE_ERR_COMPILE, "hipRTC compile of %s failed:\n%s", what, log.c_str());
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::vector<char> c(cs);
  hiprtcGetCode(prog, c.data());
  hiprtcDestroyProgram(&prog);
  mkdir(dir.c_str(), 0755);
  {
    std::string tmp = path + ".tmp" + std::to_string(getpid());
    std::ofstream f(tmp, std::ios::binary);
    if (f) {
      f.write(c.data(), (std::streamsize)c.size());
      f.close();
      if (f) rename(tmp.c_str(), path.c_str());
      else unlink(tmp.c_str());
    }
  }
  std::lock_guard<std::mutex> lk(g_mu);
  g_codes[h] = c;
  *code = c;
  return 0;
}

int compile_code(const Spec& spec, std::vector<char>* code, uint64_t* key = nullptr) {
  pypde::Tables T;
  if (pypde::make_tables(spec.N, &T)) return fail(PYPDE_ERR_ARG, "bad order N=%d", spec.N);
  std::string src = full_source(spec, T);
  if (const char* dump = getenv("PYPDE_HIP_DUMP_SRC")) {
    std::ofstream f(dump);
    f << src;
  }
  return rtc_compile(src, "pypde_ader.hip", code, key);
}

int check_spec(const Spec& s, bool secondOrder, int FLUX) {
  if (s.D < 1 || s.D > 2)
    return fail(PYPDE_ERR_UNSUPPORTED,
                "ndim=%d: only 1 and 2 are supported (the reference's node indexing, "
                "poly/evaluations.cpp:4-21, is only correct for ndim <= 2)", s.D);
  if (s.N < 1 || s.N > 6) return fail(PYPDE_ERR_ARG, "order N=%d outside 1..6", s.N);
  int nd = 1;
  for (int d = 0; d < s.D; d++) nd *= s.N;
  if (nd > 64) return fail(PYPDE_ERR_UNSUPPORTED, "N^ndim=%d > 64 nodes per cell", nd);
  if (s.V < 1 || s.V > MAXV) return fail(PYPDE_ERR_ARG, "V=%d outside 1..%d", s.V, MAXV);
  if (secondOrder && !s.useF)
    return fail(PYPDE_ERR_ARG, "secondOrder needs F (the viscous flux depends on the gradient)");
  if (FLUX < 0 || FLUX > 2)
    return fail(PYPDE_ERR_ARG, "FLUX=%d: 0 Rusanov, 1 Roe or 2 Osher (fluxes.h:6-8)", FLUX);
  if (s.mixed && (s.useB || secondOrder))
    return fail(PYPDE_ERR_UNSUPPORTED,
                "mixed precision: built for F, S (and M) systems without B or a viscous F");
  if (s.V > 8 && (FLUX != 0 || secondOrder))
    return fail(PYPDE_ERR_UNSUPPORTED,
                "V=%d: above 8 variables only the Rusanov flux and first-order F are built", s.V);
  if (!s.src.size() && (s.useF || s.useB || s.useS))
    return fail(PYPDE_ERR_ARG, "empty system source");
  return 0;
}

enum { K_WENO, K_DT, K_PRED, K_FLUX0, K_FLUX1, K_UPDATE, K_A2S, K_S2A,
       K_HARVEST, K_MLPGUESS, K_PRED_STATIC, K_PRED_STATIC2, K_EIG, K_P2R, K_PRED32,
       K_WENO_4x64, K_WENO_16x16, K_WENO_16x32, K_FLUX0_128, K_FLUX0_64, K_FLUX1_128, K_FLUX1_64,
       K_UPD1, K_UPD2, K_UPD4, K_UPD8, K_PRED_COLD, K_PRED_HOT, K_MLPPACK, K_PRED_COLD3,
       // batched ensemble kernels (modules built with PDE_BATCH only)
       K_B_WENO, K_B_WENO_4x64, K_B_WENO_16x16, K_B_WENO_16x32, K_B_DT, K_B_PRED_STATIC,
       K_B_PRED_STATIC2, K_B_PRED_COLD, K_B_PRED_COLD3, K_B_PRED_HOT, K_B_FLUX0, K_B_FLUX0_128,
       K_B_FLUX0_64, K_B_FLUX1, K_B_FLUX1_128, K_B_FLUX1_64, K_B_UPDATE, K_B_UPD1, K_B_UPD2,
       K_B_UPD4, K_B_UPD8, K_COUNT };
const char* kernel_names[K_COUNT] = {"k_weno",   "k_dt",    "k_predictor",
                                     "k_flux0",  "k_flux1", "k_update",
                                     "k_aos_to_soa", "k_soa_to_aos", "k_harvest", "k_mlp_guess",
                                     "k_predictor_static", "k_predictor_static2", "k_eig_batch",
                                     "k_planes_to_rows", "k_predictor_f32",
                                     "k_weno_4x64", "k_weno_16x16", "k_weno_16x32",
                                     "k_flux0_128", "k_flux0_64", "k_flux1_128", "k_flux1_64",
                                     "k_update_r1", "k_update_r2", "k_update_r4", "k_update_r8",
                                     "k_predictor_cold", "k_predictor_hot", "k_mlp_pack",
                                     "k_predictor_cold3",
                                     "k_weno_b", "k_weno_4x64_b", "k_weno_16x16_b", "k_weno_16x32_b",
                                     "k_dt_b", "k_predictor_static_b", "k_predictor_static2_b",
                                     "k_predictor_cold_b", "k_predictor_cold3_b", "k_predictor_hot_b",
                                     "k_flux0_b", "k_flux0_128_b", "k_flux0_64_b",
                                     "k_flux1_b", "k_flux1_128_b", "k_flux1_64_b",
                                     "k_update_b", "k_update_r1_b", "k_update_r2_b", "k_update_r4_b",
                                     "k_update_r8_b"};

// Launch variants the autotuner chooses between (every variant of a kernel
// computes identical results): WENO tiles, flux workgroup sizes, corrector
// mappings (one thread per cell and all variables, or one variable and R
// x-rows per thread), predictor schedules (static, work queue, and for the
// stiff predictor with the plain guess the cold/hot split, its cold pass
// built for two or for three wavefronts per SIMD).
constexpr int N_WENO_V = 4, N_FLUX_V = 3, N_UPD_V = 5, N_PRED_V = 4;
const int WENO_TILE[N_WENO_V][2] = {{8, 32}, {4, 64}, {16, 16}, {16, 32}};
const int WENO_KID[N_WENO_V] = {K_WENO, K_WENO_4x64, K_WENO_16x16, K_WENO_16x32};
const int FLUX_BS[N_FLUX_V] = {256, 128, 64};
const int FLUX_KID[2][N_FLUX_V] = {{K_FLUX0, K_FLUX0_128, K_FLUX0_64},
                                   {K_FLUX1, K_FLUX1_128, K_FLUX1_64}};
const int UPD_KID[N_UPD_V] = {K_UPDATE, K_UPD1, K_UPD2, K_UPD4, K_UPD8};
const int UPD_ROWS[N_UPD_V] = {0, 1, 2, 4, 8};
const int B_WENO_KID[N_WENO_V] = {K_B_WENO, K_B_WENO_4x64, K_B_WENO_16x16, K_B_WENO_16x32};
const int B_FLUX_KID[2][N_FLUX_V] = {{K_B_FLUX0, K_B_FLUX0_128, K_B_FLUX0_64},
                                     {K_B_FLUX1, K_B_FLUX1_128, K_B_FLUX1_64}};
const int B_UPD_KID[N_UPD_V] = {K_B_UPDATE, K_B_UPD1, K_B_UPD2, K_B_UPD4, K_B_UPD8};

struct MLPW {
  const float *W1t, *b1, *W2t, *b2, *W3t, *b3;
  const float *in_mean, *in_rstd, *out_mean, *out_std;
};

}  // namespace

// ------------------------------------------------------------------ context
struct pypde_ctx {
  Spec spec;
  pypde_options opts;
  Geom g;
  int D, N, V, ND, NT, CPW;
  hipModule_t mod = nullptr;
  uint64_t code_key = 0;  // the code object's cache key (kernel source and options)
  hipFunction_t fn[K_COUNT] = {};
  hipStream_t stream = nullptr;
  double *u = nullptr, *wh = nullptr, *tr = nullptr, *vol = nullptr, *ff[2] = {}, *fb[2] = {},
         *ret = nullptr, *guess = nullptr, *tmp = nullptr;
  double* trd = nullptr;    // face gradient traces (viscous F only)
  StepState* st = nullptr;
  StepState host_st;
  size_t ret_rows = 0;
  int device = 0;
  std::vector<hipEvent_t> ev;
  double ms_kernel[8] = {0};
  hipModule_t mod_guess = nullptr;  // learned-guess variant of the predictor
  hipFunction_t fn_plain[K_COUNT] = {}, fn_guess[K_COUNT] = {};
  bool use_model = false;
  int dt_lanes = 1;  // k_dt threads per w-cell (the kernels' DT_LANES)
  MLPW model;
  double* gbuf = nullptr;   // per-step MLP guesses
  void* mlp_pack = nullptr;  // the attached model's BF16 weight fragments
  float* mlp_aux = nullptr;  // ... and its FP32 normalisation and bias vectors
  double* qh = nullptr;     // full predictor output (harvest)
  int n_cu = 0;             // compute units of the device
  hipFunction_t pred_fn = nullptr;
  long pred_blocks = 0;     // persistent predictor workgroups for pred_fn
  bool pred_static = true;  // one wavefront per cell group instead of the work queue
  bool pred_w2 = false;     // the two-wavefronts-per-SIMD register budget
  bool pred_fixed = false;  // schedule fixed by PYPDE_PRED_SCHEDULE (not tuned)
  bool pred_split = false;  // cold pass over every cell, hot pass over the cells it hands on
  bool pred_cold3 = false;  // ... with the cold pass built for three wavefronts per SIMD
  unsigned int* hot = nullptr;  // the hot pass's cell list (nwc entries)
  struct {
    int weno = 0, flux = 0, update = 0;  // variant indices (WENO_KID, FLUX_KID, UPD_KID)
    bool done = false;                   // tuned, loaded or fixed
    int source = 0;                      // 0 defaults, 1 measured here, 2 tuning table, 3 forced
    double ms[4] = {0, 0, 0, 0};         // chosen weno, predictor, flux (both), corrector
  } tune;
};

namespace {

// Loaded code objects, one per (device, code-object key), shared by every
// context of the same specification (ensemble members, repeated solves) and
// unloaded when the last one is destroyed.
struct ModEntry {
  hipModule_t mod;
  int refs;
};
std::map<std::pair<int, uint64_t>, ModEntry> g_mods;
std::mutex g_mod_mu;

void release_module(int device, hipModule_t mod) {
  if (!mod) return;
  std::lock_guard<std::mutex> lk(g_mod_mu);
  for (auto it = g_mods.begin(); it != g_mods.end(); ++it)
    if (it->first.first == device && it->second.mod == mod) {
      if (--it->second.refs == 0) {
        (void)hipModuleUnload(mod);
        g_mods.erase(it);
      }
      return;
    }
  (void)hipModuleUnload(mod);
}

int load_module(pypde_ctx* c, const std::vector<char>& code, uint64_t key, hipModule_t* mod,
                hipFunction_t* fn) {
  {
    std::lock_guard<std::mutex> lk(g_mod_mu);
    auto it = g_mods.find({c->device, key});
    if (it != g_mods.end()) {
      it->second.refs++;
      *mod = it->second.mod;
    } else {
      if (hipModuleLoadData(mod, code.data()) != hipSuccess)
        return fail(PYPDE_ERR_HIP, "hipModuleLoadData failed");
      g_mods[{c->device, key}] = {*mod, 1};
    }
  }
  for (int k = 0; k < K_COUNT; k++) {
    fn[k] = nullptr;
    bool optional = ((k == K_FLUX1 || k == K_FLUX1_128 || k == K_FLUX1_64) && c->D < 2) ||
                    k == K_MLPGUESS || k == K_MLPPACK || (k == K_PRED32 && !c->spec.mixed) ||
                    k == K_PRED_COLD || k == K_PRED_HOT || k == K_PRED_COLD3 || k >= K_B_WENO;
    if (hipModuleGetFunction(&fn[k], *mod, kernel_names[k]) != hipSuccess) {
      fn[k] = nullptr;
      if (!optional) return fail(PYPDE_ERR_HIP, "kernel %s missing", kernel_names[k]);
    }
  }
  return 0;
}

int launch(pypde_ctx* c, int k, dim3 grid, dim3 block, std::vector<void*>& args) {
  if (!c->fn[k]) return fail(PYPDE_ERR_HIP, "kernel %s not loaded", kernel_names[k]);
  hipError_t e = hipModuleLaunchKernel(c->fn[k], grid.x, grid.y, grid.z, block.x, block.y,
                                       block.z, 0, c->stream, args.data(), nullptr);
  if (e != hipSuccess)
    return fail(PYPDE_ERR_HIP, "launch %s: %s", kernel_names[k], hipGetErrorString(e));
  return 0;
}

long cdiv(long a, long b) { return (a + b - 1) / b; }

// Callers built against the older, shorter pypde_stats leave out the node
// counters; struct_size says which layout the caller has.
bool stats_has_nodes(const pypde_stats* s) {
  return s->struct_size <= 0 ||
         (size_t)s->struct_size >= offsetof(pypde_stats, pred_iters_f32) + sizeof(long);
}

// Workgroups of the persistent predictor: as many as can be resident at once
// (occupancy x compute units; PYPDE_PRED_BLOCKS overrides), never more than
// there are groups of CPW cells.
long predictor_blocks(pypde_ctx* c) {
  if (c->pred_fn != c->fn[K_PRED]) {
    c->pred_fn = c->fn[K_PRED];
    int per_cu = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, c->pred_fn, 64, 0) != hipSuccess ||
        per_cu < 1)
      per_cu = 4;
    if (c->n_cu < 1) c->n_cu = 256;
    c->pred_blocks = (long)per_cu * c->n_cu;
    if (const char* e = getenv("PYPDE_PRED_BLOCKS")) {
      long v = atol(e);
      if (v > 0) c->pred_blocks = v;
    }
  }
  long groups = cdiv(c->g.nwc, c->CPW);
  return c->pred_blocks < groups ? c->pred_blocks : groups;
}

// One timestep: boundaries + WENO, CFL reduction, predictor, face fluxes,
// corrector, bookkeeping (iterator.cpp:99-148).  Every kernel returns at once
// when the device flag `done` is set, so whole chunks can be queued blind.
int enqueue_step(pypde_ctx* c, bool guess_now, std::vector<hipEvent_t>* evs) {
  Geom g = c->g;
  int rc;
  hipError_t mark_err = hipSuccess;  // first failure of a profiling event
  auto mark = [&]() {
    if (evs && mark_err == hipSuccess) {
      hipEvent_t e;
      mark_err = hipEventCreate(&e);
      if (mark_err != hipSuccess) return;
      evs->push_back(e);
      mark_err = hipEventRecord(e, c->stream);
    }
  };
  mark();
  {
    const double* uu = c->u;
    double* w = c->wh;
    const StepState* st = c->st;
    int ghost = 1, z = 0;
    long os = g.nwc;
    std::vector<void*> a = {&uu, &w, &st, &g, &ghost, &z, &z, &os};
    const int wv = c->D == 1 ? 0 : c->tune.weno;
    const int TX = WENO_TILE[wv][0], TY = WENO_TILE[wv][1];
    dim3 grid = c->D == 1 ? dim3((unsigned)cdiv(g.nw[0], 256))
                          : dim3((unsigned)cdiv(g.nw[1], TY), (unsigned)cdiv(g.nw[0], TX),
                                 (unsigned)c->V);
    if ((rc = launch(c, WENO_KID[wv], grid, dim3(256), a))) return rc;
  }
  mark();
  {
    const double* w = c->wh;
    StepState* st = c->st;
    std::vector<void*> a = {&w, &st, &g};
    // its last workgroup also turns the wave-speed maximum into dt
    if ((rc = launch(c, K_DT, dim3((unsigned)cdiv(g.nwc * c->dt_lanes, 256)), dim3(256), a)))
      return rc;
  }
  mark();
  if (c->use_model) {
    const double* w = c->wh;
    const StepState* st = c->st;
    const void* pk = c->mlp_pack;
    const float* ax = c->mlp_aux;
    double* gb = c->gbuf;
    long nwc = g.nwc;
    std::vector<void*> a = {&w, &st, &pk, &ax, &gb, &nwc};
    // four wavefronts of 32 cells per workgroup
    if ((rc = launch(c, K_MLPGUESS, dim3((unsigned)cdiv(g.nwc, 128)), dim3(256), a))) return rc;
  }
  if (c->spec.mixed) {
    // FP32 phase: from the learned guess, a caller's guess or the plain one;
    // its iterate becomes the FP64 predictor's initial guess
    const double* w = c->wh;
    const double* gin = c->use_model ? c->gbuf : (guess_now ? c->guess : nullptr);
    double* gout = c->gbuf;
    StepState* st = c->st;
    float tol32 = (float)c->opts.mixed_tol;
    int maxit32 = c->opts.mixed_maxit;
    std::vector<void*> a = {&w, &gin, &gout, &st, &g, &tol32, &maxit32};
    if ((rc = launch(c, K_PRED32, dim3((unsigned)cdiv(g.nwc, c->CPW)), dim3(64), a))) return rc;
  }
  mark();
  {
    const double* w = c->wh;
    const double* gs = (c->use_model || c->spec.mixed) ? c->gbuf : (guess_now ? c->guess : nullptr);
    double *t = c->tr, *v = c->vol, *qh = c->qh;
    StepState* st = c->st;
    double tol = c->opts.newton_tol;
    int maxit = c->opts.newton_maxit;
    double* td = c->trd;
    std::vector<void*> a = {&w, &gs, &t, &v, &qh, &st, &g, &tol, &maxit, &td};
    if (c->pred_split && c->fn[K_PRED_COLD] && c->fn[K_PRED_HOT] && !gs && c->hot) {
      unsigned int* hl = c->hot;
      std::vector<void*> b = {&w, &t, &v, &qh, &st, &g, &tol, &maxit, &hl};
      const unsigned blocks = (unsigned)cdiv(g.nwc, c->CPW);
      const int kc = (c->pred_cold3 && c->fn[K_PRED_COLD3]) ? K_PRED_COLD3 : K_PRED_COLD;
      if ((rc = launch(c, kc, dim3(blocks), dim3(64), b))) return rc;
      if ((rc = launch(c, K_PRED_HOT, dim3(blocks), dim3(64), b))) return rc;
    } else if (c->pred_static) {
      if ((rc = launch(c, c->pred_w2 ? K_PRED_STATIC2 : K_PRED_STATIC,
                       dim3((unsigned)cdiv(g.nwc, c->CPW)), dim3(64), a)))
        return rc;
    } else if ((rc = launch(c, K_PRED, dim3((unsigned)predictor_blocks(c)), dim3(64), a))) {
      return rc;
    }
  }
  mark();
  for (int d = 0; d < c->D; d++) {
    const double* t = c->tr;
    double *f = c->ff[d], *b = c->fb[d];
    const StepState* st = c->st;
    const double* td = c->trd;
    std::vector<void*> a = {&t, &f, &b, &st, &g, &td};
    long items = g.nf[d] * c->N * (c->ND / c->N);
    const int bs = FLUX_BS[c->tune.flux];
    if ((rc = launch(c, FLUX_KID[d][c->tune.flux], dim3((unsigned)cdiv(items, bs)), dim3(bs), a)))
      return rc;
    mark();
  }
  if (c->D == 1) mark();
  {
    double* uu = c->u;
    const double* v = c->vol;
    FaceBufs fbufs = {{c->ff[0], c->ff[1]}, {c->fb[0], c->fb[1]}};
    double* r = c->ret;
    StepState* st = c->st;
    std::vector<void*> a = {&uu, &v, &fbufs, &r, &st, &g};
    const int R = UPD_ROWS[c->tune.update];
    const long threads = R == 0 ? g.ncell : (long)c->V * cdiv(g.nx[0], R) * g.nx[1];
    if ((rc = launch(c, UPD_KID[c->tune.update], dim3((unsigned)cdiv(threads, 256)), dim3(256), a)))
      return rc;
  }
  mark();
  // the corrector's last workgroup advances t, count and the snapshot row
  // and folds the step's statistics: the advance stage keeps its marks and
  // times nothing
  mark();
  if (mark_err != hipSuccess)
    return fail(PYPDE_ERR_HIP, "profiling event: %s", hipGetErrorString(mark_err));
  return 0;
}

int make_spec(Spec* s, const char* src, bool useF, bool useB, bool useS, int ndim, int N, int V,
              bool STIFF, const pypde_options* o) {
  s->src = src ? src : "";
  s->useF = useF;
  s->useB = useB;
  s->useS = useS;
  s->useM = o && o->use_M;
  s->stiff = STIFF;
  s->guess = false;
  s->D = ndim;
  s->N = N;
  s->V = V;
  s->b_quirk = o ? o->b_quirk : 0;
  return 0;
}

}  // namespace

extern "C" {

void pypde_default_options(pypde_options* o) {
  memset(o, 0, sizeof *o);
  o->struct_size = sizeof *o;
  o->newton_tol = 1e-12;
  o->newton_maxit = 100;
  o->mixed_tol = 6e-6;
  o->mixed_maxit = 20;
}

const char* pypde_hip_last_error(void) { return g_err.c_str(); }

int pypde_hip_tables(int N, double* nodes, double* wghts, double* endvals, double* dervals,
                     double* weno_inv, double* sig, double* dgmat, double* dginv, double* fl_c,
                     double* fl_m) {
  pypde::Tables T;
  if (pypde::make_tables(N, &T)) return fail(PYPDE_ERR_ARG, "bad order N=%d", N);
  for (int i = 0; i < N; i++) {
    if (nodes) nodes[i] = T.x[i];
    if (wghts) wghts[i] = T.w[i];
    if (endvals) {
      endvals[i] = T.end0[i];
      endvals[N + i] = T.end1[i];
    }
    if (fl_c) fl_c[i] = T.flc[i];
    for (int j = 0; j < N; j++) {
      if (dervals) dervals[i * N + j] = T.der[i][j];
      if (sig) sig[i * N + j] = T.sig[i][j];
      if (dgmat) dgmat[i * N + j] = T.dgmat[i][j];
      if (dginv) dginv[i * N + j] = T.dginv[i][j];
      for (int s = 0; s < 4; s++)
        if (weno_inv) weno_inv[(s * N + i) * N + j] = T.winv[s][i][j];
      for (int k = 0; k < N; k++)
        if (fl_m) fl_m[(k * N + i) * N + j] = T.flm[k][i][j];
    }
  }
  return 0;
}

int pypde_hip_compile(const char* src, bool useF, bool useB, bool useS, int ndim, int N, int V,
                      bool STIFF, int FLUX, bool secondOrder, const pypde_options* opts) {
  Spec s;
  make_spec(&s, src, useF, useB, useS, ndim, N, V, STIFF, opts);
  s.flux = FLUX;
  s.second = secondOrder;
  s.mixed = opts && (size_t)opts->struct_size >= offsetof(pypde_options, mixed_maxit) + sizeof(int) &&
            opts->mixed;
  s.guess = s.mixed;
  int rc = check_spec(s, secondOrder, FLUX);
  if (rc) return rc;
  std::vector<char> code;
  if ((rc = compile_code(s, &code))) return rc;
  if (s.mixed) return 0;
  if (STIFF) {  // the learned-guess variant of the stiff predictor
    s.guess = true;
    if ((rc = compile_code(s, &code))) return rc;
  }
  return 0;
}

int pypde_hip_compile_ensemble(const char* src, bool useF, bool useB, bool useS, int ndim, int N,
                               int V, bool STIFF, int FLUX, bool secondOrder,
                               const pypde_options* opts) {
  Spec s;
  make_spec(&s, src, useF, useB, useS, ndim, N, V, STIFF, opts);
  s.flux = FLUX;
  s.second = secondOrder;
  s.batch = true;
  int rc = check_spec(s, secondOrder, FLUX);
  if (rc) return rc;
  std::vector<char> code;
  return compile_code(s, &code);
}

// Teardown has no status to return: failures of the releases are ignored.
void pypde_destroy(pypde_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (double* p : {c->u, c->wh, c->tr, c->vol, c->ff[0], c->ff[1], c->fb[0], c->fb[1], c->ret,
                    c->guess, c->tmp, c->trd})
    if (p) (void)hipFree(p);
  if (c->st) (void)hipFree(c->st);
  if (c->gbuf) (void)hipFree(c->gbuf);
  if (c->mlp_pack) (void)hipFree(c->mlp_pack);
  if (c->mlp_aux) (void)hipFree(c->mlp_aux);
  if (c->hot) (void)hipFree(c->hot);
  if (c->qh) (void)hipFree(c->qh);
  release_module(c->device, c->mod);
  release_module(c->device, c->mod_guess);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// A context; batch = true builds its code object with the batched ensemble
// kernels as well (ensemble members without a mixed-precision phase).
static pypde_ctx* create_ctx(const char* src, bool useF, bool useB, bool useS, const int* nX,
                             int ndim, const double* dX, double CFL, const int* boundaryTypes,
                             bool STIFF, int FLUX, int N, int V, bool secondOrder,
                             const pypde_options* opts_in, int* rc_out, bool batch) {
  int dummy;
  int& rc = rc_out ? *rc_out : dummy;
  pypde_options opts;
  pypde_default_options(&opts);
  if (opts_in) {
    size_t n = opts_in->struct_size > 0 && (size_t)opts_in->struct_size < sizeof opts
                   ? (size_t)opts_in->struct_size
                   : sizeof opts;
    memcpy(&opts, opts_in, n);
    opts.struct_size = sizeof opts;
    if (!(opts.newton_tol > 0)) opts.newton_tol = 1e-12;
    if (opts.newton_maxit <= 0) opts.newton_maxit = 100;
    if (!(opts.mixed_tol > 0)) opts.mixed_tol = 6e-6;
    if (opts.mixed_maxit <= 0) opts.mixed_maxit = 20;
  }
  Spec s;
  make_spec(&s, src, useF, useB, useS, ndim, N, V, STIFF, &opts);
  s.mixed = opts.mixed != 0;
  s.guess = opts.guess != nullptr || s.mixed;
  s.flux = FLUX;
  s.second = secondOrder;
  s.batch = batch && !s.mixed && !s.guess;
  if ((rc = check_spec(s, secondOrder, FLUX))) return nullptr;
  if (!nX || !dX || !boundaryTypes) {
    rc = fail(PYPDE_ERR_ARG, "null grid argument");
    return nullptr;
  }
  for (int d = 0; d < ndim; d++) {
    if (nX[d] < 1) {
      rc = fail(PYPDE_ERR_ARG, "nX[%d]=%d", d, nX[d]);
      return nullptr;
    }
    if (boundaryTypes[d] < 0 || boundaryTypes[d] > 2) {
      rc = fail(PYPDE_ERR_ARG, "boundaryTypes[%d]=%d (0 transmissive, 1 periodic, 2 reflective)",
                d, boundaryTypes[d]);
      return nullptr;
    }
    if (!(dX[d] > 0)) {
      rc = fail(PYPDE_ERR_ARG, "dX[%d] must be positive", d);
      return nullptr;
    }
    if (boundaryTypes[d] != 0 && nX[d] < N) {
      rc = fail(PYPDE_ERR_ARG, "periodic/reflective axis %d needs nX >= N (%d < %d)", d, nX[d], N);
      return nullptr;
    }
  }
  std::vector<char> code;
  uint64_t code_key = 0;
  if ((rc = compile_code(s, &code, &code_key))) return nullptr;
  pypde_ctx* c = new pypde_ctx();
  c->spec = s;
  c->code_key = code_key;
  c->opts = opts;
  c->D = ndim;
  c->N = N;
  c->V = V;
  c->ND = 1;
  for (int d = 0; d < ndim; d++) c->ND *= N;
  c->NT = c->ND * N;
  c->CPW = 64 / c->ND;
  c->dt_lanes = (ndim == 2 && V <= 5 && !secondOrder) ? 2 : 1;  // DT_PAIR in ader_kernels.h
  c->device = opts.device;
  // Predictor schedule: the autotuner picks among the static one (one
  // wavefront per group of CPW cells), the persistent work-queued kernel and
  // the stiff cold/hot split; PYPDE_PRED_SCHEDULE fixes one instead.
  c->pred_static = true;
  if (const char* e = getenv("PYPDE_PRED_SCHEDULE")) {
    c->pred_static = strcmp(e, "queue") != 0;
    c->pred_w2 = strcmp(e, "static2") == 0;
    c->pred_split = strcmp(e, "split") == 0;
    c->pred_fixed = true;
  }
  Geom& g = c->g;
  memset(&g, 0, sizeof g);
  g.ncell = 1;
  g.nwc = 1;
  for (int d = 0; d < 2; d++) {
    bool on = d < ndim;
    g.nx[d] = on ? nX[d] : 1;
    g.nw[d] = on ? nX[d] + 2 : 1;
    g.nb[d] = on ? nX[d] + 2 * N : 1;
    g.bc[d] = on ? boundaryTypes[d] : 0;
    g.dX[d] = on ? dX[d] : 1.0;
    for (int v = 0; v < MAXV; v++) g.refl[d][v] = (on && opts.reflect && v < V) ? opts.reflect[d * V + v] : 1.0;
    g.ncell *= g.nx[d];
    g.nwc *= g.nw[d];
  }
  for (int d = 0; d < ndim; d++) {
    g.nf[d] = 1;
    for (int e = 0; e < ndim; e++) g.nf[d] *= (e == d) ? nX[e] + 1 : nX[e];
  }
  g.idx32 = g.nwc < (1L << 31);
  for (int d = 0; d < 2; d++) {
    long nf = d < ndim ? g.nf[d] : 1;
    udiv_magic(nf, &g.fmul[d], &g.fsh[d]);
    udiv_magic(d == 0 ? g.nx[1] : g.nx[1] + 1, &g.rmul[d], &g.rsh[d]);
    if (nf * N * (c->ND / N) >= (1L << 31)) g.idx32 = 0;
  }
  // PYPDE_FLUX_IDX64=1: the flux kernels' 64-bit index path on any grid (for
  // testing the path that grids of 2^31 flux items or w-cells take)
  if (const char* e = getenv("PYPDE_FLUX_IDX64"))
    if (e[0] == '1') g.idx32 = 0;
  auto bail = [&](int code_) {
    rc = code_;
    pypde_destroy(c);
    return (pypde_ctx*)nullptr;
  };
  if (hipSetDevice(c->device) != hipSuccess)
    return bail(fail(PYPDE_ERR_HIP, "hipSetDevice(%d) failed: no usable GPU", c->device));
  if (hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, c->device) !=
      hipSuccess)
    c->n_cu = 0;
  if (load_module(c, code, code_key, &c->mod, c->fn_plain)) return bail(PYPDE_ERR_HIP);
  memcpy(c->fn, c->fn_plain, sizeof c->fn);
  if (hipStreamCreate(&c->stream) != hipSuccess)
    return bail(fail(PYPDE_ERR_HIP, "hipStreamCreate failed"));
  size_t NTR = 2 * (size_t)ndim * N * (c->ND / N) * V;
  auto alloc = [&](double** p, size_t n) {
    return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(double)) == hipSuccess;
  };
  bool ok = alloc(&c->u, g.ncell * V) && alloc(&c->wh, g.nwc * c->ND * V) &&
            alloc(&c->tr, g.nwc * NTR) && alloc(&c->tmp, g.ncell * V) &&
            (!secondOrder || alloc(&c->trd, g.nwc * NTR * ndim)) &&
            ((!useS && !useB) || alloc(&c->vol, g.nwc * V));
  for (int d = 0; d < ndim && ok; d++) {
    size_t fs = (size_t)g.nf[d] * N * (c->ND / N) * V;
    ok = (!useF || alloc(&c->ff[d], fs)) && (!useB || alloc(&c->fb[d], fs));
  }
  if (ok && opts.guess) {
    ok = alloc(&c->guess, g.nwc * c->NT * V);
    if (ok) {
      // host layout [cell][t][node][v] -> device planes [(t*ND+node)*V+v][cell]
      std::vector<double> h(g.nwc * c->NT * V);
      for (long cc = 0; cc < g.nwc; cc++)
        for (int p = 0; p < c->NT * V; p++) h[(size_t)p * g.nwc + cc] = opts.guess[cc * c->NT * V + p];
      ok = hipMemcpy(c->guess, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    }
  }
  if (ok && s.mixed) ok = alloc(&c->gbuf, g.nwc * c->NT * V);
  if (ok && c->fn[K_PRED_COLD])
    ok = hipMalloc((void**)&c->hot, std::max<long>(g.nwc, 1) * sizeof(unsigned int)) == hipSuccess;
  ok = ok && hipMalloc((void**)&c->st, sizeof(StepState)) == hipSuccess;
  if (!ok) return bail(fail(PYPDE_ERR_HIP, "device allocation failed"));
  memset(&c->host_st, 0, sizeof c->host_st);
  c->host_st.cfl = CFL;
  c->host_st.done = 1;
  if (hipMemcpy(c->st, &c->host_st, sizeof(StepState), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(PYPDE_ERR_HIP, "state upload failed"));
  rc = 0;
  return c;
}

pypde_ctx* pypde_create(const char* src, bool useF, bool useB, bool useS, int* nX, int ndim,
                        double* dX, double CFL, int* boundaryTypes, bool STIFF, int FLUX, int N,
                        int V, bool secondOrder, const pypde_options* opts_in, int* rc_out) {
  return create_ctx(src, useF, useB, useS, nX, ndim, dX, CFL, boundaryTypes, STIFF, FLUX, N, V,
                    secondOrder, opts_in, rc_out, false);
}

int pypde_set_state(pypde_ctx* c, const double* u) {
  HIPCHK(hipSetDevice(c->device));
  size_t n = c->g.ncell * c->V;
  HIPCHK(hipMemcpyAsync(c->tmp, u, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  std::vector<void*> a = {&c->tmp, &c->u, &c->g.ncell};
  int rc = launch(c, K_A2S, dim3((unsigned)cdiv(n, 256)), dim3(256), a);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int pypde_get_state(pypde_ctx* c, double* u) {
  HIPCHK(hipSetDevice(c->device));
  size_t n = c->g.ncell * c->V;
  std::vector<void*> a = {&c->u, &c->tmp, &c->g.ncell};
  int rc = launch(c, K_S2A, dim3((unsigned)cdiv(n, 256)), dim3(256), a);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(u, c->tmp, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

void* pypde_device_state(pypde_ctx* c) { return c->u; }
void* pypde_stream(pypde_ctx* c) { return c->stream; }

// ------------------------------------------------------------- autotuner
// Picks, per problem (system source, ndim, order, nvar, flags, grid size,
// GPU architecture), the WENO tile, the flux workgroup size, the corrector
// mapping and the predictor schedule.  Trials are real timesteps from the
// current state, each kernel bracketed by events; u and the step state are
// restored after every trial, so tuning changes nothing but the launch
// choices.  Choices persist in <code-object cache>/tuning.txt (one line per
// problem; the last line for a key wins) and in memory for the process.
// PYPDE_TUNE=0 keeps the defaults; PYPDE_TUNE_FORCE="w,f,u,p" fixes them.
namespace {
std::map<uint64_t, std::vector<int>> g_tuned;

uint64_t tune_key(pypde_ctx* c) {
  std::ostringstream k;
  k << c->spec.src << "|" << c->D << c->N << c->V << c->spec.useF << c->spec.useB << c->spec.useS
    << c->spec.useM << c->spec.stiff << c->spec.guess << c->spec.mixed << c->spec.flux
    << c->spec.second << c->spec.b_quirk << "|" << c->g.nx[0] << "x" << c->g.nx[1] << "|"
    << std::hex << c->code_key << std::dec << "|";
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) k << prop.gcnArchName;
  return fnv1a(k.str());
}

void apply_tuning(pypde_ctx* c, const std::vector<int>& v) {
  c->tune.weno = (c->D == 2 && v[0] >= 0 && v[0] < N_WENO_V) ? v[0] : 0;
  c->tune.flux = (v[1] >= 0 && v[1] < N_FLUX_V) ? v[1] : 0;
  c->tune.update = (v[2] >= 0 && v[2] < N_UPD_V) ? v[2] : 0;
  if (!c->pred_fixed) {
    c->pred_static = v[3] != 1;
    c->pred_split = (v[3] == 2 || v[3] == 3) && c->fn[K_PRED_COLD] != nullptr;
    c->pred_cold3 = v[3] == 3;
  }
}

int autotune(pypde_ctx* c) {
  if (c->tune.done) return 0;
  c->tune.done = true;
  const char* off = getenv("PYPDE_TUNE");
  if (off && strcmp(off, "0") == 0) return 0;
  if (const char* f = getenv("PYPDE_TUNE_FORCE")) {
    std::vector<int> v(4, 0);
    if (sscanf(f, "%d,%d,%d,%d", &v[0], &v[1], &v[2], &v[3]) != 4)
      return fail(PYPDE_ERR_ARG, "PYPDE_TUNE_FORCE=%s: expected w,f,u,p", f);
    apply_tuning(c, v);
    c->tune.source = 3;
    return 0;
  }
  const uint64_t key = tune_key(c);
  const std::string path = cache_dir() + "/tuning.txt";
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tuned.find(key);
    if (it == g_tuned.end()) {
      std::ifstream f(path);
      std::string line;
      while (std::getline(f, line)) {
        unsigned long long k = 0;
        std::vector<int> v(4, 0);
        if (sscanf(line.c_str(), "%llx %d %d %d %d", &k, &v[0], &v[1], &v[2], &v[3]) == 5 &&
            k == key)
          g_tuned[key] = v;
      }
      it = g_tuned.find(key);
    }
    if (it != g_tuned.end()) {
      apply_tuning(c, it->second);
      c->tune.source = 2;
      return 0;
    }
  }
  if (c->host_st.done) return 0;  // nothing to time: keep the defaults
  // back up the state, run the trials, restore after each
  const size_t ub = c->g.ncell * c->V * sizeof(double);
  StepState* st_bak = nullptr;
  HIPCHK(hipMalloc((void**)&st_bak, sizeof(StepState)));
  HIPCHK(hipMemcpyAsync(c->tmp, c->u, ub, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(st_bak, c->st, sizeof(StepState), hipMemcpyDeviceToDevice, c->stream));
  // every trial step starts from the run's state, but with the step count
  // past the first steps' dt ramp (x0.2 while count <= 5, stepper.cpp:69-71):
  // at the ramped dt the stiff predictor has fewer full Newton solves and
  // iterations than for the rest of the run, which would favour the wrong
  // variants
  StepState trial;
  HIPCHK(hipMemcpyAsync(&trial, c->st, sizeof(StepState), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (trial.count < 6) trial.count = 6;
  StepState* st_trial = nullptr;
  if (hipMalloc((void**)&st_trial, sizeof(StepState)) != hipSuccess) {
    (void)hipFree(st_bak);
    return fail(PYPDE_ERR_HIP, "autotune: allocation failed");
  }
  if (hipMemcpy(st_trial, &trial, sizeof(StepState), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpyAsync(c->st, st_trial, sizeof(StepState), hipMemcpyDeviceToDevice, c->stream) !=
          hipSuccess) {
    (void)hipFree(st_bak);
    (void)hipFree(st_trial);
    return fail(PYPDE_ERR_HIP, "autotune: trial state upload failed");
  }
  double* keep_ret = c->ret;
  const bool keep_static = c->pred_static, keep_split = c->pred_split, keep_c3 = c->pred_cold3;
  c->ret = nullptr;
  const int nw = c->D == 2 ? N_WENO_V : 1;
  const int np = c->pred_fixed ? 1 : (c->fn[K_PRED_COLD] && c->hot ? N_PRED_V : 2);
  const int ntrial = std::max(std::max(nw, N_FLUX_V), std::max(N_UPD_V, np));
  const int reps = 2;
  std::vector<std::vector<hipEvent_t>> evs(1 + reps * ntrial);
  int rc = 0;
  for (size_t k = 0; k < evs.size() && !rc; k++) {
    const int tk = k == 0 ? 0 : (int)((k - 1) % ntrial);
    c->tune.weno = tk % nw;
    c->tune.flux = tk % N_FLUX_V;
    c->tune.update = tk % N_UPD_V;
    if (!c->pred_fixed) {
      c->pred_static = (tk % np) != 1;
      c->pred_split = (tk % np) >= 2;
      c->pred_cold3 = (tk % np) == 3;
    }
    rc = enqueue_step(c, false, &evs[k]);
    if (!rc && hipMemcpyAsync(c->u, c->tmp, ub, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
      rc = fail(PYPDE_ERR_HIP, "autotune: state restore failed");
    if (!rc && hipMemcpyAsync(c->st, st_trial, sizeof(StepState), hipMemcpyDeviceToDevice,
                              c->stream) != hipSuccess)
      rc = fail(PYPDE_ERR_HIP, "autotune: state restore failed");
  }
  if (!rc && hipMemcpyAsync(c->st, st_bak, sizeof(StepState), hipMemcpyDeviceToDevice,
                            c->stream) != hipSuccess)
    rc = fail(PYPDE_ERR_HIP, "autotune: state restore failed");
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess)
    rc = fail(PYPDE_ERR_HIP, "autotune: trial steps failed");
  (void)hipFree(st_trial);
  // per variant, the fastest of its trials; event marks per step:
  // weno | dt | guess | predictor | flux0 | flux1 | update | advance
  double tw[N_WENO_V], tf[N_FLUX_V], tu[N_UPD_V], tp[N_PRED_V];
  std::fill(tw, tw + N_WENO_V, 1e30);
  std::fill(tf, tf + N_FLUX_V, 1e30);
  std::fill(tu, tu + N_UPD_V, 1e30);
  std::fill(tp, tp + N_PRED_V, 1e30);
  auto span = [&](std::vector<hipEvent_t>& e, int a, int b) {
    float ms = 0;
    if ((int)e.size() > b && hipEventElapsedTime(&ms, e[a], e[b]) == hipSuccess) return (double)ms;
    return 1e30;
  };
  for (size_t k = 1; k < evs.size() && !rc; k++) {
    const int tk = (int)((k - 1) % ntrial);
    tw[tk % nw] = std::min(tw[tk % nw], span(evs[k], 0, 1));
    tp[tk % np] = std::min(tp[tk % np], span(evs[k], 3, 4));
    tf[tk % N_FLUX_V] = std::min(tf[tk % N_FLUX_V], span(evs[k], 4, 6));
    tu[tk % N_UPD_V] = std::min(tu[tk % N_UPD_V], span(evs[k], 6, 7));
  }
  for (auto& e : evs)
    for (auto x : e) (void)hipEventDestroy(x);
  (void)hipFree(st_bak);
  c->ret = keep_ret;
  c->pred_static = keep_static;
  c->pred_split = keep_split;
  c->pred_cold3 = keep_c3;
  if (rc) return rc;
  auto best = [](const double* t, int n) {
    int b = 0;
    for (int i = 1; i < n; i++)
      if (t[i] < t[b]) b = i;
    return b;
  };
  std::vector<int> v = {best(tw, nw), best(tf, N_FLUX_V), best(tu, N_UPD_V),
                        c->pred_fixed ? (c->pred_static ? 0 : 1) : best(tp, np)};
  apply_tuning(c, v);
  c->tune.source = 1;
  c->tune.ms[0] = tw[v[0]];
  c->tune.ms[1] = tp[c->pred_fixed ? 0 : v[3]];
  c->tune.ms[2] = tf[v[1]];
  c->tune.ms[3] = tu[v[2]];
  std::lock_guard<std::mutex> lk(g_mu);
  g_tuned[key] = v;
  mkdir(cache_dir().c_str(), 0755);
  std::ofstream f(path, std::ios::app);
  if (f) {
    char line[160];
    snprintf(line, sizeof line, "%016llx %d %d %d %d   # weno %.4f pred %.4f flux %.4f update %.4f ms\n",
             (unsigned long long)key, v[0], v[1], v[2], v[3], c->tune.ms[0], c->tune.ms[1],
             c->tune.ms[2], c->tune.ms[3]);
    f << line;
  }
  return 0;
}
}  // namespace

// One timestep of each context captured into a graph: a fork event on the
// first context's stream, every context's step on its own stream, joined back.
static int capture_steps(const std::vector<pypde_ctx*>& cs, hipGraphExec_t* exec) {
  hipStream_t s0 = cs[0]->stream;
  std::vector<hipEvent_t> evs(cs.size() + 1, nullptr);
  for (auto& e : evs) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  int rc = 0;
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
  // fork and join: an error here is reported, and the capture still ends
  auto chk = [&](hipError_t e, const char* what) {
    if (!rc && e != hipSuccess) rc = fail(PYPDE_ERR_HIP, "graph capture, %s: %s", what, hipGetErrorString(e));
  };
  chk(hipEventRecord(evs[0], s0), "fork event");
  for (size_t i = 0; i < cs.size() && !rc; i++) {
    if (i > 0) chk(hipStreamWaitEvent(cs[i]->stream, evs[0], 0), "fork wait");
    if (!rc) rc = enqueue_step(cs[i], false, nullptr);
    if (!rc && i > 0) {
      chk(hipEventRecord(evs[i + 1], cs[i]->stream), "join event");
      chk(hipStreamWaitEvent(s0, evs[i + 1], 0), "join wait");
    }
  }
  hipError_t e = hipStreamEndCapture(s0, &graph);
  if (!rc && e != hipSuccess) rc = fail(PYPDE_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
  if (!rc && hipGraphInstantiate(exec, graph, nullptr, nullptr, 0) != hipSuccess)
    rc = fail(PYPDE_ERR_HIP, "graph instantiation failed");
  if (graph) (void)hipGraphDestroy(graph);
  for (auto ev : evs) (void)hipEventDestroy(ev);
  return rc;
}

static int run_loop(pypde_ctx* c, long max_steps, pypde_stats* stats, bool first_guess) {
  if (int r = autotune(c)) return r;
  auto t0 = std::chrono::steady_clock::now();
  int chunk = 4;
  bool guess_now = first_guess;
  std::vector<hipEvent_t> evs;
  hipGraphExec_t exec = nullptr;
  const bool graph = c->opts.use_graph && !c->opts.profile;
  int rc = 0;
  while (true) {
    // never queue past a step cap: the launch count equals the step count
    if (max_steps > 0) {
      long left = max_steps - (long)c->host_st.count;
      if (left < chunk) chunk = (int)std::max(1L, left);
    }
    for (int i = 0; i < chunk && !rc; i++) {
      if (graph && !guess_now) {
        if (!exec && (rc = capture_steps({c}, &exec))) break;
        if (hipGraphLaunch(exec, c->stream) != hipSuccess) rc = fail(PYPDE_ERR_HIP, "graph launch failed");
      } else {
        std::vector<hipEvent_t>* ep = c->opts.profile ? &evs : nullptr;
        rc = enqueue_step(c, guess_now, ep);
      }
      guess_now = false;
    }
    if (rc) break;
    const long long count_before = c->host_st.count;
    HIPCHK(hipMemcpyAsync(&c->host_st, c->st, sizeof(StepState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->host_st.done) break;
    if (c->host_st.count == count_before) {
      rc = fail(PYPDE_ERR_HIP, "no progress in %d steps", chunk);
      break;
    }
    if (chunk < 64) chunk *= 2;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (rc) {
    for (auto e : evs) (void)hipEventDestroy(e);
    return rc;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (c->opts.profile) {
    // 9 marks per step bracket: weno | dt | guess | predictor | flux0 | flux1 | update | advance
    for (size_t s0 = 0; s0 + 9 <= evs.size(); s0 += 9)
      for (int k = 0; k < 8; k++) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, evs[s0 + k], evs[s0 + k + 1]) == hipSuccess)
          c->ms_kernel[k] += ms;
      }
    for (auto e : evs) (void)hipEventDestroy(e);
  }
  if (stats) {
    stats->steps = c->host_st.count;
    stats->t = c->host_st.t;
    stats->pred_iters = (long)c->host_st.pred_iters;
    stats->pred_cells = (long)c->host_st.pred_cells;
    stats->pred_maxit = (long)c->host_st.pred_maxit;
    stats->pred_capped = (long)c->host_st.pred_capped;
    stats->guess_fallbacks = (long)c->host_st.pred_fallbacks;
    stats->nan_step = c->host_st.nan_step;
    stats->ms_loop = std::chrono::duration<double, std::milli>(t1 - t0).count();
    for (int k = 0; k < 8; k++) stats->ms_kernel[k] = c->ms_kernel[k];
    if (stats_has_nodes(stats)) {
      stats->pred_node_full = (long)c->host_st.pred_nfull;
      stats->pred_node_short = (long)c->host_st.pred_nshort;
      stats->pred_iters_f32 = (long)c->host_st.pred_iters32;
    }
  }
  return c->host_st.nan_step ? -c->host_st.nan_step : 0;
}

int pypde_run(pypde_ctx* c, double tf, long max_steps, int ndt, double* ret, pypde_stats* stats) {
  HIPCHK(hipSetDevice(c->device));
  if (ndt < 0) return fail(PYPDE_ERR_ARG, "ndt < 0");
  size_t rows = (size_t)ndt;
  size_t row = c->g.ncell * c->V;
  if (rows > c->ret_rows) {
    if (c->ret) (void)hipFree(c->ret);
    c->ret = nullptr;
    HIPCHK(hipMalloc((void**)&c->ret, rows * row * sizeof(double)));
    c->ret_rows = rows;
  }
  if (rows) HIPCHK(hipMemsetAsync(c->ret, 0, rows * row * sizeof(double), c->stream));
  c->g.ndt = ndt;
  StepState s;
  memset(&s, 0, sizeof s);
  s.tf = tf;
  s.cfl = c->host_st.cfl;
  s.max_steps = max_steps;
  s.done = (0.0 < tf) ? 0 : 1;
  c->host_st = s;
  HIPCHK(hipMemcpyAsync(c->st, &s, sizeof s, hipMemcpyHostToDevice, c->stream));
  for (double& m : c->ms_kernel) m = 0;
  double* keep_ret = c->ret;
  if (!rows) c->ret = nullptr;
  int rc = run_loop(c, max_steps, stats, c->guess != nullptr);
  c->ret = keep_ret;
  if (rc > 0) return rc;
  if (rows) {
    // final state into the last row (iterator.cpp:150)
    std::vector<void*> a = {&c->u, &c->tmp, &c->g.ncell};
    int r2 = launch(c, K_S2A, dim3((unsigned)cdiv(row, 256)), dim3(256), a);
    if (r2) return r2;
    HIPCHK(hipMemcpyAsync(c->ret + (rows - 1) * row, c->tmp, row * sizeof(double),
                          hipMemcpyDeviceToDevice, c->stream));
    if (ret) HIPCHK(hipMemcpyAsync(ret, c->ret, rows * row * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return rc;
}

int pypde_step(pypde_ctx* c, long steps, double tf, pypde_stats* stats) {
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(&c->host_st, c->st, sizeof(StepState), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  StepState s = c->host_st;
  s.tf = tf;
  s.max_steps = s.count + steps;
  s.done = (s.t < tf && steps > 0) ? 0 : 1;
  c->host_st = s;
  c->g.ndt = 0;
  HIPCHK(hipMemcpyAsync(c->st, &s, sizeof s, hipMemcpyHostToDevice, c->stream));
  for (double& m : c->ms_kernel) m = 0;
  double* keep_ret = c->ret;
  c->ret = nullptr;
  int rc = run_loop(c, s.max_steps, stats, false);
  c->ret = keep_ret;
  return rc;
}

// ------------------------------------------------------------ ensembles
// Stages of a batched ensemble step, each one launch per group of members.
enum { BS_WENO, BS_DT, BS_PRED, BS_FLUX0, BS_FLUX1, BS_UPDATE, BS_COUNT };

// Members advanced by the same launches: entries [first, first + count) of
// the member table, workgroup maps at map (BS_COUNT x (count + 1) entries),
// launched on `stream`.
struct EnsGroup {
  int first = 0, count = 0;
  int lead = 0;  // the smallest member index in the group
  size_t map = 0;
  unsigned int blocks[BS_COUNT] = {};
  hipStream_t stream = nullptr;
  // launch variants: those the autotuner chose for the group's largest member
  // (the work queue replaced by the static schedule)
  int weno = 0, flux = 0, update = 0;
  int pred = 0;  // 0 static, 1 static (two-wavefront budget), 2 split, 3 split (cold pass x3)
};

struct pypde_ensemble {
  std::vector<pypde_ctx*> m;
  // batched mode (members built with the batched kernels): the member table
  // (members in group order), the groups and their workgroup maps, and the
  // launch variants every member runs with
  Member* d_ms = nullptr;
  BlkMap* d_map = nullptr;
  size_t n_map = 0;
  std::vector<EnsGroup> groups;
  std::vector<int> order;  // member table entry -> member index
};

void pypde_ensemble_destroy(pypde_ensemble* e) {
  if (!e) return;
  if (!e->m.empty()) (void)hipSetDevice(e->m[0]->device);
  if (e->d_ms) (void)hipFree(e->d_ms);
  if (e->d_map) (void)hipFree(e->d_map);
  for (auto c : e->m) pypde_destroy(c);
  delete e;
}

// Groups, launch variants and workgroup maps of the batched step.  Members
// are taken largest first and packed greedily into groups of at most
// `cap` reconstructed cells (a larger member forms its own group, which runs
// the member's own single-grid step).  Measured on 2D Euler, order 3: one
// launch per stage over all 48 benchmark grids (3.1M cells) is 6 % slower
// than groups of at most 2^18 cells, as between two stages of one launch
// the data of every grid passes through HBM instead of staying in the
// last-level cache, while on 48 grids of 64 x 64 cells (one group either
// way) batching runs at 2.1x one stream per grid.  Each group runs with the
// variants the autotuner chose for its largest member.
static int batch_setup(pypde_ensemble* e, long cap) {
  const int nm = (int)e->m.size();
  std::vector<int>& order = e->order;
  order.resize(nm);
  for (int i = 0; i < nm; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return e->m[a]->g.nwc > e->m[b]->g.nwc; });
  const pypde_ctx* c0 = e->m[0];
  const int D = c0->D, N = c0->N, V = c0->V, ND = c0->ND, CPW = c0->CPW;
  e->groups.clear();
  long cells = 0;
  for (int k = 0; k < nm; k++) {
    const long n = e->m[order[k]]->g.nwc;
    if (e->groups.empty() || (cap > 0 && cells + n > cap)) {
      const pypde_ctx* big = e->m[order[k]];
      EnsGroup gr;
      gr.first = k;
      gr.stream = big->stream;
      gr.weno = D == 1 ? 0 : big->tune.weno;
      gr.flux = big->tune.flux;
      gr.update = big->tune.update;
      if (big->pred_split && big->hot && big->fn[K_B_PRED_COLD] && big->fn[K_B_PRED_HOT])
        gr.pred = (big->pred_cold3 && big->fn[K_B_PRED_COLD3]) ? 3 : 2;
      else if (big->pred_static && big->pred_w2)
        gr.pred = 1;
      e->groups.push_back(gr);
      cells = 0;
    }
    EnsGroup& cur = e->groups.back();
    cur.lead = cur.count == 0 ? order[k] : std::min(cur.lead, order[k]);
    cur.count++;
    cells += n;
  }
  std::vector<BlkMap> h;
  for (auto& gr : e->groups) {
    gr.map = h.size();
    h.resize(h.size() + (size_t)BS_COUNT * (gr.count + 1));
    for (int s = 0; s < BS_COUNT; s++) {
      unsigned long total = 0;
      for (int i = 0; i < gr.count; i++) {
        const pypde_ctx* mc = e->m[order[gr.first + i]];
        const Geom& g = mc->g;
        long gx = 1, gy = 1, gz = 1;
        switch (s) {
          case BS_WENO:
            if (D == 1) {
              gx = cdiv(g.nw[0], 256);
            } else {
              gx = cdiv(g.nw[1], WENO_TILE[gr.weno][1]);
              gy = cdiv(g.nw[0], WENO_TILE[gr.weno][0]);
              gz = V;
            }
            break;
          case BS_DT: gx = cdiv(g.nwc * mc->dt_lanes, 256); break;
          case BS_PRED: gx = cdiv(g.nwc, CPW); break;
          case BS_FLUX0:
          case BS_FLUX1: {
            const int d = s - BS_FLUX0;
            gx = d < D ? cdiv(g.nf[d] * N * (ND / N), FLUX_BS[gr.flux]) : 0;
            break;
          }
          default: {
            const int R = UPD_ROWS[gr.update];
            gx = cdiv(R == 0 ? g.ncell : (long)V * cdiv(g.nx[0], R) * g.nx[1], 256);
          }
        }
        h[gr.map + (size_t)s * (gr.count + 1) + i] = {(unsigned)total, (unsigned)gx, (unsigned)gy,
                                                       (unsigned)gz};
        total += (unsigned long)(gx * gy * gz);
      }
      if (total > 0xffffffffUL) return fail(PYPDE_ERR_ARG, "ensemble: too many workgroups per launch");
      h[gr.map + (size_t)s * (gr.count + 1) + gr.count] = {(unsigned)total, 1, 1, 1};
      gr.blocks[s] = (unsigned)total;
    }
  }
  std::vector<Member> hm(nm);
  for (int k = 0; k < nm; k++) {
    pypde_ctx* c = e->m[order[k]];
    Member& M = hm[k];
    M.g = c->g;
    M.u = c->u;
    M.wh = c->wh;
    M.tr = c->tr;
    M.vol = c->vol;
    M.trd = c->trd;
    for (int d = 0; d < 2; d++) {
      M.ff[d] = c->ff[d];
      M.fb[d] = c->fb[d];
    }
    M.st = c->st;
    M.hot = c->hot;
  }
  if (!e->d_ms) HIPCHK(hipMalloc((void**)&e->d_ms, hm.size() * sizeof(Member)));
  if (e->d_map && e->n_map < h.size()) {
    (void)hipFree(e->d_map);
    e->d_map = nullptr;
  }
  if (!e->d_map) {
    HIPCHK(hipMalloc((void**)&e->d_map, h.size() * sizeof(BlkMap)));
    e->n_map = h.size();
  }
  hipStream_t s0 = e->m[0]->stream;
  HIPCHK(hipMemcpyAsync(e->d_ms, hm.data(), hm.size() * sizeof(Member), hipMemcpyHostToDevice, s0));
  HIPCHK(hipMemcpyAsync(e->d_map, h.data(), h.size() * sizeof(BlkMap), hipMemcpyHostToDevice, s0));
  HIPCHK(hipStreamSynchronize(s0));
  return 0;
}

// One timestep of one group of members, one launch per stage, on its stream.
static int enqueue_group_step(pypde_ensemble* e, const EnsGroup& gr) {
  pypde_ctx* c = e->m[0];
  const int nm = gr.count;
  Member* ms = e->d_ms + gr.first;
  double tol = c->opts.newton_tol;
  int maxit = c->opts.newton_maxit;
  auto go = [&](int kid, int stage, int bs, bool pred) {
    if (gr.blocks[stage] == 0) return 0;
    if (!c->fn[kid]) return fail(PYPDE_ERR_HIP, "kernel %s not loaded", kernel_names[kid]);
    BlkMap* map = e->d_map + gr.map + (size_t)stage * (nm + 1);
    std::vector<void*> a = {&ms, &map, (void*)&nm};
    if (pred) {
      a.push_back(&tol);
      a.push_back(&maxit);
    }
    hipError_t err = hipModuleLaunchKernel(c->fn[kid], gr.blocks[stage], 1, 1, bs, 1, 1, 0,
                                           gr.stream, a.data(), nullptr);
    if (err != hipSuccess)
      return fail(PYPDE_ERR_HIP, "launch %s: %s", kernel_names[kid], hipGetErrorString(err));
    return 0;
  };
  int rc;
  if ((rc = go(B_WENO_KID[gr.weno], BS_WENO, 256, false))) return rc;
  if ((rc = go(K_B_DT, BS_DT, 256, false))) return rc;
  if (gr.pred >= 2) {
    if ((rc = go(gr.pred == 3 ? K_B_PRED_COLD3 : K_B_PRED_COLD, BS_PRED, 64, true))) return rc;
    if ((rc = go(K_B_PRED_HOT, BS_PRED, 64, true))) return rc;
  } else if ((rc = go(gr.pred == 1 ? K_B_PRED_STATIC2 : K_B_PRED_STATIC, BS_PRED, 64, true))) {
    return rc;
  }
  for (int d = 0; d < c->D; d++)
    if ((rc = go(B_FLUX_KID[d][gr.flux], BS_FLUX0 + d, FLUX_BS[gr.flux], false))) return rc;
  return go(B_UPD_KID[gr.update], BS_UPDATE, 256, false);
}

// One timestep of every group captured into a graph: a fork event on the
// first group's stream, every group on its own stream, joined back.
static int capture_batch(pypde_ensemble* e, hipGraphExec_t* exec) {
  const size_t ng = e->groups.size();
  // groups in the order of their first members, i.e. the caller's order (on
  // the 48-grid benchmark ensemble, 1.5 % faster than the groups sorted by
  // size either way); the capture starts on the first group's stream and
  // every other group's stream joins it through the fork event
  std::vector<size_t> gorder(ng);
  for (size_t k = 0; k < ng; k++) gorder[k] = k;
  std::stable_sort(gorder.begin(), gorder.end(), [&](size_t a, size_t b) {
    return e->groups[a].lead < e->groups[b].lead;
  });
  hipStream_t s0 = e->groups[gorder[0]].stream;
  std::vector<hipEvent_t> evs(ng + 1, nullptr);
  for (auto& ev : evs) HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  int rc = 0;
  auto chk = [&](hipError_t err, const char* what) {
    if (!rc && err != hipSuccess)
      rc = fail(PYPDE_ERR_HIP, "graph capture, %s: %s", what, hipGetErrorString(err));
  };
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
  chk(hipEventRecord(evs[0], s0), "fork event");
  for (size_t k = 0; k < ng && !rc; k++) {
    const EnsGroup& gr = e->groups[gorder[k]];
    if (k > 0) chk(hipStreamWaitEvent(gr.stream, evs[0], 0), "fork wait");
    // a group of one runs the member's own single-grid step
    if (!rc)
      rc = gr.count == 1 ? enqueue_step(e->m[e->order[gr.first]], false, nullptr)
                         : enqueue_group_step(e, gr);
    if (!rc && k > 0) {
      chk(hipEventRecord(evs[k + 1], gr.stream), "join event");
      chk(hipStreamWaitEvent(s0, evs[k + 1], 0), "join wait");
    }
  }
  hipError_t err = hipStreamEndCapture(s0, &graph);
  if (!rc && err != hipSuccess) rc = fail(PYPDE_ERR_HIP, "graph capture: %s", hipGetErrorString(err));
  if (!rc && hipGraphInstantiate(exec, graph, nullptr, nullptr, 0) != hipSuccess)
    rc = fail(PYPDE_ERR_HIP, "graph instantiation failed");
  if (graph) (void)hipGraphDestroy(graph);
  for (auto ev : evs) (void)hipEventDestroy(ev);
  return rc;
}

pypde_ensemble* pypde_ensemble_create(int n, const char* src, bool useF, bool useB, bool useS,
                                      const int* nX, int ndim, const double* dX, double CFL,
                                      const int* boundaryTypes, bool STIFF, int FLUX, int N, int V,
                                      bool secondOrder, const pypde_options* opts, int* rc_out) {
  int dummy;
  int& rc = rc_out ? *rc_out : dummy;
  if (n < 1 || !nX || !dX || !boundaryTypes || ndim < 1 || ndim > 2) {
    rc = fail(PYPDE_ERR_ARG, "ensemble: n >= 1, nX, dX, boundaryTypes and ndim 1..2 required");
    return nullptr;
  }
  if (opts && opts->guess) {
    rc = fail(PYPDE_ERR_ARG, "ensemble: per-member guesses are not supported");
    return nullptr;
  }
  auto* e = new pypde_ensemble();
  for (int i = 0; i < n; i++) {
    int nx[2] = {nX[i * ndim], ndim > 1 ? nX[i * ndim + 1] : 1};
    double dx[2] = {dX[i * ndim], ndim > 1 ? dX[i * ndim + 1] : 1.0};
    int bc[2] = {boundaryTypes[0], ndim > 1 ? boundaryTypes[1] : 0};
    pypde_ctx* c = create_ctx(src, useF, useB, useS, nx, ndim, dx, CFL, bc, STIFF, FLUX, N, V,
                              secondOrder, opts, &rc, true);
    if (!c) {
      pypde_ensemble_destroy(e);
      return nullptr;
    }
    e->m.push_back(c);
  }
  rc = 0;
  return e;
}

int pypde_ensemble_set_state(pypde_ensemble* e, int i, const double* u) {
  if (!e || i < 0 || i >= (int)e->m.size()) return fail(PYPDE_ERR_ARG, "ensemble member %d", i);
  return pypde_set_state(e->m[i], u);
}

int pypde_ensemble_get_state(pypde_ensemble* e, int i, double* u) {
  if (!e || i < 0 || i >= (int)e->m.size()) return fail(PYPDE_ERR_ARG, "ensemble member %d", i);
  return pypde_get_state(e->m[i], u);
}

int pypde_ensemble_run(pypde_ensemble* e, double tf, long max_steps, pypde_stats* stats,
                       double* ms_loop) {
  if (!e) return fail(PYPDE_ERR_ARG, "null ensemble");
  pypde_ctx* c0 = e->m[0];
  HIPCHK(hipSetDevice(c0->device));
  for (auto c : e->m) {
    StepState s;
    memset(&s, 0, sizeof s);
    s.tf = tf;
    s.cfl = c->host_st.cfl;
    s.max_steps = max_steps;
    s.done = (0.0 < tf) ? 0 : 1;
    c->host_st = s;
    c->g.ndt = 0;
    c->ret = nullptr;
    HIPCHK(hipMemcpyAsync(c->st, &s, sizeof s, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  for (auto c : e->m)
    if (int r = autotune(c)) return r;
  // one launch per stage per group of members, unless the members lack the
  // batched kernels (mixed precision) or PYPDE_ENSEMBLE=streams asks for one
  // stream per member inside the graph; PYPDE_ENSEMBLE_CELLS sets the group
  // size cap (reconstructed cells; 0: one group)
  const char* mode = getenv("PYPDE_ENSEMBLE");
  const bool batched = c0->fn[K_B_DT] && !(mode && strcmp(mode, "streams") == 0);
  long cap = 1L << 18;
  if (const char* v = getenv("PYPDE_ENSEMBLE_CELLS")) cap = atol(v);
  auto t0 = std::chrono::steady_clock::now();
  hipGraphExec_t exec = nullptr;
  int rc = batched ? batch_setup(e, cap) : 0;
  if (!rc) rc = batched ? capture_batch(e, &exec) : capture_steps(e->m, &exec);
  int chunk = 4;
  long launched = 0;
  std::vector<long long> before(e->m.size());
  for (size_t i = 0; i < e->m.size(); i++) before[i] = e->m[i]->host_st.count;
  while (!rc) {
    if (max_steps > 0 && max_steps - launched < chunk) chunk = (int)std::max(1L, max_steps - launched);
    for (int i = 0; i < chunk && !rc; i++)
      if (hipGraphLaunch(exec, c0->stream) != hipSuccess) rc = fail(PYPDE_ERR_HIP, "graph launch failed");
    launched += chunk;
    if (rc) break;
    bool all = true;
    for (auto c : e->m)
      if (hipMemcpyAsync(&c->host_st, c->st, sizeof(StepState), hipMemcpyDeviceToHost, c0->stream) !=
          hipSuccess)
        rc = fail(PYPDE_ERR_HIP, "state readback failed");
    if (hipStreamSynchronize(c0->stream) != hipSuccess) rc = fail(PYPDE_ERR_HIP, "ensemble step failed");
    if (rc) break;
    // every member still running advances by one step per launch
    for (size_t i = 0; i < e->m.size() && !rc; i++) {
      const StepState& h = e->m[i]->host_st;
      if (!h.done && h.count == before[i])
        rc = fail(PYPDE_ERR_HIP, "ensemble member %zu made no progress in %d steps", i, chunk);
      before[i] = h.count;
      all = all && h.done;
    }
    if (all) break;
    if (chunk < 64) chunk *= 2;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  auto t1 = std::chrono::steady_clock::now();
  if (rc) return rc;
  if (ms_loop) *ms_loop = std::chrono::duration<double, std::milli>(t1 - t0).count();
  int nan = 0;
  for (size_t i = 0; i < e->m.size(); i++) {
    const StepState& h = e->m[i]->host_st;
    if (h.nan_step && !nan) nan = h.nan_step;
    if (stats) {
      pypde_stats* o = stats + i;
      o->steps = h.count;
      o->t = h.t;
      o->pred_iters = (long)h.pred_iters;
      o->pred_cells = (long)h.pred_cells;
      o->pred_maxit = (long)h.pred_maxit;
      o->pred_capped = (long)h.pred_capped;
      o->guess_fallbacks = (long)h.pred_fallbacks;
      o->nan_step = h.nan_step;
      o->ms_loop = std::chrono::duration<double, std::milli>(t1 - t0).count();
      for (int k = 0; k < 8; k++) o->ms_kernel[k] = 0.0;
      if (stats_has_nodes(stats)) {
        o->pred_node_full = (long)h.pred_nfull;
        o->pred_node_short = (long)h.pred_nshort;
        o->pred_iters_f32 = (long)h.pred_iters32;
      }
    }
  }
  return nan ? -nan : 0;
}

int pde_solver_hip(const char* src, bool useF, bool useB, bool useS, double* u, double tf,
                   int* nX, int ndim, double* dX, double CFL, int* boundaryTypes, bool STIFF,
                   int FLUX, int N, int V, int ndt, bool secondOrder, double* ret,
                   const pypde_options* opts, pypde_stats* stats) {
  if (!u || !ret || ndt < 1) return fail(PYPDE_ERR_ARG, "u, ret and ndt >= 1 are required");
  int rc = 0;
  pypde_ctx* c = pypde_create(src, useF, useB, useS, nX, ndim, dX, CFL, boundaryTypes, STIFF,
                              FLUX, N, V, secondOrder, opts, &rc);
  if (!c) return rc ? rc : PYPDE_ERR_HIP;
  if ((rc = pypde_set_state(c, u))) {
    pypde_destroy(c);
    return rc;
  }
  long ms = opts ? opts->max_steps : 0;
  rc = pypde_run(c, tf, ms, ndt, ret, stats);
  if (rc <= 0) {
    int r2 = pypde_get_state(c, u);
    if (r2) rc = r2;
  }
  pypde_destroy(c);
  return rc;
}

int pypde_set_guess_model(pypde_ctx* c, const pypde_mlp* m) {
  HIPCHK(hipSetDevice(c->device));
  if (!m) {
    c->use_model = false;
    memcpy(c->fn, c->fn_plain, sizeof c->fn);  // the module the context was created with
    return 0;
  }
  if (!c->spec.stiff) return fail(PYPDE_ERR_ARG, "a guess model needs the stiff (Newton) predictor");
  if (!c->mod_guess) {
    Spec s = c->spec;
    s.guess = true;
    std::vector<char> code;
    uint64_t key = 0;
    int rc = compile_code(s, &code, &key);
    if (rc) return rc;
    if ((rc = load_module(c, code, key, &c->mod_guess, c->fn_guess))) return rc;
  }
  if (!c->gbuf) HIPCHK(hipMalloc((void**)&c->gbuf, c->g.nwc * c->NT * c->V * sizeof(double)));
  // pack the weights once: BF16 fragments in the inference kernel's order
  // (its sizes follow ader_step.h: hidden width 128)
  const int n_in = c->ND * c->V + 1, n_out = c->NT * c->V, hid = 128;
  const long nfrag = (long)(hid / 32) * ((n_in + 7) / 8) + (long)(hid / 32) * (hid / 8) +
                     (long)((n_out + 31) / 32) * (hid / 8);
  const long naux = 2L * n_in + 2L * hid + 3L * n_out;
  if (!c->mlp_pack) HIPCHK(hipMalloc(&c->mlp_pack, nfrag * 64 * 8));
  if (!c->mlp_aux) HIPCHK(hipMalloc((void**)&c->mlp_aux, naux * sizeof(float)));
  c->model = {m->W1t, m->b1, m->W2t, m->b2, m->W3t, m->b3, m->in_mean, m->in_rstd, m->out_mean,
              m->out_std};
  memcpy(c->fn, c->fn_guess, sizeof c->fn);
  {
    MLPW mw = c->model;
    void* pk = c->mlp_pack;
    float* ax = c->mlp_aux;
    std::vector<void*> a = {&mw, &pk, &ax};
    int rc = launch(c, K_MLPPACK, dim3((unsigned)cdiv(std::max(nfrag * 64, naux), 256)), dim3(256), a);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  c->use_model = true;
  return 0;
}

int pypde_eig_batch(pypde_ctx* c, int which, const double* A, const double* x, long n,
                    double* out) {
  if (!c || !A || !out || n < 0 || (which != 0 && which != 2))
    return fail(PYPDE_ERR_ARG, "pypde_eig_batch: bad arguments");
  if (which == 2 && (c->spec.flux == 0 || !x))
    return fail(PYPDE_ERR_ARG, "pypde_eig_batch: abs_apply needs a Roe or Osher context and x");
  if (n == 0) return 0;
  HIPCHK(hipSetDevice(c->device));
  const long V = c->V, nout = which == 2 ? n * V : n;
  double *dA = nullptr, *dx = nullptr, *dout = nullptr;
  int rc = 0;
  auto done = [&](int r) {
    if (dA) (void)hipFree(dA);
    if (dx) (void)hipFree(dx);
    if (dout) (void)hipFree(dout);
    return r;
  };
  if (hipMalloc((void**)&dA, n * V * V * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&dout, nout * sizeof(double)) != hipSuccess ||
      (which == 2 && hipMalloc((void**)&dx, n * V * sizeof(double)) != hipSuccess))
    return done(fail(PYPDE_ERR_HIP, "pypde_eig_batch: hipMalloc failed"));
  if (hipMemcpyAsync(dA, A, n * V * V * sizeof(double), hipMemcpyHostToDevice, c->stream) !=
          hipSuccess ||
      (which == 2 && hipMemcpyAsync(dx, x, n * V * sizeof(double), hipMemcpyHostToDevice,
                                    c->stream) != hipSuccess))
    return done(fail(PYPDE_ERR_HIP, "pypde_eig_batch: copy in failed"));
  const double* a0 = dA;
  const double* x0 = dx;
  std::vector<void*> a = {&a0, &x0, &dout, &n, &which};
  if ((rc = launch(c, K_EIG, dim3((unsigned)cdiv(n, 256)), dim3(256), a))) return done(rc);
  if (hipMemcpyAsync(out, dout, nout * sizeof(double), hipMemcpyDeviceToHost, c->stream) !=
          hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return done(fail(PYPDE_ERR_HIP, "pypde_eig_batch: copy out failed"));
  return done(0);
}

int pypde_harvest_step(pypde_ctx* c, float* X, float* Y, long ldx, long col0, pypde_stats* stats) {
  HIPCHK(hipSetDevice(c->device));
  if (!c->qh) HIPCHK(hipMalloc((void**)&c->qh, c->g.nwc * c->NT * c->V * sizeof(double)));
  int rc = pypde_step(c, 1, 1e300, stats);
  if (rc) return rc;
  const double* w = c->wh;
  const double* q = c->qh;
  const StepState* st = c->st;
  long nwc = c->g.nwc;
  std::vector<void*> a = {&w, &q, &st, &X, &Y, &ldx, &col0, &nwc};
  rc = launch(c, K_HARVEST, dim3((unsigned)cdiv(nwc, 256)), dim3(256), a);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  (void)hipFree(c->qh);
  c->qh = nullptr;
  return 0;
}

int pypde_last_guess(pypde_ctx* c, double* out) {
  HIPCHK(hipSetDevice(c->device));
  if (!c->gbuf) return fail(PYPDE_ERR_ARG, "no guess model attached");
  size_t n = c->g.nwc * c->NT * c->V;
  std::vector<double> h(n);
  HIPCHK(hipMemcpy(h.data(), c->gbuf, n * sizeof(double), hipMemcpyDeviceToHost));
  size_t per = (size_t)c->NT * c->V;
  for (long cc = 0; cc < c->g.nwc; cc++)
    for (size_t p = 0; p < per; p++) out[cc * per + p] = h[p * c->g.nwc + cc];
  return 0;
}

int pypde_get_tuning(pypde_ctx* c, int* choices, double* ms) {
  if (!c) return fail(PYPDE_ERR_ARG, "null context");
  if (choices) {
    choices[0] = c->tune.weno;
    choices[1] = c->tune.flux;
    choices[2] = c->tune.update;
    choices[3] = c->pred_split ? (c->pred_cold3 ? 3 : 2) : (c->pred_static ? 0 : 1);
    choices[4] = c->tune.source;
  }
  if (ms)
    for (int i = 0; i < 4; i++) ms[i] = c->tune.ms[i];
  return 0;
}

int pypde_geometry(pypde_ctx* c, long* nwc, int* n_in, int* n_out) {
  if (nwc) *nwc = c->g.nwc;
  if (n_in) *n_in = c->ND * c->V + 1;
  if (n_out) *n_out = c->NT * c->V;
  return 0;
}

// ---------------------------------------------------------------- training
static hipModule_t g_mlp_mod[16];
static hipFunction_t g_mlp_fn[16][6];
static const char* mlp_names[6] = {"k_gemm_bf16", "k_bias_act", "k_act_grad", "k_rowsum", "k_mse", "k_adam"};

static int mlp_fn(int k, hipFunction_t* f) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 16) return fail(PYPDE_ERR_ARG, "device %d", dev);
  static std::mutex mu;  // not g_mu: rtc_compile takes that one
  std::lock_guard<std::mutex> lk(mu);
  if (!g_mlp_mod[dev]) {
    std::vector<char> code;
    int rc = rtc_compile(MLP_SRC, "pypde_mlp.hip", &code);
    if (rc) return rc;
    HIPCHK(hipModuleLoadData(&g_mlp_mod[dev], code.data()));
    for (int i = 0; i < 6; i++) HIPCHK(hipModuleGetFunction(&g_mlp_fn[dev][i], g_mlp_mod[dev], mlp_names[i]));
  }
  *f = g_mlp_fn[dev][k];
  return 0;
}

static int mlp_launch(int k, dim3 grid, dim3 block, void** args, void* stream) {
  hipFunction_t f;
  int rc = mlp_fn(k, &f);
  if (rc) return rc;
  HIPCHK(hipModuleLaunchKernel(f, grid.x, grid.y, grid.z, block.x, block.y, block.z, 0,
                               (hipStream_t)stream, args, nullptr));
  return 0;
}

void* pypde_dev_alloc(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
    fail(PYPDE_ERR_HIP, "hipMalloc(%zu) failed", bytes);
    return nullptr;
  }
  return p;
}

int pypde_dev_free(void* p) {
  if (p) HIPCHK(hipFree(p));
  return 0;
}

int pypde_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                              : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  HIPCHK(hipMemcpy(dst, src, bytes, k));
  return 0;
}

int pypde_dev_memset(void* p, int value, size_t bytes) {
  HIPCHK(hipMemset(p, value, bytes));
  return 0;
}

int pypde_dev_sync(void) {
  HIPCHK(hipDeviceSynchronize());
  return 0;
}

int pypde_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int pypde_mlp_compile(void) {
  // compile (or fetch) the training kernels without a GPU
  std::vector<char> code;
  return rtc_compile(MLP_SRC, "pypde_mlp.hip", &code);
}

int pypde_mlp_gemm(const float* A, int lda, int transA, const float* B, int ldb, int transB,
                   float* C, int ldc, int M, int N, int K, float alpha, float beta, void* stream) {
  if (M <= 0 || N <= 0 || K <= 0) return fail(PYPDE_ERR_ARG, "gemm sizes");
  void* a[] = {&A, &lda, &transA, &B, &ldb, &transB, &C, &ldc, &M, &N, &K, &alpha, &beta};
  return mlp_launch(0, dim3((unsigned)cdiv(N, 64), (unsigned)cdiv(M, 64)), dim3(256), a, stream);
}

int pypde_mlp_bias_act(float* Z, const float* b, float* H, int rows, int cols, int act, void* stream) {
  void* a[] = {&Z, &b, &H, &rows, &cols, &act};
  return mlp_launch(1, dim3((unsigned)cdiv((long)rows * cols, 256)), dim3(256), a, stream);
}

int pypde_mlp_act_grad(float* dZ, const float* dH, const float* Z, long n, void* stream) {
  void* a[] = {&dZ, &dH, &Z, &n};
  return mlp_launch(2, dim3((unsigned)cdiv(n, 256)), dim3(256), a, stream);
}

int pypde_mlp_rowsum(const float* dZ, float* db, int rows, int cols, void* stream) {
  void* a[] = {&dZ, &db, &rows, &cols};
  return mlp_launch(3, dim3((unsigned)rows), dim3(256), a, stream);
}

int pypde_mlp_mse(const float* Y, const float* T, float* dY, float* loss, int rows, int cols,
                  void* stream) {
  void* a[] = {&Y, &T, &dY, &loss, &rows, &cols};
  return mlp_launch(4, dim3((unsigned)cdiv((long)rows * cols, 256)), dim3(256), a, stream);
}

int pypde_mlp_adam(float* w, const float* g, float* m, float* v, long n, float lr, float b1,
                   float b2, float eps, int step, void* stream) {
  float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  void* a[] = {&w, &g, &m, &v, &n, &lr, &b1, &b2, &eps, &c1, &c2};
  return mlp_launch(5, dim3((unsigned)cdiv(n, 256)), dim3(256), a, stream);
}

int weno_solver(double* ret, double* u, int* nX, int ndim, int N, int V) {
  Spec s;
  make_spec(&s, "", false, false, false, ndim, N, V, false, nullptr);
  int rc = check_spec(s, false, 0);
  if (rc) return rc;
  long nin = 1, nout = 1;
  int no[2] = {1, 1};
  for (int d = 0; d < ndim; d++) {
    no[d] = nX[d] - 2 * (N - 1);
    if (no[d] < 1) return fail(PYPDE_ERR_ARG, "nX[%d]=%d too small for N=%d", d, nX[d], N);
    nin *= nX[d];
    nout *= no[d];
  }
  std::vector<char> code;
  if ((rc = compile_code(s, &code))) return rc;
  HIPCHK(hipSetDevice(0));
  // the module and device buffers are released on every return path
  struct Release {
    hipModule_t mod = nullptr;
    double *du = nullptr, *dus = nullptr, *dw = nullptr;
    ~Release() {
      for (double* p : {du, dus, dw})
        if (p) (void)hipFree(p);
      if (mod) (void)hipModuleUnload(mod);
    }
  } rel;
  HIPCHK(hipModuleLoadData(&rel.mod, code.data()));
  hipModule_t mod = rel.mod;
  hipFunction_t fw, fa;
  HIPCHK(hipModuleGetFunction(&fw, mod, "k_weno"));
  HIPCHK(hipModuleGetFunction(&fa, mod, "k_aos_to_soa"));
  int ND = 1;
  for (int d = 0; d < ndim; d++) ND *= N;
  double *&du = rel.du, *&dus = rel.dus, *&dw = rel.dw;
  HIPCHK(hipMalloc((void**)&du, nin * V * sizeof(double)));
  HIPCHK(hipMalloc((void**)&dus, nin * V * sizeof(double)));
  HIPCHK(hipMalloc((void**)&dw, nout * ND * V * sizeof(double)));
  HIPCHK(hipMemcpy(du, u, nin * V * sizeof(double), hipMemcpyHostToDevice));
  {
    void* a[] = {&du, &dus, &nin};
    HIPCHK(hipModuleLaunchKernel(fa, (unsigned)cdiv(nin * V, 256), 1, 1, 256, 1, 1, 0, nullptr, a, nullptr));
  }
  Geom g;
  memset(&g, 0, sizeof g);
  const StepState* st = nullptr;
  int ghost = 0, n0 = nX[0], n1 = ndim > 1 ? nX[1] : 1;
  {
    void* a[] = {&dus, &dw, &st, &g, &ghost, &n0, &n1, &nout};
    unsigned gx = ndim == 1 ? (unsigned)cdiv(no[0], 256) : (unsigned)cdiv(no[1], 32);
    unsigned gy = ndim == 1 ? 1 : (unsigned)cdiv(no[0], 8);
    unsigned gz = ndim == 1 ? 1 : (unsigned)V;  // 2D: one variable per workgroup
    HIPCHK(hipModuleLaunchKernel(fw, gx, gy, gz, 256, 1, 1, 0, nullptr, a, nullptr));
  }
  std::vector<double> h(nout * ND * V);
  HIPCHK(hipMemcpy(h.data(), dw, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  // planes [node*V+v][cell] -> reference layout [cell][node][v]
  for (long cc = 0; cc < nout; cc++)
    for (int p = 0; p < ND * V; p++) ret[cc * ND * V + p] = h[(size_t)p * nout + cc];
  return 0;
}

// ------------------------------------------------ allocation-free WENO
// Device scratch of weno_solver_noalloc: one array per GPU, held until the
// process ends.  Host code never passes `scratch` itself to a kernel (in host
// code the name stands for a host-side shadow); it uses the per-device
// address that hipGetSymbolAddress returns.
}  // extern "C"

constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The scratch array's device address on `dev`, looked up once per device
// with that device current.
int scratch_address(int dev, double** out) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(PYPDE_ERR_ARG, "device %d", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return fail(PYPDE_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

bool is_device_pointer(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}
}  // namespace

extern "C" {

int weno_solver_noalloc(pypde_ctx* c, double* ret, const double* u, const int* nX, void* stream) {
  if (!c || !ret || !u || !nX) return fail(PYPDE_ERR_ARG, "weno_solver_noalloc: null argument");
  const int ndim = c->D, N = c->N, V = c->V;
  long nin = 1, nout = 1;
  int no[2] = {1, 1};
  for (int d = 0; d < ndim; d++) {
    no[d] = nX[d] - 2 * (N - 1);
    if (nX[d] < 1 || no[d] < 1)
      return fail(PYPDE_ERR_ARG, "weno_solver_noalloc: nX[%d]=%d too small for N=%d", d, nX[d], N);
    nin *= nX[d];
    nout *= no[d];
  }
  const size_t in_doubles = ((size_t)nin * V + 31) / 32 * 32;  // 256-byte aligned split
  const size_t need = (in_doubles + (size_t)nout * c->ND * V) * sizeof(double);
  if (need > NOALLOC_SCRATCH_BYTES)
    return fail(PYPDE_ERR_ARG,
                "weno_solver_noalloc: the grid needs %zu bytes of scratch (input %ld x %d "
                "doubles, output %ld x %d doubles); the limit is %zu bytes",
                need, nin, V, nout, c->ND * V, (size_t)NOALLOC_SCRATCH_BYTES);
  HIPCHK(hipSetDevice(c->device));
  if (!is_device_pointer(u) || !is_device_pointer(ret))
    return fail(PYPDE_ERR_ARG, "weno_solver_noalloc: u and ret must be device memory");
  double* sp = nullptr;
  int rc = scratch_address(c->device, &sp);
  if (rc) return rc;
  double* s_in = sp;               // u as planes [v][cell]
  double* s_out = sp + in_doubles;  // reconstruction planes [node * V + v][cell]
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  auto go = [&](int k, unsigned gx, unsigned gy, unsigned gz, void** a) -> int {
    if (!c->fn[k]) return fail(PYPDE_ERR_HIP, "kernel %s not loaded", kernel_names[k]);
    hipError_t e = hipModuleLaunchKernel(c->fn[k], gx, gy, gz, 256, 1, 1, 0, s, a, nullptr);
    if (e != hipSuccess)
      return fail(PYPDE_ERR_HIP, "launch %s: %s", kernel_names[k], hipGetErrorString(e));
    return 0;
  };
  {
    void* a[] = {&u, &s_in, &nin};
    if ((rc = go(K_A2S, (unsigned)cdiv(nin * V, 256), 1, 1, a))) return rc;
  }
  {
    Geom g;
    memset(&g, 0, sizeof g);
    const StepState* st = nullptr;
    int ghost = 0, n0 = nX[0], n1 = ndim > 1 ? nX[1] : 1;
    void* a[] = {&s_in, &s_out, &st, &g, &ghost, &n0, &n1, &nout};
    unsigned gx = ndim == 1 ? (unsigned)cdiv(no[0], 256) : (unsigned)cdiv(no[1], 32);
    unsigned gy = ndim == 1 ? 1 : (unsigned)cdiv(no[0], 8);
    unsigned gz = ndim == 1 ? 1 : (unsigned)V;
    if ((rc = go(K_WENO, gx, gy, gz, a))) return rc;
  }
  {
    void* a[] = {&s_out, &ret, &nout};
    if ((rc = go(K_P2R, (unsigned)cdiv(nout * c->ND * V, 256), 1, 1, a))) return rc;
  }
  if (!stream) HIPCHK(hipStreamSynchronize(s));
  return 0;
}

}  // extern "C"
