// Host side of libpfc_hip.so: instance sets ("engines"), launch shapes, the
// autotuner and the C ABI.  A pfc_plan is an engine of one slot; a pfc_batch
// is an engine of many same-shape slots stepped by the same launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <cstdlib>
#include <vector>

#include "../../include/pfc_hip.h"
#include "pfc_fast.hpp"

using namespace pfc;

namespace {

thread_local std::string g_err;

struct Failure {
  int code;
};

[[noreturn]] void fail(int code, const std::string& msg) {
  g_err = msg;
  throw Failure{code};
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(PFC_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return PFC_OK;
  } catch (const Failure& x) {
    return x.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return PFC_ENOMEM;
  } catch (...) {
    g_err = "unexpected C++ exception";
    return PFC_EHIP;
  }
}

// ---------------------------------------------------------------------------
// radix schedules and twiddles
// ---------------------------------------------------------------------------
bool factor(int L, Radices& pl) {
  pl.L = L;
  pl.np = 0;
  int m = L;
  auto push = [&](int r) {
    if (pl.np >= kMaxPass) return false;
    pl.r[pl.np++] = r;
    m /= r;
    return true;
  };
  while (m % 8 == 0)
    if (!push(8)) return false;
  while (m % 4 == 0)
    if (!push(4)) return false;
  while (m % 2 == 0)
    if (!push(2)) return false;
  for (int r : {3, 5, 7, 11, 13})
    while (m % r == 0)
      if (!push(r)) return false;
  return m == 1;
}

// exp(-2 pi i m / n), m = 0..n-1, computed in long double and rounded once;
// exact at the quarter points (1, -i, -1, i), where the real-FFT pack/unpack
// relies on exact zeros to keep the DC and Nyquist terms real.
std::vector<V2<double>> twiddles(int n) {
  std::vector<V2<double>> t(n);
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < n; ++m) {
    if ((4LL * m) % n == 0) {
      const int q = (int)((4LL * m) / n);
      const double c[4] = {1.0, 0.0, -1.0, 0.0}, s[4] = {0.0, -1.0, 0.0, 1.0};
      t[m] = V2<double>{c[q], s[q]};
      continue;
    }
    long double a = twopi * (long double)m / (long double)n;
    t[m] = V2<double>{(double)cosl(a), -(double)sinl(a)};
  }
  return t;
}

std::vector<V2<float>> to_f32(const std::vector<V2<double>>& v) {
  std::vector<V2<float>> o(v.size());
  for (size_t i = 0; i < v.size(); ++i) o[i] = V2<float>{(float)v[i].x, (float)v[i].y};
  return o;
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t bytes) {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (bytes == 0) return;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      fail(PFC_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    n = bytes;
  }
  template <typename X>
  void upload(const std::vector<X>& v) {
    alloc(v.size() * sizeof(X));
    hip_check(hipMemcpy(p, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice), "upload");
  }
};

int lds_limit() {
  static int lim = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 65536;
    return v;
  }();
  return lim;
}

// Launch shape of the row kernels: rows per block within the LDS budget,
// keeping at least ~512 blocks when the grid allows.
RowGeom row_geom(int Nx, int Ny, size_t csize) {
  RowGeom g{};
  g.Nx = Nx;
  g.M = Nx / 2;
  g.Nc = Nx / 2 + 1;
  g.Ny = Ny;
  g.lay = row_major(g.Nc);
  g.ld = g.Nc;
  if (!factor(g.M, g.plM)) fail(PFC_EUNSUPPORTED, "Nx/2 = " + std::to_string(g.M) + " is not 2,3,5,7,11,13-smooth");
  const size_t per_row = 2 * (size_t)g.ld * csize;
  int nl = (int)std::max<size_t>(1, (size_t)lds_limit() / per_row);
  if (per_row > (size_t)lds_limit()) fail(PFC_EUNSUPPORTED, "Nx too large for one row in LDS");
  nl = std::min(nl, 8);
  while (nl > 1 && (Ny + nl - 1) / nl < 512) --nl;
  g.nl = nl;
  return g;
}

ColGeom col_geom(int Nx, int Ny, size_t csize) {
  ColGeom g{};
  g.Nc = Nx / 2 + 1;
  g.Ny = Ny;
  g.lay = row_major(g.Nc);
  if (!factor(Ny, g.plY)) fail(PFC_EUNSUPPORTED, "Ny = " + std::to_string(Ny) + " is not 2,3,5,7,11,13-smooth");
  const int maxc = (kColRegs * 256) / Ny;
  if (maxc < 1) fail(PFC_EUNSUPPORTED, "Ny too large for the column kernel");
  int nc = std::min(maxc, 8);
  auto fits = [&](int c, int ld) { return 2 * (size_t)c * ld * csize <= (size_t)lds_limit(); };
  while (nc > 1 && !fits(nc, Ny + 1) && !fits(nc, Ny)) --nc;
  if (!fits(nc, Ny) && !fits(nc, Ny + 1)) fail(PFC_EUNSUPPORTED, "Ny too large for one column in LDS");
  while (nc > 1 && (g.Nc + nc - 1) / nc < 256) --nc;
  g.nc = nc;
  g.ld = fits(nc, Ny + 1) ? Ny + 1 : Ny;
  return g;
}

// Storage order of an engine's state arrays: 128-byte tiles of 2 kx x 4 rows
// (FP64) or 4 kx x 4 rows (FP32); row-major when Ny is not a multiple of 4.
Lay state_layout(int Nx, int Ny, int precision) {
  const int Nc = Nx / 2 + 1;
  const int ltx = precision == 64 ? 1 : 2, lty = 2;
  if (Ny % (1 << lty)) return row_major(Nc);
  return Lay{ltx, lty, (Nc + (1 << ltx) - 1) >> ltx};
}

size_t row_lds(const RowGeom& g, size_t csize) { return 2 * (size_t)g.nl * g.ld * csize; }
size_t col_lds(const ColGeom& g, size_t csize) { return 2 * (size_t)g.nc * g.ld * csize; }
int row_blocks(const RowGeom& g) { return (g.Ny + g.nl - 1) / g.nl; }
int col_blocks(const ColGeom& g) { return (g.Nc + g.nc - 1) / g.nc; }

// ---------------------------------------------------------------------------
// fast (compile-time schedule) kernels: variants, launch shapes, launchers
// ---------------------------------------------------------------------------
std::vector<SchedInfo> schedules_for(int L) {
  std::vector<SchedInfo> v;
#define PFC_X(LEN, VAR, PP, ...) \
  if (L == LEN) v.push_back(SchedInfo{LEN, VAR, PP});
  PFC_SCHEDULES(PFC_X)
#undef PFC_X
  return v;
}

struct FastShape {
  bool on = false;
  bool nt = true;   // non-temporal output stores (else ordinary)
  bool cap = false; // the kernel form held to kCapWaves waves per SIMD (FP32, <= 8 points per lane)
  int var = 0;      // schedule variant
  // the launchers' variant argument: the variant, with kCapFlag for the capped form
  int code() const { return var | (cap ? kCapFlag : 0); }
  int P = 0;        // points per lane
  int tpl = 0;      // lanes per line
  int lines = 0;    // lines per block
  int blocks = 0;
  size_t lds = 0;
};

// Row kernel shape: LDS = lines * (M + 1), padded.
FastShape row_shape(int Nx, int Ny, size_t csize, const SchedInfo& si, int lines) {
  FastShape f;
  const int M = Nx / 2;
  f.var = si.variant;
  f.P = si.P;
  f.tpl = M / si.P;
  f.lines = lines;
  f.blocks = (Ny + lines - 1) / lines;
  f.lds = (size_t)lines * row_ld(M) * csize;
  f.on = f.tpl * lines <= 256 && f.tpl * lines >= 64 && f.lds <= (size_t)lds_limit();
  return f;
}

// Column kernel shape: LDS = lines * Ny, padded.
FastShape col_shape(int Nx, int Ny, size_t csize, const SchedInfo& si, int lines) {
  FastShape f;
  const int Nc = Nx / 2 + 1;
  f.var = si.variant;
  f.P = si.P;
  f.tpl = Ny / si.P;
  f.lines = lines;
  f.blocks = (Nc + lines - 1) / lines;
  f.lds = (size_t)lines * col_ld(Ny) * csize;
  f.on = f.tpl * lines <= 256 && f.tpl * lines >= 64 && f.lds <= (size_t)lds_limit();
  return f;
}

// Candidate (variant, lines) shapes; the default is variant 0 with enough
// blocks to cover the 256 CUs.
// FP32 schedules of up to 8 points per lane also come held to kCapWaves
// waves per SIMD (see kCapWaves): both forms are candidates.
void add_shape(std::vector<FastShape>& out, FastShape f, size_t csize) {
  if (!f.on) return;
  out.push_back(f);
  if (csize == sizeof(V2<float>) && f.P <= 8) {
    f.cap = true;
    out.push_back(f);
  }
}

std::vector<FastShape> row_candidates(int Nx, int Ny, size_t csize) {
  std::vector<FastShape> out;
  for (const SchedInfo& si : schedules_for(Nx / 2))
    for (int w = 1; w <= 256; w *= 2) add_shape(out, row_shape(Nx, Ny, csize, si, w), csize);
  return out;
}

std::vector<FastShape> col_candidates(int Nx, int Ny, size_t csize) {
  std::vector<FastShape> out;
  for (const SchedInfo& si : schedules_for(Ny))
    for (int c = 1; c <= 256; c *= 2) add_shape(out, col_shape(Nx, Ny, csize, si, c), csize);
  return out;
}

FastShape pick_default(const std::vector<FastShape>& cands, int target_blocks) {
  FastShape best;
  for (const FastShape& f : cands) {
    if (f.var != 0 || f.cap) continue;
    // most lines per block that still gives >= target_blocks blocks, else the fewest lines
    if (!best.on) {
      best = f;
      continue;
    }
    if (f.blocks >= target_blocks && f.lines > best.lines) best = f;
  }
  return best;
}

inline bool row_range_launch(int M, int var, dim3 g, dim3 b, size_t lds, hipStream_t st, const V2<double>* S, int Ny,
                             const Lay& lay, const V2<double>* tw, const Slots& sl, RangePart* part) {
  return row_range_launch_f64(M, var, g, b, lds, st, S, Ny, lay, tw, sl, part);
}
inline bool row_range_launch(int M, int var, dim3 g, dim3 b, size_t lds, hipStream_t st, const V2<float>* S, int Ny,
                             const Lay& lay, const V2<double>* tw, const Slots& sl, RangePart* part) {
  return row_range_launch_f32(M, var, g, b, lds, st, S, Ny, lay, tw, sl, part);
}

template <typename T> struct FastLaunch {
  static void row(int M, const FastShape& f, int ny_inst, hipStream_t st, const V2<T>* S, V2<T>* Tout, int Ny,
                  const Lay& lay, const Nonlin& q, const V2<double>* tw, const Slots& sl, RangePart* rpart = nullptr,
                  const RangeCheck& chk = RangeCheck{}) {
    if (!row_fast_launch<T>(M, f.code(), f.nt, dim3(f.blocks, ny_inst), dim3(f.lines * f.tpl), f.lds, st, S, Tout, Ny,
                            lay, q, tw, sl, rpart, chk))
      fail(PFC_EUNSUPPORTED, "no fast row schedule");
  }
  static void col(int Ny, int mode, const FastShape& f, int ny_inst, hipStream_t st, V2<T>* S, const V2<T>* Tin,
                  V2<T>* D, int Nc, const Lay& lay, const ColTables& tb, const V2<double>* tw, const Slots& sl) {
    if (!col_fast_launch<T>(Ny, f.code(), mode, f.nt, dim3(f.blocks, ny_inst), dim3(f.lines * f.tpl), f.lds, st, S, Tin,
                            D, Nc, lay, tb, tw, sl))
      fail(PFC_EUNSUPPORTED, "no fast column schedule");
  }
};

// ---------------------------------------------------------------------------
// FP64 k-grids (the reference's host arithmetic, restated)
// ---------------------------------------------------------------------------
// kx = 2 pi * rfftfreq(Nx, dx), ky = 2 pi * fftfreq(Ny, dy)  (geometry_2d.py:83-84)
std::vector<double> rfft_k(int n, double d) {
  const double val = 1.0 / (n * d);
  std::vector<double> k(n / 2 + 1);
  for (int i = 0; i <= n / 2; ++i) k[i] = 6.283185307179586 * ((double)i * val);
  return k;
}
std::vector<double> fft_k(int n, double d) {
  const double val = 1.0 / (n * d);
  std::vector<double> k(n);
  const int N = (n - 1) / 2 + 1;
  for (int i = 0; i < n; ++i) {
    const int f = i < N ? i : i - n;
    k[i] = 6.283185307179586 * ((double)f * val);
  }
  return k;
}
std::vector<double> k_squared(const std::vector<double>& k) {
  std::vector<double> k2(k.size());
  for (size_t i = 0; i < k.size(); ++i) k2[i] = k[i] * k[i];
  return k2;
}

void validate(const pfc_params* p) {
  if (!p) fail(PFC_EINVAL, "params is NULL");
  if (p->model != PFC_MODEL_STD && p->model != PFC_MODEL_LOG) fail(PFC_EINVAL, "model must be PFC_MODEL_STD or PFC_MODEL_LOG");
  if (p->Nx < 4 || p->Ny < 2 || (p->Nx & 1)) fail(PFC_EINVAL, "Nx must be even and >= 4, Ny >= 2");
  if (!(p->Lx > 0) || !(p->Ly > 0)) fail(PFC_EINVAL, "Lx and Ly must be positive");
  if (p->precision != 64 && p->precision != 32) fail(PFC_EINVAL, "precision must be 64 or 32");
  if (p->scheme != PFC_SCHEME_ETD1 && p->scheme != PFC_SCHEME_ETDRK2) fail(PFC_EINVAL, "scheme must be ETD1 or ETDRK2");
  if (!std::isfinite(p->dt) || !(p->dt > 0)) fail(PFC_EINVAL, "dt must be positive");
  if (p->model == PFC_MODEL_LOG && p->clamp_log && !(p->minPhi > 0)) fail(PFC_EINVAL, "minPhi must be positive");
}

// ---------------------------------------------------------------------------
// Engine: `cap` same-shape instances (slots).  Per slot: state S~ (row
// spectrum), T (row spectrum of N), D (ETDRK2 carry), ETD tables, epsilon and
// M, the k = 0 pin and a status word.  Steps act on the active slot list.
// ---------------------------------------------------------------------------
struct Engine {
  pfc_params p{};
  int cap = 0, Nc = 0;
  size_t csize = 0, nc = 0;  // bytes per complex; complex elements per slot array
  size_t ntab = 0;           // doubles per slot ETD table: Nc * (Ny/2 + 1)
  Lay lay{};  // storage order of S~, T (and the diagnostic scratch)
  RowGeom rg{}, rg64{};
  ColGeom cg{}, cg64{};
  FastShape frow{}, fcol{};
  ResGeom res{};
  bool resident = false;  // small ETD1 instances: k_resident_etd1, all steps in LDS
  bool dense = false;     // ... the smallest of them: k_resident_dense
  bool tuned = false, batched = false;
  DevBuf S, Tb, D, E1, Q1, Q2;       // [cap][...]
  DevBuf E1f, Q1f, Q2f;              // FP32 engines: the tables' FP32 copies (ColTables::E1f)
  DevBuf twN, twY, twN64, twY64, kx2, ky2, dts;  // kx2 [cap][Nc], ky2 [cap][Ny], dts [cap]
  DevBuf part, out, pin, status, eps, Ms, active, diag, dgi, dgo;  // dgi/dgo: diagnostics entries / results
  DevBuf cnt;         // per-slot step counts (pfc_batch_step_counts)
  DevBuf track;       // tracker series (engine_run_tracked)
  // stabilising shift managed on the device (pfc_set_auto_shift): per slot
  // ShiftState and a table-rebuild flag, checked before every step
  bool auto_shift = false;
  ShiftParams shp{};
  DevBuf shst, rebuildf;
  // in-step decision (col_refit): [cap][2] breach flags and the parity of the
  // next step's word
  DevBuf brk;
  int par = 0;
  DevBuf rpart;  // k_row_range: per-tile (min, max)
  DevBuf vhat, extc;  // external potential: rfft2(V) and its source term, [cap][Ny][Nc] complex FP64
  std::vector<char> h_ext;  // slots with a potential
  std::vector<double> h_eps, h_M, h_Lx, h_Ly, h_dt;
  std::vector<int> h_active;
  std::vector<char> field_set;
  double* host_out = nullptr;   // pinned [4 cap]
  int* host_status = nullptr;   // pinned [cap]
  // step-loop graphs: kGraphSteps steps of the row/column launches, captured
  // per launch configuration (see engine_step)
  std::map<std::array<long long, 11>, hipGraphExec_t> graphs;
  hipStream_t cap_st = nullptr;  // private stream the graphs are captured on
  long long graph_replays = 0, graph_captures = 0;
  void drop_graphs() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
  }
  ~Engine() {
    drop_graphs();
    if (cap_st) (void)hipStreamDestroy(cap_st);
    if (host_out) (void)hipHostFree(host_out);
    if (host_status) (void)hipHostFree(host_status);
  }
  bool f64() const { return p.precision == 64; }
  int nact() const { return (int)h_active.size(); }
  bool has_ext() const { return extc.p != nullptr; }
  ColTables tables() const {
    ColTables t{E1.p, Q1.p, Q2.p, table_rows(p.Ny), (const V2<double>*)extc.p, (long long)Nc * p.Ny};
    t.E1f = (float*)E1f.p;
    t.Q1f = (float*)Q1f.p;
    t.Q2f = (float*)Q2f.p;
    t.tp.inv_ny = 1.0 / p.Ny;  // E1 = (1 + E1f) / Ny (the refit path sets the whole of tp)
    return t;
  }
  int pstride() const { return row_blocks(rg64) + col_blocks(cg64); }
  // ETD1 under the device-managed shift folds -M phi into the E1 table: the
  // row pass computes N without it, so the range check can follow the step's
  // own row pass (with the fast row shape, whose blocks are whole wavefronts,
  // the row kernel writes the range partials itself)
  bool fold_shift() const { return auto_shift && p.scheme == PFC_SCHEME_ETD1 && !resident; }
  bool fused_range() const { return fold_shift() && frow.on && (frow.lines * frow.tpl) % 64 == 0; }
  // the fused step decides inside its two launches: the row kernel flags a
  // slot whose tile left the fitted window, the fast column kernel refits it
  // and rebuilds its own columns' tables (no decision or table launch)
  bool col_refit() const {
    return fused_range() && fcol.on && !has_ext() && (fcol.lines * fcol.tpl) % 64 == 0;
  }
  Slots slots_active() const {
    return Slots{(const int*)active.p, (long long)nc, (long long)ntab, p.pin_mean ? (const double*)pin.p : nullptr,
                 (int*)status.p, (const double*)Ms.p};
  }
  // a single slot addressed directly (list = null, blockIdx.y = 0 is that slot)
  Slots slots_one(int slot) const {
    return Slots{nullptr, 0, 0, p.pin_mean ? (const double*)pin.p + slot : nullptr, (int*)status.p + slot,
                 (const double*)Ms.p + slot};
  }
  void check_slot(int slot) const {
    if (slot < 0 || slot >= cap) fail(PFC_EINVAL, "slot out of range");
  }
};

// The resident stepper holds four Ny x (Nx/2 + 1) complex arrays, the
// twiddles and the instance's ETD tables in LDS; it is used for ETD1 when
// they fit in 64 KiB (e.g. up to 28 x 28 in FP64), where the two-kernel step
// is launch- and latency-bound.  PFC_HIP_RESIDENT=0 turns it off.
constexpr size_t kResidentLds = 64 * 1024;
size_t resident_lds(int Nx, int Ny, size_t csize) {
  const int Nc = Nx / 2 + 1;
  return resident_table_bytes(Nc, table_rows(Ny)) + ((size_t)Nx + Ny + 4 * (size_t)Ny * Nc) * csize;
}
bool resident_fits(const pfc_params& p) {
  const char* env = std::getenv("PFC_HIP_RESIDENT");
  if (env && std::string(env) == "0") return false;
  const size_t lds = resident_lds(p.Nx, p.Ny, p.precision == 64 ? sizeof(V2<double>) : sizeof(V2<float>));
  return p.scheme == PFC_SCHEME_ETD1 && lds <= std::min(kResidentLds, (size_t)lds_limit());
}
// Cells of at most (Nx/2 + 1) Ny = 512 points with Ny <= 32 and Nx <= 32 take
// the dense resident stepper (three LDS phases per step instead of the FFT
// stepper's dozen).  PFC_HIP_RESIDENT_DENSE=0 keeps the FFT stepper.
bool dense_ok(const pfc_params& p) {
  const char* env = std::getenv("PFC_HIP_RESIDENT_DENSE");
  if (env && std::string(env) == "0") return false;
  const size_t lds = dense_lds_bytes(p.Nx, p.Ny, p.precision == 64 ? sizeof(V2<double>) : sizeof(V2<float>));
  return dense_fits(p.Nx, p.Ny) && lds <= std::min(kResidentLds, (size_t)lds_limit());
}

void engine_init(Engine& e, const pfc_params* params, int cap, bool batched) {
  validate(params);
  if (cap < 1) fail(PFC_EINVAL, "capacity must be >= 1");
  e.p = *params;
  e.cap = cap;
  e.batched = batched;
  const int Nx = params->Nx, Ny = params->Ny;
  e.Nc = Nx / 2 + 1;
  e.lay = state_layout(Nx, Ny, params->precision);
  e.nc = e.lay.size(Ny);  // >= Nc * Ny, the ETDRK2 carry's extent
  e.ntab = (size_t)e.Nc * table_rows(Ny);
  e.csize = e.f64() ? sizeof(V2<double>) : sizeof(V2<float>);
  e.rg = row_geom(Nx, Ny, e.csize);
  e.cg = col_geom(Nx, Ny, e.csize);
  e.rg64 = row_geom(Nx, Ny, sizeof(V2<double>));
  e.cg64 = col_geom(Nx, Ny, sizeof(V2<double>));
  e.rg.lay = e.cg.lay = e.rg64.lay = e.cg64.lay = e.lay;
  e.res = ResGeom{Nx, Nx / 2, e.Nc, Ny, table_rows(Ny), e.lay, e.rg.plM, e.cg.plY};
  e.resident = resident_fits(*params);
  e.dense = e.resident && dense_ok(*params);
  if (!e.resident && !std::getenv("PFC_HIP_GENERIC")) {
    auto rc = row_candidates(Nx, Ny, e.csize);
    auto cc = col_candidates(Nx, Ny, e.csize);
    e.frow = pick_default(rc, 512);
    e.fcol = pick_default(cc, 256);
    // PFC_HIP_SCHEDULE="rowvar:rowlines:colvar:collines" pins a schedule
    // (tests, tuning studies and profiles; a variant + 16 is its capped
    // form); PFC_HIP_SCHEDULE64 / PFC_HIP_SCHEDULE32 pin one precision's
    // plans and take precedence.  The autotuner is then skipped.
    const char* fx = std::getenv(e.f64() ? "PFC_HIP_SCHEDULE64" : "PFC_HIP_SCHEDULE32");
    if (!fx) fx = std::getenv("PFC_HIP_SCHEDULE");
    if (fx) {
      int v[4] = {-1, -1, -1, -1};
      if (std::sscanf(fx, "%d:%d:%d:%d", &v[0], &v[1], &v[2], &v[3]) == 4) {
        bool okr = false, okc = false;
        for (const FastShape& f : rc)
          if (f.code() == v[0] && f.lines == v[1]) e.frow = f, okr = true;
        for (const FastShape& f : cc)
          if (f.code() == v[2] && f.lines == v[3]) e.fcol = f, okc = true;
        if (!okr && e.frow.on) fail(PFC_EINVAL, "PFC_HIP_SCHEDULE: no such row schedule");
        if (!okc && e.fcol.on) fail(PFC_EINVAL, "PFC_HIP_SCHEDULE: no such column schedule");
        e.tuned = true;
      }
    }
    // PFC_HIP_STORES=nt|ordinary pins the output store kind (else autotuned)
    if (const char* sk = std::getenv("PFC_HIP_STORES"))
      if (sk[0] == 'n' || sk[0] == 'o') e.frow.nt = e.fcol.nt = sk[0] == 'n';
  }
  const size_t bytes = (size_t)cap * e.nc * e.csize;
  const size_t tbytes = (size_t)cap * e.ntab * sizeof(double);  // ETD tables are FP64 in every precision
  e.S.alloc(bytes);
  e.Tb.alloc(bytes);
  if (params->scheme == PFC_SCHEME_ETDRK2) e.D.alloc(bytes);
  e.E1.alloc(tbytes);
  e.Q1.alloc(tbytes);
  if (params->scheme == PFC_SCHEME_ETDRK2) e.Q2.alloc(tbytes);
  if (!e.f64()) {  // FP32 copies for the FP32 fast column kernel
    const size_t fbytes = (size_t)cap * e.ntab * sizeof(float);
    e.E1f.alloc(fbytes);
    e.Q1f.alloc(fbytes);
    if (params->scheme == PFC_SCHEME_ETDRK2) e.Q2f.alloc(fbytes);
  }
  // every slot starts in the shared box with the shared dt
  const std::vector<double> kx2 = k_squared(rfft_k(Nx, params->Lx / Nx)), ky2 = k_squared(fft_k(Ny, params->Ly / Ny));
  std::vector<double> kx2s, ky2s;
  for (int i = 0; i < cap; ++i) {
    kx2s.insert(kx2s.end(), kx2.begin(), kx2.end());
    ky2s.insert(ky2s.end(), ky2.begin(), ky2.end());
  }
  e.kx2.upload(kx2s);
  e.ky2.upload(ky2s);
  e.h_Lx.assign(cap, params->Lx);
  e.h_Ly.assign(cap, params->Ly);
  e.h_dt.assign(cap, params->dt);
  e.dts.upload(e.h_dt);
  auto tn = twiddles(Nx), ty = twiddles(Ny);
  e.twN64.upload(tn);
  e.twY64.upload(ty);
  if (e.f64()) {
    e.twN.upload(tn);
    e.twY.upload(ty);
  } else {
    e.twN.upload(to_f32(tn));
    e.twY.upload(to_f32(ty));
  }
  e.part.alloc((size_t)cap * e.pstride() * sizeof(Partial));
  e.out.alloc((size_t)cap * 4 * sizeof(double));
  e.pin.alloc((size_t)cap * sizeof(double));
  e.status.alloc((size_t)cap * sizeof(int));
  e.active.alloc((size_t)cap * sizeof(int));
  e.h_eps.assign(cap, params->epsilon);
  e.h_M.assign(cap, params->M);
  e.eps.upload(e.h_eps);
  e.Ms.upload(e.h_M);
  hip_check(hipMemset(e.pin.p, 0, e.pin.n), "memset");
  hip_check(hipMemset(e.status.p, 0, e.status.n), "memset");
  hip_check(hipMemset(e.S.p, 0, bytes), "memset");
  e.field_set.assign(cap, 0);
  hip_check(hipHostMalloc((void**)&e.host_out, (size_t)cap * 4 * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
  hip_check(hipHostMalloc((void**)&e.host_status, (size_t)cap * sizeof(int), hipHostMallocDefault), "hipHostMalloc");
  // every slot starts with tables for the shared epsilon and M, and is active
  e.h_active.resize(cap);
  for (int i = 0; i < cap; ++i) e.h_active[i] = i;
  e.active.upload(e.h_active);
}

TabParams tab_params(const Engine& e) {
  return TabParams{e.p.Gamma, e.p.beta, e.p.q0, 0.0, 0.0, e.p.dt, e.p.smallZTol,
                   1.0 / e.p.Ny, 1.0 / ((double)e.p.Nx * (double)e.p.Ny), e.Nc, e.p.Ny, table_rows(e.p.Ny),
                   e.fold_shift() ? 1 : 0};
}

// ETD tables of slots [slot0, slot0 + n), built in FP64 on the device
void engine_tables(Engine& e, int slot0, int n, hipStream_t st) {
  const TabParams tp = tab_params(e);
  const int blocks = (int)std::min<size_t>((e.ntab + 255) / 256, 1024);
  const SlotGrid sg{(const double*)e.kx2.p, (const double*)e.ky2.p, (const double*)e.dts.p};
  hipLaunchKernelGGL(k_tables<double>, dim3(blocks, n), dim3(256), 0, st, (double*)e.E1.p, (double*)e.Q1.p,
                     (double*)e.Q2.p, (long long)e.ntab, sg, tp, (const double*)e.eps.p, (const double*)e.Ms.p,
                     slot0, (const int*)nullptr, (const int*)nullptr, (float*)e.E1f.p, (float*)e.Q1f.p,
                     (float*)e.Q2f.p);
  if (e.has_ext()) {
    const long long cs = (long long)e.Nc * e.p.Ny;
    const int cb = (int)std::min<long long>((cs + 255) / 256, 1024);
    hipLaunchKernelGGL(k_ext_table, dim3(cb, n), dim3(256), 0, st, (V2<double>*)e.extc.p,
                       (const V2<double>*)e.vhat.p, cs, sg, tp, (const double*)e.eps.p, (const double*)e.Ms.p, slot0);
  }
  hip_check(hipGetLastError(), "table launch");
}

Nonlin nonlin_of(const Engine& e) {
  Nonlin q{};
  q.model = e.p.model;
  q.clamp = e.p.clamp_log;
  q.g = e.p.g;
  q.minPhi = e.p.minPhi;
  q.logMin = std::log(e.p.minPhi > 0 ? e.p.minPhi : 1e-300);
  q.M = 0.0;  // per slot, from Slots.M
  return q;
}

// The device-managed stabilising shift, before a step: the min and max of
// every active slot's field (with the fast row shape k_row_range, the step's
// own inverse row transforms, so the range is that of the values the
// nonlinearity sees; else k_fe_row<MINMAX>, the arithmetic of pfc_reduce's),
// the range check and refit (k_shift_update), and the ETD tables (and
// external-potential source) of the refitted slots only.  No host synchronisation; the launches can be
// recorded in a step graph.
template <typename T>
void shift_check(Engine& e, hipStream_t st, bool row_partials = false) {
  const int ny = e.nact();
  if (ny == 0) return;
  const Slots sl = e.slots_active();
  const FeParams fp{e.p.model, e.p.beta, e.p.q0, e.p.epsilon, e.p.g, (const double*)e.kx2.p,
                    (const double*)e.ky2.p, (const double*)e.eps.p};
  if (row_partials) {
    // the step's row kernel wrote one range pair per row tile (fused_range)
    hipLaunchKernelGGL(k_shift_update<RangePart>, dim3(ny), dim3(256), 0, st, (const RangePart*)e.rpart.p,
                       e.frow.blocks, sl, e.frow.blocks, e.shp, (ShiftState*)e.shst.p, (double*)e.Ms.p,
                       (int*)e.rebuildf.p);
  } else if (e.frow.on) {
    // the fast row shape's inverse transforms
    const FastShape& f = e.frow;
    if (!row_range_launch(e.p.Nx / 2, f.var, dim3(f.blocks, ny), dim3(f.lines * f.tpl), f.lds, st,
                          (const V2<T>*)e.S.p, e.p.Ny, e.lay, (const V2<double>*)e.twN64.p, sl, (RangePart*)e.rpart.p))
      fail(PFC_EUNSUPPORTED, "no fast row schedule");
    hipLaunchKernelGGL(k_shift_update<RangePart>, dim3(ny), dim3(256), 0, st, (const RangePart*)e.rpart.p, f.blocks,
                       sl, f.blocks, e.shp, (ShiftState*)e.shst.p, (double*)e.Ms.p, (int*)e.rebuildf.p);
  } else {
    Partial* part = (Partial*)e.part.p;
    const int rb = row_blocks(e.rg64), ps = e.pstride();
    const size_t rl = std::max(row_lds(e.rg64, sizeof(V2<double>)), 256 * sizeof(Partial));
    hipLaunchKernelGGL((k_fe_row<T, true>), dim3(rb, ny), dim3(256), rl, st, (const V2<T>*)e.S.p, e.rg64, fp,
                       (const V2<double>*)e.twN64.p, part, sl, ps);
    hipLaunchKernelGGL(k_shift_update<Partial>, dim3(ny), dim3(256), 0, st, (const Partial*)part, rb, sl, ps, e.shp,
                       (ShiftState*)e.shst.p, (double*)e.Ms.p, (int*)e.rebuildf.p);
  }
  const TabParams tp = tab_params(e);
  const SlotGrid sg{(const double*)e.kx2.p, (const double*)e.ky2.p, (const double*)e.dts.p};
  const int blocks = (int)std::min<size_t>((e.ntab + 255) / 256, 1024);
  hipLaunchKernelGGL(k_tables<double>, dim3(blocks, ny), dim3(256), 0, st, (double*)e.E1.p, (double*)e.Q1.p,
                     (double*)e.Q2.p, (long long)e.ntab, sg, tp, (const double*)e.eps.p, (const double*)e.Ms.p, 0,
                     (const int*)e.active.p, (const int*)e.rebuildf.p, (float*)e.E1f.p, (float*)e.Q1f.p,
                     (float*)e.Q2f.p);
  if (e.has_ext()) {
    const long long cs = (long long)e.Nc * e.p.Ny;
    const int cb = (int)std::min<long long>((cs + 255) / 256, 1024);
    hipLaunchKernelGGL(k_ext_table, dim3(cb, ny), dim3(256), 0, st, (V2<double>*)e.extc.p,
                       (const V2<double>*)e.vhat.p, cs, sg, tp, (const double*)e.eps.p, (const double*)e.Ms.p, 0,
                       (const int*)e.active.p, (const int*)e.rebuildf.p);
  }
}

// one nonlinear row pass and one column pass (MODE) over `ny` instances, on
// the engine's arrays or on caller-given ones of the same shape
template <typename T>
struct Steps {
  using V = V2<T>;
  static void row(Engine& e, hipStream_t st, const Slots& sl, int ny, RangePart* rpart = nullptr,
                  const RangeCheck& chk = RangeCheck{}) {
    row(e, st, sl, ny, (const V*)e.S.p, (V*)e.Tb.p, rpart, chk);
  }
  static void col(Engine& e, int mode, hipStream_t st, const Slots& sl, int ny, bool refit = false) {
    col(e, mode, st, sl, ny, (V*)e.S.p, (const V*)e.Tb.p, (V*)e.D.p, refit);
  }
  // chk: the in-step decision's tile test (fused range shape only)
  static void row(Engine& e, hipStream_t st, const Slots& sl, int ny, const V* S, V* Tb,
                  RangePart* rpart = nullptr, const RangeCheck& chk = RangeCheck{}) {
    const Nonlin q = nonlin_of(e);
    Slots sr = sl;
    if (e.fold_shift()) sr.M = nullptr;  // the shift acts through E1
    if (e.frow.on)
      FastLaunch<T>::row(e.p.Nx / 2, e.frow, ny, st, S, Tb, e.p.Ny, e.lay, q, (const V2<double>*)e.twN64.p, sr,
                         rpart, chk);
    else
      hipLaunchKernelGGL((k_row_nonlin<T>), dim3(row_blocks(e.rg), ny), dim3(256), row_lds(e.rg, sizeof(V)), st, S,
                         Tb, e.rg, q, (const V*)e.twN.p, sr);
  }
  // refit (col_refit engines, ETD1): the column kernel reads the step's
  // breach flags and refits a flagged slot itself (shift, tables, fit)
  static void col(Engine& e, int mode, hipStream_t st, const Slots& sl, int ny, V* S, const V* Tb, V* D,
                  bool refit = false) {
    ColTables tb = e.tables();
    if (refit) {
      if (mode != COL_ETD1 || !e.col_refit()) fail(PFC_EUNSUPPORTED, "in-step refit on a path without it");
      tb.brk = (const int*)e.brk.p;
      tb.par = e.par;
      tb.part = (const RangePart*)e.rpart.p;
      tb.nrow = tb.pstride = e.frow.blocks;
      tb.st = (ShiftState*)e.shst.p;
      tb.Ms = (double*)e.Ms.p;
      tb.g = e.shp.g;
      tb.minPhi = e.shp.minPhi;
      tb.eps = (const double*)e.eps.p;
      tb.sg = SlotGrid{(const double*)e.kx2.p, (const double*)e.ky2.p, (const double*)e.dts.p};
      tb.tp = tab_params(e);
    }
    if (e.fcol.on) {
      FastLaunch<T>::col(e.p.Ny, mode, e.fcol, ny, st, S, Tb, D, e.Nc, e.lay, tb, (const V2<double>*)e.twY64.p, sl);
      return;
    }
    const dim3 g(col_blocks(e.cg), ny);
    const size_t lds = col_lds(e.cg, sizeof(V));
    const V* tw = (const V*)e.twY.p;
    if (mode == COL_ETD1)
      hipLaunchKernelGGL((k_col_etd<T, COL_ETD1>), g, dim3(256), lds, st, S, Tb, D, e.cg, tb, tw, sl);
    else if (mode == COL_RK2_STAGE1)
      hipLaunchKernelGGL((k_col_etd<T, COL_RK2_STAGE1>), g, dim3(256), lds, st, S, Tb, D, e.cg, tb, tw, sl);
    else
      hipLaunchKernelGGL((k_col_etd<T, COL_RK2_STAGE2>), g, dim3(256), lds, st, S, Tb, D, e.cg, tb, tw, sl);
  }
  // nsteps ETD1 steps of every active slot, one workgroup each, in launches
  // of at most kResidentChunk steps
  // counts (device, per active slot) or null: every slot nsteps
  static void resident(Engine& e, long long nsteps, hipStream_t st, const Slots& sl, int ny, V* S,
                       const long long* counts = nullptr) {
    constexpr long long kResidentChunk = 4096;
    ColTables tb = e.tables();
    if (e.dense) {
      const int nsp = e.Nc * e.p.Ny;
      const size_t lds = dense_lds_bytes(e.p.Nx, e.p.Ny, sizeof(V));
      // the smallest compiled row and angle bounds that hold this cell
      const int rows = e.p.Ny <= 16 ? 16 : e.p.Ny <= 24 ? 24 : 32;
      const int angles = e.p.Nx / 2 - 1 <= 7 ? 7 : 15;
      for (long long done = 0; done < nsteps; done += kResidentChunk) {
        const long long n = std::min(kResidentChunk, nsteps - done);
#define PFC_DENSE(NY, MH)                                                                                 \
  hipLaunchKernelGGL((k_resident_dense<T, NY, MH>), dim3(ny), dim3((nsp + 63) / 64 * 64), lds, st, S, e.res, \
                     nonlin_of(e), tb, sl, n, counts, done)
        if (rows == 16 && angles == 7)
          PFC_DENSE(16, 7);
        else if (rows == 16)
          PFC_DENSE(16, 15);
        else if (rows == 24 && angles == 7)
          PFC_DENSE(24, 7);
        else if (rows == 24)
          PFC_DENSE(24, 15);
        else if (angles == 7)
          PFC_DENSE(32, 7);
        else
          PFC_DENSE(32, 15);
#undef PFC_DENSE
      }
      return;
    }
    const size_t lds = resident_lds(e.p.Nx, e.p.Ny, sizeof(V));
    static const int threads = [] {
      const char* v = std::getenv("PFC_HIP_RESIDENT_THREADS");
      return v && std::atoi(v) == 256 ? 256 : 512;
    }();
    for (long long done = 0; done < nsteps; done += kResidentChunk) {
      const long long n = std::min(kResidentChunk, nsteps - done);
#define PFC_RESIDENT(BS)                                                                                          \
  hipLaunchKernelGGL((k_resident_etd1<T, BS>), dim3(ny), dim3(BS), lds, st, S, e.res, nonlin_of(e), tb,          \
                     (const V*)e.twN.p, (const V*)e.twY.p, sl, n, counts, done)
      if (threads == 512)
        PFC_RESIDENT(512);
      else
        PFC_RESIDENT(256);
#undef PFC_RESIDENT
    }
  }
  static void step(Engine& e, long long nsteps, hipStream_t st) {
    const int ny = e.nact();
    if (ny == 0) return;
    const Slots sl = e.slots_active();
    if (e.resident) {
      if (e.auto_shift) {
        for (long long s = 0; s < nsteps; ++s) {
          shift_check<T>(e, st);
          resident(e, 1, st, sl, ny, (V*)e.S.p);
        }
      } else {
        resident(e, nsteps, st, sl, ny, (V*)e.S.p);
      }
      hip_check(hipGetLastError(), "step launch");
      return;
    }
    loop(e, nsteps, st);
  }
  // nsteps of the row/column launch sequence (also what a step graph holds)
  static void loop(Engine& e, long long nsteps, hipStream_t st) {
    const int ny = e.nact();
    const Slots sl = e.slots_active();
    const bool fused = e.fused_range();
    for (long long s = 0; s < nsteps; ++s) {
      if (fused) {
        // row pass (N without the shift, and the field's range), decision and
        // rebuild, then the column pass with the refitted E1; with col_refit
        // the decision and rebuild happen inside the two launches (breach
        // flags of this step's parity)
        if (e.col_refit()) {
          const RangeCheck chk{(const ShiftState*)e.shst.p, (int*)e.brk.p, e.par, e.shp.floor_tol, e.shp.ceil_tol};
          row(e, st, sl, ny, (RangePart*)e.rpart.p, chk);
          col(e, COL_ETD1, st, sl, ny, true);
          e.par ^= 1;
        } else {
          row(e, st, sl, ny, (RangePart*)e.rpart.p);
          shift_check<T>(e, st, true);
          col(e, COL_ETD1, st, sl, ny);
        }
        continue;
      }
      if (e.auto_shift) shift_check<T>(e, st);
      row(e, st, sl, ny);
      if (e.p.scheme == PFC_SCHEME_ETD1) {
        col(e, COL_ETD1, st, sl, ny);
      } else {
        col(e, COL_RK2_STAGE1, st, sl, ny);
        row(e, st, sl, ny);
        col(e, COL_RK2_STAGE2, st, sl, ny);
      }
    }
    hip_check(hipGetLastError(), "step launch");
  }
};

// ---------------------------------------------------------------------------
// autotuner: at an engine's first step, time every (schedule variant, lines
// per block) candidate of the fast row and column kernels on scratch copies
// of the live state and keep the fastest.  Choices are cached per (device,
// kernel kind, lengths, precision, mode, instances per launch) for the life of
// the process.  PFC_HIP_TUNE=0 keeps the defaults.
// ---------------------------------------------------------------------------
std::mutex g_tune_mu;
std::map<std::array<int, 7>, std::pair<int, int>> g_tune;  // -> (variant, lines)

template <typename F>
float time_launches(hipStream_t st, int reps, F&& launch) {
  hipEvent_t a = nullptr, b = nullptr;
  hip_check(hipEventCreate(&a), "event create");
  hip_check(hipEventCreate(&b), "event create");
  launch();
  hip_check(hipEventRecord(a, st), "event record");
  for (int i = 0; i < reps; ++i) launch();
  hip_check(hipEventRecord(b, st), "event record");
  hip_check(hipEventSynchronize(b), "event sync");
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, a, b), "event elapsed");
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  hip_check(hipGetLastError(), "tuning launch");
  return ms / reps;
}

int bucket(int n) {
  int b = 1;
  while (b < n) b *= 2;
  return b;
}

template <typename T>
void autotune(Engine& e, hipStream_t st) {
  e.tuned = true;
  const char* env = std::getenv("PFC_HIP_TUNE");
  if (e.resident || (env && env[0] == '0')) return;
  if (e.has_ext()) e.fcol.on = false;  // the fast column kernel has no potential term
  using V = V2<T>;
  const int ny = e.nact();
  if (ny == 0) return;
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  const int Nx = e.p.Nx, Ny = e.p.Ny, M = Nx / 2, Nc = e.Nc;
  // launches per timing; every candidate is timed in two rounds (forward and
  // reverse order) and keeps its faster one, so a transient stall of the
  // device does not decide the choice
  const int reps = ny > 8 ? 4 : 20;
  auto pick_fastest = [&](const std::vector<FastShape>& cands, auto&& launch) {
    std::vector<float> t(cands.size(), 1e30f);
    for (int round = 0; round < 2; ++round)
      for (size_t j = 0; j < cands.size(); ++j) {
        const size_t i = round ? cands.size() - 1 - j : j;
        t[i] = std::min(t[i], time_launches(st, reps, [&] { launch(cands[i]); }));
      }
    return cands[std::min_element(t.begin(), t.end()) - t.begin()];
  };
  const Slots sl = e.slots_active();
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto apply = [](FastShape& cur, const std::vector<FastShape>& cands, std::pair<int, int> vl) {
    for (const FastShape& f : cands)
      if (f.code() == vl.first && f.lines == vl.second) cur = f;
  };
  // scratch arrays for every slot index the launches may touch
  const size_t bytes = (size_t)e.cap * e.nc * sizeof(V);
  if (e.frow.on) {
    const std::array<int, 7> key{dev, 0, M, Ny, e.p.precision, e.p.model * 2 + e.p.clamp_log, bucket(ny)};
    auto cands = row_candidates(Nx, Ny, e.csize);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      apply(e.frow, cands, it->second);
    } else if (cands.size() > 1) {
      DevBuf T2;
      T2.alloc(bytes);
      const Nonlin q = nonlin_of(e);
      const FastShape pick = pick_fastest(cands, [&](const FastShape& f) {
        FastLaunch<T>::row(M, f, ny, st, (const V*)e.S.p, (V*)T2.p, Ny, e.lay, q, (const V2<double>*)e.twN64.p, sl);
      });
      e.frow = pick;
      g_tune[key] = {pick.code(), pick.lines};
    }
  }
  if (e.fcol.on) {
    const int mode = e.p.scheme == PFC_SCHEME_ETD1 ? COL_ETD1 : COL_RK2_STAGE1;
    const std::array<int, 7> key{dev, 1, Ny, Nc, e.p.precision, mode, bucket(ny)};
    auto cands = col_candidates(Nx, Ny, e.csize);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      apply(e.fcol, cands, it->second);
    } else if (cands.size() > 1) {
      DevBuf S2, T2, D2;
      S2.alloc(bytes);
      T2.alloc(bytes);
      D2.alloc(bytes);
      hip_check(hipMemcpyAsync(S2.p, e.S.p, bytes, hipMemcpyDeviceToDevice, st), "copy");
      hip_check(hipMemcpyAsync(T2.p, e.S.p, bytes, hipMemcpyDeviceToDevice, st), "copy");
      ColTables tb = e.tables();
      const FastShape pick = pick_fastest(cands, [&](const FastShape& f) {
        FastLaunch<T>::col(Ny, mode, f, ny, st, (V*)S2.p, (const V*)T2.p, (V*)D2.p, Nc, e.lay, tb,
                           (const V2<double>*)e.twY64.p, sl);
      });
      e.fcol = pick;
      g_tune[key] = {pick.code(), pick.lines};
      hip_check(hipStreamSynchronize(st), "sync after tuning");
    }
  }
  // Store kind of both kernels' outputs: non-temporal stores stream the
  // arrays out while the kernel runs (the next kernel reads them from other
  // XCDs), ordinary stores let L2 merge the partial 128-byte lines of tiles
  // narrower than a line.  Which wins depends on the shape and the working
  // set, and only shows across the kernel boundary, so the two are timed as
  // a step loop (row, column[, row, column]) on scratch copies of the state.
  const char* ntenv = std::getenv("PFC_HIP_STORES");  // "nt" or "ordinary" pins the kind
  if (ntenv && (ntenv[0] == 'n' || ntenv[0] == 'o')) {
    e.frow.nt = e.fcol.nt = ntenv[0] == 'n';
  } else if (e.frow.on && e.fcol.on) {
    const std::array<int, 7> key{dev, 2, M, Ny, e.p.precision, e.p.scheme, bucket(ny)};
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      e.frow.nt = e.fcol.nt = it->second.first != 0;
    } else {
      DevBuf S2, T2, D2;
      S2.alloc(bytes);
      T2.alloc(bytes);
      D2.alloc(bytes);
      const Nonlin q = nonlin_of(e);
      ColTables tb = e.tables();
      const bool rk2 = e.p.scheme == PFC_SCHEME_ETDRK2;
      auto step = [&](bool nt) {
        FastShape fr = e.frow, fc = e.fcol;
        fr.nt = fc.nt = nt;
        auto row = [&] {
          FastLaunch<T>::row(M, fr, ny, st, (const V*)S2.p, (V*)T2.p, Ny, e.lay, q, (const V2<double>*)e.twN64.p, sl);
        };
        auto col = [&](int mode) {
          FastLaunch<T>::col(Ny, mode, fc, ny, st, (V*)S2.p, (const V*)T2.p, (V*)D2.p, Nc, e.lay, tb,
                             (const V2<double>*)e.twY64.p, sl);
        };
        row();
        col(rk2 ? COL_RK2_STAGE1 : COL_ETD1);
        if (rk2) {
          row();
          col(COL_RK2_STAGE2);
        }
      };
      float t[2] = {1e30f, 1e30f};
      for (int round = 0; round < 2; ++round)
        for (int j = 0; j < 2; ++j) {
          const int nt = round ? 1 - j : j;
          hip_check(hipMemcpyAsync(S2.p, e.S.p, bytes, hipMemcpyDeviceToDevice, st), "copy");
          t[nt] = std::min(t[nt], time_launches(st, ny > 8 ? 4 : 8, [&] { step(nt != 0); }));
        }
      const bool nt = t[1] <= t[0];
      e.frow.nt = e.fcol.nt = nt;
      g_tune[key] = {nt ? 1 : 0, 0};
      hip_check(hipStreamSynchronize(st), "sync after tuning");
    }
  }
}

// ---------------------------------------------------------------------------
// step-loop graphs: a call of n >= kGraphSteps steps on the two-kernel path
// replays a HIP graph of kGraphSteps steps n / kGraphSteps times and runs the
// rest eagerly, so the host enqueues one graph instead of 2 (ETD1) or 4
// (ETDRK2) kernels per step.  A graph is captured on a private stream when a
// call of n >= kGraphCaptureSteps steps meets a launch configuration without
// one (at least eight replays to pay for the capture and instantiation; a
// serving loop's short calls on changing active sets never capture), and is
// replayed by any later call of n >= kGraphSteps with that configuration.
// Graphs are kept for the engine's life: their nodes hold the engine's own
// arrays, the active-list pointer and
// the schedule, and read slot ids, tables, pins and shifts from device
// memory when they run, so it stays valid as those change.  What changes a
// node (the active count, the schedule, an external potential's arrays) is in
// the key.  The nodes are exactly the eager launches, so results are
// bit-identical.  PFC_HIP_GRAPH=0 steps eagerly.
// ---------------------------------------------------------------------------
constexpr long long kGraphSteps = 32;
constexpr long long kGraphCaptureSteps = 8 * kGraphSteps;
constexpr size_t kGraphCacheMax = 16;

bool graphs_enabled() {
  static const bool on = [] {
    const char* v = std::getenv("PFC_HIP_GRAPH");
    return !(v && v[0] == '0');
  }();
  return on;
}

// the cached graph of the engine's launch configuration; without one,
// captured when `capture`, else null
hipGraphExec_t step_graph(Engine& e, bool capture) {
  // par: the in-step decision's parity at the graph's first step (an even
  // number of steps returns it there)
  const std::array<long long, 11> key{e.nact(),          e.frow.on, e.frow.code(), e.frow.lines,
                                      e.fcol.on,         e.fcol.code(), e.fcol.lines,
                                      (long long)(intptr_t)e.extc.p, (long long)(intptr_t)e.D.p, e.auto_shift,
                                      e.par};
  auto it = e.graphs.find(key);
  if (it != e.graphs.end()) return it->second;
  if (!capture) return nullptr;
  if (e.graphs.size() >= kGraphCacheMax) e.drop_graphs();
  if (!e.cap_st) hip_check(hipStreamCreateWithFlags(&e.cap_st, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipStreamBeginCapture(e.cap_st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  hipGraph_t g = nullptr;
  try {
    if (e.f64())
      Steps<double>::loop(e, kGraphSteps, e.cap_st);
    else
      Steps<float>::loop(e, kGraphSteps, e.cap_st);
  } catch (...) {
    (void)hipStreamEndCapture(e.cap_st, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  hip_check(hipStreamEndCapture(e.cap_st, &g), "hipStreamEndCapture");
  hipGraphExec_t x = nullptr;
  const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  hip_check(ie, "hipGraphInstantiate");
  e.graphs[key] = x;
  ++e.graph_captures;
  return x;
}

// `capture`: a launch configuration without a graph may get one (by
// default when this call alone has enough steps to pay for it)
void engine_step(Engine& e, long long nsteps, hipStream_t st, int capture = -1) {
  if (nsteps < 0) fail(PFC_EINVAL, "nsteps < 0");
  for (int s : e.h_active)
    if (!e.field_set[s]) fail(PFC_EINVAL, "a stepped slot has no field (pfc_set_field / pfc_batch_set_field)");
  if (!e.tuned && !e.resident) {
    if (e.f64())
      autotune<double>(e, st);
    else
      autotune<float>(e, st);
  }
  const hipGraphExec_t g = !e.resident && e.nact() > 0 && nsteps >= kGraphSteps && graphs_enabled()
                               ? step_graph(e, capture < 0 ? nsteps >= kGraphCaptureSteps : capture != 0)
                               : nullptr;
  if (g) {
    const long long reps = nsteps / kGraphSteps;
    for (long long r = 0; r < reps; ++r) hip_check(hipGraphLaunch(g, st), "hipGraphLaunch");
    e.graph_replays += reps;
    nsteps -= reps * kGraphSteps;
  }
  if (e.f64())
    Steps<double>::step(e, nsteps, st);
  else
    Steps<float>::step(e, nsteps, st);
}

// counts[i] steps for the i-th active slot (resident engines only)
void engine_step_counts(Engine& e, const long long* counts, hipStream_t st) {
  const int n = e.nact();
  if (n == 0) return;
  if (!counts) fail(PFC_EINVAL, "NULL counts");
  if (!e.resident) fail(PFC_EUNSUPPORTED, "per-slot step counts need the resident stepper (small ETD1 grids)");
  if (e.auto_shift) fail(PFC_EUNSUPPORTED, "per-slot step counts with a device-managed shift");
  for (int s : e.h_active)
    if (!e.field_set[s]) fail(PFC_EINVAL, "a stepped slot has no field (pfc_set_field / pfc_batch_set_field)");
  long long most = 0;
  for (int i = 0; i < n; ++i) {
    if (counts[i] < 0) fail(PFC_EINVAL, "negative step count");
    most = std::max(most, counts[i]);
  }
  if (e.cnt.n < (size_t)n * sizeof(long long)) {
    hip_check(hipStreamSynchronize(st), "sync before scratch growth");
    e.cnt.alloc((size_t)e.cap * sizeof(long long));
  }
  hip_check(hipMemcpyAsync(e.cnt.p, counts, (size_t)n * sizeof(long long), hipMemcpyHostToDevice, st), "copy");
  const Slots sl = e.slots_active();
  if (e.f64())
    Steps<double>::resident(e, most, st, sl, n, (V2<double>*)e.S.p, (const long long*)e.cnt.p);
  else
    Steps<float>::resident(e, most, st, sl, n, (V2<float>*)e.S.p, (const long long*)e.cnt.p);
  hip_check(hipGetLastError(), "step launch");
  hip_check(hipStreamSynchronize(st), "sync");  // the counts must outlive the copy
}

// Average launch time of each kernel kind of the engine's current schedule:
// n launches back to back on scratch copies of the live state, one event pair
// per kind (kind 0 row pass, 1 first column pass, 2 ETDRK2 second column
// pass).  The engine's own arrays are not touched.
template <typename T>
void kernel_times(Engine& e, long long n, hipStream_t st, double ms[3], long long calls[3]) {
  using V = V2<T>;
  for (int k = 0; k < 3; ++k) ms[k] = 0.0, calls[k] = 0;
  const int ny = e.nact();
  if (ny == 0 || n <= 0) return;
  const Slots sl = e.slots_active();
  const size_t bytes = (size_t)e.cap * e.nc * sizeof(V);
  DevBuf S2, T2, D2;
  S2.alloc(bytes);
  T2.alloc(bytes);
  D2.alloc(bytes);
  hip_check(hipMemcpyAsync(S2.p, e.S.p, bytes, hipMemcpyDeviceToDevice, st), "copy");
  hip_check(hipMemcpyAsync(T2.p, e.Tb.p, bytes, hipMemcpyDeviceToDevice, st), "copy");
  if (e.D.p) hip_check(hipMemcpyAsync(D2.p, e.D.p, bytes, hipMemcpyDeviceToDevice, st), "copy");
  if (e.resident) {  // kind 0: one resident launch per step
    ms[0] = (double)time_launches(st, (int)n, [&] { Steps<T>::resident(e, 1, st, sl, ny, (V*)S2.p); }) * n;
    calls[0] = n;
    hip_check(hipStreamSynchronize(st), "sync");
    return;
  }
  const bool rk2 = e.p.scheme == PFC_SCHEME_ETDRK2;
  for (int k = 0; k < (rk2 ? 3 : 2); ++k) {
    const float m = time_launches(st, (int)n, [&] {
      if (k == 0)
        Steps<T>::row(e, st, sl, ny, (const V*)S2.p, (V*)T2.p);
      else
        Steps<T>::col(e, k == 1 ? (rk2 ? COL_RK2_STAGE1 : COL_ETD1) : COL_RK2_STAGE2, st, sl, ny, (V*)S2.p,
                      (const V*)T2.p, (V*)D2.p);
    });
    ms[k] = (double)m * n;
    calls[k] = n;
  }
  hip_check(hipStreamSynchronize(st), "sync");
}

// FP64 reductions of `ny` instances (the active list, or one slot when
// Sbase is offset to it and sl.list is null); results land in e.out[4 i ..];
// pin_out (device, per slot) receives the means.
template <typename T>
void launch_reduce(Engine& e, const void* Sbase, const double* eps, const double* kx2, const double* ky2,
                   hipStream_t st, const Slots& sl, int ny, double* pin_out, double* out = nullptr) {
  using V = V2<T>;
  const V* S = (const V*)Sbase;
  FeParams fp{e.p.model, e.p.beta, e.p.q0, e.p.epsilon, e.p.g, kx2, ky2, eps};
  Partial* part = (Partial*)e.part.p;
  const int rb = row_blocks(e.rg64), cb = col_blocks(e.cg64), ps = e.pstride();
  const size_t red = 256 * sizeof(Partial);
  const size_t rl = std::max(row_lds(e.rg64, sizeof(V2<double>)), red);
  const size_t cl = std::max(col_lds(e.cg64, sizeof(V2<double>)), red);
  hipLaunchKernelGGL((k_fe_row<T>), dim3(rb, ny), dim3(256), rl, st, S, e.rg64, fp, (const V2<double>*)e.twN64.p,
                     part, sl, ps);
  hipLaunchKernelGGL((k_fe_col<T>), dim3(cb, ny), dim3(256), cl, st, S, e.cg64, fp, e.p.Nx,
                     (const V2<double>*)e.twY64.p, part + rb, sl, ps);
  hipLaunchKernelGGL(k_fe_final, dim3(ny), dim3(256), 0, st, part, rb, cb, e.p.Nx, e.p.Ny,
                     out ? out : (double*)e.out.p, pin_out, sl, ps);
  hip_check(hipGetLastError(), "reduce launch");
}

// reduce one slot; optionally capture its mean as the slot's k = 0 pin
void reduce_slot(Engine& e, int slot, hipStream_t st, bool set_pin) {
  const Slots one{nullptr, 0, 0, nullptr, nullptr, nullptr};
  const void* S = (const char*)e.S.p + (size_t)slot * e.nc * e.csize;
  double* pin_out = set_pin ? (double*)e.pin.p + slot : nullptr;
  const double* eps = (const double*)e.eps.p + slot;
  const double* kx2 = (const double*)e.kx2.p + (size_t)slot * e.Nc;
  const double* ky2 = (const double*)e.ky2.p + (size_t)slot * e.p.Ny;
  if (e.f64())
    launch_reduce<double>(e, S, eps, kx2, ky2, st, one, 1, pin_out);
  else
    launch_reduce<float>(e, S, eps, kx2, ky2, st, one, 1, pin_out);
}

// a new field or new slot parameters: the device-managed shift refits at the
// next step (the host-driven split resets its window on a new field too)
void reset_shift_fit(Engine& e, int slot, hipStream_t st) {
  if (!e.auto_shift) return;
  hip_check(hipMemsetAsync((ShiftState*)e.shst.p + slot, 0, sizeof(ShiftState), st), "memset");
  hip_check(hipMemsetAsync((int*)e.brk.p + 2 * slot, 0, 2 * sizeof(int), st), "memset");
}

void engine_set_field(Engine& e, int slot, const void* field, int dtype, hipStream_t st) {
  e.check_slot(slot);
  if (!field) fail(PFC_EINVAL, "NULL field");
  if (dtype != 64 && dtype != 32) fail(PFC_EINVAL, "dtype must be 64 or 32");
  const int rb = row_blocks(e.rg);
  const size_t rl = row_lds(e.rg, e.csize);
  const size_t off = (size_t)slot * e.nc;
  if (e.f64()) {
    const double inv = 1.0 / e.p.Nx;
    V2<double>* S = (V2<double>*)e.S.p + off;
    if (dtype == 64)
      hipLaunchKernelGGL((k_row_r2c<double, double>), dim3(rb), dim3(256), rl, st, (const double*)field, S, e.rg,
                         (const V2<double>*)e.twN.p, inv, (size_t)0, (size_t)0);
    else
      hipLaunchKernelGGL((k_row_r2c<double, float>), dim3(rb), dim3(256), rl, st, (const float*)field, S, e.rg,
                         (const V2<double>*)e.twN.p, inv, (size_t)0, (size_t)0);
  } else {
    const float inv = 1.0f / e.p.Nx;
    V2<float>* S = (V2<float>*)e.S.p + off;
    if (dtype == 64)
      hipLaunchKernelGGL((k_row_r2c<float, double>), dim3(rb), dim3(256), rl, st, (const double*)field, S, e.rg,
                         (const V2<float>*)e.twN.p, inv, (size_t)0, (size_t)0);
    else
      hipLaunchKernelGGL((k_row_r2c<float, float>), dim3(rb), dim3(256), rl, st, (const float*)field, S, e.rg,
                         (const V2<float>*)e.twN.p, inv, (size_t)0, (size_t)0);
  }
  hip_check(hipGetLastError(), "set_field launch");
  reduce_slot(e, slot, st, true);
  e.field_set[slot] = 1;
  reset_shift_fit(e, slot, st);
}

void engine_get_field(Engine& e, int slot, void* field, int dtype, hipStream_t st) {
  e.check_slot(slot);
  if (!field) fail(PFC_EINVAL, "NULL field");
  if (dtype != 64 && dtype != 32) fail(PFC_EINVAL, "dtype must be 64 or 32");
  const int rb = row_blocks(e.rg);
  const size_t rl = row_lds(e.rg, e.csize);
  const size_t off = (size_t)slot * e.nc;
  if (e.f64()) {
    const V2<double>* S = (const V2<double>*)e.S.p + off;
    if (dtype == 64)
      hipLaunchKernelGGL((k_row_c2r<double, double>), dim3(rb), dim3(256), rl, st, S, (double*)field, e.rg,
                         (const V2<double>*)e.twN.p, 1.0, (size_t)0, (size_t)0, nullptr);
    else
      hipLaunchKernelGGL((k_row_c2r<double, float>), dim3(rb), dim3(256), rl, st, S, (float*)field, e.rg,
                         (const V2<double>*)e.twN.p, 1.0, (size_t)0, (size_t)0, nullptr);
  } else {
    const V2<float>* S = (const V2<float>*)e.S.p + off;
    if (dtype == 64)
      hipLaunchKernelGGL((k_row_c2r<float, double>), dim3(rb), dim3(256), rl, st, S, (double*)field, e.rg,
                         (const V2<float>*)e.twN.p, 1.0f, (size_t)0, (size_t)0, nullptr);
    else
      hipLaunchKernelGGL((k_row_c2r<float, float>), dim3(rb), dim3(256), rl, st, S, (float*)field, e.rg,
                         (const V2<float>*)e.twN.p, 1.0f, (size_t)0, (size_t)0, nullptr);
  }
  hip_check(hipGetLastError(), "get_field launch");
}

// fields of the active slots, in active order, into out[nact][Ny][Nx]
template <typename T, typename Tout>
void launch_get_fields(Engine& e, Tout* out, hipStream_t st) {
  const dim3 grid(row_blocks(e.rg), e.nact());
  hipLaunchKernelGGL((k_row_c2r<T, Tout>), grid, dim3(256), row_lds(e.rg, e.csize), st, (const V2<T>*)e.S.p, out, e.rg,
                     (const V2<T>*)e.twN.p, T(1), e.nc, (size_t)e.p.Ny * e.p.Nx, (const int*)e.active.p);
}

void engine_get_fields(Engine& e, void* out, int dtype, hipStream_t st) {
  if (!out) fail(PFC_EINVAL, "NULL out");
  if (dtype != 64 && dtype != 32) fail(PFC_EINVAL, "dtype must be 64 or 32");
  for (int s : e.h_active)
    if (!e.field_set[s]) fail(PFC_EINVAL, "an active slot has no field");
  if (e.nact() == 0) return;
  if (e.f64()) {
    if (dtype == 64) launch_get_fields<double, double>(e, (double*)out, st);
    else launch_get_fields<double, float>(e, (float*)out, st);
  } else {
    if (dtype == 64) launch_get_fields<float, double>(e, (double*)out, st);
    else launch_get_fields<float, float>(e, (float*)out, st);
  }
  hip_check(hipGetLastError(), "get_fields launch");
}

// spectral diagnostics of the active slots (k_diagnose): out[nact][4 + 2 nent]
constexpr int kMaxDiagEntries = 256;
size_t diag_lds(int Nx, int Ny) { return 2 * (size_t)Ny * (Nx / 2 + 1) * sizeof(V2<double>) + 3 * 256 * sizeof(DD); }

void engine_diagnose(Engine& e, const int* ky, const int* kx, int nent, double* out, hipStream_t st) {
  if (!out || (nent > 0 && (!ky || !kx))) fail(PFC_EINVAL, "NULL argument");
  if (nent < 0 || nent > kMaxDiagEntries) fail(PFC_EINVAL, "entry count out of range (0..256)");
  for (int i = 0; i < nent; ++i)
    if (ky[i] < 0 || ky[i] >= e.p.Ny || kx[i] < 0 || kx[i] >= e.Nc) fail(PFC_EINVAL, "entry outside the half spectrum");
  for (int s : e.h_active)
    if (!e.field_set[s]) fail(PFC_EINVAL, "an active slot has no field");
  const size_t lds = diag_lds(e.p.Nx, e.p.Ny);
  if (lds > (size_t)lds_limit()) fail(PFC_EUNSUPPORTED, "grid too large for the device diagnostics");
  const int n = e.nact();
  if (n == 0) return;
  const int per = 4 + 2 * nent;
  const size_t ibytes = 2 * (size_t)std::max(nent, 1) * sizeof(int), obytes = (size_t)n * per * sizeof(double);
  if (e.dgi.n < ibytes || e.dgo.n < obytes) {
    hip_check(hipStreamSynchronize(st), "sync before scratch growth");
    if (e.dgi.n < ibytes) e.dgi.alloc(ibytes);
    if (e.dgo.n < obytes) e.dgo.alloc(obytes);
  }
  std::vector<int> idx(2 * (size_t)std::max(nent, 1), 0);
  std::copy(ky, ky + nent, idx.begin());
  std::copy(kx, kx + nent, idx.begin() + nent);
  hip_check(hipMemcpyAsync(e.dgi.p, idx.data(), ibytes, hipMemcpyHostToDevice, st), "copy");
  const int* dky = (const int*)e.dgi.p;
  DiagParams dp{e.p.Nx, e.Nc, e.p.Ny, e.lay, e.cg64.plY, nent, dky, dky + nent, (const double*)e.kx2.p,
                (const double*)e.ky2.p};
  const Slots sl = e.slots_active();
  if (e.f64())
    hipLaunchKernelGGL(k_diagnose<double>, dim3(n), dim3(256), lds, st, (const V2<double>*)e.S.p, dp,
                       (const V2<double>*)e.twY64.p, sl, (double*)e.dgo.p);
  else
    hipLaunchKernelGGL(k_diagnose<float>, dim3(n), dim3(256), lds, st, (const V2<float>*)e.S.p, dp,
                       (const V2<double>*)e.twY64.p, sl, (double*)e.dgo.p);
  hip_check(hipGetLastError(), "diagnose launch");
  hip_check(hipMemcpyAsync(out, e.dgo.p, obytes, hipMemcpyDeviceToHost, st), "copy");
  hip_check(hipStreamSynchronize(st), "sync");  // idx and out must outlive the copies
}

// (epsilon, M) of one slot; rebuilds its tables on `st`
void engine_set_slot(Engine& e, int slot, double epsilon, double M, hipStream_t st) {
  e.check_slot(slot);
  if (!std::isfinite(epsilon) || !std::isfinite(M)) fail(PFC_EINVAL, "epsilon and M must be finite");
  e.h_eps[slot] = epsilon;
  e.h_M[slot] = M;
  hip_check(hipMemcpyAsync((double*)e.eps.p + slot, &e.h_eps[slot], sizeof(double), hipMemcpyHostToDevice, st), "copy");
  hip_check(hipMemcpyAsync((double*)e.Ms.p + slot, &e.h_M[slot], sizeof(double), hipMemcpyHostToDevice, st), "copy");
  engine_tables(e, slot, 1, st);
  reset_shift_fit(e, slot, st);
  // the host values must outlive the asynchronous copies
  hip_check(hipStreamSynchronize(st), "sync");
}

// Device-managed stabilising shift of every slot (logPFC), checked before
// each step; `on` = 0 returns M to the host (pfc_set_shift / set_slot).
void engine_set_auto_shift(Engine& e, int on, double floor_tol, double ceil_tol, hipStream_t st) {
  if (on < 0 || on > 2) fail(PFC_EINVAL, "on must be 0, 1 or 2");
  if (on && e.p.model != PFC_MODEL_LOG) fail(PFC_EINVAL, "the stabilising shift is a logPFC option");
  if (on && !(floor_tol >= 0 && ceil_tol >= 0)) fail(PFC_EINVAL, "tolerances must be >= 0");
  if (on && !(e.p.minPhi > 0)) fail(PFC_EINVAL, "minPhi must be positive");
  if (on == 2 && !e.auto_shift) fail(PFC_EINVAL, "on = 2 changes the tolerances of a device-managed shift that is on");
  hip_check(hipStreamSynchronize(st), "sync");
  if (on == 2) {
    // new tolerances, the fit kept; graph nodes carry the tolerances by value
    const ShiftParams nshp{e.p.g, e.p.minPhi, floor_tol, ceil_tol};
    if (std::memcmp(&nshp, &e.shp, sizeof nshp) != 0) e.drop_graphs();
    e.shp = nshp;
    return;
  }
  if (on && !e.shst.p) {
    e.shst.alloc((size_t)e.cap * sizeof(ShiftState));
    e.rebuildf.alloc((size_t)e.cap * sizeof(int));
    e.brk.alloc((size_t)e.cap * 2 * sizeof(int));
    e.rpart.alloc((size_t)e.cap * e.p.Ny * sizeof(RangePart));  // at most Ny row tiles per instance
  }
  if (on) {
    hip_check(hipMemset(e.shst.p, 0, e.shst.n), "memset");
    hip_check(hipMemset(e.rebuildf.p, 0, e.rebuildf.n), "memset");
    hip_check(hipMemset(e.brk.p, 0, e.brk.n), "memset");
  } else if (e.auto_shift) {
    // the host takes M back: its copy follows the device's
    hip_check(hipMemcpy(e.h_M.data(), e.Ms.p, (size_t)e.cap * sizeof(double), hipMemcpyDeviceToHost), "copy");
  }
  const ShiftParams nshp{e.p.g, e.p.minPhi, floor_tol, ceil_tol};
  if (std::memcmp(&nshp, &e.shp, sizeof nshp) != 0) e.drop_graphs();
  e.shp = nshp;
  const bool folded = e.fold_shift();
  e.auto_shift = on != 0;
  if (e.fold_shift() != folded) {
    // E1 with or without the folded shift, for every slot
    engine_tables(e, 0, e.cap, st);
    hip_check(hipStreamSynchronize(st), "sync");
  }
}

// out = {M, lo, hi} of `slot` (lo = hi = NaN before the first fit); synchronises
void engine_get_shift(Engine& e, int slot, double out[3], hipStream_t st) {
  e.check_slot(slot);
  double M = 0.0;
  hip_check(hipMemcpyAsync(&M, (const double*)e.Ms.p + slot, sizeof(double), hipMemcpyDeviceToHost, st), "copy");
  ShiftState w{NAN, NAN, 0};
  if (e.shst.p)
    hip_check(hipMemcpyAsync(&w, (const ShiftState*)e.shst.p + slot, sizeof(ShiftState), hipMemcpyDeviceToHost, st),
              "copy");
  hip_check(hipStreamSynchronize(st), "sync");
  out[0] = M;
  out[1] = w.fitted ? w.lo : NAN;
  out[2] = w.fitted ? w.hi : NAN;
}

void engine_set_slot_box(Engine& e, int slot, double Lx, double Ly, double dt, hipStream_t st) {
  e.check_slot(slot);
  if (!(std::isfinite(Lx) && Lx > 0) || !(std::isfinite(Ly) && Ly > 0)) fail(PFC_EINVAL, "Lx and Ly must be positive");
  if (!(std::isfinite(dt) && dt > 0)) fail(PFC_EINVAL, "dt must be positive");
  const std::vector<double> kx2 = k_squared(rfft_k(e.p.Nx, Lx / e.p.Nx)), ky2 = k_squared(fft_k(e.p.Ny, Ly / e.p.Ny));
  e.h_Lx[slot] = Lx;
  e.h_Ly[slot] = Ly;
  e.h_dt[slot] = dt;
  hip_check(hipMemcpyAsync((double*)e.kx2.p + (size_t)slot * e.Nc, kx2.data(), kx2.size() * sizeof(double),
                           hipMemcpyHostToDevice, st), "copy");
  hip_check(hipMemcpyAsync((double*)e.ky2.p + (size_t)slot * e.p.Ny, ky2.data(), ky2.size() * sizeof(double),
                           hipMemcpyHostToDevice, st), "copy");
  hip_check(hipMemcpyAsync((double*)e.dts.p + slot, &e.h_dt[slot], sizeof(double), hipMemcpyHostToDevice, st), "copy");
  engine_tables(e, slot, 1, st);
  // the host vectors must outlive the asynchronous copies
  hip_check(hipStreamSynchronize(st), "sync");
}

// cache of transform-only twiddles/geometry for pfc_rfft2 / pfc_irfft2
struct FftEntry {
  RowGeom rg{};
  ColGeom cg{};
  DevBuf twN, twY;
};
std::mutex g_fft_mu;
std::map<std::tuple<int, int, int, int>, std::unique_ptr<FftEntry>> g_fft;
// pfc_irfft2's working copy of the spectrum, one per (device, stream): a
// stream's calls are ordered on it, so calls on different streams never share
// one, and growth waits only for the stream that uses it
std::map<std::pair<int, hipStream_t>, std::unique_ptr<DevBuf>> g_fft_scratch;

FftEntry* fft_entry(int Ny, int Nx, int dtype) {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  auto key = std::make_tuple(dev, Ny, Nx, dtype);
  auto it = g_fft.find(key);
  if (it != g_fft.end()) return it->second.get();
  auto e = std::make_unique<FftEntry>();
  const size_t cs = dtype == 64 ? sizeof(V2<double>) : sizeof(V2<float>);
  e->rg = row_geom(Nx, Ny, cs);
  e->cg = col_geom(Nx, Ny, cs);
  auto tn = twiddles(Nx), ty = twiddles(Ny);
  if (dtype == 64) {
    e->twN.upload(tn);
    e->twY.upload(ty);
  } else {
    e->twN.upload(to_f32(tn));
    e->twY.upload(to_f32(ty));
  }
  FftEntry* raw = e.get();
  g_fft[key] = std::move(e);
  return raw;
}

// External potential of one slot: V_hat = rfft2(V) (FP64) and its source
// term; V = null removes it.  The first potential switches the engine to the
// generic column kernel, which carries the term (the resident stepper does too).
void engine_set_external(Engine& e, int slot, const void* V, int dtype, hipStream_t st) {
  e.check_slot(slot);
  if (V && dtype != 64 && dtype != 32) fail(PFC_EINVAL, "dtype must be 64 or 32");
  const size_t per = (size_t)e.Nc * e.p.Ny;
  if (!V && !e.has_ext()) return;
  if (!e.has_ext()) {
    hip_check(hipStreamSynchronize(st), "sync before allocation");
    e.vhat.alloc((size_t)e.cap * per * sizeof(V2<double>));
    e.extc.alloc((size_t)e.cap * per * sizeof(V2<double>));
    hip_check(hipMemset(e.vhat.p, 0, e.vhat.n), "memset");
    hip_check(hipMemset(e.extc.p, 0, e.extc.n), "memset");
    e.h_ext.assign(e.cap, 0);
    e.fcol.on = false;
  }
  V2<double>* vh = (V2<double>*)e.vhat.p + (size_t)slot * per;
  if (!V) {
    hip_check(hipMemsetAsync(vh, 0, per * sizeof(V2<double>), st), "memset");
  } else {
    std::lock_guard<std::mutex> lk(g_fft_mu);
    FftEntry* fe = fft_entry(e.p.Ny, e.p.Nx, 64);
    const size_t nreal = (size_t)e.p.Ny * e.p.Nx;
    if (dtype == 64)
      hipLaunchKernelGGL((k_row_r2c<double, double>), dim3(row_blocks(fe->rg)), dim3(256), row_lds(fe->rg, 16), st,
                         (const double*)V, vh, fe->rg, (const V2<double>*)fe->twN.p, 1.0, nreal, per);
    else
      hipLaunchKernelGGL((k_row_r2c<double, float>), dim3(row_blocks(fe->rg)), dim3(256), row_lds(fe->rg, 16), st,
                         (const float*)V, vh, fe->rg, (const V2<double>*)fe->twN.p, 1.0, nreal, per);
    hipLaunchKernelGGL((k_col_fft<double, false>), dim3(col_blocks(fe->cg)), dim3(256), col_lds(fe->cg, 16), st, vh,
                       fe->cg, (const V2<double>*)fe->twY.p, 1.0, per);
    hip_check(hipGetLastError(), "potential transform launch");
  }
  e.h_ext[slot] = V ? 1 : 0;
  engine_tables(e, slot, 1, st);
  hip_check(hipStreamSynchronize(st), "sync");
}

void engine_set_active(Engine& e, const int* slots, int n, hipStream_t st) {
  if (n < 0 || n > e.cap) fail(PFC_EINVAL, "active count out of range");
  if (n > 0 && !slots) fail(PFC_EINVAL, "NULL slot list");
  std::vector<int> seen(e.cap, 0), list(slots, slots + n);
  for (int s : list) {
    e.check_slot(s);
    if (seen[s]++) fail(PFC_EINVAL, "duplicate slot in the active list");
  }
  e.h_active = list;
  // the in-step decision's flags: between steps every word has been read, so
  // clearing them all restarts a slot rejoining the list at either parity
  if (e.brk.p) hip_check(hipMemsetAsync(e.brk.p, 0, e.brk.n, st), "memset");
  if (n) hip_check(hipMemcpyAsync(e.active.p, e.h_active.data(), n * sizeof(int), hipMemcpyHostToDevice, st), "copy");
  hip_check(hipStreamSynchronize(st), "sync");
}

// reductions of the active list into out[4 i ..] (host); synchronises
void engine_reduce_active(Engine& e, double* out, hipStream_t st) {
  const int ny = e.nact();
  if (ny == 0) return;
  const Slots sl = e.slots_active();
  if (e.f64())
    launch_reduce<double>(e, e.S.p, (const double*)e.eps.p, (const double*)e.kx2.p, (const double*)e.ky2.p, st, sl,
                          ny, nullptr);
  else
    launch_reduce<float>(e, e.S.p, (const double*)e.eps.p, (const double*)e.kx2.p, (const double*)e.ky2.p, st, sl,
                         ny, nullptr);
  hip_check(hipMemcpyAsync(e.host_out, e.out.p, (size_t)ny * 4 * sizeof(double), hipMemcpyDeviceToHost, st), "copy");
  hip_check(hipStreamSynchronize(st), "sync");
  std::memcpy(out, e.host_out, (size_t)ny * 4 * sizeof(double));
}

void engine_reduce_one(Engine& e, int slot, double out[4], hipStream_t st) {
  e.check_slot(slot);
  reduce_slot(e, slot, st, false);
  hip_check(hipMemcpyAsync(e.host_out, e.out.p, 4 * sizeof(double), hipMemcpyDeviceToHost, st), "copy");
  hip_check(hipStreamSynchronize(st), "sync");
  std::memcpy(out, e.host_out, 4 * sizeof(double));
}

// Tracker: `nsteps` steps of the active slots with the FP64 reductions
// (free energy per site, mean, min, max) of every active slot written to a
// device series after each `every` steps, then one copy to the host and one
// synchronisation: out[rec][i][4] for rec < nsteps / every and the i-th
// active slot.  Steps past the last full block run without a record.
void engine_run_tracked(Engine& e, long long nsteps, long long every, double* out, hipStream_t st) {
  if (nsteps < 0) fail(PFC_EINVAL, "nsteps < 0");
  if (every < 1) fail(PFC_EINVAL, "every must be >= 1");
  if (!out && nsteps >= every) fail(PFC_EINVAL, "NULL out");
  const int ny = e.nact();
  const long long nrec = nsteps / every;
  const size_t bytes = (size_t)nrec * ny * 4 * sizeof(double);
  if (e.track.n < bytes) {
    hip_check(hipStreamSynchronize(st), "sync before tracker growth");
    e.track.alloc(bytes);
  }
  // graphs pay for their capture over the whole call, not one block
  const int capture = nsteps >= kGraphCaptureSteps ? 1 : 0;
  const Slots sl = e.slots_active();
  for (long long r = 0; r < nrec; ++r) {
    engine_step(e, every, st, capture);
    if (ny == 0) continue;
    double* dst = (double*)e.track.p + (size_t)r * ny * 4;
    if (e.f64())
      launch_reduce<double>(e, e.S.p, (const double*)e.eps.p, (const double*)e.kx2.p, (const double*)e.ky2.p, st,
                            sl, ny, nullptr, dst);
    else
      launch_reduce<float>(e, e.S.p, (const double*)e.eps.p, (const double*)e.kx2.p, (const double*)e.ky2.p, st,
                           sl, ny, nullptr, dst);
  }
  engine_step(e, nsteps - nrec * every, st, capture);
  if (bytes) hip_check(hipMemcpyAsync(out, e.track.p, bytes, hipMemcpyDeviceToHost, st), "copy");
  hip_check(hipStreamSynchronize(st), "sync");
}

void engine_status(Engine& e, int* flags, hipStream_t st) {
  hip_check(hipMemcpyAsync(e.host_status, e.status.p, e.cap * sizeof(int), hipMemcpyDeviceToHost, st), "copy");
  hip_check(hipStreamSynchronize(st), "sync");
  std::memcpy(flags, e.host_status, e.cap * sizeof(int));
  hip_check(hipMemsetAsync(e.status.p, 0, e.cap * sizeof(int), st), "clear status");
}

template <typename T>
void launch_filter(Engine& e, int slot, int mode, void* out0, void* out1, int dtype, hipStream_t st) {
  using V = V2<T>;
  if (e.diag.n < 2 * e.nc * sizeof(V)) {
    hip_check(hipStreamSynchronize(st), "sync before scratch growth");
    e.diag.alloc(2 * e.nc * sizeof(V));
  }
  V* d0 = (V*)e.diag.p;
  V* d1 = d0 + e.nc;
  const V* S = (const V*)e.S.p + (size_t)slot * e.nc;
  V* Tb = (V*)e.Tb.p + (size_t)slot * e.nc;
  const V* twN = (const V*)e.twN.p;
  const V* twY = (const V*)e.twY.p;
  // the shift as the step sees it: the device copy, which k_shift_update
  // refits without touching the host's
  FilterParams fp{e.p.beta, e.p.q0, e.h_eps[slot], 6.283185307179586, (const double*)e.Ms.p + slot, e.p.Nx,
                  (const double*)e.kx2.p + (size_t)slot * e.Nc, (const double*)e.ky2.p + (size_t)slot * e.p.Ny};
  const int rb = row_blocks(e.rg), cb = col_blocks(e.cg);
  const size_t rl = row_lds(e.rg, sizeof(V)), cl = col_lds(e.cg, sizeof(V));
  if (mode == FILTER_MU) {
    Slots one = e.slots_one(slot);
    hipLaunchKernelGGL((k_row_nonlin<T>), dim3(rb), dim3(256), rl, st, S, Tb, e.rg, nonlin_of(e), twN, one);
    hipLaunchKernelGGL((k_col_filter<T, FILTER_MU>), dim3(cb), dim3(256), cl, st, S, Tb, d0, d1, e.cg, fp, twY);
  } else {
    hipLaunchKernelGGL((k_col_filter<T, FILTER_LOCAL>), dim3(cb), dim3(256), cl, st, S, Tb, d0, d1, e.cg, fp, twY);
  }
  const T scale = T(1.0 / ((double)e.p.Nx * (double)e.p.Ny));
  const int nout = mode == FILTER_MU ? 2 : 1;
  for (int o = 0; o < nout; ++o) {
    const V* src = o == 0 ? d0 : d1;
    void* dst = o == 0 ? out0 : out1;
    if (dtype == 64)
      hipLaunchKernelGGL((k_row_c2r<T, double>), dim3(rb), dim3(256), rl, st, src, (double*)dst, e.rg, twN, scale,
                         (size_t)0, (size_t)0, nullptr);
    else
      hipLaunchKernelGGL((k_row_c2r<T, float>), dim3(rb), dim3(256), rl, st, src, (float*)dst, e.rg, twN, scale,
                         (size_t)0, (size_t)0, nullptr);
  }
  hip_check(hipGetLastError(), "diagnostic launch");
}

size_t engine_bytes(const Engine& e) {
  size_t b = 0;
  for (const DevBuf* d : {&e.S, &e.Tb, &e.D, &e.E1, &e.Q1, &e.Q2, &e.E1f, &e.Q1f, &e.Q2f, &e.twN, &e.twY, &e.twN64, &e.twY64, &e.kx2, &e.ky2,
                          &e.dts, &e.part, &e.out, &e.pin, &e.status, &e.eps, &e.Ms, &e.active, &e.diag, &e.dgi, &e.dgo, &e.vhat,
                          &e.extc, &e.cnt, &e.shst, &e.rebuildf, &e.brk, &e.rpart})
    b += d->n;
  return b;
}


}  // namespace

struct pfc_plan {
  Engine e;
};

struct pfc_batch {
  Engine e;
};

// ---------------------------------------------------------------------------
// groups: batches of different grid sizes stepped by shared launches
// (k_row_group / k_col_group), one per pass and register class with members.
// The descriptor lists live in one device array allocated at creation; they
// depend only on each member's active count (everything else they hold is
// fixed for a batch's life), so they are rebuilt, after the stream has
// drained, only when an active count changes.
// ---------------------------------------------------------------------------
struct GroupLaunch {
  int nmem = 0, blocks = 0;
  size_t lds = 0, off = 0;  // off: first descriptor of this launch in the array
};

struct pfc_group {
  std::vector<pfc_batch*> mem;  // not owned
  std::vector<int> order;       // members by decreasing instance size (launch order)
  int precision = 0, scheme = 0;
  bool nt = false;              // output stores of the per-pass launches: non-temporal (else ordinary)
  DevBuf desc;                  // GroupMember<T>[chunks][2 passes][2 classes][members]
  std::vector<long long> built; // active counts the descriptors hold
  // Members are stepped in chunks whose arrays and tables fit `budget` bytes
  // (sized to the Infinity Cache): a step graph runs each chunk's 32 steps in
  // turn, so a chunk's working set can stay cached across its steps (chunks
  // are independent, so the order changes no result); 0 = one chunk
  size_t budget = 0;
  struct Chunk {
    GroupLaunch rows[kGroupClasses], cols[kGroupClasses];
    // every member has its row and column transforms in one class, and both
    // classes have members: the two classes' launch chains are independent,
    // and a step graph runs them as two parallel branches
    bool split = false;
  };
  std::vector<Chunk> chunks;
  std::map<std::vector<long long>, hipGraphExec_t> graphs;
  hipStream_t cap_st = nullptr, cap_st2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  long long captures = 0, replays = 0;
  // Persistent form (pfc_group_set_flow): each register class's members step
  // in one k_flow launch per `flow_steps` steps (0: the per-pass launches
  // above).  Per class: the members' row and column descriptors (fdesc),
  // each member's first instance counter (fidx), the step's ticket entries
  // (ftab), the ticket, the per-instance counters and the abort flag (fcnt);
  // the two classes run on the caller's stream and flow_st2 side by side.
  int flow_steps = 0;
  struct FlowClass {
    int nmem = 0, per_step = 0, ninst = 0, blocks = 0;
    size_t lds = 0;
    size_t desc_off = 0, idx_off = 0, cnt_off = 0;  // bytes into fdesc; ints into fidx; words into fcnt
    size_t tab_off = 0;                               // entries into ftab
    size_t abort_off = 0;                             // word of the class's abort flag in fcnt
  };
  FlowClass flow[kGroupClasses];
  DevBuf fdesc, fidx, fcnt, ftab;
  std::vector<long long> flow_built;
  hipStream_t flow_st2 = nullptr;
  hipEvent_t flow_fork = nullptr, flow_join = nullptr;
  long long flow_launches = 0;
  ~pfc_group() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (cap_st) (void)hipStreamDestroy(cap_st);
    if (cap_st2) (void)hipStreamDestroy(cap_st2);
    if (flow_fork) (void)hipEventDestroy(flow_fork);
    if (flow_join) (void)hipEventDestroy(flow_join);
    if (flow_st2) (void)hipStreamDestroy(flow_st2);
  }
};

namespace {

void group_check_member(const Engine& e) {
  int P = 0;
  if (e.resident)
    fail(PFC_EUNSUPPORTED, "a group member runs the resident stepper (small ETD1 grid): step it on its own");
  if (e.has_ext()) fail(PFC_EUNSUPPORTED, "a group member has an external potential");
  if (e.auto_shift) fail(PFC_EUNSUPPORTED, "a group member manages its stabilising shift on the device");
  if (group_case(e.p.Nx / 2, &P) < 0 || group_case(e.p.Ny, &P) < 0)
    fail(PFC_EUNSUPPORTED, "group members need Nx/2 and Ny in {32, 48, 64, 96, 128, 192, 256, 384, 512}");
}

std::vector<long long> group_counts(const pfc_group& g) {
  std::vector<long long> k;
  for (const pfc_batch* b : g.mem) k.push_back(b->e.nact());
  return k;
}

// bytes one step of member `e` works on: its active slots' state arrays
// (S~, T and the ETDRK2 carry) and FP64 ETD tables
size_t group_working_set(const Engine& e) {
  const size_t k = e.p.scheme == PFC_SCHEME_ETDRK2 ? 3 : 2;
  return (size_t)e.nact() * k * (e.nc * e.csize + e.ntab * sizeof(double));
}

template <typename T>
void group_build(pfc_group& g, hipStream_t st) {
  const size_t n = g.mem.size();
  // chunks in launch order (largest instances first), greedily up to the budget
  std::vector<std::vector<int>> parts;
  size_t cur = 0;
  for (int i : g.order) {
    const Engine& e = g.mem[i]->e;
    if (e.nact() == 0) continue;
    const size_t ws = group_working_set(e);
    if (parts.empty() || (g.budget && cur > 0 && cur + ws > g.budget)) {
      parts.emplace_back();
      cur = 0;
    }
    parts.back().push_back(i);
    cur += ws;
  }
  g.chunks.assign(parts.size(), pfc_group::Chunk{});
  std::vector<GroupMember<T>> all(parts.size() * 2 * kGroupClasses * n);
  for (size_t k = 0; k < parts.size(); ++k) {
    pfc_group::Chunk& ch = g.chunks[k];
    ch.split = true;
    for (int i : parts[k]) {
      const Engine& e = g.mem[i]->e;
      int P = 0, rc = 0, cc = 0;
      group_case(e.p.Nx / 2, &P, &rc);
      group_case(e.p.Ny, &P, &cc);
      ch.split = ch.split && rc == cc;
    }
    for (int pass = 0; pass < 2; ++pass)
      for (int c = 0; c < kGroupClasses; ++c) {
        GroupLaunch& L = pass == 0 ? ch.rows[c] : ch.cols[c];
        L = GroupLaunch{};
        L.off = ((k * 2 + pass) * kGroupClasses + c) * n;
        for (int i : parts[k]) {
          Engine& e = g.mem[i]->e;
          const int ny = e.nact();
          const int len = pass == 0 ? e.p.Nx / 2 : e.p.Ny;
          int P = 0, cls = 0;
          const int kcase = group_case(len, &P, &cls);
          if (cls != c) continue;
          GroupMember<T> m{};
          m.S = (V2<T>*)e.S.p;
          m.Tb = (V2<T>*)e.Tb.p;
          m.D = (V2<T>*)e.D.p;
          m.tw = (const V2<double>*)(pass == 0 ? e.twN64.p : e.twY64.p);
          m.lay = e.lay;
          m.q = nonlin_of(e);
          m.sl = e.slots_active();
          m.tb = e.tables();
          m.Ny = e.p.Ny;
          m.Nc = e.Nc;
          m.kcase = kcase;
          const int per = kGroupThreads / (len / P);  // rows (columns) per block
          m.ntile = pass == 0 ? (e.p.Ny + per - 1) / per : (e.Nc + per - 1) / per;
          m.blk0 = L.blocks;
          L.blocks += m.ntile * ny;
          L.lds = std::max(L.lds, (size_t)per * (pass == 0 ? row_ld(len) : col_ld(len)) * sizeof(V2<T>));
          all[L.off + L.nmem++] = m;
        }
      }
    ch.split = ch.split && ch.rows[0].nmem > 0 && ch.rows[1].nmem > 0;
  }
  // launches (or graphs) still reading the old descriptors finish first
  hip_check(hipStreamSynchronize(st), "sync before group rebuild");
  if (!all.empty())
    hip_check(hipMemcpy(g.desc.p, all.data(), all.size() * sizeof(GroupMember<T>), hipMemcpyHostToDevice), "copy");
  g.built = group_counts(g);
}

template <typename T>
void group_launch(int kind, int cls, bool nt, int blocks, size_t lds, hipStream_t st, const GroupMember<T>* mem,
                  int nmem) {
  if constexpr (sizeof(T) == 8)
    group_launch_f64(kind, cls, nt, blocks, lds, st, mem, nmem);
  else
    group_launch_f32(kind, cls, nt, blocks, lds, st, mem, nmem);
}

// one pass (kind 0: rows, else the column mode kind - 1) of chunk `k` over
// class `cls` (-1: every class with members)
template <typename T>
void group_pass(pfc_group& g, size_t k, int kind, hipStream_t st, int cls = -1) {
  const auto* base = (const GroupMember<T>*)g.desc.p;
  const GroupLaunch* L = kind == 0 ? g.chunks[k].rows : g.chunks[k].cols;
  for (int c = 0; c < kGroupClasses; ++c)
    if (L[c].nmem && (cls < 0 || cls == c))
      group_launch<T>(kind, c, g.nt, L[c].blocks, L[c].lds, st, base + L[c].off, L[c].nmem);
}

template <typename T>
void group_loop(pfc_group& g, size_t k, long long nsteps, hipStream_t st, int cls = -1) {
  for (long long s = 0; s < nsteps; ++s) {
    group_pass<T>(g, k, 0, st, cls);
    if (g.scheme == PFC_SCHEME_ETD1) {
      group_pass<T>(g, k, 1 + COL_ETD1, st, cls);
    } else {
      group_pass<T>(g, k, 1 + COL_RK2_STAGE1, st, cls);
      group_pass<T>(g, k, 0, st, cls);
      group_pass<T>(g, k, 1 + COL_RK2_STAGE2, st, cls);
    }
  }
  hip_check(hipGetLastError(), "group step launch");
}

// nsteps steps of every chunk, chunk after chunk
template <typename T>
void group_loop_chunks(pfc_group& g, long long nsteps, hipStream_t st) {
  for (size_t k = 0; k < g.chunks.size(); ++k) group_loop<T>(g, k, nsteps, st);
}

// kGraphSteps steps of every chunk captured on g.cap_st, chunk after chunk;
// within a chunk with independent classes, class 1's chain forks onto a
// second stream and joins at the chunk's end (two parallel graph branches)
template <typename T>
void group_capture_steps(pfc_group& g) {
  for (size_t k = 0; k < g.chunks.size(); ++k) {
    if (!g.chunks[k].split) {
      group_loop<T>(g, k, kGraphSteps, g.cap_st);
      continue;
    }
    if (!g.cap_st2) hip_check(hipStreamCreateWithFlags(&g.cap_st2, hipStreamNonBlocking), "hipStreamCreate");
    if (!g.fork) hip_check(hipEventCreateWithFlags(&g.fork, hipEventDisableTiming), "hipEventCreate");
    if (!g.join) hip_check(hipEventCreateWithFlags(&g.join, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventRecord(g.fork, g.cap_st), "hipEventRecord");
    hip_check(hipStreamWaitEvent(g.cap_st2, g.fork, 0), "hipStreamWaitEvent");
    group_loop<T>(g, k, kGraphSteps, g.cap_st2, 1);
    group_loop<T>(g, k, kGraphSteps, g.cap_st, 0);
    hip_check(hipEventRecord(g.join, g.cap_st2), "hipEventRecord");
    hip_check(hipStreamWaitEvent(g.cap_st, g.join, 0), "hipStreamWaitEvent");
  }
}

// ---- persistent form: k_flow -------------------------------------------------
// every member's row and column lengths in one register class (a member's
// chain of passes then lives in one launch)
void flow_check(const pfc_group& g) {
  for (const pfc_batch* b : g.mem) {
    int P = 0, rc = 0, cc = 0;
    group_case(b->e.p.Nx / 2, &P, &rc);
    group_case(b->e.p.Ny, &P, &cc);
    if (rc != cc)
      fail(PFC_EUNSUPPORTED, "persistent group step: a member's row (Nx/2) and column (Ny) lengths are in different "
                             "register classes");
  }
}

template <typename T>
void flow_build(pfc_group& g, hipStream_t st) {
  const size_t n = g.mem.size();
  const int nph = g.scheme == PFC_SCHEME_ETD1 ? 2 : 4;
  const bool rk2 = nph == 4;
  int dev = 0, ncu = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "device attribute");
  std::vector<GroupMember<T>> desc(kGroupClasses * 2 * n);
  std::vector<int> idx(kGroupClasses * (n + 1));
  std::vector<FlowEntry> tab;
  size_t cnt_words = 0;
  for (int c = 0; c < kGroupClasses; ++c) {
    pfc_group::FlowClass& F = g.flow[c];
    F = pfc_group::FlowClass{};
    F.desc_off = (size_t)c * 2 * n * sizeof(GroupMember<T>);
    F.idx_off = (size_t)c * (n + 1);
    std::vector<int> mems;
    for (int i : g.order) {
      const Engine& e = g.mem[i]->e;
      int P = 0, rc = 0;
      group_case(e.p.Nx / 2, &P, &rc);
      if (rc == c && e.nact() > 0) mems.push_back(i);
    }
    F.nmem = (int)mems.size();
    if (F.nmem == 0) continue;
    GroupMember<T>* rows = desc.data() + (size_t)c * 2 * n;
    GroupMember<T>* cols = rows + n;
    int* cnt0 = idx.data() + F.idx_off;
    cnt0[0] = 1;  // word 0 is the ticket
    for (int k = 0; k < F.nmem; ++k) {
      Engine& e = g.mem[mems[k]]->e;
      for (int pass = 0; pass < 2; ++pass) {
        const int len = pass == 0 ? e.p.Nx / 2 : e.p.Ny;
        int P = 0;
        GroupMember<T> m{};
        m.S = (V2<T>*)e.S.p;
        m.Tb = (V2<T>*)e.Tb.p;
        m.D = (V2<T>*)e.D.p;
        m.tw = (const V2<double>*)(pass == 0 ? e.twN64.p : e.twY64.p);
        m.lay = e.lay;
        m.q = nonlin_of(e);
        m.sl = e.slots_active();
        m.tb = e.tables();
        m.Ny = e.p.Ny;
        m.Nc = e.Nc;
        m.kcase = group_case(len, &P);
        const int per = kGroupThreads / (len / P);  // rows (columns) per block
        m.ntile = pass == 0 ? (e.p.Ny + per - 1) / per : (e.Nc + per - 1) / per;
        m.blk0 = 0;
        F.lds = std::max(F.lds, (size_t)per * (pass == 0 ? row_ld(len) : col_ld(len)) * sizeof(V2<T>));
        (pass == 0 ? rows : cols)[k] = m;
      }
      cnt0[k + 1] = cnt0[k] + e.nact();
    }
    long long at = 0;  // tickets per step
    for (int p = 0; p < nph; ++p)
      for (int k = 0; k < F.nmem; ++k)
        at += (long long)g.mem[mems[k]]->e.nact() * ((p & 1) == 0 ? rows[k].ntile : cols[k].ntile);
    if (at > (1LL << 30)) fail(PFC_EUNSUPPORTED, "persistent group step: too many tiles per step");
    F.per_step = (int)at;
    // the step's tickets in work-list order
    F.tab_off = tab.size();
    for (int p = 0; p < nph; ++p)
      for (int k = 0; k < F.nmem; ++k) {
        const int nr = rows[k].ntile, nc = cols[k].ntile;
        const int ntile = (p & 1) == 0 ? nr : nc;
        for (int yi = 0; yi < cnt0[k + 1] - cnt0[k]; ++yi)
          for (int t = 0; t < ntile; ++t) {
            FlowEntry en{};
            en.p = p;
            en.m = k;
            en.yi = yi;
            en.tile = t;
            en.cidx = cnt0[k] + yi;
            en.pre = (unsigned)((p >> 1) * (nr + nc) + (p & 1) * nr);
            en.per_inst = (unsigned)(nph / 2 * (nr + nc));
            tab.push_back(en);
          }
      }
    if ((long long)(tab.size() - F.tab_off) != at) fail(PFC_EUNSUPPORTED, "flow work list size mismatch");
    F.ninst = cnt0[F.nmem] - 1;
    F.cnt_off = cnt_words;
    cnt_words += (size_t)(1 + F.ninst) * kFlowCntStride;
    const int occ = g.precision == 64 ? flow_occupancy_f64(c, rk2, F.lds) : flow_occupancy_f32(c, rk2, F.lds);
    if (occ < 1) fail(PFC_EUNSUPPORTED, "persistent group step: the kernel does not fit a CU");
    F.blocks = occ * ncu;
  }
  // launches still reading the old plan finish first
  hip_check(hipStreamSynchronize(st), "sync before flow rebuild");
  if (g.fdesc.n < desc.size() * sizeof(GroupMember<T>)) g.fdesc.alloc(desc.size() * sizeof(GroupMember<T>));
  hip_check(hipMemcpy(g.fdesc.p, desc.data(), desc.size() * sizeof(GroupMember<T>), hipMemcpyHostToDevice), "copy");
  g.fidx.upload(idx);
  g.ftab.upload(tab);
  // counters, then one abort flag per class on its own 256-byte line
  for (int c = 0; c < kGroupClasses; ++c) g.flow[c].abort_off = cnt_words + (size_t)c * kFlowCntStride;
  const size_t words = cnt_words + (size_t)kGroupClasses * kFlowCntStride;
  // all of it zero: the new layout's abort flags may sit where the old
  // layout kept counters
  if (g.fcnt.n < words * sizeof(unsigned)) g.fcnt.alloc(words * sizeof(unsigned));
  hip_check(hipMemset(g.fcnt.p, 0, g.fcnt.n), "memset");
  g.flow_built = group_counts(g);
}

template <typename T>
void flow_launch(pfc_group& g, int c, long long steps, hipStream_t st) {
  const pfc_group::FlowClass& F = g.flow[c];
  const size_t n = g.mem.size();
  const int nph = g.scheme == PFC_SCHEME_ETD1 ? 2 : 4;
  if ((unsigned long long)steps * (unsigned long long)F.per_step >= (1ULL << 31))
    fail(PFC_EUNSUPPORTED, "persistent group step: too many tiles per launch");
  FlowPlan<T> f{};
  const GroupMember<T>* base = (const GroupMember<T>*)((const char*)g.fdesc.p + F.desc_off);
  f.rows = base;
  f.cols = base + n;
  f.cnt0 = (const int*)g.fidx.p + F.idx_off;
  f.tab = (const FlowEntry*)g.ftab.p + F.tab_off;
  unsigned* words = (unsigned*)g.fcnt.p;
  f.cnt = words + F.cnt_off;
  f.ticket = f.cnt;
  f.abort = (int*)(words + F.abort_off);
  f.nmem = F.nmem;
  f.nph = nph;
  f.per_step = F.per_step;
  f.total = (unsigned)(steps * F.per_step);
  f.nwords = (1 + F.ninst) * kFlowCntStride;
  // PFC_HIP_FLOW_SPIN: polls of one dependency wait before the launch aborts
  // (tests force the abort path with 0; the default bounds a wait to ~1 s)
  const char* spin = std::getenv("PFC_HIP_FLOW_SPIN");
  f.spin_limit = spin ? std::max(0, std::atoi(spin)) : kFlowSpinLimit;
  const int blocks = (int)std::max(1LL, std::min<long long>(F.blocks, (long long)f.total));
  if constexpr (sizeof(T) == 8)
    flow_launch_f64(c, nph == 4, blocks, F.lds, st, f);
  else
    flow_launch_f32(c, nph == 4, blocks, F.lds, st, f);
  ++g.flow_launches;
}

// nsteps steps in launches of at most flow_steps steps per class; the two
// classes' launch chains are independent and run side by side (class 1 on
// flow_st2, forked from and joined back to `st`)
template <typename T>
void flow_step(pfc_group& g, long long nsteps, hipStream_t st) {
  const std::vector<long long> counts = group_counts(g);
  if (counts != g.flow_built) flow_build<T>(g, st);
  if (nsteps == 0) return;
  const bool two = g.flow[0].nmem > 0 && g.flow[1].nmem > 0;
  hipStream_t st1 = st;
  if (two) {
    if (!g.flow_st2) hip_check(hipStreamCreateWithFlags(&g.flow_st2, hipStreamNonBlocking), "hipStreamCreate");
    if (!g.flow_fork) hip_check(hipEventCreateWithFlags(&g.flow_fork, hipEventDisableTiming), "hipEventCreate");
    if (!g.flow_join) hip_check(hipEventCreateWithFlags(&g.flow_join, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventRecord(g.flow_fork, st), "hipEventRecord");
    hip_check(hipStreamWaitEvent(g.flow_st2, g.flow_fork, 0), "hipStreamWaitEvent");
    st1 = g.flow_st2;
  }
  for (long long done = 0; done < nsteps; done += g.flow_steps) {
    const long long k = std::min<long long>(g.flow_steps, nsteps - done);
    if (g.flow[0].nmem) flow_launch<T>(g, 0, k, st);
    if (g.flow[1].nmem) flow_launch<T>(g, 1, k, st1);
  }
  if (two) {
    hip_check(hipEventRecord(g.flow_join, g.flow_st2), "hipEventRecord");
    hip_check(hipStreamWaitEvent(st, g.flow_join, 0), "hipStreamWaitEvent");
  }
  hip_check(hipGetLastError(), "flow launch");
}

template <typename T>
void group_step(pfc_group& g, long long nsteps, hipStream_t st) {
  if (nsteps < 0) fail(PFC_EINVAL, "nsteps < 0");
  for (const pfc_batch* b : g.mem) {
    group_check_member(b->e);
    for (int s : b->e.h_active)
      if (!b->e.field_set[s]) fail(PFC_EINVAL, "a stepped slot has no field (pfc_batch_set_field)");
  }
  if (g.flow_steps > 0) {
    flow_step<T>(g, nsteps, st);
    return;
  }
  const std::vector<long long> counts = group_counts(g);
  if (counts != g.built) group_build<T>(g, st);
  if (g.chunks.empty() || nsteps == 0) return;
  if (nsteps >= kGraphSteps && graphs_enabled()) {
    hipGraphExec_t x = nullptr;
    auto it = g.graphs.find(counts);
    if (it != g.graphs.end()) {
      x = it->second;
    } else if (nsteps >= kGraphCaptureSteps) {
      if (!g.cap_st) hip_check(hipStreamCreateWithFlags(&g.cap_st, hipStreamNonBlocking), "hipStreamCreate");
      hip_check(hipStreamBeginCapture(g.cap_st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      hipGraph_t gr = nullptr;
      try {
        group_capture_steps<T>(g);
      } catch (...) {
        (void)hipStreamEndCapture(g.cap_st, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        throw;
      }
      hip_check(hipStreamEndCapture(g.cap_st, &gr), "hipStreamEndCapture");
      const hipError_t ie = hipGraphInstantiate(&x, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      hip_check(ie, "hipGraphInstantiate");
      g.graphs[counts] = x;
      ++g.captures;
    }
    if (x) {
      const long long reps = nsteps / kGraphSteps;
      for (long long r = 0; r < reps; ++r) hip_check(hipGraphLaunch(x, st), "hipGraphLaunch");
      g.replays += reps;
      nsteps -= reps * kGraphSteps;
    }
  }
  group_loop_chunks<T>(g, nsteps, st);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
// ---- an irfft2 that allocates nothing -----------------------------------------
// pfc_irfft2_noalloc keeps its twiddles and its working copy of the spectrum in
// this file-scope device array instead of library allocations.  Each device
// holds its own copy until the process ends; host code reaches it only through
// the address hipGetSymbolAddress gives on that device (the array's name is a
// host-side stand-in there, not a device address).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;

struct NoAllocScratch {
  std::once_flag once;
  unsigned char* ptr = nullptr;
  hipError_t err = hipSuccess;
};
NoAllocScratch g_noalloc[kNoAllocMaxDevices];

// the current device's copy of `scratch`, looked up once per device
unsigned char* noalloc_scratch() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kNoAllocMaxDevices) fail(PFC_EUNSUPPORTED, "device index beyond the static scratch table");
  NoAllocScratch& d = g_noalloc[dev];
  std::call_once(d.once, [&] {
    void* p = nullptr;
    d.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    d.ptr = static_cast<unsigned char*>(p);
  });
  if (d.err != hipSuccess || !d.ptr)
    fail(PFC_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(d.err));
  return d.ptr;
}

// tw[m] = exp(-2 pi i m / n), the quarter points exact; rounded once to T
template <typename T>
__global__ void __launch_bounds__(256) k_twiddle_table(V2<T>* __restrict__ tw, int n) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double c, sn;
  if ((4LL * m) % n == 0) {
    const int q = (int)((4LL * m) / n);
    c = q == 0 ? 1.0 : (q == 2 ? -1.0 : 0.0);
    sn = q == 1 ? 1.0 : (q == 3 ? -1.0 : 0.0);
  } else {
    sincospi(2.0 * (double)m / (double)n, &sn, &c);
  }
  tw[m] = V2<T>{T(c), T(-sn)};
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

extern "C" {

const char* pfc_last_error(void) { return g_err.c_str(); }

int pfc_version(void) { return 2; }

int pfc_create(const pfc_params* params, pfc_plan** out) {
  return guarded([&] {
    if (!out) fail(PFC_EINVAL, "out is NULL");
    *out = nullptr;
    auto pl = std::make_unique<pfc_plan>();
    engine_init(pl->e, params, 1, false);
    engine_tables(pl->e, 0, 1, nullptr);
    hip_check(hipStreamSynchronize(nullptr), "sync");
    *out = pl.release();
  });
}

int pfc_destroy(pfc_plan* plan) {
  return guarded([&] { delete plan; });
}

int pfc_set_field(pfc_plan* pl, const void* field, int dtype, void* stream) {
  return guarded([&] {
    if (!pl) fail(PFC_EINVAL, "NULL plan");
    engine_set_field(pl->e, 0, field, dtype, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_get_field(pfc_plan* pl, void* field, int dtype, void* stream) {
  return guarded([&] {
    if (!pl) fail(PFC_EINVAL, "NULL plan");
    engine_get_field(pl->e, 0, field, dtype, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_step(pfc_plan* pl, long long nsteps, void* stream) {
  return guarded([&] {
    if (!pl) fail(PFC_EINVAL, "NULL plan");
    engine_step(pl->e, nsteps, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_reduce(pfc_plan* pl, double out[4], void* stream) {
  return guarded([&] {
    if (!pl || !out) fail(PFC_EINVAL, "NULL plan or out");
    engine_reduce_one(pl->e, 0, out, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_set_shift(pfc_plan* pl, double M) {
  return guarded([&] {
    if (!pl) fail(PFC_EINVAL, "NULL plan");
    hip_check(hipDeviceSynchronize(), "sync before table rebuild");
    engine_set_slot(pl->e, 0, pl->e.h_eps[0], M, nullptr);
  });
}

int pfc_set_auto_shift(pfc_plan* pl, int on, double floor_tol, double ceil_tol, void* stream) {
  return guarded([&] {
    if (!pl) fail(PFC_EINVAL, "NULL plan");
    engine_set_auto_shift(pl->e, on, floor_tol, ceil_tol, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_get_shift(pfc_plan* pl, double out[3], void* stream) {
  return guarded([&] {
    if (!pl || !out) fail(PFC_EINVAL, "NULL plan or out");
    engine_get_shift(pl->e, 0, out, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_set_auto_shift(pfc_batch* b, int on, double floor_tol, double ceil_tol, void* stream) {
  return guarded([&] {
    if (!b) fail(PFC_EINVAL, "NULL batch");
    engine_set_auto_shift(b->e, on, floor_tol, ceil_tol, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_get_shift(pfc_batch* b, int slot, double out[3], void* stream) {
  return guarded([&] {
    if (!b || !out) fail(PFC_EINVAL, "NULL batch or out");
    engine_get_shift(b->e, slot, out, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_get_status(pfc_plan* pl, int* flags, void* stream) {
  return guarded([&] {
    if (!pl || !flags) fail(PFC_EINVAL, "NULL plan or flags");
    engine_status(pl->e, flags, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_chemical_potential(pfc_plan* pl, void* mu, void* local_mu, int dtype, void* stream) {
  return guarded([&] {
    if (!pl || !mu || !local_mu) fail(PFC_EINVAL, "NULL argument");
    if (dtype != 64 && dtype != 32) fail(PFC_EINVAL, "dtype must be 64 or 32");
    if (!pl->e.field_set[0]) fail(PFC_EINVAL, "pfc_set_field has not been called");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (pl->e.f64())
      launch_filter<double>(pl->e, 0, FILTER_MU, mu, local_mu, dtype, st);
    else
      launch_filter<float>(pl->e, 0, FILTER_MU, mu, local_mu, dtype, st);
  });
}

int pfc_local_field(pfc_plan* pl, void* out, int dtype, void* stream) {
  return guarded([&] {
    if (!pl || !out) fail(PFC_EINVAL, "NULL argument");
    if (dtype != 64 && dtype != 32) fail(PFC_EINVAL, "dtype must be 64 or 32");
    if (!pl->e.field_set[0]) fail(PFC_EINVAL, "pfc_set_field has not been called");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (pl->e.f64())
      launch_filter<double>(pl->e, 0, FILTER_LOCAL, out, nullptr, dtype, st);
    else
      launch_filter<float>(pl->e, 0, FILTER_LOCAL, out, nullptr, dtype, st);
  });
}

int pfc_set_external_potential(pfc_plan* pl, const void* V, int dtype, void* stream) {
  return guarded([&] {
    if (!pl) fail(PFC_EINVAL, "NULL plan");
    engine_set_external(pl->e, 0, V, dtype, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_diagnose(pfc_plan* pl, const int* ky, const int* kx, int nent, double* out, void* stream) {
  return guarded([&] {
    if (!pl) fail(PFC_EINVAL, "NULL plan");
    engine_diagnose(pl->e, ky, kx, nent, out, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_step_profile(pfc_plan* pl, long long nsteps, void* stream, double ms[3], long long calls[3]) {
  return guarded([&] {
    if (!pl || !ms || !calls) fail(PFC_EINVAL, "NULL argument");
    if (nsteps < 0) fail(PFC_EINVAL, "nsteps < 0");
    if (!pl->e.field_set[0]) fail(PFC_EINVAL, "pfc_set_field has not been called");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (!pl->e.tuned) {
      if (pl->e.f64())
        autotune<double>(pl->e, st);
      else
        autotune<float>(pl->e, st);
    }
    if (pl->e.f64())
      kernel_times<double>(pl->e, nsteps, st, ms, calls);
    else
      kernel_times<float>(pl->e, nsteps, st, ms, calls);
  });
}

int pfc_plan_schedule(const pfc_plan* pl, int info[8]) {
  return guarded([&] {
    if (!pl || !info) fail(PFC_EINVAL, "NULL argument");
    const FastShape* fs[2] = {&pl->e.frow, &pl->e.fcol};
    for (int k = 0; k < 2; ++k) {
      info[4 * k + 0] = pl->e.dense ? -2 : pl->e.resident ? -1 : fs[k]->on ? 1 : 0;
      info[4 * k + 1] = fs[k]->code();
      info[4 * k + 2] = fs[k]->P;
      info[4 * k + 3] = fs[k]->lines;
    }
  });
}

int pfc_plan_stores(const pfc_plan* pl, int* nt) {
  return guarded([&] {
    if (!pl || !nt) fail(PFC_EINVAL, "NULL argument");
    *nt = (pl->e.resident || !(pl->e.frow.on && pl->e.fcol.on)) ? -1 : (pl->e.frow.nt ? 1 : 0);
  });
}

int pfc_run_tracked(pfc_plan* pl, long long nsteps, long long every, double* out, void* stream) {
  return guarded([&] {
    if (!pl) fail(PFC_EINVAL, "NULL plan");
    engine_run_tracked(pl->e, nsteps, every, out, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_run_tracked(pfc_batch* b, long long nsteps, long long every, double* out, void* stream) {
  return guarded([&] {
    if (!b) fail(PFC_EINVAL, "NULL batch");
    engine_run_tracked(b->e, nsteps, every, out, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_plan_graph_stats(const pfc_plan* pl, long long out[2]) {
  return guarded([&] {
    if (!pl || !out) fail(PFC_EINVAL, "NULL argument");
    out[0] = pl->e.graph_captures;
    out[1] = pl->e.graph_replays;
  });
}

int pfc_batch_graph_stats(const pfc_batch* b, long long out[2]) {
  return guarded([&] {
    if (!b || !out) fail(PFC_EINVAL, "NULL argument");
    out[0] = b->e.graph_captures;
    out[1] = b->e.graph_replays;
  });
}

int pfc_plan_bytes(const pfc_plan* pl, long long* bytes) {
  return guarded([&] {
    if (!pl || !bytes) fail(PFC_EINVAL, "NULL plan or bytes");
    *bytes = (long long)engine_bytes(pl->e);
  });
}

// ---- batches ---------------------------------------------------------------

int pfc_batch_create(const pfc_params* shared, int capacity, pfc_batch** out) {
  return guarded([&] {
    if (!out) fail(PFC_EINVAL, "out is NULL");
    *out = nullptr;
    auto b = std::make_unique<pfc_batch>();
    engine_init(b->e, shared, capacity, true);
    engine_tables(b->e, 0, capacity, nullptr);
    hip_check(hipStreamSynchronize(nullptr), "sync");
    *out = b.release();
  });
}

int pfc_batch_destroy(pfc_batch* b) {
  return guarded([&] { delete b; });
}

int pfc_batch_set_slot(pfc_batch* b, int slot, double epsilon, double M, void* stream) {
  return guarded([&] {
    if (!b) fail(PFC_EINVAL, "NULL batch");
    engine_set_slot(b->e, slot, epsilon, M, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_set_slot_box(pfc_batch* b, int slot, double Lx, double Ly, double dt, void* stream) {
  return guarded([&] {
    if (!b) fail(PFC_EINVAL, "NULL batch");
    engine_set_slot_box(b->e, slot, Lx, Ly, dt, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_set_field(pfc_batch* b, int slot, const void* field, int dtype, void* stream) {
  return guarded([&] {
    if (!b) fail(PFC_EINVAL, "NULL batch");
    engine_set_field(b->e, slot, field, dtype, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_get_field(pfc_batch* b, int slot, void* field, int dtype, void* stream) {
  return guarded([&] {
    if (!b) fail(PFC_EINVAL, "NULL batch");
    engine_get_field(b->e, slot, field, dtype, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_set_external_potential(pfc_batch* b, int slot, const void* V, int dtype, void* stream) {
  return guarded([&] {
    if (!b) fail(PFC_EINVAL, "NULL batch");
    engine_set_external(b->e, slot, V, dtype, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_get_fields(pfc_batch* b, void* fields, int dtype, void* stream) {
  return guarded([&] {
    if (!b) fail(PFC_EINVAL, "NULL batch");
    engine_get_fields(b->e, fields, dtype, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_step_counts(pfc_batch* b, const long long* counts, void* stream) {
  return guarded([&] {
    if (!b) fail(PFC_EINVAL, "NULL batch");
    engine_step_counts(b->e, counts, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_diagnose(pfc_batch* b, const int* ky, const int* kx, int nent, double* out, void* stream) {
  return guarded([&] {
    if (!b) fail(PFC_EINVAL, "NULL batch");
    engine_diagnose(b->e, ky, kx, nent, out, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_set_active(pfc_batch* b, const int* slots, int n, void* stream) {
  return guarded([&] {
    if (!b) fail(PFC_EINVAL, "NULL batch");
    engine_set_active(b->e, slots, n, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_step(pfc_batch* b, long long nsteps, void* stream) {
  return guarded([&] {
    if (!b) fail(PFC_EINVAL, "NULL batch");
    engine_step(b->e, nsteps, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_reduce(pfc_batch* b, double* out, void* stream) {
  return guarded([&] {
    if (!b || !out) fail(PFC_EINVAL, "NULL batch or out");
    engine_reduce_active(b->e, out, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_get_status(pfc_batch* b, int* flags, void* stream) {
  return guarded([&] {
    if (!b || !flags) fail(PFC_EINVAL, "NULL batch or flags");
    engine_status(b->e, flags, reinterpret_cast<hipStream_t>(stream));
  });
}

int pfc_batch_bytes(const pfc_batch* b, long long* bytes) {
  return guarded([&] {
    if (!b || !bytes) fail(PFC_EINVAL, "NULL batch or bytes");
    *bytes = (long long)engine_bytes(b->e);
  });
}

int pfc_group_create(pfc_batch* const* batches, int n, pfc_group** out) {
  return guarded([&] {
    if (!out) fail(PFC_EINVAL, "out is NULL");
    *out = nullptr;
    if (n < 1 || !batches) fail(PFC_EINVAL, "a group needs at least one batch");
    auto g = std::make_unique<pfc_group>();
    for (int i = 0; i < n; ++i) {
      pfc_batch* b = batches[i];
      if (!b) fail(PFC_EINVAL, "NULL batch");
      for (const pfc_batch* o : g->mem)
        if (o == b) fail(PFC_EINVAL, "a batch appears twice in the group");
      group_check_member(b->e);
      if (i == 0) {
        g->precision = b->e.p.precision;
        g->scheme = b->e.p.scheme;
      } else if (b->e.p.precision != g->precision || b->e.p.scheme != g->scheme) {
        fail(PFC_EINVAL, "group members must share precision and scheme");
      }
      g->mem.push_back(b);
    }
    g->order.resize(n);
    for (int i = 0; i < n; ++i) g->order[i] = i;
    // largest instances first: their tiles are dispatched first, the small
    // members' tiles fill in behind them
    std::stable_sort(g->order.begin(), g->order.end(), [&](int a, int b) {
      const Engine &ea = g->mem[a]->e, &eb = g->mem[b]->e;
      return (long long)ea.p.Nx * ea.p.Ny > (long long)eb.p.Nx * eb.p.Ny;
    });
    const size_t dsize = g->precision == 64 ? sizeof(GroupMember<double>) : sizeof(GroupMember<float>);
    g->desc.alloc((size_t)n * 2 * kGroupClasses * n * dsize);  // at most one chunk per member
    // cache budget per chunk in MB (PFC_HIP_GROUP_CACHE_MB; default 0 = one
    // chunk: on the 256-instance sweep, 100 and 160 MB chunks ran 8 % slower
    // than one chunk, the lost overlap of the two class chains outweighing
    // the cached working set)
    const char* cm = std::getenv("PFC_HIP_GROUP_CACHE_MB");
    g->budget = (size_t)((cm ? std::max(0.0, std::atof(cm)) : 0.0) * 1024 * 1024);
    // output stores of the per-pass launches: ordinary by default (on the
    // 256-instance sweep's sizes they measured 3-25 % faster than streaming
    // ones); PFC_HIP_GROUP_STORES=nt streams them
    const char* gs = std::getenv("PFC_HIP_GROUP_STORES");
    g->nt = gs && gs[0] == 'n';
    g->built.assign(n, -1);
    *out = g.release();
  });
}

int pfc_group_destroy(pfc_group* g) {
  return guarded([&] { delete g; });
}

int pfc_group_step(pfc_group* g, long long nsteps, void* stream) {
  return guarded([&] {
    if (!g) fail(PFC_EINVAL, "NULL group");
    const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (g->precision == 64)
      group_step<double>(*g, nsteps, st);
    else
      group_step<float>(*g, nsteps, st);
  });
}

int pfc_group_set_flow(pfc_group* g, int steps_per_launch) {
  return guarded([&] {
    if (!g) fail(PFC_EINVAL, "NULL group");
    if (steps_per_launch < 0) fail(PFC_EINVAL, "steps_per_launch < 0");
    if (steps_per_launch > 0) flow_check(*g);
    g->flow_steps = steps_per_launch;
    g->flow_built.clear();  // the next step builds the plan
  });
}

int pfc_group_flow_stats(const pfc_group* g, long long out[4]) {
  return guarded([&] {
    if (!g || !out) fail(PFC_EINVAL, "NULL argument");
    out[0] = g->flow_launches;
    out[1] = g->flow[0].blocks;
    out[2] = g->flow[1].blocks;
    out[3] = g->flow_steps;
  });
}

int pfc_group_graph_stats(const pfc_group* g, long long out[2]) {
  return guarded([&] {
    if (!g || !out) fail(PFC_EINVAL, "NULL argument");
    out[0] = g->captures;
    out[1] = g->replays;
  });
}

int pfc_group_chunks(const pfc_group* g, int* nchunks) {
  return guarded([&] {
    if (!g || !nchunks) fail(PFC_EINVAL, "NULL argument");
    *nchunks = (int)g->chunks.size();
  });
}

int pfc_slot_bytes(const pfc_params* params, long long* bytes) {
  return guarded([&] {
    validate(params);
    if (!bytes) fail(PFC_EINVAL, "NULL bytes");
    const size_t nc = state_layout(params->Nx, params->Ny, params->precision).size(params->Ny);
    const size_t nt = (size_t)(params->Nx / 2 + 1) * table_rows(params->Ny);
    const size_t cs = params->precision == 64 ? 16 : 8, ts = sizeof(double);
    const bool rk2 = params->scheme == PFC_SCHEME_ETDRK2;
    const size_t grid = (size_t)(params->Nx / 2 + 1) + params->Ny + 1;  // the slot's k-grids and dt
    const size_t tf = params->precision == 64 ? 0 : sizeof(float);      // FP32 engines' table copies
    *bytes = (long long)(nc * cs * (rk2 ? 3 : 2) + nt * (ts + tf) * (rk2 ? 3 : 2) + grid * ts);
  });
}

// ---- plain transforms --------------------------------------------------------

int pfc_rfft2(const void* in, void* out, int batch, int Ny, int Nx, int dtype, void* stream) {
  return guarded([&] {
    if (!in || !out) fail(PFC_EINVAL, "NULL buffer");
    if (batch < 1 || Nx < 4 || (Nx & 1) || Ny < 2) fail(PFC_EINVAL, "bad shape");
    if (dtype != 64 && dtype != 32) fail(PFC_EINVAL, "dtype must be 64 or 32");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(g_fft_mu);
    FftEntry* e = fft_entry(Ny, Nx, dtype);
    const size_t nreal = (size_t)Ny * Nx, ncplx = (size_t)Ny * (Nx / 2 + 1);
    dim3 rgrid(row_blocks(e->rg), batch), cgrid(col_blocks(e->cg), batch);
    if (dtype == 64) {
      hipLaunchKernelGGL((k_row_r2c<double, double>), rgrid, dim3(256), row_lds(e->rg, 16), st, (const double*)in,
                         (V2<double>*)out, e->rg, (const V2<double>*)e->twN.p, 1.0, nreal, ncplx);
      hipLaunchKernelGGL((k_col_fft<double, false>), cgrid, dim3(256), col_lds(e->cg, 16), st, (V2<double>*)out,
                         e->cg, (const V2<double>*)e->twY.p, 1.0, ncplx);
    } else {
      hipLaunchKernelGGL((k_row_r2c<float, float>), rgrid, dim3(256), row_lds(e->rg, 8), st, (const float*)in,
                         (V2<float>*)out, e->rg, (const V2<float>*)e->twN.p, 1.0f, nreal, ncplx);
      hipLaunchKernelGGL((k_col_fft<float, false>), cgrid, dim3(256), col_lds(e->cg, 8), st, (V2<float>*)out, e->cg,
                         (const V2<float>*)e->twY.p, 1.0f, ncplx);
    }
    hip_check(hipGetLastError(), "rfft2 launch");
  });
}

int pfc_irfft2(const void* in, void* out, int batch, int Ny, int Nx, int dtype, void* stream) {
  return guarded([&] {
    if (!in || !out) fail(PFC_EINVAL, "NULL buffer");
    if (batch < 1 || Nx < 4 || (Nx & 1) || Ny < 2) fail(PFC_EINVAL, "bad shape");
    if (dtype != 64 && dtype != 32) fail(PFC_EINVAL, "dtype must be 64 or 32");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(g_fft_mu);
    FftEntry* e = fft_entry(Ny, Nx, dtype);
    const size_t nreal = (size_t)Ny * Nx, ncplx = (size_t)Ny * (Nx / 2 + 1);
    const size_t cs = dtype == 64 ? 16 : 8;
    const size_t need = (size_t)batch * ncplx * cs;
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::unique_ptr<DevBuf>& sbuf = g_fft_scratch[std::make_pair(dev, st)];
    if (!sbuf) sbuf = std::make_unique<DevBuf>();
    if (sbuf->n < need) {
      hip_check(hipStreamSynchronize(st), "sync before scratch growth");
      sbuf->alloc(need);
    }
    DevBuf* scr = sbuf.get();
    hip_check(hipMemcpyAsync(scr->p, in, need, hipMemcpyDeviceToDevice, st), "copy");
    dim3 rgrid(row_blocks(e->rg), batch), cgrid(col_blocks(e->cg), batch);
    const double scale = 1.0 / ((double)Nx * (double)Ny);
    if (dtype == 64) {
      hipLaunchKernelGGL((k_col_fft<double, true>), cgrid, dim3(256), col_lds(e->cg, 16), st,
                         (V2<double>*)scr->p, e->cg, (const V2<double>*)e->twY.p, 1.0, ncplx);
      hipLaunchKernelGGL((k_row_c2r<double, double>), rgrid, dim3(256), row_lds(e->rg, 16), st,
                         (const V2<double>*)scr->p, (double*)out, e->rg, (const V2<double>*)e->twN.p, scale,
                         ncplx, nreal, nullptr);
    } else {
      hipLaunchKernelGGL((k_col_fft<float, true>), cgrid, dim3(256), col_lds(e->cg, 8), st, (V2<float>*)scr->p,
                         e->cg, (const V2<float>*)e->twY.p, 1.0f, ncplx);
      hipLaunchKernelGGL((k_row_c2r<float, float>), rgrid, dim3(256), row_lds(e->rg, 8), st,
                         (const V2<float>*)scr->p, (float*)out, e->rg, (const V2<float>*)e->twN.p,
                         (float)scale, ncplx, nreal, nullptr);
    }
    hip_check(hipGetLastError(), "irfft2 launch");
  });
}

int pfc_irfft2_noalloc(const void* in, void* out, int batch, int Ny, int Nx, int dtype, void* stream) {
  return guarded([&] {
    if (!in || !out) fail(PFC_EINVAL, "NULL buffer");
    if (batch < 1 || Nx < 4 || (Nx & 1) || Ny < 2) fail(PFC_EINVAL, "bad shape");
    if (dtype != 64 && dtype != 32) fail(PFC_EINVAL, "dtype must be 64 or 32");
    const size_t cs = dtype == 64 ? 16 : 8;
    const size_t ncplx = (size_t)Ny * (Nx / 2 + 1), nreal = (size_t)Ny * Nx;
    // layout of the static scratch: roots of Nx, roots of Ny, the spectrum copy
    const size_t offY = align256((size_t)Nx * cs);
    const size_t offS = offY + align256((size_t)Ny * cs);
    const size_t need = offS + (size_t)batch * ncplx * cs;
    if (need > kNoAllocScratchBytes)
      fail(PFC_EINVAL, "batch x Ny x (Nx/2+1) spectrum plus roots needs " + std::to_string(need) +
                           " bytes; the static scratch holds " + std::to_string(kNoAllocScratchBytes));
    const RowGeom rg = row_geom(Nx, Ny, cs);
    const ColGeom cg = col_geom(Nx, Ny, cs);
    unsigned char* base = noalloc_scratch();
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hip_check(hipMemcpyAsync(base + offS, in, (size_t)batch * ncplx * cs, hipMemcpyDeviceToDevice, st), "copy");
    dim3 rgrid(row_blocks(rg), batch), cgrid(col_blocks(cg), batch);
    const double scale = 1.0 / ((double)Nx * (double)Ny);
    if (dtype == 64) {
      using V = V2<double>;
      V* twN = reinterpret_cast<V*>(base);
      V* twY = reinterpret_cast<V*>(base + offY);
      V* sp = reinterpret_cast<V*>(base + offS);
      hipLaunchKernelGGL((k_twiddle_table<double>), dim3((Nx + 255) / 256), dim3(256), 0, st, twN, Nx);
      hipLaunchKernelGGL((k_twiddle_table<double>), dim3((Ny + 255) / 256), dim3(256), 0, st, twY, Ny);
      hipLaunchKernelGGL((k_col_fft<double, true>), cgrid, dim3(256), col_lds(cg, 16), st, sp, cg,
                         (const V*)twY, 1.0, ncplx);
      hipLaunchKernelGGL((k_row_c2r<double, double>), rgrid, dim3(256), row_lds(rg, 16), st, (const V*)sp,
                         (double*)out, rg, (const V*)twN, scale, ncplx, nreal, nullptr);
    } else {
      using V = V2<float>;
      V* twN = reinterpret_cast<V*>(base);
      V* twY = reinterpret_cast<V*>(base + offY);
      V* sp = reinterpret_cast<V*>(base + offS);
      hipLaunchKernelGGL((k_twiddle_table<float>), dim3((Nx + 255) / 256), dim3(256), 0, st, twN, Nx);
      hipLaunchKernelGGL((k_twiddle_table<float>), dim3((Ny + 255) / 256), dim3(256), 0, st, twY, Ny);
      hipLaunchKernelGGL((k_col_fft<float, true>), cgrid, dim3(256), col_lds(cg, 8), st, sp, cg, (const V*)twY,
                         1.0f, ncplx);
      hipLaunchKernelGGL((k_row_c2r<float, float>), rgrid, dim3(256), row_lds(rg, 8), st, (const V*)sp,
                         (float*)out, rg, (const V*)twN, (float)scale, ncplx, nreal, nullptr);
    }
    hip_check(hipGetLastError(), "irfft2_noalloc launch");
  });
}

}  // extern "C"
