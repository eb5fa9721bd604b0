// Host side of libpfc_hip.so: plans, FP64 tables, launch shapes, C ABI.
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
  int var = 0;      // schedule variant
  int P = 0;        // points per lane
  int tpl = 0;      // lanes per line
  int lines = 0;    // lines per block
  int blocks = 0;
  size_t lds = 0;
};

// Row kernel shape: LDS = twiddles (Nx) + lines * (M + 1).
FastShape row_shape(int Nx, int Ny, size_t csize, const SchedInfo& si, int lines) {
  FastShape f;
  const int M = Nx / 2;
  f.var = si.variant;
  f.P = si.P;
  f.tpl = M / si.P;
  f.lines = lines;
  f.blocks = (Ny + lines - 1) / lines;
  f.lds = ((size_t)Nx + (size_t)lines * row_ld(M)) * csize;
  f.on = f.tpl * lines <= 256 && f.tpl * lines >= 64 && f.lds <= (size_t)lds_limit();
  return f;
}

// Column kernel shape: LDS = twiddles (Ny) + lines * Ny.
FastShape col_shape(int Nx, int Ny, size_t csize, const SchedInfo& si, int lines) {
  FastShape f;
  const int Nc = Nx / 2 + 1;
  f.var = si.variant;
  f.P = si.P;
  f.tpl = Ny / si.P;
  f.lines = lines;
  f.blocks = (Nc + lines - 1) / lines;
  f.lds = ((size_t)Ny + (size_t)lines * col_ld(Ny)) * csize;
  f.on = f.tpl * lines <= 256 && f.tpl * lines >= 64 && f.lds <= (size_t)lds_limit();
  return f;
}

// Candidate (variant, lines) shapes; the default is variant 0 with enough
// blocks to cover the 256 CUs.
std::vector<FastShape> row_candidates(int Nx, int Ny, size_t csize) {
  std::vector<FastShape> out;
  for (const SchedInfo& si : schedules_for(Nx / 2))
    for (int w = 1; w <= 256; w *= 2) {
      FastShape f = row_shape(Nx, Ny, csize, si, w);
      if (f.on) out.push_back(f);
    }
  return out;
}

std::vector<FastShape> col_candidates(int Nx, int Ny, size_t csize) {
  std::vector<FastShape> out;
  for (const SchedInfo& si : schedules_for(Ny))
    for (int c = 1; c <= 256; c *= 2) {
      FastShape f = col_shape(Nx, Ny, csize, si, c);
      if (f.on) out.push_back(f);
    }
  return out;
}

FastShape pick_default(const std::vector<FastShape>& cands, int target_blocks) {
  FastShape best;
  for (const FastShape& f : cands) {
    if (f.var != 0) continue;
    // most lines per block that still gives >= target_blocks blocks, else the fewest lines
    if (!best.on) {
      best = f;
      continue;
    }
    if (f.blocks >= target_blocks && f.lines > best.lines) best = f;
  }
  return best;
}

template <typename T> struct FastLaunch;
template <> struct FastLaunch<double> {
  static void row(int M, const FastShape& f, hipStream_t st, const V2<double>* S, V2<double>* T, int Ny, const Nonlin& q,
                  const V2<double>* tw, int* status) {
    if (!row_fast_launch_f64(M, f.var, dim3(f.blocks), dim3(f.lines * f.tpl), f.lds, st, S, T, Ny, q, tw, status))
      fail(PFC_EUNSUPPORTED, "no fast row schedule");
  }
  static void col(int Ny, int mode, const FastShape& f, hipStream_t st, V2<double>* S, const V2<double>* T, V2<double>* D,
                  int Nc, const ColTables& tb, const V2<double>* tw) {
    if (!col_fast_launch_f64(Ny, f.var, mode, dim3(f.blocks), dim3(f.lines * f.tpl), f.lds, st, S, T, D, Nc, tb,
                             tw))
      fail(PFC_EUNSUPPORTED, "no fast column schedule");
  }
};
template <> struct FastLaunch<float> {
  static void row(int M, const FastShape& f, hipStream_t st, const V2<float>* S, V2<float>* T, int Ny, const Nonlin& q,
                  const V2<float>* tw, int* status) {
    if (!row_fast_launch_f32(M, f.var, dim3(f.blocks), dim3(f.lines * f.tpl), f.lds, st, S, T, Ny, q, tw, status))
      fail(PFC_EUNSUPPORTED, "no fast row schedule");
  }
  static void col(int Ny, int mode, const FastShape& f, hipStream_t st, V2<float>* S, const V2<float>* T, V2<float>* D,
                  int Nc, const ColTables& tb, const V2<float>* tw) {
    if (!col_fast_launch_f32(Ny, f.var, mode, dim3(f.blocks), dim3(f.lines * f.tpl), f.lds, st, S, T, D, Nc, tb,
                             tw))
      fail(PFC_EUNSUPPORTED, "no fast column schedule");
  }
};

// ---------------------------------------------------------------------------
// FP64 k-grids and ETD tables (the reference's host arithmetic, restated)
// ---------------------------------------------------------------------------
// kx = 2 pi * rfftfreq(Nx, dx), ky = 2 pi * fftfreq(Ny, dy)  (geometry_2d.py:154-155)
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

// ETDRK2 coefficient dt*phi2(z), phi2 = (e^z - 1 - z)/z^2 (Cox & Matthews 2002):
// Taylor series for |z| < 0.1, expm1 otherwise.
double etd2_coeff(double z, double dt) {
  if (std::fabs(z) < 0.1) {
    double s = 0.0, t = 0.5;
    for (int j = 0; j < 16; ++j) {
      s += t;
      t = t * z / (j + 3);
    }
    return dt * s;
  }
  return dt * ((std::expm1(z) - z) / (z * z));
}

struct Tables {
  std::vector<double> E1, Q1, Q2;  // column-major [Nc][Ny]
};

// E = exp(L dt), P = (E-1)/L or dt(1 + z/2 + z^2/6) (kernels.py:55-71) with
// L = Gamma lap (beta (q0^2 + lap)^2 + eps + M), lap = -(kx^2 + ky^2)
// (stdpfc_cpu/simulation.py:95-108; M as logpfc_extv/simulation.py:100-116).
// Stored pre-scaled for the row-spectrum state: E1 = E/Ny, Q1 = P Gamma lap/(Nx Ny).
Tables build_tables(const pfc_params& p, const std::vector<double>& kx, const std::vector<double>& ky, double M) {
  const int Nc = (int)kx.size(), Ny = (int)ky.size();
  const double inv_ny = 1.0 / Ny, inv_n = 1.0 / ((double)p.Nx * (double)Ny);
  Tables t;
  t.E1.resize((size_t)Nc * Ny);
  t.Q1.resize((size_t)Nc * Ny);
  if (p.scheme == PFC_SCHEME_ETDRK2) t.Q2.resize((size_t)Nc * Ny);
  for (int c = 0; c < Nc; ++c) {
    for (int j = 0; j < Ny; ++j) {
      const double lap = -(kx[c] * kx[c] + ky[j] * ky[j]);
      const double s = p.q0 * p.q0 + lap;
      double lmu = p.beta * (s * s) + p.epsilon;
      if (M != 0.0) lmu = lmu + M;
      const double L = p.Gamma * lap * lmu;
      const double z = L * p.dt;
      const double E = std::exp(z);
      const double P = std::fabs(z) > p.smallZTol ? (E - 1) / L : p.dt * (1 + z / 2 + z * z / 6);
      const size_t i = (size_t)c * Ny + j;
      t.E1[i] = E * inv_ny;
      t.Q1[i] = P * (p.Gamma * lap) * inv_n;
      if (!t.Q2.empty()) t.Q2[i] = etd2_coeff(z, p.dt) * (p.Gamma * lap) * inv_n;
    }
  }
  return t;
}

template <typename X>
std::vector<X> cast_vec(const std::vector<double>& v) {
  std::vector<X> o(v.size());
  for (size_t i = 0; i < v.size(); ++i) o[i] = (X)v[i];
  return o;
}

}  // namespace

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
struct pfc_plan {
  pfc_params p{};
  int Nc = 0;
  size_t csize = 0;  // bytes per complex of the bulk precision
  RowGeom rg{}, rg64{};
  ColGeom cg{}, cg64{};
  FastShape frow{}, fcol{};      // compile-time-schedule kernels, when the sizes have one
  bool tuned = false;            // autotuner has run (or was skipped) for this plan
  DevBuf S, Tb, D;               // state, row spectrum of N, RK2 k-space carry
  DevBuf E1, Q1, Q2;             // ETD tables (bulk precision)
  DevBuf twN, twY;               // twiddles (bulk precision)
  DevBuf twN64, twY64;           // FP64 twiddles for the reductions
  DevBuf kx2, ky2;               // FP64 k^2 grids for the free energy
  DevBuf part;                   // reduction partials
  DevBuf diag;                   // two row spectra for the chemical-potential diagnostics
  DevBuf scal;                   // [0..3] reduce out, [4] pin, int status at [5]
  std::vector<double> kx, ky;
  double M = 0.0;
  bool field_set = false;
  double* host_out = nullptr;    // pinned host mirror of scal[0..3]
  ~pfc_plan() {
    if (host_out) (void)hipHostFree(host_out);
  }
  bool f64() const { return p.precision == 64; }
  double* pin() const { return reinterpret_cast<double*>(scal.p) + 4; }
  int* status() const { return reinterpret_cast<int*>(reinterpret_cast<double*>(scal.p) + 5); }
};

namespace {

void upload_tables(pfc_plan* pl) {
  Tables t = build_tables(pl->p, pl->kx, pl->ky, pl->M);
  if (pl->f64()) {
    pl->E1.upload(t.E1);
    pl->Q1.upload(t.Q1);
    if (!t.Q2.empty()) pl->Q2.upload(t.Q2);
  } else {
    pl->E1.upload(cast_vec<float>(t.E1));
    pl->Q1.upload(cast_vec<float>(t.Q1));
    if (!t.Q2.empty()) pl->Q2.upload(cast_vec<float>(t.Q2));
  }
}

Nonlin nonlin_of(const pfc_plan* pl) {
  Nonlin q{};
  q.model = pl->p.model;
  q.clamp = pl->p.clamp_log;
  q.g = pl->p.g;
  q.minPhi = pl->p.minPhi;
  q.logMin = std::log(pl->p.minPhi > 0 ? pl->p.minPhi : 1e-300);
  q.M = pl->M;
  return q;
}

// Optional per-launch event timing (pfc_step_profile): kind 0 row, 1 col, 2 col stage 2.
struct Prof {
  hipEvent_t ev[2] = {nullptr, nullptr};
  double ms[3] = {0, 0, 0};
  long long calls[3] = {0, 0, 0};
  hipStream_t st = nullptr;
  void begin() { hip_check(hipEventRecord(ev[0], st), "event record"); }
  void end(int kind) {
    hip_check(hipEventRecord(ev[1], st), "event record");
    hip_check(hipEventSynchronize(ev[1]), "event sync");
    float t = 0.f;
    hip_check(hipEventElapsedTime(&t, ev[0], ev[1]), "event elapsed");
    ms[kind] += t;
    calls[kind] += 1;
  }
};

template <typename T>
void launch_step(pfc_plan* pl, long long nsteps, hipStream_t st, Prof* prof = nullptr) {
  using V = V2<T>;
  V* S = reinterpret_cast<V*>(pl->S.p);
  V* Tb = reinterpret_cast<V*>(pl->Tb.p);
  V* D = reinterpret_cast<V*>(pl->D.p);
  const V* twN = reinterpret_cast<const V*>(pl->twN.p);
  const V* twY = reinterpret_cast<const V*>(pl->twY.p);
  ColTables tb{pl->E1.p, pl->Q1.p, pl->Q2.p, pl->p.pin_mean ? pl->pin() : nullptr};
  const Nonlin q = nonlin_of(pl);
  const size_t rl = row_lds(pl->rg, sizeof(V)), cl = col_lds(pl->cg, sizeof(V));
  const int rb = row_blocks(pl->rg), cb = col_blocks(pl->cg);
  auto B = [&]() { if (prof) prof->begin(); };
  auto E = [&](int k) { if (prof) prof->end(k); };
  const int M = pl->p.Nx / 2, Ny = pl->p.Ny, Nc = pl->Nc;
  auto row = [&]() {
    B();
    if (pl->frow.on)
      FastLaunch<T>::row(M, pl->frow, st, S, Tb, Ny, q, twN, pl->status());
    else
      hipLaunchKernelGGL((k_row_nonlin<T>), dim3(rb), dim3(256), rl, st, S, Tb, pl->rg, q, twN, pl->status());
    E(0);
  };
  auto col = [&](auto mode_tag, int kind) {
    constexpr int MODE = decltype(mode_tag)::value;
    B();
    if (pl->fcol.on)
      FastLaunch<T>::col(Ny, MODE, pl->fcol, st, S, Tb, D, Nc, tb, twY);
    else
      hipLaunchKernelGGL((k_col_etd<T, MODE>), dim3(cb), dim3(256), cl, st, S, Tb, D, pl->cg, tb, twY);
    E(kind);
  };
  for (long long s = 0; s < nsteps; ++s) {
    row();
    if (pl->p.scheme == PFC_SCHEME_ETD1) {
      col(std::integral_constant<int, COL_ETD1>{}, 1);
    } else {
      col(std::integral_constant<int, COL_RK2_STAGE1>{}, 1);
      row();
      col(std::integral_constant<int, COL_RK2_STAGE2>{}, 2);
    }
  }
  hip_check(hipGetLastError(), "step launch");
}

// ---------------------------------------------------------------------------
// autotuner: at a plan's first step, time every (schedule variant, lines per
// block) candidate of the fast row and column kernels on scratch copies of the
// live state and keep the fastest.  Choices are cached per (device, kernel
// kind, lengths, precision, mode) for the life of the process.  PFC_HIP_TUNE=0
// keeps the defaults.
// ---------------------------------------------------------------------------
std::mutex g_tune_mu;
std::map<std::array<int, 6>, std::pair<int, int>> g_tune;  // -> (variant, lines)

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

template <typename T>
void autotune(pfc_plan* pl, hipStream_t st) {
  pl->tuned = true;
  const char* env = std::getenv("PFC_HIP_TUNE");
  if (env && env[0] == '0') return;
  using V = V2<T>;
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  const int Nx = pl->p.Nx, Ny = pl->p.Ny, M = Nx / 2, Nc = pl->Nc;
  const size_t nc = (size_t)Nc * Ny;
  const int reps = 5;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto apply = [](FastShape& cur, const std::vector<FastShape>& cands, std::pair<int, int> vl) {
    for (const FastShape& f : cands)
      if (f.var == vl.first && f.lines == vl.second) cur = f;
  };
  if (pl->frow.on) {
    const std::array<int, 6> key{dev, 0, M, Ny, pl->p.precision, pl->p.model * 2 + pl->p.clamp_log};
    auto cands = row_candidates(Nx, Ny, pl->csize);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      apply(pl->frow, cands, it->second);
    } else if (cands.size() > 1) {
      DevBuf T2;
      T2.alloc(nc * sizeof(V));
      const Nonlin q = nonlin_of(pl);
      float best = 1e30f;
      FastShape pick = pl->frow;
      for (const FastShape& f : cands) {
        float ms = time_launches(st, reps, [&] {
          FastLaunch<T>::row(M, f, st, (const V*)pl->S.p, (V*)T2.p, Ny, q, (const V*)pl->twN.p, pl->status());
        });
        if (ms < best) best = ms, pick = f;
      }
      pl->frow = pick;
      g_tune[key] = {pick.var, pick.lines};
    }
  }
  if (pl->fcol.on) {
    const int mode = pl->p.scheme == PFC_SCHEME_ETD1 ? COL_ETD1 : COL_RK2_STAGE1;
    const std::array<int, 6> key{dev, 1, Ny, Nc, pl->p.precision, mode};
    auto cands = col_candidates(Nx, Ny, pl->csize);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      apply(pl->fcol, cands, it->second);
    } else if (cands.size() > 1) {
      DevBuf S2, T2, D2;
      S2.alloc(nc * sizeof(V));
      T2.alloc(nc * sizeof(V));
      D2.alloc(nc * sizeof(V));
      hip_check(hipMemcpyAsync(S2.p, pl->S.p, nc * sizeof(V), hipMemcpyDeviceToDevice, st), "copy");
      hip_check(hipMemcpyAsync(T2.p, pl->S.p, nc * sizeof(V), hipMemcpyDeviceToDevice, st), "copy");
      ColTables tb{pl->E1.p, pl->Q1.p, pl->Q2.p, pl->p.pin_mean ? pl->pin() : nullptr};
      float best = 1e30f;
      FastShape pick = pl->fcol;
      for (const FastShape& f : cands) {
        float ms = time_launches(st, reps, [&] {
          FastLaunch<T>::col(Ny, mode, f, st, (V*)S2.p, (const V*)T2.p, (V*)D2.p, Nc, tb, (const V*)pl->twY.p);
        });
        if (ms < best) best = ms, pick = f;
      }
      pl->fcol = pick;
      g_tune[key] = {pick.var, pick.lines};
      hip_check(hipStreamSynchronize(st), "sync after tuning");
    }
  }
}

// reductions of the current state; pin_out (device) receives the mean if non-null
template <typename T>
void launch_reduce(pfc_plan* pl, hipStream_t st, double* pin_out) {
  using V = V2<T>;
  const V* S = reinterpret_cast<const V*>(pl->S.p);
  FeParams fp{pl->p.model, pl->p.beta, pl->p.q0, pl->p.epsilon, pl->p.g,
              reinterpret_cast<const double*>(pl->kx2.p), reinterpret_cast<const double*>(pl->ky2.p)};
  Partial* part = reinterpret_cast<Partial*>(pl->part.p);
  const int rb = row_blocks(pl->rg64), cb = col_blocks(pl->cg64);
  const size_t red = 256 * sizeof(Partial);
  const size_t rl = std::max(row_lds(pl->rg64, sizeof(V2<double>)), red);
  const size_t cl = std::max(col_lds(pl->cg64, sizeof(V2<double>)), red);
  hipLaunchKernelGGL((k_fe_row<T>), dim3(rb), dim3(256), rl, st, S, pl->rg64, fp,
                     reinterpret_cast<const V2<double>*>(pl->twN64.p), part);
  hipLaunchKernelGGL((k_fe_col<T>), dim3(cb), dim3(256), cl, st, S, pl->cg64, fp, pl->p.Nx,
                     reinterpret_cast<const V2<double>*>(pl->twY64.p), part + rb);
  hipLaunchKernelGGL(k_fe_final, dim3(1), dim3(256), 0, st, part, rb, cb, pl->p.Nx, pl->p.Ny,
                     reinterpret_cast<double*>(pl->scal.p), pin_out);
  hip_check(hipGetLastError(), "reduce launch");
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

// cache of transform-only twiddles/geometry for pfc_rfft2 / pfc_irfft2
struct FftEntry {
  RowGeom rg{};
  ColGeom cg{};
  DevBuf twN, twY, scratch;
};
std::mutex g_fft_mu;
std::map<std::tuple<int, int, int, int>, std::unique_ptr<FftEntry>> g_fft;

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

}  // namespace

namespace {

template <typename T>
void launch_filter(pfc_plan* pl, int mode, void* out0, void* out1, int dtype, hipStream_t st) {
  using V = V2<T>;
  const size_t nc = (size_t)pl->Nc * pl->p.Ny;
  if (pl->diag.n < 2 * nc * sizeof(V)) pl->diag.alloc(2 * nc * sizeof(V));
  V* d0 = reinterpret_cast<V*>(pl->diag.p);
  V* d1 = d0 + nc;
  const V* S = reinterpret_cast<const V*>(pl->S.p);
  V* Tb = reinterpret_cast<V*>(pl->Tb.p);
  const V* twN = reinterpret_cast<const V*>(pl->twN.p);
  const V* twY = reinterpret_cast<const V*>(pl->twY.p);
  FilterParams fp{pl->p.beta, pl->p.q0, pl->p.epsilon, pl->M, 6.283185307179586, pl->p.Nx,
                  reinterpret_cast<const double*>(pl->kx2.p), reinterpret_cast<const double*>(pl->ky2.p)};
  const int rb = row_blocks(pl->rg), cb = col_blocks(pl->cg);
  const size_t rl = row_lds(pl->rg, sizeof(V)), cl = col_lds(pl->cg, sizeof(V));
  if (mode == FILTER_MU) {
    hipLaunchKernelGGL((k_row_nonlin<T>), dim3(rb), dim3(256), rl, st, S, Tb, pl->rg, nonlin_of(pl), twN,
                       pl->status());
    hipLaunchKernelGGL((k_col_filter<T, FILTER_MU>), dim3(cb), dim3(256), cl, st, S, Tb, d0, d1, pl->cg, fp, twY);
  } else {
    hipLaunchKernelGGL((k_col_filter<T, FILTER_LOCAL>), dim3(cb), dim3(256), cl, st, S, Tb, d0, d1, pl->cg, fp,
                       twY);
  }
  const T scale = T(1.0 / ((double)pl->p.Nx * (double)pl->p.Ny));
  const int nout = mode == FILTER_MU ? 2 : 1;
  for (int o = 0; o < nout; ++o) {
    const V* src = o == 0 ? d0 : d1;
    void* dst = o == 0 ? out0 : out1;
    if (dtype == 64)
      hipLaunchKernelGGL((k_row_c2r<T, double>), dim3(rb), dim3(256), rl, st, src, (double*)dst, pl->rg, twN, scale,
                         (size_t)0, (size_t)0);
    else
      hipLaunchKernelGGL((k_row_c2r<T, float>), dim3(rb), dim3(256), rl, st, src, (float*)dst, pl->rg, twN, scale,
                         (size_t)0, (size_t)0);
  }
  hip_check(hipGetLastError(), "diagnostic launch");
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* pfc_last_error(void) { return g_err.c_str(); }

int pfc_version(void) { return 1; }

int pfc_create(const pfc_params* params, pfc_plan** out) {
  return guarded([&] {
    if (!out) fail(PFC_EINVAL, "out is NULL");
    *out = nullptr;
    validate(params);
    auto pl = std::make_unique<pfc_plan>();
    pl->p = *params;
    pl->M = params->M;
    const int Nx = params->Nx, Ny = params->Ny;
    pl->Nc = Nx / 2 + 1;
    pl->csize = pl->f64() ? sizeof(V2<double>) : sizeof(V2<float>);
    pl->rg = row_geom(Nx, Ny, pl->csize);
    pl->cg = col_geom(Nx, Ny, pl->csize);
    pl->rg64 = row_geom(Nx, Ny, sizeof(V2<double>));
    pl->cg64 = col_geom(Nx, Ny, sizeof(V2<double>));
    if (!std::getenv("PFC_HIP_GENERIC")) {
      auto rc = row_candidates(Nx, Ny, pl->csize);
      auto cc = col_candidates(Nx, Ny, pl->csize);
      pl->frow = pick_default(rc, 512);
      pl->fcol = pick_default(cc, 256);
      // PFC_HIP_SCHEDULE="rowvar:rowlines:colvar:collines" pins a schedule
      // (tests and tuning studies); the autotuner is then skipped.
      if (const char* fx = std::getenv("PFC_HIP_SCHEDULE")) {
        int v[4] = {-1, -1, -1, -1};
        if (std::sscanf(fx, "%d:%d:%d:%d", &v[0], &v[1], &v[2], &v[3]) == 4) {
          bool okr = false, okc = false;
          for (const FastShape& f : rc)
            if (f.var == v[0] && f.lines == v[1]) pl->frow = f, okr = true;
          for (const FastShape& f : cc)
            if (f.var == v[2] && f.lines == v[3]) pl->fcol = f, okc = true;
          if (!okr && pl->frow.on) fail(PFC_EINVAL, "PFC_HIP_SCHEDULE: no such row schedule");
          if (!okc && pl->fcol.on) fail(PFC_EINVAL, "PFC_HIP_SCHEDULE: no such column schedule");
          pl->tuned = true;
        }
      }
    }
    const size_t nc = (size_t)pl->Nc * Ny;
    pl->S.alloc(nc * pl->csize);
    pl->Tb.alloc(nc * pl->csize);
    if (params->scheme == PFC_SCHEME_ETDRK2) pl->D.alloc(nc * pl->csize);
    pl->kx = rfft_k(Nx, params->Lx / Nx);
    pl->ky = fft_k(Ny, params->Ly / Ny);
    upload_tables(pl.get());
    auto tn = twiddles(Nx), ty = twiddles(Ny);
    pl->twN64.upload(tn);
    pl->twY64.upload(ty);
    if (pl->f64()) {
      pl->twN.upload(tn);
      pl->twY.upload(ty);
    } else {
      pl->twN.upload(to_f32(tn));
      pl->twY.upload(to_f32(ty));
    }
    std::vector<double> kx2(pl->kx.size()), ky2(pl->ky.size());
    for (size_t i = 0; i < kx2.size(); ++i) kx2[i] = pl->kx[i] * pl->kx[i];
    for (size_t i = 0; i < ky2.size(); ++i) ky2[i] = pl->ky[i] * pl->ky[i];
    pl->kx2.upload(kx2);
    pl->ky2.upload(ky2);
    pl->part.alloc((size_t)(row_blocks(pl->rg64) + col_blocks(pl->cg64)) * sizeof(Partial));
    pl->scal.alloc(8 * sizeof(double));
    hip_check(hipMemset(pl->scal.p, 0, 8 * sizeof(double)), "memset");
    hip_check(hipHostMalloc((void**)&pl->host_out, 8 * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
    hip_check(hipMemset(pl->S.p, 0, nc * pl->csize), "memset");
    *out = pl.release();
  });
}

int pfc_destroy(pfc_plan* plan) {
  return guarded([&] { delete plan; });
}

int pfc_set_field(pfc_plan* pl, const void* field, int dtype, void* stream) {
  return guarded([&] {
    if (!pl || !field) fail(PFC_EINVAL, "NULL plan or field");
    if (dtype != 64 && dtype != 32) fail(PFC_EINVAL, "dtype must be 64 or 32");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int rb = row_blocks(pl->rg);
    const size_t rl = row_lds(pl->rg, pl->csize);
    if (pl->f64()) {
      const double inv = 1.0 / pl->p.Nx;
      if (dtype == 64)
        hipLaunchKernelGGL((k_row_r2c<double, double>), dim3(rb), dim3(256), rl, st, (const double*)field,
                           (V2<double>*)pl->S.p, pl->rg, (const V2<double>*)pl->twN.p, inv, (size_t)0, (size_t)0);
      else
        hipLaunchKernelGGL((k_row_r2c<double, float>), dim3(rb), dim3(256), rl, st, (const float*)field,
                           (V2<double>*)pl->S.p, pl->rg, (const V2<double>*)pl->twN.p, inv, (size_t)0, (size_t)0);
      hip_check(hipGetLastError(), "set_field launch");
      launch_reduce<double>(pl, st, pl->pin());
    } else {
      const float inv = 1.0f / pl->p.Nx;
      if (dtype == 64)
        hipLaunchKernelGGL((k_row_r2c<float, double>), dim3(rb), dim3(256), rl, st, (const double*)field,
                           (V2<float>*)pl->S.p, pl->rg, (const V2<float>*)pl->twN.p, inv, (size_t)0, (size_t)0);
      else
        hipLaunchKernelGGL((k_row_r2c<float, float>), dim3(rb), dim3(256), rl, st, (const float*)field,
                           (V2<float>*)pl->S.p, pl->rg, (const V2<float>*)pl->twN.p, inv, (size_t)0, (size_t)0);
      hip_check(hipGetLastError(), "set_field launch");
      launch_reduce<float>(pl, st, pl->pin());
    }
    pl->field_set = true;
  });
}

int pfc_get_field(pfc_plan* pl, void* field, int dtype, void* stream) {
  return guarded([&] {
    if (!pl || !field) fail(PFC_EINVAL, "NULL plan or field");
    if (dtype != 64 && dtype != 32) fail(PFC_EINVAL, "dtype must be 64 or 32");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int rb = row_blocks(pl->rg);
    const size_t rl = row_lds(pl->rg, pl->csize);
    if (pl->f64()) {
      if (dtype == 64)
        hipLaunchKernelGGL((k_row_c2r<double, double>), dim3(rb), dim3(256), rl, st, (const V2<double>*)pl->S.p,
                           (double*)field, pl->rg, (const V2<double>*)pl->twN.p, 1.0, (size_t)0, (size_t)0);
      else
        hipLaunchKernelGGL((k_row_c2r<double, float>), dim3(rb), dim3(256), rl, st, (const V2<double>*)pl->S.p,
                           (float*)field, pl->rg, (const V2<double>*)pl->twN.p, 1.0, (size_t)0, (size_t)0);
    } else {
      if (dtype == 64)
        hipLaunchKernelGGL((k_row_c2r<float, double>), dim3(rb), dim3(256), rl, st, (const V2<float>*)pl->S.p,
                           (double*)field, pl->rg, (const V2<float>*)pl->twN.p, 1.0f, (size_t)0, (size_t)0);
      else
        hipLaunchKernelGGL((k_row_c2r<float, float>), dim3(rb), dim3(256), rl, st, (const V2<float>*)pl->S.p,
                           (float*)field, pl->rg, (const V2<float>*)pl->twN.p, 1.0f, (size_t)0, (size_t)0);
    }
    hip_check(hipGetLastError(), "get_field launch");
  });
}

int pfc_step(pfc_plan* pl, long long nsteps, void* stream) {
  return guarded([&] {
    if (!pl) fail(PFC_EINVAL, "NULL plan");
    if (nsteps < 0) fail(PFC_EINVAL, "nsteps < 0");
    if (!pl->field_set) fail(PFC_EINVAL, "pfc_set_field has not been called");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (!pl->tuned) {
      if (pl->f64())
        autotune<double>(pl, st);
      else
        autotune<float>(pl, st);
    }
    if (pl->f64())
      launch_step<double>(pl, nsteps, st);
    else
      launch_step<float>(pl, nsteps, st);
  });
}

int pfc_reduce(pfc_plan* pl, double out[4], void* stream) {
  return guarded([&] {
    if (!pl || !out) fail(PFC_EINVAL, "NULL plan or out");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (pl->f64())
      launch_reduce<double>(pl, st, nullptr);
    else
      launch_reduce<float>(pl, st, nullptr);
    hip_check(hipMemcpyAsync(pl->host_out, pl->scal.p, 4 * sizeof(double), hipMemcpyDeviceToHost, st), "copy");
    hip_check(hipStreamSynchronize(st), "sync");
    std::memcpy(out, pl->host_out, 4 * sizeof(double));
  });
}

int pfc_set_shift(pfc_plan* pl, double M) {
  return guarded([&] {
    if (!pl) fail(PFC_EINVAL, "NULL plan");
    if (!std::isfinite(M)) fail(PFC_EINVAL, "M must be finite");
    hip_check(hipDeviceSynchronize(), "sync before table rebuild");
    pl->M = M;
    upload_tables(pl);
  });
}

int pfc_get_status(pfc_plan* pl, int* flags, void* stream) {
  return guarded([&] {
    if (!pl || !flags) fail(PFC_EINVAL, "NULL plan or flags");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    int v = 0;
    hip_check(hipMemcpyAsync(&v, pl->status(), sizeof(int), hipMemcpyDeviceToHost, st), "copy");
    hip_check(hipStreamSynchronize(st), "sync");
    hip_check(hipMemsetAsync(pl->status(), 0, sizeof(int), st), "clear status");
    *flags = v;
  });
}

int pfc_chemical_potential(pfc_plan* pl, void* mu, void* local_mu, int dtype, void* stream) {
  return guarded([&] {
    if (!pl || !mu || !local_mu) fail(PFC_EINVAL, "NULL argument");
    if (dtype != 64 && dtype != 32) fail(PFC_EINVAL, "dtype must be 64 or 32");
    if (!pl->field_set) fail(PFC_EINVAL, "pfc_set_field has not been called");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (pl->f64())
      launch_filter<double>(pl, FILTER_MU, mu, local_mu, dtype, st);
    else
      launch_filter<float>(pl, FILTER_MU, mu, local_mu, dtype, st);
  });
}

int pfc_local_field(pfc_plan* pl, void* out, int dtype, void* stream) {
  return guarded([&] {
    if (!pl || !out) fail(PFC_EINVAL, "NULL argument");
    if (dtype != 64 && dtype != 32) fail(PFC_EINVAL, "dtype must be 64 or 32");
    if (!pl->field_set) fail(PFC_EINVAL, "pfc_set_field has not been called");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (pl->f64())
      launch_filter<double>(pl, FILTER_LOCAL, out, nullptr, dtype, st);
    else
      launch_filter<float>(pl, FILTER_LOCAL, out, nullptr, dtype, st);
  });
}

int pfc_step_profile(pfc_plan* pl, long long nsteps, void* stream, double ms[3], long long calls[3]) {
  Prof prof;
  int rc = guarded([&] {
    if (!pl || !ms || !calls) fail(PFC_EINVAL, "NULL argument");
    if (nsteps < 0) fail(PFC_EINVAL, "nsteps < 0");
    if (!pl->field_set) fail(PFC_EINVAL, "pfc_set_field has not been called");
    prof.st = reinterpret_cast<hipStream_t>(stream);
    if (!pl->tuned) {
      if (pl->f64())
        autotune<double>(pl, prof.st);
      else
        autotune<float>(pl, prof.st);
    }
    hip_check(hipEventCreate(&prof.ev[0]), "event create");
    hip_check(hipEventCreate(&prof.ev[1]), "event create");
    if (pl->f64())
      launch_step<double>(pl, nsteps, prof.st, &prof);
    else
      launch_step<float>(pl, nsteps, prof.st, &prof);
    for (int k = 0; k < 3; ++k) {
      ms[k] = prof.ms[k];
      calls[k] = prof.calls[k];
    }
  });
  for (hipEvent_t e : prof.ev)
    if (e) (void)hipEventDestroy(e);
  return rc;
}

int pfc_plan_schedule(const pfc_plan* pl, int info[8]) {
  return guarded([&] {
    if (!pl || !info) fail(PFC_EINVAL, "NULL argument");
    const FastShape* fs[2] = {&pl->frow, &pl->fcol};
    for (int k = 0; k < 2; ++k) {
      info[4 * k + 0] = fs[k]->on ? 1 : 0;
      info[4 * k + 1] = fs[k]->var;
      info[4 * k + 2] = fs[k]->P;
      info[4 * k + 3] = fs[k]->lines;
    }
  });
}

int pfc_plan_bytes(const pfc_plan* pl, long long* bytes) {
  return guarded([&] {
    if (!pl || !bytes) fail(PFC_EINVAL, "NULL plan or bytes");
    size_t b = 0;
    for (const DevBuf* d : {&pl->S, &pl->Tb, &pl->D, &pl->E1, &pl->Q1, &pl->Q2, &pl->twN, &pl->twY, &pl->twN64,
                            &pl->twY64, &pl->kx2, &pl->ky2, &pl->part, &pl->scal, &pl->diag})
      b += d->n;
    *bytes = (long long)b;
  });
}

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
      hipLaunchKernelGGL((k_col_fft<double, false>), cgrid, dim3(256), col_lds(e->cg, 16), st, (V2<double>*)out, e->cg,
                         (const V2<double>*)e->twY.p, 1.0, ncplx);
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
    if (e->scratch.n < need) {
      hip_check(hipStreamSynchronize(st), "sync before scratch growth");
      e->scratch.alloc(need);
    }
    hip_check(hipMemcpyAsync(e->scratch.p, in, need, hipMemcpyDeviceToDevice, st), "copy");
    dim3 rgrid(row_blocks(e->rg), batch), cgrid(col_blocks(e->cg), batch);
    const double scale = 1.0 / ((double)Nx * (double)Ny);
    if (dtype == 64) {
      hipLaunchKernelGGL((k_col_fft<double, true>), cgrid, dim3(256), col_lds(e->cg, 16), st, (V2<double>*)e->scratch.p,
                         e->cg, (const V2<double>*)e->twY.p, 1.0, ncplx);
      hipLaunchKernelGGL((k_row_c2r<double, double>), rgrid, dim3(256), row_lds(e->rg, 16), st,
                         (const V2<double>*)e->scratch.p, (double*)out, e->rg, (const V2<double>*)e->twN.p, scale, ncplx,
                         nreal);
    } else {
      hipLaunchKernelGGL((k_col_fft<float, true>), cgrid, dim3(256), col_lds(e->cg, 8), st, (V2<float>*)e->scratch.p,
                         e->cg, (const V2<float>*)e->twY.p, 1.0f, ncplx);
      hipLaunchKernelGGL((k_row_c2r<float, float>), rgrid, dim3(256), row_lds(e->rg, 8), st,
                         (const V2<float>*)e->scratch.p, (float*)out, e->rg, (const V2<float>*)e->twN.p, (float)scale,
                         ncplx, nreal);
    }
    hip_check(hipGetLastError(), "irfft2 launch");
  });
}

}  // extern "C"
