// Kernels of the fused ETD step (generic mixed-radix path).
//
// State layout in HBM.  An instance of Ny x Nx real points is held as the
// row spectrum S~[y][kx] = (1/Nx) * sum_x n[y][x] exp(-2 pi i kx x / Nx),
// kx = 0..Nx/2 (Nc = Nx/2+1 complex per row), stored in the line-tiled order
// of `Lay` below.  Its column FFT is n_hat/Nx, with n_hat =
// numpy.fft.rfft2(n) (ky in fftfreq order).
//
// One ETD1 step = two kernels:
//   row kernel : S~ row -> C2R (length-Nx/2 complex IFFT + pack) -> n
//                -> N(n) pointwise -> R2C -> T row      (reads S~, writes T)
//   col kernel : column tile of S~ and T -> forward column FFTs -> A = n_hat/Nx,
//                B = N_hat -> out = E1 A + Q1 B (k=0 pinned) -> inverse column FFT
//                -> S~ (in place)                        (reads S~ and T, writes S~)
// with E1 = E/Ny and Q1 = P Gamma lap / (Nx Ny) precomputed in FP64 on the host.
// ETDRK2 adds a second row pass and a second column pass; the first column
// pass also writes D = out1 - Q2 B1 (column-major), the second forms
// out2 = D + Q2 B2.
#pragma once
#include "pfc_fft.hpp"

namespace pfc {

enum { MODEL_STD = 0, MODEL_LOG = 1 };

struct Nonlin {
  int model;
  int clamp;      // logPFC: ln(phi) -> ln(minPhi) where phi <= minPhi
  double g;
  double minPhi;
  double logMin;  // ln(minPhi)
  double M;       // stabilising shift: N -= M phi
};

template <typename T>
__device__ __forceinline__ T nonlin_mu(T f, const Nonlin& q) {
  // std: g n^2 + n^3 ; log: ln(phi) + g phi^2 + phi^3 - M phi
  T g = T(q.g);
  T poly = g * (f * f) + (f * f * f);
  if (q.model == MODEL_STD) return poly;
  T lg;
  if (q.clamp) {
    lg = (f > T(q.minPhi)) ? log(f) : T(q.logMin);
  } else {
    lg = log(f);
  }
  return lg + poly - T(q.M) * f;
}

// The nonlinearity about the conserved mean m: N(m + d) - N(m), the same
// function up to a constant, which only the k = 0 mode sees and the update's
// Gamma lap removes.  With d the fluctuation (small where the mean is large)
// every term keeps d's relative precision: the polynomial in Horner form
// about m, the log as log1p(d / m).  Coefficients (FP64, rounded once):
// std  N(m + d) - N(m) = d (c1 + d (c2 + d)),  c1 = 2 g m + 3 m^2, c2 = g + 3 m;
// log  the same polynomial + log1p(d / m) - M d, clamped like nonlin_mu
//      (log1p in Kahan's form, log1p_kahan).
struct NonlinAbout {
  double m, c1, c2, inv_m, clamp_lg;  // clamp_lg = ln(minPhi) - ln(m)
};

__device__ __forceinline__ NonlinAbout nonlin_about(const Nonlin& q, double m) {
  NonlinAbout a;
  a.m = m;
  a.c1 = 2.0 * q.g * m + 3.0 * m * m;
  a.c2 = q.g + 3.0 * m;
  const bool lg = q.model != MODEL_STD && m > 0.0;  // the log terms (logPFC) only
  a.inv_m = lg ? 1.0 / m : 0.0;
  a.clamp_lg = lg ? q.logMin - log(m) : 0.0;
  return a;
}

// log(1 + x) to a few ulp: with u = 1 + x rounded, log(u) x / (u - 1) (W.
// Kahan's formulation; the quotient corrects the rounding of u), x where u
// rounds to 1.  In FP32 log(u) is the hardware log2 times ln 2 (u = phi / m
// is never subnormal where it is used, so the library logf's subnormal
// scaling is dead weight) and the quotient uses the hardware reciprocal:
// ~10 VALU instructions against ~120 for the library log1pf.  Measured on
// gfx950 over 4M arguments in (-0.999, 3): at most 4.0 ulp from the FP64
// log1p, mean 0.56 ulp (the library-logf form: 4.3, 0.60).
__device__ __forceinline__ float log1p_kahan(float x) {
  const float u = 1.0f + x;
  return u == 1.0f ? x
                   : __builtin_amdgcn_logf(u) * ((x * 0.693147180559945309f) * __builtin_amdgcn_rcpf(u - 1.0f));
}
__device__ __forceinline__ double log1p_kahan(double x) { return log1p(x); }

template <typename T>
__device__ __forceinline__ T nonlin_mu_about(T d, const Nonlin& q, const NonlinAbout& a) {
  const T poly = d * (T(a.c1) + d * (T(a.c2) + d));
  if (q.model == MODEL_STD) return poly;
  T lg = log1p_kahan(d * T(a.inv_m));
  if (q.clamp && !(T(a.m) + d > T(q.minPhi))) lg = T(a.clamp_lg);
  return lg + poly - T(q.M) * d;
}

// Instances stepped together by one launch.  blockIdx.y indexes the active
// list (or is the slot itself when the list is null); a slot's arrays sit
// `sstride` (state) / `tstride` (tables) elements from slot 0's.
struct Slots {
  const int* list;    // active slot ids in device memory, or null
  long long sstride;  // complex elements between slots' S~, T and D
  long long tstride;  // elements between slots' ETD tables
  const double* pin;  // per-slot pinned k = 0 value, or null (no pin)
  int* status;        // per-slot status flags, or null
  const double* M;    // per-slot stabilising shift, or null (use Nonlin.M)
};

__device__ __forceinline__ int slot_of(const Slots& s, int y) { return s.list ? s.list[y] : y; }

// Storage order of a (Ny, Nc) half spectrum.  A tile of TX consecutive kx of
// TY consecutive rows fills one 128-byte line (TX * TY elements); the tiles of
// one TY-row group follow each other in kx order, and the groups in y.  A
// block of TY rows (row pass) or of TX columns (column pass) then moves whole
// cache lines, which a row-major array only gives the row pass.  TX = TY = 1
// is row-major with row pitch lpg.
struct Lay {
  int ltx, lty;  // log2 TX, log2 TY
  int lpg;       // tiles per row group: ceil(Nc / TX)
  __host__ __device__ size_t at(int y, int kx) const {
    const size_t tile = (size_t)(y >> lty) * lpg + (kx >> ltx);
    return (tile << (ltx + lty)) + ((size_t)(y & ((1 << lty) - 1)) << ltx) + (kx & ((1 << ltx) - 1));
  }
  // elements of one array of Ny rows (Ny a multiple of TY)
  __host__ __device__ size_t size(int Ny) const { return (size_t)Ny * lpg << ltx; }
};

__host__ __device__ inline Lay row_major(int Nc) { return Lay{0, 0, Nc}; }

struct RowGeom {
  int Nx, M, Nc, Ny;
  Lay lay;  // storage of the half spectrum
  int nl;   // rows per block
  int ld;   // LDS line stride (complex elements), >= Nc
  Radices plM;
};

struct ColGeom {
  int Nc, Ny;
  Lay lay;  // storage of the half spectrum
  int nc;   // columns per block
  int ld;   // LDS line stride, >= Ny
  Radices plY;
};

// ---------------------------------------------------------------------------
// row-side helpers
// ---------------------------------------------------------------------------
// Load rows of the half spectrum (Nc per row) into LDS and form the packed
// length-M spectrum Z[k] = (X[k] + conj X[M-k]) + i w^-k (X[k] - conj X[M-k]),
// whose unnormalised inverse is x[2m] + i x[2m+1] when X = rfft(x)/Nx.
// The imaginary parts of X[0] and X[M] are dropped, as C2R transforms do.
template <typename T, typename Tin>
__device__ __forceinline__ void row_load_pack(const V2<Tin>* __restrict__ S, V2<T>* stage, V2<T>* out,
                                              const RowGeom& g, int row0, int nrows,
                                              const V2<T>* __restrict__ twN) {
  using V = V2<T>;
  const int Nc = g.Nc, M = g.M, ld = g.ld;
  for (int e = threadIdx.x; e < nrows * Nc; e += blockDim.x) {
    const int line = e / Nc;
    const int k = e - line * Nc;
    V2<Tin> v = S[g.lay.at(row0 + line, k)];
    stage[line * ld + k] = V{T(v.x), T(v.y)};
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nrows * M; e += blockDim.x) {
    const int line = e / M;
    const int k = e - line * M;
    V xk = stage[line * ld + k];
    V xm = stage[line * ld + (M - k)];
    if (k == 0) {
      xk.y = T(0);
      xm.y = T(0);
    }
    V s = V{xk.x + xm.x, xk.y - xm.y};  // X[k] + conj X[M-k]
    V d = V{xk.x - xm.x, xk.y + xm.y};  // X[k] - conj X[M-k]
    V w = twN[k];                        // exp(-2 pi i k / Nx)
    V wd = cmul(cconj(w), d);            // w^-k d
    out[line * ld + k] = V{s.x - wd.y, s.y + wd.x};  // s + i wd
  }
  __syncthreads();
}

// Unpack the length-M forward FFT Z of packed real data into the half
// spectrum Y[k] = 1/2 [(Z[k] + conj Z[M-k]) - i w^k (Z[k] - conj Z[M-k])], k=0..M.
template <typename T>
__device__ __forceinline__ V2<T> row_unpack(const V2<T>* Z, int k, int M, const V2<T>* __restrict__ twN) {
  using V = V2<T>;
  V zk = Z[k == M ? 0 : k];
  V zm = Z[k == 0 ? 0 : M - k];
  V s = V{zk.x + zm.x, zk.y - zm.y};
  V d = V{zk.x - zm.x, zk.y + zm.y};
  V wd = cmul(twN[k], d);
  // s - i wd
  return V{T(0.5) * (s.x + wd.y), T(0.5) * (s.y - wd.x)};
}

// The same unpack from the two spectrum values and w = exp(-2 pi i k / Nx).
template <typename T>
__device__ __forceinline__ V2<T> row_unpack_pair(V2<T> zk, V2<T> zm, V2<T> w) {
  using V = V2<T>;
  V s = V{zk.x + zm.x, zk.y - zm.y};
  V d = V{zk.x - zm.x, zk.y + zm.y};
  V wd = cmul(w, d);
  return V{T(0.5) * (s.x + wd.y), T(0.5) * (s.y - wd.x)};
}

__device__ __forceinline__ int xcd_remap(int b, int nb) {
  // Workgroups are dealt round-robin to the 8 XCDs; give each XCD a contiguous
  // run of tiles so neighbouring columns share that XCD's L2.
  const int q = nb >> 3, r = nb & 7;
  const int x = b & 7, l = b >> 3;
  return x < r ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
}

// ---------------------------------------------------------------------------
// row kernel: S~ -> n -> N(n) -> T
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_row_nonlin(const V2<T>* __restrict__ S, V2<T>* __restrict__ Tout,
                                                    RowGeom g, Nonlin q, const V2<T>* __restrict__ twN, Slots sl) {
  using V = V2<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* b0 = reinterpret_cast<V*>(smem_raw);
  V* b1 = b0 + g.nl * g.ld;
  // XCD-aware order over (instance, tile), as in the fast kernels
  const int ntile = gridDim.x;
  const int lin = xcd_remap(blockIdx.x + blockIdx.y * ntile, ntile * gridDim.y);
  const int yi = lin / ntile;
  const int tile = lin - yi * ntile;
  const int slot = slot_of(sl, yi);
  S += (size_t)slot * sl.sstride;
  Tout += (size_t)slot * sl.sstride;
  if (sl.M) q.M = sl.M[slot];
  const int row0 = tile * g.nl;
  const int nrows = min(g.nl, g.Ny - row0);
  row_load_pack<T, T>(S, b1, b0, g, row0, nrows, twN);
  V* z = fft_lines<T, true>(b0, b1, g.plM, nrows, g.ld, twN, 2);
  bool bad = false;
  for (int e = threadIdx.x; e < nrows * g.M; e += blockDim.x) {
    const int line = e / g.M;
    const int m = e - line * g.M;
    V v = z[line * g.ld + m];
    bad |= !(isfinite(v.x) && isfinite(v.y));
    z[line * g.ld + m] = V{nonlin_mu<T>(v.x, q), nonlin_mu<T>(v.y, q)};
  }
  if (bad && sl.status) atomicOr(sl.status + slot, 1);
  __syncthreads();
  V* other = (z == b0) ? b1 : b0;
  V* Z = fft_lines<T, false>(z, other, g.plM, nrows, g.ld, twN, 2);
  for (int e = threadIdx.x; e < nrows * g.Nc; e += blockDim.x) {
    const int line = e / g.Nc;
    const int k = e - line * g.Nc;
    Tout[g.lay.at(row0 + line, k)] = row_unpack<T>(Z + line * g.ld, k, g.M, twN);
  }
}

// ---------------------------------------------------------------------------
// column kernel: the ETD combine between forward and inverse column FFTs
// ---------------------------------------------------------------------------
enum { COL_ETD1 = 0, COL_RK2_STAGE1 = 1, COL_RK2_STAGE2 = 2 };

// The ETD coefficients depend on k^2 only, and ky^2 of rows ky and Ny - ky
// are the same double (fftfreq's negative frequencies negate exactly), so a
// column's table stores rows 0..Ny/2 and row ky reads row ky_half(ky).
__host__ __device__ __forceinline__ int ky_half(int ky, int Ny) { return ky <= Ny - ky ? ky : Ny - ky; }
__host__ __device__ __forceinline__ int table_rows(int Ny) { return Ny / 2 + 1; }

// E = exp(L dt), P = (E - 1)/L where |L dt| > smallZTol else dt (1 + z/2 + z^2/6)
// (the ETD1 coefficients of kernels.py:55-71), L = Gamma lap (beta (q0^2 + lap)^2
// + eps + M), lap = -(kx^2 + ky^2); P2 = dt (e^z - 1 - z)/z^2 (ETDRK2).
// Stored column-major [Nc][Ny], pre-scaled for the row-spectrum state:
// E1 = E/Ny, Q1 = P Gamma lap/(Nx Ny), Q2 = P2 Gamma lap/(Nx Ny); with `fold`
// (ETD1 under the device-managed shift) E1 = (E - M P Gamma lap)/Ny, which
// takes the split's -M phi term out of the row pass.
struct TabParams {
  double Gamma, beta, q0, epsilon, M, dt, smallZTol;
  double inv_ny, inv_n;
  int Nc, Ny;
  int hy;    // stored rows per column: Ny/2 + 1 (see ky_half)
  int fold;  // ETD1 with the shift folded into E1 (the row pass computes N without -M phi)
};

// Per-slot box and time step: slot s reads kx^2 at kx2 + s*Nc, ky^2 at
// ky2 + s*Ny and its dt at dt[s].
struct SlotGrid {
  const double* kx2;
  const double* ky2;
  const double* dt;
};

struct RangePart;
struct ShiftState;

struct ColTables {
  const void* E1;  // column-major [Nc][hy] per slot, FP64 in every precision
  const void* Q1;
  const void* Q2;
  int hy;          // rows stored per column (table_rows(Ny))
  // external-potential source of the first stage, P lap V_hat / (Nx Ny), FP64
  // [Ny][Nc] per slot (cstride apart), or null: no potential
  const V2<double>* C;
  long long cstride;
  // ETD1 under the device-managed shift, decided inside the step: the row
  // kernel raised brk[2 slot + par] where a row tile left the fitted window
  // (or the slot is unfitted); then the fast column kernel takes the field's
  // range from every row tile's pair (part: nrow pairs per active instance,
  // pstride apart), refits M, recomputes its own columns' E1, Q1 (and Q2)
  // from eps, sg and tp with it, uses them and writes them back in place, and
  // tile 0 stores the new fit to st and Ms.  brk null: tables read as stored.
  const int* brk;
  int par;
  const RangePart* part;
  int nrow, pstride;
  ShiftState* st;
  double* Ms;
  double g, minPhi;
  const double* eps;
  SlotGrid sg;
  TabParams tp;
  // FP32 copies of the tables for FP32 engines' fast column kernel (null in
  // FP64 engines): E1f = Ny E1 - 1, so that E1 = (1 + E1f) / Ny in FP64 is
  // within an FP32 rounding of E - 1 (absolute error <= 6e-8 |E - 1| / Ny,
  // where a rounded E1 would be off by up to 6e-8 E / Ny: for the slow modes
  // E is ~1 and E - 1 tiny), Q1f = Q1, Q2f = Q2 rounded once.  Half the
  // bytes of the FP64 tables, which the kernel otherwise reads for every
  // point of the spectrum.
  float* E1f = nullptr;
  float* Q1f = nullptr;
  float* Q2f = nullptr;
};

constexpr int kColRegs = 16;  // A values held per thread between the two forward FFTs


template <typename T>
__device__ __forceinline__ void col_load(const V2<T>* __restrict__ src, V2<T>* buf, const ColGeom& g, int c0,
                                         int ncols) {
  for (int e = threadIdx.x; e < ncols * g.Ny; e += blockDim.x) {
    const int y = e / ncols;
    const int line = e - y * ncols;
    buf[line * g.ld + y] = src[g.lay.at(y, c0 + line)];
  }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_col_etd(V2<T>* __restrict__ S, const V2<T>* __restrict__ Tin,
                                                 V2<T>* __restrict__ D, ColGeom g, ColTables tb,
                                                 const V2<T>* __restrict__ twY, Slots sl) {
  using V = V2<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* b0 = reinterpret_cast<V*>(smem_raw);
  V* b1 = b0 + g.nc * g.ld;
  const int ntile = gridDim.x;
  const int lin = xcd_remap(blockIdx.x + blockIdx.y * ntile, ntile * gridDim.y);
  const int yi = lin / ntile;
  const int slot = slot_of(sl, yi);
  const int c0 = (lin - yi * ntile) * g.nc;
  S += (size_t)slot * sl.sstride;
  Tin += (size_t)slot * sl.sstride;
  if (D) D += (size_t)slot * sl.sstride;
  const int ncols = min(g.nc, g.Nc - c0);
  const int nel = ncols * g.Ny;
  const double* E1 = reinterpret_cast<const double*>(tb.E1) + (size_t)slot * sl.tstride;
  const double* Q1 = reinterpret_cast<const double*>(tb.Q1) + (size_t)slot * sl.tstride;
  const double* Q2 = reinterpret_cast<const double*>(tb.Q2) + (size_t)slot * sl.tstride;

  V a[kColRegs];
  if (MODE != COL_RK2_STAGE2) {
    col_load<T>(S, b0, g, c0, ncols);
    __syncthreads();
    V* A = fft_lines<T, false>(b0, b1, g.plY, ncols, g.ld, twY, 1);
#pragma unroll
    for (int i = 0; i < kColRegs; ++i) {
      const int e = threadIdx.x + i * blockDim.x;
      if (e < nel) {
        const int line = e / g.Ny;
        const int ky = e - line * g.Ny;
        a[i] = A[line * g.ld + ky];
      }
    }
    __syncthreads();
  }
  col_load<T>(Tin, b0, g, c0, ncols);
  __syncthreads();
  V* B = fft_lines<T, false>(b0, b1, g.plY, ncols, g.ld, twY, 1);
  const double* pinp = sl.pin ? sl.pin + slot : nullptr;
  const double pin = pinp ? *pinp : 0.0;
#pragma unroll
  for (int i = 0; i < kColRegs; ++i) {
    const int e = threadIdx.x + i * blockDim.x;
    if (e < nel) {
      const int line = e / g.Ny;
      const int ky = e - line * g.Ny;
      const size_t di = (size_t)(c0 + line) * g.Ny + ky;                      // carry
      const size_t ti = (size_t)(c0 + line) * tb.hy + ky_half(ky, g.Ny);  // tables
      V b = B[line * g.ld + ky];
      const double bx = b.x, by = b.y;
      double ox, oy;
      if (MODE == COL_RK2_STAGE2) {
        const double q2 = Q2[ti];
        V d = D[di];
        ox = double(d.x) + q2 * bx;
        oy = double(d.y) + q2 * by;
      } else {
        const double e1 = E1[ti], q1 = Q1[ti];
        ox = e1 * double(a[i].x) + q1 * bx;
        oy = e1 * double(a[i].y) + q1 * by;
        if (tb.C) {
          const V2<double> c = tb.C[(size_t)slot * tb.cstride + (size_t)ky * g.Nc + (c0 + line)];
          ox += c.x;
          oy += c.y;
        }
        if (MODE == COL_RK2_STAGE1) {
          const double q2 = Q2[ti];
          D[di] = V{T(ox - q2 * bx), T(oy - q2 * by)};
        }
      }
      if (pinp && c0 + line == 0 && ky == 0) ox = pin, oy = 0.0;
      B[line * g.ld + ky] = V{T(ox), T(oy)};
    }
  }
  __syncthreads();
  V* other = (B == b0) ? b1 : b0;
  V* R = fft_lines<T, true>(B, other, g.plY, ncols, g.ld, twY, 1);
  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const int y = e / ncols;
    const int line = e - y * ncols;
    V v = R[line * g.ld + y];
    // The kx = 0 and kx = Nx/2 columns of a real field's row spectrum are
    // real.  Rounding leaves ~1e-17 imaginary parts there; the C2R of the next
    // row pass never sees them, so nothing damps them and unstable modes would
    // grow them as exp(L dt)^n.  Storing only the real part is the projection
    // onto real fields that irfft2 performs every step in the reference.
    if (c0 + line == 0 || c0 + line == g.Nc - 1) v.y = T(0);
    S[g.lay.at(y, c0 + line)] = v;
  }
}

// ---------------------------------------------------------------------------
// resident stepper: small instances, one workgroup each, many ETD1 steps per
// launch with the whole instance in LDS
// ---------------------------------------------------------------------------
// A grid of a few thousand points is latency-bound on the two-kernel step:
// every step pays two launches and two HBM round trips of the state for a
// few microseconds of arithmetic.  Here a workgroup loads its instance's row
// spectrum, twiddles and ETD tables into LDS once, runs `nsteps` ETD1 steps
// (the arithmetic of k_row_nonlin and k_col_etd: the same Stockham passes,
// pack/unpack, nonlinearity and FP64 combine) and stores the state once.
// The phase-diagram driver's cells (e.g. 28 x 24) are this case.
//
// LDS: two regions of 2 Ny Nc complex values.  The state S~ ([y][k], row
// pitch Nc) and T~ share one region; the other is the FFT ping-pong
// partner.  Column transforms run in place of the [y][k] arrays (element
// stride Nc), and the S~ and T~ columns go through one set of passes.
struct ResGeom {
  int Nx, M, Nc, Ny;  // Nx, Nx/2, Nx/2 + 1, Ny
  int hy;             // table rows per column
  Lay lay;            // storage of S~ in device memory
  Radices plM, plY;
};

// n / d for 0 <= n < 2^20 and 1 <= d < 2^20, from inv, a float reciprocal of
// d: the estimate n * inv is within 0.25 of the quotient (relative error below
// 2^-22), so its truncation is the quotient or one off either way, and one
// correction each way makes it exact.  Index math of the resident stepper
// runs on runtime lengths, where a 32-bit division is a ~20-instruction
// sequence; this is five.
__device__ __forceinline__ int qdiv(int n, int d, float inv) {
  int q = __float2int_rz((float)n * inv);
  const int r = n - q * d;
  q += r >= d ? 1 : 0;
  q -= r < 0 ? 1 : 0;
  return q;
}

// One Stockham pass (sk_pass's arithmetic) over nl lines of length L in
// LDS: line l starts at (l % ng) * ld + (l / ng) * gs, elements es apart.
template <typename T, int R, bool INV>
__device__ __forceinline__ void rs_pass(const V2<T>* __restrict__ src, V2<T>* __restrict__ dst, int L, int Ns,
                                        int nl, int ng, int ld, int gs, int es, const V2<T>* __restrict__ tw,
                                        int tws) {
  using V = V2<T>;
  const int nb = L / R;
  const int total = nb * nl;
  const int twstep = (L / (Ns * R)) * tws;
  const int rootstep = nb * tws;
  const float inb = 1.0f / (float)nb, ing = 1.0f / (float)ng, ins = 1.0f / (float)Ns;
  for (int w = threadIdx.x; w < total; w += blockDim.x) {
    const int line = qdiv(w, nb, inb);
    const int j = w - line * nb;
    const int grp = qdiv(line, ng, ing);
    const int off = (line - grp * ng) * ld + grp * gs;
    const V* s = src + off;
    V* d = dst + off;
    const int k = j - qdiv(j, Ns, ins) * Ns;
    V v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[(j + r * nb) * es];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        V t = tw[k * r * twstep];
        if (INV) t.y = -t.y;
        v[r] = cmul(v[r], t);
      }
    }
    Dft<T, R, INV>::run(v, tw, rootstep);
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[(base + r * Ns) * es] = v[r];
  }
}

template <typename T, bool INV>
__device__ V2<T>* rs_fft(V2<T>* buf0, V2<T>* buf1, const Radices& pl, int nl, int ng, int ld, int gs, int es,
                         const V2<T>* __restrict__ tw, int tws) {
  V2<T>* src = buf0;
  V2<T>* dst = buf1;
  int Ns = 1;
  for (int p = 0; p < pl.np; ++p) {
    const int R = pl.r[p];
    switch (R) {
      case 2: rs_pass<T, 2, INV>(src, dst, pl.L, Ns, nl, ng, ld, gs, es, tw, tws); break;
      case 3: rs_pass<T, 3, INV>(src, dst, pl.L, Ns, nl, ng, ld, gs, es, tw, tws); break;
      case 4: rs_pass<T, 4, INV>(src, dst, pl.L, Ns, nl, ng, ld, gs, es, tw, tws); break;
      case 5: rs_pass<T, 5, INV>(src, dst, pl.L, Ns, nl, ng, ld, gs, es, tw, tws); break;
      case 7: rs_pass<T, 7, INV>(src, dst, pl.L, Ns, nl, ng, ld, gs, es, tw, tws); break;
      case 8: rs_pass<T, 8, INV>(src, dst, pl.L, Ns, nl, ng, ld, gs, es, tw, tws); break;
      case 11: rs_pass<T, 11, INV>(src, dst, pl.L, Ns, nl, ng, ld, gs, es, tw, tws); break;
      default: rs_pass<T, 13, INV>(src, dst, pl.L, Ns, nl, ng, ld, gs, es, tw, tws); break;
    }
    __syncthreads();
    Ns *= R;
    V2<T>* t = src;
    src = dst;
    dst = t;
  }
  return src;
}

// LDS bytes of the resident stepper: tables (FP64, padded to 16 B), twiddles
// and the two state regions.
__host__ __device__ inline size_t resident_table_bytes(int Nc, int hy) {
  return ((size_t)2 * Nc * hy * sizeof(double) + 15) & ~(size_t)15;
}

// `nsteps` steps per workgroup, or, with `counts` (per active slot, in
// active order), min(counts[i] - done, nsteps) steps for the i-th: each
// instance can take its own number of steps in one launch.
template <typename T, int BS>
__global__ void __launch_bounds__(BS) k_resident_etd1(V2<T>* __restrict__ S, ResGeom g, Nonlin q, ColTables tb,
                                                      const V2<T>* __restrict__ twN, const V2<T>* __restrict__ twY,
                                                      Slots sl, long long nsteps, const long long* __restrict__ counts,
                                                      long long done) {
  if (counts) nsteps = max(0LL, min(nsteps, counts[blockIdx.x] - done));
  if (nsteps == 0) return;
  using V = V2<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int M = g.M, Nc = g.Nc, Ny = g.Ny, nsp = Ny * Nc, nt = Nc * g.hy;
  double* tE = reinterpret_cast<double*>(smem_raw);
  double* tQ = tE + nt;
  V* twn = reinterpret_cast<V*>(smem_raw + resident_table_bytes(Nc, g.hy));
  V* twy = twn + g.Nx;
  V* P0 = twy + Ny;  // region 0: [S~ | T~] while the state lives here
  V* P1 = P0 + 2 * nsp;
  const int slot = slot_of(sl, blockIdx.x);
  S += (size_t)slot * sl.sstride;
  if (sl.M) q.M = sl.M[slot];
  const double* E1 = reinterpret_cast<const double*>(tb.E1) + (size_t)slot * sl.tstride;
  const double* Q1 = reinterpret_cast<const double*>(tb.Q1) + (size_t)slot * sl.tstride;
  const bool pinned = sl.pin != nullptr;
  const double pin = pinned ? sl.pin[slot] : 0.0;
  for (int e = threadIdx.x; e < nt; e += blockDim.x) {
    tE[e] = E1[e];
    tQ[e] = Q1[e];
  }
  for (int e = threadIdx.x; e < g.Nx; e += blockDim.x) twn[e] = twN[e];
  for (int e = threadIdx.x; e < Ny; e += blockDim.x) twy[e] = twY[e];
  const float iM = 1.0f / (float)M, iNc = 1.0f / (float)Nc;
  for (int e = threadIdx.x; e < nsp; e += blockDim.x) {
    const int y = qdiv(e, Nc, iNc);
    P0[e] = S[g.lay.at(y, e - y * Nc)];
  }
  __syncthreads();
  V* X = P0;  // the region holding S~ (and T~ after the row pass)
  V* Y = P1;
  bool bad = false;
  for (long long step = 0; step < nsteps; ++step) {
    // row pass: S~ -> n -> N(n) -> T~.  The pack also stores the real
    // projection of the kx = 0 and kx = Nx/2 columns (k_col_etd stores it
    // after its inverse transform; the column pass below reads it here).
    for (int e = threadIdx.x; e < Ny * M; e += blockDim.x) {
      const int line = qdiv(e, M, iM);
      const int k = e - line * M;
      V xk = X[line * Nc + k];
      V xm = X[line * Nc + (M - k)];
      if (k == 0) {
        xk.y = T(0);
        xm.y = T(0);
        X[line * Nc] = xk;
        X[line * Nc + M] = xm;
      }
      V sm = V{xk.x + xm.x, xk.y - xm.y};
      V df = V{xk.x - xm.x, xk.y + xm.y};
      V wd = cmul(cconj(twn[k]), df);
      Y[line * Nc + k] = V{sm.x - wd.y, sm.y + wd.x};
    }
    __syncthreads();
    V* z = rs_fft<T, true>(Y, Y + nsp, g.plM, Ny, Ny, Nc, 0, 1, twn, 2);
    for (int e = threadIdx.x; e < Ny * M; e += blockDim.x) {
      const int line = qdiv(e, M, iM);
      const int m = e - line * M;
      V v = z[line * Nc + m];
      bad |= !(isfinite(v.x) && isfinite(v.y));
      z[line * Nc + m] = V{nonlin_mu<T>(v.x, q), nonlin_mu<T>(v.y, q)};
    }
    __syncthreads();
    V* Z = rs_fft<T, false>(z, z == Y ? Y + nsp : Y, g.plM, Ny, Ny, Nc, 0, 1, twn, 2);
    for (int e = threadIdx.x; e < nsp; e += blockDim.x) {
      const int line = qdiv(e, Nc, iNc);
      X[nsp + e] = row_unpack<T>(Z + line * Nc, e - line * Nc, M, twn);
    }
    __syncthreads();
    // column pass: the Nc columns of S~ and of T~ (2 Nc lines, stride Nc)
    // in one forward transform, out = E1 A + Q1 B (k = 0 pinned) into A's
    // place, inverse transform of those Nc columns
    V* W = rs_fft<T, false>(X, Y, g.plY, 2 * Nc, Nc, 1, nsp, Nc, twy, 1);
    for (int e = threadIdx.x; e < nsp; e += blockDim.x) {
      const int ky = qdiv(e, Nc, iNc);
      const int line = e - ky * Nc;
      const int ti = line * g.hy + ky_half(ky, Ny);
      const V a = W[e], b = W[nsp + e];
      const double e1 = tE[ti], q1 = tQ[ti];
      double ox = e1 * double(a.x) + q1 * double(b.x);
      double oy = e1 * double(a.y) + q1 * double(b.y);
      if (tb.C) {
        const V2<double> c = tb.C[(size_t)slot * tb.cstride + e];
        ox += c.x;
        oy += c.y;
      }
      if (pinned && line == 0 && ky == 0) ox = pin, oy = 0.0;
      W[e] = V{T(ox), T(oy)};
    }
    __syncthreads();
    X = rs_fft<T, true>(W, W == X ? Y : X, g.plY, Nc, Nc, 1, 0, Nc, twy, 1);
    Y = X == P0 ? P1 : P0;
  }
  for (int e = threadIdx.x; e < nsp; e += blockDim.x) {
    const int y = qdiv(e, Nc, iNc);
    const int k = e - y * Nc;
    V v = X[e];
    if (k == 0 || k == Nc - 1) v.y = T(0);
    S[g.lay.at(y, k)] = v;
  }
  if (bad && sl.status) atomicOr(sl.status + slot, 1);
}

// ---------------------------------------------------------------------------
// dense resident stepper: the smallest cells, three LDS phases per step
// ---------------------------------------------------------------------------
// For cells of a few hundred points the FFT stepper above is a chain of 12
// short LDS phases (pack, two row passes each way, nonlinearity, unpack, two
// column passes each way, combine), several of them with fewer work items
// than one wave (a radix-7 pass of a 28-point row has 48).  Here one thread
// owns one (row y, index j) pair, j = 0..Nx/2, and a step is three phases
// with every thread busy:
//   1. the two real-space values at x = j and x = Nx - j of row y from the
//      half spectrum (cosine and sine sums over kx, sharing one set of
//      angles 2 pi kx j / Nx), and their nonlinearity;
//   2. T~(y, j) = sum_x N(y, x) exp(-2 pi i j x / Nx), pairing x with Nx - x
//      (the same angles);
//   3. the whole column step at once: the column transform, the ETD1
//      combine and the inverse column transform form, per column j, the map
//      S~'(., j) = A_j S~(., j) + B_j T~(., j) with A_j, B_j the circulant
//      matrices of the inverse transforms of E1(j, .) and Q1(j, .).  E1 and Q1
//      depend on ky only through ky^2, so these are real and symmetric:
//      A_j[y][y'] = a_j((y - y') mod Ny), a_j(d) = sum_ky E1(j, ky) cos(2 pi ky d / Ny).
// The angles of phases 1-2 and the thread's row of A_j and B_j are set up
// once per launch and held in registers.  The k = 0 pin drops the ky = 0
// term from column 0's A and B and adds the pinned value to every row of
// that column (the unnormalised inverse transform of a lone k = 0 value);
// the external-potential source adds its column-inverse transform, also set
// up once.  In exact arithmetic this is the FFT stepper's step; the results
// differ from it by rounding only.  LDS arrays are column-major ([j][y]) so
// that a wave's lanes (consecutive y) touch consecutive words.
//
// NY and MH bound Ny and Nx/2 - 1 at compile time: the sums run over NY rows
// and MH angles with zero coefficients past Ny and Nx/2 - 1 and their loads
// clamped to a valid element, so every loop body is branch-free and the
// compiler can issue a phase's LDS loads ahead of its FMAs.
constexpr int kDenseThreads = 512;  // (Nx/2 + 1) Ny work items at most
constexpr int kDenseNyMax = 32;
constexpr int kDenseMhMax = 15;

__host__ __device__ inline bool dense_fits(int Nx, int Ny) {
  return Ny <= kDenseNyMax && Nx / 2 - 1 <= kDenseMhMax && (Nx / 2 + 1) * Ny <= kDenseThreads;
}

// LDS bytes of the dense stepper: the circulant coefficients and the trig
// tables (FP64), two state buffers and T~ (complex), the nonlinearity (real).
__host__ __device__ inline size_t dense_lds_bytes(int Nx, int Ny, size_t csize) {
  const size_t nsp = (size_t)(Nx / 2 + 1) * Ny;
  return (2 * nsp + 2 * (size_t)(Nx + Ny)) * sizeof(double) + 3 * nsp * csize + (size_t)Nx * Ny * (csize / 2);
}

template <typename T, int NY, int MH>
__global__ void __launch_bounds__(kDenseThreads) k_resident_dense(V2<T>* __restrict__ S, ResGeom g, Nonlin q,
                                                                  ColTables tb, Slots sl, long long nsteps,
                                                                  const long long* __restrict__ counts,
                                                                  long long done) {
  if (counts) nsteps = max(0LL, min(nsteps, counts[blockIdx.x] - done));
  if (nsteps == 0) return;
  using V = V2<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Nx = g.Nx, M = g.M, Nc = g.Nc, Ny = g.Ny, nsp = Nc * Ny;
  double* ca = reinterpret_cast<double*>(smem_raw);  // [Nc][Ny] a_j(d)
  double* cb = ca + nsp;                             // [Nc][Ny] b_j(d)
  double* cx = cb + nsp;                             // cos, sin (2 pi m / Nx), m < Nx
  double* sx = cx + Nx;
  double* cy = sx + Nx;                              // cos, sin (2 pi m / Ny), m < Ny
  double* sy = cy + Ny;
  V* Sa = reinterpret_cast<V*>(sy + Ny);             // state [j][y], two buffers
  V* Sb = Sa + nsp;
  V* Tt = Sb + nsp;                                  // T~ [j][y]
  T* Nb = reinterpret_cast<T*>(Tt + nsp);            // N [x][y]
  const int tid = threadIdx.x;
  const bool live = tid < nsp;
  const int j = live ? tid / Ny : 0;
  const int y = live ? tid - j * Ny : 0;
  const int slot = slot_of(sl, blockIdx.x);
  S += (size_t)slot * sl.sstride;
  if (sl.M) q.M = sl.M[slot];
  const double* E1 = reinterpret_cast<const double*>(tb.E1) + (size_t)slot * sl.tstride;
  const double* Q1 = reinterpret_cast<const double*>(tb.Q1) + (size_t)slot * sl.tstride;
  const bool pinned = sl.pin != nullptr;
  const double pin = pinned ? sl.pin[slot] : 0.0;

  // set-up: trig tables, the state (column-major, kx = 0 and Nx/2 columns
  // projected on the reals)
  for (int m = tid; m < Nx; m += blockDim.x) {
    cx[m] = cospi(2.0 * m / Nx);
    sx[m] = sinpi(2.0 * m / Nx);
  }
  for (int m = tid; m < Ny; m += blockDim.x) {
    cy[m] = cospi(2.0 * m / Ny);
    sy[m] = sinpi(2.0 * m / Ny);
  }
  for (int e = tid; e < nsp; e += blockDim.x) {
    const int yy = e / Nc;
    const int k = e - yy * Nc;
    V v = S[g.lay.at(yy, k)];
    if (k == 0 || k == M) v.y = T(0);
    Sa[k * Ny + yy] = v;
  }
  __syncthreads();
  // circulant coefficients of every column
  for (int e = tid; e < nsp; e += blockDim.x) {
    const int k = e / Ny;
    const int d = e - k * Ny;
    double a = 0.0, b = 0.0;
    for (int ky = 0; ky < Ny; ++ky) {
      if (pinned && k == 0 && ky == 0) continue;
      const int ti = k * g.hy + ky_half(ky, Ny);
      const double c = cy[(ky * d) % Ny];
      a = fma(E1[ti], c, a);
      b = fma(Q1[ti], c, b);
    }
    ca[e] = a;
    cb[e] = b;
  }
  // this thread's angles 2 pi k j / Nx, k = 1..MH (zero past Nx/2 - 1)
  T cj[MH], sj[MH];
#pragma unroll
  for (int k = 0; k < MH; ++k) {
    const int m = ((k + 1) * j) % Nx;
    cj[k] = k + 1 < M ? T(cx[m]) : T(0);
    sj[k] = k + 1 < M ? T(sx[m]) : T(0);
  }
  // the external-potential source's column-inverse transform at (y, j)
  double srcx = 0.0, srcy = 0.0;
  if (tb.C && live) {
    const V2<double>* C = tb.C + (size_t)slot * tb.cstride;
    for (int ky = 0; ky < Ny; ++ky) {
      if (pinned && j == 0 && ky == 0) continue;
      const V2<double> c = C[ky * Nc + j];
      const int m = (ky * y) % Ny;
      srcx += c.x * cy[m] - c.y * sy[m];
      srcy += c.x * sy[m] + c.y * cy[m];
    }
  }
  if (pinned && j == 0) srcx += pin;
  if (j == 0 || j == M) srcy = 0.0;
  __syncthreads();
  // this thread's rows of A_j and B_j: r[y'] = coefficient at (y - y') mod Ny
  // (zero past Ny)
  double ra[NY], rb[NY];
#pragma unroll
  for (int yp = 0; yp < NY; ++yp) {
    int d = y - yp;
    d += d < 0 ? Ny : 0;
    const bool in = live && yp < Ny;
    ra[yp] = in ? ca[j * Ny + d] : 0.0;
    rb[yp] = in ? cb[j * Ny + d] : 0.0;
  }
  const T sgn = (j & 1) ? T(-1) : T(1);  // (-1)^j
  const bool edge = j == 0 || j == M;
  bool bad = false;
  V* X = Sa;
  V* Y = Sb;
  for (long long step = 0; step < nsteps; ++step) {
    // phase 1: n(y, j) and n(y, Nx - j), then the nonlinearity
    if (live) {
      const T base = X[y].x + sgn * X[M * Ny + y].x;
      T P = T(0), Qs = T(0);
#pragma unroll
      for (int k = 0; k < MH; ++k) {
        const int kk = k + 1 < M ? k + 1 : 0;
        const V s = X[kk * Ny + y];
        P = fma(s.x, cj[k], P);
        Qs = fma(s.y, sj[k], Qs);
      }
      const T n0 = base + T(2) * (P - Qs);
      const T n1 = base + T(2) * (P + Qs);
      bad |= !(isfinite(n0) && isfinite(n1));
      Nb[j * Ny + y] = nonlin_mu<T>(n0, q);
      if (!edge) Nb[(Nx - j) * Ny + y] = nonlin_mu<T>(n1, q);
    }
    __syncthreads();
    // phase 2: T~(y, j)
    if (live) {
      T re = Nb[y] + sgn * Nb[M * Ny + y], im = T(0);
#pragma unroll
      for (int k = 0; k < MH; ++k) {
        const bool in = k + 1 < M;
        const T a = Nb[(in ? k + 1 : 0) * Ny + y], b = Nb[(in ? Nx - k - 1 : 0) * Ny + y];
        re = fma(a + b, cj[k], re);
        im = fma(b - a, sj[k], im);
      }
      Tt[j * Ny + y] = V{re, edge ? T(0) : im};
    }
    __syncthreads();
    // phase 3: the column step as one circulant map per column
    if (live) {
      double ox = srcx, oy = srcy;
      const V* sc = X + j * Ny;
      const V* tc = Tt + j * Ny;
#pragma unroll
      for (int yp = 0; yp < NY; ++yp) {
        const int yy = yp < Ny ? yp : 0;
        const V s = sc[yy], t = tc[yy];
        ox = fma(ra[yp], double(s.x), ox);
        oy = fma(ra[yp], double(s.y), oy);
        ox = fma(rb[yp], double(t.x), ox);
        oy = fma(rb[yp], double(t.y), oy);
      }
      Y[j * Ny + y] = V{T(ox), edge ? T(0) : T(oy)};
    }
    __syncthreads();
    V* t = X;
    X = Y;
    Y = t;
  }
  for (int e = tid; e < nsp; e += blockDim.x) {
    const int yy = e / Nc;
    const int k = e - yy * Nc;
    S[g.lay.at(yy, k)] = X[k * Ny + yy];
  }
  if (bad && sl.status) atomicOr(sl.status + slot, 1);
}

// ---------------------------------------------------------------------------
// spectral diagnostics of the phase-diagram driver (one workgroup per slot)
// ---------------------------------------------------------------------------
// From the row spectrum S~ alone: A = FFT_y of each column in FP64 in LDS,
// n_hat(ky, kx) = Nx A[kx][ky] = rfft2(n).  Per slot, out[0..3] =
//   mean  = n_hat(0, 0) / N,
//   m0    = sum_k |n_hat / N|^2            (= mean of n^2, Parseval),
//   m1,m2 = sum_k |n_hat / N|^2 |k|^2, |k|^4 (the slot's own k-grid),
// sums over the full spectrum (Hermitian weight 2 off the kx = 0 and Nx/2
// columns), double-double in a fixed order; then out[4 + 2 e .. 5 + 2 e] =
// n_hat at the e-th requested (ky, kx).  These give the driver's 3-mode
// fitness and amplitudes and the frozen-field moments without moving the
// field off the device.
struct DiagParams {
  int Nx, Nc, Ny;
  Lay lay;
  Radices plY;
  int nent;           // requested entries
  const int* ky;      // [nent] in 0..Ny-1
  const int* kx;      // [nent] in 0..Nc-1
  const double* kx2;  // [slot][Nc]
  const double* ky2;  // [slot][Ny]
};

template <typename T>
__global__ void __launch_bounds__(256) k_diagnose(const V2<T>* __restrict__ S, DiagParams p,
                                                  const V2<double>* __restrict__ twY, Slots sl, double* __restrict__ out) {
  using V = V2<double>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Nc = p.Nc, Ny = p.Ny, nsp = Nc * Ny;
  V* b0 = reinterpret_cast<V*>(smem_raw);
  V* b1 = b0 + nsp;
  DD* red = reinterpret_cast<DD*>(b1 + nsp);  // [3][blockDim.x]
  const int slot = slot_of(sl, blockIdx.x);
  S += (size_t)slot * sl.sstride;
  const double* kx2 = p.kx2 + (size_t)slot * Nc;
  const double* ky2 = p.ky2 + (size_t)slot * Ny;
  out += (size_t)blockIdx.x * (4 + 2 * p.nent);
  for (int e = threadIdx.x; e < nsp; e += blockDim.x) {
    const int y = e / Nc;
    const int k = e - y * Nc;
    const V2<T> v = S[p.lay.at(y, k)];
    b0[k * Ny + y] = V{double(v.x), double(v.y)};
  }
  __syncthreads();
  const V* A = fft_lines<double, false>(b0, b1, p.plY, Nc, Ny, twY, 1);
  DD m[3] = {DD{0.0, 0.0}, DD{0.0, 0.0}, DD{0.0, 0.0}};
  for (int e = threadIdx.x; e < nsp; e += blockDim.x) {
    const int k = e / Ny;
    const int y = e - k * Ny;
    const V a = A[e];
    const double w = (k == 0 || 2 * k == p.Nx) ? 1.0 : 2.0;
    const double pw = w * (a.x * a.x + a.y * a.y);
    const double k2 = kx2[k] + ky2[y];
    m[0] = dd_add(m[0], pw);
    m[1] = dd_add(m[1], pw * k2);
    m[2] = dd_add(m[2], pw * (k2 * k2));
  }
  for (int i = 0; i < 3; ++i) red[i * blockDim.x + threadIdx.x] = m[i];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int i = 0; i < 3; ++i) {
        DD* r = red + i * blockDim.x + threadIdx.x;
        r[0] = dd_add(r[0], r[s]);
      }
    __syncthreads();
  }
  // |n_hat / N|^2 = |A|^2 / Ny^2
  const double inv = 1.0 / ((double)Ny * (double)Ny);
  if (threadIdx.x == 0) {
    out[0] = A[0].x / Ny;
    for (int i = 0; i < 3; ++i) out[1 + i] = (red[i * blockDim.x].hi + red[i * blockDim.x].lo) * inv;
  }
  for (int e = threadIdx.x; e < p.nent; e += blockDim.x) {
    const V a = A[p.kx[e] * Ny + p.ky[e]];
    out[4 + 2 * e] = a.x * p.Nx;
    out[5 + 2 * e] = a.y * p.Nx;
  }
}

// ---------------------------------------------------------------------------
// plain transforms (field upload/download and the backend rfft2/irfft2)
// ---------------------------------------------------------------------------
// real rows (Nx) -> half spectrum rows, times `scale`.  blockIdx.y = batch.
template <typename T, typename Tin>
__global__ void __launch_bounds__(256) k_row_r2c(const Tin* __restrict__ x, V2<T>* __restrict__ S, RowGeom g,
                                                 const V2<T>* __restrict__ twN, T scale, size_t in_bstride,
                                                 size_t out_bstride) {
  using V = V2<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* b0 = reinterpret_cast<V*>(smem_raw);
  V* b1 = b0 + g.nl * g.ld;
  // XCD-aware order over (instance, tile), as in the fast kernels
  const int ntile = gridDim.x;
  const int lin = xcd_remap(blockIdx.x + blockIdx.y * ntile, ntile * gridDim.y);
  const int yi = lin / ntile;
  const int tile = lin - yi * ntile;
  x += yi * in_bstride;
  S += yi * out_bstride;
  const int row0 = tile * g.nl;
  const int nrows = min(g.nl, g.Ny - row0);
  for (int e = threadIdx.x; e < nrows * g.M; e += blockDim.x) {
    const int line = e / g.M;
    const int m = e - line * g.M;
    const Tin* xr = x + (size_t)(row0 + line) * g.Nx;
    b0[line * g.ld + m] = V{T(xr[2 * m]), T(xr[2 * m + 1])};
  }
  __syncthreads();
  V* Z = fft_lines<T, false>(b0, b1, g.plM, nrows, g.ld, twN, 2);
  for (int e = threadIdx.x; e < nrows * g.Nc; e += blockDim.x) {
    const int line = e / g.Nc;
    const int k = e - line * g.Nc;
    V y = row_unpack<T>(Z + line * g.ld, k, g.M, twN);
    S[g.lay.at(row0 + line, k)] = V{y.x * scale, y.y * scale};
  }
}

// half spectrum rows -> real rows, unnormalised inverse times `scale`.
template <typename T, typename Tout>
__global__ void __launch_bounds__(256) k_row_c2r(const V2<T>* __restrict__ S, Tout* __restrict__ x, RowGeom g,
                                                 const V2<T>* __restrict__ twN, T scale, size_t in_bstride,
                                                 size_t out_bstride, const int* __restrict__ list) {
  using V = V2<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* b0 = reinterpret_cast<V*>(smem_raw);
  V* b1 = b0 + g.nl * g.ld;
  // XCD-aware order over (instance, tile), as in the fast kernels
  const int ntile = gridDim.x;
  const int lin = xcd_remap(blockIdx.x + blockIdx.y * ntile, ntile * gridDim.y);
  const int yi = lin / ntile;
  const int tile = lin - yi * ntile;
  S += (size_t)(list ? list[yi] : yi) * in_bstride;  // input instance: list entry or index
  x += yi * out_bstride;
  const int row0 = tile * g.nl;
  const int nrows = min(g.nl, g.Ny - row0);
  row_load_pack<T, T>(S, b1, b0, g, row0, nrows, twN);
  V* z = fft_lines<T, true>(b0, b1, g.plM, nrows, g.ld, twN, 2);
  for (int e = threadIdx.x; e < nrows * g.M; e += blockDim.x) {
    const int line = e / g.M;
    const int m = e - line * g.M;
    V v = z[line * g.ld + m];
    Tout* xr = x + (size_t)(row0 + line) * g.Nx;
    xr[2 * m] = Tout(v.x * scale);
    xr[2 * m + 1] = Tout(v.y * scale);
  }
}

// column FFT of a (Ny, Nc) complex array stored as g.lay, in place, times `scale`.
template <typename T, bool INV>
__global__ void __launch_bounds__(256) k_col_fft(V2<T>* __restrict__ S, ColGeom g, const V2<T>* __restrict__ twY,
                                                 T scale, size_t bstride) {
  using V = V2<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* b0 = reinterpret_cast<V*>(smem_raw);
  V* b1 = b0 + g.nc * g.ld;
  // XCD-aware order over (instance, tile), as in the fast kernels
  const int ntile = gridDim.x;
  const int lin = xcd_remap(blockIdx.x + blockIdx.y * ntile, ntile * gridDim.y);
  const int yi = lin / ntile;
  const int tile = lin - yi * ntile;
  S += yi * bstride;
  const int c0 = tile * g.nc;
  const int ncols = min(g.nc, g.Nc - c0);
  col_load<T>(S, b0, g, c0, ncols);
  __syncthreads();
  V* R = fft_lines<T, INV>(b0, b1, g.plY, ncols, g.ld, twY, 1);
  for (int e = threadIdx.x; e < ncols * g.Ny; e += blockDim.x) {
    const int y = e / ncols;
    const int line = e - y * ncols;
    V v = R[line * g.ld + y];
    S[g.lay.at(y, c0 + line)] = V{v.x * scale, v.y * scale};
  }
}

// ---------------------------------------------------------------------------
// diagnostics: chemical potential and Gaussian-smoothed fields
// ---------------------------------------------------------------------------
// Column pass producing row spectra ready for k_row_c2r (scale 1/(Nx Ny)):
//   FILTER_MU    out0 = IFFT_y(mu_hat), out1 = IFFT_y(G mu_hat),
//                mu_hat = L_mu n_hat + N_hat, n_hat = Nx FFT_y(S~), N_hat = FFT_y(T)
//   FILTER_LOCAL out0 = IFFT_y(G n_hat)
// with L_mu = beta (q0^2 + lap)^2 + eps + M and G = exp(-k^2 w^2 / 2) evaluated in FP64.
enum { FILTER_MU = 0, FILTER_LOCAL = 1 };

struct FilterParams {
  double beta, q0, epsilon, w;
  const double* M;  // the slot's stabilising shift in device memory (the value the step uses)
  int Nx;
  const double* kx2;
  const double* ky2;
};

template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_col_filter(const V2<T>* __restrict__ S, const V2<T>* __restrict__ Tin,
                                                    V2<T>* __restrict__ out0, V2<T>* __restrict__ out1, ColGeom g,
                                                    FilterParams fp, const V2<T>* __restrict__ twY) {
  using V = V2<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* b0 = reinterpret_cast<V*>(smem_raw);
  V* b1 = b0 + g.nc * g.ld;
  // XCD-aware order over (instance, tile), as in the fast kernels
  const int ntile = gridDim.x;
  const int lin = xcd_remap(blockIdx.x + blockIdx.y * ntile, ntile * gridDim.y);
  const int yi = lin / ntile;
  const int tile = lin - yi * ntile;
  const int c0 = tile * g.nc;
  const int ncols = min(g.nc, g.Nc - c0);
  const int nel = ncols * g.Ny;
  V a[kColRegs];
  col_load<T>(S, b0, g, c0, ncols);
  __syncthreads();
  V* A = fft_lines<T, false>(b0, b1, g.plY, ncols, g.ld, twY, 1);
#pragma unroll
  for (int i = 0; i < kColRegs; ++i) {
    const int e = threadIdx.x + i * blockDim.x;
    if (e < nel) {
      const int line = e / g.Ny;
      const int ky = e - line * g.Ny;
      a[i] = A[line * g.ld + ky];
    }
  }
  __syncthreads();
  V* B = nullptr;
  if (MODE == FILTER_MU) {
    col_load<T>(Tin, b0, g, c0, ncols);
    __syncthreads();
    B = fft_lines<T, false>(b0, b1, g.plY, ncols, g.ld, twY, 1);
  } else {
    B = b0;
  }
  // mu_hat (or G n_hat) goes to B; the G-filtered copy of mu_hat stays in
  // registers for the second output
  V gm[kColRegs];
#pragma unroll
  for (int i = 0; i < kColRegs; ++i) {
    const int e = threadIdx.x + i * blockDim.x;
    if (e < nel) {
      const int line = e / g.Ny;
      const int ky = e - line * g.Ny;
      const int kx = c0 + line;
      const double k2 = fp.kx2[kx] + fp.ky2[ky];
      const double t = fp.q0 * fp.q0 - k2;
      const double lmu = fp.beta * (t * t) + fp.epsilon + *fp.M;
      const double gg = exp(-k2 * fp.w * fp.w / 2);
      const double nx = double(fp.Nx);
      if (MODE == FILTER_MU) {
        V b = B[line * g.ld + ky];
        V mu = V{T(lmu * nx * a[i].x + b.x), T(lmu * nx * a[i].y + b.y)};
        B[line * g.ld + ky] = mu;
        gm[i] = V{T(gg * mu.x), T(gg * mu.y)};
      } else {
        B[line * g.ld + ky] = V{T(gg * nx * a[i].x), T(gg * nx * a[i].y)};
      }
    }
  }
  __syncthreads();
  V* other = (B == b0) ? b1 : b0;
  V* R = fft_lines<T, true>(B, other, g.plY, ncols, g.ld, twY, 1);
  for (int e = threadIdx.x; e < nel; e += blockDim.x) {
    const int y = e / ncols;
    const int line = e - y * ncols;
    out0[g.lay.at(y, c0 + line)] = R[line * g.ld + y];
  }
  if (MODE == FILTER_MU) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kColRegs; ++i) {
      const int e = threadIdx.x + i * blockDim.x;
      if (e < nel) {
        const int line = e / g.Ny;
        const int ky = e - line * g.Ny;
        b0[line * g.ld + ky] = gm[i];
      }
    }
    __syncthreads();
    V* R2 = fft_lines<T, true>(b0, b1, g.plY, ncols, g.ld, twY, 1);
    for (int e = threadIdx.x; e < nel; e += blockDim.x) {
      const int y = e / ncols;
      const int line = e - y * ncols;
      out1[g.lay.at(y, c0 + line)] = R2[line * g.ld + y];
    }
  }
}

// ---------------------------------------------------------------------------
// ETD tables, built on the device in FP64 and stored in the bulk precision
// ---------------------------------------------------------------------------
__device__ __forceinline__ double etd2_phi_dt(double z, double dt) {
  if (fabs(z) < 0.1) {
    double s = 0.0, t = 0.5;
    for (int j = 0; j < 16; ++j) {
      s += t;
      t = t * z / (j + 3);
    }
    return dt * s;
  }
  return dt * ((expm1(z) - z) / (z * z));
}

// E = exp(L dt) and P = (E - 1)/L (dt (1 + z/2 + z^2/6) for |z| <= smallZTol,
// z = L dt) of L = Gamma lap (beta (q0^2 + lap)^2 + eps + M): the reference's
// kernel_etd (discretization/spectral/kernels.py:55-71) in FP64.
struct EtdCoef {
  double L, z, E, P;
};
__device__ __forceinline__ EtdCoef etd_coef(double lap, const TabParams& p) {
  const double sq = p.q0 * p.q0 + lap;
  double lmu = p.beta * (sq * sq) + p.epsilon;
  if (p.M != 0.0) lmu = lmu + p.M;
  EtdCoef c;
  c.L = p.Gamma * lap * lmu;
  c.z = c.L * p.dt;
  c.E = exp(c.z);
  c.P = fabs(c.z) > p.smallZTol ? (c.E - 1) / c.L : p.dt * (1 + c.z / 2 + c.z * c.z / 6);
  return c;
}

// One stored table entry at lap = -(kx^2 + ky^2): E1, Q1 and (q2) Q2 as
// described above; k_tables and the fast ETD1 column kernel's in-place
// rebuild both compute their entries here.
struct TabEntry {
  double e1, q1, q2;
  double em1;  // Ny e1 - 1 (the E1f form)
};
__device__ __forceinline__ TabEntry tab_entry(double lap, const TabParams& p, bool q2) {
  const EtdCoef k = etd_coef(lap, p);
  // folded shift: E phi + P Gamma lap (N - M phi) = (E - M P Gamma lap) phi + P Gamma lap N
  const double e = p.fold ? k.E - p.M * k.P * (p.Gamma * lap) : k.E;
  TabEntry t;
  t.e1 = e * p.inv_ny;
  t.em1 = e - 1.0;
  t.q1 = k.P * (p.Gamma * lap) * p.inv_n;
  t.q2 = q2 ? etd2_phi_dt(k.z, p.dt) * (p.Gamma * lap) * p.inv_n : 0.0;
  return t;
}

// blockIdx.y = slot - slot0; per-slot epsilon, M, box and dt come from
// eps[], Ms[] and the SlotGrid.
template <typename T>
__global__ void __launch_bounds__(256) k_tables(T* __restrict__ E1, T* __restrict__ Q1, T* __restrict__ Q2,
                                                long long tstride, SlotGrid sg, TabParams p,
                                                const double* __restrict__ eps, const double* __restrict__ Ms,
                                                int slot0, const int* __restrict__ list = nullptr,
                                                const int* __restrict__ rebuild = nullptr, float* __restrict__ E1f = nullptr,
                                                float* __restrict__ Q1f = nullptr, float* __restrict__ Q2f = nullptr) {
  // slots slot0 + blockIdx.y, or list[blockIdx.y]; with `rebuild`, only the
  // slots whose flag is set
  const int slot = list ? list[blockIdx.y] : slot0 + (int)blockIdx.y;
  if (rebuild && !rebuild[slot]) return;
  p.epsilon = eps[slot];
  p.M = Ms[slot];
  p.dt = sg.dt[slot];
  const double* __restrict__ kx2 = sg.kx2 + (size_t)slot * p.Nc;
  const double* __restrict__ ky2 = sg.ky2 + (size_t)slot * p.Ny;
  E1 += (size_t)slot * tstride;
  Q1 += (size_t)slot * tstride;
  if (Q2) Q2 += (size_t)slot * tstride;
  if (E1f) {
    E1f += (size_t)slot * tstride;
    Q1f += (size_t)slot * tstride;
    if (Q2f) Q2f += (size_t)slot * tstride;
  }
  const size_t n = (size_t)p.Nc * p.hy;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i / p.hy);
    const int j = (int)(i - (size_t)c * p.hy);
    const TabEntry t = tab_entry(-(kx2[c] + ky2[j]), p, Q2 != nullptr);
    E1[i] = T(t.e1);
    Q1[i] = T(t.q1);
    if (Q2) Q2[i] = T(t.q2);
    if (E1f) {
      E1f[i] = float(t.em1);
      Q1f[i] = float(t.q1);
      if (Q2f) Q2f[i] = float(t.q2);
    }
  }
}

// External-potential source of slots [slot0, slot0 + gridDim.y):
// C = P lap V_hat / (Nx Ny) over the full [Ny][Nc] half spectrum (V_hat = rfft2
// of the slot's potential; logpfc_extv/simulation.py:557-577 adds
// P lap V_hat to the update, without Gamma).
static __global__ void __launch_bounds__(256) k_ext_table(V2<double>* __restrict__ C,
                                                          const V2<double>* __restrict__ Vhat, long long cstride,
                                                          SlotGrid sg, TabParams p, const double* __restrict__ eps,
                                                          const double* __restrict__ Ms, int slot0,
                                                          const int* __restrict__ list = nullptr,
                                                          const int* __restrict__ rebuild = nullptr) {
  const int slot = list ? list[blockIdx.y] : slot0 + (int)blockIdx.y;  // as in k_tables
  if (rebuild && !rebuild[slot]) return;
  p.epsilon = eps[slot];
  p.M = Ms[slot];
  p.dt = sg.dt[slot];
  const double* __restrict__ kx2 = sg.kx2 + (size_t)slot * p.Nc;
  const double* __restrict__ ky2 = sg.ky2 + (size_t)slot * p.Ny;
  C += (size_t)slot * cstride;
  Vhat += (size_t)slot * cstride;
  const int n = p.Nc * p.Ny;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int ky = i / p.Nc;
    const int kx = i - ky * p.Nc;
    const double lap = -(kx2[kx] + ky2[ky]);
    const double w = etd_coef(lap, p).P * lap * p.inv_n;
    const V2<double> v = Vhat[i];
    C[i] = V2<double>{w * v.x, w * v.y};
  }
}

// ---------------------------------------------------------------------------
// FP64 reductions: free energy per site, mean, min, max
// ---------------------------------------------------------------------------
struct FeParams {
  int model;
  double beta, q0, epsilon, g;
  const double* kx2;  // [slot][Nc]  kx^2 (slot as indexed by Slots)
  const double* ky2;  // [slot][Ny]  ky^2
  const double* eps;  // per-slot epsilon (indexed like Slots), or null: use `epsilon`
};

// per-block partials: [0] sum n, [1] sum f_local, [2] sum w |A|^2 Lf  (each DD),
// plus min and max of n.
struct Partial {
  DD s[3];
  double mn, mx;
};

__device__ __forceinline__ double local_fe(double f, const FeParams& p) {
  double poly = (p.epsilon / 2.0) * (f * f) + (p.g / 3.0) * (f * f * f) + 0.25 * (f * f * f * f);
  if (p.model == MODEL_STD) return poly;
  return f * log(f) + poly;
}

// `red` is LDS scratch of blockDim.x entries (blockDim.x a power of two).
__device__ __forceinline__ void block_reduce_partial(Partial& mine, Partial* red, Partial* out) {
  // fixed-order tree over the block: deterministic for a given launch shape
  red[threadIdx.x] = mine;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      Partial o = red[threadIdx.x + s];
      Partial m = red[threadIdx.x];
      for (int i = 0; i < 3; ++i) m.s[i] = dd_add(m.s[i], o.s[i]);
      m.mn = fmin(m.mn, o.mn);
      m.mx = fmax(m.mx, o.mx);
      red[threadIdx.x] = m;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// rows of S~ (storage Tin) -> n in FP64 -> sums of n and of the local FE
// density, min and max; MINMAX: min and max only (the stabilised split's
// range check), from the same FP64 row transforms
template <typename Tin, bool MINMAX = false>
__global__ void __launch_bounds__(256) k_fe_row(const V2<Tin>* __restrict__ S, RowGeom g, FeParams p,
                                                const V2<double>* __restrict__ twN, Partial* __restrict__ part,
                                                Slots sl, int pstride) {
  using V = V2<double>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* b0 = reinterpret_cast<V*>(smem_raw);
  V* b1 = b0 + g.nl * g.ld;
  // XCD-aware order over (instance, row block): the blocks sharing a
  // line-tiled row group's 128-byte lines meet in one XCD's L2.  A block's
  // partial sums still go to its row block's index, so the fixed reduction
  // order (and the result) does not depend on which block ran it.
  const int ntile = gridDim.x;
  const int lin = xcd_remap(blockIdx.x + blockIdx.y * ntile, ntile * gridDim.y);
  const int yi = lin / ntile;
  const int tile = lin - yi * ntile;
  const int slot = slot_of(sl, yi);
  S += (size_t)slot * sl.sstride;
  if (p.eps) p.epsilon = p.eps[slot];
  part += (size_t)yi * pstride;
  const int row0 = tile * g.nl;
  const int nrows = min(g.nl, g.Ny - row0);
  row_load_pack<double, Tin>(S, b1, b0, g, row0, nrows, twN);
  V* z = fft_lines<double, true>(b0, b1, g.plM, nrows, g.ld, twN, 2);
  Partial mine;
  for (int i = 0; i < 3; ++i) mine.s[i] = DD{0.0, 0.0};
  mine.mn = INFINITY;
  mine.mx = -INFINITY;
  for (int e = threadIdx.x; e < nrows * g.M; e += blockDim.x) {
    const int line = e / g.M;
    const int m = e - line * g.M;
    V v = z[line * g.ld + m];
    if constexpr (!MINMAX) {
      mine.s[0] = dd_add(mine.s[0], v.x);
      mine.s[0] = dd_add(mine.s[0], v.y);
      mine.s[1] = dd_add(mine.s[1], local_fe(v.x, p));
      mine.s[1] = dd_add(mine.s[1], local_fe(v.y, p));
    }
    mine.mn = fmin(mine.mn, fmin(v.x, v.y));
    mine.mx = fmax(mine.mx, fmax(v.x, v.y));
  }
  __syncthreads();
  block_reduce_partial(mine, reinterpret_cast<Partial*>(smem_raw), part + tile);
}

// columns of S~ -> A = n_hat/Nx in FP64 -> sum over the half spectrum of
// w(kx) |A|^2 (beta/2)(q0^2 - k^2)^2, w = 1 on kx = 0 and kx = Nx/2, else 2.
template <typename Tin>
__global__ void __launch_bounds__(256) k_fe_col(const V2<Tin>* __restrict__ S, ColGeom g, FeParams p, int Nx,
                                                const V2<double>* __restrict__ twY, Partial* __restrict__ part,
                                                Slots sl, int pstride) {
  using V = V2<double>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* b0 = reinterpret_cast<V*>(smem_raw);
  V* b1 = b0 + g.nc * g.ld;
  // XCD-aware order, partials at the column block's index (as in k_fe_row)
  const int ntile = gridDim.x;
  const int lin = xcd_remap(blockIdx.x + blockIdx.y * ntile, ntile * gridDim.y);
  const int yi = lin / ntile;
  const int tile = lin - yi * ntile;
  const int slot = slot_of(sl, yi);
  S += (size_t)slot * sl.sstride;
  const double* __restrict__ kx2 = p.kx2 + (size_t)slot * g.Nc;
  const double* __restrict__ ky2 = p.ky2 + (size_t)slot * g.Ny;
  part += (size_t)yi * pstride;
  const int c0 = tile * g.nc;
  const int ncols = min(g.nc, g.Nc - c0);
  for (int e = threadIdx.x; e < ncols * g.Ny; e += blockDim.x) {
    const int y = e / ncols;
    const int line = e - y * ncols;
    V2<Tin> v = S[g.lay.at(y, c0 + line)];
    b0[line * g.ld + y] = V{double(v.x), double(v.y)};
  }
  __syncthreads();
  V* A = fft_lines<double, false>(b0, b1, g.plY, ncols, g.ld, twY, 1);
  Partial mine;
  for (int i = 0; i < 3; ++i) mine.s[i] = DD{0.0, 0.0};
  mine.mn = INFINITY;
  mine.mx = -INFINITY;
  for (int e = threadIdx.x; e < ncols * g.Ny; e += blockDim.x) {
    const int line = e / g.Ny;
    const int ky = e - line * g.Ny;
    const int kx = c0 + line;
    V a = A[line * g.ld + ky];
    const double lap = -(kx2[kx] + ky2[ky]);
    const double t = p.q0 * p.q0 + lap;
    const double lf = (p.beta / 2.0) * (t * t);
    const double w = (kx == 0 || 2 * kx == Nx) ? 1.0 : 2.0;
    mine.s[2] = dd_add(mine.s[2], w * lf * (a.x * a.x + a.y * a.y));
  }
  __syncthreads();
  block_reduce_partial(mine, reinterpret_cast<Partial*>(smem_raw), part + tile);
}

// one block per active instance (blockIdx.x indexes the active list):
// fixed-order sum of its row and column partials into out[4*i ..], i the
// position in the active list:
// free energy per site, mean, min, max.  pin_out[slot] receives the mean.
static __global__ void __launch_bounds__(256) k_fe_final(const Partial* __restrict__ part, int nrow, int ncol, int Nx,
                                                  int Ny, double* __restrict__ out, double* __restrict__ pin_out,
                                                  Slots sl, int pstride) {
  part += (size_t)blockIdx.x * pstride;
  out += 4 * (size_t)blockIdx.x;
  if (pin_out) pin_out += slot_of(sl, blockIdx.x);
  Partial mine;
  for (int i = 0; i < 3; ++i) mine.s[i] = DD{0.0, 0.0};
  mine.mn = INFINITY;
  mine.mx = -INFINITY;
  for (int b = threadIdx.x; b < nrow + ncol; b += blockDim.x) {
    Partial o = part[b];
    for (int i = 0; i < 3; ++i) mine.s[i] = dd_add(mine.s[i], o.s[i]);
    mine.mn = fmin(mine.mn, o.mn);
    mine.mx = fmax(mine.mx, o.mx);
  }
  __shared__ Partial red[256];
  __shared__ Partial res;
  block_reduce_partial(mine, red, &res);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double N = double(Nx) * double(Ny);
    const double grad = (res.s[2].hi + res.s[2].lo) / (double(Ny) * double(Ny));
    const double loc = (res.s[1].hi + res.s[1].lo) / N;
    const double mean = (res.s[0].hi + res.s[0].lo) / N;
    out[0] = grad + loc;
    out[1] = mean;
    out[2] = res.mn;
    out[3] = res.mx;
    if (pin_out) *pin_out = mean;
  }
}

// The stabilised split's range check on the device, once per step before it
// (logpfc_extv/simulation.py:375-405 with OptimizeM :332-373): from the
// field's min and max, a slot that has no fit yet, or whose minimum fell
// below its fitted floor or whose maximum rose above its fitted ceiling by
// more than the relative tolerances, is refitted: lo = min if min > 0 else
// minPhi, hi = max, M = (N(hi) - N(lo)) / (hi - lo) with N(phi) = g phi^2 +
// phi^3 + ln phi, written to the slot's shift and flagged for a table
// rebuild: k_shift_update, one block per active slot, over row partials of
// the field's min and max (from k_fe_row<MINMAX>, as pfc_reduce reports them,
// or from k_row_range, the step's own inverse row transforms).
struct ShiftState {
  double lo, hi;  // fitted window
  int fitted;
};

struct ShiftParams {
  double g, minPhi, floor_tol, ceil_tol;
};

// min and max over a block of at most 256 threads (any count: the fast row
// shapes' blocks hold lines x lanes threads); every thread passes its own and
// receives the block's
__device__ __forceinline__ void block_minmax(double& mn, double& mx) {
  __shared__ double rmn[256], rmx[256];
  __syncthreads();  // earlier readers of the arrays are done
  for (int i = threadIdx.x; i < 256; i += blockDim.x) {
    rmn[i] = INFINITY;
    rmx[i] = -INFINITY;
  }
  __syncthreads();
  rmn[threadIdx.x] = mn;
  rmx[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    for (int i = threadIdx.x; i < s; i += blockDim.x) {
      rmn[i] = fmin(rmn[i], rmn[i + s]);
      rmx[i] = fmax(rmx[i], rmx[i + s]);
    }
    __syncthreads();
  }
  mn = rmn[0];
  mx = rmx[0];
}

// min and max over a block of whole wavefronts (at most four): shuffles
// within each wavefront, then one LDS exchange; the result is thread 0's
__device__ __forceinline__ void block_minmax_waves(double& mn, double& mx) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mn = fmin(mn, __shfl_xor(mn, o, 64));
    mx = fmax(mx, __shfl_xor(mx, o, 64));
  }
  __shared__ double wmn[4], wmx[4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    wmn[w] = mn;
    wmx[w] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
      mn = fmin(mn, wmn[i]);
      mx = fmax(mx, wmx[i]);
    }
}

// Whether a field range [mn, mx] calls for a refit of the fit w: unfitted,
// or the range left the fitted window by more than the tolerances.  Each
// test is monotone in mn (mx), so the range of a whole field breaches exactly
// when the range of one of its tiles does.
__device__ __forceinline__ bool shift_breach(const ShiftState& w, double mn, double mx, double floor_tol,
                                             double ceil_tol) {
  if (!w.fitted) return true;
  return (w.lo - mn) / w.lo > floor_tol || (mx - w.hi) / w.hi > ceil_tol;
}

// The shift fitted to [mn, mx]: the chord slope of the nonlinearity's
// non-log part plus log between lo = max(mn, minPhi) and mx.
__device__ __forceinline__ double shift_fit(double mn, double mx, double g, double minPhi, double& lo) {
  lo = mn > 0.0 ? mn : minPhi;
  const double n_hi = g * mx * mx + mx * mx * mx + log(mx);
  const double n_lo = g * lo * lo + lo * lo * lo + log(lo);
  return (n_hi - n_lo) / (mx - lo);
}

// the refit decision for one slot from its field's min and max (one thread)
__device__ __forceinline__ void shift_refit(double mn, double mx, int slot, const ShiftParams& sp,
                                            ShiftState* __restrict__ st, double* __restrict__ Ms,
                                            int* __restrict__ rebuild) {
  const bool refit = shift_breach(st[slot], mn, mx, sp.floor_tol, sp.ceil_tol);
  if (refit) {
    double lo;
    Ms[slot] = shift_fit(mn, mx, sp.g, sp.minPhi, lo);
    st[slot] = ShiftState{lo, mx, 1};
  }
  rebuild[slot] = refit ? 1 : 0;
}

// The row kernel's share of the in-step decision: every row tile tests its
// own range against the slot's fit and raises brk[2 slot + par]; tile 0
// clears the other word, which the next step (the other parity) raises.
struct RangeCheck {
  const ShiftState* st;  // null: no check (range pairs only)
  int* brk;
  int par;
  double floor_tol, ceil_tol;
};

// per-tile min and max of the field (k_row_range)
struct RangePart {
  double mn, mx;
};

// min and max over a block of whole wavefronts (at most four), returned to
// every thread
__device__ __forceinline__ void block_minmax_all(double& mn, double& mx) {
  block_minmax_waves(mn, mx);
  __shared__ double bmn, bmx;
  if (threadIdx.x == 0) {
    bmn = mn;
    bmx = mx;
  }
  __syncthreads();
  mn = bmn;
  mx = bmx;
}

// PartT: Partial (k_fe_row<MINMAX>) or RangePart (k_row_range), `nrow` per
// instance, `pstride` apart
template <class PartT>
__global__ void __launch_bounds__(256) k_shift_update(const PartT* __restrict__ part, int nrow, Slots sl, int pstride,
                                                      ShiftParams sp, ShiftState* __restrict__ st,
                                                      double* __restrict__ Ms, int* __restrict__ rebuild) {
  part += (size_t)blockIdx.x * pstride;
  double mn = INFINITY, mx = -INFINITY;
#pragma unroll 8
  for (int b = threadIdx.x; b < nrow; b += blockDim.x) {  // up to 8 loads per lane in flight at once
    mn = fmin(mn, part[b].mn);
    mx = fmax(mx, part[b].mx);
  }
  block_minmax_waves(mn, mx);  // 256 threads: whole wavefronts
  if (threadIdx.x == 0) shift_refit(mn, mx, slot_of(sl, blockIdx.x), sp, st, Ms, rebuild);
}

}  // namespace pfc
