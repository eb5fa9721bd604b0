// Register-resident Stockham FFT kernels for compile-time line lengths.
//
// A line of L complex points is owned by TPL = L/P lanes; lane l holds the P
// points l + TPL*q (q = 0..P-1) in registers ("strided layout").  A radix-R
// Stockham pass gives each lane the butterflies j = l + TPL*i (i < P/R); its
// inputs j + r*L/R are exactly registers i + r*P/R of the strided layout, and
// in the last pass (Ns = L/R) its outputs j + r*Ns land on the same
// registers.  So the first pass reads registers, the last pass writes
// registers, and only intermediate passes exchange through LDS.  The forward
// transform therefore ends in the layout the inverse starts from, which is
// what lets the column kernel apply the ETD combine in registers between the
// forward and the inverse column FFTs.
#pragma once
#include "pfc_kernels.hpp"

namespace pfc {

// ---------------------------------------------------------------------------
// radix-16 = 4 x 4 with the internal w16 twiddles
// ---------------------------------------------------------------------------
template <typename T, bool INV> struct Dft<T, 16, INV> {
  using V = V2<T>;
  __device__ __forceinline__ static void run(V* v, const V*, int) {
    // X[k1 + 4 k2] = sum_n2 w4^(n2 k2) [ w16^(n2 k1) sum_n1 x[4 n1 + n2] w4^(n1 k1) ]
    const T c1 = T(0.92387953251128675612818318939678828682);  // cos(pi/8)
    const T s1 = T(0.38268343236508977172845998403039886676);  // sin(pi/8)
    const T h = T(0.70710678118654752440084436210484903928);
    V a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      V t[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
      Dft<T, 4, INV>::run(t, nullptr, 0);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = t[k1];
    }
    // twiddle w16^(n2 k1): exponents 0..9
    auto tw = [&](V x, int e) -> V {
      // w16^e = exp(-+ i pi e / 8)
      T c, s;
      switch (e) {
        case 0: return x;
        case 1: c = c1; s = s1; break;
        case 2: c = h; s = h; break;
        case 3: c = s1; s = c1; break;
        case 4: c = T(0); s = T(1); break;
        case 6: c = -h; s = h; break;
        case 9: c = -c1; s = -s1; break;
        default: c = T(0); s = T(0); break;
      }
      // forward: multiply by (c - i s); inverse: (c + i s)
      if (INV) return V{x.x * c - x.y * s, x.x * s + x.y * c};
      return V{x.x * c + x.y * s, x.y * c - x.x * s};
    };
#pragma unroll
    for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
      for (int k1 = 1; k1 < 4; ++k1) a[n2][k1] = tw(a[n2][k1], n2 * k1);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      V t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
      Dft<T, 4, INV>::run(t, nullptr, 0);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  }
};

// ---------------------------------------------------------------------------
// compile-time schedules
// ---------------------------------------------------------------------------
template <int L> struct FastSched;  // P points per lane and the radix sequence
#define PFC_SCHED(LEN, PP, ...)                                   \
  template <> struct FastSched<LEN> {                             \
    static constexpr int P = PP;                                  \
    static constexpr int R[] = {__VA_ARGS__};                     \
    static constexpr int NP = sizeof(R) / sizeof(int);            \
  };
PFC_SCHED(32, 8, 8, 4)
PFC_SCHED(48, 12, 4, 4, 3)
PFC_SCHED(64, 8, 8, 8)
PFC_SCHED(96, 12, 4, 4, 2, 3)
PFC_SCHED(128, 8, 8, 8, 2)
PFC_SCHED(192, 12, 4, 4, 4, 3)
PFC_SCHED(256, 8, 8, 8, 4)
PFC_SCHED(384, 12, 4, 4, 4, 2, 3)
PFC_SCHED(512, 8, 8, 8, 8)
PFC_SCHED(768, 12, 4, 4, 4, 4, 3)
PFC_SCHED(1024, 16, 16, 8, 8)
PFC_SCHED(2048, 16, 16, 16, 8)
#undef PFC_SCHED

// product of the first p radices
template <int L> __host__ __device__ constexpr int sched_ns(int p) {
  int ns = 1;
  for (int i = 0; i < p; ++i) ns *= FastSched<L>::R[i];
  return ns;
}

// One pass: butterflies from registers, results back to registers (in
// butterfly-local order: register i + r*NB holds output r of butterfly i).
template <typename T, int L, int PI, bool INV>
__device__ __forceinline__ void reg_pass(V2<T>* v, int lane, const V2<T>* __restrict__ tw, int tws) {
  using V = V2<T>;
  constexpr int P = FastSched<L>::P;
  constexpr int R = FastSched<L>::R[PI];
  constexpr int NS = sched_ns<L>(PI);
  constexpr int TPL = L / P;
  constexpr int NB = P / R;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    V x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = v[i + r * NB];
    if (NS > 1) {
      const int j = lane + TPL * i;
      const int k = j % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        V t = tw[(k * r * (L / (NS * R))) * tws];
        if (INV) t.y = -t.y;
        x[r] = cmul(x[r], t);
      }
    }
    Dft<T, R, INV>::run(x, tw, (L / R) * tws);
#pragma unroll
    for (int r = 0; r < R; ++r) v[i + r * NB] = x[r];
  }
}

// Write pass PI's outputs to their Stockham positions in `line` (LDS).
template <typename T, int L, int PI>
__device__ __forceinline__ void reg_scatter(const V2<T>* v, V2<T>* line, int lane) {
  constexpr int P = FastSched<L>::P;
  constexpr int R = FastSched<L>::R[PI];
  constexpr int NS = sched_ns<L>(PI);
  constexpr int TPL = L / P;
  constexpr int NB = P / R;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int j = lane + TPL * i;
    const int k = j % NS;
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) line[base + r * NS] = v[i + r * NB];
  }
}

template <typename T, int L>
__device__ __forceinline__ void reg_gather(V2<T>* v, const V2<T>* line, int lane) {
  constexpr int P = FastSched<L>::P;
  constexpr int TPL = L / P;
#pragma unroll
  for (int q = 0; q < P; ++q) v[q] = line[lane + TPL * q];
}

// Full FFT of the lane's strided registers, exchanging through `line`
// (L elements of LDS private to this line).  All threads of the block must
// call it (it contains block barriers).
template <typename T, int L, bool INV, int PI = 0>
__device__ __forceinline__ void reg_fft(V2<T>* v, V2<T>* line, int lane, const V2<T>* __restrict__ tw, int tws) {
  reg_pass<T, L, PI, INV>(v, lane, tw, tws);
  if constexpr (PI + 1 < FastSched<L>::NP) {
    __syncthreads();  // earlier readers of `line` are done
    reg_scatter<T, L, PI>(v, line, lane);
    __syncthreads();
    reg_gather<T, L>(v, line, lane);
    reg_fft<T, L, INV, PI + 1>(v, line, lane, tw, tws);
  }
}

// ---------------------------------------------------------------------------
// fast row kernel: S~ -> C2R -> N(n) -> R2C -> T, one row per TPL lanes
// ---------------------------------------------------------------------------
template <typename T, int M>
__global__ void __launch_bounds__(256) k_row_fast(const V2<T>* __restrict__ S, V2<T>* __restrict__ Tout, int Ny,
                                                  Nonlin q, const V2<T>* __restrict__ twN,
                                                  int* __restrict__ status) {
  using V = V2<T>;
  constexpr int P = FastSched<M>::P;
  constexpr int TPL = M / P;
  constexpr int Nc = M + 1;
  constexpr int LD = M + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int W = blockDim.x / TPL;
  const int wl = threadIdx.x / TPL;
  const int lane = threadIdx.x - wl * TPL;
  const int row = blockIdx.x * W + wl;
  const bool live = row < Ny;
  V* line = reinterpret_cast<V*>(smem_raw) + wl * LD;

  if (live) {
    const V* src = S + (size_t)row * Nc;
#pragma unroll
    for (int qq = 0; qq < P; ++qq) line[lane + TPL * qq] = src[lane + TPL * qq];
    if (lane == 0) line[M] = src[M];
  }
  __syncthreads();
  V v[P];
#pragma unroll
  for (int qq = 0; qq < P; ++qq) {
    const int k = lane + TPL * qq;
    V xk = line[k];
    V xm = line[M - k];
    if (k == 0) {
      xk.y = T(0);
      xm.y = T(0);
    }
    V s = V{xk.x + xm.x, xk.y - xm.y};
    V d = V{xk.x - xm.x, xk.y + xm.y};
    V w = twN[k];
    V wd = cmul(cconj(w), d);
    v[qq] = V{s.x - wd.y, s.y + wd.x};
  }
  reg_fft<T, M, true>(v, line, lane, twN, 2);
  bool bad = false;
#pragma unroll
  for (int qq = 0; qq < P; ++qq) {
    bad |= !(isfinite(v[qq].x) && isfinite(v[qq].y));
    v[qq] = V{nonlin_mu<T>(v[qq].x, q), nonlin_mu<T>(v[qq].y, q)};
  }
  if (bad && live) atomicOr(status, 1);
  reg_fft<T, M, false>(v, line, lane, twN, 2);
  __syncthreads();
#pragma unroll
  for (int qq = 0; qq < P; ++qq) line[lane + TPL * qq] = v[qq];
  __syncthreads();
  if (live) {
    V* dst = Tout + (size_t)row * Nc;
#pragma unroll
    for (int qq = 0; qq < P; ++qq) {
      const int k = lane + TPL * qq;
      dst[k] = row_unpack<T>(line, k, M, twN);
    }
    if (lane == 0) dst[M] = row_unpack<T>(line, M, M, twN);
  }
}

// ---------------------------------------------------------------------------
// fast column kernel: forward column FFTs of S~ and T, ETD combine in
// registers, inverse column FFT, S~ written in place.  C columns per block.
// ---------------------------------------------------------------------------
template <typename T, int NY, int MODE>
__global__ void __launch_bounds__(256) k_col_fast(V2<T>* __restrict__ S, const V2<T>* __restrict__ Tin,
                                                  V2<T>* __restrict__ D, int Nc, ColTables tb,
                                                  const V2<T>* __restrict__ twY) {
  using V = V2<T>;
  constexpr int P = FastSched<NY>::P;
  constexpr int TPL = NY / P;
  constexpr int LD = NY;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds = reinterpret_cast<V*>(smem_raw);
  const int C = blockDim.x / TPL;
  const int wl = threadIdx.x / TPL;
  const int lane = threadIdx.x - wl * TPL;
  const int ntile = (Nc + C - 1) / C;
  const int c0 = xcd_remap(blockIdx.x, ntile) * C;
  const int ncols = min(C, Nc - c0);
  const int col = c0 + wl;
  const bool live = wl < ncols;
  V* line = lds + wl * LD;
  const T* E1 = reinterpret_cast<const T*>(tb.E1);
  const T* Q1 = reinterpret_cast<const T*>(tb.Q1);
  const T* Q2 = reinterpret_cast<const T*>(tb.Q2);

  // coalesced tile load: consecutive threads take consecutive columns of a row
  auto tile_load = [&](const V* __restrict__ src) {
    for (int e = threadIdx.x; e < ncols * NY; e += blockDim.x) {
      const int y = e / ncols;
      const int c = e - y * ncols;
      lds[c * LD + y] = src[(size_t)y * Nc + c0 + c];
    }
  };

  V a[P], b[P];
  if (MODE != COL_RK2_STAGE2) {
    tile_load(S);
    __syncthreads();
    reg_gather<T, NY>(a, line, lane);
    reg_fft<T, NY, false>(a, line, lane, twY, 1);
    __syncthreads();
  }
  tile_load(Tin);
  __syncthreads();
  reg_gather<T, NY>(b, line, lane);
  reg_fft<T, NY, false>(b, line, lane, twY, 1);

  const double pin = tb.pin ? *tb.pin : 0.0;
  if (live) {
    const size_t tbase = (size_t)col * NY;
#pragma unroll
    for (int qq = 0; qq < P; ++qq) {
      const int ky = lane + TPL * qq;
      const size_t ti = tbase + ky;
      V out;
      if (MODE == COL_RK2_STAGE2) {
        const T q2 = Q2[ti];
        V d = D[ti];
        out = V{d.x + q2 * b[qq].x, d.y + q2 * b[qq].y};
      } else {
        const T e1 = E1[ti], q1 = Q1[ti];
        out = V{e1 * a[qq].x + q1 * b[qq].x, e1 * a[qq].y + q1 * b[qq].y};
        if (MODE == COL_RK2_STAGE1) {
          const T q2 = Q2[ti];
          D[ti] = V{out.x - q2 * b[qq].x, out.y - q2 * b[qq].y};
        }
      }
      if (tb.pin && col == 0 && ky == 0) out = V{T(pin), T(0)};
      b[qq] = out;
    }
  }
  reg_fft<T, NY, true>(b, line, lane, twY, 1);
  __syncthreads();
  // real projection of the kx = 0 and kx = Nx/2 columns (see k_col_etd)
  const bool realcol = (col == 0 || col == Nc - 1);
#pragma unroll
  for (int qq = 0; qq < P; ++qq) {
    V x = b[qq];
    if (realcol) x.y = T(0);
    line[lane + TPL * qq] = x;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ncols * NY; e += blockDim.x) {
    const int y = e / ncols;
    const int c = e - y * ncols;
    S[(size_t)y * Nc + c0 + c] = lds[c * LD + y];
  }
}

}  // namespace pfc
