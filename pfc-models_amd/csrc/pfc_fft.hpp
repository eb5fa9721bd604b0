// Device building blocks for the fused pseudo-spectral PFC step on gfx950.
//
// Complex lines live in LDS; a mixed-radix Stockham FFT (radices 2,3,4,5,7,8,
// 11,13) ping-pongs them between two LDS buffers, one radix-R butterfly per
// thread per work item.  The Stockham ordering (output index
// (j/Ns)*Ns*R + j%Ns + r*Ns) sorts itself, so no bit reversal pass is needed.
// Twiddles come from a per-length table exp(-2 pi i m / (L*tws)) built on the
// host in FP64 (then rounded to the storage type), so every root is correctly
// rounded instead of being produced by device sincos.
#pragma once
#include <hip/hip_runtime.h>

namespace pfc {

// Complex values are 2-element clang vectors (x = real, y = imaginary):
// copies are single 8/16-byte vector loads and stores of a first-class value.
// Struct types (HIP's double2/float2 included) are copied with memcpy across
// address spaces, which keeps register arrays from being promoted out of
// scratch memory.
template <typename T> struct Vec2 { using type = T __attribute__((ext_vector_type(2))); };
template <typename T> using V2 = typename Vec2<T>::type;

template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return V{a.x + b.x, a.y + b.y}; }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return V{a.x - b.x, a.y - b.y}; }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
  return V{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename V> __device__ __forceinline__ V cconj(V a) { return V{a.x, -a.y}; }
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename V> __device__ __forceinline__ V cmul_mi(V a) {
  return INV ? V{-a.y, a.x} : V{a.y, -a.x};
}
template <typename V, typename S> __device__ __forceinline__ V cscale(V a, S s) { return V{a.x * s, a.y * s}; }

constexpr int kMaxPass = 20;

// Radix schedule of one FFT length (product of r[0..np-1] == L).
struct Radices {
  int L;
  int np;
  int r[kMaxPass];
};

// ---------------------------------------------------------------------------
// small DFTs (forward: exp(-2 pi i / R); INV: conjugate roots)
// ---------------------------------------------------------------------------
template <typename T, int R, bool INV> struct Dft;

template <typename T, bool INV> struct Dft<T, 2, INV> {
  using V = V2<T>;
  __device__ __forceinline__ static void run(V* v, const V*, int) {
    V a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename T, bool INV> struct Dft<T, 4, INV> {
  using V = V2<T>;
  __device__ __forceinline__ static void run(V* v, const V*, int) {
    V t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    V t2 = cadd(v[1], v[3]), t3 = cmul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <typename T, bool INV> struct Dft<T, 8, INV> {
  using V = V2<T>;
  __device__ __forceinline__ static void run(V* v, const V*, int) {
    const T h = T(0.70710678118654752440084436210484903928);  // 1/sqrt(2)
    V e[4] = {v[0], v[2], v[4], v[6]};
    V o[4] = {v[1], v[3], v[5], v[7]};
    Dft<T, 4, INV>::run(e, nullptr, 0);
    Dft<T, 4, INV>::run(o, nullptr, 0);
    // w8^1 = (1 -+ i)/sqrt2, w8^2 = -+i, w8^3 = (-1 -+ i)/sqrt2
    V o1 = INV ? V{(o[1].x - o[1].y) * h, (o[1].x + o[1].y) * h}
               : V{(o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h};
    V o2 = cmul_mi<INV>(o[2]);
    V o3 = INV ? V{-(o[3].x + o[3].y) * h, (o[3].x - o[3].y) * h}
               : V{(o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

template <typename T, bool INV> struct Dft<T, 3, INV> {
  using V = V2<T>;
  __device__ __forceinline__ static void run(V* v, const V*, int) {
    const T s3 = T(0.86602540378443864676372317075293618347);  // sqrt(3)/2
    V s = cadd(v[1], v[2]);
    V d = csub(v[1], v[2]);
    V a = v[0];
    V m = V{a.x - T(0.5) * s.x, a.y - T(0.5) * s.y};
    // forward: X1 = m - i s3 d, X2 = m + i s3 d
    V id = INV ? V{-d.y * s3, d.x * s3} : V{d.y * s3, -d.x * s3};
    v[0] = cadd(a, s);
    v[1] = cadd(m, id);
    v[2] = csub(m, id);
  }
};

// Odd prime radices: direct O(R^2) DFT with roots read from the twiddle table
// (w_R^m = table[m * step]).  Used for 5, 7, 11, 13.
template <typename T, int R, bool INV> struct Dft {
  using V = V2<T>;
  __device__ __forceinline__ static void run(V* v, const V* tw, int step) {
    V out[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      V acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        V w = tw[((r * s) % R) * step];
        if (INV) w.y = -w.y;
        acc = cadd(acc, cmul(v[r], w));
      }
      out[s] = acc;
    }
#pragma unroll
    for (int s = 0; s < R; ++s) v[s] = out[s];
  }
};

// ---------------------------------------------------------------------------
// one Stockham pass over `nl` lines of length L (line stride ld) in LDS
// ---------------------------------------------------------------------------
template <typename T, int R, bool INV>
__device__ __forceinline__ void sk_pass(const V2<T>* __restrict__ src, V2<T>* __restrict__ dst, int L, int Ns,
                                        int nl, int ld, const V2<T>* __restrict__ tw, int tws) {
  using V = V2<T>;
  const int nb = L / R;
  const int total = nb * nl;
  const int twstep = (L / (Ns * R)) * tws;  // table stride of one (k*r) unit
  const int rootstep = nb * tws;            // table stride of w_R
  for (int w = threadIdx.x; w < total; w += blockDim.x) {
    const int line = w / nb;
    const int j = w - line * nb;
    const V* s = src + line * ld;
    V* d = dst + line * ld;
    const int k = j % Ns;
    V v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[j + r * nb];
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
    for (int r = 0; r < R; ++r) d[base + r * Ns] = v[r];
  }
}

// Full FFT of `nl` lines held in buf0; returns the buffer holding the result.
// Callers must __syncthreads() between writing buf0 and calling this.
template <typename T, bool INV>
__device__ V2<T>* fft_lines(V2<T>* buf0, V2<T>* buf1, const Radices& pl, int nl, int ld,
                            const V2<T>* __restrict__ tw, int tws) {
  V2<T>* src = buf0;
  V2<T>* dst = buf1;
  int Ns = 1;
  for (int p = 0; p < pl.np; ++p) {
    const int R = pl.r[p];
    switch (R) {
      case 2: sk_pass<T, 2, INV>(src, dst, pl.L, Ns, nl, ld, tw, tws); break;
      case 3: sk_pass<T, 3, INV>(src, dst, pl.L, Ns, nl, ld, tw, tws); break;
      case 4: sk_pass<T, 4, INV>(src, dst, pl.L, Ns, nl, ld, tw, tws); break;
      case 5: sk_pass<T, 5, INV>(src, dst, pl.L, Ns, nl, ld, tw, tws); break;
      case 7: sk_pass<T, 7, INV>(src, dst, pl.L, Ns, nl, ld, tw, tws); break;
      case 8: sk_pass<T, 8, INV>(src, dst, pl.L, Ns, nl, ld, tw, tws); break;
      case 11: sk_pass<T, 11, INV>(src, dst, pl.L, Ns, nl, ld, tw, tws); break;
      default: sk_pass<T, 13, INV>(src, dst, pl.L, Ns, nl, ld, tw, tws); break;
    }
    __syncthreads();
    Ns *= R;
    V2<T>* t = src;
    src = dst;
    dst = t;
  }
  return src;
}

// ---------------------------------------------------------------------------
// compensated (double-double) accumulation for the FP64 reductions
// ---------------------------------------------------------------------------
struct DD {
  double hi, lo;
};

__device__ __forceinline__ DD dd_add(DD a, double b) {
  // Knuth TwoSum of a.hi + b, low parts carried
  double s = a.hi + b;
  double bb = s - a.hi;
  double err = (a.hi - (s - bb)) + (b - bb);
  double lo = a.lo + err;
  double hi = s + lo;
  return DD{hi, lo - (hi - s)};
}

__device__ __forceinline__ DD dd_add(DD a, DD b) {
  double s = a.hi + b.hi;
  double bb = s - a.hi;
  double err = (a.hi - (s - bb)) + (b.hi - bb);
  double lo = a.lo + b.lo + err;
  double hi = s + lo;
  return DD{hi, lo - (hi - s)};
}

}  // namespace pfc
