// Register-resident Stockham FFT kernels for compile-time schedules.
//
// A schedule Sch<L, P, R...> transforms lines of L complex points with
// TPL = L/P lanes per line; lane l holds the P points l + TPL*q (q < P) in
// registers ("strided layout").  A radix-R Stockham pass gives each lane the
// butterflies j = l + TPL*i (i < P/R); their inputs j + r*L/R are registers
// i + r*P/R of the strided layout, and in the last pass (Ns = L/R) the outputs
// j + r*Ns land on the same registers.  So the first pass reads registers, the
// last pass writes registers, and only intermediate passes exchange through
// LDS.  The forward transform therefore ends in the layout the inverse starts
// from, which lets the column kernel apply the ETD combine in registers
// between its forward and inverse column FFTs.
//
// When a line fits in one wavefront (TPL <= 64, TPL | 64) its LDS exchanges
// need only wave-level ordering; otherwise they use block barriers.
#pragma once
#include <type_traits>

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
    auto tw = [&](V x, int e) -> V {
      // w16^e = exp(-+ i pi e / 8), e in {1, 2, 3, 4, 6, 9}
      T c, s;
      switch (e) {
        case 1: c = c1; s = s1; break;
        case 2: c = h; s = h; break;
        case 3: c = s1; s = c1; break;
        case 4: c = T(0); s = T(1); break;
        case 6: c = -h; s = h; break;
        default: c = -c1; s = -s1; break;  // 9
      }
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
// schedules
// ---------------------------------------------------------------------------
template <int L_, int P_, int... Rs> struct Sch {
  static constexpr int L = L_;
  static constexpr int P = P_;
  static constexpr int TPL = L_ / P_;
  static constexpr int NP = sizeof...(Rs);
  static constexpr int R[] = {Rs...};
  static constexpr bool WAVE = (TPL <= 64) && (64 % TPL == 0);
  static constexpr int ns(int p) {
    int n = 1;
    for (int i = 0; i < p; ++i) n *= R[i];
    return n;
  }
  static constexpr bool valid() {
    int prod = 1;
    for (int i = 0; i < NP; ++i) {
      if (P % R[i] != 0) return false;
      prod *= R[i];
    }
    return prod == L && L % P == 0;
  }
};

// Ordering of LDS traffic between the lanes of one line: a wave-level fence
// when the line's lanes share one wavefront, else a block barrier.
template <bool BLOCK> __device__ __forceinline__ void line_sync() {
  if constexpr (!BLOCK) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  } else {
    __syncthreads();
  }
}

// Stores of arrays that the next kernel reads on other XCDs (S~ and T cross
// the row/column transpose): non-temporal, so they stream out while the
// kernel runs instead of waiting as dirty L2 lines for the end-of-kernel
// write-back.  (Measured faster per step also for tiles narrower than a
// 128-byte line, whose partial lines are then written separately.)
template <typename V> __device__ __forceinline__ void store_stream(V* p, V v) { __builtin_nontemporal_store(v, p); }
// The store of a row or column kernel's output: non-temporal (NT) or
// ordinary.  Ordinary stores leave the lines in L2, which merges the partial
// 128-byte lines of tiles narrower than a line before they reach memory; the
// autotuner picks the kind per engine from a timed step loop.
template <bool NT, typename V> __device__ __forceinline__ void store_out(V* p, V v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// State accesses (S~, T, the ETDRK2 carry) of a tile.  MEM 0: ordinary loads
// and store_out<NT>.  MEM 1 (the persistent group step, k_flow, where one
// launch's tiles hand their results to each other): device-scope relaxed
// atomic loads and stores (sc1), which read from and write through to the
// device's coherence point, so a tile that waits for another tile's counter
// sees its stores without an L2 write-back or invalidate.  FP64 complex
// values move as two FP64 atomics (an element-wise bit cast of the vector
// miscompiles here), FP32 ones as one 64-bit atomic.
template <int MEM, typename T> __device__ __forceinline__ V2<T> ld_state(const V2<T>* p) {
  if constexpr (MEM == 0) {
    return *p;
  } else if constexpr (sizeof(T) == 8) {
    const T* q = reinterpret_cast<const T*>(p);
    const T a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const T b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return V2<T>{a, b};
  } else {
    const unsigned long long* q = reinterpret_cast<const unsigned long long*>(p);
    return __builtin_bit_cast(V2<T>, __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
}
template <int MEM, bool NT, typename T> __device__ __forceinline__ void st_state(V2<T>* p, V2<T> v) {
  if constexpr (MEM == 0) {
    store_out<NT>(p, v);
  } else if constexpr (sizeof(T) == 8) {
    T* q = reinterpret_cast<T*>(p);
    __hip_atomic_store(q, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), __builtin_bit_cast(unsigned long long, v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Wavefront shuffles for the row kernel's pack and unpack.  Both pair the
// element k = lane + TPL q of a line with M - k, which for a line inside one
// wavefront lives in lane (TPL - lane) mod TPL of the same line, register
// P - 1 - q (lane 0: register P - q, its own): a fixed lane permutation, done
// with ds_bpermute pulls (no LDS store, no barrier) instead of a round trip
// through the line's LDS.  Each pull moves one dword per lane.
constexpr bool kRowShuffle = true;

template <typename T> __device__ __forceinline__ V2<T> pull_lane(V2<T> x, int addr) {
  constexpr int n = sizeof(V2<T>) / 4;
  union {
    V2<T> v;
    int w[n];
  } a, b;
  a.v = x;
#pragma unroll
  for (int i = 0; i < n; ++i) b.w[i] = __builtin_amdgcn_ds_bpermute(addr, a.w[i]);
  return b.v;
}

// m[q] = element M - (lane + TPL q) of the line (q >= 1 or lane > 0); for
// lane 0, q = 0 the caller substitutes element M (or 0) itself.
template <typename T, class S>
__device__ __forceinline__ void mirror_pull(const V2<T> (&v)[S::P], V2<T> (&m)[S::P], int lane) {
  constexpr int P = S::P, TPL = S::TPL;
  const int wl = __lane_id();
  const int src = wl - lane + ((TPL - lane) & (TPL - 1));
  const int addr = src << 2;
#pragma unroll
  for (int qq = 0; qq < P; ++qq) {
    const V2<T> far = pull_lane<T>(v[P - 1 - qq], addr);
    m[qq] = (lane == 0) ? v[(P - qq) % P] : far;
  }
}

// LDS line index with one pad slot per 16 elements: Stockham scatters write
// at strides of R (or R*Ns) elements, which without padding put every lane of
// a wavefront on the same bank.
__host__ __device__ constexpr int lpad(int i) { return i + (i >> 4); }
// padded LDS line lengths of the row kernel (M + 1 points) and column kernel (Ny)
__host__ __device__ constexpr int row_ld(int M) { return lpad(M) + 2; }
__host__ __device__ constexpr int col_ld(int Ny) { return lpad(Ny - 1) + 2; }

// Twiddles of the lane's butterflies in passes 1..NP-1, held in registers:
// pass PI contributes (P / R) * (R - 1) roots w_L^(k r L / (NS R)), k = j % NS.
// The same roots serve the forward and (conjugated) inverse transforms.
template <class S> __host__ __device__ constexpr int tw_offset(int pi) {
  int n = 0;
  for (int p = 1; p < pi; ++p) n += (S::P / S::R[p]) * (S::R[p] - 1);
  return n;
}
template <class S> __host__ __device__ constexpr int tw_count() { return tw_offset<S>(S::NP) > 0 ? tw_offset<S>(S::NP) : 1; }

// w[r - 1] = u^r, r = 1..R-1, by binary splitting (u^r = u^(r/2) u^(r - r/2)):
// at most ceil(log2 R) rounded products separate u^r from the loaded root.
template <int R>
__device__ __forceinline__ void root_powers(V2<double> u, V2<double>* w) {
  w[0] = u;
#pragma unroll
  for (int r = 2; r < R; ++r) w[r - 1] = cmul(w[r / 2 - 1], w[r - r / 2 - 1]);
}

// The lane's roots for passes PI..NP-1 from an FP64 table exp(-2 pi i m /
// (L * tws)): one load per butterfly (its base root w^(k L / (NS R))), the
// other R - 2 powers computed in FP64 and rounded once to T.
template <typename T, class S, int PI = 1>
__device__ __forceinline__ void tw_load(V2<T>* w, int lane, const V2<double>* __restrict__ tw, int tws) {
  if constexpr (PI < S::NP) {
    constexpr int R = S::R[PI];
    constexpr int NS = S::ns(PI);
    constexpr int NB = S::P / R;
    constexpr int OFF = tw_offset<S>(PI);
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int k = (lane + S::TPL * i) % NS;
      V2<double> pw[R - 1];
      root_powers<R>(tw[(k * (S::L / (NS * R))) * tws], pw);
#pragma unroll
      for (int r = 1; r < R; ++r) w[OFF + i * (R - 1) + r - 1] = V2<T>{T(pw[r - 1].x), T(pw[r - 1].y)};
    }
    tw_load<T, S, PI + 1>(w, lane, tw, tws);
  }
}

template <typename T, class S, int PI, bool INV>
__device__ __forceinline__ void reg_pass(V2<T>* v, const V2<T>* w) {
  using V = V2<T>;
  constexpr int R = S::R[PI];
  constexpr int NS = S::ns(PI);
  constexpr int NB = S::P / R;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    V x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = v[i + r * NB];
    if constexpr (NS > 1) {
      constexpr int OFF = tw_offset<S>(PI);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        V t = w[OFF + i * (R - 1) + r - 1];
        if (INV) t.y = -t.y;
        x[r] = cmul(x[r], t);
      }
    }
    Dft<T, R, INV>::run(x, nullptr, 0);
#pragma unroll
    for (int r = 0; r < R; ++r) v[i + r * NB] = x[r];
  }
}

template <typename T, class S, int PI>
__device__ __forceinline__ void reg_scatter(const V2<T>* v, V2<T>* line, int lane) {
  constexpr int R = S::R[PI];
  constexpr int NS = S::ns(PI);
  constexpr int NB = S::P / R;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int j = lane + S::TPL * i;
    const int k = j % NS;
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) line[lpad(base + r * NS)] = v[i + r * NB];
  }
}

template <typename T, class S>
__device__ __forceinline__ void reg_gather(V2<T>* v, const V2<T>* line, int lane) {
#pragma unroll
  for (int q = 0; q < S::P; ++q) v[q] = line[lpad(lane + S::TPL * q)];
}

// Full FFT of the lane's strided registers, exchanging through `line` (S::L
// LDS elements owned by this line); `w` holds the lane's roots (tw_load).
// BLOCK: the line's lanes span several wavefronts.  Every lane of the block
// must call it.
template <typename T, class S, bool INV, bool BLOCK, int PI = 0>
__device__ __forceinline__ void reg_fft(V2<T>* v, V2<T>* line, int lane, const V2<T>* w) {
  reg_pass<T, S, PI, INV>(v, w);
  if constexpr (PI + 1 < S::NP) {
    line_sync<BLOCK>();  // earlier readers of `line` are done
    reg_scatter<T, S, PI>(v, line, lane);
    line_sync<BLOCK>();
    reg_gather<T, S>(v, line, lane);
    reg_fft<T, S, INV, BLOCK, PI + 1>(v, line, lane, w);
  }
}

// ---------------------------------------------------------------------------
// fast row kernel: S~ -> C2R -> N(n) -> R2C -> T; W rows per block
// LDS: W lines of M + 1 (padded).  Roots come from the FP64 table
// exp(-2 pi i m / Nx), m < Nx, into registers.
// ---------------------------------------------------------------------------
// The C2R half of the row pass for one line: loads row `rowc` of S~ (the
// lane's P points and the kx = M point), folds the half spectrum into the
// M-point complex sequence whose inverse FFT interleaves the row's even and
// odd real samples, and runs that FFT.  On return v[q] = (n[2k], n[2k+1]),
// k = lane + TPL q, and wk, wM, wt hold the roots the forward half needs.
// m0: the conserved mean, taken off the kx = 0 coefficient (the row's
// average) so that the transform carries the fluctuation n - m0 only (0: none).
template <typename T, class S, int MEM = 0>
__device__ __forceinline__ void row_c2r(const V2<T>* __restrict__ Sg, const Lay& lay, int rowc, int lane, V2<T>* line,
                                        const V2<double>* __restrict__ twN, V2<T> (&v)[S::P], V2<T> (&wk)[S::P],
                                        V2<T>& wM, V2<T> (&wt)[tw_count<S>()], double m0 = 0.0) {
  using V = V2<T>;
  constexpr int M = S::L, P = S::P, TPL = S::TPL;
  V x[P];
#pragma unroll
  for (int qq = 0; qq < P; ++qq) x[qq] = ld_state<MEM, T>(Sg + lay.at(rowc, lane + TPL * qq));
  const V xM = ld_state<MEM, T>(Sg + lay.at(rowc, M));
  // roots: w_Nx^k for the (un)packing of k = lane + TPL q, as w_Nx^lane times
  // the wave-uniform w_Nx^(TPL q) (scalar loads), in FP64; and the FFT roots
  const V2<double> wl0 = twN[lane];
#pragma unroll
  for (int qq = 0; qq < P; ++qq) {
    const V2<double> c = cmul(wl0, twN[TPL * qq]);
    wk[qq] = V{T(c.x), T(c.y)};
  }
  wM = V{T(twN[M].x), T(twN[M].y)};
  tw_load<T, S>(wt, lane, twN, 2);
  // keep every load above in flight together: no arithmetic is scheduled
  // ahead of the last one
  __builtin_amdgcn_sched_barrier(0);
  if (lane == 0) x[0].x = T(double(x[0].x) - m0);
  V xmr[P];
  if constexpr (kRowShuffle && S::WAVE) {
    mirror_pull<T, S>(x, xmr, lane);
    if (lane == 0) xmr[0] = xM;
  } else {
#pragma unroll
    for (int qq = 0; qq < P; ++qq) line[lpad(lane + TPL * qq)] = x[qq];
    line[lpad(M)] = xM;  // every lane stores the same value: no branch for the load to sink into
    line_sync<!S::WAVE>();
#pragma unroll
    for (int qq = 0; qq < P; ++qq) xmr[qq] = line[lpad(M - (lane + TPL * qq))];
  }
#pragma unroll
  for (int qq = 0; qq < P; ++qq) {
    const int k = lane + TPL * qq;
    V xk = x[qq];
    V xm = xmr[qq];
    if (k == 0) {
      xk.y = T(0);
      xm.y = T(0);
    }
    V s = V{xk.x + xm.x, xk.y - xm.y};
    V d = V{xk.x - xm.x, xk.y + xm.y};
    V wd = cmul(cconj(wk[qq]), d);
    v[qq] = V{s.x - wd.y, s.y + wd.x};
  }
  reg_fft<T, S, true, !S::WAVE>(v, line, lane, wt);
}

// Row tile `tile` (blockDim.x / TPL rows) of active instance `yi`; every
// thread of the block calls it.
// RANGE (a block of whole wavefronts): the min and max of the tile's field,
// the values the nonlinearity sees, are also written to `rtile`: the range
// check of the stabilised split folded into the ETD1 tables.  A compile-time
// choice: the ordinary kernels keep exactly the code the group kernels share.
template <typename T, class S, bool RANGE = false, bool NT = true, int MEM = 0>
__device__ __forceinline__ void row_fast_tile(const V2<T>* __restrict__ Sg, V2<T>* __restrict__ Tout, int Ny,
                                              const Lay& lay, Nonlin q, const V2<double>* __restrict__ twN,
                                              const Slots& sl, int yi, int tile, unsigned char* smem_raw, int tid,
                                              RangePart* rtile = nullptr, RangeCheck chk = RangeCheck{}) {
  using V = V2<T>;
  constexpr int M = S::L, P = S::P, TPL = S::TPL, LD = row_ld(M);
  const int slot = slot_of(sl, yi);
  Sg += (size_t)slot * sl.sstride;
  Tout += (size_t)slot * sl.sstride;
  if (sl.M) q.M = sl.M[slot];
  const int wl = tid / TPL;
  const int lane = tid - wl * TPL;
  const int row = tile * (blockDim.x / TPL) + wl;
  const bool live = row < Ny;
  V* line = reinterpret_cast<V*>(smem_raw) + wl * LD;

  // Loads are unconditional (rows past Ny read row Ny - 1; results are not
  // stored): a load under a branch makes the join wait for it, which would
  // serialise it with the root loads.
  // FP32 bulk with a pinned mean m: the transforms carry n - m and the
  // nonlinearity is taken about m (nonlin_mu_about), so FP32 round-off scales
  // with the fluctuation, not with the mean (FP64 needs no such care)
  constexpr bool kAbout = sizeof(T) == 4;
  const bool about_m = kAbout && sl.pin;
  const double m0 = about_m ? sl.pin[slot] : 0.0;
  V v[P], wk[P], wM;
  V wt[tw_count<S>()];
  row_c2r<T, S, MEM>(Sg, lay, live ? row : Ny - 1, lane, line, twN, v, wk, wM, wt, m0);
  // the lane's min and max now; the block's after the stores, where the
  // reduction overlaps their drain
  // (in T, then once in FP64 about m0: m0 + x is monotone in x, so this is
  // the FP64 min and max of m0 + v exactly)
  double rmn = INFINITY, rmx = -INFINITY;
  ShiftState wfit{};
  if constexpr (RANGE) {
    if (chk.st) wfit = chk.st[slot];  // the fit the tile is tested against, loaded early
    if (live) {
      T lmn = T(INFINITY), lmx = T(-INFINITY);
#pragma unroll
      for (int qq = 0; qq < P; ++qq) {
        lmn = fmin(lmn, fmin(v[qq].x, v[qq].y));
        lmx = fmax(lmx, fmax(v[qq].x, v[qq].y));
      }
      rmn = m0 + (double)lmn;
      rmx = m0 + (double)lmx;
    }
  }
  bool bad = false;
  NonlinAbout about{};
  if constexpr (kAbout) about = nonlin_about(q, m0);
#pragma unroll
  for (int qq = 0; qq < P; ++qq) {
    bad |= !(isfinite(v[qq].x) && isfinite(v[qq].y));
    if (about_m)
      v[qq] = V{nonlin_mu_about<T>(v[qq].x, q, about), nonlin_mu_about<T>(v[qq].y, q, about)};
    else
      v[qq] = V{nonlin_mu<T>(v[qq].x, q), nonlin_mu<T>(v[qq].y, q)};
  }
  if (bad && live && sl.status) atomicOr(sl.status + slot, 1);
  reg_fft<T, S, false, !S::WAVE>(v, line, lane, wt);
  // unpack Y[k] from Z[k] (in registers) and Z[M-k] (pulled from its lane,
  // or through LDS when the line spans wavefronts)
  V zmr[P];
  if constexpr (kRowShuffle && S::WAVE) {
    mirror_pull<T, S>(v, zmr, lane);
  } else {
    line_sync<!S::WAVE>();
#pragma unroll
    for (int qq = 0; qq < P; ++qq) line[lpad(lane + TPL * qq)] = v[qq];
    line_sync<!S::WAVE>();
#pragma unroll
    for (int qq = 0; qq < P; ++qq) {
      const int k = lane + TPL * qq;
      zmr[qq] = line[lpad(k == 0 ? 0 : M - k)];
    }
  }
  if (live) {
#pragma unroll
    for (int qq = 0; qq < P; ++qq) {
      const int k = lane + TPL * qq;
      st_state<MEM, NT, T>(Tout + lay.at(row, k), row_unpack_pair<T>(v[qq], zmr[qq], wk[qq]));
    }
    // Z[0] is lane 0's register 0
    if (lane == 0) st_state<MEM, NT, T>(Tout + lay.at(row, M), row_unpack_pair<T>(v[0], v[0], wM));
  }
  if constexpr (RANGE) {
    block_minmax_waves(rmn, rmx);
    if (tid == 0) {
      *rtile = RangePart{rmn, rmx};
      if (chk.st) {
        if (shift_breach(wfit, rmn, rmx, chk.floor_tol, chk.ceil_tol)) chk.brk[2 * slot + chk.par] = 1;
        if (tile == 0) chk.brk[2 * slot + (chk.par ^ 1)] = 0;
      }
    }
  }
}

// RANGE: rpart holds [instance][tile] range pairs (row_fast_tile)
// WAVES > 0: the kernel is compiled for that many waves per SIMD (its
// registers held to 512 / WAVES); 0 leaves the compiler's choice.  FP32
// schedules of up to 8 points per lane come in both forms, the autotuner
// picks (a 2048-point FP32 column at its own 171 VGPRs fits two one-column
// blocks per CU, so its 1025 columns take three rounds; held to 168 they
// take two, with a few spilled registers).
constexpr int kCapWaves = 3;
constexpr int kCapFlag = 16;  // in a launcher's `var`: the WAVES = kCapWaves form

template <typename T, class S, bool RANGE, bool NT, int WAVES = 0>
__global__ void __launch_bounds__(256, (WAVES > 0 ? WAVES : 1)) k_row_fast(const V2<T>* __restrict__ Sg, V2<T>* __restrict__ Tout, int Ny,
                                                  Lay lay, Nonlin q, const V2<double>* __restrict__ twN,
                                                  Slots sl, RangePart* rpart, RangeCheck chk) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // XCD-aware order over (instance, row block): the blocks that share a row
  // group's 128-byte lines run on one XCD and meet in its L2
  const int ntile = gridDim.x;
  const int lin = xcd_remap(blockIdx.x + blockIdx.y * ntile, ntile * gridDim.y);
  const int yi = lin / ntile, tile = lin - yi * ntile;
  row_fast_tile<T, S, RANGE, NT>(Sg, Tout, Ny, lay, q, twN, sl, yi, tile, smem_raw, threadIdx.x,
                                 RANGE ? rpart + (size_t)yi * ntile + tile : nullptr, chk);
}

// ---------------------------------------------------------------------------
// range of the field for the stabilised split's check, one launch before each
// step: the row pass's inverse transforms (row_c2r, the values the step's
// nonlinearity will see) and the min and max of each row tile, written to
// part[instance][tile]; k_shift_update<RangePart> then decides per slot.
// ---------------------------------------------------------------------------
template <typename T, class S>
__global__ void __launch_bounds__(256) k_row_range(const V2<T>* __restrict__ Sg, int Ny, Lay lay,
                                                   const V2<double>* __restrict__ twN, Slots sl,
                                                   RangePart* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using V = V2<T>;
  constexpr int P = S::P, TPL = S::TPL, LD = row_ld(S::L);
  const int ntile = gridDim.x;
  const int lin = xcd_remap(blockIdx.x + blockIdx.y * ntile, ntile * gridDim.y);
  const int yi = lin / ntile, tile = lin - yi * ntile;
  Sg += (size_t)slot_of(sl, yi) * sl.sstride;
  const int wl = threadIdx.x / TPL;
  const int lane = threadIdx.x - wl * TPL;
  const int row = tile * (blockDim.x / TPL) + wl;
  const bool live = row < Ny;
  V* line = reinterpret_cast<V*>(smem_raw) + wl * LD;
  V v[P], wk[P], wM;
  V wt[tw_count<S>()];
  row_c2r<T, S>(Sg, lay, live ? row : Ny - 1, lane, line, twN, v, wk, wM, wt);
  double mn = INFINITY, mx = -INFINITY;
  if (live) {
#pragma unroll
    for (int qq = 0; qq < P; ++qq) {
      mn = fmin(mn, fmin((double)v[qq].x, (double)v[qq].y));
      mx = fmax(mx, fmax((double)v[qq].x, (double)v[qq].y));
    }
  }
  block_minmax(mn, mx);
  if (threadIdx.x == 0) part[(size_t)yi * ntile + tile] = RangePart{mn, mx};
}

// ---------------------------------------------------------------------------
// fast column kernel: forward column FFTs of S~ and T, ETD combine in
// registers, inverse column FFT, S~ written in place.  C columns per block.
// LDS: C lines of Ny (padded); roots from the FP64 table exp(-2 pi i m / Ny)
// in registers.
// ---------------------------------------------------------------------------
// Column tile `tile` (blockDim.x / TPL columns) of active instance `yi`;
// every thread of the block calls it.
// Where a block's threads sit in column tile `tile` of active instance `yi`.
// Thread t is transform lane t / C of column t % C of the tile, holding rows
// lane + TPL q.  The lanes of one load or store then cover C columns of
// consecutive rows (whole 128-byte lines when C is a multiple of the
// layout's tile width), and the loaded registers are already in the
// transforms' strided order: no LDS transpose on the way in or out.  Lanes
// past the last column of a partial tile shadow that column (in bounds,
// results discarded) so loads and barriers stay uniform.
template <typename T> struct ColPos {
  V2<T>* Sg;          // the instance's S~ (and the output)
  const V2<T>* Tin;   // its T
  V2<T>* D;           // its ETDRK2 carry, or null
  int slot, col, lane, wl;
  bool live;
};

template <typename T, class S>
__device__ __forceinline__ ColPos<T> col_pos(V2<T>* __restrict__ Sg, const V2<T>* __restrict__ Tin,
                                             V2<T>* __restrict__ D, int Nc, const Slots& sl, int yi, int tile,
                                             int tid) {
  ColPos<T> c;
  const int C = blockDim.x / S::TPL;  // columns per block, a power of two
  const int lgC = __ffs(C) - 1;
  c.wl = tid & (C - 1);
  c.lane = tid >> lgC;
  c.slot = slot_of(sl, yi);
  const int c0 = tile * C;
  c.Sg = Sg + (size_t)c.slot * sl.sstride;
  c.Tin = Tin + (size_t)c.slot * sl.sstride;
  c.D = D ? D + (size_t)c.slot * sl.sstride : nullptr;
  const int ncols = min(C, Nc - c0);
  c.live = c.wl < ncols;
  c.col = c0 + (c.live ? c.wl : ncols - 1);
  return c;
}

// ETD1 under the device-managed shift, decided in the step: where the row
// kernel flagged the slot, every block takes the field's range from all row
// tiles' pairs, refits M (the same value in every block) and first
// recomputes its columns' table entries with it (as k_tables computes them),
// before anything reads them; tile 0 stores the new fit.  The body then
// reads the entries back (each lane reads entries it wrote, or the identical
// value written for the mirrored row).
template <typename T, class S, int MODE>
__device__ __forceinline__ void col_refit(const ColPos<T>& c, const ColTables& tb, const Slots& sl, int yi,
                                          int tile, int tid) {
  constexpr int NY = S::L, P = S::P, TPL = S::TPL;
  if (MODE == COL_ETD1 && tb.brk && tb.brk[2 * c.slot + tb.par]) {
    const RangePart* __restrict__ pp = tb.part + (size_t)yi * tb.pstride;
    double mn = INFINITY, mx = -INFINITY;
    for (int b = tid; b < tb.nrow; b += blockDim.x) {
      mn = fmin(mn, pp[b].mn);
      mx = fmax(mx, pp[b].mx);
    }
    block_minmax_all(mn, mx);  // whole wavefronts: the host checks the block size
    double lo;
    TabParams p = tb.tp;
    p.M = shift_fit(mn, mx, tb.g, tb.minPhi, lo);
    p.epsilon = tb.eps[c.slot];
    p.dt = tb.sg.dt[c.slot];
    const double kx2 = tb.sg.kx2[(size_t)c.slot * p.Nc + c.col];
    const double* __restrict__ ky2 = tb.sg.ky2 + (size_t)c.slot * p.Ny;
    const size_t hb = (size_t)c.col * tb.hy + (size_t)c.slot * sl.tstride;
    double* E1w = const_cast<double*>(reinterpret_cast<const double*>(tb.E1)) + hb;
    double* Q1w = const_cast<double*>(reinterpret_cast<const double*>(tb.Q1)) + hb;
    double* Q2w = tb.Q2 ? const_cast<double*>(reinterpret_cast<const double*>(tb.Q2)) + hb : nullptr;
    // the FP32 copies (FP32 engines) at the same places
    float* E1fw = tb.E1f ? tb.E1f + hb : nullptr;
    float* Q1fw = tb.E1f ? tb.Q1f + hb : nullptr;
    float* Q2fw = tb.E1f && tb.Q2f ? tb.Q2f + hb : nullptr;
    if (c.live)
      for (int qq = 0; qq < P; ++qq) {
        const int j = ky_half(c.lane + TPL * qq, NY);
        const TabEntry t = tab_entry(-(kx2 + ky2[j]), p, Q2w != nullptr);
        E1w[j] = t.e1;
        Q1w[j] = t.q1;
        if (Q2w) Q2w[j] = t.q2;
        if (E1fw) {
          E1fw[j] = float(t.em1);
          Q1fw[j] = float(t.q1);
          if (Q2fw) Q2fw[j] = float(t.q2);
        }
      }
    if (tile == 0 && tid == 0) {
      tb.st[c.slot] = ShiftState{lo, mx, 1};
      tb.Ms[c.slot] = p.M;
    }
    asm volatile("" ::: "memory");  // the table loads that follow read memory, not these values
  }
}

// The tile's column of S~ (not in the second ETDRK2 stage) and of T.
template <typename T, class S, int MODE, int MEM>
__device__ __forceinline__ void col_load_ab(V2<T> (&a)[S::P], V2<T> (&b)[S::P], const ColPos<T>& c,
                                            const Lay& lay) {
#pragma unroll
  for (int qq = 0; qq < S::P; ++qq) {
    const size_t off = lay.at(c.lane + S::TPL * qq, c.col);
    if constexpr (MODE != COL_RK2_STAGE2) a[qq] = ld_state<MEM, T>(c.Sg + off);
    b[qq] = ld_state<MEM, T>(c.Tin + off);
  }
}

// Everything after the loads: forward column FFTs of S~ and T, the ETD
// combine in FP64 (tables FP64 in every precision, rounded once to the bulk
// type), the inverse column FFT and the store of S~.  `wt`: the lane's roots.
template <typename T, class S, int MODE, bool NT, int MEM>
__device__ __forceinline__ void col_body(V2<T> (&a)[S::P], V2<T> (&b)[S::P], const V2<T>* wt, const ColPos<T>& c,
                                         int Nc, const Lay& lay, const ColTables& tb, const Slots& sl,
                                         unsigned char* smem_raw) {
  using V = V2<T>;
  constexpr int NY = S::L, P = S::P, TPL = S::TPL, LD = col_ld(NY);
  V* line = reinterpret_cast<V*>(smem_raw) + c.wl * LD;
  const int lane = c.lane, col = c.col;
  const double* E1 = reinterpret_cast<const double*>(tb.E1) + (size_t)c.slot * sl.tstride;
  const double* Q1 = reinterpret_cast<const double*>(tb.Q1) + (size_t)c.slot * sl.tstride;
  const double* Q2 = reinterpret_cast<const double*>(tb.Q2) + (size_t)c.slot * sl.tstride;
  // FP32 bulk with a pinned mean m: the kx = 0 column (the rows' averages)
  // is transformed as S~ - m, its k = 0 coefficient pinned to 0 and m added
  // back at the store, so the transforms carry the fluctuation only (the
  // stage-2 carry D is already about m).  FP64 pins k = 0 to m directly.
  constexpr bool kAbout = sizeof(T) == 4;
  const double* pinp = sl.pin ? sl.pin + c.slot : nullptr;
  const double pin = pinp ? *pinp : 0.0;
  const double mcol = (kAbout && pinp && col == 0) ? pin : 0.0;
  if constexpr (MODE != COL_RK2_STAGE2) {
    if constexpr (kAbout) {
      if (mcol != 0.0) {  // the kx = 0 column only
#pragma unroll
        for (int qq = 0; qq < P; ++qq) a[qq].x = T(double(a[qq].x) - mcol);
      }
    }
    reg_fft<T, S, false, true>(a, line, lane, wt);
  }
  // ETD tables (and the stage-2 carry) for this lane's ky, loaded now so they
  // arrive while the second forward transform runs
  const size_t tbase = (size_t)col * NY;         // carry, [Nc][Ny]
  const size_t hbase = (size_t)col * tb.hy;      // tables, [Nc][Ny/2 + 1]
  // FP32 bulk reads the tables' FP32 copies (tb.E1f, ...; the host launches
  // no FP32 column kernel without them): half the bytes, widened to FP64 at
  // the combine, E1 as (1 + E1f) / Ny
  constexpr bool kF32Tab = sizeof(T) == 4;
  using TT = std::conditional_t<kF32Tab, float, double>;
  const size_t soff = (size_t)c.slot * sl.tstride;
  const TT* E1t = kF32Tab ? reinterpret_cast<const TT*>(tb.E1f + soff) : reinterpret_cast<const TT*>(E1);
  const TT* Q1t = kF32Tab ? reinterpret_cast<const TT*>(tb.Q1f + soff) : reinterpret_cast<const TT*>(Q1);
  const TT* Q2t = kF32Tab ? reinterpret_cast<const TT*>(tb.Q2f + soff) : reinterpret_cast<const TT*>(Q2);
  TT e1r[P], q1r[P], q2r[P];
  V dr[P];
#pragma unroll
  for (int qq = 0; qq < P; ++qq) {
    const size_t ti = hbase + ky_half(lane + TPL * qq, NY);
    if constexpr (MODE == COL_RK2_STAGE2) {
      q2r[qq] = Q2t[ti];
      dr[qq] = ld_state<MEM, T>(c.D + tbase + lane + TPL * qq);
    } else {
      e1r[qq] = E1t[ti];
      q1r[qq] = Q1t[ti];
      if constexpr (MODE == COL_RK2_STAGE1) q2r[qq] = Q2t[ti];
    }
  }
  __builtin_amdgcn_sched_barrier(0);  // table loads issued before the second transform
  reg_fft<T, S, false, true>(b, line, lane, wt);
  const double inv_ny = tb.tp.inv_ny;
  auto e1_of = [&](TT v) { return kF32Tab ? (1.0 + double(v)) * inv_ny : double(v); };

  if (c.live) {
#pragma unroll
    for (int qq = 0; qq < P; ++qq) {
      const int ky = lane + TPL * qq;
      const double bx = b[qq].x, by = b[qq].y;
      double ox, oy;
      if constexpr (MODE == COL_RK2_STAGE2) {
        const double q2 = double(q2r[qq]);
        ox = double(dr[qq].x) + q2 * bx;
        oy = double(dr[qq].y) + q2 * by;
      } else {
        const double e1 = e1_of(e1r[qq]), q1 = double(q1r[qq]);
        ox = e1 * double(a[qq].x) + q1 * bx;
        oy = e1 * double(a[qq].y) + q1 * by;
        if constexpr (MODE == COL_RK2_STAGE1) {
          const double q2 = double(q2r[qq]);
          st_state<MEM, false, T>(c.D + tbase + ky, V{T(ox - q2 * bx), T(oy - q2 * by)});
        }
      }
      if (pinp && col == 0 && ky == 0) ox = kAbout ? 0.0 : pin, oy = 0.0;  // FP32: the fluctuation's mean, 0
      b[qq] = V{T(ox), T(oy)};
    }
  }
  reg_fft<T, S, true, true>(b, line, lane, wt);
  if (c.live) {
    // real projection of the kx = 0 and kx = Nx/2 columns (see k_col_etd)
    const bool realcol = (col == 0 || col == Nc - 1);
#pragma unroll
    for (int qq = 0; qq < P; ++qq) {
      V xv = b[qq];
      if (realcol) xv.y = T(0);
      if constexpr (kAbout)
        if (mcol != 0.0) xv.x = T(double(xv.x) + mcol);
      st_state<MEM, NT, T>(c.Sg + lay.at(lane + TPL * qq, col), xv);
    }
  }
}

// Column tile `tile` (blockDim.x / TPL columns) of active instance `yi`:
// forward column FFTs of S~ and T, ETD combine in registers, inverse column
// FFT, S~ written in place.  Every thread of the block calls it.
template <typename T, class S, int MODE, bool NT = true, int MEM = 0>
__device__ __forceinline__ void col_fast_tile(V2<T>* __restrict__ Sg, const V2<T>* __restrict__ Tin,
                                              V2<T>* __restrict__ D, int Nc, const Lay& lay, const ColTables& tb,
                                              const V2<double>* __restrict__ twY, const Slots& sl, int yi, int tile,
                                              unsigned char* smem_raw, int tid) {
  const ColPos<T> c = col_pos<T, S>(Sg, Tin, D, Nc, sl, yi, tile, tid);
  col_refit<T, S, MODE>(c, tb, sl, yi, tile, tid);
  V2<T> a[S::P], b[S::P];
  col_load_ab<T, S, MODE, MEM>(a, b, c, lay);  // in flight while the first transform runs
  V2<T> wt[tw_count<S>()];
  tw_load<T, S>(wt, c.lane, twY, 1);
  // every load above is issued before the first transform starts
  __builtin_amdgcn_sched_barrier(0);
  col_body<T, S, MODE, NT, MEM>(a, b, wt, c, Nc, lay, tb, sl, smem_raw);
}

template <typename T, class S, int MODE, bool NT, int WAVES = 0>
__global__ void __launch_bounds__(256, (WAVES > 0 ? WAVES : 1)) k_col_fast(V2<T>* __restrict__ Sg, const V2<T>* __restrict__ Tin,
                                                  V2<T>* __restrict__ D, int Nc, Lay lay, ColTables tb,
                                                  const V2<double>* __restrict__ twY, Slots sl) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // XCD-aware order over (instance, tile): consecutive tiles of an instance
  // land on the same XCD
  const int ntile = gridDim.x;
  const int lin = xcd_remap(blockIdx.x + blockIdx.y * ntile, ntile * gridDim.y);
  const int yi = lin / ntile;
  col_fast_tile<T, S, MODE, NT>(Sg, Tin, D, Nc, lay, tb, twY, sl, yi, lin - yi * ntile, smem_raw, threadIdx.x);
}

// ---------------------------------------------------------------------------
// schedule variants per length (the autotuner chooses among them)
// X(L, VARIANT, P, radices...)
// ---------------------------------------------------------------------------
#define PFC_SCHEDULES(X)              \
  X(32, 0, 8, 8, 4)                   \
  X(48, 0, 12, 4, 4, 3)               \
  X(64, 0, 8, 8, 8)                   \
  X(64, 1, 4, 4, 4, 4)                \
  X(96, 0, 12, 4, 4, 2, 3)            \
  X(96, 1, 12, 3, 4, 4, 2)            \
  X(128, 0, 8, 8, 8, 2)               \
  X(128, 1, 4, 4, 4, 4, 2)            \
  X(192, 0, 12, 4, 4, 4, 3)           \
  X(192, 1, 12, 3, 4, 4, 4)           \
  X(256, 0, 8, 8, 8, 4)               \
  X(256, 1, 4, 4, 4, 4, 4)            \
  X(256, 2, 16, 16, 16)               \
  X(384, 0, 12, 4, 4, 4, 2, 3)        \
  X(512, 0, 8, 8, 8, 8)               \
  X(512, 1, 4, 4, 4, 4, 4, 2)         \
  X(512, 2, 16, 16, 16, 2)            \
  X(768, 0, 12, 4, 4, 4, 4, 3)        \
  X(1024, 0, 16, 16, 8, 8)            \
  X(1024, 1, 8, 8, 8, 8, 2)           \
  X(1024, 2, 4, 4, 4, 4, 4, 4)        \
  X(2048, 0, 16, 16, 16, 8)           \
  X(2048, 1, 8, 8, 8, 8, 4)

// host-visible description of a schedule
struct SchedInfo {
  int L, variant, P;
};

// Launchers, instantiated in the pfc_inst_*.hip translation units.  Each
// returns false if (L, variant) has no schedule.
// Per (bulk type, store kind) launchers, instantiated in the pfc_inst_*.hip
// translation units (one per kind, precision and store kind, so the library
// builds in parallel).  Each returns false if (L, variant) has no schedule.
template <typename T, bool NT>
bool row_fast_launch_impl(int M, int var, dim3 grid, dim3 block, size_t lds, hipStream_t st, const V2<T>* S,
                          V2<T>* Tout, int Ny, const Lay& lay, const Nonlin& q, const V2<double>* tw, const Slots& sl,
                          RangePart* rpart, const RangeCheck& chk) {
#define PFC_X(LEN, VAR, PP, ...)                                                                                 \
  if (M == LEN && (var & (kCapFlag - 1)) == VAR) {                                                               \
    using Sc = Sch<LEN, PP, __VA_ARGS__>;                                                                        \
    if constexpr (sizeof(T) == 4 && PP <= 8) {                                                                   \
      if (var & kCapFlag) {                                                                                      \
        if (rpart)                                                                                               \
          hipLaunchKernelGGL((k_row_fast<T, Sc, true, NT, kCapWaves>), grid, block, lds, st, S, Tout, Ny, lay, q, \
                             tw, sl, rpart, chk);                                                                \
        else                                                                                                     \
          hipLaunchKernelGGL((k_row_fast<T, Sc, false, NT, kCapWaves>), grid, block, lds, st, S, Tout, Ny, lay,   \
                             q, tw, sl, rpart, chk);                                                             \
        return true;                                                                                             \
      }                                                                                                          \
    }                                                                                                            \
    if (rpart)                                                                                                   \
      hipLaunchKernelGGL((k_row_fast<T, Sc, true, NT>), grid, block, lds, st, S, Tout, Ny, lay, q, tw, sl, rpart, \
                         chk);                                                                                   \
    else                                                                                                         \
      hipLaunchKernelGGL((k_row_fast<T, Sc, false, NT>), grid, block, lds, st, S, Tout, Ny, lay, q, tw, sl,      \
                         rpart, chk);                                                                            \
    return true;                                                                                                 \
  }
  PFC_SCHEDULES(PFC_X)
#undef PFC_X
  return false;
}

template <typename T, bool NT>
bool col_fast_launch_impl(int Ny, int var, int mode, dim3 grid, dim3 block, size_t lds, hipStream_t st, V2<T>* S,
                          const V2<T>* Tin, V2<T>* D, int Nc, const Lay& lay, const ColTables& tb,
                          const V2<double>* tw, const Slots& sl) {
#define PFC_X(LEN, VAR, PP, ...)                                                                                 \
  if (Ny == LEN && (var & (kCapFlag - 1)) == VAR) {                                                              \
    using Sc = Sch<LEN, PP, __VA_ARGS__>;                                                                        \
    if constexpr (sizeof(T) == 4 && PP <= 8) {                                                                   \
      if (var & kCapFlag) {                                                                                      \
        if (mode == COL_ETD1)                                                                                    \
          hipLaunchKernelGGL((k_col_fast<T, Sc, COL_ETD1, NT, kCapWaves>), grid, block, lds, st, S, Tin, D, Nc,   \
                             lay, tb, tw, sl);                                                                   \
        else if (mode == COL_RK2_STAGE1)                                                                         \
          hipLaunchKernelGGL((k_col_fast<T, Sc, COL_RK2_STAGE1, NT, kCapWaves>), grid, block, lds, st, S, Tin, D, \
                             Nc, lay, tb, tw, sl);                                                               \
        else                                                                                                     \
          hipLaunchKernelGGL((k_col_fast<T, Sc, COL_RK2_STAGE2, NT, kCapWaves>), grid, block, lds, st, S, Tin, D, \
                             Nc, lay, tb, tw, sl);                                                               \
        return true;                                                                                             \
      }                                                                                                          \
    }                                                                                                            \
    if (mode == COL_ETD1)                                                                                        \
      hipLaunchKernelGGL((k_col_fast<T, Sc, COL_ETD1, NT>), grid, block, lds, st, S, Tin, D, Nc, lay, tb, tw, sl); \
    else if (mode == COL_RK2_STAGE1)                                                                             \
      hipLaunchKernelGGL((k_col_fast<T, Sc, COL_RK2_STAGE1, NT>), grid, block, lds, st, S, Tin, D, Nc, lay, tb, tw, \
                         sl);                                                                                    \
    else                                                                                                         \
      hipLaunchKernelGGL((k_col_fast<T, Sc, COL_RK2_STAGE2, NT>), grid, block, lds, st, S, Tin, D, Nc, lay, tb, tw, \
                         sl);                                                                                    \
    return true;                                                                                                 \
  }
  PFC_SCHEDULES(PFC_X)
#undef PFC_X
  return false;
}

#define PFC_LAUNCH_DECL(EXT, T, NT)                                                                              \
  EXT template bool row_fast_launch_impl<T, NT>(int, int, dim3, dim3, size_t, hipStream_t, const V2<T>*, V2<T>*,  \
                                                int, const Lay&, const Nonlin&, const V2<double>*, const Slots&,   \
                                                RangePart*, const RangeCheck&);                                    \
  EXT template bool col_fast_launch_impl<T, NT>(int, int, int, dim3, dim3, size_t, hipStream_t, V2<T>*,          \
                                                const V2<T>*, V2<T>*, int, const Lay&, const ColTables&,           \
                                                const V2<double>*, const Slots&);
#ifndef PFC_INST_UNIT
PFC_LAUNCH_DECL(extern, double, true)
PFC_LAUNCH_DECL(extern, double, false)
PFC_LAUNCH_DECL(extern, float, true)
PFC_LAUNCH_DECL(extern, float, false)
#endif

template <typename T>
inline bool row_fast_launch(int M, int var, bool nt, dim3 grid, dim3 block, size_t lds, hipStream_t st,
                            const V2<T>* S, V2<T>* Tout, int Ny, const Lay& lay, const Nonlin& q,
                            const V2<double>* tw, const Slots& sl, RangePart* rpart, const RangeCheck& chk) {
  return nt ? row_fast_launch_impl<T, true>(M, var, grid, block, lds, st, S, Tout, Ny, lay, q, tw, sl, rpart, chk)
            : row_fast_launch_impl<T, false>(M, var, grid, block, lds, st, S, Tout, Ny, lay, q, tw, sl, rpart, chk);
}
template <typename T>
inline bool col_fast_launch(int Ny, int var, int mode, bool nt, dim3 grid, dim3 block, size_t lds, hipStream_t st,
                            V2<T>* S, const V2<T>* Tin, V2<T>* D, int Nc, const Lay& lay, const ColTables& tb,
                            const V2<double>* tw, const Slots& sl) {
  return nt ? col_fast_launch_impl<T, true>(Ny, var, mode, grid, block, lds, st, S, Tin, D, Nc, lay, tb, tw, sl)
            : col_fast_launch_impl<T, false>(Ny, var, mode, grid, block, lds, st, S, Tin, D, Nc, lay, tb, tw, sl);
}

bool row_range_launch_f64(int M, int var, dim3 grid, dim3 block, size_t lds, hipStream_t st, const V2<double>* S,
                          int Ny, const Lay& lay, const V2<double>* tw, const Slots& sl, RangePart* part);
bool row_range_launch_f32(int M, int var, dim3 grid, dim3 block, size_t lds, hipStream_t st, const V2<float>* S,
                          int Ny, const Lay& lay, const V2<double>* tw, const Slots& sl, RangePart* part);

// ---------------------------------------------------------------------------
// multi-shape step launches: batches of different grid sizes ("members")
// advance together.  A launch covers the row (or column) tiles of every
// member whose transform length is in its register class: class 0 holds the
// 8-points-per-lane schedules, class 1 the 12-points-per-lane ones (lengths
// with a factor 3), which need more registers; one kernel over both would run
// every block at the larger register count (a third fewer waves per SIMD for
// class 0).  So a step takes at most two row and two column launches,
// whatever the number of sizes.  Block b of a launch is tile (b - blk0) %
// tiles of active instance (b - blk0) / tiles of the member whose block range
// holds b; every block has 256 threads, so a member's rows (columns) per
// block are 256 / TPL.  The transforms are the variant-0 schedules of each
// length (the same radix sequence as variant 0 of PFC_SCHEDULES, so a
// member's results equal its batch stepped alone with that variant).
// X(CASE, CLASS, L, P, radices...)
// ---------------------------------------------------------------------------
#define PFC_GROUP_SCHEDULES(X)       \
  X(0, 0, 32, 8, 8, 4)               \
  X(1, 1, 48, 12, 4, 4, 3)           \
  X(2, 0, 64, 8, 8, 8)               \
  X(3, 1, 96, 12, 4, 4, 2, 3)        \
  X(4, 0, 128, 8, 8, 8, 2)           \
  X(5, 1, 192, 12, 4, 4, 4, 3)       \
  X(6, 0, 256, 8, 8, 8, 4)           \
  X(7, 1, 384, 12, 4, 4, 4, 2, 3)    \
  X(8, 0, 512, 8, 8, 8, 8)

constexpr int kGroupThreads = 256;
constexpr int kGroupClasses = 2;

// schedule case of a transform length, or -1; *P = points per lane, *cls = register class
inline int group_case(int L, int* P, int* cls = nullptr) {
#define PFC_X(CASE, CLS, LEN, PP, ...) \
  if (L == LEN) {                      \
    *P = PP;                           \
    if (cls) *cls = CLS;               \
    return CASE;                       \
  }
  PFC_GROUP_SCHEDULES(PFC_X)
#undef PFC_X
  return -1;
}

// one member's part of a launch: `kcase` is the schedule case of the pass's
// transform (row: Nx/2, column: Ny), `ntile` its tiles per instance and
// `blk0` its first block in the launch
template <typename T> struct GroupMember {
  V2<T>* S;   // row spectra S~, [cap][...] in the member's layout
  V2<T>* Tb;  // row spectra of N
  V2<T>* D;   // ETDRK2 carry, or null
  const V2<double>* tw;  // FP64 roots: exp(-2 pi i m / Nx), m < Nx (rows) or exp(-2 pi i m / Ny) (columns)
  Lay lay;
  Nonlin q;
  Slots sl;      // the member's active slots
  ColTables tb;  // the member's ETD tables
  int Ny, Nc;
  int kcase, ntile, blk0;
};

template <typename T, int CLASS, bool NT = true>
__global__ void __launch_bounds__(kGroupThreads) k_row_group(const GroupMember<T>* __restrict__ mem, int nmem) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // members are laid out in block order; XCD-aware order as in k_row_fast
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  int m = 0;
  while (m + 1 < nmem && lin >= mem[m + 1].blk0) ++m;
  const GroupMember<T> g = mem[m];
  const int local = lin - g.blk0;
  const int yi = local / g.ntile;
  const int tile = local - yi * g.ntile;
  switch (g.kcase) {
#define PFC_X(CASE, CLS, LEN, PP, ...)                                                                           \
  case CASE:                                                                                                     \
    if constexpr (CLS == CLASS)                                                                                  \
      row_fast_tile<T, Sch<LEN, PP, __VA_ARGS__>, false, NT>(g.S, g.Tb, g.Ny, g.lay, g.q, g.tw, g.sl, yi, tile,   \
                                                             smem_raw, threadIdx.x);                             \
    break;
    PFC_GROUP_SCHEDULES(PFC_X)
#undef PFC_X
    default:
      break;
  }
}

template <typename T, int MODE, int CLASS, bool NT = true>
__global__ void __launch_bounds__(kGroupThreads) k_col_group(const GroupMember<T>* __restrict__ mem, int nmem) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  int m = 0;
  while (m + 1 < nmem && lin >= mem[m + 1].blk0) ++m;
  const GroupMember<T> g = mem[m];
  const int local = lin - g.blk0;
  const int yi = local / g.ntile;
  const int tile = local - yi * g.ntile;
  switch (g.kcase) {
#define PFC_X(CASE, CLS, LEN, PP, ...)                                                                          \
  case CASE:                                                                                                    \
    if constexpr (CLS == CLASS)                                                                                 \
      col_fast_tile<T, Sch<LEN, PP, __VA_ARGS__>, MODE, NT>(g.S, g.Tb, g.D, g.Nc, g.lay, g.tb, g.tw, g.sl, yi,   \
                                                            tile, smem_raw, threadIdx.x);                       \
    break;
    PFC_GROUP_SCHEDULES(PFC_X)
#undef PFC_X
    default:
      break;
  }
}

// Launchers (pfc_inst_group*.hip): kind 0 = row pass, else the column pass of
// mode kind - 1 (COL_ETD1, COL_RK2_STAGE1, COL_RK2_STAGE2); cls = register
// class; nt: non-temporal (else ordinary) output stores.
void group_launch_f64(int kind, int cls, bool nt, int blocks, size_t lds, hipStream_t st,
                      const GroupMember<double>* mem, int nmem);
void group_launch_f32(int kind, int cls, bool nt, int blocks, size_t lds, hipStream_t st,
                      const GroupMember<float>* mem, int nmem);

// ---------------------------------------------------------------------------
// persistent multi-shape step ("flow"): one launch advances the members of
// one register class by several steps.  Its work list is the sequence
//   for each step, for each phase (row pass, column pass; ETDRK2: row,
//   column stage 1, row, column stage 2), for each member (largest first),
//   for each active instance, for each of the phase's tiles
// and blocks take positions in it ("tickets") from a counter, in order.  A
// tile first waits until its instance has completed every tile of the phases
// before its own (a per-instance counter), all of them earlier tickets, then
// runs the same tile function as the per-pass launches, with device-coherent
// state accesses (st_state / ld_state, MEM 1), and counts itself done.  Each
// instance thus follows its own chain of passes while other instances' tiles
// keep the device busy: no launch boundary and no device-wide drain between
// passes or steps.
// Progress needs no co-resident blocks: the unfinished ticket with the
// lowest number depends only on finished tickets, and the block holding it
// runs it (the ticket a block prefetches is above the one it runs), so the
// launch completes with any number of blocks resident, alongside other
// kernels.  A wait that exceeds its bound (never expected) raises `abort`,
// which stops every block; k_flow_mark then flags the members' slots
// PFC_STATUS_ABORTED.
// ---------------------------------------------------------------------------
constexpr int kFlowCntStride = 64;      // counters 256 bytes apart
constexpr int kFlowSpinLimit = 1 << 21;  // default polls of one wait (~1 s) before the launch aborts
constexpr int kStatusAborted = 2;        // PFC_STATUS_ABORTED

// A ticket's place within a step, precomputed on the host (32 bytes, one
// scalar load): phase, member, active instance, tile, the instance's counter
// index, the counter value its step's earlier phases add up to, and the
// instance's tiles per step.
struct FlowEntry {
  int p, m, yi, tile, cidx;
  unsigned pre, per_inst;
  int pad;
};

template <typename T> struct FlowPlan {
  const GroupMember<T>* rows;  // [nmem] row-pass descriptors
  const GroupMember<T>* cols;  // [nmem] column-pass descriptors
  const int* cnt0;             // [nmem + 1]: first instance counter of each member
  const FlowEntry* tab;        // [per_step]: each ticket's place in a step (one scalar load to decode)
  unsigned* ticket;            // next ticket (cnt[0]), zero at launch
  unsigned* cnt;               // per-instance completed tiles, kFlowCntStride apart, zero at launch
  int* abort;                  // a wait exceeded its bound (cleared by k_flow_mark)
  int nmem, nph, per_step;
  unsigned total;              // tickets of the launch: steps * per_step
  int nwords;                  // words from cnt[0] that k_flow_reset clears
  int spin_limit;              // polls of one wait before the launch aborts
};

// before a flow launch: the ticket and the instance counters to zero (a fill
// kernel, not a memset node, so the sequence may sit in a captured graph)
template <typename U>
__global__ void k_flow_reset(U* cnt, int nwords) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += gridDim.x * blockDim.x) cnt[i] = U(0);
}

// thread 0: wait until *c >= need; false if the launch aborted
__device__ __forceinline__ bool flow_wait(const unsigned* c, unsigned need, int* abort, int limit) {
  for (int i = 0;; ++i) {
    if (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need) return true;
    if ((i & 31) == 31 && __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    if (i >= limit) {
      __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

template <typename T, int CLASS>
__device__ __forceinline__ void flow_row(const GroupMember<T>& g, int yi, int tile, unsigned char* smem, int tid) {
  switch (g.kcase) {
#define PFC_X(CASE, CLS, LEN, PP, ...)                                                                           \
  case CASE:                                                                                                     \
    if constexpr (CLS == CLASS)                                                                                  \
      row_fast_tile<T, Sch<LEN, PP, __VA_ARGS__>, false, true, 1>(g.S, g.Tb, g.Ny, g.lay, g.q, g.tw, g.sl, yi,    \
                                                                  tile, smem, tid);                              \
    break;
    PFC_GROUP_SCHEDULES(PFC_X)
#undef PFC_X
    default:
      break;
  }
}

template <typename T, int CLASS, int MODE>
__device__ __forceinline__ void flow_col(const GroupMember<T>& g, int yi, int tile, unsigned char* smem, int tid) {
  switch (g.kcase) {
#define PFC_X(CASE, CLS, LEN, PP, ...)                                                                           \
  case CASE:                                                                                                     \
    if constexpr (CLS == CLASS)                                                                                  \
      col_fast_tile<T, Sch<LEN, PP, __VA_ARGS__>, MODE, true, 1>(g.S, g.Tb, g.D, g.Nc, g.lay, g.tb, g.tw, g.sl,   \
                                                                 yi, tile, smem, tid);                           \
    break;
    PFC_GROUP_SCHEDULES(PFC_X)
#undef PFC_X
    default:
      break;
  }
}

// Waves per SIMD a flow launch is compiled for: the register budget of the
// per-pass kernels of the same class (class 0: 8 points per lane, class 1:
// 12), so the persistent kernel keeps their occupancy.
template <typename T, int CLASS> constexpr int flow_waves() {
  return sizeof(T) == 4 ? (CLASS == 0 ? 3 : 2) : (CLASS == 0 ? 2 : 1);
}

// A ticket's place in the work list (every field block-uniform).
struct FlowTicket {
  unsigned s;     // step within the launch
  int p, m;       // phase, member
  int yi, tile;   // active instance, tile
  unsigned need;  // the instance's counter value that makes the tile ready
  unsigned* c;    // the instance's counter
};

// The step's quotient is computed in vector registers; readfirstlane returns
// it to a scalar one, so the entry (and the member's descriptor found from
// it) is read with scalar loads.
template <typename T> __device__ __forceinline__ FlowTicket flow_decode(const FlowPlan<T>& f, unsigned t) {
  FlowTicket k;
  k.s = __builtin_amdgcn_readfirstlane(t / (unsigned)f.per_step);
  const FlowEntry e = f.tab[t - k.s * (unsigned)f.per_step];
  k.p = e.p;
  k.m = e.m;
  k.yi = e.yi;
  k.tile = e.tile;
  // the instance's tiles of every earlier phase: s whole steps, then the
  // phases of this step before p
  k.need = k.s * e.per_inst + e.pre;
  k.c = f.cnt + (size_t)e.cidx * kFlowCntStride;
  return k;
}

// Thread 0 keeps tickets two deep: while the block runs ticket t it knows
// the next one (tn) and has the one after in flight (atomicAdd, waited for
// only at the tile's end); tn's counter is polled under the tile's final
// store drain, so a ready tile starts without a round trip to memory.
template <typename T, int CLASS, bool RK2>
__global__ void __launch_bounds__(kGroupThreads, (flow_waves<T, CLASS>())) k_flow(FlowPlan<T> f) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ unsigned s_tk[2];
  __shared__ int s_go;
  if (threadIdx.x == 0) {
    s_tk[0] = atomicAdd(f.ticket, 1u);
    s_tk[1] = atomicAdd(f.ticket, 1u);
  }
  __syncthreads();
  // tickets are block-uniform: readfirstlane keeps them (and the descriptors
  // found from them) in scalar registers
  unsigned t = __builtin_amdgcn_readfirstlane(s_tk[0]);
  unsigned tn = __builtin_amdgcn_readfirstlane(s_tk[1]);
  __syncthreads();  // every thread has read s_tk before thread 0 writes it again
  unsigned early = 0;  // thread 0: t's counter as polled during the previous tile
  while (t < f.total) {
    const FlowTicket k = flow_decode(f, t);
    unsigned nxt = 0;  // thread 0: the ticket after tn (in flight)
    if (threadIdx.x == 0) {
      const bool ok = k.need == 0 || early >= k.need || flow_wait(k.c, k.need, f.abort, f.spin_limit);
      s_go = ok;
      if (ok) nxt = atomicAdd(f.ticket, 1u);
    }
    __syncthreads();
    if (!s_go) break;
    // the thread index, opaque to the compiler in each iteration: otherwise
    // it hoists every schedule's lane arithmetic out of the loop and keeps
    // all of it live at once (past 256 VGPRs)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    // the descriptor stays in memory: fields load where the tile uses them
    // (a copy would hold the whole struct, ~100 dwords, in registers)
    if ((k.p & 1) == 0) {
      flow_row<T, CLASS>(f.rows[k.m], k.yi, k.tile, smem_raw, tid);
    } else {
      if constexpr (RK2) {
        if (k.p == 1)
          flow_col<T, CLASS, COL_RK2_STAGE1>(f.cols[k.m], k.yi, k.tile, smem_raw, tid);
        else
          flow_col<T, CLASS, COL_RK2_STAGE2>(f.cols[k.m], k.yi, k.tile, smem_raw, tid);
      } else {
        flow_col<T, CLASS, COL_ETD1>(f.cols[k.m], k.yi, k.tile, smem_raw, tid);
      }
    }
    // tn's counter, polled now so its latency hides under the drain below
    unsigned e2 = 0;
    if (threadIdx.x == 0 && tn < f.total)
      e2 = __hip_atomic_load(flow_decode(f, tn).c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // every thread's stores have completed (written through) before the
    // tile counts as done
    __builtin_amdgcn_s_waitcnt(0);
    if (threadIdx.x == 0) s_tk[0] = nxt;
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(k.c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    t = tn;
    tn = __builtin_amdgcn_readfirstlane(s_tk[0]);
    early = e2;
  }
}

// After a flow launch: when it aborted, flag every member's active slots
// and clear the flag for the next launch.  One block.
template <typename T>
__global__ void k_flow_mark(const GroupMember<T>* __restrict__ rows, const int* __restrict__ cnt0, int nmem,
                            int* abort) {
  if (*abort == 0) return;
  for (int m = 0; m < nmem; ++m) {
    const Slots& sl = rows[m].sl;
    for (int yi = threadIdx.x; yi < cnt0[m + 1] - cnt0[m]; yi += blockDim.x)
      if (sl.status) atomicOr(sl.status + slot_of(sl, yi), kStatusAborted);
  }
  __syncthreads();
  if (threadIdx.x == 0) *abort = 0;
}

// Launchers (pfc_inst_group*.hip): a flow launch of class `cls` (rk2: the
// ETDRK2 phases), and its resident blocks per CU at `lds` bytes of LDS.
void flow_launch_f64(int cls, bool rk2, int blocks, size_t lds, hipStream_t st, const FlowPlan<double>& f);
void flow_launch_f32(int cls, bool rk2, int blocks, size_t lds, hipStream_t st, const FlowPlan<float>& f);
int flow_occupancy_f64(int cls, bool rk2, size_t lds);
int flow_occupancy_f32(int cls, bool rk2, size_t lds);

}  // namespace pfc
