// Register-resident batched DCT-I / DST-I line transforms for smooth lengths.
//
// Each complex line (two real lines packed as real and imaginary parts, evenly
// or oddly extended to L = 2(N-1)) is held in registers, V values per thread,
// across all Stockham passes.  Between passes the values move through LDS, the
// real parts first and then the imaginary parts, so the LDS footprint is one
// real value per point (half of a complex buffer) and two workgroups fit on a
// CU at the extended-grid sizes.  The first pass reads its inputs straight
// from HBM in the layout it needs (element j + r L/R for butterfly j), and the
// last pass writes its outputs straight to HBM from the same layout, so HBM is
// touched once per element in each direction.  Twiddles come from a two-level
// table in LDS: w^m = T1[m / 64] * T0[m % 64].
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fft_lds.h"

namespace hosfft2 {

using hosfft::C2;
using hosfft::cmul;
using hosfft::mk;

constexpr int kMaxPass2 = 8;
constexpr int hosnwt_kind_cos = 0;
constexpr int hosnwt_kind_sin = 1;

template <typename T>
struct Plan2 {
  int L;
  int npass;
  int radix[kMaxPass2];
  int ntl;              // threads per complex line = L / V
  int lp;               // LDS stride per line (reals), padded
  const C2<T>* tw0;     // 64 entries: w^m, m < 64
  const C2<T>* tw1;     // ceil(L/64) entries: w^(64 m)
};

__host__ __device__ __forceinline__ int padx(int p) { return p + (p >> 3); }

// LDS position of element p in the exchange after a pass of radix R and
// stride Ns.  The writing pass puts lane j at (j/Ns) R Ns + j%Ns + r Ns, the
// reading pass takes aligned runs of consecutive p, and one LDS cycle serves
// 256 bytes: a run of W = 32 doubles or 64 floats.
//  x lines (one line per lane run):
//    Ns = 1        p + p/W         stride-R writes and the runs both spread
//    1 < Ns < W    p + Ns p/(R Ns) the Ns-long write runs land Ns banks apart
//    Ns >= W       p               writes are runs already
//  y tiles (lane = position-major, line-minor; element p of line cl at
//  xpos(p) * ncl + cl): p + p/8 spreads both patterns over a run of W/ncl.
template <typename T, bool YP, int R, int Ns>
__device__ __forceinline__ int xpos(int p) {
  constexpr int W = 256 / (int)sizeof(T);
  if constexpr (YP) return p + (p >> 3);
  else if constexpr (Ns == 1) return p + p / W;
  else if constexpr (Ns < W) return p + Ns * (p / (R * Ns));
  else return p;
}

template <int A, int... B>
struct hosfft2_first {
  static constexpr int value = A;
};
template <int... Rs>
struct hosfft2_last;
template <int A>
struct hosfft2_last<A> {
  static constexpr int value = A;
};
template <int A, int... B>
struct hosfft2_last<A, B...> {
  static constexpr int value = hosfft2_last<B...>::value;
};
template <int... Rs>
struct hosfft2_prod {
  static constexpr int value = (Rs * ... * 1);
};

template <typename T>
__device__ __forceinline__ C2<T> twiddle(const C2<T>* t0, const C2<T>* t1, int m) {
  return cmul<T>(t1[m >> 6], t0[m & 63]);
}

// Butterflies of radix R on the V/R groups held by this thread.
template <typename T, int V, int R>
__device__ __forceinline__ void pass_compute(C2<T>* v, int tl, int ntl, int Ns, int L, const C2<T>* t0,
                                             const C2<T>* t1) {
  const int twstep = L / (Ns * R);
#pragma unroll
  for (int b = 0; b < V / R; ++b) {
    const int j = tl + b * ntl;
    const int k = j % Ns;
    if (Ns > 1) {
      // one table lookup per butterfly; higher powers by repeated products
      const C2<T> w1 = twiddle<T>(t0, t1, twstep * k);
      C2<T> w = w1;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        v[b * R + r] = cmul<T>(v[b * R + r], w);
        if (r + 1 < R) w = cmul<T>(w, w1);
      }
    }
    hosfft::Bfly<T, R>::run(v + b * R);
  }
}

// Write this pass's outputs (component c of each value) to LDS at their
// Stockham positions (j - k) R + k + r Ns.
template <typename T, int V, int R, int Ns, bool YP>
__device__ __forceinline__ void pass_write(const C2<T>* v, T* line, int es, int tl, int ntl, int comp) {
#pragma unroll
  for (int b = 0; b < V / R; ++b) {
    const int j = tl + b * ntl;
    const int k = j % Ns;
    const int o = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r)
      line[(YP ? es : 1) * xpos<T, YP, R, Ns>(o + r * Ns)] = comp ? v[b * R + r].y : v[b * R + r].x;
  }
}

// Read the next pass's inputs (radix R2): element j + r L/R2 for j = tl + b ntl.
template <typename T, int V, int R2, int RW, int NsW, bool YP>
__device__ __forceinline__ void pass_read(C2<T>* v, const T* line, int es, int tl, int ntl, int L, int comp) {
  const int stride = L / R2;
#pragma unroll
  for (int b = 0; b < V / R2; ++b) {
    const int j = tl + b * ntl;
#pragma unroll
    for (int r = 0; r < R2; ++r) {
      const T x = line[(YP ? es : 1) * xpos<T, YP, RW, NsW>(j + r * stride)];
      if (comp)
        v[b * R2 + r].y = x;
      else
        v[b * R2 + r].x = x;
    }
  }
}

// All passes of a compile-time radix sequence.  On entry v holds the
// first-pass layout, on exit the last-pass layout (value b*R_last + r at
// position tl + b ntl + r L/R_last).  `line` is this line's LDS region; every
// thread of the workgroup must call.
template <typename T, int V, bool YP, int L, int Ns, int R, int... Rest>
__device__ __forceinline__ void run_passes(C2<T>* v, T* line, int es, int tl, const C2<T>* t0, const C2<T>* t1) {
  constexpr int ntl = L / V;
  pass_compute<T, V, R>(v, tl, ntl, Ns, L, t0, t1);
  if constexpr (sizeof...(Rest) > 0) {
    constexpr int R2 = hosfft2_first<Rest...>::value;
    pass_write<T, V, R, Ns, YP>(v, line, es, tl, ntl, 0);
    __syncthreads();
    pass_read<T, V, R2, R, Ns, YP>(v, line, es, tl, ntl, L, 0);
    __syncthreads();
    pass_write<T, V, R, Ns, YP>(v, line, es, tl, ntl, 1);
    __syncthreads();
    pass_read<T, V, R2, R, Ns, YP>(v, line, es, tl, ntl, L, 1);
    __syncthreads();
    run_passes<T, V, YP, L, Ns * R, Rest...>(v, line, es, tl, t0, t1);
  }
}

// FP32 lines exchange both components at once: a complex value is one 8-byte
// LDS access, at the positions of the 8-byte scheme (W = 32), so each exchange
// takes one write, one read and two barriers instead of two of each and four.
// (FP64 keeps one component at a time: a whole complex line of 6144 points
// would hold 96 KiB of LDS and one workgroup per CU.)
template <typename T, int V, int R, int Ns, bool YP>
__device__ __forceinline__ void pass_write_cx(const C2<T>* v, C2<T>* line, int es, int tl, int ntl) {
#pragma unroll
  for (int b = 0; b < V / R; ++b) {
    const int j = tl + b * ntl;
    const int k = j % Ns;
    const int o = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) line[(YP ? es : 1) * xpos<double, YP, R, Ns>(o + r * Ns)] = v[b * R + r];
  }
}

template <typename T, int V, int R2, int RW, int NsW, bool YP>
__device__ __forceinline__ void pass_read_cx(C2<T>* v, const C2<T>* line, int es, int tl, int ntl, int L) {
  const int stride = L / R2;
#pragma unroll
  for (int b = 0; b < V / R2; ++b) {
    const int j = tl + b * ntl;
#pragma unroll
    for (int r = 0; r < R2; ++r) v[b * R2 + r] = line[(YP ? es : 1) * xpos<double, YP, RW, NsW>(j + r * stride)];
  }
}

template <typename T, int V, bool YP, int L, int Ns, int R, int... Rest>
__device__ __forceinline__ void run_passes_cx(C2<T>* v, C2<T>* line, int es, int tl, const C2<T>* t0,
                                              const C2<T>* t1) {
  constexpr int ntl = L / V;
  pass_compute<T, V, R>(v, tl, ntl, Ns, L, t0, t1);
  if constexpr (sizeof...(Rest) > 0) {
    constexpr int R2 = hosfft2_first<Rest...>::value;
    pass_write_cx<T, V, R, Ns, YP>(v, line, es, tl, ntl);
    __syncthreads();
    pass_read_cx<T, V, R2, R, Ns, YP>(v, line, es, tl, ntl, L);
    __syncthreads();
    run_passes_cx<T, V, YP, L, Ns * R, Rest...>(v, line, es, tl, t0, t1);
  }
}

// Position of value i (= b*R + r) in the layout of radix R.
template <int R>
__device__ __forceinline__ int layout_pos(int i, int tl, int ntl, int L) {
  const int b = i / R, r = i - (i / R) * R;
  return tl + b * ntl + r * (L / R);
}

// N is a compile-time constant at the call (L / 2 + 1), so that the mirror and
// edge tests of most values fold.  A value outside the input extent, or of a
// line beyond the batch, is loaded as zero (the load's offset lies beyond the
// field's buffer resource).
// P2 (y tiles whose line pairs all hold data): the two columns of a line pair
// are adjacent, so each row's pair is one load.
template <typename T, int V, int R, bool YP, bool P2, class S>
__device__ __forceinline__ void load_line(C2<T>* v, const S& src, int tl, int ntl, int L, int N, int kind, int dir,
                                          int n_in, int la, int lb, bool ha, bool hb, const T* mult) {
  // phase 1: issue every load (branch-free); phase 2: multipliers, signs
  decltype(src.raw(0, 0)) ra[V], rb[V];
#pragma unroll
  for (int i = 0; i < V; ++i) {
    const int p = layout_pos<R>(i, tl, ntl, L);
    const int e = (p >= N) ? L - p : p;
    const bool in = e < n_in;
    if constexpr (YP && P2) {
      const auto q = src.raw2_if(la, e, in);
      ra[i] = q.x;
      rb[i] = q.y;
    } else if (YP) {
      ra[i] = src.raw_if(la, e, in & ha);
      rb[i] = src.raw_if(lb, e, in & hb);
    } else {
      ra[i] = src.raw_if(e, la, in & ha);
      rb[i] = src.raw_if(e, lb, in & hb);
    }
  }
#pragma unroll
  for (int i = 0; i < V; ++i) {
    const int p = layout_pos<R>(i, tl, ntl, L);
    const bool mir = p >= N;
    const int e = mir ? L - p : p;
    const bool edge = (e == 0 || e == N - 1);
    T a = (T)ra[i], b = (T)rb[i];
    if (mult) {
      // multipliers of both real lines, precomputed in LDS: [2][N].  With a
      // source of the transform type the product is formed in that type: the
      // product of two floats is exact in double, so rounding it once to float
      // gives the float product; a double source is multiplied in double and
      // rounded once.
      const int ec = e < n_in ? e : 0;
      if constexpr (std::is_same<decltype(src.raw(0, 0)), T>::value) {
        a = mult[ec] * ra[i];
        b = mult[N + ec] * rb[i];
      } else {
        a = (T)((double)mult[ec] * (double)ra[i]);
        b = (T)((double)mult[N + ec] * (double)rb[i]);
      }
    }
    T f = (dir < 0 && !edge) ? (T)0.5 : (T)1;
    if (mir && kind == hosnwt_kind_sin) f = -f;
    const bool zero = kind == hosnwt_kind_sin && edge;  // sine ends are zero whatever they hold
    v[i] = mk<T>(zero ? (T)0 : a * f, zero ? (T)0 : b * f);
  }
}

template <typename T, int V, int R, bool YP, bool P2, class D>
__device__ __forceinline__ void store_line(const C2<T>* v, const D& dst, int tl, int ntl, int L, int N, int kind,
                                           int dir, int n_out, int la, int lb, bool ha, bool hb) {
  if constexpr (YP && P2) {
    // both columns of every line pair: one access per row
#pragma unroll
    for (int i0 = 0; i0 < V; i0 += 8) {
      typename D::Pre2 q[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int i = i0 + c;
        if (i >= V) break;
        const int k = layout_pos<R>(i, tl, ntl, L);
        q[c] = dst.pre2(la, k, k < n_out);
      }
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int i = i0 + c;
        if (i >= V) break;
        const int k = layout_pos<R>(i, tl, ntl, L);
        const C2<T> Z = v[i];
        T A, B;
        if (kind == hosnwt_kind_cos) {
          A = Z.x;
          B = Z.y;
        } else {
          A = -Z.y;
          B = Z.x;
          if (k == 0 || k == N - 1) A = B = (T)0;
        }
        if (dir > 0) {
          const T f = (k == 0 || k == N - 1) ? (T)(1.0 / (double)L) : (T)(2.0 / (double)L);
          A *= f;
          B *= f;
        }
        dst.put2_if(la, k, A, B, q[c], k < n_out);
      }
    }
    return;
  }
  const T inv = (T)(1.0 / (double)L);
  // outputs whose loads are issued together: 8, or 4 for FP64 destinations
  // that read back several values per output (fewer spills at two waves per SIMD)
  constexpr int CH = (sizeof(T) == 8 && sizeof(typename D::Pre) > 2 * sizeof(T)) ? 4 : 8;
#pragma unroll
  for (int i0 = 0; i0 < V; i0 += CH) {
    typename D::Pre qa[CH], qb[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int i = i0 + c;
      if (i >= V) break;
      const int k = layout_pos<R>(i, tl, ntl, L);
      const bool ok = k < n_out;
      if (YP) {
        qa[c] = dst.pre(la, k, ok && ha);
        qb[c] = dst.pre(lb, k, ok && hb);
      } else {
        qa[c] = dst.pre(k, la, ok && ha);
        qb[c] = dst.pre(k, lb, ok && hb);
      }
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int i = i0 + c;
      if (i >= V) break;
      const int k = layout_pos<R>(i, tl, ntl, L);
      const bool ok = k < n_out;
      const C2<T> Z = v[i];
      T A, B;
      if (kind == hosnwt_kind_cos) {
        A = Z.x;
        B = Z.y;
      } else {
        A = -Z.y;
        B = Z.x;
        if (k == 0 || k == N - 1) A = B = (T)0;
      }
      if (dir > 0) {
        const T f = (k == 0 || k == N - 1) ? inv : (T)2 * inv;
        A *= f;
        B *= f;
      }
      // (outputs not stored take the offset beyond the destination's resource)
      if (YP) {
        dst.put_if(la, k, A, qa[c], ok && ha);
        dst.put_if(lb, k, B, qb[c], ok && hb);
      } else {
        dst.put_if(k, la, A, qa[c], ok && ha);
        dst.put_if(k, lb, B, qb[c], ok && hb);
      }
    }
  }
}

// Line kernel for a compile-time radix sequence Rs (L = prod Rs).
// YP = false: x pass, ncl complex lines per workgroup, threads of a line
// contiguous; complex line q packs real lines 2q, 2q+1 of the batch.
// YP = true: y pass over TX = 2*ncl adjacent columns, lines interleaved across
// threads so that a row segment of TX columns is one coalesced access; the
// column tiles are visited in an XCD-aware order so that neighbouring tiles
// share an L2.  Spectral multipliers (kth, -kx, -ky) are evaluated once per
// element in a rolled loop into LDS before the loads.
// Workgroups of up to 1024 threads (lines of more than 512 threads each take
// a workgroup of their own).  FP64 instances are held to two waves per SIMD
// (at most 256 registers; a 1024-thread workgroup already needs four).
// FP32 bounds (x: free, y: 4 waves; 3 for destinations that read back more
// than one value per output) measured best among x 1-3 and y 3-4 waves on the
// 3-D mixed step with the first compiled sequences.
// Independent line transforms sharing one launch (blockIdx.y = item): same
// length, direction and functor types; kind, extents, source and destination
// per item.  A workgroup whose lines lie beyond its item's extent exits at
// once, before any barrier.
constexpr int kMaxBatch = 8;
template <class S, class D>
struct LineBatch {
  int n = 0;
  int kind[kMaxBatch] = {};
  int n_in[kMaxBatch] = {};
  int n_out[kMaxBatch] = {};
  int nlines[kMaxBatch] = {};
  S src[kMaxBatch];
  D dst[kMaxBatch];
};

// NC > 0 fixes the lines per workgroup at compile time (the y tiles: NC complex
// lines of 2 columns interleaved in LDS), so that the LDS positions of the
// exchanges are shifts and constant offsets rather than products with a runtime
// stride; NC = 0 takes ncl from the launch (x passes).
template <typename T, int V, bool YP, class S, class D, int NC, int... Rs>
__global__ void __launch_bounds__(hosfft2_prod<Rs...>::value / V > 512
                                      ? 1024
                                      : ((YP || hosfft2_prod<Rs...>::value / V > 256) ? 512 : 256),
                                  sizeof(T) == 8 || (YP && V >= 24) ? 2 : (!YP ? 1 : (sizeof(typename D::Pre) > 2 * sizeof(T) ? 3 : 4)))
    r2r_seq_kernel(int N, int dir, int ncl_arg, LineBatch<S, D> bt,
                                                                 Plan2<T> pl) {
  const int ncl = NC > 0 ? NC : ncl_arg;
  constexpr int L = hosfft2_prod<Rs...>::value;
  constexpr int NL = L / 2 + 1;  // points per line (the launch checks N == NL)
  (void)N;
  const int item = blockIdx.y;
  const int kind = bt.kind[item], n_in = bt.n_in[item], n_out = bt.n_out[item], nlines = bt.nlines[item];
  const S& src = bt.src[item];
  const D& dst = bt.dst[item];
  constexpr int ntl = L / V;
  constexpr int R0 = hosfft2_first<Rs...>::value;
  constexpr int Rl = hosfft2_last<Rs...>::value;
  // FP32: complex exchanges (pl.lp complex values per line); FP64: one component
  // at a time (pl.lp reals per line)
  constexpr bool CX = sizeof(T) == 4;
  constexpr int XW = CX ? 2 : 1;  // reals per exchanged element
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* lds = reinterpret_cast<T*>(smem_raw);
  C2<T>* t0 = reinterpret_cast<C2<T>*>(lds + (size_t)ncl * pl.lp * XW);
  C2<T>* t1 = t0 + 64;
  const int tid = threadIdx.x;
  int cl, tl;
  if (YP) {
    cl = tid % ncl;
    tl = tid / ncl;
  } else {
    cl = tid / ntl;
    tl = tid - cl * ntl;
  }
  // (lane ranges the compiler may use to fold the mirror tests and LDS positions)
  __builtin_assume(tl >= 0 && tl < ntl);
  __builtin_assume(cl >= 0 && cl < ncl);
  int grp = blockIdx.x;
  if (YP) {
    const int nb = gridDim.x;
    if ((nb & 7) == 0) grp = (blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3);
  }
  if (2 * grp * ncl >= nlines) return;  // uniform over the workgroup
  const int la = 2 * (grp * ncl + cl), lb = la + 1;
  const bool ha = la < nlines, hb = lb < nlines;
  // exchange buffer of this lane's line: x lines are contiguous regions of lp
  // values; the lines of a y tile are interleaved element by element
  T* line = YP ? lds + cl : lds + (size_t)cl * pl.lp;
  C2<T>* line2 = YP ? reinterpret_cast<C2<T>*>(lds) + cl : reinterpret_cast<C2<T>*>(lds) + (size_t)cl * pl.lp;
  const T* mult = nullptr;
  if (src.mul != 0) {
    // multipliers of the two real lines of every complex line: line[q*N + e]
    // (real line q of the workgroup: complex line q / 2, half q % 2; the
    // threads stride over the elements, so no index is divided)
    for (int q = 0; q < 2 * ncl; ++q) {
      const int ln = 2 * grp * ncl + q;
      T* m = lds + (size_t)(q >> 1) * pl.lp * XW + (q & 1) * NL;
      for (int e = tid; e < n_in; e += blockDim.x) m[e] = (T)(YP ? src.mulval(ln, e) : src.mulval(e, ln));
    }
    __syncthreads();
    mult = lds + (size_t)cl * pl.lp * XW;
  }
  C2<T> v[V];
  // y tiles whose line pairs all hold data (all but an item's last tile) take
  // one access per row and line pair (uniform over the workgroup)
  const bool pairs = YP && 2 * (grp + 1) * ncl <= nlines;
  if (pairs && 2 * (grp + 1) * ncl <= src.nx)
    load_line<T, V, R0, YP, true>(v, src, tl, ntl, L, NL, kind, dir, n_in, la, lb, ha, hb, mult);
  else
    load_line<T, V, R0, YP, false>(v, src, tl, ntl, L, NL, kind, dir, n_in, la, lb, ha, hb, mult);
  for (int m = tid; m < 64; m += blockDim.x) t0[m] = pl.tw0[m];
  for (int m = tid; m < (L + 63) / 64; m += blockDim.x) t1[m] = pl.tw1[m];
  __syncthreads();
  if constexpr (CX) {
    (void)line;
    run_passes_cx<T, V, YP, L, 1, Rs...>(v, line2, ncl, tl, t0, t1);
  } else {
    (void)line2;
    run_passes<T, V, YP, L, 1, Rs...>(v, line, ncl, tl, t0, t1);
  }
  if (pairs)
    store_line<T, V, Rl, YP, true>(v, dst, tl, ntl, L, NL, kind, dir, n_out, la, lb, ha, hb);
  else
    store_line<T, V, Rl, YP, false>(v, dst, tl, ntl, L, NL, kind, dir, n_out, la, lb, ha, hb);
}

}  // namespace hosfft2
