// Batched DCT-I / DST-I transforms on column-major 2-D grids (x fastest).
//
// A 2-D transform is two line passes.  The x pass reads whole contiguous
// x-lines; the y pass reads tiles of TX adjacent columns so every row segment
// is a coalesced TX-wide load.  Two real lines are packed as the real and
// imaginary parts of one complex line, extended evenly (cos) or oddly (sin) to
// length 2(N-1), and transformed by one LDS FFT; the even/odd symmetry makes
// both halves recoverable from the real and imaginary parts of the result.
//
// Normalisations follow fourier_FFTW.f90:178-229:
//   forward (space -> Fourier): out_k * s2f_k, s2f = 1/(2(N-1)) at the ends,
//                               1/(N-1) inside;
//   inverse (Fourier -> space): in_j * f2s_j, f2s = 1 at the ends, 1/2 inside,
// so that the inverse evaluates sum_k a_k cos(k pi x / L) at the nodes.
// Sine transforms use interior points only with the end values forced to
// zero (fourier_FFTW.f90:284-299).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "fft_lds.h"
#include "fft_reg.h"

namespace hosnwt {

using hosfft::C2;
using hosfft::FftPlan;

#define HIP_OK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " +   \
                               __FILE__ + ":" + std::to_string(__LINE__) + " (" #expr ")");  \
  } while (0)

enum Kind { KCOS = 0, KSIN = 1 };
enum Dir { FWD = 1, INV = -1 };
enum Mul { MUL_NONE = 0, MUL_NEG_KX = 1, MUL_NEG_KY = 2, MUL_KTH = 3, MUL_KTH2 = 4 };

// ---------------------------------------------------------------------------
// source / destination functors
// ---------------------------------------------------------------------------
// Fields are read and written through buffer resources: the address is the
// field's base (scalar registers) plus a 32-bit byte offset, and the hardware
// returns zero for a load and drops a store whose offset lies at or beyond the
// resource's extent.  An element outside the valid extents takes the offset
// kBufOff (2^31), beyond every extent, so that masked loads and stores need
// neither a branch nor a select on the value.  Every field a transform touches
// holds fewer than 2^31 bytes (checked where a context is created).
constexpr unsigned kBufOff = 0x80000000u;
constexpr int kBufFlags = 0x00020000;  // descriptor word 3 of a raw buffer on gfx9-family chips

template <typename U>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const U* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<U*>(p), (short)0, (int)bytes, kBufFlags);
}
template <typename U>
__device__ __forceinline__ U buf_load(__amdgpu_buffer_rsrc_t r, unsigned off) {
  static_assert(sizeof(U) == 4 || sizeof(U) == 8, "4- or 8-byte elements");
  if constexpr (sizeof(U) == 4)
    return __builtin_bit_cast(U, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  else
    return __builtin_bit_cast(U, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
// two adjacent elements in one access (4-byte aligned for FP32 pairs: the
// chip's buffer accesses need dword alignment only)
template <typename U>
struct BufPair {
  U x, y;
};
template <typename U>
__device__ __forceinline__ BufPair<U> buf_load2(__amdgpu_buffer_rsrc_t r, unsigned off) {
  if constexpr (sizeof(U) == 4)
    return __builtin_bit_cast(BufPair<U>, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  else
    return __builtin_bit_cast(BufPair<U>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
template <typename U>
__device__ __forceinline__ void buf_store2(__amdgpu_buffer_rsrc_t r, unsigned off, BufPair<U> v) {
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  typedef unsigned u4 __attribute__((ext_vector_type(4)));
  if constexpr (sizeof(U) == 4)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, off, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, off, 0, 0);
}
template <typename U>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, unsigned off, U v) {
  static_assert(sizeof(U) == 4 || sizeof(U) == 8, "4- or 8-byte elements");
  if constexpr (sizeof(U) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
  } else {
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, off, 0, 0);
  }
}
// byte offset of element (ix, iy) at leading dimension ld (24-bit factors)
template <typename U>
__device__ __forceinline__ unsigned elem_off(int ix, int iy, int ld) {
  return ((unsigned)ix + __umul24((unsigned)ld, (unsigned)iy)) * (unsigned)sizeof(U);
}

// Real field on a column-major grid, zero outside [0,nx) x [0,ny); optional
// spectral multiplier evaluated from the mode indices (kx = ix*pix, ky = iy*piy).
// U is the stored type, T the transform type.
template <typename T, typename U = T>
struct Src {
  const U* __restrict__ p;
  int ld, nx, ny;
  int mul, jpow;
  double pix, piy;
  // the valid extent in bytes: elements ix < nx of rows iy < ny
  __device__ __forceinline__ unsigned extent() const {
    return (nx > 0 && ny > 0) ? ((unsigned)ld * (unsigned)(ny - 1) + (unsigned)nx) * (unsigned)sizeof(U) : 0u;
  }
  // Branch-free load: an element outside the extents reads as zero (its offset
  // is beyond the resource), so that a thread can keep many loads in flight.
  __device__ __forceinline__ U raw(int ix, int iy) const { return raw_if(ix, iy, true); }
  // the same, zero also where ok is false
  __device__ __forceinline__ U raw_if(int ix, int iy, bool ok) const {
    ok = ok & (ix < nx) & (iy < ny);
    return buf_load<U>(buf_rsrc(p, extent()), ok ? elem_off<U>(ix, iy, ld) : kBufOff);
  }
  // elements (ix, iy) and (ix + 1, iy) in one access; both zero where ok is false.
  // The caller guarantees ix + 1 < nx.
  __device__ __forceinline__ BufPair<U> raw2_if(int ix, int iy, bool ok) const {
    ok = ok & (iy < ny);
    return buf_load2<U>(buf_rsrc(p, extent()), ok ? elem_off<U>(ix, iy, ld) : kBufOff);
  }
  // Spectral multiplier of mode (ix, iy).
  __device__ __forceinline__ double mulval(int ix, int iy) const {
    if (mul == MUL_NONE) return 1.0;
    if (mul == MUL_NEG_KX) return -(double)ix * pix;
    if (mul == MUL_NEG_KY) return -(double)iy * piy;
    // kth: k^j tanh(k) for odd j, k^j for even j (HOS-NWT.f90:311-327);
    // kth2: k^j / tanh(k) for odd j (0 at k = 0), k^j for even j (:316-324)
    const double kx = (double)ix * pix, ky = (double)iy * piy;
    const double k = sqrt(kx * kx + ky * ky);
    double kj = k;
    for (int q = 1; q < jpow; ++q) kj *= k;
    if (jpow & 1) {
      const double th = (k < 20.0 ? tanh(k) : 1.0);
      if (mul == MUL_KTH2) return (ix == 0 && iy == 0) ? 0.0 : kj / th;
      kj *= th;
    }
    return kj;
  }
  __device__ __forceinline__ T apply(U v, int ix, int iy) const {
    if (mul == MUL_NONE) return (T)v;
    return (T)(mulval(ix, iy) * (double)v);
  }
  __device__ __forceinline__ T get(int ix, int iy) const { return apply(raw(ix, iy), ix, iy); }
};

template <typename T, typename U = T>
inline Src<T, U> src_plain(const U* p, int ld, int nx, int ny) {
  return Src<T, U>{p, ld, nx, ny, MUL_NONE, 0, 0.0, 0.0};
}
template <typename T, typename U = T>
inline Src<T, U> src_mul(const U* p, int ld, int nx, int ny, int mul, int jpow, double pix, double piy) {
  return Src<T, U>{p, ld, nx, ny, mul, jpow, pix, piy};
}

// Plain store: p[ix + ld*iy] = v (mode 0) or += v (mode 1).
// Stores go through a two-phase interface: pre() issues the loads an output
// needs (branch-free), put() combines and stores; put_if() stores only when ok
// (an output that is not stored takes the offset beyond the resource).  The
// destination's resource spans 2^31 - 1 bytes from its base: the offsets of
// valid outputs lie in the field, and kBufOff lies beyond.
// Perm: rows (hx > 0) and/or columns (hy > 0) de-interleaved by parity, odd
// indices first: index i is stored at i / 2 when odd and at h + i / 2 when even
// (a product with a matrix that vanishes where the parities of its two indices
// agree then reads each parity's block contiguously).
template <typename T, bool Perm = false>
struct Dst {
  T* __restrict__ p;
  int ld, mode;
  int hx = 0, hy = 0;
  static constexpr bool kPerm = Perm;
  struct Pre {
    T old;
  };
  __device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc() const { return buf_rsrc(p, 0x7fffffffu); }
  __device__ __forceinline__ unsigned off(int ix, int iy) const {
    if constexpr (Perm) {
      if (hx) ix = (ix & 1) ? (ix >> 1) : hx + (ix >> 1);
      if (hy) iy = (iy & 1) ? (iy >> 1) : hy + (iy >> 1);
    }
    return elem_off<T>(ix, iy, ld);
  }
  __device__ __forceinline__ Pre pre(int ix, int iy, bool ok) const {
    Pre q;
    q.old = (mode == 0) ? (T)0 : buf_load<T>(rsrc(), ok ? off(ix, iy) : kBufOff);
    return q;
  }
  __device__ __forceinline__ void put_if(int ix, int iy, T v, const Pre& q, bool ok) const {
    buf_store<T>(rsrc(), ok ? off(ix, iy) : kBufOff, (mode == 0) ? v : q.old + v);
  }
  __device__ __forceinline__ void put(int ix, int iy, T v, const Pre& q) const { put_if(ix, iy, v, q, true); }
  __device__ __forceinline__ void put(int ix, int iy, T v) const { put(ix, iy, v, pre(ix, iy, true)); }
  // outputs (ix, iy) and (ix + 1, iy) in one access (two where rows are permuted)
  struct Pre2 {
    BufPair<T> old;
  };
  __device__ __forceinline__ Pre2 pre2(int ix, int iy, bool ok) const {
    Pre2 q;
    if constexpr (Perm) {
      q.old = BufPair<T>{pre(ix, iy, ok).old, pre(ix + 1, iy, ok).old};
    } else {
      q.old = (mode == 0) ? BufPair<T>{(T)0, (T)0} : buf_load2<T>(rsrc(), ok ? elem_off<T>(ix, iy, ld) : kBufOff);
    }
    return q;
  }
  __device__ __forceinline__ void put2_if(int ix, int iy, T a, T b, const Pre2& q, bool ok) const {
    if constexpr (Perm) {
      put_if(ix, iy, a, Pre{q.old.x}, ok);
      put_if(ix + 1, iy, b, Pre{q.old.y}, ok);
    } else {
      const BufPair<T> v = (mode == 0) ? BufPair<T>{a, b} : BufPair<T>{q.old.x + a, q.old.y + b};
      buf_store2<T>(rsrc(), ok ? elem_off<T>(ix, iy, ld) : kBufOff, v);
    }
  }
};
template <typename T>
using DstP = Dst<T, true>;

// ---------------------------------------------------------------------------
// line kernels
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T f2s_fac(int j, int N) { return (j == 0 || j == N - 1) ? (T)1 : (T)0.5; }

// x pass: workgroup `pair` transforms lines 2*pair and 2*pair+1 (< nlines).
template <typename T, int VPT, class S, class D>
__global__ void __launch_bounds__(256) r2r_x_kernel(int N, int kind, int dir, int n_in, int n_out, int nlines, S src, D dst,
                             FftPlan<T> pl) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C2<T>* buf = reinterpret_cast<C2<T>*>(smem_raw);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int la = 2 * blockIdx.x, lb = la + 1;
  const bool hb = lb < nlines;
  const int L = 2 * (N - 1);
  for (int j = tid; j < N; j += nt) {
    T a = (T)0, b = (T)0;
    if (j < n_in) {
      a = src.get(j, la);
      if (hb) b = src.get(j, lb);
    }
    if (dir == INV) {
      const T f = f2s_fac<T>(j, N);
      a *= f;
      b *= f;
    }
    C2<T> z = hosfft::mk<T>(a, b);
    if (kind == KSIN && (j == 0 || j == N - 1)) z = hosfft::mk<T>((T)0, (T)0);
    buf[j] = z;
    if (j >= 1 && j <= N - 2) buf[L - j] = (kind == KCOS) ? z : hosfft::mk<T>(-z.x, -z.y);
  }
  __syncthreads();
  hosfft::fft_lines<T, VPT>(buf, 1, pl.L, pl, tid, nt);
  const T inv = (T)(1.0 / (double)L);
  for (int k = tid; k < n_out; k += nt) {
    const C2<T> Z = buf[k];
    T A, B;
    if (kind == KCOS) {
      A = Z.x;
      B = Z.y;
    } else {
      A = -Z.y;
      B = Z.x;
      if (k == 0 || k == N - 1) A = B = (T)0;
    }
    if (dir == FWD) {
      const T f = (k == 0 || k == N - 1) ? inv : (T)2 * inv;
      A *= f;
      B *= f;
    }
    dst.put(k, la, A);
    if (hb) dst.put(k, lb, B);
  }
}

// y pass: workgroup handles columns [c0, c0+TX) (< ncols), packed in pairs.
template <typename T, int VPT, class S, class D>
__global__ void __launch_bounds__(256) r2r_y_kernel(int N, int kind, int dir, int n_in, int n_out, int ncols, int TX, S src, D dst,
                             FftPlan<T> pl) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C2<T>* buf = reinterpret_cast<C2<T>*>(smem_raw);
  T* bufr = reinterpret_cast<T*>(smem_raw);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int c0 = blockIdx.x * TX;
  const int npairs = TX / 2;
  const int ls = pl.L;
  const int L = 2 * (N - 1);
  for (int idx = tid; idx < TX * N; idx += nt) {
    const int c = idx % TX, j = idx / TX;
    const int col = c0 + c;
    T v = (T)0;
    if (j < n_in && col < ncols) v = src.get(col, j);
    if (dir == INV) v *= f2s_fac<T>(j, N);
    if (kind == KSIN && (j == 0 || j == N - 1)) v = (T)0;
    bufr[2 * ((c >> 1) * ls + j) + (c & 1)] = v;
  }
  __syncthreads();
  for (int idx = tid; N > 2 && idx < npairs * (N - 2); idx += nt) {
    const int q = idx / (N - 2), j = 1 + idx % (N - 2);
    const C2<T> z = buf[q * ls + j];
    buf[q * ls + L - j] = (kind == KCOS) ? z : hosfft::mk<T>(-z.x, -z.y);
  }
  __syncthreads();
  hosfft::fft_lines<T, VPT>(buf, npairs, ls, pl, tid, nt);
  const T inv = (T)(1.0 / (double)L);
  for (int idx = tid; idx < TX * n_out; idx += nt) {
    const int c = idx % TX, k = idx / TX;
    const int col = c0 + c;
    if (col >= ncols) continue;
    const C2<T> Z = buf[(c >> 1) * ls + k];
    T v;
    if (kind == KCOS) {
      v = (c & 1) ? Z.y : Z.x;
    } else {
      v = (c & 1) ? Z.x : -Z.y;
      if (k == 0 || k == N - 1) v = (T)0;
    }
    if (dir == FWD) v *= (k == 0 || k == N - 1) ? inv : (T)2 * inv;
    dst.put(col, k, v);
  }
}

// ---------------------------------------------------------------------------
// host-side plans
// ---------------------------------------------------------------------------
// Extended-grid line lengths with a compiled radix sequence (values per thread V).
inline bool seq_supported(int L) {
  switch (L) {
    case 12288: case 8192: case 6144: case 3072: case 1536: case 768: case 4096: case 2048: case 1024: case 512:
      return true;
    default: return false;
  }
}
inline int seq_values(int L) {
  switch (L) {
    case 12288: case 6144: case 3072: case 1536: case 768: return 24;
    case 8192: case 4096: case 2048: return 16;
    default: return 8;
  }
}
// Second compiled schedule of a length (schedule 1): half the values per
// thread, twice the threads per line (up to 1024), radices 4 (and 6 / 3 / 2)
// where the values per thread are not a multiple of 8.
inline int seq_alt_values(int L) {
  switch (L) {
    case 12288: case 6144: case 3072: case 1536: case 768: return 12;
    case 8192: case 4096: case 2048: return 8;
    case 1024: case 512: return 4;
    default: return 0;
  }
}
// The radix sequences compiled for each length and schedule (launch_reg
// instantiates exactly these; launch_seq checks the match).
inline std::vector<int> seq_radices(int L, bool alt) {
  switch (L) {
    case 12288: return alt ? std::vector<int>{4, 4, 4, 4, 4, 4, 3} : std::vector<int>{8, 8, 8, 8, 3};
    case 8192: return {8, 8, 8, 8, 2};
    case 6144: return alt ? std::vector<int>{4, 4, 4, 4, 4, 6} : std::vector<int>{8, 8, 8, 4, 3};
    case 3072: return alt ? std::vector<int>{4, 4, 4, 4, 4, 3} : std::vector<int>{8, 8, 8, 2, 3};
    case 1536: return alt ? std::vector<int>{4, 4, 4, 4, 6} : std::vector<int>{8, 8, 8, 3};
    case 768: return alt ? std::vector<int>{4, 4, 4, 4, 3} : std::vector<int>{8, 8, 4, 3};
    case 4096: return {8, 8, 8, 8};
    case 2048: return {8, 8, 8, 4};
    case 1024: return alt ? std::vector<int>{4, 4, 4, 4, 4} : std::vector<int>{8, 8, 8, 2};
    case 512: return alt ? std::vector<int>{4, 4, 4, 4, 2} : std::vector<int>{8, 8, 8};
    default: return {};
  }
}
// LDS values per line of the register engine: one past the largest padded
// position any exchange of the sequence uses (hosfft2::xpos for x lines,
// hosfft2::padx for the interleaved y tiles), rounded up to even.
inline int seq_lds_stride(int L, const std::vector<int>& rs, int elem_bytes) {
  const int W = 256 / elem_bytes;
  int mx = hosfft2::padx(L - 1);
  int Ns = 1;
  for (size_t i = 0; i + 1 < rs.size(); ++i) {  // the last pass writes no exchange
    const int R = rs[i], p = L - 1;
    int x = p;
    if (Ns == 1) x = p + p / W;
    else if (Ns < W) x = p + Ns * (p / (R * Ns));
    mx = std::max(mx, x);
    Ns *= R;
  }
  const int lp = mx + 1;
  return lp + (lp & 1);
}
// Compiled lengths in increasing order (extended grids are padded to one).
constexpr int kSeqLengths[] = {512, 768, 1024, 1536, 2048, 3072, 4096, 6144, 8192, 12288};


inline bool smooth235(long L) {
  for (int r : {2, 3, 5})
    while (L % r == 0) L /= r;
  return L == 1;
}

inline void host_fft_pow2(std::vector<std::complex<double>>& a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    for (size_t i = 0; i < n; i += len) {
      for (size_t k = 0; k < len / 2; ++k) {
        const double ang = -2.0 * M_PI * (double)k / (double)len;
        const std::complex<double> w(std::cos(ang), std::sin(ang));
        const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Long x lines (beyond the register engine's compiled lengths and one LDS
// workgroup): the same packed-pair DCT-I/DST-I, with the complex FFT done as
// Stockham radix passes between two HBM buffers, one launch per pass, each
// butterfly on its own thread, so one line spreads over the whole chip.
// The line pairs of several items of a batch share each launch: blockIdx.y is
// the pair across the group, first[i] the first pair of item i.
// ws holds 2 * L * (pairs of the group) complex values; pair g uses [g*L, (g+1)*L).
// ---------------------------------------------------------------------------
struct GPairs {
  int n;
  int first[hosfft2::kMaxBatch + 1];
  __device__ __forceinline__ int item(int g) const {
    int it = 0;
    while (it + 1 < n && g >= first[it + 1]) ++it;
    return it;
  }
};

template <typename T, class S, class D>
__global__ void __launch_bounds__(256) k_gline_load(int N, int dir, int L, hosfft2::LineBatch<S, D> bt, GPairs gp,
                                                    C2<T>* __restrict__ ws) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y;
  const int it = gp.item(g);
  const int pair = g - gp.first[it];
  if (j >= N) return;
  const int kind = bt.kind[it], n_in = bt.n_in[it], nlines = bt.nlines[it];
  const S& src = bt.src[it];
  const int la = 2 * pair, lb = la + 1;
  T a = (T)0, b = (T)0;
  if (j < n_in) {
    a = src.get(j, la);
    if (lb < nlines) b = src.get(j, lb);
  }
  if (dir == INV) {
    const T f = f2s_fac<T>(j, N);
    a *= f;
    b *= f;
  }
  C2<T> z = hosfft::mk<T>(a, b);
  if (kind == KSIN && (j == 0 || j == N - 1)) z = hosfft::mk<T>((T)0, (T)0);
  C2<T>* line = ws + (size_t)g * L;
  line[j] = z;
  if (j >= 1 && j <= N - 2) line[L - j] = (kind == KCOS) ? z : hosfft::mk<T>(-z.x, -z.y);
}

// One Stockham pass (radix R, current sub-transform length Ns): butterfly j reads
// j + r*(L/R), twiddles input r by W_{Ns R}^{r (j%Ns)} and writes
// (j/Ns)*Ns*R + (j%Ns) + r*Ns (the LDS engine's stockham_pass, through HBM).
template <typename T, int R>
__global__ void __launch_bounds__(256) k_gline_pass(int L, int Ns, const C2<T>* __restrict__ in,
                                                    C2<T>* __restrict__ out, const C2<T>* __restrict__ tw) {
  const int nb = L / R;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nb) return;
  const size_t off = (size_t)blockIdx.y * L;
  C2<T> v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = in[off + j + r * nb];
  const int k = j % Ns;
  if (Ns > 1) {
    const C2<T> w1 = tw[(L / (Ns * R)) * k];
    C2<T> w = w1;
#pragma unroll
    for (int r = 1; r < R; ++r) {
      v[r] = hosfft::cmul<T>(v[r], w);
      if (r + 1 < R) w = hosfft::cmul<T>(w, w1);
    }
  }
  hosfft::Bfly<T, R>::run(v);
  C2<T>* dst = out + off + (size_t)(j - k) * R + k;
#pragma unroll
  for (int r = 0; r < R; ++r) dst[r * Ns] = v[r];
}

template <typename T, class S, class D>
__global__ void __launch_bounds__(256) k_gline_store(int N, int dir, int L, const C2<T>* __restrict__ ws,
                                                     hosfft2::LineBatch<S, D> bt, GPairs gp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = blockIdx.y;
  const int it = gp.item(g);
  const int pair = g - gp.first[it];
  const int kind = bt.kind[it], n_out = bt.n_out[it], nlines = bt.nlines[it];
  const D& dst = bt.dst[it];
  if (k >= n_out) return;
  const int la = 2 * pair, lb = la + 1;
  const C2<T> Z = ws[(size_t)g * L + k];
  T A, B;
  if (kind == KCOS) {
    A = Z.x;
    B = Z.y;
  } else {
    A = -Z.y;
    B = Z.x;
    if (k == 0 || k == N - 1) A = B = (T)0;
  }
  if (dir == FWD) {
    const T inv = (T)(1.0 / (double)L);
    const T f = (k == 0 || k == N - 1) ? inv : (T)2 * inv;
    A *= f;
    B *= f;
  }
  dst.put(k, la, A);
  if (lb < nlines) dst.put(k, lb, B);
}

// Device tables for one logical line length (complex FFT length Lorig).
template <typename T>
struct LinePlan {
  FftPlan<T> dev{};
  hosfft2::Plan2<T> dev2{};
  int V2 = 0;  // values per thread of the register-resident engine (0: not usable)
  int sched = 0;  // register-engine schedule: 0 the first compiled sequence, 1 the second
  int glob = 0;  // 1: line too long for both in-workgroup engines, multi-launch HBM passes
  std::vector<void*> allocs;
  // Register-resident plan: lengths 2^a 3^b with radices 8/4/2/3 dividing V.
  void build2(int L) {
    long r = L;
    int a = 0, b = 0;
    while (r % 2 == 0) { r /= 2; ++a; }
    while (r % 3 == 0) { r /= 3; ++b; }
    if (r != 1) return;
    (void)a;
    (void)b;
    if (!seq_supported(L)) return;
    const bool alt = sched == 1 && seq_alt_values(L) > 0;
    const int V = alt ? seq_alt_values(L) : seq_values(L);
    const std::vector<int> rs = seq_radices(L, alt);
    if (rs.empty() || (int)rs.size() > hosfft2::kMaxPass2) throw std::runtime_error("no radix sequence for this length");
    dev2.L = L;
    dev2.npass = (int)rs.size();
    for (size_t i = 0; i < rs.size(); ++i) dev2.radix[i] = rs[i];
    dev2.ntl = L / V;
    // FP32 lines exchange complex values (8 bytes), FP64 lines one component (8 bytes)
    dev2.lp = seq_lds_stride(L, rs, 8);
    std::vector<C2<T>> t0(64), t1((L + 63) / 64);
    for (int m = 0; m < 64; ++m) {
      const double ang = -2.0 * M_PI * (double)m / (double)L;
      t0[m].x = (T)std::cos(ang);
      t0[m].y = (T)std::sin(ang);
    }
    for (int m = 0; m < (L + 63) / 64; ++m) {
      const double ang = -2.0 * M_PI * (double)(64L * m % L) / (double)L;
      t1[m].x = (T)std::cos(ang);
      t1[m].y = (T)std::sin(ang);
    }
    dev2.tw0 = upload(t0);
    dev2.tw1 = upload(t1);
    V2 = V;
  }
  void build(int Lorig) {
    build2(Lorig);
    dev.Lorig = Lorig;
    long L = Lorig;
    dev.bluestein = smooth235(L) ? 0 : 1;
    if (dev.bluestein) {
      L = 1;
      while (L < 2L * Lorig - 1) L <<= 1;
    }
    dev.L = (int)L;
    // radix sequence: 8s first, then 4/2, 3, 5
    long r = L;
    int np = 0;
    while (r % 8 == 0) { dev.radix[np++] = 8; r /= 8; }
    while (r % 4 == 0) { dev.radix[np++] = 4; r /= 4; }
    while (r % 2 == 0) { dev.radix[np++] = 2; r /= 2; }
    while (r % 3 == 0) { dev.radix[np++] = 3; r /= 3; }
    while (r % 5 == 0) { dev.radix[np++] = 5; r /= 5; }
    if (r != 1 || np > hosfft::kMaxPass) throw std::runtime_error("unsupported FFT length");
    dev.npass = np;
    std::vector<C2<T>> tw(L);
    for (long m = 0; m < L; ++m) {
      const double a = -2.0 * M_PI * (double)m / (double)L;
      tw[m].x = (T)std::cos(a);
      tw[m].y = (T)std::sin(a);
    }
    dev.tw = upload(tw);
    if (dev.bluestein) {
      std::vector<std::complex<double>> c(Lorig), b(L, 0.0);
      for (long j = 0; j < Lorig; ++j) {
        const long long q = ((long long)j * j) % (2LL * Lorig);
        const double a = -M_PI * (double)q / (double)Lorig;
        c[j] = std::complex<double>(std::cos(a), std::sin(a));
      }
      for (long m = 0; m < Lorig; ++m) {
        b[m] = std::conj(c[m]);
        if (m > 0) b[L - m] = std::conj(c[m]);
      }
      host_fft_pow2(b);
      std::vector<C2<T>> ct(Lorig), bt(L);
      for (long j = 0; j < Lorig; ++j) {
        ct[j].x = (T)c[j].real();
        ct[j].y = (T)c[j].imag();
      }
      for (long j = 0; j < L; ++j) {
        bt[j].x = (T)(b[j].real() / (double)L);
        bt[j].y = (T)(b[j].imag() / (double)L);
      }
      dev.chirp = upload(ct);
      dev.bhat = upload(bt);
    }
  }
  const C2<T>* upload(const std::vector<C2<T>>& v) {
    void* d = nullptr;
    HIP_OK(hipMalloc(&d, v.size() * sizeof(C2<T>)));
    HIP_OK(hipMemcpy(d, v.data(), v.size() * sizeof(C2<T>), hipMemcpyHostToDevice));
    allocs.push_back(d);
    return reinterpret_cast<const C2<T>*>(d);
  }
  void release() {
    for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
  }
};

// Schedule of one grid's register-engine transforms: the compiled sequence of
// the x and of the y lines (0 or 1, see seq_alt_values) and the y tile width
// (columns per workgroup; 0: the default).
struct GridSched {
  int xs = 0, ys = 0, tx = 0;
  bool operator==(const GridSched& o) const { return xs == o.xs && ys == o.ys && tx == o.tx; }
};

// Fields on an n1 x n2 grid are addressed through buffer resources with 32-bit
// byte offsets (Src, Dst): such a grid holds fewer than 2^31 bytes per FP64 field.
inline void check_buffer_grid(long n1, long n2) {
  if (n1 * n2 >= (1L << 31) / 8)
    throw std::runtime_error("grid " + std::to_string(n1) + " x " + std::to_string(n2) +
                             ": a field of 2^31 bytes or more is beyond the transforms' 32-bit offsets");
}

// Transforms on one N1 x N2 grid (N2 == 1: 2-D tank, x lines only).
template <typename T>
struct Grid2 {
  int N1 = 0, N2 = 1;
  int TX = 8;
  GridSched sched{};
  LinePlan<T> px, py;
  void build(int n1, int n2, GridSched sc = {}) {
    check_buffer_grid(n1, n2);
    N1 = n1;
    N2 = n2;
    px.sched = sc.xs;
    py.sched = sc.ys;
    px.build(2 * (N1 - 1));
    if (N2 > 1) py.build(2 * (N2 - 1));
    px.glob = (!px.V2 && !lds_fits(px, 32) && !px.dev.bluestein) ? 1 : 0;
    TX = sizeof(T) == 8 ? 8 : 16;
    // keep the y tile within the LDS budget
    while (TX > 2 && N2 > 1 &&
           ((size_t)(TX / 2) * py.dev.L * sizeof(C2<T>) > 96 * 1024 || (TX / 2) * py.dev.L > 256 * hosfft::kValsPerThread))
      TX /= 2;
    if (N2 > 1 && py.V2) {
      // y tile width: 8 columns by default.  A lone 3073 x 769 pass is 3-13 %
      // faster with 16 (longer row segments), but the batched launches of a
      // right-hand side run 2-7 % faster with 8, whose smaller LDS footprint
      // keeps two tiles per CU (FP32: 8 and 16 equal on the full step)
      TX = sc.tx > 0 ? sc.tx : 8;
      while (TX > 2 && ((size_t)(TX / 2) * py.dev2.lp * 8 > 128 * 1024 || (TX / 2) * py.dev2.ntl > 512))
        TX /= 2;
    }
    sched = GridSched{px.V2 && px.sched == 1 && seq_alt_values(px.dev2.L) ? 1 : 0,
                      py.V2 && py.sched == 1 && seq_alt_values(py.dev2.L) ? 1 : 0, N2 > 1 ? TX : 0};
  }
  static bool lds_fits(const LinePlan<T>& pl, int vpt) {
    return (size_t)pl.dev.L * sizeof(C2<T>) <= 160 * 1024 && (long)pl.dev.L <= 256L * vpt;
  }
  // Throws when a line is longer than the engines hold: the register engine
  // (compiled lengths), one LDS workgroup of the Stockham/Bluestein engine, or
  // (x lines of 2,3,5-smooth length) the multi-launch HBM passes.
  // allow_glob: the caller passes the HBM workspace to every transform on this grid.
  void check_lengths(bool allow_glob) const {
    if (!(px.V2 || lds_fits(px, 32) || (allow_glob && px.glob)) || (N2 > 1 && !(py.V2 || lds_fits(py, 16))))
      throw std::runtime_error("grid " + std::to_string(N1) + " x " + std::to_string(N2) +
                               ": line transform longer than the supported lengths");
  }
  // T values of the HBM workspace of long x lines (0 when not needed)
  // line pairs the HBM workspace holds: one item's lines, and in a 2-D tank (one
  // line per item) every item of a batch, so that a batch takes one set of launches
  int gline_ws_pairs() const { return px.glob ? std::max((N2 + 1) / 2, N2 == 1 ? hosfft2::kMaxBatch : 1) : 0; }
  size_t gline_ws_elems() const { return (size_t)4 * px.dev.L * gline_ws_pairs(); }
  void release() {
    px.release();
    py.release();
  }
};

// Threads for `vals` complex values per workgroup at `vpt` values per thread.
inline int pick_threads(long vals, int vpt = hosfft::kValsPerThread) {
  long nt = (vals + vpt - 1) / vpt;
  nt = ((nt + 63) / 64) * 64;
  if (nt < 64) nt = 64;
  if (nt > 256) throw std::runtime_error("line batch too large for one workgroup");
  return (int)nt;
}

constexpr size_t kLdsMax = 160 * 1024;

template <typename K>
inline void allow_lds(K kernel, size_t bytes) {
  static std::map<const void*, size_t> granted;
  static std::mutex mu;  // contexts may be driven from several host threads
  if (bytes <= 64 * 1024) return;
  std::lock_guard<std::mutex> lk(mu);
  const void* key = reinterpret_cast<const void*>(kernel);
  if (bytes > granted[key]) {
    HIP_OK(hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    granted[key] = bytes;
  }
}

using hosfft2::kMaxBatch;
using hosfft2::LineBatch;

template <class S, class D>
inline int batch_max_lines(const LineBatch<S, D>& b) {
  int m = 0;
  for (int i = 0; i < b.n; ++i) m = std::max(m, (b.n_out[i] > 0) ? b.nlines[i] : 0);
  return m;
}

// Measurement hook: while a PassLog is installed on the calling host thread, each
// register-engine launch of x lines of N[0] points or y lines of N[1] points is
// bracketed by two events on its stream, with its algorithmic bytes.
struct PassLog {
  int N[2] = {0, 0};
  size_t cap = 0;
  std::vector<hipEvent_t> ev;  // 2 per launch, created before the measured interval
  std::vector<double> bytes;
  std::vector<int> pass;   // 0 x, 1 y
  std::vector<int> items;  // transforms in the launch
  bool full = false;
  PassLog() = default;
  PassLog(const PassLog&) = delete;
  PassLog& operator=(const PassLog&) = delete;
  ~PassLog() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};
inline thread_local PassLog* t_pass_log = nullptr;

// Algorithmic HBM bytes of one launch: every valid source element read once
// (zeros beyond a source's extents are not read), every output written once
// (read and written where the destination accumulates).
template <bool YP, class S, class D>
double batch_bytes(const LineBatch<S, D>& b, int N) {
  double s = 0.0;
  for (int i = 0; i < b.n; ++i) {
    if (b.nlines[i] <= 0 || b.n_out[i] <= 0) continue;
    const S& src = b.src[i];
    const long lines_in = std::max(0, std::min(b.nlines[i], YP ? src.nx : src.ny));
    const long per_line = std::max(0, std::min(N, YP ? src.ny : src.nx));
    s += (double)lines_in * per_line * sizeof(*src.p);
    s += (double)b.nlines[i] * b.n_out[i] * sizeof(*b.dst[i].p) * (b.dst[i].mode ? 2 : 1);
  }
  return s;
}

template <typename T, int V, bool YP, class S, class D, int NC, int... Rs>
void launch_seq_nc(const LinePlan<T>& pl, int ncl, int N, int dir, const LineBatch<S, D>& b, hipStream_t st) {
  static_assert(hosfft2::hosfft2_prod<Rs...>::value % V == 0, "V must divide the line length");
  {
    constexpr int rs[] = {Rs...};
    bool same = hosfft2::hosfft2_prod<Rs...>::value == pl.dev2.L && (int)sizeof...(Rs) == pl.dev2.npass && V == pl.V2;
    for (int i = 0; same && i < (int)sizeof...(Rs); ++i) same = rs[i] == pl.dev2.radix[i];
    if (!same) throw std::runtime_error("compiled radix sequence does not match the plan of length " +
                                        std::to_string(pl.dev2.L));
  }
  // exchange elements of 8 bytes: a component in FP64, a complex value in FP32
  const size_t lds = (size_t)ncl * pl.dev2.lp * 8 + (size_t)(64 + (pl.dev2.L + 63) / 64) * sizeof(C2<T>);
  if (2 * (N - 1) != pl.dev2.L) throw std::runtime_error("line plan does not match N=" + std::to_string(N));
  const int nt = ncl * pl.dev2.ntl;
  const int nl = batch_max_lines(b);
  if (nl <= 0) return;
  int groups = (nl + 2 * ncl - 1) / (2 * ncl);
  if (YP) groups = (groups + 7) / 8 * 8;  // XCD-aware tile order needs a multiple of 8
  auto kern = hosfft2::r2r_seq_kernel<T, V, YP, S, D, NC, Rs...>;
  allow_lds(kern, lds);
  PassLog* lg = t_pass_log;
  hipEvent_t* ev = nullptr;
  if (lg && lg->N[YP ? 1 : 0] == N) {
    if (lg->bytes.size() < lg->cap) {
      ev = &lg->ev[2 * lg->bytes.size()];
      lg->bytes.push_back(batch_bytes<YP>(b, N));
      lg->pass.push_back(YP ? 1 : 0);
      int items = 0;
      for (int i = 0; i < b.n; ++i) items += (b.nlines[i] > 0 && b.n_out[i] > 0) ? 1 : 0;
      lg->items.push_back(items);
      HIP_OK(hipEventRecord(ev[0], st));
    } else {
      lg->full = true;
    }
  }
  hipLaunchKernelGGL(kern, dim3(groups, b.n), dim3(nt), lds, st, N, dir, ncl, b, pl.dev2);
  HIP_OK(hipGetLastError());
  if (ev) HIP_OK(hipEventRecord(ev[1], st));
}

// y tiles: the lines per workgroup (tile width / 2) as a compile-time constant
template <typename T, int V, bool YP, class S, class D, int... Rs>
void launch_seq(const LinePlan<T>& pl, int ncl, int N, int dir, const LineBatch<S, D>& b, hipStream_t st) {
  if constexpr (YP) {
    switch (ncl) {
      case 1: launch_seq_nc<T, V, YP, S, D, 1, Rs...>(pl, ncl, N, dir, b, st); return;
      case 2: launch_seq_nc<T, V, YP, S, D, 2, Rs...>(pl, ncl, N, dir, b, st); return;
      case 4: launch_seq_nc<T, V, YP, S, D, 4, Rs...>(pl, ncl, N, dir, b, st); return;
      case 8: launch_seq_nc<T, V, YP, S, D, 8, Rs...>(pl, ncl, N, dir, b, st); return;
      default: throw std::runtime_error("y tile of " + std::to_string(2 * ncl) + " columns is not compiled");
    }
  } else {
    launch_seq_nc<T, V, YP, S, D, 0, Rs...>(pl, ncl, N, dir, b, st);
  }
}

// The compiled schedules: length, values per thread, radix sequence.
template <typename T, bool YP, class S, class D>
void launch_reg(const LinePlan<T>& pl, int ncl, int N, int dir, const LineBatch<S, D>& b, hipStream_t st) {
  const bool alt = pl.V2 == seq_alt_values(pl.dev2.L) && pl.V2 != seq_values(pl.dev2.L);
  switch (pl.dev2.L) {
    case 12288:
      if (alt) launch_seq<T, 12, YP, S, D, 4, 4, 4, 4, 4, 4, 3>(pl, ncl, N, dir, b, st);
      else launch_seq<T, 24, YP, S, D, 8, 8, 8, 8, 3>(pl, ncl, N, dir, b, st);
      break;
    case 8192:
      if (alt) launch_seq<T, 8, YP, S, D, 8, 8, 8, 8, 2>(pl, ncl, N, dir, b, st);
      else launch_seq<T, 16, YP, S, D, 8, 8, 8, 8, 2>(pl, ncl, N, dir, b, st);
      break;
    case 6144:
      if (alt) launch_seq<T, 12, YP, S, D, 4, 4, 4, 4, 4, 6>(pl, ncl, N, dir, b, st);
      else launch_seq<T, 24, YP, S, D, 8, 8, 8, 4, 3>(pl, ncl, N, dir, b, st);
      break;
    case 3072:
      if (alt) launch_seq<T, 12, YP, S, D, 4, 4, 4, 4, 4, 3>(pl, ncl, N, dir, b, st);
      else launch_seq<T, 24, YP, S, D, 8, 8, 8, 2, 3>(pl, ncl, N, dir, b, st);
      break;
    case 1536:
      if (alt) launch_seq<T, 12, YP, S, D, 4, 4, 4, 4, 6>(pl, ncl, N, dir, b, st);
      else launch_seq<T, 24, YP, S, D, 8, 8, 8, 3>(pl, ncl, N, dir, b, st);
      break;
    case 768:
      if (alt) launch_seq<T, 12, YP, S, D, 4, 4, 4, 4, 3>(pl, ncl, N, dir, b, st);
      else launch_seq<T, 24, YP, S, D, 8, 8, 4, 3>(pl, ncl, N, dir, b, st);
      break;
    case 4096:
      if (alt) launch_seq<T, 8, YP, S, D, 8, 8, 8, 8>(pl, ncl, N, dir, b, st);
      else launch_seq<T, 16, YP, S, D, 8, 8, 8, 8>(pl, ncl, N, dir, b, st);
      break;
    case 2048:
      if (alt) launch_seq<T, 8, YP, S, D, 8, 8, 8, 4>(pl, ncl, N, dir, b, st);
      else launch_seq<T, 16, YP, S, D, 8, 8, 8, 4>(pl, ncl, N, dir, b, st);
      break;
    case 1024:
      if (alt) launch_seq<T, 4, YP, S, D, 4, 4, 4, 4, 4>(pl, ncl, N, dir, b, st);
      else launch_seq<T, 8, YP, S, D, 8, 8, 8, 2>(pl, ncl, N, dir, b, st);
      break;
    case 512:
      if (alt) launch_seq<T, 4, YP, S, D, 4, 4, 4, 4, 2>(pl, ncl, N, dir, b, st);
      else launch_seq<T, 8, YP, S, D, 8, 8, 8>(pl, ncl, N, dir, b, st);
      break;
    default: throw std::runtime_error("no compiled radix sequence for L=" + std::to_string(pl.dev2.L));
  }
}

// x passes of every item of a batch (lines of item i: nlines[i] x-lines).
template <typename T, class S, class D>
void launch_x(const LinePlan<T>& pl, int N, int dir, const LineBatch<S, D>& b, hipStream_t st, T* gws = nullptr,
              int ws_pairs = 0) {
  const int nl = batch_max_lines(b);
  if (nl <= 0) return;
  if (pl.glob) {
    if (!gws) throw std::runtime_error("long x lines need the HBM workspace");
    const int L = pl.dev.L;
    if (L != 2 * (N - 1)) throw std::runtime_error("long x line plan does not match N");
    C2<T>* w0 = reinterpret_cast<C2<T>*>(gws);
    int i = 0;
    while (i < b.n) {
      // a group of consecutive items whose line pairs fit the workspace
      LineBatch<S, D> sub;
      GPairs gp{};
      int tot = 0, max_out = 0;
      for (; i < b.n; ++i) {
        if (b.nlines[i] <= 0 || b.n_out[i] <= 0) continue;
        const int np = (b.nlines[i] + 1) / 2;
        if (np > ws_pairs) throw std::runtime_error("long x lines: workspace too small for one item");
        if (sub.n > 0 && tot + np > ws_pairs) break;
        const int k = sub.n++;
        sub.kind[k] = b.kind[i];
        sub.n_in[k] = b.n_in[i];
        sub.n_out[k] = b.n_out[i];
        sub.nlines[k] = b.nlines[i];
        sub.src[k] = b.src[i];
        sub.dst[k] = b.dst[i];
        gp.first[k] = tot;
        tot += np;
        max_out = std::max(max_out, b.n_out[i]);
      }
      if (sub.n == 0) break;
      gp.n = sub.n;
      gp.first[sub.n] = tot;
      C2<T>* in = w0;
      C2<T>* out = w0 + (size_t)L * tot;
      hipLaunchKernelGGL((k_gline_load<T, S, D>), dim3((N + 255) / 256, tot), dim3(256), 0, st, N, dir, L, sub, gp,
                         in);
      int Ns = 1;
      for (int p = 0; p < pl.dev.npass; ++p) {
        const int R = pl.dev.radix[p];
        const dim3 g((L / R + 255) / 256, tot);
        switch (R) {
          case 8: hipLaunchKernelGGL((k_gline_pass<T, 8>), g, dim3(256), 0, st, L, Ns, in, out, pl.dev.tw); break;
          case 4: hipLaunchKernelGGL((k_gline_pass<T, 4>), g, dim3(256), 0, st, L, Ns, in, out, pl.dev.tw); break;
          case 2: hipLaunchKernelGGL((k_gline_pass<T, 2>), g, dim3(256), 0, st, L, Ns, in, out, pl.dev.tw); break;
          case 3: hipLaunchKernelGGL((k_gline_pass<T, 3>), g, dim3(256), 0, st, L, Ns, in, out, pl.dev.tw); break;
          case 5: hipLaunchKernelGGL((k_gline_pass<T, 5>), g, dim3(256), 0, st, L, Ns, in, out, pl.dev.tw); break;
          default: throw std::runtime_error("long x line: unsupported radix");
        }
        Ns *= R;
        std::swap(in, out);
      }
      hipLaunchKernelGGL((k_gline_store<T, S, D>), dim3((max_out + 255) / 256, tot), dim3(256), 0, st, N, dir, L,
                         (const C2<T>*)in, sub, gp);
      HIP_OK(hipGetLastError());
    }
    return;
  }
  if (pl.V2) {
    int ncl = 1;
    while (2 * ncl * pl.dev2.ntl <= 256 && 2 * ncl <= (nl + 1) / 2) ncl *= 2;
    launch_reg<T, false>(pl, ncl, N, dir, b, st);
    return;
  }
  const size_t lds = (size_t)pl.dev.L * sizeof(C2<T>);
  if (lds > kLdsMax) throw std::runtime_error("x line exceeds LDS: N=" + std::to_string(N));
  for (int i = 0; i < b.n; ++i) {
    if (b.nlines[i] <= 0 || b.n_out[i] <= 0) continue;
    if (pl.dev.L <= 256 * 16) {
      const int nt = pick_threads(pl.dev.L, 16);
      auto kern = r2r_x_kernel<T, 16, S, D>;
      allow_lds(kern, lds);
      hipLaunchKernelGGL(kern, dim3((b.nlines[i] + 1) / 2), dim3(nt), lds, st, N, b.kind[i], dir, b.n_in[i],
                         b.n_out[i], b.nlines[i], b.src[i], b.dst[i], pl.dev);
    } else {
      const int nt = pick_threads(pl.dev.L, 32);
      auto kern = r2r_x_kernel<T, 32, S, D>;
      allow_lds(kern, lds);
      hipLaunchKernelGGL(kern, dim3((b.nlines[i] + 1) / 2), dim3(nt), lds, st, N, b.kind[i], dir, b.n_in[i],
                         b.n_out[i], b.nlines[i], b.src[i], b.dst[i], pl.dev);
    }
    HIP_OK(hipGetLastError());
  }
}

// y passes of every item (item i: nlines[i] columns of n_in[i] valid rows).
template <typename T, class S, class D>
void launch_y(const LinePlan<T>& pl, int TX, int N, int dir, const LineBatch<S, D>& b, hipStream_t st) {
  if (batch_max_lines(b) <= 0) return;
  if (pl.V2) {
    launch_reg<T, true>(pl, TX / 2, N, dir, b, st);
    return;
  }
  const size_t lds = (size_t)(TX / 2) * pl.dev.L * sizeof(C2<T>);
  if (lds > kLdsMax) throw std::runtime_error("y tile exceeds LDS: N=" + std::to_string(N));
  const int nt = pick_threads((long)(TX / 2) * pl.dev.L, 16);
  auto kern = r2r_y_kernel<T, 16, S, D>;
  allow_lds(kern, lds);
  for (int i = 0; i < b.n; ++i) {
    if (b.nlines[i] <= 0 || b.n_out[i] <= 0) continue;
    hipLaunchKernelGGL(kern, dim3((b.nlines[i] + TX - 1) / TX), dim3(nt), lds, st, N, b.kind[i], dir, b.n_in[i],
                       b.n_out[i], b.nlines[i], TX, b.src[i], b.dst[i], pl.dev);
    HIP_OK(hipGetLastError());
  }
}

// Independent 2-D transforms on one grid in one direction (shared functor
// types): item i maps src[i] (valid extents src.nx x src.ny) to dst[i] with
// output extents nout1[i] x nout2[i], through the intermediate tmp[i]
// (nout1[i] * min(src.ny, N2) values).
template <typename T, class S, class D>
struct Batch2 {
  int n = 0;
  int kx[kMaxBatch] = {}, ky[kMaxBatch] = {}, nout1[kMaxBatch] = {}, nout2[kMaxBatch] = {};
  S src[kMaxBatch];
  D dst[kMaxBatch];
  T* tmp[kMaxBatch] = {};
  void add(int kx_, int ky_, const S& s, int o1, int o2, const D& d, T* t) {
    if (n >= kMaxBatch) throw std::runtime_error("transform batch full");
    kx[n] = kx_;
    ky[n] = ky_;
    src[n] = s;
    nout1[n] = o1;
    nout2[n] = o2;
    dst[n] = d;
    tmp[n] = t;
    ++n;
  }
};

// yfirst: the y pass runs first (on the src.nx columns that hold data) and the
// x pass, whose lines are contiguous, writes the destination; tmp[i] then
// holds min(src.nx, N1) x nout2[i] values.
template <typename T, class D>
void launch_second(bool y, const Grid2<T>& g, int dir, const LineBatch<Src<T, T>, D>& b, hipStream_t st,
                   T* gws = nullptr) {
  if (y)
    launch_y<T>(g.py, g.TX, g.N2, dir, b, st);
  else
    launch_x<T>(g.px, g.N1, dir, b, st, gws, g.gline_ws_pairs());
}

template <typename T, class S, class D>
void r2r2d_batch(const Grid2<T>& g, int dir, const Batch2<T, S, D>& b, hipStream_t st, bool yfirst = false,
                 T* gws = nullptr) {
  if (b.n == 0) return;
  if (g.N2 == 1) {
    LineBatch<S, D> x;
    x.n = b.n;
    for (int i = 0; i < b.n; ++i) {
      x.kind[i] = b.kx[i];
      x.n_in[i] = b.src[i].nx;
      x.n_out[i] = b.nout1[i];
      x.nlines[i] = 1;
      x.src[i] = b.src[i];
      x.dst[i] = b.dst[i];
    }
    launch_x<T>(g.px, g.N1, dir, x, st, gws, g.gline_ws_pairs());
    return;
  }
  LineBatch<S, Dst<T>> p1;
  LineBatch<Src<T, T>, D> p2;
  p1.n = p2.n = b.n;
  for (int i = 0; i < b.n; ++i) {
    if (!yfirst) {
      const int nl = b.src[i].ny < g.N2 ? b.src[i].ny : g.N2;
      p1.kind[i] = b.kx[i];
      p1.n_in[i] = b.src[i].nx;
      p1.n_out[i] = b.nout1[i];
      p1.nlines[i] = nl;
      p1.dst[i] = Dst<T>{b.tmp[i], b.nout1[i], 0};
      p2.kind[i] = b.ky[i];
      p2.n_in[i] = nl;
      p2.n_out[i] = b.nout2[i];
      p2.nlines[i] = b.nout1[i];
      p2.src[i] = src_plain<T, T>(b.tmp[i], b.nout1[i], b.nout1[i], nl);
    } else {
      const int nc = b.src[i].nx < g.N1 ? b.src[i].nx : g.N1;
      p1.kind[i] = b.ky[i];
      p1.n_in[i] = b.src[i].ny;
      p1.n_out[i] = b.nout2[i];
      p1.nlines[i] = nc;
      p1.dst[i] = Dst<T>{b.tmp[i], nc, 0};
      p2.kind[i] = b.kx[i];
      p2.n_in[i] = nc;
      p2.n_out[i] = b.nout1[i];
      p2.nlines[i] = b.nout2[i];
      p2.src[i] = src_plain<T, T>(b.tmp[i], nc, nc, b.nout2[i]);
    }
    p1.src[i] = b.src[i];
    p2.dst[i] = b.dst[i];
  }
  if (!yfirst) {
    // (permuted destinations are written by x passes only)
    if constexpr (D::kPerm) throw std::runtime_error("a permuted destination needs the y pass first");
    else {
      launch_x<T>(g.px, g.N1, dir, p1, st, gws, g.gline_ws_pairs());
      launch_second<T, D>(true, g, dir, p2, st, gws);
    }
  } else {
    launch_y<T>(g.py, g.TX, g.N2, dir, p1, st);
    launch_x<T>(g.px, g.N1, dir, p2, st, gws, g.gline_ws_pairs());
  }
}

// Single 2-D transform: src (valid extents src.nx, src.ny) -> dst with output
// extents nout1 x nout2.  tmp must hold nout1 * src.ny values.
template <typename T, class S, class D>
void r2r2d(const Grid2<T>& g, int kx, int ky, int dir, const S& src, int nout1, int nout2, const D& dst, T* tmp,
           hipStream_t st, T* gws = nullptr) {
  Batch2<T, S, D> b;
  b.add(kx, ky, src, nout1, nout2, dst, tmp);
  r2r2d_batch<T>(g, dir, b, st, false, gws);
}

}  // namespace hosnwt
