// LDS-resident batched complex FFT engine for CDNA4 (gfx950).
//
// One workgroup transforms `nlines` complex lines that live in LDS with a
// stride of `lstride` elements.  Smooth lengths (radices 2,3,4,5,8) run as a
// Stockham auto-sort sequence: each pass stages its butterfly inputs in
// registers, synchronises, and writes the outputs to their sorted positions,
// so a single LDS buffer suffices.  Lengths with other prime factors run as
// Bluestein chirp-z convolutions over a power-of-two length P >= 2L-1.
//
// Forward DFT convention: X_k = sum_j x_j exp(-2 pi i j k / L).
#pragma once
#include <hip/hip_runtime.h>

namespace hosfft {

template <typename T> struct cpx_t;
template <> struct cpx_t<double> { using type = double2; };
template <> struct cpx_t<float> { using type = float2; };
template <typename T> using C2 = typename cpx_t<T>::type;

template <typename T>
__device__ __forceinline__ C2<T> mk(T x, T y) {
  C2<T> r;
  r.x = x;
  r.y = y;
  return r;
}
template <typename T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by -i
template <typename T>
__device__ __forceinline__ C2<T> cmnj(C2<T> a) { return mk<T>(a.y, -a.x); }
template <typename T>
__device__ __forceinline__ C2<T> cconj(C2<T> a) { return mk<T>(a.x, -a.y); }
template <typename T>
__device__ __forceinline__ C2<T> cscale(C2<T> a, T s) { return mk<T>(a.x * s, a.y * s); }

constexpr int kMaxPass = 16;
// Complex values a thread stages per pass (VPT template parameter): the host
// sizes the workgroup so that nlines * L <= VPT * blockDim.x.
constexpr int kValsPerThread = 16;

// Device-side plan of one line length.  For Bluestein plans `L` is the
// power-of-two convolution length P and `Lorig` the logical length.
template <typename T>
struct FftPlan {
  int L;
  int Lorig;
  int npass;
  int bluestein;
  int radix[kMaxPass];
  const C2<T>* tw;     // exp(-2 pi i m / L), m < L
  const C2<T>* chirp;  // exp(-pi i j^2 / Lorig), j < Lorig   (Bluestein only)
  const C2<T>* bhat;   // FFT_P(conj chirp, wrapped) / P       (Bluestein only)
};

// ---------------------------------------------------------------------------
// small DFT butterflies (forward)
// ---------------------------------------------------------------------------
template <typename T, int R>
struct Bfly;

template <typename T>
struct Bfly<T, 2> {
  __device__ static void run(C2<T>* v) {
    C2<T> a = v[0], b = v[1];
    v[0] = cadd<T>(a, b);
    v[1] = csub<T>(a, b);
  }
};

template <typename T>
struct Bfly<T, 4> {
  __device__ static void run(C2<T>* v) {
    C2<T> s02 = cadd<T>(v[0], v[2]), d02 = csub<T>(v[0], v[2]);
    C2<T> s13 = cadd<T>(v[1], v[3]), d13 = cmnj<T>(csub<T>(v[1], v[3]));
    v[0] = cadd<T>(s02, s13);
    v[2] = csub<T>(s02, s13);
    v[1] = cadd<T>(d02, d13);
    v[3] = csub<T>(d02, d13);
  }
};

template <typename T>
struct Bfly<T, 8> {
  __device__ static void run(C2<T>* v) {
    const T r2 = (T)0.70710678118654752440084436210484904;
    C2<T> a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      a[r] = cadd<T>(v[r], v[r + 4]);
      b[r] = csub<T>(v[r], v[r + 4]);
    }
    // b[r] *= W8^r
    b[1] = mk<T>((b[1].x + b[1].y) * r2, (b[1].y - b[1].x) * r2);
    b[2] = cmnj<T>(b[2]);
    b[3] = mk<T>((b[3].y - b[3].x) * r2, -(b[3].x + b[3].y) * r2);
    Bfly<T, 4>::run(a);
    Bfly<T, 4>::run(b);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v[2 * q] = a[q];
      v[2 * q + 1] = b[q];
    }
  }
};

template <typename T>
struct Bfly<T, 3> {
  __device__ static void run(C2<T>* v) {
    const T c = (T)-0.5;
    const T s = (T)0.86602540378443864676372317075293618;
    C2<T> t = cadd<T>(v[1], v[2]), d = csub<T>(v[1], v[2]);
    C2<T> m = mk<T>(v[0].x + c * t.x, v[0].y + c * t.y);
    v[0] = cadd<T>(v[0], t);
    v[1] = mk<T>(m.x + s * d.y, m.y - s * d.x);
    v[2] = mk<T>(m.x - s * d.y, m.y + s * d.x);
  }
};

template <typename T>
struct Bfly<T, 5> {
  __device__ static void run(C2<T>* v) {
    const T c1 = (T)0.30901699437494742410229341718281906;
    const T c2 = (T)-0.80901699437494742410229341718281906;
    const T s1 = (T)0.95105651629515357211643933337938214;
    const T s2 = (T)0.58778525229247312916870595463907277;
    C2<T> t1 = cadd<T>(v[1], v[4]), d1 = csub<T>(v[1], v[4]);
    C2<T> t2 = cadd<T>(v[2], v[3]), d2 = csub<T>(v[2], v[3]);
    C2<T> m1 = mk<T>(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    C2<T> m2 = mk<T>(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    // n1 = s1 d1 + s2 d2 ; n2 = s2 d1 - s1 d2 ; X = m -/+ i n
    C2<T> n1 = mk<T>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
    C2<T> n2 = mk<T>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
    v[0] = cadd<T>(cadd<T>(v[0], t1), t2);
    v[1] = mk<T>(m1.x + n1.y, m1.y - n1.x);
    v[4] = mk<T>(m1.x - n1.y, m1.y + n1.x);
    v[2] = mk<T>(m2.x + n2.y, m2.y - n2.x);
    v[3] = mk<T>(m2.x - n2.y, m2.y + n2.x);
  }
};

// Radix 6 as the prime-factor (Good-Thomas) split 6 = 2 x 3, which needs no
// inner twiddles: input n = (3 n1 + 2 n2) mod 6, two 3-point transforms over
// n2, then 2-point ones over n1; output k = (3 k1 + 4 k2) mod 6.
template <typename T>
struct Bfly<T, 6> {
  __device__ static void run(C2<T>* v) {
    C2<T> a[3] = {v[0], v[2], v[4]}, b[3] = {v[3], v[5], v[1]};
    Bfly<T, 3>::run(a);
    Bfly<T, 3>::run(b);
    v[0] = cadd<T>(a[0], b[0]);
    v[3] = csub<T>(a[0], b[0]);
    v[4] = cadd<T>(a[1], b[1]);
    v[1] = csub<T>(a[1], b[1]);
    v[2] = cadd<T>(a[2], b[2]);
    v[5] = csub<T>(a[2], b[2]);
  }
};

// One Stockham pass of radix R on `nlines` lines of length L.
// Inputs of butterfly j sit at j + r*(L/R); outputs go to
// (j/Ns)*Ns*R + (j%Ns) + r*Ns, after twiddling input r by W_{Ns R}^{r (j%Ns)}.
template <typename T, int R, int VPT>
__device__ __forceinline__ void stockham_pass(C2<T>* buf, int L, int Ns, int nlines, int lstride,
                                              const C2<T>* __restrict__ tw, int tid, int nt) {
  constexpr int BMAX = (VPT + R - 1) / R;
  const int nb = L / R;
  const int total = nb * nlines;
  const int twstep = L / (Ns * R);
  C2<T> v[BMAX][R];
#pragma unroll
  for (int b = 0; b < BMAX; ++b) {
    const int idx = tid + b * nt;
    if (idx < total) {
      const int line = idx / nb;
      const int j = idx - line * nb;
      const C2<T>* src = buf + line * lstride + j;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = src[r * nb];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BMAX; ++b) {
    const int idx = tid + b * nt;
    if (idx < total) {
      const int line = idx / nb;
      const int j = idx - line * nb;
      const int k = j % Ns;
      if (Ns > 1) {
        const C2<T> w1 = tw[twstep * k];
        C2<T> w = w1;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          v[b][r] = cmul<T>(v[b][r], w);
          if (r + 1 < R) w = cmul<T>(w, w1);
        }
      }
      Bfly<T, R>::run(v[b]);
      C2<T>* dst = buf + line * lstride + (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * Ns] = v[b][r];
    }
  }
  __syncthreads();
}

// Smooth-length FFT of all lines (in place, natural order out).
template <typename T, int VPT>
__device__ void fft_smooth(C2<T>* buf, int nlines, int lstride, const FftPlan<T>& pl, int tid, int nt) {
  int Ns = 1;
  for (int p = 0; p < pl.npass; ++p) {
    const int R = pl.radix[p];
    switch (R) {
      case 8: stockham_pass<T, 8, VPT>(buf, pl.L, Ns, nlines, lstride, pl.tw, tid, nt); break;
      case 4: stockham_pass<T, 4, VPT>(buf, pl.L, Ns, nlines, lstride, pl.tw, tid, nt); break;
      case 2: stockham_pass<T, 2, VPT>(buf, pl.L, Ns, nlines, lstride, pl.tw, tid, nt); break;
      case 3: stockham_pass<T, 3, VPT>(buf, pl.L, Ns, nlines, lstride, pl.tw, tid, nt); break;
      case 5: stockham_pass<T, 5, VPT>(buf, pl.L, Ns, nlines, lstride, pl.tw, tid, nt); break;
      default: break;
    }
    Ns *= R;
  }
}

// FFT of logical length pl.Lorig.  On entry each line holds Lorig values
// (lstride >= pl.L); on exit it holds the Lorig DFT outputs.  Must be called
// by the whole workgroup after a barrier.
template <typename T, int VPT>
__device__ void fft_lines(C2<T>* buf, int nlines, int lstride, const FftPlan<T>& pl, int tid, int nt) {
  if (!pl.bluestein) {
    fft_smooth<T, VPT>(buf, nlines, lstride, pl, tid, nt);
    return;
  }
  const int P = pl.L, Lo = pl.Lorig;
  for (int idx = tid; idx < nlines * P; idx += nt) {
    const int line = idx / P, j = idx - line * P;
    C2<T>* e = buf + line * lstride + j;
    *e = (j < Lo) ? cmul<T>(*e, pl.chirp[j]) : mk<T>((T)0, (T)0);
  }
  __syncthreads();
  fft_smooth<T, VPT>(buf, nlines, lstride, pl, tid, nt);
  for (int idx = tid; idx < nlines * P; idx += nt) {
    const int line = idx / P, j = idx - line * P;
    C2<T>* e = buf + line * lstride + j;
    *e = cconj<T>(cmul<T>(*e, pl.bhat[j]));
  }
  __syncthreads();
  fft_smooth<T, VPT>(buf, nlines, lstride, pl, tid, nt);
  for (int idx = tid; idx < nlines * Lo; idx += nt) {
    const int line = idx / Lo, j = idx - line * Lo;
    C2<T>* e = buf + line * lstride + j;
    *e = cmul<T>(cconj<T>(*e), pl.chirp[j]);
  }
  __syncthreads();
}

}  // namespace hosfft
