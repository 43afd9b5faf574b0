// MI355X HOS-NWT time-stepping core: device-resident RK4(3) stepper, one
// solveHOS2 right-hand side per stage, first-order wavemaker potential.
//
// Reference path (LHEEA/HOS-NWT sources/): runge4_adapt2
// (timeintegration/runge_kutta.f90:81-360) -> solveHOS2 (main/resol_HOS.f90:270-656)
// -> phisxy_etaxy / HOSphis_modes (resol_HOS.f90:665-1078), dealias / filter_ext
// (main/dealiasing.f90), solveHOS_SWEET_1 / phiadd_SL (main/resol_wmkr.f90),
// pos_1st (main/wavemaking.f90:259-304), FFTW r2r transforms
// (fourier/fourier_FFTW.f90).
//
// The right-hand side runs the necessary transform sequence: identity round
// trips (S2F after F2S), truncations whose result is only summed into a
// filter_ext'ed total, and the dead SWENSE/PRESENT(c) branches are removed;
// each removal is exact in exact arithmetic because truncation is a linear
// projection.  The extended grid is the smallest N' >= ((M+1) n)/2 with a
// 2,3,5-smooth 2(N'-1): with dealiasing factor (M+1)/2 the retained modes of
// every product do not depend on the extended grid size (verified against the
// FP64 oracle), so the padding is exact up to rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>
#include <limits>

#include "../../include/hosnwt.h"
#include "gemm.h"
#include "r2r.h"
#include "mlp.h"

using namespace hosnwt;

namespace {

thread_local std::string g_last_error;

constexpr double kCut = 44.0;  // wavemaker sums: terms with k_add*x > kCut are below 1e-19
constexpr int L_ADD = 2;       // common_vars.f90:47

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_k(int i1, int i2, double pix, double piy) {
  const double kx = i1 * pix, ky = i2 * piy;
  return sqrt(kx * kx + ky * ky);
}
__device__ __forceinline__ double wave_omega(double k) { return sqrt(k * tanh(k)); }

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wmax(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
// block reductions for blockDim.x <= 1024 (multiple of 64)
__device__ double block_sum(double v) {
  __shared__ double red[16];
  v = wsum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = v;
  __syncthreads();
  const int nw = (blockDim.x + 63) >> 6;
  double r = 0.0;
  if (threadIdx.x < 64) {
    r = (l < nw) ? red[l] : 0.0;
    r = wsum(r);
  }
  __syncthreads();
  return r;
}
__device__ double block_max(double v) {
  __shared__ double red[16];
  v = wmax(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = v;
  __syncthreads();
  const int nw = (blockDim.x + 63) >> 6;
  double r = 0.0;
  if (threadIdx.x < 64) {
    r = (l < nw) ? red[l] : 0.0;
    r = wmax(r);
  }
  __syncthreads();
  return r;
}

inline unsigned nblk(size_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// spectral-state kernels (double unless noted)
// ---------------------------------------------------------------------------
// Ramp alpha(t), alpha', alpha'' (resol_wmkr.f90:1268-1397); false for an unknown ramp type.
__host__ __device__ inline bool ramp_hd(int iramp, double Tr, double Ts, double T_mono, double t, double& a,
                                        double& at, double& att) {
  if (iramp == 11) {
    const double Tsw = 3.0 / sqrt(5.0 / 9.81);
    if (t <= Tsw) {
      a = t / Tsw; at = 1.0 / Tsw; att = 0.0;
    } else {
      a = 1.0; at = 0.0; att = 0.0;
    }
    return true;
  }
  if (iramp == 33) {
    const double Tsw = -log(0.1) / (10.0 * T_mono);
    a = 1.0 - exp(-Tsw * t);
    at = (1.0 - a) * Tsw;
    att = (1.0 - a) * Tsw * Tsw;
    return true;
  }
  if (t <= Tr) {
    switch (iramp) {
      case 0: a = 1.0; at = 0.0; att = 0.0; return true;
      case 1: a = t / Tr; at = 1.0 / Tr; att = 0.0; return true;
      case 2: { const double s = t / Tr; a = s * s; at = 2.0 * s / Tr; att = 2.0 / (Tr * Tr); return true; }
      case 21: {
        const double T2 = Tr * Tr;
        if (t <= Tr / 2.0) { a = 2.0 * t * t / T2; at = 4.0 * t / T2; att = 4.0 / T2; }
        else { const double ts = t / Tr - 1.0; a = 1.0 - 2.0 * ts * ts; at = -4.0 * ts / Tr; att = -4.0 / T2; }
        return true;
      }
      case 4: { const double s = t / Tr, s2 = s * s; a = s2 * s2; at = 4.0 * s2 * s / Tr; att = 12.0 * s2 / (Tr * Tr); return true; }
      case 41: {
        const double T2 = Tr * Tr, rap = 8.0 * t / Tr, rap2 = t * t / T2, rap22 = 3.0 * rap2;
        a = (rap22 - rap + 6.0) * rap2;
        at = (4.0 * rap22 - 3.0 * rap + 12.0) * t / T2;
        att = (12.0 * rap22 - 6.0 * rap + 12.0) / T2;
        return true;
      }
      default: return false;
    }
  }
  if (t >= Ts + Tr) { a = at = att = 0.0; return true; }
  if (t >= Ts) {
    const double ti = Ts + Tr - t;
    switch (iramp) {
      case 1: a = ti / Tr; at = -1.0 / Tr; att = 0.0; return true;
      case 2: { const double s = ti / Tr; a = s * s; at = -2.0 * s / Tr; att = 2.0 / (Tr * Tr); return true; }
      case 21: {
        const double T2 = Tr * Tr;
        if (ti <= Tr / 2.0) { a = 2.0 * ti * ti / T2; at = -4.0 * ti / T2; att = 4.0 / T2; }
        else { const double ts = ti / Tr - 1.0; a = 1.0 - 2.0 * ts * ts; at = 4.0 * ts / Tr; att = -4.0 / T2; }
        return true;
      }
      case 4: { const double s = ti / Tr, s2 = s * s; a = s2 * s2; at = -4.0 * s2 * s / Tr; att = 12.0 * s2 / (Tr * Tr); return true; }
      case 41: {
        const double T2 = Tr * Tr, rap = 8.0 * ti / Tr, rap2 = ti * ti / T2, rap22 = 3.0 * rap2;
        a = (rap22 - rap + 6.0) * rap2;
        at = -(4.0 * rap22 - 3.0 * rap + 12.0) * ti / T2;
        att = (12.0 * rap22 - 6.0 * rap + 12.0) / T2;
        return true;
      }
      default: a = 1.0; at = 0.0; att = 0.0; return true;
    }
  }
  a = 1.0; at = 0.0; att = 0.0;
  return true;
}

// Device step scalars of the graph-captured attempt (hos_advance_graph): the
// controller's state and, per RK stage s, the stage time, the integrating-factor
// shift and the ramp and paddle values, so that a captured attempt reads its
// time-dependent scalars from memory instead of kernel arguments.
enum DevScal {
  DS_T = 0, DS_H, DS_HL, DS_TNEXT, DS_TOLER, DS_STAB, DS_DOUT, DS_NACC, DS_NREJ, DS_ACTIVE, DS_ACCEPT,
  DS_ENDRAMP, DS_ERR, DS_GALPHA, DS_GPOS, DS_GALPHAT, DS_SUB, DS_STAGE0 = 24, DS_COUNT = DS_STAGE0 + 40
};
enum StageScal { SS_T = 0, SS_DT, SS_A, SS_AT, SS_ATT, SS_POS, SS_STRIDE = 8 };
__device__ __forceinline__ double stage_val(const double* dsc, int s, int k) { return dsc[DS_STAGE0 + SS_STRIDE * s + k]; }

// GtoF + division by omega at the start of solveHOS2 (resol_HOS.f90:344-351);
// also partial sums of eta_hat for the mean-volume fix (resol_HOS.f90:353-364).
template <typename T>
__global__ void k_gtof_in(int n1, int n2, double pix, double piy, double dt, const double* __restrict__ G1,
                          const double* __restrict__ G2, T* eta_hat, T* phis_hat, double* partial,
                          double* phis64, const double* dsc, int stage) {
  if (dsc) dt = stage_val(dsc, stage, SS_DT);
  const size_t n = (size_t)n1 * n2;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < n) {
    const int i1 = (int)(i % n1), i2 = (int)(i / n1);
    const double w = wave_omega(wave_k(i1, i2, pix, piy));
    const double c = cos(-w * dt), sn = sin(-w * dt);
    const double F1 = G1[i] * c - G2[i] * sn;
    const double F2 = G2[i] * c + G1[i] * sn;
    eta_hat[i] = (T)F1;
    const double ph = (i == 0) ? F2 : F2 / w;
    phis_hat[i] = (T)ph;
    if (phis64) phis64[i] = ph;
    s = F1;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Mean-volume mode of eta from the paddle-corner value (resol_HOS.f90:358-364,
// runge_kutta.f90:306-312): a single block sums the partials.
template <typename T>
__global__ void k_volfix(int npart, const double* partial, T* eta_hat, double* eta_hat_d, int mode, double alpha,
                         double pos_top, double xlen, double d_hinge, const double* dsc) {
  if (dsc) {
    alpha = dsc[DS_GALPHA];
    pos_top = dsc[DS_GPOS];
  }
  double s = 0.0;
  for (int i = threadIdx.x; i < npart; i += blockDim.x) s += partial[i];
  s = block_sum(s);
  if (threadIdx.x == 0) {
    const double x = alpha * pos_top / xlen;
    double v = 0.0;
    if (mode == 1)
      v = x + s * x;
    else if (mode == 2)
      v = 0.5 * alpha * (1 - d_hinge) * pos_top / xlen + s * x;
    if (eta_hat) eta_hat[0] = (T)v;
    if (eta_hat_d) eta_hat_d[0] = v;
  }
}

template <typename T>
struct PtrSet {
  T* p[10];
};

// One level iH of the HOS recursion from its inverse transforms v_j (item q of
// lv holds j = iH - q), accumulated in the reference's order j = iH..1:
// phim = 0 - sum v_j eta^j/j!, phizm(iH) = 0 + sum v_j eta^(j-1)/(j-1)!
// (resol_HOS.f90:736-772).  The level's transforms then share one launch
// (measured on 1024 x 256, M = 5: 92 -> 102 FP64 and 130 -> 139 mixed steps/s
// against accumulating in each transform's x-pass epilogue, one launch each).
// The powers eta^j/j! (j < M; eta^M only feeds the dead last-order phimext) are
// formed in registers from eta as (eta^j/j!)/(j+1)*eta (resol_HOS.f90:974-982),
// ascending; the loops run
// over a fixed bound so that every index is known at compile time.
constexpr int kMaxOrder = 10;
template <typename T>
__device__ __forceinline__ void hos_level_point(size_t i, int iH, int M, const PtrSet<T>& lv, const T* __restrict__ eta,
                                                T* phim, T* phizm) {
  const bool do_phim = iH < M;
  const int top = min(iH, M - 1);  // highest power read at this level
  T pw[kMaxOrder];
#pragma unroll
  for (int j = 0; j < kMaxOrder; ++j) pw[j] = (T)0;
  pw[0] = (T)1;
  if (top >= 1) pw[1] = eta[i];
#pragma unroll
  for (int j = 1; j < kMaxOrder - 1; ++j)
    if (j + 1 <= top) pw[j + 1] = pw[j] / (T)(j + 1) * pw[1];
  T pm = (T)0, pz = (T)0;
#pragma unroll
  for (int j = kMaxOrder - 1; j >= 1; --j) {  // j = iH..1: item iH - j of lv
    if (j <= iH) {
      const T v = lv.p[iH - j][i];
      const T e = do_phim ? pw[j] : (T)0;
      const T e1 = (j >= 2) ? pw[j - 1] : (T)1;
      pm = pm - v * e;
      pz = pz + v * e1;
    }
  }
  if (do_phim) phim[i] = pm;
  phizm[i] = pz;
}
template <typename T>
__global__ void k_hos_level(size_t n, int iH, int M, PtrSet<T> lv, const T* __restrict__ eta, T* phim, T* phizm) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  hos_level_point(i, iH, M, lv, eta, phim, phizm);
}
// The same with 16 bytes of each field per thread (4 FP32 or 2 FP64 points, the
// loads issued together); every field 16-byte aligned (checked by the caller), the
// last partial group point by point.
template <typename T>
__global__ void k_hos_level_vec(size_t n, int iH, int M, PtrSet<T> lv, const T* __restrict__ eta, T* phim, T* phizm) {
  constexpr int W = 16 / sizeof(T);
  typedef T V __attribute__((ext_vector_type(W)));
  const size_t i0 = (blockIdx.x * (size_t)blockDim.x + threadIdx.x) * W;
  if (i0 >= n) return;
  if (i0 + W > n) {
    for (size_t i = i0; i < n; ++i) hos_level_point(i, iH, M, lv, eta, phim, phizm);
    return;
  }
  const bool do_phim = iH < M;
  const int top = min(iH, M - 1);
  // (eta is read only where a power of it enters: with M < 2 it is not allocated)
  V ev = {};
  if (top >= 1) ev = *reinterpret_cast<const V*>(eta + i0);
  V vv[kMaxOrder];
#pragma unroll
  for (int j = 1; j < kMaxOrder; ++j)
    if (j <= iH) vv[j] = *reinterpret_cast<const V*>(lv.p[iH - j] + i0);
  V pmv, pzv;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    T pw[kMaxOrder];
#pragma unroll
    for (int j = 0; j < kMaxOrder; ++j) pw[j] = (T)0;
    pw[0] = (T)1;
    if (top >= 1) pw[1] = ev[w];
#pragma unroll
    for (int j = 1; j < kMaxOrder - 1; ++j)
      if (j + 1 <= top) pw[j + 1] = pw[j] / (T)(j + 1) * pw[1];
    T pm = (T)0, pz = (T)0;
#pragma unroll
    for (int j = kMaxOrder - 1; j >= 1; --j) {
      if (j <= iH) {
        const T v = vv[j][w];
        const T e = do_phim ? pw[j] : (T)0;
        const T e1 = (j >= 2) ? pw[j - 1] : (T)1;
        pm = pm - v * e;
        pz = pz + v * e1;
      }
    }
    pmv[w] = pm;
    pzv[w] = pz;
  }
  if (do_phim) *reinterpret_cast<V*>(phim + i0) = pmv;
  *reinterpret_cast<V*>(phizm + i0) = pzv;
}

// out = sum_{j0 <= j <= M} z_j (zero when j0 > M)
template <typename T>
__global__ void k_sum_orders(size_t n, int M, int j0, PtrSet<T> z, T* out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  T s = (j0 <= M) ? z.p[j0][i] : (T)0;
  for (int j = j0 + 1; j <= M; ++j) s += z.p[j][i];
  out[i] = s;
}

template <typename T>
__global__ void k_narrow(size_t n, const double* __restrict__ in, T* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = (T)in[i];
}

// Free-surface products on the extended grid (resol_HOS.f90:376-392, :534-555 and
// the phiz2 / geta2phiz / geta2phiz2 assembly of :840-905):
//   X = -grad(phis).grad(eta) + geta2phiz - grad(eta).grad(phi_add)
//   Y = 1/2 (-|grad phis|^2 + geta2phiz2 + phiz2)
//       - (grad(phis).grad(phi_add) + 1/2 |grad phi_add|^2)
// With p = M none of the reference's dealias calls inside these sums is active
// before the final filter_ext, so all terms are pointwise.
// The products of a sine-type slope with a cosine-type paddle-potential
// gradient (eta_x phi_add,x, phis_x phi_add,x and their y twins) are not
// band-limited in the cosine basis, so their filtered value depends on the
// extended grid; they are returned separately (XA, YA: sine in x; XB, YB:
// sine in y) and converted exactly to the reference grid afterwards.
template <typename T>
__global__ void k_fsbc(size_t n, int M, PtrSet<T> zz, const T* __restrict__ px, const T* __restrict__ py,
                       const T* __restrict__ ex, const T* __restrict__ ey,
                       const T* __restrict__ ax, const T* __restrict__ ay, const T* __restrict__ az, T* X, T* Y,
                       T* XA, T* YA, T* XB, T* YB, int j0, T* zsum) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  T z[10];
  for (int j = 1; j <= M; ++j) z[j] = zz.p[j][i];
  // sum_{j0 <= j <= M} z_j as k_sum_orders forms it
  T zs = (j0 <= M) ? z[j0] : (T)0;
  for (int j = j0 + 1; j <= M; ++j) zs += z[j];
  zsum[i] = zs;
  const T vpx = px[i], vpy = py[i], vex = ex[i], vey = ey[i];
  const T g = vex * vex + vey * vey;  // |grad eta|^2
  T sgz = (T)0;
  for (int j = 1; j <= M - 3; ++j) sgz += z[j];
  if (M >= 3) sgz += z[M - 2 >= 1 ? M - 2 : 1];
  T t1 = (T)0, gg2 = (T)0;
  for (int j = 1; j <= (M - 2) / 2; ++j) {
    const T t = z[j] * z[j];
    t1 += t;
    gg2 += g * t;
  }
  for (int j = 2; j <= M - 3; ++j) {
    const int qm = min(M - 2 - j, j - 1);
    for (int q = 1; q <= qm; ++q) {
      const T t = (T)2 * z[j] * z[q];
      t1 += t;
      gg2 += g * t;
    }
  }
  if (M >= 2) t1 += z[M / 2] * z[M / 2];
  for (int j = 2; j <= M - 1; ++j) {
    const int q0 = max(0, min(M - 2 - j, j - 1)) + 1, q1 = min(M - j, j - 1);
    for (int q = q0; q <= q1; ++q) t1 += (T)2 * z[j] * z[q];
  }
  const T vax = ax[i], vay = ay[i], vaz = az[i];
  const T deta = -vpx * vex - vpy * vey + g * sgz;
  const T dphis = -vpx * vpx - vpy * vpy + gg2 + t1;
  X[i] = deta;
  Y[i] = (T)0.5 * dphis - (T)0.5 * (vax * vax + vay * vay + vaz * vaz);
  XA[i] = -vex * vax;
  YA[i] = -vpx * vax;
  if (XB) {
    XB[i] = -vey * vay;
    YB[i] = -vpy * vay;
  }
}

// RA = T1 PA in a 2-D tank (n1 x K1 times K1 x 2): a matrix-vector product with two
// right-hand sides.  T1[k][s] vanishes when s + k is even, so only the other half is
// stored: T1h[k + n1 j] = T1[k][2j + (k even ? 1 : 0)], j < K1/2.  The product is
// split over K in chunks of kc (even) so that the matrix streams through many
// workgroups; chunk s writes its sums to part[s][c][m] and k_add_mixed adds the
// chunks in order.
constexpr int kT1Parts = 64;     // chunks at most
constexpr int kT1MaxChunk = 256;  // k values per chunk at most
template <typename T>
__global__ void k_t1_partial(int n1, int K1, int kc, const T* __restrict__ T1h, const T* __restrict__ PA,
                             T* part) {
  __shared__ T pa[2][kT1MaxChunk];
  const int s = blockIdx.y, k0 = s * kc, kn = min(kc, K1 - k0);
  for (int j = threadIdx.x; j < 2 * kc; j += blockDim.x) {
    const int cc = j / kc, kk = j % kc;
    pa[cc][kk] = (kk < kn) ? PA[k0 + kk + (size_t)K1 * cc] : (T)0;
  }
  __syncthreads();
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n1) return;
  const int par = (m & 1) ? 0 : 1;  // parity of the nonzero columns of row m
  const T* col = T1h + m + (size_t)n1 * (k0 / 2);
  T a0 = (T)0, a1 = (T)0;
#pragma unroll 8
  for (int jj = 0; jj < kn / 2; ++jj) {
    const T t = col[(size_t)n1 * jj];
    a0 += t * pa[0][2 * jj + par];
    a1 += t * pa[1][2 * jj + par];
  }
  part[(size_t)(2 * s) * n1 + m] = a0;
  part[(size_t)(2 * s + 1) * n1 + m] = a1;
}

// D1 += RA_X + RB_X, D2 += RA_Y + RB_Y (exact reference-grid filter of the
// mixed-parity products), and Q = D1 + P.  With nparts > 0 (2-D) RA comes as the
// chunk sums of k_t1_partial, added here in chunk order.
template <typename T>
__global__ void k_add_mixed(int n1, int n2, const T* RA, const T* RB, T* D1, T* D2, const T* P, T* Q,
                            int nparts, int psplit) {
  const size_t n = (size_t)n1 * n2;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int i1 = (int)(i % n1), i2 = (int)(i / n1);
  T a, b;
  if (nparts > 0) {
    // the chunks' partial sums in chunk order; their loads are issued eight
    // chunks at a time (the sums keep the order)
    a = RA[i];
    b = RA[n1 + i];
    int s = 1;
    for (; s + 8 <= nparts; s += 8) {
      T pa[8], pb[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        pa[q] = RA[(size_t)(2 * (s + q)) * n1 + i];
        pb[q] = RA[(size_t)(2 * (s + q) + 1) * n1 + i];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        a += pa[q];
        b += pb[q];
      }
    }
    for (; s < nparts; ++s) {
      a += RA[(size_t)(2 * s) * n1 + i];
      b += RA[(size_t)(2 * s + 1) * n1 + i];
    }
  } else if (psplit) {
    // rows of RA in parity order: even i1 first
    const size_t r = ((i1 & 1) ? (n1 >> 1) + (i1 >> 1) : (i1 >> 1)) + (size_t)n1 * i2;
    a = RA[r];
    b = RA[n + r];
  } else {
    a = RA[i];
    b = RA[n + i];
  }
  if (RB) {
    // with psplit the columns of RB in parity order: even i2 first
    const int c2 = psplit ? ((i2 & 1) ? (n2 >> 1) + (i2 >> 1) : (i2 >> 1)) : i2;
    a += RB[i1 + (size_t)2 * n1 * c2];
    b += RB[n1 + i1 + (size_t)2 * n1 * c2];
  }
  const T d1 = D1[i] + a;
  D1[i] = d1;
  D2[i] += b;
  Q[i] = d1 + P[i];
}

template <typename T>
__global__ void k_add3(size_t n, const T* a, const T* b, T* out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + b[i];
}

// dphin (physical) = F2S(D1+P) + phiz_add ; Z = -phit_add - nu dphin
template <typename T>
__global__ void k_dphin(size_t n, const T* __restrict__ fq, const double* __restrict__ phiz_add,
                        const double* __restrict__ phit_add, const double* __restrict__ nu, T* Z) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dphin = (double)fq[i] + phiz_add[i];
  Z[i] = (T)(-phit_add[i] - (nu ? nu[i] : 0.0) * dphin);
}

// Mixed path, the same Z in two parts: the first-order part
// -phit_add - nu (phiz_add + wlin), wlin = F2S(kth1 phis_hat) (FP64), and the
// beach term of the orders >= 2, -nu F2S(D1 + P) (type T, P without order 1).
template <typename T>
__global__ void k_beach(size_t n, const T* __restrict__ fq, const double* __restrict__ nu, T* Z) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) Z[i] = (T)(-(nu ? nu[i] : 0.0) * (double)fq[i]);
}
__global__ void k_forcing(size_t n, const double* __restrict__ phiz_add, const double* __restrict__ phit_add,
                          const double* __restrict__ wlin, const double* __restrict__ nu, double* Z) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) Z[i] = -phit_add[i] - (nu ? nu[i] : 0.0) * (phiz_add[i] + wlin[i]);
}

// NL1 = D1 + P - phis_hat kth1 + S2F(phiz_add) ; NL2 = D2 + S2F(Z) ; expon
// (resol_HOS.f90:623-644, variablechange.f90:37-80); k_G2(1,1) = 0
// (runge_kutta.f90:150,199,205).  phis_hat null: P holds only orders >= 2, whose
// first-order part P(1) = kth1 phis_hat cancels the subtracted term exactly;
// Zh2: optional second part of S2F(Z) (FP64).
// filt: optional mode weights of the free-surface filters (resol_HOS.f90:613-621).
template <typename T, typename TL>
__global__ void k_expon_out(int n1, int n2, double pix, double piy, double dt, const T* D1, const T* P,
                            const T* phis_hat, const TL* phizh, const T* D2, const T* Zh, const double* Zh2,
                            const double* filt, double* dG1, double* dG2, const double* dsc, int stage) {
  if (dsc) dt = stage_val(dsc, stage, SS_DT);
  const size_t n = (size_t)n1 * n2;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int i1 = (int)(i % n1), i2 = (int)(i / n1);
  const double k = wave_k(i1, i2, pix, piy);
  const double w = wave_omega(k);
  const double kth1 = k * tanh(k);
  const double f = filt ? filt[i] : 1.0;
  const double lin = phis_hat ? (double)phis_hat[i] * kth1 : 0.0;
  const double nl1 = ((D1 ? (double)D1[i] : 0.0) + (double)P[i] - lin + (double)phizh[i]) * f;
  double z = (double)Zh[i];
  if (Zh2) z += Zh2[i];
  const double nl2 = ((D2 ? (double)D2[i] : 0.0) + z) * f * w;
  const double c = cos(w * dt), s = sin(w * dt);
  dG1[i] = nl1 * c - nl2 * s;
  dG2[i] = (i == 0) ? 0.0 : nl2 * c + nl1 * s;
}

// ---------------------------------------------------------------------------
// wavemaker kernels (double)
// ---------------------------------------------------------------------------
// sum over harmonics of pos_ini(i2,f) e^{i w t} (i w)^d, d = 0,1,2, times g(z)
// (wavemaking.f90:259-304).  One block per y-node, threads over frequencies.
// pos arrays are (n3, n2).
__global__ void k_pos(int n2, int n3, int nf, double t, const double* __restrict__ freq,
                      const double* __restrict__ pre, const double* __restrict__ pim, const double* __restrict__ gz,
                      double* X, double* Xt, double* Xtt, const double* dsc = nullptr, int stage = 0) {
  if (dsc) t = stage_val(dsc, stage, SS_T);
  const int i2 = blockIdx.x;
  double s0 = 0, s1 = 0, s2 = 0;
  for (int f = threadIdx.x; f < nf; f += blockDim.x) {
    const double w = 2.0 * M_PI * freq[f];
    double sn, cs;
    sincos(w * t, &sn, &cs);
    const double pr = pre[i2 + (size_t)n2 * f], pi = pim[i2 + (size_t)n2 * f];
    const double rr = pr * cs - pi * sn, ri = pr * sn + pi * cs;  // pos e^{iwt}
    s0 += rr;
    s1 += -w * ri;      // Re(i w z) = -w Im z
    s2 += -w * w * rr;  // Re(-w^2 z)
  }
  __shared__ double tot[3];
  s0 = block_sum(s0);
  if (threadIdx.x == 0) tot[0] = s0;
  s1 = block_sum(s1);
  if (threadIdx.x == 0) tot[1] = s1;
  s2 = block_sum(s2);
  if (threadIdx.x == 0) tot[2] = s2;
  __syncthreads();
  for (int i3 = threadIdx.x; i3 < n3; i3 += blockDim.x) {
    const size_t o = i3 + (size_t)n3 * i2;
    X[o] = gz[i3] * tot[0];
    Xt[o] = gz[i3] * tot[1];
    Xtt[o] = gz[i3] * tot[2];
  }
}

// Paddle RHS columns and their C2 continuation (resol_wmkr.f90:148-156 and
// match_surface_wmk :1655-1741): f1 = c1x X + c1t Xt + c1tt Xtt,
// f2 = c2x X + c2t Xt + c2tt Xtt on i3 < n3, continued to n3_add.
struct Cont {
  double m[3][4];  // c_r = m[0][r] f + m[1][r] df + m[2][r] d2f
};
// dsc: the right-hand side's coefficients (alpha'', 2 alpha', alpha; alpha'', alpha', 0) of stage
// `stage` from the device step scalars
__global__ void k_paddle_rhs(int n2, int n3, int n3a, double delz, const double* X, const double* Xt,
                             const double* Xtt, double c1x, double c1t, double c1tt, double c2x, double c2t,
                             double c2tt, Cont cm, double* f1, double* f2, const double* dsc = nullptr,
                             int stage = 0) {
  if (dsc) {
    const double al = stage_val(dsc, stage, SS_A), alt = stage_val(dsc, stage, SS_AT),
                 altt = stage_val(dsc, stage, SS_ATT);
    c1x = altt;
    c1t = 2.0 * alt;
    c1tt = al;
    c2x = altt;
    c2t = alt;
    c2tt = 0.0;
  }
  // one thread per (i3, i2) of the paddle grid (i3 fastest: coalesced); the
  // continuation nodes recompute the column's last four values
  const size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (idx >= (size_t)n3a * n2) return;
  const int i3 = (int)(idx % n3a), i2 = (int)(idx / n3a);
  for (int which = 0; which < 2; ++which) {
    double* f = which ? f2 : f1;
    if (!f) continue;
    const double cx = which ? c2x : c1x, ct = which ? c2t : c1t, ctt = which ? c2tt : c1tt;
    // evaluation order of the reference: - a Xtt - 2 a' Xt - a'' X
    auto val = [&](int q) {
      const size_t o = q + (size_t)n3 * i2;
      return -ctt * Xtt[o] - ct * Xt[o] - cx * X[o];
    };
    if (i3 < n3) {
      f[idx] = val(i3);
      continue;
    }
    const double fm3 = val(n3 - 4), fm2 = val(n3 - 3), fm1 = val(n3 - 2), f0 = val(n3 - 1);
    const double df = (fm2 - 4.0 * fm1 + 3.0 * f0) / (2.0 * delz);
    const double d2f = (-fm3 + 4.0 * fm2 - 5.0 * fm1 + 2.0 * f0) / (delz * delz);
    double c[4];
    for (int r = 0; r < 4; ++r) c[r] = cm.m[0][r] * f0 + cm.m[1][r] * df + cm.m[2][r] * d2f;
    const double xx = (double)i3 / (double)(n3 - 1) - (double)L_ADD;
    const double xx2 = xx * xx;
    f[idx] = (((c[0] * xx2 + c[1]) * xx2 + c[2]) * xx2 + c[3]) * xx * exp(-6.0 * (xx + 1.0));
  }
}

// space_2_Fourier_add('cos','cos') then / k_add tanh(k_add xlen)
// (fourier_FFTW.f90:432-447, resol_wmkr.f90:160-171): the z REDFT01 and y
// REDFT00 are dense products with fixed matrices (gemm.h); this kernel applies
// the normalisation s_2_f_add and the division.
__global__ void k_add_scale(int n2, int n3a, int nfields, double xlen, double piy, const double* wz,
                            const double* in, double* out1, double* out2) {
  const size_t na = (size_t)n3a * n2;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= na * nfields) return;
  const int w = (int)(i / na);
  const size_t r = i - (size_t)w * na;
  const int k = (int)(r % n3a), m = (int)(r / n3a);
  const double kxa = (2 * k + 1) * M_PI / (2.0 * L_ADD);
  const double ky = m * piy;
  const double ka = sqrt(kxa * kxa + ky * ky);
  const double div = ka * tanh(ka * xlen);
  double ys2f = 1.0;
  if (n2 > 1) ys2f = ((m == 0 || m == n2 - 1) ? 1.0 : 2.0) / (2.0 * (n2 - 1));
  const double f = wz ? wz[k] : 1.0;  // paddle z-mode filter (resol_wmkr.f90:158-160)
  (w ? out2 : out1)[r] = in[i] * (ys2f / (double)n3a) * f / div;
}

// csh_add_x and k_add_sh_add_x / k (HOS-NWT.f90:364-405) for one (x, k).
__device__ __forceinline__ void csh_sh(double k, double x, double xlen, double& csh, double& sh) {
  const double k2 = 2.0 * k;
  const double kx3 = k2 * xlen;
  const double e3 = (kx3 <= 700.0) ? exp(-kx3) : 0.0;
  const double kx1 = k * x;
  const double e1 = (kx1 <= 700.0) ? exp(-kx1) / (e3 + 1.0) : 0.0;
  const double kx2 = k2 * (xlen - x);
  const double e2 = (kx2 <= 700.0) ? exp(-kx2) : 0.0;
  const double e12 = (kx1 + kx2 <= 700.0) ? e1 * e2 : 0.0;
  csh = e1 + e12;
  sh = e1 - e12;
}

// Additional potential on the free surface (resol_wmkr.f90:1149-1175) and,
// from the same sums, its first-order part at z = 0 (phiadd_1st_SWEET
// :485-532): with B_f(i1,i3,i2) = sum_j2 A_f(i1,j2,i3) cos|sin(ky(j2) y(i2)),
//   phi  = sum_i3 cos(kx_add (eta+1)) B_phi,  phiz = sum_i3 sin(kx_add (eta+1)) B_z
// and at z = 0 the same sums with cos(kx_add), sin(kx_add).  Workgroup
// (i1, i3-chunk), threads over y-nodes (k_phiadd3d: y-node pairs); A_f of a y-mode chunk is staged in LDS;
// cos(ky y) advances by rotation.  Modes with k_add x > kCut contribute below
// 1e-19 and are skipped.  Partial sums per chunk go to `part`
// [chunk][field][i1 + n1 i2]; k_phiadd_reduce adds the chunks in order.
constexpr int kPhiCh = 256;
constexpr int kPhiMaxPer = 4;
constexpr int kPhiI3 = 8;   // wavemaker modes per workgroup
constexpr int kPhiF = 7;    // phi, phit, phix, phiy, phiz, phit(z=0), phiz(z=0)
__device__ __forceinline__ double kx_add_of(int i3) { return (2 * i3 + 1) * M_PI / (2.0 * L_ADD); }

// Node pairs: cos(pi j (n2-1-i2)/(n2-1)) = (-1)^j cos(pi j i2/(n2-1)) and
// sin(pi j (n2-1-i2)/(n2-1)) = -(-1)^j sin(pi j i2/(n2-1)), so one rotation chain
// over j, summed separately over even and odd j, gives y-node i2 (even + odd) and
// its mirror n2-1-i2 (even - odd; odd - even for the sine sum): half the work of
// one chain per node.  Each thread takes kPhi3Per pairs.
constexpr int kPhi3Per = 2;
template <typename T>
__global__ void __launch_bounds__(256) k_phiadd3d(int n1, int n2, int n3a, double xlen, double piy, double dx,
                                                  const double* __restrict__ a, const double* __restrict__ da,
                                                  const T* __restrict__ eta, double* part) {
  __shared__ double A[5][kPhiCh];
  const int i1 = blockIdx.x, ch = blockIdx.y, tid = threadIdx.x, nt = blockDim.x;
  const double x = i1 * dx;
  const int i3a = ch * kPhiI3, i3b = min(n3a, i3a + kPhiI3);
  if (x * kx_add_of(i3a) > kCut) return;  // the reduction skips this chunk too
  const size_t ns = (size_t)n1 * n2;
  const int np = (n2 + 1) / 2;  // pairs (i2, n2-1-i2); the middle node of an odd n2 is its own mirror
  double out[kPhiF][2][kPhi3Per];
  double ep1[2][kPhi3Per];
#pragma unroll
  for (int q = 0; q < kPhi3Per; ++q) {
    const int pr = tid + q * nt;
    ep1[0][q] = (pr < np) ? (double)eta[i1 + (size_t)n1 * pr] + 1.0 : 1.0;
    ep1[1][q] = (pr < np) ? (double)eta[i1 + (size_t)n1 * (n2 - 1 - pr)] + 1.0 : 1.0;
#pragma unroll
    for (int f = 0; f < kPhiF; ++f) out[f][0][q] = out[f][1][q] = 0.0;
  }
  const double th = 1.0 / (double)(n2 - 1);
  for (int i3 = i3a; i3 < i3b; ++i3) {
    const double kxa = kx_add_of(i3);
    if (x * kxa > kCut) break;
    int jmax = n2;
    if (x > 0.0) {
      const double r = kCut / x;
      jmax = min(n2, (int)floor(sqrt(r * r - kxa * kxa) / piy) + 1);
    }
    double Be[5][kPhi3Per], Bo[5][kPhi3Per];
#pragma unroll
    for (int f = 0; f < 5; ++f)
#pragma unroll
      for (int q = 0; q < kPhi3Per; ++q) Be[f][q] = Bo[f][q] = 0.0;
    for (int j0 = 0; j0 < jmax; j0 += kPhiCh) {  // j0 is even: the parity of j2 is that of jj
      const int jn = min(kPhiCh, jmax - j0);
      for (int jj = tid; jj < jn; jj += nt) {
        const int j2 = j0 + jj;
        const double ky = j2 * piy;
        const double k = sqrt(kxa * kxa + ky * ky);
        double csh, sh;
        csh_sh(k, x, xlen, csh, sh);
        const double av = a[i3 + (size_t)n3a * j2];
        A[0][jj] = av * csh;
        A[1][jj] = da[i3 + (size_t)n3a * j2] * csh;
        A[2][jj] = -av * sh * k;
        A[3][jj] = -av * csh * ky;
        A[4][jj] = -av * csh * kxa;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < kPhi3Per; ++q) {
        const int i2 = tid + q * nt;
        if (i2 >= np) continue;
        // (c, s) = (cos, sin)(pi j2 i2 / (n2-1)), advanced by the rotation (cr, sr)
        const double ang = (double)(((long)j0 * i2) % (2 * (n2 - 1))) * th;
        double c = cospi(ang), sn = sinpi(ang);
        const double cr = cospi(i2 * th), sr = sinpi(i2 * th);
        double e0 = Be[0][q], e1 = Be[1][q], e2 = Be[2][q], e3 = Be[3][q], e4 = Be[4][q];
        double o0 = Bo[0][q], o1 = Bo[1][q], o2 = Bo[2][q], o3 = Bo[3][q], o4 = Bo[4][q];
        int jj = 0;
        for (; jj + 1 < jn; jj += 2) {
          e0 += A[0][jj] * c;
          e1 += A[1][jj] * c;
          e2 += A[2][jj] * c;
          e3 += A[3][jj] * sn;
          e4 += A[4][jj] * c;
          double cn = c * cr - sn * sr;
          sn = sn * cr + c * sr;
          c = cn;
          o0 += A[0][jj + 1] * c;
          o1 += A[1][jj + 1] * c;
          o2 += A[2][jj + 1] * c;
          o3 += A[3][jj + 1] * sn;
          o4 += A[4][jj + 1] * c;
          cn = c * cr - sn * sr;
          sn = sn * cr + c * sr;
          c = cn;
        }
        if (jj < jn) {
          e0 += A[0][jj] * c;
          e1 += A[1][jj] * c;
          e2 += A[2][jj] * c;
          e3 += A[3][jj] * sn;
          e4 += A[4][jj] * c;
        }
        Be[0][q] = e0; Be[1][q] = e1; Be[2][q] = e2; Be[3][q] = e3; Be[4][q] = e4;
        Bo[0][q] = o0; Bo[1][q] = o1; Bo[2][q] = o2; Bo[3][q] = o3; Bo[4][q] = o4;
      }
      __syncthreads();
    }
    double s0, c0;
    sincos(kxa, &s0, &c0);
#pragma unroll
    for (int q = 0; q < kPhi3Per; ++q) {
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const double sg = side ? -1.0 : 1.0;
        double B[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) B[f] = Be[f][q] + sg * Bo[f][q];
        if (side) B[3] = Bo[3][q] - Be[3][q];
        double sn, cs;
        sincos(kxa * ep1[side][q], &sn, &cs);
        out[0][side][q] += cs * B[0];
        out[1][side][q] += cs * B[1];
        out[2][side][q] += cs * B[2];
        out[3][side][q] += cs * B[3];
        out[4][side][q] += sn * B[4];
        out[5][side][q] += c0 * B[1];
        out[6][side][q] += s0 * B[4];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kPhi3Per; ++q) {
    const int i2 = tid + q * nt;
    if (i2 >= np) continue;
    const int m = n2 - 1 - i2;
    const size_t o = i1 + (size_t)n1 * i2, om = i1 + (size_t)n1 * m;
#pragma unroll
    for (int f = 0; f < kPhiF; ++f) {
      part[((size_t)ch * kPhiF + f) * ns + o] = out[f][0][q];
      if (m != i2) part[((size_t)ch * kPhiF + f) * ns + om] = out[f][1][q];
    }
  }
}

// Sum the active chunks of each x-node in order and write the fields.
__global__ void k_phiadd_reduce(int n1, int n2, int nch, double dx, const double* __restrict__ part, double* phi,
                                double* phit, double* phix, double* phiy, double* phiz, double* phit1,
                                double* phiz1) {
  const size_t ns = (size_t)n1 * n2;
  const size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (o >= ns) return;
  const double x = (double)(o % n1) * dx;
  double acc[kPhiF] = {0, 0, 0, 0, 0, 0, 0};
  for (int ch = 0; ch < nch; ++ch) {
    if (x * kx_add_of(ch * kPhiI3) > kCut) break;
    const double* pc = part + (size_t)ch * kPhiF * ns;
    for (int f = 0; f < kPhiF; ++f) acc[f] += pc[f * ns + o];
  }
  phi[o] = acc[0];
  phit[o] = acc[1];
  phix[o] = acc[2];
  phiy[o] = acc[3];
  phiz[o] = acc[4];
  phit1[o] = acc[5];
  phiz1[o] = acc[6];
}

// 2-D tank (n2 = 1): one wave per x-node, its lanes over the wavemaker modes
// (lane, lane + 64, ...), the six sums reduced across the wave in a fixed
// order (bit-reproducible), all seven fields.
template <typename T>
__global__ void __launch_bounds__(256) k_phiadd2d(int n1, int n3a, double xlen, double dx, const double* __restrict__ a,
                                                  const double* __restrict__ da, const T* __restrict__ eta,
                                                  double* o_phi, double* o_phit, double* o_phix, double* o_phiy,
                                                  double* o_phiz, double* o_phit1, double* o_phiz1) {
  const int lane = threadIdx.x & 63;
  const int i1 = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i1 >= n1) return;  // uniform over the wave
  const double x = i1 * dx;
  const double ep1 = (double)eta[i1] + 1.0;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int i3 = lane; i3 < n3a; i3 += 64) {
    const double kxa = kx_add_of(i3);
    if (x * kxa > kCut) break;  // the modes increase with i3
    double csh, sh;
    csh_sh(kxa, x, xlen, csh, sh);
    double sn, cs, s0, c0;
    sincos(kxa * ep1, &sn, &cs);
    sincos(kxa, &s0, &c0);
    const double av = a[i3], dav = da[i3];
    acc[0] += cs * (av * csh);
    acc[1] += cs * (dav * csh);
    acc[2] += cs * (-av * sh * kxa);
    acc[3] += sn * (-av * csh * kxa);
    acc[4] += c0 * (dav * csh);
    acc[5] += s0 * (-av * csh * kxa);
  }
#pragma unroll
  for (int f = 0; f < 6; ++f)
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) acc[f] += __shfl_xor(acc[f], d, 64);
  if (lane == 0) {
    o_phi[i1] = acc[0];
    o_phit[i1] = acc[1];
    o_phix[i1] = acc[2];
    o_phiy[i1] = 0.0;
    o_phiz[i1] = acc[3];
    o_phit1[i1] = acc[4];
    o_phiz1[i1] = acc[5];
  }
}

// ---------------------------------------------------------------------------
// second-order wavemaker (i_wmk = 2) kernels
// ---------------------------------------------------------------------------
// Wavemaker sums at z = 0 (phiadd_1st_SWEET, resol_wmkr.f90:485-532, and
// phiadd_2nd, :836-857), same workgroup layout as k_phiadd3d.  With
// S_f(i1,i3,i2) = sum_j2 A_f(i1,j2,i3) cos|sin(ky(j2) y(i2)) and c0, s0 =
// cos, sin(kx_add(i3)):
//   SET 0 (11 fields, from a, da): phit = c0 S(da csh), phiz = -s0 kx S(a csh),
//     phix = -c0 S(a k sh), phiy = -c0 S_sin(a ky csh), phizz = -c0 kx^2 S(a csh),
//     phixt = -c0 S(da k sh), phiyt = -c0 S_sin(da ky csh), phizt = -s0 kx S(da csh),
//     phixx = c0 S(a k^2 csh), phixz = s0 kx S(a k sh), phiyz = s0 S_sin(a ky csh)
//   SET 1 (2 fields): phit, phiz as above.
// A 2-D tank (n2 = 1) has the single mode j2 = 0.
template <int SET>
struct Z0Set;
template <>
struct Z0Set<0> {
  static constexpr int NB = 7, NF = 11;
};
template <>
struct Z0Set<1> {
  static constexpr int NB = 2, NF = 2;
};

// The y sums run over mirror-node pairs as in k_phiadd3d (even + odd for node i2,
// even - odd for n2-1-i2; odd - even for the sine sums).
template <int SET>
__global__ void __launch_bounds__(256) k_phiadd_z0(int n1, int n2, int n3a, double xlen, double piy, double dx,
                                                   const double* __restrict__ a, const double* __restrict__ da,
                                                   double* part) {
  constexpr int NB = Z0Set<SET>::NB, NF = Z0Set<SET>::NF;
  // the S sums taken with sin(ky y) (the y derivatives)
  constexpr bool kSine[7] = {false, false, false, true, false, true, false};
  __shared__ double A[NB][kPhiCh];
  const int i1 = blockIdx.x, ch = blockIdx.y, tid = threadIdx.x, nt = blockDim.x;
  const double x = i1 * dx;
  const int i3a = ch * kPhiI3, i3b = min(n3a, i3a + kPhiI3);
  if (x * kx_add_of(i3a) > kCut) return;  // the reduction skips this chunk too
  const size_t ns = (size_t)n1 * n2;
  const int np = (n2 + 1) / 2;  // pairs (i2, n2-1-i2); the middle node of an odd n2 is its own mirror
  double out[NF][2][kPhi3Per];
#pragma unroll
  for (int q = 0; q < kPhi3Per; ++q)
#pragma unroll
    for (int f = 0; f < NF; ++f) out[f][0][q] = out[f][1][q] = 0.0;
  const double th = (n2 > 1) ? 1.0 / (double)(n2 - 1) : 0.0;
  for (int i3 = i3a; i3 < i3b; ++i3) {
    const double kxa = kx_add_of(i3);
    if (x * kxa > kCut) break;
    int jmax = n2;
    if (x > 0.0 && n2 > 1) {
      const double r = kCut / x;
      jmax = min(n2, (int)floor(sqrt(r * r - kxa * kxa) / piy) + 1);
    }
    double Se[NB][kPhi3Per], So[NB][kPhi3Per];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int q = 0; q < kPhi3Per; ++q) Se[b][q] = So[b][q] = 0.0;
    for (int j0 = 0; j0 < jmax; j0 += kPhiCh) {  // j0 is even: the parity of j2 is that of jj
      const int jn = min(kPhiCh, jmax - j0);
      for (int jj = tid; jj < jn; jj += nt) {
        const int j2 = j0 + jj;
        const double ky = j2 * piy;
        const double k = sqrt(kxa * kxa + ky * ky);
        double csh, sh;
        csh_sh(k, x, xlen, csh, sh);
        const double av = a[i3 + (size_t)n3a * j2], dv = da[i3 + (size_t)n3a * j2];
        A[0][jj] = av * csh;
        A[1][jj] = dv * csh;
        if constexpr (SET == 0) {
          A[2][jj] = av * sh * k;
          A[3][jj] = av * csh * ky;
          A[4][jj] = dv * sh * k;
          A[5][jj] = dv * csh * ky;
          A[6][jj] = av * csh * k * k;
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < kPhi3Per; ++q) {
        const int i2 = tid + q * nt;
        if (i2 >= np) continue;
        const double ang = (n2 > 1) ? (double)(((long)j0 * i2) % (2 * (n2 - 1))) * th : 0.0;
        double c = cospi(ang), sn = sinpi(ang);
        const double cr = cospi(i2 * th), sr = sinpi(i2 * th);
        double e[NB], o[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          e[b] = Se[b][q];
          o[b] = So[b][q];
        }
        int jj = 0;
        for (; jj + 1 < jn; jj += 2) {
#pragma unroll
          for (int b = 0; b < NB; ++b) e[b] += A[b][jj] * (kSine[b] ? sn : c);
          double cn = c * cr - sn * sr;
          sn = sn * cr + c * sr;
          c = cn;
#pragma unroll
          for (int b = 0; b < NB; ++b) o[b] += A[b][jj + 1] * (kSine[b] ? sn : c);
          cn = c * cr - sn * sr;
          sn = sn * cr + c * sr;
          c = cn;
        }
        if (jj < jn) {
#pragma unroll
          for (int b = 0; b < NB; ++b) e[b] += A[b][jj] * (kSine[b] ? sn : c);
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          Se[b][q] = e[b];
          So[b][q] = o[b];
        }
      }
      __syncthreads();
    }
    double s0, c0;
    sincos(kxa, &s0, &c0);
#pragma unroll
    for (int q = 0; q < kPhi3Per; ++q) {
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        double S[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b)
          S[b] = side == 0 ? Se[b][q] + So[b][q] : (kSine[b] ? So[b][q] - Se[b][q] : Se[b][q] - So[b][q]);
        out[0][side][q] += c0 * S[1];
        out[1][side][q] += -s0 * kxa * S[0];
        if constexpr (SET == 0) {
          out[2][side][q] += -c0 * S[2];
          out[3][side][q] += -c0 * S[3];
          out[4][side][q] += -c0 * kxa * kxa * S[0];
          out[5][side][q] += -c0 * S[4];
          out[6][side][q] += -c0 * S[5];
          out[7][side][q] += -s0 * kxa * S[1];
          out[8][side][q] += c0 * S[6];
          out[9][side][q] += s0 * kxa * S[2];
          out[10][side][q] += s0 * S[3];
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kPhi3Per; ++q) {
    const int i2 = tid + q * nt;
    if (i2 >= np) continue;
    const int m = n2 - 1 - i2;
    const size_t o = i1 + (size_t)n1 * i2, om = i1 + (size_t)n1 * m;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      part[((size_t)ch * NF + f) * ns + o] = out[f][0][q];
      if (m != i2) part[((size_t)ch * NF + f) * ns + om] = out[f][1][q];
    }
  }
}

struct Ptr11 {
  double* p[11];
};

// Sum the active chunks of each x-node in order (nf fields).
__global__ void k_phiadd_reduce_n(int n1, int n2, int nch, int nf, double dx, const double* __restrict__ part,
                                  Ptr11 outp) {
  const size_t ns = (size_t)n1 * n2;
  const size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (o >= ns) return;
  const double x = (double)(o % n1) * dx;
  double acc[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int ch = 0; ch < nch; ++ch) {
    if (x * kx_add_of(ch * kPhiI3) > kCut) break;
    const double* pc = part + (size_t)ch * nf * ns;
    for (int f = 0; f < nf; ++f) acc[f] += pc[f * ns + o];
  }
  for (int f = 0; f < nf; ++f)
    if (outp.p[f]) outp.p[f][o] = acc[f];
}

// y and z derivatives of the paddle displacement and their time derivatives
// (dposdy_1st, dposdz_1st: wavemaking.f90:306-406): sums over harmonics of
// dposdy_ini(i2,f) e^{iwt} g(z), pos_ini(i2,f) e^{iwt} g'(z), times 1 and i w.
__global__ void k_dpos(int n2, int n3, int nf, double t, const double* __restrict__ freq,
                       const double* __restrict__ pre, const double* __restrict__ pim,
                       const double* __restrict__ dyre, const double* __restrict__ dyim,
                       const double* __restrict__ gz, const double* __restrict__ dgz, double* Xy, double* Xyt,
                       double* Xz, double* Xzt) {
  const int i2 = blockIdx.x;
  double y0 = 0, y1 = 0, z0 = 0, z1 = 0;
  for (int f = threadIdx.x; f < nf; f += blockDim.x) {
    const double w = 2.0 * M_PI * freq[f];
    double sn, cs;
    sincos(w * t, &sn, &cs);
    const size_t o = i2 + (size_t)n2 * f;
    const double yr = dyre ? dyre[o] : 0.0, yi = dyim ? dyim[o] : 0.0;
    const double ry = yr * cs - yi * sn, iy = yr * sn + yi * cs;
    const double rz = pre[o] * cs - pim[o] * sn, iz = pre[o] * sn + pim[o] * cs;
    y0 += ry;
    y1 += -w * iy;
    z0 += rz;
    z1 += -w * iz;
  }
  __shared__ double tot[4];
  y0 = block_sum(y0);
  if (threadIdx.x == 0) tot[0] = y0;
  y1 = block_sum(y1);
  if (threadIdx.x == 0) tot[1] = y1;
  z0 = block_sum(z0);
  if (threadIdx.x == 0) tot[2] = z0;
  z1 = block_sum(z1);
  if (threadIdx.x == 0) tot[3] = z1;
  __syncthreads();
  for (int i3 = threadIdx.x; i3 < n3; i3 += blockDim.x) {
    const size_t o = i3 + (size_t)n3 * i2;
    Xy[o] = gz[i3] * tot[0];
    Xyt[o] = gz[i3] * tot[1];
    Xz[o] = dgz[i3] * tot[2];
    Xzt[o] = dgz[i3] * tot[3];
  }
}

// Modal coefficients of the first-order additional-potential derivatives on the
// paddle (phiadd_1st_SWEET :535-547), stacked [xx, xxt, y, yt, z, zt] (n3a x n2
// each): a k_add^2, da k_add^2, -a ky, -da ky, -a kx_add, -da kx_add.
__global__ void k_paddle_prep6(int n2, int n3a, double piy, const double* __restrict__ a,
                               const double* __restrict__ da, double* out) {
  const size_t na = (size_t)n3a * n2;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= na) return;
  const int i3 = (int)(i % n3a), j2 = (int)(i / n3a);
  const double kxa = kx_add_of(i3), ky = j2 * piy;
  const double k2 = kxa * kxa + ky * ky;
  out[i] = a[i] * k2;
  out[na + i] = da[i] * k2;
  out[2 * na + i] = -a[i] * ky;
  out[3 * na + i] = -da[i] * ky;
  out[4 * na + i] = -a[i] * kxa;
  out[5 * na + i] = -da[i] * kxa;
}

// Third-order wavemaker (wmk_3rdorder, resol_wmkr.f90:982-990): modal
// coefficients of the first-order additional potential's third derivatives on
// the paddle, stacked [xxx, xxxt, xy, xyt, xz, xzt] in the kinds of
// k_paddle_prep6: -a k^2 kt, -da k^2 kt, a ky kt, da ky kt, a kx_add kt,
// da kx_add kt with kt = k tanh(k xlen).  The free-surface part is a sine series
// in x and vanishes on the paddle.
__global__ void k_paddle_prep6_x3(int n2, int n3a, double piy, double xlen, const double* __restrict__ a,
                                  const double* __restrict__ da, double* out) {
  const size_t na = (size_t)n3a * n2;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= na) return;
  const int i3 = (int)(i % n3a), j2 = (int)(i / n3a);
  const double kxa = kx_add_of(i3), ky = j2 * piy;
  const double k = sqrt(kxa * kxa + ky * ky);
  const double kt = k * tanh(k * xlen);
  const double k2 = k * k;
  out[i] = -a[i] * k2 * kt;
  out[na + i] = -da[i] * k2 * kt;
  out[2 * na + i] = a[i] * ky * kt;
  out[3 * na + i] = da[i] * ky * kt;
  out[4 * na + i] = a[i] * kxa * kt;
  out[5 * na + i] = da[i] * kxa * kt;
}

// First-order free-surface modes evaluated on the paddle x = 0 at the nodes
// z(i3) (d1st_for_2nd, resol_wmkr.f90:686-700, tables HOS-NWT.f90:342-355),
// added to the stacked z-physical / y-modal fields: one workgroup per (i3, j2),
// block sums over i1 in a fixed order.  cosh(k z')/cosh k and k sinh(k z')/cosh k
// (z' = z + 1) are evaluated in overflow-free exponential form.
__global__ void k_fs_on_paddle(int n1, int n2, int n3, int n3a, double delz, double pix, double piy,
                               const double* __restrict__ a1, const double* __restrict__ da1, double* stacked) {
  const int i3 = blockIdx.x, j2 = blockIdx.y;
  const size_t na = (size_t)n3a * n2;
  const double zp = i3 * delz;  // cx_add (HOS-NWT.f90:242)
  const double ky = j2 * piy;
  double sxx = 0, sxxt = 0, sy = 0, syt = 0, sz = 0, szt = 0;
  for (int i1 = threadIdx.x; i1 < n1; i1 += blockDim.x) {
    const double kx = i1 * pix;
    const double k = sqrt(kx * kx + ky * ky);
    const double e = exp(k * (zp - 1.0)) / (1.0 + exp(-2.0 * k));
    const double e2 = exp(-2.0 * k * zp);
    const double ch = e * (1.0 + e2), sh = k * e * (1.0 - e2);
    const size_t o = i1 + (size_t)n1 * j2;
    const double av = a1[o], dv = da1[o];
    sxx -= av * ch * kx * kx;
    sxxt -= dv * ch * kx * kx;
    sy -= av * ch * ky;
    syt -= dv * ch * ky;
    sz += av * sh;
    szt += dv * sh;
  }
  const double v[6] = {sxx, sxxt, sy, syt, sz, szt};
  __shared__ double tot[6];
  for (int f = 0; f < 6; ++f) {
    const double r = block_sum(v[f]);
    if (threadIdx.x == 0) tot[f] = r;
  }
  __syncthreads();
  if (threadIdx.x < 6) stacked[threadIdx.x * na + i3 + (size_t)n3a * j2] += tot[threadIdx.x];
}

// Modal inputs of the first-order free-surface fields (d1st_for_2nd :711-721):
// [phix (sin-cos), phiy (cos-sin), etax (sin-cos), etay (cos-sin), phizt, phizz
// (cos-cos), phixz (sin-cos), phiyz (cos-sin)].
__global__ void k_fs_modes8(int n1, int n2, double pix, double piy, const double* __restrict__ a1,
                            const double* __restrict__ da1, double* out) {
  const size_t ns = (size_t)n1 * n2;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const int i1 = (int)(i % n1), i2 = (int)(i / n1);
  const double kx = i1 * pix, ky = i2 * piy;
  const double k = sqrt(kx * kx + ky * ky);
  const double kth1 = k * tanh(k);
  const double av = a1[i], dv = da1[i];
  out[i] = -av * kx;
  out[ns + i] = -av * ky;
  out[2 * ns + i] = dv * kx;
  out[3 * ns + i] = dv * ky;
  out[4 * ns + i] = dv * kth1;
  out[5 * ns + i] = av * k * k;
  out[6 * ns + i] = -av * kx * kth1;
  out[7 * ns + i] = -av * ky * kth1;
}

// First-order free-surface fields (d1st_for_2nd :739-777) from the transformed
// modes T (stacked as in k_fs_modes8) and the z = 0 wavemaker fields; written in
// place into T: [phix1st, phiy1st, etax1st, etay1st, phizt1st, phizz1st].
struct Z0Fields {
  const double *phix, *phiy, *phizz, *phixt, *phiyt, *phizt, *phixz, *phiyz;
};
__global__ void k_fs_fields(size_t ns, double* T, Z0Fields z, const double* __restrict__ phiz1st,
                            const double* __restrict__ nu, const double* __restrict__ nux,
                            const double* __restrict__ nuy, double iabs) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const double phix = T[i] + z.phix[i];
  const double phiy = T[ns + i] + z.phiy[i];
  const double phizt = T[4 * ns + i] + z.phizt[i];
  const double phizz = T[5 * ns + i] + z.phizz[i];
  const double phixz = T[6 * ns + i] + z.phixz[i];
  const double phiyz = T[7 * ns + i] + z.phiyz[i];
  const double nv = nu ? nu[i] : 0.0, nx = nux ? nux[i] : 0.0, ny = nuy ? nuy[i] : 0.0;
  const double etax = T[2 * ns + i] - z.phixt[i] - iabs * (nx * phiz1st[i] + nv * phixz);
  const double etay = T[3 * ns + i] - z.phiyt[i] - iabs * (ny * phiz1st[i] + nv * phiyz);
  T[i] = phix;
  T[ns + i] = phiy;
  T[2 * ns + i] = etax;
  T[3 * ns + i] = etay;
  T[4 * ns + i] = phizt;
  T[5 * ns + i] = phizz;
}

// C2 continuation of a paddle column above the paddle (match_surface_wmk,
// resol_wmkr.f90:1655-1741): rows n3 .. n3a-1 from the last four rows.
__device__ __forceinline__ void continue_column(double* col, int n3, int n3a, double delz, const Cont& cm) {
  const double fm3 = col[n3 - 4], fm2 = col[n3 - 3], fm1 = col[n3 - 2], f0 = col[n3 - 1];
  const double df = (fm2 - 4.0 * fm1 + 3.0 * f0) / (2.0 * delz);
  const double d2f = (-fm3 + 4.0 * fm2 - 5.0 * fm1 + 2.0 * f0) / (delz * delz);
  double c[4];
  for (int r = 0; r < 4; ++r) c[r] = cm.m[0][r] * f0 + cm.m[1][r] * df + cm.m[2][r] * d2f;
  for (int i3 = n3; i3 < n3a; ++i3) {
    const double xx = (double)i3 / (double)(n3 - 1) - (double)L_ADD;
    const double xx2 = xx * xx;
    col[i3] = (((c[0] * xx2 + c[1]) * xx2 + c[2]) * xx2 + c[3]) * xx * exp(-6.0 * (xx + 1.0));
  }
}

// Second-order paddle condition (solveHOS_SWEET_2, resol_wmkr.f90:264-277, ifree = 0):
//   d2 = -alpha (-X P_xxt - X_t P_xx + X_y P_yt + X_yt P_y + X_z P_zt + X_zt P_z)
//        - alpha' (-X P_xx + X_y P_y + X_z P_z)
// on i3 < n3, then the continuation.  P = the stacked paddle fields (n3a x n2,
// rows < n3 used), X.. (n3 x n2).
__global__ void k_paddle_bc2(int n2, int n3, int n3a, double delz, double al, double alt, const double* P,
                             const double* X, const double* Xt, const double* Xy, const double* Xyt,
                             const double* Xz, const double* Xzt, Cont cm, double* d2) {
  const int i2 = blockIdx.x * blockDim.x + threadIdx.x;
  if (i2 >= n2) return;
  const size_t na = (size_t)n3a * n2;
  double* col = d2 + (size_t)n3a * i2;
  for (int i3 = 0; i3 < n3; ++i3) {
    const size_t q = i3 + (size_t)n3a * i2, o = i3 + (size_t)n3 * i2;
    const double pxx = P[q], pxxt = P[na + q], py = P[2 * na + q], pyt = P[3 * na + q], pz = P[4 * na + q],
                 pzt = P[5 * na + q];
    const double noramp = -X[o] * pxx + Xy[o] * py + Xz[o] * pz;
    col[i3] = -al * (-X[o] * pxxt - Xt[o] * pxx + Xy[o] * pyt + Xyt[o] * py + Xz[o] * pzt + Xzt[o] * pz) -
              alt * noramp;
  }
  continue_column(col, n3, n3a, delz, cm);
}

// Third-order paddle condition (solveHOS_SWEET_3, resol_wmkr.f90:366-393), summed
// in the reference's order alpha + alphat + alpha^2 + alpha alphat:
//   -alpha (-X_t P2xx - X P2xxt + X_yt P2y + X_y P2yt + X_zt P2z + X_z P2zt)
//   - alpha' (-X P2xx + X_y P2y + X_z P2z)
//   - alpha^2 (-X^2/2 P1xxxt - X_t X P1xxx + X_t X_y P1xy + X X_y P1xyt + X X_yt P1xy
//              + X_t X_z P1xz + X X_z P1xzt + X X_zt P1xz)
//   - alpha alpha' (-X^2 P1xxx + 2 X X_y P1xy + 2 X X_z P1xz)
// P1 = first-order third derivatives [xxx, xxxt, xy, xyt, xz, xzt], P2 = second-order
// fields [xx, xxt, y, yt, z, zt] (n3a x n2 each, rows < n3 used), then the continuation.
__global__ void k_paddle_bc3(int n2, int n3, int n3a, double delz, double al, double alt, const double* P1,
                             const double* P2, const double* X, const double* Xt, const double* Xy,
                             const double* Xyt, const double* Xz, const double* Xzt, Cont cm, double* d3) {
  const int i2 = blockIdx.x * blockDim.x + threadIdx.x;
  if (i2 >= n2) return;
  const size_t na = (size_t)n3a * n2;
  double* col = d3 + (size_t)n3a * i2;
  for (int i3 = 0; i3 < n3; ++i3) {
    const size_t q = i3 + (size_t)n3a * i2, o = i3 + (size_t)n3 * i2;
    const double x = X[o], xt = Xt[o], xy = Xy[o], xyt = Xyt[o], xz = Xz[o], xzt = Xzt[o];
    const double xxx = P1[q], xxxt = P1[na + q], pxy = P1[2 * na + q], pxyt = P1[3 * na + q],
                 pxz = P1[4 * na + q], pxzt = P1[5 * na + q];
    const double xx = P2[q], xxt = P2[na + q], py = P2[2 * na + q], pyt = P2[3 * na + q], pz = P2[4 * na + q],
                 pzt = P2[5 * na + q];
    const double t_al = -al * (-xt * xx - x * xxt + xyt * py + xy * pyt + xzt * pz + xz * pzt);
    const double t_alt = -alt * (-x * xx + xy * py + xz * pz);
    const double t_al2 = -(al * al) * (-(x * x) / 2.0 * xxxt - xt * x * xxx + xt * xy * pxy + x * xy * pxyt +
                                       x * xyt * pxy + xt * xz * pxz + x * xz * pxzt + x * xzt * pxz);
    const double t_alalt = -al * alt * (-(x * x) * xxx + 2.0 * x * xy * pxy + 2.0 * x * xz * pxz);
    col[i3] = t_al + t_alt + t_al2 + t_alalt;
  }
  continue_column(col, n3, n3a, delz, cm);
}

// Second-order free-surface conditions (solveHOS_SWEET_2 :302-318):
//   deta2 = phiz2nd - phix1st etax1st - phiy1st etay1st + eta1st phizz1st
//   tmp   = -eta2nd - phit_add_2 - (phix1st^2 + phiy1st^2 + phiz1st^2)/2
//           - eta1st phizt1st - iabsnb nu deta2
// with phiz2nd = F2S(a2nd kth1) + phiz_add_2; F = [phix1st, phiy1st, etax1st,
// etay1st, phizt1st, phizz1st] (stacked n1 x n2).
__global__ void k_second_fs(size_t ns, const double* __restrict__ w2, const double* __restrict__ phiz2,
                            const double* __restrict__ phit2, const double* __restrict__ F,
                            const double* __restrict__ phiz1st, const double* __restrict__ eta1,
                            const double* __restrict__ eta2, const double* __restrict__ nu, double iabs,
                            double* deta2, double* tmp) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const double phix = F[i], phiy = F[ns + i], etax = F[2 * ns + i], etay = F[3 * ns + i], phizt = F[4 * ns + i],
               phizz = F[5 * ns + i];
  const double d = (w2[i] + phiz2[i]) - phix * etax - phiy * etay + eta1[i] * phizz;
  deta2[i] = d;
  const double pz = phiz1st[i];
  tmp[i] = -eta2[i] - phit2[i] - 0.5 * (phix * phix + phiy * phiy + pz * pz) - eta1[i] * phizt -
           iabs * (nu ? nu[i] : 0.0) * d;
}

// Free-surface elevation at probe positions (output.f90:266-276 with
// mat_coscos, initial_condition.f90:502-510): eta(x, y) = sum_{i1,i2}
// eta_hat(i1,i2) cos(kx(i1) x) cos(ky(i2) y); one workgroup per probe, the
// cosines staged in LDS (n1 + n2 doubles), a fixed-order block reduction.
__global__ void __launch_bounds__(256) k_probes(int n1, int n2, double pix, double piy, const double* __restrict__ xp,
                                                const double* __restrict__ yp, const double* __restrict__ eta_hat,
                                                double* out) {
  extern __shared__ double cxy[];
  const int j = blockIdx.x;
  const double x = xp[j], y = yp[j];
  for (int i = threadIdx.x; i < n1; i += blockDim.x) cxy[i] = cos((i * pix) * x);
  for (int i = threadIdx.x; i < n2; i += blockDim.x) cxy[n1 + i] = cos((i * piy) * y);
  __syncthreads();
  double acc = 0.0;
  const size_t ns = (size_t)n1 * n2;
  for (size_t o = threadIdx.x; o < ns; o += blockDim.x) {
    const int i1 = (int)(o % n1), i2 = (int)(o / n1);
    acc += eta_hat[o] * cxy[i1] * cxy[n1 + i2];
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) out[j] = acc;
}

// First-order free surface RHS (resol_wmkr.f90:186-193 via phidphiz :583-636):
// deta1 = F2S(a1 kth1) + phiz_add_1 ; tmp = -eta1 - phit_add_1 - iabsnb nu deta1.
__global__ void k_first_order(size_t n, const double* w1, const double* phiz1, const double* phit1,
                              const double* eta1, const double* nu, double iabs, double* deta1, double* tmp) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = w1[i] + phiz1[i];
  deta1[i] = d;
  tmp[i] = -eta1[i] - phit1[i] - iabs * (nu ? nu[i] : 0.0) * d;
}

// ---------------------------------------------------------------------------
// RK kernels (double)
// ---------------------------------------------------------------------------
struct KSet {
  const double* k[5];
  double c[5];
  int nk;
};
// out = base + h sum_s c_s k_s for every state field in one launch (blockIdx.y = field)
constexpr int kMaxAxpy = 10;
struct AxpyBatch {
  int n;
  double* out[kMaxAxpy];
  const double* base[kMaxAxpy];
  size_t len[kMaxAxpy];
  KSet ks[kMaxAxpy];
};
// (at most 9 fields: G1, G2, a1st_add, a1st, eta1st, the three second-order fields, a3rd_add)
static_assert(kMaxAxpy >= 9, "AxpyBatch holds every RK state field");
// Stage scalars of the attempt from the committed time and step proposal in
// dsc (one block): the step hl = min(h, t_next - t), per stage the time
// t + hl c_s (two roundings, as stage_time), the shift hl c_s, the ramp and the
// paddle-top displacement g(0) sum_f Re(pos_ini(1,f) e^{i w t}).
__global__ void __launch_bounds__(256) k_stage_scalars(double* dsc, int iramp, double Tr, double Ts, double T_mono,
                                                       int nf, const double* __restrict__ freq,
                                                       const double* __restrict__ pre0,
                                                       const double* __restrict__ pim0, double gz_top) {
  const double cst[5] = {0.0, 0.5, 0.5, 1.0, 1.0};
  const double t = dsc[DS_T], h = dsc[DS_H], tn = dsc[DS_TNEXT];
  const double hl = (tn - t < h) ? tn - t : h;
  for (int st = 0; st < 5; ++st) {
    const double dt = __dmul_rn(hl, cst[st]);
    const double ts = __dadd_rn(t, dt);
    double sum = 0.0;
    for (int f = threadIdx.x; f < nf; f += blockDim.x) {
      const double w = 2.0 * M_PI * freq[f];
      double sn, cs;
      sincos(w * ts, &sn, &cs);
      sum += pre0[f] * cs - pim0[f] * sn;
    }
    sum = block_sum(sum);
    if (threadIdx.x == 0) {
      double a = 1.0, at = 0.0, att = 0.0;
      ramp_hd(iramp, Tr, Ts, T_mono, ts, a, at, att);
      double* o = dsc + DS_STAGE0 + SS_STRIDE * st;
      o[SS_T] = ts;
      o[SS_DT] = dt;
      o[SS_A] = a;
      o[SS_AT] = at;
      o[SS_ATT] = att;
      o[SS_POS] = gz_top * sum;
    }
  }
  if (threadIdx.x == 0) {
    dsc[DS_HL] = hl;
    dsc[DS_ACTIVE] = (t < tn) ? 1.0 : 0.0;
  }
}

// The accept / reject controller of the inner loop (HOS-NWT.f90:518-563) on the
// device: the embedded error from the attempt's maxima (runge_kutta.f90:272-279),
// the decision, the end-of-ramp flag, the new time and step proposal.
__global__ void k_control(double* dsc, const double* __restrict__ red, int npair, int iramp, double Tramp) {
  const double tiny = 2.220446049250313e-16;
  if (!(dsc[DS_ACTIVE] > 0.0)) {
    dsc[DS_ACCEPT] = 0.0;
    dsc[DS_SUB] = 0.0;
    return;
  }
  double e = 0.0;
  for (int q = 0; q < npair; ++q) e = fmax(e, red[2 * q] / (red[2 * q + 1] + tiny));
  const double toler = dsc[DS_TOLER], stab = dsc[DS_STAB], dout = dsc[DS_DOUT];
  double t = dsc[DS_T], h = dsc[DS_H];
  const double hl = dsc[DS_HL];
  double corr;
  int endramp = (int)dsc[DS_ENDRAMP];
  if (e > toler && h > 0.001 * 0.95 * stab) {
    corr = 0.9 * pow(e / toler, -(1.0 / 4.0));
    dsc[DS_NREJ] += 1.0;
    if (fabs(Tramp - (t + hl)) <= tiny && Tramp > t) endramp = 0;
    dsc[DS_ACCEPT] = 0.0;
    dsc[DS_SUB] = 0.0;
  } else {
    const bool sub = iramp == 1 && Tramp - (t + hl) <= tiny && Tramp > t && endramp != 1;
    if (sub) endramp = 1;
    dsc[DS_SUB] = sub ? 1.0 : 0.0;
    dsc[DS_ACCEPT] = 1.0;
    t = t + hl;
    corr = (fabs(e) < tiny) ? 2.0 : 0.9 * pow(e / toler, -(1.0 / 4.0));
    dsc[DS_NACC] += 1.0;
  }
  h = h * corr;
  h = fmin(fmin(h, 0.95 * stab), dout);
  h = fmax(fmax(h, 0.0001 * 0.95 * stab), 0.0001 * dout);
  dsc[DS_T] = t;
  dsc[DS_H] = h;
  dsc[DS_ERR] = e;
  dsc[DS_ENDRAMP] = (double)endramp;
}

// the accept branch's copy-back (HOS-NWT.f90:534-543) of every state field, when
// the controller accepted (blockIdx.y = field)
constexpr int kMaxCopy = 12;
struct CopyBatch {
  int n;
  double* dst[kMaxCopy];
  const double* src[kMaxCopy];
  size_t len[kMaxCopy];
};
__global__ void k_commit_if(const double* dsc, CopyBatch b) {
  if (!(dsc[DS_ACCEPT] > 0.0)) return;
  const int f = blockIdx.y;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < b.len[f]) b.dst[f][i] = b.src[f][i];
}
// the end-of-ramp correction a1st_add -= a_add_ramp (HOS-NWT.f90:546-550) when flagged
__global__ void k_sub_if(size_t n, double* a, const double* b, const double* dsc) {
  if (!(dsc[DS_SUB] > 0.0)) return;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] -= b[i];
}

// wavemaker globals of the evaluation at stage `stage` (alpha, pos1st(n3,1), alpha')
__global__ void k_wmk_glob(double* dsc, int stage) {
  dsc[DS_GALPHA] = stage_val(dsc, stage, SS_A);
  dsc[DS_GPOS] = stage_val(dsc, stage, SS_POS);
  dsc[DS_GALPHAT] = stage_val(dsc, stage, SS_AT);
}
__global__ void k_axpyk_batch(double h, AxpyBatch b, const double* dsc = nullptr) {
  if (dsc) h = dsc[DS_HL];
  const int f = blockIdx.y;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= b.len[f]) return;
  const KSet& ks = b.ks[f];
  double acc = 0.0;
  for (int s = 0; s < ks.nk; ++s) acc = acc + ks.k[s][i] * ks.c[s];
  b.out[f][i] = b.base[f][i] + h * acc;
}
__global__ void k_reduce_max(int n, const double* p, double* out) {
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, p[i]);
  m = block_max(m);
  if (threadIdx.x == 0) *out = m;
}
// X_new = X + h sum b k ; err = |(sum e k - sum b k) h| ; block maxima of err and
// |X_new|, for every error-carrying field in one launch (blockIdx.y = field f);
// field f's block maxima go to pmax_err/pmax_x + f * stride.
constexpr int kMaxFinal = 5;
struct FinalBatch {
  int n;
  size_t stride;
  double* out[kMaxFinal];
  const double* base[kMaxFinal];
  size_t len[kMaxFinal];
  KSet kb[kMaxFinal], ke[kMaxFinal];
};
// (G1, G2 and the first-, second- and third-order paddle modes)
static_assert(kMaxFinal >= 5, "FinalBatch holds every error-carrying field");
__global__ void k_final_err_batch(double h, FinalBatch b, double* pmax_err, double* pmax_x,
                                  const double* dsc = nullptr) {
  if (dsc) h = dsc[DS_HL];
  const int f = blockIdx.y;
  const size_t n = b.len[f];
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  // blocks past this field's length write nothing: every thread of a block takes the same branch
  if (blockIdx.x * (size_t)blockDim.x >= n) return;
  double me = 0.0, mx = 0.0;
  if (i < n) {
    const KSet& kb = b.kb[f];
    const KSet& ke = b.ke[f];
    double db = 0.0, de = 0.0;
    for (int s = 0; s < kb.nk; ++s) db = db + kb.k[s][i] * kb.c[s];
    for (int s = 0; s < ke.nk; ++s) de = de + ke.k[s][i] * ke.c[s];
    const double xn = b.base[f][i] + h * db;
    b.out[f][i] = xn;
    me = fabs((de - db) * h);
    mx = fabs(xn);
  }
  me = block_max(me);
  mx = block_max(mx);
  if (threadIdx.x == 0) {
    pmax_err[f * b.stride + blockIdx.x] = me;
    pmax_x[f * b.stride + blockIdx.x] = mx;
  }
}
// Block r reduces field r / 2's maxima, of the error for even r and of |X_new| for odd r,
// into out[r]; nblocks[f] block maxima per field.
struct FinalCounts {
  int nb[kMaxFinal];
};
__global__ void k_reduce_max_batch(size_t stride, FinalCounts c, const double* pmax_err, const double* pmax_x,
                                   double* out) {
  const int r = blockIdx.x, f = r >> 1;
  const double* p = ((r & 1) ? pmax_x : pmax_err) + f * stride;
  double m = 0.0;
  for (int i = threadIdx.x; i < c.nb[f]; i += blockDim.x) m = fmax(m, p[i]);
  m = block_max(m);
  if (threadIdx.x == 0) out[r] = m;
}

// End of runge4_adapt2 (runge_kutta.f90:281-299): GtoF, /omega, phis(1,1)
// restored, deta = GtoF(k_s) + phis omega^2; partial sums for the volume fix.
__global__ void k_step_end(int n1, int n2, double pix, double piy, double h, const double* G1, const double* G2,
                           const double* kG1, const double* kG2, const double* phis_hat_cur, double* eta_hat,
                           double* phis_hat, double* deta_hat, double* partial, const double* dsc = nullptr) {
  if (dsc) h = dsc[DS_HL];
  const size_t n = (size_t)n1 * n2;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < n) {
    const double phis0 = phis_hat_cur[0];
    const int i1 = (int)(i % n1), i2 = (int)(i / n1);
    const double w = wave_omega(wave_k(i1, i2, pix, piy));
    const double c = cos(-w * h), sn = sin(-w * h);
    const double F1 = G1[i] * c - G2[i] * sn;
    const double F2 = G2[i] * c + G1[i] * sn;
    const double D1 = kG1[i] * c - kG2[i] * sn;
    const double ph = (i == 0) ? phis0 : F2 / w;
    eta_hat[i] = F1;
    phis_hat[i] = ph;
    deta_hat[i] = (i == 0) ? D1 : D1 + ph * w * w;
    s = F1;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// G1 = eta_hat, G2 = phis_hat * omega at the start of runge4_adapt2
// (runge_kutta.f90:128-141); G2(1,1) = 0 since omega(1,1) = 0.
__global__ void k_step_begin(int n1, int n2, double pix, double piy, const double* eta_hat,
                             const double* phis_hat, double* G1, double* G2) {
  const size_t n = (size_t)n1 * n2;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int i1 = (int)(i % n1), i2 = (int)(i / n1);
  const double w = wave_omega(wave_k(i1, i2, pix, piy));
  G1[i] = eta_hat[i];
  G2[i] = phis_hat[i] * w;
}

// a = a1 + a2 (+ a3), d = d1 + d2 (+ d3): total additional-potential modes
// (resol_HOS.f90:498-499); a3, d3 null below a third-order wavemaker
__global__ void k_sum_orders3(size_t n, const double* a1, const double* a2, const double* a3, const double* d1,
                              const double* d2, const double* d3, double* a, double* d) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double av = a1[i] + a2[i], dv = d1[i] + d2[i];
  if (a3) av = av + a3[i];
  if (d3) dv = dv + d3[i];
  a[i] = av;
  d[i] = dv;
}

__global__ void k_mulw(size_t n, const double* __restrict__ w, double* __restrict__ x) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] *= w[i];
}

__global__ void k_sub(size_t n, double* a, const double* b) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] -= b[i];
}

template <typename A, typename B>
__global__ void k_convert(size_t n, const A* in, B* out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = (B)in[i];
}

__global__ void k_absmax2(size_t n, const void* a, const void* b, int is_f32, double* pa, double* pb) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  double ma = 0.0, mb = 0.0;
  if (i < n) {
    if (is_f32) {
      ma = fabs((double)((const float*)a)[i]);
      mb = fabs((double)((const float*)b)[i]);
    } else {
      ma = fabs(((const double*)a)[i]);
      mb = fabs(((const double*)b)[i]);
    }
  }
  ma = block_max(ma);
  mb = block_max(mb);
  if (threadIdx.x == 0) {
    pa[blockIdx.x] = ma;
    pb[blockIdx.x] = mb;
  }
}


// ---------------------------------------------------------------------------
// per-order spectral band energies (features of the learned order selection)
// ---------------------------------------------------------------------------
// Record of one right-hand side: rec[b] = sum of eta_hat^2 over the modes of
// band b; rec[64 j + b] = sum of (k tanh k a_j)^2, a_j the modal potential of
// order j at z = 0 (order 1: phis_hat; order j >= 2: aHOS(j) of HOSphis_modes,
// resol_HOS.f90:720-782), i.e. the band energy of the linear part of W^(j);
// rec[kFeatM] = the order of the evaluation, rec[kFeatEta] = sum of eta_hat^2.
// Band b holds the modes with b <= 64 |k| / |k|max < b + 1 (the last band
// closed), |k|max = |k(n1-1, n2-1)|.
constexpr int kFeatBands = 64;
constexpr int kFeatOrders = 8;
constexpr int kFeatM = (kFeatOrders + 1) * kFeatBands;  // 576
constexpr int kFeatEta = kFeatM + 1;
constexpr int kFeatRec = kFeatM + 2;
constexpr int kFeatBlockModes = 512;
constexpr int kFeatGrid = 256;  // blocks of the band-energy kernel (grid-stride)

template <typename T>
struct OrderFields {
  const T* a[kFeatOrders + 1];  // a[j], j = 1..nord
  int ld[kFeatOrders + 1];
};

// Deterministic: each block sums its modes band by band in a fixed order, and
// the blocks' partials are added in block order (k_band_reduce).
template <typename T>
__global__ void __launch_bounds__(256) k_band_energy(int n1, int n2, double pix, double piy, int nord,
                                                     OrderFields<T> f, const T* __restrict__ eta_hat,
                                                     double* __restrict__ part) {
  __shared__ int band[kFeatBlockModes];
  __shared__ double val[kFeatOrders + 1][kFeatBlockModes];
  const double kx1 = (n1 - 1) * pix, ky1 = (n2 - 1) * piy;
  const double kmax = sqrt(kx1 * kx1 + ky1 * ky1);
  const size_t ns = (size_t)n1 * n2;
  const int nq = (nord + 1) * kFeatBands;
  double acc[3] = {0.0, 0.0, 0.0};  // q = threadIdx.x + 256 r (nq <= 576)
  double eta_tot = 0.0;
  for (size_t base = (size_t)blockIdx.x * kFeatBlockModes; base < ns; base += (size_t)gridDim.x * kFeatBlockModes) {
    for (int e = threadIdx.x; e < kFeatBlockModes; e += blockDim.x) {
      const size_t m = base + e;
      int b = -1;
      if (m < ns) {
        const int i1 = (int)(m % n1), i2 = (int)(m / n1);
        const double k = wave_k(i1, i2, pix, piy);
        const double kth = k * tanh(k);
        b = kmax > 0.0 ? (int)(kFeatBands * k / kmax) : 0;
        b = b < kFeatBands ? b : kFeatBands - 1;
        const double ev = (double)eta_hat[m];
        val[0][e] = ev * ev;
        for (int j = 1; j <= nord; ++j) {
          const double w = kth * (double)f.a[j][i1 + (size_t)f.ld[j] * i2];
          val[j][e] = w * w;
        }
      }
      band[e] = b;
    }
    __syncthreads();
    for (int r = 0; r < 3; ++r) {
      const int q = threadIdx.x + 256 * r;
      if (q >= nq) break;
      const int j = q / kFeatBands, b = q % kFeatBands;
      double sum = 0.0;
      for (int e = 0; e < kFeatBlockModes; ++e)
        if (band[e] == b) sum += val[j][e];
      acc[r] += sum;
    }
    if (threadIdx.x == 0)
      for (int e = 0; e < kFeatBlockModes; ++e)
        if (band[e] >= 0) eta_tot += val[0][e];
    __syncthreads();
  }
  double* out = part + (size_t)blockIdx.x * kFeatRec;
  for (int r = 0; r < 3; ++r) {
    const int q = threadIdx.x + 256 * r;
    if (q < kFeatM) out[q] = (q < nq) ? acc[r] : 0.0;
  }
  if (threadIdx.x == 0) {
    out[kFeatM] = 0.0;
    out[kFeatEta] = eta_tot;
  }
}

__global__ void k_band_reduce(int nblk, int order, const double* __restrict__ part, double* __restrict__ rec) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= kFeatRec) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[(size_t)b * kFeatRec + q];
  rec[q] = (q == kFeatM) ? (double)order : s;
}


// MLP input rows of one feature record (row r predicts order r + 2 from orders
// 1..r+1; rows beyond the record's order, or beyond order 6, are zero):
//   block 0     s(E_eta(b) / E_eta)
//   block j     s(E_j(b) / E_eta), j = 1..r+1   (zero above)
//   block 7     one-hot of r+1 at 448 + r + 1, (log10 E_eta + 10) / 10 at 456
// with s(x) = (log10(max(x, 1e-30)) + 15) / 5; the target of row r is
// s(E_{r+2}(b) / E_eta).  hosnwt/learned.py builds the same rows for training.
constexpr int kRowIn = 512;
constexpr int kPredOrders = 6;  // predictions for orders 2..7
__device__ __forceinline__ double feat_scale(double x) { return (log10(fmax(x, 1e-30)) + 15.0) / 5.0; }
__global__ void k_feat_rows(const double* __restrict__ rec, int m_run, __bf16* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 64 * kRowIn) return;
  const int r = i / kRowIn, c = i % kRowIn;
  const int mp = r + 1;  // orders available to this row
  const int blk = c / kFeatBands, b = c % kFeatBands;
  const double et = rec[kFeatEta];
  double v = 0.0;
  if (mp <= m_run && mp <= kPredOrders) {
    const double inv = et > 0.0 ? 1.0 / et : 0.0;
    if (blk == 0) {
      v = feat_scale(rec[b] * inv);
    } else if (blk <= kPredOrders) {
      v = (blk <= mp) ? feat_scale(rec[blk * kFeatBands + b] * inv) : 0.0;
    } else {
      if (b == mp) v = 1.0;
      if (b == 8) v = (log10(fmax(et, 1e-30)) + 10.0) / 10.0;
    }
  }
  x[i] = (__bf16)(float)v;
}

// sum (a - b)^2 and sum b^2 over n values: block partials, then one block
__global__ void k_diff_norm(size_t n, const double* __restrict__ a, const double* __restrict__ b,
                            double* __restrict__ part) {
  double d = 0.0, s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double x = a[i] - b[i];
    d += x * x;
    s += b[i] * b[i];
  }
  d = block_sum(d);
  s = block_sum(s);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = d;
    part[2 * blockIdx.x + 1] = s;
  }
}
__global__ void k_diff_reduce(int nb, const double* __restrict__ part, double* __restrict__ out) {
  double d = 0.0, s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    d += part[2 * i];
    s += part[2 * i + 1];
  }
  d = block_sum(d);
  s = block_sum(s);
  if (threadIdx.x == 0) {
    out[0] = d;
    out[1] = s;
  }
}

// ---------------------------------------------------------------------------
// dipole wavemaker (igeom = 3): phi_dipoles (resol_wmkr.f90:1492-1647)
// ---------------------------------------------------------------------------
// Components sorted by dipole depth: those of depth z[c] are [start[c], start[c+1]).
// m holds mu_x re, mu_x im, mu_z re, mu_z im, each nf long; w = 2 pi f; cs the cosine
// and sine of w t of the evaluation (k_dip_phase).
constexpr int kDipMaxDepth = 8;
struct DipSet {
  int nf = 0, nz = 0, n_side = 0;
  double x_dip = 0.0, xlen = 0.0, dx = 0.0;
  int start[kDipMaxDepth + 1] = {};
  double z[kDipMaxDepth] = {};
  const double* w = nullptr;
  const double* m = nullptr;
  double* cs = nullptr;
};

// e^{i w t} of every component, as the reference forms EXP(i TWOPI freq time) (:1628)
__global__ void k_dip_phase(int nf, double t, const double* __restrict__ w, double* __restrict__ cs) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  double s, c;
  sincos(w[f] * t, &s, &c);
  cs[f] = c;
  cs[nf + f] = s;
}

// One base-grid node per thread at z = eta.  Each image term is linear in the moments:
// with r2 = 1 / (dx^2 + dz^2), sg the x-moment sign of the image (-1 for odd wall
// images) and up = +1 (dipole level) or -1 (bottom image, z moment mirrored),
//   resu    += sg mu_x r2 dx + up mu_z r2 dz
//   resu_dx += sg mu_x r2 (1 - 2 dx^2 r2) - 2 up mu_z dx dz r2^2
//   resu_dz += -2 sg mu_x dx dz r2^2 + up mu_z r2 (1 - 2 dz^2 r2)
// so the six real image sums of a depth are formed once and every component of that
// depth adds Re(resu e^{i w t}) and the like.  phi_t = alpha' phi + alpha dphi/dt with
// dphi/dt = i w phi; phiy is not formed by the reference (zero); beta = 1, beta_x = 0.
template <typename T>
__global__ void __launch_bounds__(256) k_dipoles(int n1, int ns, DipSet d, double al, double alt,
                                                 const T* __restrict__ eta, double* __restrict__ phi,
                                                 double* __restrict__ phit, double* __restrict__ phix,
                                                 double* __restrict__ phiy, double* __restrict__ phiz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const double x = (double)(i % n1) * d.dx;
  const double e = (double)eta[i];
  const int nf = d.nf;
  double P = 0.0, Pt = 0.0, Px = 0.0, Pz = 0.0;
  for (int c = 0; c < d.nz; ++c) {
    const double zu = e - d.z[c], zd = e - (-2.0 - d.z[c]);
    double A1 = 0.0, A2 = 0.0, B1 = 0.0, B2 = 0.0, C1 = 0.0, C2 = 0.0;
    auto term = [&](double ddx, double sg, double dz, double up) {
      const double r2 = 1.0 / (ddx * ddx + dz * dz);
      A1 += sg * r2 * ddx;
      A2 += up * r2 * dz;
      B1 += sg * r2 * (1.0 - 2.0 * ddx * ddx * r2);
      B2 += -2.0 * ddx * r2 * (up * r2 * dz);
      C1 += -2.0 * dz * r2 * (sg * r2 * ddx);
      C2 += up * r2 * (1.0 - 2.0 * dz * dz * r2);
    };
    term(x - d.x_dip, 1.0, zu, 1.0);
    term(x - d.x_dip, 1.0, zd, -1.0);
    for (int j = 1; j <= d.n_side; ++j) {
      // image abscissae of dipoles_ini (resol_wmkr.f90:1466-1481)
      const double right = (j % 2 == 0) ? d.x_dip + d.xlen * j : -d.x_dip + d.xlen * (j + 1);
      const double left = (j % 2 == 0) ? d.x_dip - d.xlen * j : -d.x_dip - d.xlen * (j - 1);
      const double sg = (j % 2 == 1) ? -1.0 : 1.0;
      term(x - left, sg, zu, 1.0);
      term(x - right, sg, zu, 1.0);
      term(x - right, sg, zd, -1.0);
      term(x - left, sg, zd, -1.0);
    }
    for (int f = d.start[c]; f < d.start[c + 1]; ++f) {
      const double m1r = d.m[f], m1i = d.m[nf + f], m2r = d.m[2 * nf + f], m2i = d.m[3 * nf + f];
      const double co = d.cs[f], si = d.cs[nf + f];
      const double rr = m1r * A1 + m2r * A2, ri = m1i * A1 + m2i * A2;
      P += rr * co - ri * si;
      Pt -= d.w[f] * (rr * si + ri * co);
      const double xr = m1r * B1 + m2r * B2, xi = m1i * B1 + m2i * B2;
      Px += xr * co - xi * si;
      const double zr = m1r * C1 + m2r * C2, zi = m1i * C1 + m2i * C2;
      Pz += zr * co - zi * si;
    }
  }
  const double tp = 2.0 * M_PI;
  P /= tp;
  Pt /= tp;
  Px /= tp;
  Pz /= tp;
  phi[i] = P * al;
  phit[i] = P * alt + Pt * al;
  phix[i] = Px * al;
  phiz[i] = Pz * al;
  phiy[i] = 0.0;
}

// ---------------------------------------------------------------------------
// SWENSE output (velocities.f90:40-784) helpers
// ---------------------------------------------------------------------------
// out = s A B + t C over n values (B, C may be null: B = 1, C = 0); out may alias A, B or C.
__global__ void k_ew(size_t n, double s, const double* A, const double* B, double t, const double* C,
                     double* out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = s * A[i];
  if (B) v *= B[i];
  out[i] = C ? v + t * C[i] : v;
}

// dst (N1 x N2, zero beyond n1 x n2) = src (n1 x n2): extend (variables.f90:243-266)
__global__ void k_extend(int n1, int n2, int N1, int N2, const double* __restrict__ src, double* __restrict__ dst) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)N1 * N2) return;
  const int i1 = (int)(i % N1), i2 = (int)(i / N1);
  dst[i] = (i1 < n1 && i2 < n2) ? src[i1 + (size_t)n1 * i2] : 0.0;
}

// Full time derivatives of init_velocity_comp (velocities.f90:586-611) from a right-hand
// side evaluated with dt = 0 (so GtoF is the identity): deta = dG1 + phis omega^2,
// dphis = dG2 / omega - eta, and at the mean mode dphis = (the free-surface RHS's mean
// mode, dphisrk_1) - eta, deta = dG1.
__global__ void k_swense_dt(int n1, int n2, double pix, double piy, const double* dG1, const double* dG2,
                            const double* eta, const double* phis, const double* d11, double* deta,
                            double* dphis) {
  const size_t n = (size_t)n1 * n2;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i == 0) {
    deta[0] = dG1[0];
    dphis[0] = d11[0] - eta[0];
    return;
  }
  const double w = wave_omega(wave_k((int)(i % n1), (int)(i / n1), pix, piy));
  deta[i] = dG1[i] + phis[i] * w * w;
  dphis[i] = dG2[i] / w - eta[i];
}

// the mean mode of the free-surface RHS before the omega scaling (resol_HOS.f90:640):
// (D2 + Z)(1,1) of k_expon_out, times its filter weight
__global__ void k_mean_mode(const double* D2, const double* Zh, const double* Zh2, const double* filt,
                            double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double z = Zh[0];
  if (Zh2) z += Zh2[0];
  out[0] = ((D2 ? D2[0] : 0.0) + z) * (filt ? filt[0] : 1.0);
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

class Arena {
 public:
  template <typename T>
  T* get(size_t n) {
    void* p = nullptr;
    HIP_OK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    HIP_OK(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
    ptrs_.push_back(p);
    bytes_ += n * sizeof(T);
    return reinterpret_cast<T*>(p);
  }
  ~Arena() {
    for (void* p : ptrs_) (void)hipFree(p);
  }
  size_t bytes() const { return bytes_; }

 private:
  std::vector<void*> ptrs_;
  size_t bytes_ = 0;
};

// Extended grid size N' >= N: the smallest N' whose line length 2(N'-1) has a
// compiled register-engine sequence, when that costs at most a third more
// points; otherwise the smallest N' with a 2,3,5-smooth 2(N'-1).  Any N' >= N
// leaves the retained modes unchanged (dealiasing factor (M+1)/2).
int pad_smooth(int N) {
  const long need = 2L * (N - 1);
  for (int L : kSeqLengths)
    if (L >= need && 3L * L <= 4L * need) return L / 2 + 1;
  int n = N;
  while (!smooth235(2L * (n - 1))) ++n;
  return n;
}

// ramp (resol_wmkr.f90:1268-1397)
// Dealiasing parameters and extended grid of the order-m right-hand side: the
// context's own p1, p2 at m = mHOS, p1 = p2 = m (3-D) below it, i.e. the grid of a
// context created with mHOS = m (common_vars.f90:38-39 with p = M).
void order_grid(const hos_params& p, int m, int& p1, int& p2, int& N1b, int& N2b) {
  p1 = (m == p.mHOS) ? p.p1 : m;
  p2 = (m == p.mHOS) ? p.p2 : (p.n2 > 1 ? m : 1);
  const int Nd1 = ((p1 + 1) * p.n1) / 2;
  const int Nd2 = p.n2 > 1 ? ((p2 + 1) * p.n2) / 2 : 1;
  N1b = pad_smooth(std::max(Nd1, 2));
  N2b = p.n2 > 1 ? pad_smooth(Nd2) : 1;
}

void ramp_eval(const hos_params& p, double t, double& a, double& at, double& att) {
  if (!ramp_hd(p.iramp, p.Tramp, p.T_stop, p.T_mono, t, a, at, att)) throw std::runtime_error("Unknown ramp");
}

// paddle shape g(z) (wavemaking.f90:184-215)
double gzed(double z, int igeom, double d) {
  const double delta = std::min(d, (1.0 - d) / 10.0);
  if (igeom == 1) return 1.0;
  if (z >= d - 1.0 + delta) return (z + 1.0 - d) / (1.0 - d);
  if (z <= d - 1.0 - delta) return 0.0;
  const double u = (z - (d - 1.0)) / delta;
  return delta / (2.0 * (1.0 - d)) *
         (5.0 / 16.0 + u + 15.0 / 16.0 * (u * u - std::pow(u, 4) / 3.0 + std::pow(u, 6) / 15.0));
}

// vertical derivative of the paddle shape (wavemaking.f90:222-250)
double dgzeddz(double z, int igeom, double d) {
  const double delta = std::min(d, (1.0 - d) / 10.0);
  if (igeom == 1) return 0.0;
  if (z >= d - 1.0 + delta) return 1.0 / (1.0 - d);
  if (z <= d - 1.0 - delta) return 0.0;
  const double u = (z - (d - 1.0)) / delta;
  return 1.0 / (1.0 - d) * (0.5 + 15.0 / 16.0 * (u - 2.0 / 3.0 * u * u * u + u * u * u * u * u / 5.0));
}

// Inverse of the 4x4 system P(-1)=f, P'(-1)=df, P''(-1)=d2f, P(-2/3)=0 for
// P(x) = (c1 x^7 + c2 x^5 + c3 x^3 + c4 x) e^{-6(x+1)}  (match_surface_wmk).
Cont continuation_matrix() {
  double A[4][8] = {};
  const double x0 = -1.0, a = 6.0, xb = -2.0 / 3.0;
  const int pw[4] = {7, 5, 3, 1};
  for (int c = 0; c < 4; ++c) {
    const int n = pw[c];
    const double q = std::pow(x0, n), q1 = n * std::pow(x0, n - 1), q2 = n >= 2 ? n * (n - 1) * std::pow(x0, n - 2) : 0;
    A[0][c] = q;
    A[1][c] = q1 - a * q;
    A[2][c] = q2 - 2 * a * q1 + a * a * q;
    A[3][c] = std::pow(xb, n) * std::exp(-a * (xb + 1.0));
  }
  for (int r = 0; r < 4; ++r) A[r][4 + r] = 1.0;
  for (int c = 0; c < 4; ++c) {  // Gauss-Jordan with partial pivoting
    int piv = c;
    for (int r = c + 1; r < 4; ++r)
      if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
    for (int k = 0; k < 8; ++k) std::swap(A[c][k], A[piv][k]);
    const double d = A[c][c];
    for (int k = 0; k < 8; ++k) A[c][k] /= d;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const double f = A[r][c];
      for (int k = 0; k < 8; ++k) A[r][k] -= f * A[c][k];
    }
  }
  Cont cm{};
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 3; ++k) cm.m[k][r] = A[r][4 + k];  // c_r = sum_k inv[r][k] rhs_k
  return cm;
}

// Sine-to-cosine conversion on the reference's N-point extended grid:
// T[k][s] = S2F_N(sin(s pi x / L) at the N nodes)_k for k < n, s < K
//         = s2f_k (C(s+k) + C(s-k)),  C(q) = cot(pi q / (2(N-1))) for odd q, 0 for even q,
// from sum_{j=1}^{N-2} sin(pi q j/(N-1)) = cot(pi q/(2(N-1))) (q odd).  Stored
// column-major (ld = n).
std::vector<double> sin_to_cos_matrix(int N, int n, int K) {
  const int M = N - 1;
  // C(q) for q = s +- k in (-n, K + n): one evaluation per argument
  const long q0 = -(long)n, nq = (long)K + 2L * n;
  std::vector<double> Ct(nq);
  for (long j = 0; j < nq; ++j) {
    const long q = q0 + j;
    Ct[j] = (q % 2 == 0) ? 0.0 : 1.0 / std::tan(M_PI * (double)q / (2.0 * M));
  }
  auto C = [&](long q) { return Ct[q - q0]; };
  std::vector<double> t((size_t)n * K);
  for (int s = 0; s < K; ++s)
    for (int k = 0; k < n; ++k) {
      const double s2f = (k == 0 || k == N - 1) ? 1.0 / (2.0 * M) : 1.0 / M;
      t[k + (size_t)n * s] = s2f * (C((long)s + k) + C((long)s - k));
    }
  return t;
}

// The two nonzero parity blocks of an n x K map t (column-major, t[k + n s] = 0
// where s + k is even; n and K even), each k-major for the products: block 0
// maps the odd inputs s = 2j + 1 to the even outputs k = 2i, block 1 the even
// inputs s = 2j to the odd outputs k = 2i + 1; element (i, j) of a block at
// j + (K / 2) i.
std::vector<double> parity_halves_kmajor(const std::vector<double>& t, int n, int K) {
  const int h = n / 2, hk = K / 2;
  std::vector<double> out((size_t)2 * h * hk);
  for (int b = 0; b < 2; ++b)
    for (int i = 0; i < h; ++i)
      for (int j = 0; j < hk; ++j) {
        const int k = 2 * i + b, s = 2 * j + 1 - b;
        out[(size_t)b * h * hk + j + (size_t)hk * i] = t[k + (size_t)n * s];
      }
  return out;
}

// C (m x n) = A (m x k) B (k x n), C column-major, on the matrix cores (gemm.h);
// ak / bk: the operand is stored k-contiguous (A[kk + lda i], B[kk + ldb j])
template <typename U>
inline void gemm(hipStream_t st, bool ak, bool bk, int m, int n, int k, const U* A, int lda, const U* B, int ldb,
                 U* Cm, int ldc) {
  hosmm::mm<U>(st, ak, bk, m, n, k, A, lda, 0, B, ldb, 0, Cm, ldc, 0);
}
// C_b (m x n) = A_b (m x n, row-contiguous) B (n x n, k-major) for b < batch,
// A_b and C_b stride s apart
template <typename U>
inline void gemm_yb(hipStream_t st, int m, int n, const U* A, long s, const U* B, U* Cm, int batch) {
  hosmm::mm<U>(st, false, true, m, n, n, A, m, s, B, n, 0, Cm, m, s, batch);
}
// y(m) = AT^T x for AT stored k x m (leading dimension k): one workgroup per
// output, each a contiguous dot product summed in a fixed order in FP64
template <typename U>
__global__ void __launch_bounds__(256) k_gemv_t(int k, const U* __restrict__ AT, const U* __restrict__ x,
                                                U* __restrict__ y) {
  const U* col = AT + (size_t)k * blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < k; i += blockDim.x) acc += (double)col[i] * (double)x[i];
  acc = block_sum(acc);
  if (threadIdx.x == 0) y[blockIdx.x] = (U)acc;
}
template <typename U>
inline void gemv_t(hipStream_t st, int m, int k, const U* AT, const U* x, U* y) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_gemv_t<U>, dim3(m), dim3(256), 0, st, k, AT, x, y);
  HIP_OK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// spectral filters (filtering.f90:334-525)
// ---------------------------------------------------------------------------
// Weight of modes i = 1..n along one direction for coefficient c, as
// filtering_base applies it: c >= 0 keeps the first INT(n c) modes (all when
// c >= 0.9999); -1 zeroes the last mode (the integer quotient i/n); -2 sinc;
// -3 raised cosine; -4 the degree-7 smoothstep of the raised cosine; -5 an
// erfc profile centred at 3n/4 (its 0/0 at the centre is taken at the limit
// 1/2); other negatives the exponential cut-off of i > INT(n |c|).
std::vector<double> filter_weights(int n, double c) {
  const int nf = (c >= 0.9999) ? n : std::abs((int)(n * c));
  const double tiny = std::numeric_limits<double>::epsilon();
  const double tp = 2.0 * M_PI;
  std::vector<double> w(n, 1.0);
  for (int i = 1; i <= n; ++i) {
    double v = 1.0;
    if (c >= 0.0) {
      v = (i > nf) ? 0.0 : 1.0;
    } else if (std::fabs(c + 1.0) < tiny) {
      v = 1.0 - (double)(i / n);
    } else if (std::fabs(c + 2.0) < tiny) {
      const double a = tp * i / n;
      v = std::sin(a) / a;
    } else if (std::fabs(c + 3.0) < tiny) {
      v = (1.0 + std::cos(tp * i / n)) / 2.0;
    } else if (std::fabs(c + 4.0) < tiny) {
      const double s0 = (1.0 + std::cos(tp * i / n)) / 2.0;
      v = std::pow(s0, 4) * (35.0 - 84.0 * s0 + 70.0 * s0 * s0 - 20.0 * s0 * s0 * s0);
    } else if (std::fabs(c + 5.0) < tiny) {
      const double s0 = ((double)i - 0.75 * n) / (n + 1.0);
      const double q = (4.0 / 3.0 * s0) * (4.0 / 3.0 * s0);
      const double r = (q > 0.0) ? -std::log(1.0 - q) / q : 1.0;
      v = 0.5 * std::erfc(2.0 * std::sqrt(10.0) * s0 * std::sqrt(r));
    } else if (i > nf) {
      const double d = tp * (double)(i / n) - tp * (double)(nf / n);
      v = std::exp(-2.0 * d * d * d * d);
    }
    w[i - 1] = v;
  }
  return w;
}

// A zero coefficient (a zero-initialised hos_params) means off, like 1.
inline bool filter_active(double c) { return c < 1.0 && c != 0.0; }

// 2-D weights of filtering_x (coefficient cx) then filtering_y (cy) on an
// n1 x n2 modal array; the exponential y cut-off touches only the last x mode
// (filtering.f90:513-517).
std::vector<double> filter_weights_2d(int n1, int n2, double cx, double cy) {
  std::vector<double> w((size_t)n1 * n2, 1.0);
  if (filter_active(cx)) {
    const std::vector<double> wx = filter_weights(n1, cx);
    for (int i2 = 0; i2 < n2; ++i2)
      for (int i1 = 0; i1 < n1; ++i1) w[i1 + (size_t)n1 * i2] *= wx[i1];
  }
  if (n2 > 1 && filter_active(cy)) {
    const double tiny = std::numeric_limits<double>::epsilon();
    bool expo = cy < 0.0;
    for (double t : {1.0, 2.0, 3.0, 4.0, 5.0})
      if (std::fabs(cy + t) < tiny) expo = false;
    const std::vector<double> wy = filter_weights(n2, cy);
    for (int i2 = 0; i2 < n2; ++i2)
      for (int i1 = expo ? n1 - 1 : 0; i1 < n1; ++i1) w[i1 + (size_t)n1 * i2] *= wy[i2];
  }
  return w;
}

// ---------------------------------------------------------------------------
// base-grid transforms as dense products
// ---------------------------------------------------------------------------
// At the reference sizes the base-grid lengths 2(n-1) carry large primes
// (1024: 2046 = 2.3.11.31, 256: 510 = 2.3.5.17), so an FFT needs Bluestein's
// doubled length; on a grid this small the transform is latency-bound, and the
// N x N DCT-I/DST-I matrix product (FP64/FP32 MFMA, gemm.h) is both faster and
// exact to rounding.  Column-major N x N, out_r = sum_c M[r + N c] in_c, with
// the normalisations of fourier_FFTW.f90:178-229 folded in:
//   cos forward  a_r = s2f_r (x_0 + (-1)^r x_{N-1} + 2 sum_{0<c<N-1} x_c cos(pi r c/(N-1)))
//   cos inverse  x_r = sum_c a_c cos(pi r c/(N-1))
//   sin forward  a_r = 2/(N-1) sum_{0<c<N-1} x_c sin(pi r c/(N-1)), zero end rows
//   sin inverse  x_r = sum_{0<c<N-1} a_c sin(pi r c/(N-1)), zero end rows
std::vector<double> r2r_matrix(int N, int kind, int dir) {
  std::vector<double> m((size_t)N * N, 0.0);
  if (N == 1) {
    m[0] = (kind == KCOS) ? 1.0 : 0.0;
    return m;
  }
  const long P = 2L * (N - 1);
  // the angle depends on r c mod P only: one trigonometric evaluation per residue
  std::vector<double> tab(P);
  for (long k = 0; k < P; ++k) {
    const double ang = M_PI * (double)k / (double)(N - 1);
    tab[k] = (kind == KCOS) ? std::cos(ang) : std::sin(ang);
  }
  for (int c = 0; c < N; ++c) {
    const bool ec = (c == 0 || c == N - 1);
    long k = 0;  // r c mod P, stepped by c
    for (int r = 0; r < N; ++r, k = (k + c) % P) {
      const bool er = (r == 0 || r == N - 1);
      double v;
      if (kind == KCOS)
        v = (dir == FWD) ? (er ? 1.0 : 2.0) / (double)P * (ec ? 1.0 : 2.0) * tab[k] : tab[k];
      else
        v = (er || ec) ? 0.0 : ((dir == FWD) ? 2.0 / (double)(N - 1) : 1.0) * tab[k];
      m[r + (size_t)N * c] = v;
    }
  }
  return m;
}

// r2r_matrix on the device from the host table of the P = 2(N-1) residues:
// stored element e of the N x N matrix is row r, column c (transposed: e = c + N r).
template <typename U>
__global__ void k_r2r_matrix(int N, int kind, int dir, int transposed, const double* __restrict__ tab, U* out) {
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e >= (size_t)N * N) return;
  const long P = 2L * (N - 1);
  const int a = (int)(e % N), b = (int)(e / N);
  const int r = transposed ? b : a, c = transposed ? a : b;
  const double t = tab[((long)r * c) % P];
  const bool er = (r == 0 || r == N - 1), ec = (c == 0 || c == N - 1);
  double v;
  if (kind == KCOS)
    v = (dir == FWD) ? (er ? 1.0 : 2.0) / (double)P * (ec ? 1.0 : 2.0) * t : t;
  else
    v = (er || ec) ? 0.0 : ((dir == FWD) ? 2.0 / (double)(N - 1) : 1.0) * t;
  out[e] = (U)v;
}

// Dense path applies when a base-grid direction has no FFT plan without
// Bluestein and the matrices stay small (N <= 4096).
inline bool dense_base_grid(int n1, int n2) {
  auto dense = [](int n) { return n > 1 && !seq_supported(2 * (n - 1)); };
  return n1 <= 4096 && n2 <= 4096 && (dense(n1) || (n2 > 1 && dense(n2)));
}

template <typename U, typename S>
__global__ void k_dense_in(int n1, int n2, S src, U* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)n1 * n2) return;
  const int i1 = (int)(i % n1), i2 = (int)(i / n1);
  out[i] = (U)src.get(i1, i2);
}

// The DCT-I / DST-I matrices of both directions satisfy M[r][n-1-k] = sig_r M[r][k]
// (sig_r = (-1)^r for cos, -(-1)^r for sin: cos(pi r k / (n-1)) is symmetric in r
// and k, and the end weights and normalisations are symmetric), so for even n
// the product's even rows need only x_k + sig_0 x_{n-1-k} (k < n/2) and its odd
// rows x_k - sig_0 x_{n-1-k}: the staged source holds those two halves (even-row
// half first) and two half-size products replace the full one.
template <typename U, class S>
__global__ void k_dense_fold(int n1, int n2, int sig0, S src, U* __restrict__ out) {
  const int h = n1 / 2;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)h * n2) return;
  const int k = (int)(i % h), i2 = (int)(i / h);
  const U a = (U)src.get(k, i2), b = (U)src.get(n1 - 1 - k, i2);
  out[k + (size_t)n1 * i2] = sig0 > 0 ? a + b : a - b;
  out[h + k + (size_t)n1 * i2] = sig0 > 0 ? a - b : a + b;
}
// k_dense_fold for up to four sources in one launch (blockIdx.y = source),
// source f staged at out + f n1 n2
template <class S>
struct FoldSet {
  S src[4];
};
template <typename U, class S>
__global__ void k_dense_fold_multi(int n1, int n2, int sig0, FoldSet<S> fs, U* __restrict__ out) {
  const int h = n1 / 2;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)h * n2) return;
  const S& src = fs.src[blockIdx.y];
  U* o = out + (size_t)blockIdx.y * n1 * n2;
  const int k = (int)(i % h), i2 = (int)(i / h);
  const U a = (U)src.get(k, i2), b = (U)src.get(n1 - 1 - k, i2);
  o[k + (size_t)n1 * i2] = sig0 > 0 ? a + b : a - b;
  o[h + k + (size_t)n1 * i2] = sig0 > 0 ? a - b : a + b;
}

// the even rows (b = 0) and the odd rows (b = 1) of a k-major matrix, restricted
// to k < n/2: out[b h h + k + h i] = M[2i + b][k]
template <typename U>
__global__ void k_parity_rows(int n1, const U* __restrict__ M, U* __restrict__ out) {
  const int h = n1 / 2;
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= (size_t)2 * h * h) return;
  const int k = (int)(i % h), r = (int)((i / h) % h), b = (int)(i / ((size_t)h * h));
  out[i] = M[k + (size_t)n1 * (2 * r + b)];
}

template <typename U>
struct DenseGrid {
  int n1 = 0, n2 = 1;
  U* mx[2][2] = {};  // [kind][0 forward, 1 inverse]
  U* mxh[2][2] = {};  // [kind][direction]: the parity halves of mx (k_parity_rows), when fold
  bool fold = false;
  U* my[2][2] = {};
  U* stage = nullptr;  // source after multiplier / conversion
  U* mid = nullptr;    // after the x product
};

// the transpose of a column-major r x c matrix (column-major c x r)
std::vector<double> transposed(const std::vector<double>& a, int r, int c) {
  std::vector<double> t(a.size());
  for (int j = 0; j < c; ++j)
    for (int i = 0; i < r; ++i) t[j + (size_t)c * i] = a[i + (size_t)r * j];
  return t;
}

// W[j + n m] = s nu[m] M[j + n m]: the rows of a stored n x n matrix scaled
__global__ void k_scale_rows(int n, double s, const double* __restrict__ nu, const double* __restrict__ M, double* W) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < (size_t)n * n) W[i] = s * nu[i / n] * M[i];
}

template <typename U>
U* upload_matrix(Arena& ar, const std::vector<double>& m) {
  std::vector<U> h(m.begin(), m.end());
  U* d = ar.get<U>(h.size());
  HIP_OK(hipMemcpy(d, h.data(), h.size() * sizeof(U), hipMemcpyHostToDevice));
  return d;
}

// The N x N matrix of r2r_matrix built on the device (N >= 2); synchronous.
template <typename U>
U* device_matrix(Arena& ar, int N, int kind, int dir, bool transposed) {
  const long P = 2L * (N - 1);
  std::vector<double> tab(P);
  for (long k = 0; k < P; ++k) {
    const double ang = M_PI * (double)k / (double)(N - 1);
    tab[k] = (kind == KCOS) ? std::cos(ang) : std::sin(ang);
  }
  double* dt = ar.get<double>(P);
  HIP_OK(hipMemcpy(dt, tab.data(), P * sizeof(double), hipMemcpyHostToDevice));
  U* m = ar.get<U>((size_t)N * N);
  hipLaunchKernelGGL(k_r2r_matrix<U>, dim3(nblk((size_t)N * N)), dim3(256), 0, 0, N, kind, dir, transposed ? 1 : 0,
                     (const double*)dt, m);
  HIP_OK(hipGetLastError());
  HIP_OK(hipDeviceSynchronize());
  return m;
}

template <typename U>
U* dense_matrix(Arena& ar, int N, int kind, int dir, bool transposed) {
  if (N == 1) return upload_matrix<U>(ar, r2r_matrix(1, kind, dir));
  return device_matrix<U>(ar, N, kind, dir, transposed);
}

template <typename U>
void build_dense(DenseGrid<U>& g, Arena& ar, int n1, int n2) {
  check_buffer_grid(n1, n2);
  g.n1 = n1;
  g.n2 = n2;
  // stored transposed (k-major): row r of the matrix is contiguous, the layout
  // of the constant operand of the products and of the 2-D tank's GEMV
  for (int k = 0; k < 2; ++k)
    for (int d = 0; d < 2; ++d) {
      g.mx[k][d] = dense_matrix<U>(ar, n1, k, d == 0 ? FWD : INV, true);
      if (n2 > 1) g.my[k][d] = dense_matrix<U>(ar, n2, k, d == 0 ? FWD : INV, true);
    }
  // (n1 a multiple of 8: the half products keep 16-byte operand rows)
  g.fold = n2 > 1 && n1 % 8 == 0;
  if (g.fold) {
    const size_t hh = (size_t)(n1 / 2) * (n1 / 2);
    for (int k = 0; k < 2; ++k)
      for (int d = 0; d < 2; ++d) {
        g.mxh[k][d] = ar.get<U>(2 * hh);
        hipLaunchKernelGGL(k_parity_rows<U>, dim3((unsigned)((2 * hh + 255) / 256)), dim3(256), 0, 0, n1,
                           (const U*)g.mx[k][d], g.mxh[k][d]);
        HIP_OK(hipGetLastError());
      }
    HIP_OK(hipDeviceSynchronize());
  }
  g.stage = ar.get<U>((size_t)n1 * n2);
  g.mid = ar.get<U>((size_t)n1 * n2);
}

// out (nout1 x nout2, plain store) = My[ky] (Mx[kx] src)^T restricted to the
// requested output extents.
template <typename U, class S>
void dense2d(const DenseGrid<U>& g, hipStream_t st, int kx, int ky, int dir, const S& src,
             int nout1, int nout2, const Dst<U>& d) {
  if (d.mode != 0) throw std::runtime_error("dense base transform: plain stores only");
  const int n1 = g.n1, n2 = g.n2, di = (dir == FWD) ? 0 : 1;
  const size_t n = (size_t)n1 * n2;
  if (g.fold && nout1 == n1 && n2 > 1) {
    // folded source, then the even and odd rows' half products written to every
    // other row of mid (one launch)
    const int h = n1 / 2;
    hipLaunchKernelGGL((k_dense_fold<U, S>), dim3((unsigned)(((size_t)h * n2 + 255) / 256)), dim3(256), 0, st, n1,
                       n2, kx == KCOS ? 1 : -1, src, g.stage);
    hosmm::mm<U>(st, true, true, h, n2, h, (const U*)g.mxh[kx][di], h, (long)h * h, (const U*)g.stage, n1, h, g.mid,
                 n1, 1, 2, 0, 2);
    gemm(st, false, true, nout1, nout2, n2, (const U*)g.mid, n1, g.my[ky][di], n2, d.p, d.ld);
    return;
  }
  // a plain source of the product's type and layout is read in place
  const U* in = g.stage;
  bool staged = true;
  if constexpr (std::is_same<S, Src<U, U>>::value)
    if (src.mul == MUL_NONE && src.ld == n1 && src.nx == n1 && src.ny == n2 && (const void*)src.p != d.p) {
      in = src.p;
      staged = false;
    }
  if (staged) {
    hipLaunchKernelGGL((k_dense_in<U, S>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n1, n2, src, g.stage);
    HIP_OK(hipGetLastError());
  }
  if (n2 == 1) {
    gemv_t(st, nout1, n1, g.mx[kx][di], in, d.p);
    return;
  }
  gemm(st, true, true, nout1, n2, n1, g.mx[kx][di], n1, in, n1, g.mid, n1);
  gemm(st, false, true, nout1, nout2, n2, (const U*)g.mid, n1, g.my[ky][di], n2, d.p, d.ld);
}

// y_f[m] = sum_k AT[k + n m] x_f[k], f < nv: dense FP64 base transforms of a 2-D
// tank for several vectors with one read of the matrix (one workgroup per output,
// the vectors from L2), each sum in a fixed order.
constexpr int kMaxGemv = 3;
struct GemvSet {
  int nv;
  const double* x[kMaxGemv];
  double* y[kMaxGemv];
};
__global__ void __launch_bounds__(256) k_gemv_multi(int n, const double* __restrict__ AT, GemvSet g) {
  const int m = blockIdx.x;
  const double* col = AT + (size_t)n * m;
  double acc[kMaxGemv] = {0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const double a = col[k];
#pragma unroll
    for (int f = 0; f < kMaxGemv; ++f)
      if (f < g.nv) acc[f] += a * g.x[f][k];
  }
#pragma unroll
  for (int f = 0; f < kMaxGemv; ++f) {
    if (f < g.nv) {  // the same for every thread
      const double sum = block_sum(acc[f]);
      if (threadIdx.x == 0) g.y[f][m] = sum;
    }
  }
}

}  // namespace

// ===========================================================================
// engine: the right-hand side in transform type T
// ===========================================================================
// ---------------------------------------------------------------------------
// extended-grid schedule autotuner
// ---------------------------------------------------------------------------
// A schedule (GridSched: the compiled radix sequence of the x and of the y
// lines, the y tile width) is chosen per extended grid and precision by timed
// trials when the first engine of that grid is built (Engine::tune), and kept
// in a process-wide table so that every context of one shape runs the same
// arithmetic (contexts of one shape give bit-identical results).  Modes
// (HOSNWT_TUNE): 1 (default) trials, 0 the default schedule; HOSNWT_FFT_SCHED
// ="xs,ys[,tx]" forces one; HOSNWT_TUNE_CACHE names a text file the table is
// read from at first use and new choices are appended to, so that later
// processes repeat them without trials.
struct TuneKey {
  int n1, n2, bytes, dev;
  bool operator<(const TuneKey& o) const {
    return std::tie(n1, n2, bytes, dev) < std::tie(o.n1, o.n2, o.bytes, o.dev);
  }
};
struct TuneRec {
  GridSched s{};
  double us_default = 0.0, us_best = 0.0;  // one trial batch, default and chosen schedule
  int trials = 0;                          // candidates timed (0: from the cache file or forced)
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneRec> g_tune;
bool g_tune_loaded = false;

// the forced schedule, if any
bool env_sched(GridSched& g) {
  if (const char* e = std::getenv("HOSNWT_FFT_SCHED")) {
    g = GridSched{};
    if (std::sscanf(e, "%d,%d,%d", &g.xs, &g.ys, &g.tx) < 2) throw std::runtime_error("HOSNWT_FFT_SCHED: xs,ys[,tx]");
    return true;
  }
  return false;
}
bool tune_enabled() {
  const char* e = std::getenv("HOSNWT_TUNE");
  return !(e && std::atoi(e) == 0);
}
// Looks the grid up (loading the cache file once); false when it must be tuned.
bool tune_lookup(const TuneKey& k, TuneRec& out) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (!g_tune_loaded) {
    g_tune_loaded = true;
    if (const char* f = std::getenv("HOSNWT_TUNE_CACHE")) {
      if (FILE* fp = std::fopen(f, "r")) {
        TuneKey q{};
        TuneRec r{};
        while (std::fscanf(fp, "%d %d %d %d %d %d %d %lf %lf", &q.n1, &q.n2, &q.bytes, &q.dev, &r.s.xs, &r.s.ys,
                           &r.s.tx, &r.us_default, &r.us_best) == 9) {
          r.trials = 0;
          g_tune[q] = r;
        }
        std::fclose(fp);
      }
    }
  }
  auto it = g_tune.find(k);
  if (it == g_tune.end()) return false;
  out = it->second;
  return true;
}
// Records a tuned grid; the first record of a key wins (a concurrent tuning of
// the same grid adopts it), and is appended to the cache file.
TuneRec tune_store(const TuneKey& k, const TuneRec& r) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto ins = g_tune.emplace(k, r);
  if (ins.second) {
    if (const char* f = std::getenv("HOSNWT_TUNE_CACHE")) {
      if (FILE* fp = std::fopen(f, "a")) {
        std::fprintf(fp, "%d %d %d %d %d %d %d %.3f %.3f\n", k.n1, k.n2, k.bytes, k.dev, r.s.xs, r.s.ys, r.s.tx,
                     r.us_default, r.us_best);
        std::fclose(fp);
      }
    }
  }
  return ins.first->second;
}

struct EngineBase {
  virtual ~EngineBase() = default;
  // S2 carries the second-order wavemaker stage (null below i_wmk = 2)
  struct S2 {
    const double *a2, *e2, *aad2;  // a2nd (modal), eta2nd (physical), a2nd_add (modal)
    double *ka2, *ke2, *kad2;      // their derivatives
    const double* aad3;            // a3rd_add (modal), null below i_wmk = 3
    double* kad3;                  // its derivative
  };
  virtual void rhs(double t, double dt, const double* G1, const double* G2, const double* a1, const double* e1,
                   const double* aad, double* kG1, double* kG2, double* ka1, double* ke1, double* kad1,
                   double* kadr, const S2* s2) = 0;
  virtual void slope_max(double& sx, double& sy) = 0;
  // FP64 base-grid cos-cos transform on the context stream (host-facing state I/O)
  virtual void base_cc64(int dir, const double* in, double* out) = 0;
  // paddle displacement and velocity at the top node (i3 = n3) for every y node,
  // from the last wavemaker evaluation (host arrays of n2)
  virtual void paddle_top(double* x, double* xt) = 0;
  // average duration (ms) and algorithmic HBM bytes of one extended-grid x pass
  // and one y pass (full inverse cos-cos transform), timed with HIP events
  virtual void time_big(int reps, int items, double* out4) = 0;
  // SWENSE output (init_velocity_comp + HOSvel2_SWEET + HOSvel_SWEET) after a right-hand
  // side at the committed state: velocity and phi_t modes at z = 0 (out[0..3], n1 x n2),
  // the modal d eta / dt (out[4]); device pointers
  virtual void swense(int meta2, double** out) = 0;
  // the extended grid's schedule: {x line length, x values per thread, x sequence,
  // y line length, y values per thread, y sequence, y tile width, candidates timed}
  // and the trial batch's time (us) with the default and the chosen schedule
  virtual void schedule(int* s8, double* us2) const = 0;
};

struct hos_ctx {
  hos_params p{};
  int device = 0;  // the HIP device the context was created on
  int n1, n2, n3, n3a, M, N1b, N2b;
  double pix, piy, dx, dy, delz;
  bool wmk;
  // dipole wavemaker (igeom = 3): the components' moments and depths on the device
  bool dip = false;
  DipSet dips{};
  // an additional potential phi_add enters the free-surface conditions
  bool add_pot() const { return wmk || dip; }
  hipStream_t st = nullptr;
  Arena ar;
  Grid2<double> gs64;
  // right-hand-side engines by HOS order (built on first use); eng is the active one,
  // of order M on the extended grid N1b x N2b
  std::unique_ptr<EngineBase> engines[9];
  EngineBase* eng = nullptr;
  // per-launch pass timing between hos_pass_log_begin and hos_pass_log_end
  std::unique_ptr<PassLog> plog;
  // host copies for the paddle-top displacement pos1st(n3,1)
  std::vector<double> freq, pre0, pim0;
  // host copies of the wavemaker inputs (pos_ini, dposdy_ini: n2 x n_freq; d nu/dx,
  // d nu/dy: n1 x n2; empty when absent), from which every engine is built
  std::vector<double> h_pos_re, h_pos_im, h_dy_re, h_dy_im, h_nux, h_nuy;
  // the beach nu(x) when it does not depend on y (empty otherwise)
  std::vector<double> h_nu_x;
  double gz_top = 1.0;
  // wavemaker globals of the last evaluation (alpha, pos1st(n3,1))
  double last_alpha = 0.0, last_pos_top = 0.0, last_alphat = 0.0;
  // device state
  double *eta_hat, *phis_hat, *a1st_add, *a_add_ramp, *a1st, *eta1st;
  double *eta_hat_t, *phis_hat_t, *a1st_add_t, *a_add_ramp_t, *a1st_t, *eta1st_t, *deta_hat, *deta_hat_t;
  double *G1, *G2, *G1m, *G2m, *aadm, *a1m, *e1m;
  double *kG1[5], *kG2[5], *kad1[5], *kadr[5], *ka1[5], *ke1[5];
  // second-order wavemaker state (i_wmk = 2; null otherwise): paddle modes, modal and
  // physical free-surface parts, trial copies, stage derivatives and stage values
  bool wmk2 = false;
  double *a2nd_add = nullptr, *a2nd = nullptr, *eta2nd = nullptr;
  double *a2nd_add_t = nullptr, *a2nd_t = nullptr, *eta2nd_t = nullptr;
  double *ka2[5] = {}, *ke2[5] = {}, *kad2[5] = {};
  double *a2m = nullptr, *e2m = nullptr, *aad2m = nullptr;
  // third-order wavemaker paddle modes (i_wmk = 3; null otherwise)
  bool wmk3 = false;
  double *a3rd_add = nullptr, *a3rd_add_t = nullptr, *aad3m = nullptr, *kad3[5] = {};
  double *partial, *pmax_e, *pmax_x, *red6;
  size_t pmax_stride;
  double* red_host = nullptr;  // pinned copy of red6 (the attempt's error maxima)
  // learned order selection (hos_advance_learned): current order, floor after a
  // reversion, accepted attempts since the last judge, the phase of the attempt in
  // flight (0 at the selected order, 1 its full-order judge), the checkpoint of the
  // last successful judge, and the pinned copies of the attempt's feature record,
  // predictions and judged difference norms
  struct Learn {
    bool ready = false;
    int m = 0, floor = 0, since = 0, phase = 0, m_attempt = 0, full_run = 0;
    double hl = 0.0, err_m = 0.0, pred_rate = 0.0;
    // judged / estimated drift rate: the judge's calibration of the selection's estimate
    double kappa = 1.0;
    double t_ck = 0.0, h_ck = 0.0, alpha_ck = 0.0, alphat_ck = 0.0, pos_ck = 0.0;
    int endramp_ck = 0;
    std::vector<double*> ck;
    double *judge_eta = nullptr, *dpart = nullptr, *dn = nullptr;
    double *rec_host = nullptr, *dn_host = nullptr;
    float* pred_host = nullptr;
    hosmlp::Workspace ws;
    hos_learned_stats stats{};
  } learn;
  ~hos_ctx() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (dsc_host) (void)hipHostFree(dsc_host);
    if (red_host) (void)hipHostFree(red_host);
    if (learn.rec_host) (void)hipHostFree(learn.rec_host);
    if (learn.dn_host) (void)hipHostFree(learn.dn_host);
    if (learn.pred_host) (void)hipHostFree(learn.pred_host);
    if (learn.ws.x) hosmlp::workspace_free(learn.ws);
  }
  // band-energy records of the learned order selection: feat_mode 0 off, 1 the
  // first right-hand side of each attempt into record 0, 2 every right-hand side
  // into a ring of feat_cap records (feat_n written since the last take)
  int feat_mode = 0, feat_cap = 0, feat_n = 0, cur_stage = 0;
  // graph-captured attempts (hos_advance_graph): the device step scalars (DS_COUNT
  // doubles), the paddle-top tables for them, dsc_active = dsc while an attempt is
  // being captured (the kernels then read t, h and the ramp from memory)
  double *dsc = nullptr, *dsc_active = nullptr, *freq_d = nullptr, *pre0_d = nullptr, *pim0_d = nullptr;
  double* dsc_host = nullptr;  // pinned copy of dsc
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<const void*> graph_sig;  // state pointers and engine the graph was captured with
  double *feat = nullptr, *feat_part = nullptr;
  double *phys_tmp, *nu_d;
  double* filt_fs = nullptr;  // free-surface mode filter weights (n1*n2), null when off
  double* filt_z = nullptr;   // paddle z-mode filter weights (n3_add), null when off
  double *phi_add_last;
  hos_stats stats{};
  bool have_trial = false;
  int endramp = 0;
  // rk constants (runge_kutta.f90:46-65)
  double A[5][5] = {{0}, {0.5}, {0, 0.5}, {0, 0, 1.0}, {-1.0, 2.0}};
  double B[5] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0, 0.0};
  double C[5] = {0.0, 0.5, 0.5, 1.0, 1.0};
  double E[5] = {1.0 / 6.0, 2.0 / 3.0, 0.0, 0.0, 1.0 / 6.0};

  double pos_top(double t) const {
    double s = 0.0;
    for (size_t f = 0; f < freq.size(); ++f) {
      const double w = 2.0 * M_PI * freq[f];
      s += pre0[f] * std::cos(w * t) - pim0[f] * std::sin(w * t);
    }
    return gz_top * s;
  }
  int volmode() const {
    if (!wmk) return 0;
    return p.igeom == 1 ? 1 : (p.igeom == 2 ? 2 : 0);
  }
};

template <typename T>
struct Engine : EngineBase {
  hos_ctx* c;
  // HOS order of this engine and its extended grid (a context holds one engine per
  // order it runs at; the modal RK state is shared, only the extended grid depends on M)
  int M, N1b, N2b, p1, p2;
  Grid2<T> gb;   // extended grid
  Grid2<T> gsT;  // base grid in T
  size_t nb, ns, na;
  // base (T)
  T *eta_hat, *phis_hat, *eta_phys, *P, *D1, *D2, *Q, *Zp, *Zh, *phixh, *phiyh, *phizh, *tmpS;
  // big (T)
  T* eta_ext = nullptr;  // eta on the extended grid (M >= 2; the HOS levels form its powers)
  std::vector<T*> phizm, aHOS;
  std::vector<T*> lvl;  // outputs of one HOS level's transforms
  std::vector<T*> tmpX;  // x-pass intermediates of batched transforms
  double* slope_part = nullptr;  // block maxima of |eta_x|, |eta_y| (slope_max)
  T *etax, *etay, *phisx, *phisy, *phim, *phixa, *phiya, *phiza, *X, *Y, *tmpB;
  // wavemaker / first order (double)
  double *Xp, *Xpt, *Xptt, *f1, *f2, *fz1, *fz2, *gz, *Wz, *Wy, *freq, *pre, *pim, *part;
  int nch = 1;
  double *phi_add, *phix_add, *phiy_add, *phiz_add, *phit_add, *phit1, *phiz1, *w1, *tmp1;
  T *phix_addT, *phiy_addT, *phiz_addT;
  // Mixed path: the first-order terms of the free-surface conditions stay out of
  // FP32.  The linear vertical velocity kth1 phis_hat cancels exactly against the
  // term the change of variables subtracts, so P keeps orders >= 2 only; the
  // paddle forcing (phiz_add, phit_add) and the beach's first-order term
  // nu F2S(kth1 phis_hat) are transformed in FP64 on the wavemaker stream.  FP32
  // rounding then enters through the products of order >= 2 only.  From rest the
  // fields are first-order, and an FP32 linear part sets an h-independent floor
  // (~2.5e-7 relative) under the embedded error estimate, which a 1e-7 tolerance
  // never passes.
  static constexpr bool kSplit = sizeof(T) != sizeof(double);
  double *phis64 = nullptr, *wlin = nullptr, *Zf64 = nullptr, *Zh64 = nullptr, *phizh64 = nullptr;
  Cont cont;
  // mixed-parity products: spectra, conversion matrices, results
  int K1 = 0, K2 = 0;
  T *XA, *YA, *XB, *YB, *PA, *PB, *RA, *RB, *T1, *T2;
  int t1_chunk = 0, t1_parts = 0;  // 2-D: K chunk of the T1 product and the number of chunks
  bool psplit = false;             // the sine-to-cosine products split by parity
  // The wavemaker chain (paddle motion, paddle-grid products, wavemaker sums,
  // first-order and gradient transforms) depends only on the base-grid eta, so
  // it runs on a second stream beside the HOS recursion, with its own scratch;
  // helpers use the current stream / scratch (cs, ctmp, ctmpX).
  hipStream_t sw = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join2 = nullptr;
  bool fo_late = false;  // this evaluation's first_order_rest is pending
  T* tmpBw = nullptr;
  std::vector<T*> tmpXw;
  hipStream_t cs = nullptr;
  T* ctmp = nullptr;
  std::vector<T*>* ctmpX = nullptr;
  // HBM workspaces of long extended x lines, one per stream (null when unused)
  T *gwsM = nullptr, *gwsS = nullptr, *cgws = nullptr;
  void use_main() {
    cs = c->st;
    ctmp = tmpB;
    ctmpX = &tmpX;
    cgws = gwsM;
  }
  void use_side() {
    cs = sw;
    ctmp = tmpBw;
    ctmpX = &tmpXw;
    cgws = gwsS;
  }
  // second-order wavemaker (i_wmk = 2): paddle derivatives, stacked paddle
  // fields, inverse paddle-grid matrices, z = 0 sums and first-order fields
  double *dyre = nullptr, *dyim = nullptr, *dgz = nullptr, *nux = nullptr, *nuy = nullptr;
  double *Xdy = nullptr, *Xdyt = nullptr, *Xdz = nullptr, *Xdzt = nullptr;
  double *stk6 = nullptr, *stkz = nullptr, *Pw = nullptr, *Pw3 = nullptr;
  double *Cz = nullptr, *Sz = nullptr, *Cy = nullptr, *Sy = nullptr;
  double *part2 = nullptr, *zf[11] = {};
  double *modes8 = nullptr, *T8 = nullptr, *a1f = nullptr;
  double *phit2 = nullptr, *phiz2 = nullptr, *w2 = nullptr, *tmp2 = nullptr, *atot = nullptr, *dtot = nullptr;
  // mixed path: the beach operator of the orders >= 2 (n1 x n1, column-major), or null
  T* Bnu = nullptr;
  // base grid as dense products (lengths without a Bluestein-free FFT plan)
  bool dense = false;
  DenseGrid<T> dT;
  DenseGrid<double> d64;
  // 3-D dense base grid: the wavemaker chain's base transforms of one kind and
  // direction go through one x product for all fields and one strided batch of
  // y products (kMulti fields at most); staging, intermediate and output scratch
  static constexpr int kMulti = 4;
  double* multi = nullptr;
  DenseGrid<double>& grid64() {
    if constexpr (sizeof(T) == sizeof(double))
      return dT;
    else
      return d64;
  }

  Engine(hos_ctx* ctx, int order) : c(ctx), M(order) {
    Arena& ar = c->ar;
    order_grid(c->p, order, p1, p2, N1b, N2b);
    gb.build(N1b, N2b);
    gb.check_lengths(true);
    gsT.build(c->n1, c->n2);
    nb = (size_t)N1b * N2b;
    ns = (size_t)c->n1 * c->n2;
    na = (size_t)c->n3a * c->n2;
    eta_hat = ar.get<T>(ns); phis_hat = ar.get<T>(ns); eta_phys = ar.get<T>(ns);
    P = ar.get<T>(ns); D1 = ar.get<T>(ns); D2 = ar.get<T>(ns); Q = ar.get<T>(ns);
    // (outputs of one batched base-grid product lie ns apart: one launch writes them)
    Zp = ar.get<T>(ns); Zh = ar.get<T>(ns); phixh = ar.get<T>(3 * ns); phiyh = phixh + ns;
    phizh = phiyh + ns; tmpS = ar.get<T>(ns);
    phix_addT = ar.get<T>(ns); phiy_addT = ar.get<T>(ns); phiz_addT = ar.get<T>(ns);
    if (kSplit) {
      phis64 = ar.get<double>(ns);
      Zf64 = ar.get<double>(ns); phizh64 = ar.get<double>(2 * ns); Zh64 = phizh64 + ns;
    }
    if (M >= 2) eta_ext = ar.get<T>(nb);
    phizm.assign(M + 1, nullptr);
    aHOS.assign(M + 1, nullptr);
    for (int j = 1; j <= M; ++j) {
      phizm[j] = ar.get<T>(nb);
      if (j >= 2) aHOS[j] = ar.get<T>(nb);
    }
    for (int i = 0; i < kMaxBatch && i < std::max(M, 7); ++i) tmpX.push_back(ar.get<T>(nb));
    for (int i = 0; i < M; ++i) lvl.push_back(ar.get<T>(nb));
    etax = ar.get<T>(nb); etay = ar.get<T>(nb); phisx = ar.get<T>(nb); phisy = ar.get<T>(nb);
    slope_part = ar.get<double>(2 * (size_t)nblk(nb));
    phim = ar.get<T>(nb); phixa = ar.get<T>(nb); phiya = ar.get<T>(nb);
    phiza = ar.get<T>(nb); X = ar.get<T>(nb); Y = ar.get<T>(nb); tmpB = ar.get<T>(nb);
    // wavemaker
    const int n2 = c->n2, n3 = c->n3, n3a = c->n3a;
    Xp = ar.get<double>((size_t)n3 * n2); Xpt = ar.get<double>((size_t)n3 * n2); Xptt = ar.get<double>((size_t)n3 * n2);
    f1 = ar.get<double>(2 * na); f2 = f1 + na; fz1 = ar.get<double>(2 * na); fz2 = ar.get<double>(2 * na);
    nch = (n3a + kPhiI3 - 1) / kPhiI3;
    part = (n2 > 1) ? ar.get<double>((size_t)nch * kPhiF * ns) : nullptr;
    // phi_add is read only by hos_get_physical, after the join of the evaluation that
    // wrote it: the context's last-evaluation field is this buffer itself
    phi_add = c->phi_add_last;
    phix_add = ar.get<double>(ns); phiy_add = ar.get<double>(ns);
    phiz_add = ar.get<double>(ns); phit_add = ar.get<double>(ns); phit1 = ar.get<double>(ns);
    phiz1 = ar.get<double>(ns); w1 = ar.get<double>(kSplit ? 2 * ns : ns); tmp1 = ar.get<double>(ns);
    if (kSplit) wlin = w1 + ns;
    // REDFT01 along z and REDFT00 along y as column-major matrices
    std::vector<double> h_gz(n3), h_wz((size_t)n3a * n3a), h_wy((size_t)n2 * n2);
    for (int i3 = 0; i3 < n3; ++i3) h_gz[i3] = gzed(i3 * c->delz - 1.0, c->p.igeom, c->p.d_hinge);
    for (int k = 0; k < n3a; ++k)
      for (int j = 0; j < n3a; ++j)
        h_wz[k + (size_t)n3a * j] = (j == 0 ? 1.0 : 2.0) * std::cos(M_PI * j * (2 * k + 1) / (2.0 * n3a));
    for (int j = 0; j < n2; ++j)
      for (int m = 0; m < n2; ++m)
        h_wy[j + (size_t)n2 * m] =
            n2 > 1 ? ((j == 0 || j == n2 - 1) ? 1.0 : 2.0) * std::cos(M_PI * (double)(((long)j * m) % (2 * (n2 - 1))) /
                                                                       (double)(n2 - 1))
                   : 1.0;
    gz = upload(h_gz);
    Wz = upload(transposed(h_wz, n3a, n3a));  // k-major
    Wy = upload(h_wy);
    HIP_OK(hipStreamCreateWithFlags(&sw, hipStreamNonBlocking));
    HIP_OK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&ev_join2, hipEventDisableTiming));
    tmpBw = ar.get<T>(nb);
    for (int i = 0; i < 4; ++i) tmpXw.push_back(ar.get<T>(nb));
    if (const size_t gw = gb.gline_ws_elems()) {
      gwsM = ar.get<T>(gw);
      gwsS = ar.get<T>(gw);
    }
    use_main();
    dense = dense_base_grid(c->n1, c->n2);
    if (!dense) {
      gsT.check_lengths(false);
      c->gs64.check_lengths(false);
    }
    if (dense) {
      build_dense<T>(dT, ar, c->n1, c->n2);
      if (sizeof(T) != sizeof(double)) build_dense<double>(d64, ar, c->n1, c->n2);
      if (c->n2 > 1) multi = ar.get<double>((size_t)2 * kMulti * ns);
    }
    const int nf = c->p.n_freq;
    // (from the context's host copies: engines of other orders are built after hos_create)
    freq = upload(c->freq);
    pre = upload(c->h_pos_re);
    pim = upload(c->h_pos_im);
    cont = continuation_matrix();
    if (c->n2 > 1 && c->n2 > kPhiMaxPer * 256) throw std::runtime_error("n2 too large for the wavemaker kernel");
    if (c->wmk2) {
      const int n1 = c->n1;
      std::vector<double> hdgz(n3);
      for (int i3 = 0; i3 < n3; ++i3) hdgz[i3] = dgzeddz(i3 * c->delz - 1.0, c->p.igeom, c->p.d_hinge);
      dgz = upload(hdgz);
      if (!c->h_dy_re.empty() && nf) {
        dyre = upload(c->h_dy_re);
        dyim = upload(c->h_dy_im);
      }
      if (!c->h_nux.empty()) nux = upload(c->h_nux);
      if (!c->h_nuy.empty()) nuy = upload(c->h_nuy);
      Xdy = ar.get<double>((size_t)n3 * n2); Xdyt = ar.get<double>((size_t)n3 * n2);
      Xdz = ar.get<double>((size_t)n3 * n2); Xdzt = ar.get<double>((size_t)n3 * n2);
      stk6 = ar.get<double>(6 * na); stkz = ar.get<double>(6 * na); Pw = ar.get<double>(6 * na);
      // inverse paddle-grid transforms (fourier_FFTW.f90:451-492 with f_2_s_add):
      // z: REDFT10 / RODFT10 (node k, mode m), y: REDFT00 / RODFT00
      std::vector<double> hcz((size_t)n3a * n3a), hsz((size_t)n3a * n3a), hcy((size_t)n2 * n2, 0.0),
          hsy((size_t)n2 * n2, 0.0);
      for (int m = 0; m < n3a; ++m)
        for (int k = 0; k < n3a; ++k) {
          hcz[k + (size_t)n3a * m] = std::cos(M_PI * (2 * m + 1) * k / (2.0 * n3a));
          hsz[k + (size_t)n3a * m] = std::sin(M_PI * (2 * m + 1) * (k + 1) / (2.0 * n3a));
        }
      if (n2 > 1)
        for (int m = 0; m < n2; ++m)
          for (int k = 0; k < n2; ++k) {
            const double ang = M_PI * (double)(((long)k * m) % (2 * (n2 - 1))) / (double)(n2 - 1);
            hcy[k + (size_t)n2 * m] = std::cos(ang);
            const bool interior = k > 0 && k < n2 - 1 && m > 0 && m < n2 - 1;
            hsy[k + (size_t)n2 * m] = interior ? std::sin(ang) : 0.0;
          }
      // z matrices k-major; the y matrices are symmetric (functions of k m)
      Cz = upload(transposed(hcz, n3a, n3a)); Sz = upload(transposed(hsz, n3a, n3a));
      Cy = upload(hcy); Sy = upload(hsy);
      part2 = ar.get<double>((size_t)nch * 11 * ns);
      zf[0] = phit1;
      zf[1] = phiz1;
      for (int f = 2; f < 11; ++f) zf[f] = ar.get<double>(ns);
      modes8 = ar.get<double>(8 * ns); T8 = ar.get<double>(8 * ns); a1f = ar.get<double>(ns);
      phit2 = ar.get<double>(ns); phiz2 = ar.get<double>(ns); w2 = ar.get<double>(ns); tmp2 = ar.get<double>(ns);
      atot = ar.get<double>(na); dtot = ar.get<double>(na);
      if (c->wmk3) Pw3 = ar.get<double>(6 * na);
      (void)n1;
    }
    // reference extended grid (common_vars.f90:38-39) for the exact mixed-parity filter
    const int n1 = c->n1;
    const int Nd1 = ((p1 + 1) * n1) / 2, Nd2 = ((p2 + 1) * n2) / 2;
    XA = ar.get<T>(nb); YA = ar.get<T>(nb);
    XB = YB = PB = RB = T2 = nullptr;
    if (M >= 2 && c->add_pot()) {
      if (M < 3) throw std::runtime_error("the wavemaker coupling needs mHOS >= 3 (or mHOS = 1)");
      K1 = std::min(2 * n1, N1b);
      PA = ar.get<T>((size_t)K1 * 2 * n2);
      // 2-D: k_t1_partial in FP64 at every size, and in FP32 from n1 = 2048 up (a
      // general two-column product measured: 2.3-2.7x slower in FP64 at n1 = 4096; in
      // FP32 4 % faster at n1 = 1024, equal at 2048, 1.7x slower at 4096)
      // (beyond kT1MaxChunk per chunk, i.e. n1 > 8192, the dense product below)
      int chunk = std::max(32, (K1 + kT1Parts - 1) / kT1Parts);
      chunk += chunk % 2;  // chunks start at even s
      if (n2 == 1 && (sizeof(T) == 8 || n1 >= 2048) && chunk <= kT1MaxChunk) {  // the chunk sums of k_t1_partial
        if (K1 % 2) throw std::runtime_error("2-D mixed-parity product: odd K");
        t1_chunk = chunk;
        t1_parts = (K1 + t1_chunk - 1) / t1_chunk;
        RA = ar.get<T>((size_t)t1_parts * 2 * n1);
        const std::vector<double> full = sin_to_cos_matrix(Nd1, n1, K1);
        std::vector<double> half((size_t)n1 * (K1 / 2));
        for (int j = 0; j < K1 / 2; ++j)
          for (int k = 0; k < n1; ++k) half[k + (size_t)n1 * j] = full[k + (size_t)n1 * (2 * j + (k % 2 ? 0 : 1))];
        T1 = upload_t(half);
      } else {
        RA = ar.get<T>((size_t)n1 * 2 * n2);
        // the map vanishes where s + k is even: with even n1 and n2 (multiples of 4,
        // for the products' vector loads) each output parity takes half the inputs
        psplit = n1 % 4 == 0 && (n2 == 1 || n2 % 4 == 0);
        if (psplit)
          T1 = upload_t(parity_halves_kmajor(sin_to_cos_matrix(Nd1, n1, K1), n1, K1));
        else
          T1 = upload_t(transposed(sin_to_cos_matrix(Nd1, n1, K1), n1, K1));  // k-major
      }
      if (n2 > 1) {
        K2 = std::min(2 * n2, N2b);
        XB = ar.get<T>(nb); YB = ar.get<T>(nb);
        PB = ar.get<T>((size_t)2 * n1 * K2);
        RB = ar.get<T>((size_t)2 * n1 * n2);
        if (psplit)
          T2 = upload_t(parity_halves_kmajor(sin_to_cos_matrix(Nd2, n2, K2), n2, K2));
        else
          T2 = upload_t(transposed(sin_to_cos_matrix(Nd2, n2, K2), n2, K2));  // k-major
      }
    } else {
      PA = RA = T1 = nullptr;
    }
    // Mixed path, 3-D dense base grid, beach of x alone: the beach term of the
    // orders >= 2, S2F(-nu F2S(Q)), is one product Bnu Q with the n1 x n1 matrix
    // Bnu = X_fwd diag(-nu) X_inv (the y transforms of F2S and S2F are each other's
    // inverse and nu does not depend on y), formed once in FP64 on the matrix cores.
    if constexpr (kSplit) {
      if (dense && n2 > 1 && !c->h_nu_x.empty()) {
        const size_t nn = (size_t)n1 * n1;
        double *nu = nullptr, *W = nullptr, *B64 = nullptr;
        HIP_OK(hipMalloc(&nu, n1 * sizeof(double)));
        HIP_OK(hipMalloc(&W, nn * sizeof(double)));
        HIP_OK(hipMalloc(&B64, nn * sizeof(double)));
        HIP_OK(hipMemcpy(nu, c->h_nu_x.data(), n1 * sizeof(double), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_scale_rows, dim3(nblk(nn)), dim3(256), 0, c->st, n1, -1.0, (const double*)nu,
                           (const double*)d64.mx[KCOS][1], W);
        gemm(c->st, true, false, n1, n1, n1, (const double*)d64.mx[KCOS][0], n1, (const double*)W, n1, B64, n1);
        Bnu = ar.get<T>(nn);
        hipLaunchKernelGGL(k_narrow<T>, dim3(nblk(nn)), dim3(256), 0, c->st, nn, (const double*)B64, Bnu);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(c->st));
        (void)hipFree(nu);
        (void)hipFree(W);
        (void)hipFree(B64);
      }
    }
    tune();
  }

  // The extended grid's schedule (see tune_lookup): forced, from the table, or
  // by trials.  A trial is a batch of four independent inverse cos-cos
  // transforms of the grid (the shape of a right-hand side's batched
  // launches) on scratch buffers, timed with HIP events on the context stream:
  // one warm-up and three timed batches per candidate, the best of the three.
  TuneRec tune_rec{};
  void tune() {
    GridSched forced{};
    if (env_sched(forced)) {
      rebuild(forced);
      tune_rec = TuneRec{gb.sched, 0.0, 0.0, 0};
      return;
    }
    const TuneKey key{N1b, N2b, (int)sizeof(T), c->device};
    TuneRec rec{};
    if (tune_lookup(key, rec)) {
      rebuild(rec.s);
      tune_rec = rec;
      return;
    }
    if (!tune_enabled() || !gb.px.V2) {  // nothing compiled to choose between (HBM or LDS line engines)
      tune_rec = TuneRec{gb.sched, 0.0, 0.0, 0};
      return;
    }
    std::vector<GridSched> cand;
    const bool xalt = seq_alt_values(gb.px.dev2.L) > 0;
    const bool yalt = N2b > 1 && gb.py.V2 && seq_alt_values(gb.py.dev2.L) > 0;
    const std::vector<int> txs = (N2b > 1 && gb.py.V2) ? std::vector<int>{8, 4, 16} : std::vector<int>{0};
    for (int xs = 0; xs <= (xalt ? 1 : 0); ++xs)
      for (int ys = 0; ys <= (yalt ? 1 : 0); ++ys)
        for (int tx : txs) cand.push_back(GridSched{xs, ys, tx});
    hipEvent_t ev[5];
    for (auto& e : ev) HIP_OK(hipEventCreate(&e));
    double best = 1e30, def = 0.0;
    GridSched pick = cand[0];
    std::vector<GridSched> seen;
    for (const GridSched& g : cand) {
      rebuild(g);
      bool dup = false;  // the tile width may clamp two candidates to one schedule
      for (const GridSched& q : seen) dup = dup || (q == gb.sched);
      if (dup) continue;
      seen.push_back(gb.sched);
      Batch2<T, Src<T, T>, Dst<T>> b;
      T* out[4] = {tmpX[4], tmpX[5], tmpX[6], Y};
      for (int i = 0; i < 4; ++i)
        b.add(KCOS, KCOS, src_plain<T>(X, N1b, N1b, N2b), N1b, N2b, Dst<T>{out[i], N1b, 0}, tmpX[i]);
      r2r2d_batch<T>(gb, INV, b, c->st, false, gwsM);
      for (int r = 0; r < 4; ++r) {
        if (r > 0) r2r2d_batch<T>(gb, INV, b, c->st, false, gwsM);
        HIP_OK(hipEventRecord(ev[r], c->st));
      }
      r2r2d_batch<T>(gb, INV, b, c->st, false, gwsM);
      HIP_OK(hipEventRecord(ev[4], c->st));
      HIP_OK(hipEventSynchronize(ev[4]));
      double t = 1e30;
      for (int r = 0; r < 4; ++r) {
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, ev[r], ev[r + 1]));
        t = std::min(t, 1000.0 * ms);
      }
      if (seen.size() == 1) def = t;
      if (t < best) {
        best = t;
        pick = gb.sched;
      }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    rec = tune_store(key, TuneRec{pick, def, best, (int)seen.size()});
    rebuild(rec.s);
    tune_rec = rec;
  }
  void schedule(int* s8, double* us2) const override {
    s8[0] = gb.px.V2 ? gb.px.dev2.L : gb.px.dev.Lorig;
    s8[1] = gb.px.V2;
    s8[2] = gb.sched.xs;
    s8[3] = N2b > 1 ? (gb.py.V2 ? gb.py.dev2.L : gb.py.dev.Lorig) : 0;
    s8[4] = N2b > 1 ? gb.py.V2 : 0;
    s8[5] = gb.sched.ys;
    s8[6] = N2b > 1 ? gb.TX : 0;
    s8[7] = tune_rec.trials;
    us2[0] = tune_rec.us_default;
    us2[1] = tune_rec.us_best;
  }
  // rebuilds the extended grid's plans for schedule g (the stream is idle)
  void rebuild(const GridSched& g) {
    if (g == gb.sched) return;
    HIP_OK(hipStreamSynchronize(c->st));
    gb.release();
    gb.build(N1b, N2b, g);
    gb.check_lengths(true);
  }
  ~Engine() override {
    if (sw) (void)hipStreamSynchronize(sw);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (ev_join2) (void)hipEventDestroy(ev_join2);
    if (sw) (void)hipStreamDestroy(sw);
  }

  T* upload_t(const std::vector<double>& v) {
    std::vector<T> h(v.begin(), v.end());
    T* d = c->ar.get<T>(h.size());
    HIP_OK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
  }

  double* upload(const std::vector<double>& v) {
    double* d = c->ar.get<double>(v.size());
    if (!v.empty()) HIP_OK(hipMemcpy(d, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    return d;
  }

  // ---- transform helpers ------------------------------------------------
  template <class S, class D>
  void big(int kx, int ky, int dir, const S& s, int o1, int o2, const D& d) {
    r2r2d<T>(gb, kx, ky, dir, s, o1, o2, d, ctmp, cs, cgws);
    c->stats.n_big++;
  }
  // independent extended-grid transforms in one launch per pass
  template <class S, class D>
  void bigB(int dir, const Batch2<T, S, D>& b, bool yfirst = false) {
    r2r2d_batch<T>(gb, dir, b, cs, yfirst, cgws);
    c->stats.n_big += b.n;
  }
  // a forward cos-cos transform to truncated extents, y pass first: the y pass runs
  // on all N1b columns to o2 rows, then o2 x lines instead of N2b (fewer FFT points
  // when o2 < N2b)
  template <class S, class D>
  void bigF(const S& s, int o1, int o2, const D& d) {
    Batch2<T, S, D> b;
    b.add(KCOS, KCOS, s, o1, o2, d, ctmp);
    bigB(FWD, b, true);
  }
  // base-grid transforms use tmpB (large enough) as their intermediate
  template <class S, class D>
  void base(int kx, int ky, int dir, const S& s, int o1, int o2, const D& d) {
    if (dense)
      dense2d<T>(dT, cs, kx, ky, dir, s, o1, o2, d);
    else
      r2r2d<T>(gsT, kx, ky, dir, s, o1, o2, d, ctmp, cs);
    c->stats.n_base++;
  }
  template <class S, class D>
  void base64(int kx, int ky, int dir, const S& s, int o1, int o2, const D& d) {
    if (dense) {
      if constexpr (sizeof(T) == sizeof(double))
        dense2d<double>(dT, cs, kx, ky, dir, s, o1, o2, d);
      else
        dense2d<double>(d64, cs, kx, ky, dir, s, o1, o2, d);
    } else {
      r2r2d<double>(c->gs64, kx, ky, dir, s, o1, o2, d, c->phys_tmp, cs);
    }
    c->stats.n_base++;
  }

  void paddle_top(double* x, double* xt) override {
    const int n2 = c->n2, n3 = c->n3;
    HIP_OK(hipDeviceSynchronize());
    HIP_OK(hipMemcpy2D(x, sizeof(double), Xp + (n3 - 1), (size_t)n3 * sizeof(double), sizeof(double), n2,
                       hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy2D(xt, sizeof(double), Xpt + (n3 - 1), (size_t)n3 * sizeof(double), sizeof(double), n2,
                       hipMemcpyDeviceToHost));
  }

  // out[f] = the cos-cos base transform of src[f], f < k (dense 3-D grid, stream cs)
  template <typename U>
  void dense_multi(DenseGrid<U>& g, int dir, int k, const Src<U, double>* src, U* const* out) {
    const int n1 = c->n1, n2 = c->n2, di = (dir == FWD) ? 0 : 1;
    if (k > kMulti || !multi) throw std::runtime_error("dense_multi: no scratch for this batch");
    U* stage = reinterpret_cast<U*>(multi);
    U* mid = stage + (size_t)kMulti * ns;
    if (g.fold) {
      // folded sources and the two half products (see k_dense_fold)
      const int h = n1 / 2;
      static_assert(kMulti <= 4, "FoldSet holds four sources");
      FoldSet<Src<U, double>> fs{};
      for (int f = 0; f < k; ++f) fs.src[f] = src[f];
      hipLaunchKernelGGL((k_dense_fold_multi<U, Src<U, double>>), dim3(nblk((size_t)h * n2), k), dim3(256), 0, cs, n1,
                         n2, 1, fs, stage);
      hosmm::mm<U>(cs, true, true, h, k * n2, h, (const U*)g.mxh[KCOS][di], h, (long)h * h, (const U*)stage, n1, h,
                   mid, n1, 1, 2, 0, 2);
    } else {
      for (int f = 0; f < k; ++f)
        hipLaunchKernelGGL((k_dense_in<U, Src<U, double>>), dim3(nblk(ns)), dim3(256), 0, cs, n1, n2, src[f],
                           stage + (size_t)f * ns);
      gemm(cs, true, true, n1, k * n2, n1, (const U*)g.mx[KCOS][di], n1, (const U*)stage, n1, mid, n1);
    }
    // the y products write each run of outputs that lie ns apart in one launch
    for (int f0 = 0; f0 < k;) {
      int f1 = f0 + 1;
      while (f1 < k && out[f1] == out[f0] + (size_t)(f1 - f0) * ns) ++f1;
      gemm_yb(cs, n1, n2, mid + (size_t)f0 * ns, (long)ns, g.my[KCOS][di], out[f0], f1 - f0);
      f0 = f1;
    }
    c->stats.n_base += k;
  }

  void base_cc64(int dir, const double* in, double* out) override {
    use_main();
    base64(KCOS, KCOS, dir, src_plain<double, double>(in, c->n1, c->n1, c->n2), c->n1, c->n2,
           Dst<double>{out, c->n1, 0});
  }

  void rhs(double t, double dt, const double* G1, const double* G2, const double* a1, const double* e1,
           const double* aad, double* kG1, double* kG2, double* ka1, double* ke1, double* kad1,
           double* kadr, const S2* s2) override {
    hos_ctx& C = *c;
    hipStream_t st = C.st;
    const int n1 = C.n1, n2 = C.n2;
    const double pix = C.pix, piy = C.piy;
    const int ldb = N1b;
    C.stats.n_rhs++;
    // --- GtoF, /omega, volume fix (uses the previous wavemaker evaluation) ---
    const unsigned nbs = nblk(ns);
    hipLaunchKernelGGL(k_gtof_in<T>, dim3(nbs), dim3(256), 0, st, n1, n2, pix, piy, dt, G1, G2, eta_hat,
                       phis_hat, C.partial, phis64, (const double*)C.dsc_active, C.cur_stage);
    hipLaunchKernelGGL(k_volfix<T>, dim3(1), dim3(256), 0, st, (int)nbs, C.partial, eta_hat, (double*)nullptr,
                       C.volmode(), C.last_alpha, C.last_pos_top, C.p.xlen, C.p.d_hinge, (const double*)C.dsc_active);
    // --- physical eta on the base grid ---
    base(KCOS, KCOS, INV, src_plain<T>(eta_hat, n1, n1, n2), n1, n2, Dst<T>{eta_phys, n1, 0});
    // --- the wavemaker chain runs on the side stream from here: it needs only the
    // base-grid eta (and the stage inputs); it is enqueued after the main stream's
    // work up to the join, so that the main stream does not wait for the host
    // while the host enqueues the chain ---
    HIP_OK(hipEventRecord(ev_fork, st));
    // --- phisxy_etaxy (resol_HOS.f90:924-1078): the extended-grid eta here, with the
    // first HOS level's one transform (kth1 phis).  The slopes and phis gradients
    // are read only after the join (the free-surface products): in 3-D they go to
    // the side stream, which has room beside the HOS levels; a 2-D tank, bound by
    // kernel latency, keeps them in this launch (one launch fewer per pass) ---
    const bool slopes_side = n2 > 1;
    auto slopes = [&](Batch2<T, Src<T, T>, Dst<T>>& b, std::vector<T*>& tx) {
      b.add(KSIN, KCOS, src_mul<T>(eta_hat, n1, n1, n2, MUL_NEG_KX, 0, pix, piy), N1b, N2b, Dst<T>{etax, ldb, 0},
            tx[b.n]);
      b.add(KSIN, KCOS, src_mul<T>(phis_hat, n1, n1, n2, MUL_NEG_KX, 0, pix, piy), N1b, N2b, Dst<T>{phisx, ldb, 0},
            tx[b.n]);
      if (n2 > 1) {
        b.add(KCOS, KSIN, src_mul<T>(eta_hat, n1, n1, n2, MUL_NEG_KY, 0, pix, piy), N1b, N2b, Dst<T>{etay, ldb, 0},
              tx[b.n]);
        b.add(KCOS, KSIN, src_mul<T>(phis_hat, n1, n1, n2, MUL_NEG_KY, 0, pix, piy), N1b, N2b,
              Dst<T>{phisy, ldb, 0}, tx[b.n]);
      }
    };
    {
      Batch2<T, Src<T, T>, Dst<T>> b;
      b.add(KCOS, KCOS, src_plain<T>(eta_hat, n1, n1, n2), N1b, N2b, Dst<T>{eta_ext ? eta_ext : tmpB2(), ldb, 0},
            tmpX[0]);
      b.add(KCOS, KCOS, src_mul<T>(phis_hat, n1, n1, n2, MUL_KTH, 1, pix, piy), N1b, N2b, Dst<T>{lvl[0], ldb, 0},
            tmpX[1]);
      if (!slopes_side) slopes(b, tmpX);
      bigB(INV, b);
    }
    // --- HOSphis_modes recursion (resol_HOS.f90:720-782) ---
    for (int iH = 1; iH <= M; ++iH) {
      // the level's transforms in one launch (x passes first: 3 % faster in FP64
      // at 1024 x 256 than y first), then the accumulation in the reference's
      // order (k_hos_level)
      Batch2<T, Src<T, T>, Dst<T>> b;
      PtrSet<T> lv{};
      for (int j = iH; j >= 1 && iH > 1; --j) {
        const int j2 = iH - j + 1;
        lv.p[b.n] = lvl[b.n];
        if (j2 == 1) {
          b.add(KCOS, KCOS, src_mul<T>(phis_hat, n1, n1, n2, MUL_KTH, j, pix, piy), N1b, N2b,
                Dst<T>{lvl[b.n], ldb, 0}, tmpX[b.n]);
        } else {
          const bool trunc = (j2 == M);
          b.add(KCOS, KCOS,
                src_mul<T>(aHOS[j2], trunc ? n1 : ldb, trunc ? n1 : N1b, trunc ? n2 : N2b, MUL_KTH, j, pix, piy),
                N1b, N2b, Dst<T>{lvl[b.n], ldb, 0}, tmpX[b.n]);
        }
      }
      if (iH == 1)
        lv.p[0] = lvl[0];  // transformed with the slopes
      else
        bigB(INV, b);
      {
        bool al = (uintptr_t)eta_ext % 16 == 0 && (uintptr_t)phizm[iH] % 16 == 0;
        if (iH < M) al = al && (uintptr_t)phim % 16 == 0;
        for (int q = 0; q < iH; ++q) al = al && (uintptr_t)lv.p[q] % 16 == 0;
        constexpr int W = 16 / sizeof(T);
        // (FP32: four points per thread; FP64 measured no faster with two)
        if (al && sizeof(T) == 4)
          hipLaunchKernelGGL(k_hos_level_vec<T>, dim3(nblk((nb + W - 1) / W)), dim3(256), 0, st, nb, iH, M, lv,
                             (const T*)eta_ext, phim, phizm[iH]);
        else
          hipLaunchKernelGGL(k_hos_level<T>, dim3(nblk(nb)), dim3(256), 0, st, nb, iH, M, lv, (const T*)eta_ext, phim,
                             phizm[iH]);
      }
      if (iH < M) {
        if (iH + 1 == M)
          bigF(src_plain<T>(phim, ldb, N1b, N2b), n1, n2, Dst<T>{aHOS[iH + 1], n1, 0});
        else
          big(KCOS, KCOS, FWD, src_plain<T>(phim, ldb, N1b, N2b), N1b, N2b, Dst<T>{aHOS[iH + 1], ldb, 0});
      }
    }
    if (C.feat_mode == 2 || (C.feat_mode == 1 && C.cur_stage == 0)) record_features();
    // phiz = filter_ext(sum_j phizm(j)) -> modal P (orders >= 2 on the mixed path)
    PtrSet<T> zz{};
    for (int j = 1; j <= M; ++j) zz.p[j] = phizm[j];
    // (the order sum goes to lvl[0], free after the last level)
    // (with M >= 2 k_fsbc forms the sum beside the products, which read the same fields)
    T* zsum = lvl[0];
    const int j0 = kSplit ? 2 : 1;
    if (M < 2)
      hipLaunchKernelGGL(k_sum_orders<T>, dim3(nblk(nb)), dim3(256), 0, st, nb, M, j0, zz, zsum);
    // --- the wavemaker chain (solveHOS_SWEET_1 + phiadd_SL), then the join ---
    HIP_OK(hipStreamWaitEvent(sw, ev_fork, 0));
    {
      struct Side {
        Engine* e;
        ~Side() { e->use_main(); }
      } side{this};
      use_side();
      if (slopes_side) {
        Batch2<T, Src<T, T>, Dst<T>> b;
        slopes(b, *ctmpX);
        bigB(INV, b);
      }
      fo_late = false;
      wavemaker(t, a1, e1, aad, ka1, ke1, kad1, kadr, s2);
    }
    HIP_OK(hipEventRecord(ev_join, sw));
    HIP_OK(hipStreamWaitEvent(st, ev_join, 0));
    // the first-order conditions continue on the wavemaker stream beside the
    // free-surface products; the main stream waits for them at the end
    if (fo_late) {
      struct Side {
        Engine* e;
        ~Side() { e->use_main(); }
      } side{this};
      use_side();
      first_order_rest(e1, a1, ka1, ke1);
      HIP_OK(hipEventRecord(ev_join2, sw));
    }
    // --- free-surface products, filtered to the base modes; P, the products and
    // the mixed-parity spectra in one batch (y passes first: the x pass writes the
    // spectra's rows and columns in the parity order of the sine-to-cosine products) ---
    {
      Batch2<T, Src<T, T>, DstP<T>> b;
      b.add(KCOS, KCOS, src_plain<T>(zsum, ldb, N1b, N2b), n1, n2, DstP<T>{P, n1, 0}, tmpX[0]);
      if (M >= 2) {
        // (in a 2-D tank etay, phisy and phiya hold zeros from allocation on: the
        // arena clears every buffer and only the n2 > 1 transforms write them)
        hipLaunchKernelGGL(k_fsbc<T>, dim3(nblk(nb)), dim3(256), 0, st, nb, M, zz, phisx, phisy, etax, etay,
                           phixa, phiya, phiza, X, Y, XA, YA, XB, YB, j0, zsum);
        b.add(KCOS, KCOS, src_plain<T>(X, ldb, N1b, N2b), n1, n2, DstP<T>{D1, n1, 0}, tmpX[1]);
        b.add(KCOS, KCOS, src_plain<T>(Y, ldb, N1b, N2b), n1, n2, DstP<T>{D2, n1, 0}, tmpX[2]);
        if (C.add_pot()) mixed_parity_spectra(b);
      }
      bigB(FWD, b, true);
    }
    if (M >= 2) {
      if (C.add_pot())
        mixed_parity_finish();  // with Q = D1 + P
      else
        hipLaunchKernelGGL(k_add3<T>, dim3(nbs), dim3(256), 0, st, ns, (const T*)D1, (const T*)P, Q);
    } else {
      HIP_OK(hipMemcpyAsync(Q, P, ns * sizeof(T), hipMemcpyDeviceToDevice, st));
    }
    // dphin and the final base-grid RHS
    if (!Bnu) base(KCOS, KCOS, INV, src_plain<T>(Q, n1, n1, n2), n1, n2, Dst<T>{tmpS, n1, 0});
    if constexpr (kSplit) {
      // beach term of the orders >= 2 in T; the first-order modes Zh64 come from the
      // wavemaker stream
      if (Bnu) {
        gemm(cs, false, true, n1, n2, n1, (const T*)Bnu, n1, (const T*)Q, n1, Zh, n1);
        C.stats.n_base += 2;  // the two base transforms it stands for
      } else {
        hipLaunchKernelGGL(k_beach<T>, dim3(nbs), dim3(256), 0, st, ns, (const T*)tmpS, (const double*)C.nu_d, Zp);
        base(KCOS, KCOS, FWD, src_plain<T>(Zp, n1, n1, n2), n1, n2, Dst<T>{Zh, n1, 0});
      }
      if (fo_late) HIP_OK(hipStreamWaitEvent(st, ev_join2, 0));  // Zh64, ka1, ke1
      hipLaunchKernelGGL((k_expon_out<T, double>), dim3(nbs), dim3(256), 0, st, n1, n2, pix, piy, dt,
                         (const T*)(M >= 2 ? D1 : nullptr), (const T*)P, (const T*)nullptr,
                         (const double*)phizh64, (const T*)(M >= 2 ? D2 : nullptr), (const T*)Zh,
                         (const double*)Zh64, (const double*)C.filt_fs, kG1, kG2, (const double*)C.dsc_active, C.cur_stage);
    } else {
      hipLaunchKernelGGL(k_dphin<T>, dim3(nbs), dim3(256), 0, st, ns, (const T*)tmpS, (const double*)phiz_add,
                         (const double*)phit_add, (const double*)C.nu_d, Zp);
      base(KCOS, KCOS, FWD, src_plain<T>(Zp, n1, n1, n2), n1, n2, Dst<T>{Zh, n1, 0});
      if (fo_late) HIP_OK(hipStreamWaitEvent(st, ev_join2, 0));  // ka1, ke1
      hipLaunchKernelGGL((k_expon_out<T, T>), dim3(nbs), dim3(256), 0, st, n1, n2, pix, piy, dt,
                         (const T*)(M >= 2 ? D1 : nullptr), (const T*)P, (const T*)phis_hat, (const T*)phizh,
                         (const T*)(M >= 2 ? D2 : nullptr), (const T*)Zh, (const double*)nullptr,
                         (const double*)C.filt_fs, kG1, kG2, (const double*)C.dsc_active, C.cur_stage);
    }
  }

  // band energies of this evaluation's orders 1..M (k_band_energy) into the
  // context's feature record
  void record_features() {
    hos_ctx& C = *c;
    OrderFields<T> f{};
    f.a[1] = phis_hat;
    f.ld[1] = C.n1;
    for (int j = 2; j <= M && j <= kFeatOrders; ++j) {
      f.a[j] = aHOS[j];
      f.ld[j] = (j == M) ? C.n1 : N1b;  // the last order is stored truncated
    }
    int slot = 0;
    if (C.feat_mode == 2) slot = C.feat_n++ % C.feat_cap;
    const int nord = std::min(M, kFeatOrders);
    hipLaunchKernelGGL(k_band_energy<T>, dim3(kFeatGrid), dim3(256), 0, C.st, C.n1, C.n2, C.pix, C.piy, nord, f,
                       (const T*)eta_hat, C.feat_part);
    hipLaunchKernelGGL(k_band_reduce, dim3((kFeatRec + 255) / 256), dim3(256), 0, C.st, kFeatGrid, M,
                       (const double*)C.feat_part, C.feat + (size_t)slot * kFeatRec);
    HIP_OK(hipGetLastError());
  }

  // filter_ext of the mixed-parity products on the reference grid: exact sine
  // spectra on the padded grid (added to the caller's batch), then the
  // reference grid's sine-to-cosine map (mixed_parity_finish).
  void mixed_parity_spectra(Batch2<T, Src<T, T>, DstP<T>>& b) {
    hos_ctx& C = *c;
    const int n1 = C.n1, n2 = C.n2, ldb = N1b;
    // with psplit the sine rows of PA and the sine columns of PB go in parity order
    const int ha = psplit ? K1 / 2 : 0, hb = psplit ? K2 / 2 : 0;
    b.add(KSIN, KCOS, src_plain<T>(XA, ldb, N1b, N2b), K1, n2, DstP<T>{PA, K1, 0, ha, 0}, tmpX[b.n]);
    b.add(KSIN, KCOS, src_plain<T>(YA, ldb, N1b, N2b), K1, n2, DstP<T>{PA + (size_t)K1 * n2, K1, 0, ha, 0},
          tmpX[b.n]);
    if (n2 > 1) {
      b.add(KCOS, KSIN, src_plain<T>(XB, ldb, N1b, N2b), n1, K2, DstP<T>{PB, 2 * n1, 0, 0, hb}, tmpX[b.n]);
      b.add(KCOS, KSIN, src_plain<T>(YB, ldb, N1b, N2b), n1, K2, DstP<T>{PB + n1, 2 * n1, 0, 0, hb}, tmpX[b.n]);
    }
  }
  // (and Q = D1 + P)
  void mixed_parity_finish() {
    hos_ctx& C = *c;
    const int n1 = C.n1, n2 = C.n2;
    if (t1_parts > 0) {
      hipLaunchKernelGGL(k_t1_partial<T>, dim3(nblk(n1), t1_parts), dim3(256), 0, cs, n1, K1, t1_chunk,
                         (const T*)T1, (const T*)PA, RA);
    } else if (psplit) {
      // each parity class of the outputs from the other class of the inputs: two
      // half-size products per map in one launch (batch 0: even outputs from odd
      // inputs, batch 1: odd from even), rows of RA and columns of RB in that order
      const int h1 = n1 / 2, hk1 = K1 / 2;
      hosmm::mm<T>(cs, true, true, h1, 2 * n2, hk1, (const T*)T1, hk1, (long)hk1 * h1, (const T*)PA, K1, hk1, RA,
                   n1, h1, 2);
      if (n2 > 1) {
        const int h2 = n2 / 2, hk2 = K2 / 2;
        hosmm::mm<T>(cs, false, true, 2 * n1, h2, hk2, (const T*)PB, 2 * n1, (long)2 * n1 * hk2, (const T*)T2, hk2,
                     (long)hk2 * h2, RB, 2 * n1, (long)2 * n1 * h2, 2);
      }
    } else {
      gemm(cs, true, true, n1, 2 * n2, K1, (const T*)T1, K1, (const T*)PA, K1, RA, n1);
      if (n2 > 1) gemm(cs, false, true, 2 * n1, n2, K2, (const T*)PB, 2 * n1, (const T*)T2, K2, RB, 2 * n1);
    }
    // (64-thread blocks: a 2-D tank's few outputs spread over more CUs)
    hipLaunchKernelGGL(k_add_mixed<T>, dim3(nblk(ns, 64)), dim3(64), 0, cs, n1, n2, (const T*)RA,
                       (const T*)(n2 > 1 ? RB : nullptr), D1, D2, (const T*)P, Q, t1_parts, psplit ? 1 : 0);
  }

  T* tmpB2() { return X; }

  // ---- SWENSE output on the reference's extended grid ----------------------------------
  // The H2 re-expansions of HOSvel2_SWEET keep content beyond the base modes that depends
  // on the extended grid (aliasing of the reference's own products), so they run on the
  // reference's grid Nd = (M+1) n / 2 exactly, in FP64, with dense transform matrices on
  // the matrix cores (k_mm); every field follows the reference's sequence.
  struct SwGrid {
    int N1 = 0, N2 = 1;
    DenseGrid<double> g;
    std::vector<double*> ep, ep2, phizm, phizm2, aH[4];
    double *etax, *etay, *phisx, *phisy, *g2, *pMm1, *pMm2, *p2Mm2, *phiz2, *ge2p2, *t1, *t2, *phizi, *pm;
    double *vit[4], *pa[4], *pb[4];
    double *b0, *b1, *b2, *b3, *b4, *d11;  // base-grid work (n1 x n2) and the mean mode
    double* out[5];
  };
  std::unique_ptr<SwGrid> swg;

  void swense(int meta2, double** out_dev) override {
    if constexpr (sizeof(T) != sizeof(double)) {
      (void)meta2;
      (void)out_dev;
      throw std::runtime_error("the SWENSE output runs in FP64 contexts (precision 0)");
    } else {
      hos_ctx& C = *c;
      use_main();
      hipStream_t st = cs;
      const int n1 = C.n1, n2 = C.n2;
      const size_t nsb = (size_t)n1 * n2;
      const double pix = C.pix, piy = C.piy;
      if (!swg) {
        auto g = std::make_unique<SwGrid>();
        g->N1 = ((M + 1) * n1) / 2;
        g->N2 = n2 > 1 ? ((M + 1) * n2) / 2 : 1;
        if (g->N1 > 16384 || g->N2 > 16384) throw std::runtime_error("SWENSE output: extended grid beyond 16384");
        build_dense(g->g, C.ar, g->N1, g->N2);
        const size_t nd = (size_t)g->N1 * g->N2;
        auto F = [&]() { return C.ar.get<double>(nd); };
        for (int j = 0; j <= M; ++j) {
          g->ep.push_back(F());
          g->ep2.push_back(F());
          g->phizm.push_back(F());
          g->phizm2.push_back(F());
          for (int f = 0; f < 4; ++f) g->aH[f].push_back(F());
        }
        g->etax = F(); g->etay = F(); g->phisx = F(); g->phisy = F(); g->g2 = F();
        g->pMm1 = F(); g->pMm2 = F(); g->p2Mm2 = F(); g->phiz2 = F(); g->ge2p2 = F();
        g->t1 = F(); g->t2 = F(); g->phizi = F(); g->pm = F();
        for (int f = 0; f < 4; ++f) {
          g->vit[f] = F();
          g->pa[f] = F();
          g->pb[f] = F();
        }
        g->b0 = C.ar.get<double>(nsb); g->b1 = C.ar.get<double>(nsb); g->b2 = C.ar.get<double>(nsb);
        g->b3 = C.ar.get<double>(nsb); g->b4 = C.ar.get<double>(nsb);
        g->d11 = C.ar.get<double>(1);
        for (int f = 0; f < 5; ++f) g->out[f] = C.ar.get<double>(nsb);
        swg = std::move(g);
      }
      SwGrid& G = *swg;
      const int N1 = G.N1, N2 = G.N2;
      const size_t nd = (size_t)N1 * N2;
      const int KX[4] = {KSIN, KCOS, KCOS, KCOS}, KY[4] = {KCOS, KSIN, KCOS, KCOS};
      auto ew = [&](double* out, double s1, const double* A, const double* B, double t1, const double* Cc,
                    size_t n) {
        hipLaunchKernelGGL(k_ew, dim3(nblk(n)), dim3(256), 0, st, n, s1, A, B, t1, Cc, out);
      };
      auto zero = [&](double* p, size_t n) { ew(p, 0.0, p, nullptr, 0.0, nullptr, n); };
      // dense transforms on the reference grid
      auto T2 = [&](int kx, int ky, int dir, const Src<double, double>& src, int o1, int o2, double* out) {
        dense2d<double>(G.g, st, kx, ky, dir, src, o1, o2, Dst<double>{out, N1, 0});
      };
      auto ext_src = [&](const double* a, int ld, int nx, int ny, int mul = MUL_NONE, int jpow = 0) {
        return src_mul<double, double>(a, ld, nx, ny, mul, jpow, pix, piy);
      };
      // dealias (dealiasing.f90:39-92) in place
      auto dealias = [&](int order, double* x, int kx, int ky) {
        const bool ax = M != 1 && (order - 1) % (M - 1) == 0;
        const int p2 = n2 > 1 ? M : 1;
        const bool ay = n2 > 1 && p2 != 1 && (order - 1) % (p2 - 1) == 0;
        if (!ax && !ay) return;
        const int o1 = ax ? n1 : N1, o2 = ay ? n2 : N2;
        T2(kx, ky, FWD, ext_src(x, N1, N1, N2), o1, o2, G.pm);
        T2(kx, ky, INV, ext_src(G.pm, N1, o1, o2), N1, N2, x);
      };
      // filter_ext (dealiasing.f90:96-134): base-grid physical field
      auto filter_ext = [&](const double* x, int kx, int ky, double* out) {
        T2(kx, ky, FWD, ext_src(x, N1, N1, N2), n1, n2, G.pm);
        base64(kx, ky, INV, src_plain<double, double>(G.pm, N1, n1, n2), n1, n2, Dst<double>{out, n1, 0});
      };
      // --- init_velocity_comp (velocities.f90:501-784): the right-hand side at the
      // committed state has just run; its time derivatives
      const double* filtp = C.filt_fs;
      hipLaunchKernelGGL(k_mean_mode, dim3(1), dim3(64), 0, st, (const double*)(M >= 2 ? D2 : nullptr),
                         (const double*)Zh, (const double*)nullptr, filtp, G.d11);
      double* fst = G.out[4];
      // (init_velocity_comp's eta3, phis3 are the committed state's modes, :579-580)
      hipLaunchKernelGGL(k_swense_dt, dim3(nblk(nsb)), dim3(256), 0, st, n1, n2, pix, piy,
                         (const double*)C.kG1[0], (const double*)C.kG2[0], (const double*)C.eta_hat,
                         (const double*)C.phis_hat, (const double*)G.d11, fst, G.b0);
      double* dphis3 = G.b1;  // physical d phis / dt
      base64(KCOS, KCOS, INV, src_plain<double, double>(G.b0, n1, n1, n2), n1, n2, Dst<double>{dphis3, n1, 0});
      // phisxy_etaxy on the reference grid (resol_HOS.f90:924-1078)
      T2(KCOS, KCOS, INV, ext_src(eta_hat, n1, n1, n2), N1, N2, G.ep[1]);
      for (int j = 1; j < M; ++j) {
        ew(G.ep[j + 1], 1.0 / (j + 1), G.ep[j], G.ep[1], 0.0, nullptr, nd);
        dealias(j + 1, G.ep[j + 1], KCOS, KCOS);
      }
      T2(KSIN, KCOS, INV, ext_src(eta_hat, n1, n1, n2, MUL_NEG_KX), N1, N2, G.etax);
      T2(KSIN, KCOS, INV, ext_src(phis_hat, n1, n1, n2, MUL_NEG_KX), N1, N2, G.phisx);
      if (n2 > 1) {
        T2(KCOS, KSIN, INV, ext_src(eta_hat, n1, n1, n2, MUL_NEG_KY), N1, N2, G.etay);
        T2(KCOS, KSIN, INV, ext_src(phis_hat, n1, n1, n2, MUL_NEG_KY), N1, N2, G.phisy);
      } else {
        zero(G.etay, nd);
        zero(G.phisy, nd);
      }
      ew(G.g2, 1.0, G.etax, G.etax, 0.0, nullptr, nd);
      ew(G.g2, 1.0, G.etay, G.etay, 1.0, G.g2, nd);
      dealias(2, G.g2, KCOS, KCOS);
      // HOSphis_modes on the reference grid (resol_HOS.f90:665-915): phizm by order
      hipLaunchKernelGGL(k_extend, dim3(nblk(nd)), dim3(256), 0, st, n1, n2, N1, N2, (const double*)phis_hat,
                         G.aH[0][1]);
      for (int iH = 1; iH <= M; ++iH) {
        zero(G.t1, nd);
        zero(G.phizm[iH], nd);
        for (int j = iH; j >= 1; --j) {
          const int jj = (j == 1) ? 1 : j, j2 = (j == 1) ? iH : iH - (j - 1);
          if (j == 1 && iH > 1) continue;  // the order-1 term below
          T2(KCOS, KCOS, INV, ext_src(G.aH[0][j2], N1, N1, N2, MUL_KTH, jj), N1, N2, G.phizi);
          ew(G.t1, -1.0, G.phizi, G.ep[jj], 1.0, G.t1, nd);
          if (jj > 1)
            ew(G.phizm[iH], 1.0, G.phizi, G.ep[jj - 1], 1.0, G.phizm[iH], nd);
          else
            ew(G.phizm[iH], 1.0, G.phizi, nullptr, 1.0, G.phizm[iH], nd);
        }
        if (iH > 1) {
          T2(KCOS, KCOS, INV, ext_src(G.aH[0][iH], N1, N1, N2, MUL_KTH, 1), N1, N2, G.phizi);
          ew(G.t1, -1.0, G.phizi, G.ep[1], 1.0, G.t1, nd);
          ew(G.phizm[iH], 1.0, G.phizi, nullptr, 1.0, G.phizm[iH], nd);
          dealias(iH, G.phizm[iH], KCOS, KCOS);
        }
        if (iH < M) {
          dealias(iH + 1, G.t1, KCOS, KCOS);
          T2(KCOS, KCOS, FWD, ext_src(G.t1, N1, N1, N2), N1, N2, G.aH[0][iH + 1]);
        }
      }
      // phizMm2, phizMm1, phiz2Mm2 (resol_HOS.f90:785-823)
      zero(G.pMm2, nd);
      for (int j = 1; j <= M - 2; ++j) {
        ew(G.pMm2, 1.0, G.phizm[j], nullptr, 1.0, G.pMm2, nd);
        ew(G.phizm2[j], 1.0, G.phizm[j], nullptr, 0.0, nullptr, nd);
      }
      if (M >= 2)
        ew(G.pMm1, 1.0, G.phizm[M - 1], nullptr, 1.0, G.pMm2, nd);
      else
        zero(G.pMm1, nd);
      zero(G.p2Mm2, nd);
      if (M >= 4) {
        const int m3d2 = std::max(1, (M - 3) / 2), d2m1 = std::max(1, M / 2 - 1);
        for (int j = 1; j <= m3d2; ++j) {
          ew(G.p2Mm2, 2.0, G.phizm2[j], G.phizm2[M - 2 - j], 1.0, G.p2Mm2, nd);
          ew(G.phizm2[j + 1], 1.0, G.phizm2[j + 1], nullptr, 1.0, G.phizm2[j], nd);
        }
        ew(G.p2Mm2, 1.0, G.phizm2[d2m1], G.phizm2[d2m1], 1.0, G.p2Mm2, nd);
      }
      // phiz2 and gradeta2 phiz2 with their dealiasing (resol_HOS.f90:851-905)
      zero(G.phiz2, nd);
      zero(G.ge2p2, nd);
      for (int j = 1; j <= (M - 2) / 2; ++j) {
        ew(G.t2, 1.0, G.phizm[j], G.phizm[j], 0.0, nullptr, nd);
        dealias(2 * j, G.t2, KCOS, KCOS);
        ew(G.phiz2, 1.0, G.t2, nullptr, 1.0, G.phiz2, nd);
        ew(G.t2, 1.0, G.g2, G.t2, 0.0, nullptr, nd);
        dealias(2 * j + 2, G.t2, KCOS, KCOS);
        ew(G.ge2p2, 1.0, G.t2, nullptr, 1.0, G.ge2p2, nd);
      }
      for (int j = 2; j <= M - 3; ++j)
        for (int q = 1; q <= std::min(M - 2 - j, j - 1); ++q) {
          ew(G.t2, 2.0, G.phizm[j], G.phizm[q], 0.0, nullptr, nd);
          dealias(j + q, G.t2, KCOS, KCOS);
          ew(G.phiz2, 1.0, G.t2, nullptr, 1.0, G.phiz2, nd);
          ew(G.t2, 1.0, G.g2, G.t2, 0.0, nullptr, nd);
          dealias(j + q + 2, G.t2, KCOS, KCOS);
          ew(G.ge2p2, 1.0, G.t2, nullptr, 1.0, G.ge2p2, nd);
        }
      if (M >= 2) {
        ew(G.t2, 1.0, G.phizm[M / 2], G.phizm[M / 2], 0.0, nullptr, nd);
        dealias(2 * (M / 2), G.t2, KCOS, KCOS);
        ew(G.phiz2, 1.0, G.t2, nullptr, 1.0, G.phiz2, nd);
      }
      for (int j = 2; j <= M - 1; ++j)
        for (int q = std::max(0, std::min(M - 2 - j, j - 1)) + 1; q <= std::min(M - j, j - 1); ++q) {
          ew(G.t2, 2.0, G.phizm[j], G.phizm[q], 0.0, nullptr, nd);
          dealias(j + q, G.t2, KCOS, KCOS);
          ew(G.phiz2, 1.0, G.t2, nullptr, 1.0, G.phiz2, nd);
        }
      // free-surface velocities and phi_t at z = eta (velocities.f90:656-773); the
      // etapmext of init_velocity_comp's right-hand side start the H2 levels
      double* pze = G.aH[1][1];  // extended phiz (the base modal P)
      double* pxe = G.aH[2][1];  // extended grad phi_add
      double* pye = G.aH[3][1];
      double* pza = G.vit[3];
      T2(KCOS, KCOS, INV, ext_src(P, n1, n1, n2), N1, N2, pze);
      base64(KCOS, KCOS, FWD, src_plain<double, double>(phix_add, n1, n1, n2), n1, n2, Dst<double>{G.b2, n1, 0});
      T2(KCOS, KCOS, INV, ext_src(G.b2, n1, n1, n2), N1, N2, pxe);
      if (n2 > 1) {
        base64(KCOS, KCOS, FWD, src_plain<double, double>(phiy_add, n1, n1, n2), n1, n2,
               Dst<double>{G.b2, n1, 0});
        T2(KCOS, KCOS, INV, ext_src(G.b2, n1, n1, n2), N1, N2, pye);
      } else {
        zero(pye, nd);
      }
      base64(KCOS, KCOS, FWD, src_plain<double, double>(phiz_add, n1, n1, n2), n1, n2, Dst<double>{G.b2, n1, 0});
      T2(KCOS, KCOS, INV, ext_src(G.b2, n1, n1, n2), N1, N2, pza);
      // detaphiz = geta2phiz2 + phiz2 - phisx (etax phizMm2)_d - phisy (etay phizMm2)_d
      //            - phix_add (etax phizMm1)_d - phiy_add (etay phizMm1)_d + phiz_add phiz
      double* det = G.vit[2];
      ew(det, 1.0, G.ge2p2, nullptr, 1.0, G.phiz2, nd);
      ew(G.t2, 1.0, G.etax, G.pMm2, 0.0, nullptr, nd);
      dealias(2, G.t2, KSIN, KCOS);
      ew(det, -1.0, G.phisx, G.t2, 1.0, det, nd);
      ew(G.t2, 1.0, G.etay, G.pMm2, 0.0, nullptr, nd);
      dealias(2, G.t2, KCOS, KSIN);
      ew(det, -1.0, G.phisy, G.t2, 1.0, det, nd);
      ew(G.t2, 1.0, G.etax, G.pMm1, 0.0, nullptr, nd);
      dealias(2, G.t2, KSIN, KCOS);
      ew(det, -1.0, pxe, G.t2, 1.0, det, nd);
      ew(G.t2, 1.0, G.etay, G.pMm1, 0.0, nullptr, nd);
      dealias(2, G.t2, KCOS, KSIN);
      ew(det, -1.0, pye, G.t2, 1.0, det, nd);
      ew(det, 1.0, pza, pze, 1.0, det, nd);
      double* phim[4] = {G.b2, G.b3, G.b4, G.b0};  // x, y, z, t at z = eta (base, physical)
      filter_ext(det, KCOS, KCOS, G.b0);
      ew(G.b0, -1.0, G.b0, nullptr, 1.0, dphis3, nsb);
      base64(KCOS, KCOS, INV, src_plain<double, double>(P, n1, n1, n2), n1, n2, Dst<double>{G.b4, n1, 0});
      ew(G.t2, -1.0, G.etax, G.pMm1, 1.0, G.phisx, nd);
      dealias(2, G.t2, KSIN, KCOS);
      filter_ext(G.t2, KSIN, KCOS, G.b2);
      if (n2 > 1) {
        ew(G.t2, -1.0, G.etay, G.pMm1, 1.0, G.phisy, nd);
        dealias(2, G.t2, KCOS, KSIN);
        filter_ext(G.t2, KCOS, KSIN, G.b3);
      } else {
        zero(G.b3, nsb);
      }
      // --- HOSvel2_SWEET (velocities.f90:223-492): meta2 - 1 levels from z = eta down
      for (int f = 0; f < 4; ++f) {
        base64(KX[f], KY[f], FWD, src_plain<double, double>(phim[f], n1, n1, n2), n1, n2,
               Dst<double>{G.b1, n1, 0});
        hipLaunchKernelGGL(k_extend, dim3(nblk(nd)), dim3(256), 0, st, n1, n2, N1, N2, (const double*)G.b1,
                           G.aH[f][1]);
        T2(KX[f], KY[f], INV, ext_src(G.b1, n1, n1, n2), N1, N2, G.vit[f]);
      }
      std::vector<double*>* ea = &G.ep;   // powers of the level left
      std::vector<double*>* eb = &G.ep2;  // powers of the level reached
      auto level = [&](const std::vector<double*>* ela, const std::vector<double*>* elb, int shift) {
        for (int iH = 1; iH <= M; ++iH) {
          for (int f = 0; f < 4; ++f) {
            zero(G.pa[f], nd);
            if (elb) zero(G.pb[f], nd);
          }
          for (int j = iH; j >= 1; --j) {
            if (j == 1 && iH > 1) continue;
            const int jj = j, j2 = iH - (j - 1);
            for (int f = 0; f < 4; ++f) {
              T2(KX[f], KY[f], INV, ext_src(G.aH[f][j2], N1, N1, N2, f == 2 ? MUL_KTH2 : MUL_KTH, jj), N1, N2,
                 G.phizi);
              ew(G.pa[f], -1.0, G.phizi, (*ela)[jj], 1.0, G.pa[f], nd);
              if (elb) ew(G.pb[f], 1.0, G.phizi, (*elb)[jj], 1.0, G.pb[f], nd);
            }
          }
          if (iH > 1)
            for (int f = 0; f < 4; ++f) {
              T2(KX[f], KY[f], INV, ext_src(G.aH[f][iH], N1, N1, N2, f == 2 ? MUL_KTH2 : MUL_KTH, 1), N1, N2,
                 G.phizi);
              ew(G.pa[f], -1.0, G.phizi, (*ela)[1], 1.0, G.pa[f], nd);
              if (elb) ew(G.pb[f], 1.0, G.phizi, (*elb)[1], 1.0, G.pb[f], nd);
            }
          for (int f = 0; f < 4; ++f) {
            ew(G.vit[f], 1.0, G.pa[f], nullptr, 1.0, G.vit[f], nd);
            if (elb) ew(G.vit[f], 1.0, G.pb[f], nullptr, 1.0, G.vit[f], nd);
            dealias(iH + shift, G.vit[f], KX[f], KY[f]);
            if (iH < M) {
              dealias(iH + 1, G.pa[f], KX[f], KY[f]);
              T2(KX[f], KY[f], FWD, ext_src(G.pa[f], N1, N1, N2), N1, N2, G.aH[f][iH + 1]);
            }
          }
        }
      };
      for (int l = 1; l < meta2; ++l) {
        // eta2 = eta (meta2 - l) / meta2 and its powers (dealiased at order j, :296-315)
        ew(G.b1, (double)(meta2 - l) / (double)meta2, C.eta_hat, nullptr, 0.0, nullptr, nsb);
        T2(KCOS, KCOS, INV, ext_src(G.b1, n1, n1, n2), N1, N2, (*eb)[1]);
        for (int j = 1; j < M; ++j) {
          ew((*eb)[j + 1], 1.0 / (j + 1), (*eb)[j], (*eb)[1], 0.0, nullptr, nd);
          dealias(j, (*eb)[j + 1], KCOS, KCOS);
        }
        level(ea, eb, 0);
        for (int f = 0; f < 4; ++f) T2(KX[f], KY[f], FWD, ext_src(G.vit[f], N1, N1, N2), N1, N2, G.aH[f][1]);
        std::swap(ea, eb);
      }
      // --- HOSvel_SWEET (velocities.f90:40-216): from the last level to z = 0
      for (int f = 0; f < 4; ++f) T2(KX[f], KY[f], FWD, ext_src(G.vit[f], N1, N1, N2), N1, N2, G.aH[f][1]);
      level(ea, nullptr, 1);
      // filter_ext then space_2_Fourier (:177-185): the base sine transforms drop the end modes
      for (int f = 0; f < 4; ++f) {
        filter_ext(G.vit[f], KX[f], KY[f], G.b1);
        base64(KX[f], KY[f], FWD, src_plain<double, double>(G.b1, n1, n1, n2), n1, n2,
               Dst<double>{G.out[f], n1, 0});
      }
      for (int f = 0; f < 5; ++f) out_dev[f] = G.out[f];
      HIP_OK(hipGetLastError());
    }
  }

  void wavemaker(double t, const double* a1, const double* e1, const double* aad, double* ka1, double* ke1,
                 double* kad1, double* kadr, const S2* s2) {
    hos_ctx& C = *c;
    hipStream_t st = cs;
    const int n1 = C.n1, n2 = C.n2, n3 = C.n3, n3a = C.n3a;
    const int ldb = N1b;
    if (C.dip) {
      dipole_wavemaker(t, ka1, ke1, kad1, kadr);
      return;
    }
    if (!C.wmk) {
      HIP_OK(hipMemsetAsync(phi_add, 0, ns * sizeof(double), st));
      HIP_OK(hipMemsetAsync(phix_add, 0, ns * sizeof(double), st));
      HIP_OK(hipMemsetAsync(phiy_add, 0, ns * sizeof(double), st));
      HIP_OK(hipMemsetAsync(phiz_add, 0, ns * sizeof(double), st));
      HIP_OK(hipMemsetAsync(phit_add, 0, ns * sizeof(double), st));
      HIP_OK(hipMemsetAsync(phizh, 0, ns * sizeof(T), st));
      if constexpr (kSplit) surface_forcing_modes();  // the beach's first-order term
      HIP_OK(hipMemsetAsync(phixa, 0, nb * sizeof(T), st));
      HIP_OK(hipMemsetAsync(phiya, 0, nb * sizeof(T), st));
      HIP_OK(hipMemsetAsync(phiza, 0, nb * sizeof(T), st));
      if (ka1) HIP_OK(hipMemsetAsync(ka1, 0, ns * sizeof(double), st));
      if (ke1) HIP_OK(hipMemsetAsync(ke1, 0, ns * sizeof(double), st));
      if (kad1) HIP_OK(hipMemsetAsync(kad1, 0, na * sizeof(double), st));
      if (kadr) HIP_OK(hipMemsetAsync(kadr, 0, na * sizeof(double), st));
      return;
    }
    if (s2) {
      second_order_wavemaker(t, a1, e1, aad, ka1, ke1, kad1, kadr, *s2);
      return;
    }
    double al = 0.0, alt = 0.0, altt = 0.0;
    const double* dsc = C.dsc_active;
    if (dsc) {
      // the wavemaker globals of this evaluation, read by the next volume fix
      hipLaunchKernelGGL(k_wmk_glob, dim3(1), dim3(1), 0, st, C.dsc_active, C.cur_stage);
    } else {
      ramp_eval(C.p, t, al, alt, altt);
      C.last_alpha = al;
      C.last_alphat = alt;
      C.last_pos_top = C.pos_top(t);
    }
    hipLaunchKernelGGL(k_pos, dim3(n2), dim3(256), 0, st, n2, n3, C.p.n_freq, t, (const double*)freq,
                       (const double*)pre, (const double*)pim, (const double*)gz, Xp, Xpt, Xptt, dsc, C.cur_stage);
    // f1 = -alpha X_tt - 2 alpha' X_t - alpha'' X ; f2 = -alpha' X_t - alpha'' X
    hipLaunchKernelGGL(k_paddle_rhs, dim3(nblk((size_t)n3a * n2)), dim3(256), 0, st, n2, n3, n3a, C.delz, (const double*)Xp,
                       (const double*)Xpt, (const double*)Xptt, altt, 2.0 * alt, al, altt, alt, 0.0, cont, f1, f2,
                       dsc, C.cur_stage);
    paddle_modes(2, kad1, kadr);
    // additional potential at z = eta and its first-order part at z = 0
    phiadd(aad, kad1, eta_phys);
    if (multi) {
      first_order_gradients();
      fo_late = true;  // first_order_rest after the join (rhs)
    } else {
      first_order_single(e1, a1, ka1, ke1);
    }
    {
      Batch2<T, Src<T, T>, Dst<T>> b;
      b.add(KCOS, KCOS, src_plain<T>(phixh, n1, n1, n2), N1b, N2b, Dst<T>{phixa, ldb, 0}, (*ctmpX)[0]);
      if (n2 > 1) b.add(KCOS, KCOS, src_plain<T>(phiyh, n1, n1, n2), N1b, N2b, Dst<T>{phiya, ldb, 0}, (*ctmpX)[1]);
      b.add(KCOS, KCOS, src_plain<T>(phizh, n1, n1, n2), N1b, N2b, Dst<T>{phiza, ldb, 0}, (*ctmpX)[2]);
      bigB(INV, b);
    }
  }

  // Dipole wavemaker (igeom = 3): the ramp, then phi_dipoles at z = eta
  // (resol_HOS.f90:494-496).  There is no paddle state: the first-order free-surface
  // and paddle-mode derivatives stay zero, and phiy_add is zero (the reference never
  // sets it on this path).  grad phi_add goes to the extended grid as for the paddle.
  void dipole_wavemaker(double t, double* ka1, double* ke1, double* kad1, double* kadr) {
    hos_ctx& C = *c;
    hipStream_t st = cs;
    const int n1 = C.n1, n2 = C.n2;
    const int ldb = N1b;
    if (C.dsc_active) throw std::runtime_error("graph-captured attempts need a paddle wavemaker");
    double al, alt, altt;
    ramp_eval(C.p, t, al, alt, altt);
    C.last_alpha = al;
    C.last_alphat = alt;
    C.last_pos_top = 0.0;
    const DipSet& d = C.dips;
    hipLaunchKernelGGL(k_dip_phase, dim3(nblk(d.nf)), dim3(256), 0, st, d.nf, t, d.w, d.cs);
    hipLaunchKernelGGL((k_dipoles<T>), dim3(nblk(ns)), dim3(256), 0, st, n1, (int)ns, d, al, alt,
                       (const T*)eta_phys, phi_add, phit_add, phix_add, phiy_add, phiz_add);
    HIP_OK(hipGetLastError());
    if (ka1) HIP_OK(hipMemsetAsync(ka1, 0, ns * sizeof(double), st));
    if (ke1) HIP_OK(hipMemsetAsync(ke1, 0, ns * sizeof(double), st));
    if (kad1) HIP_OK(hipMemsetAsync(kad1, 0, na * sizeof(double), st));
    if (kadr) HIP_OK(hipMemsetAsync(kadr, 0, na * sizeof(double), st));
    base(KCOS, KCOS, FWD, src_plain<T, double>(phix_add, n1, n1, n2), n1, n2, Dst<T>{phixh, n1, 0});
    HIP_OK(hipMemsetAsync(phiyh, 0, ns * sizeof(T), st));
    HIP_OK(hipMemsetAsync(phiya, 0, nb * sizeof(T), st));
    surface_forcing_modes();
    Batch2<T, Src<T, T>, Dst<T>> b;
    b.add(KCOS, KCOS, src_plain<T>(phixh, n1, n1, n2), N1b, N2b, Dst<T>{phixa, ldb, 0}, (*ctmpX)[0]);
    b.add(KCOS, KCOS, src_plain<T>(phizh, n1, n1, n2), N1b, N2b, Dst<T>{phiza, ldb, 0}, (*ctmpX)[2]);
    bigB(INV, b);
  }

  // First-order free-surface RHS (phidphiz) and the modes of grad phi_add, one
  // base transform at a time.
  // Dense FP64 cos-cos base transforms of a 2-D tank for nv <= 3 vectors with one
  // read of the matrix; sources with a multiplier go through the grid's two staging
  // vectors.
  void dense_multi_2d(int dir, int nv, const Src<double, double>* src, double* const* out) {
    DenseGrid<double>& g = grid64();
    const int n1 = c->n1, di = (dir == FWD) ? 0 : 1;
    if (c->n2 != 1 || nv > kMaxGemv) throw std::runtime_error("dense_multi_2d: 2-D, three vectors at most");
    double* stage[2] = {g.stage, g.mid};
    int used = 0;
    GemvSet gs{};
    gs.nv = nv;
    for (int f = 0; f < nv; ++f) {
      const Src<double, double>& sf = src[f];
      if (sf.mul == MUL_NONE && sf.ld == n1 && sf.nx == n1 && sf.ny == 1 && (const void*)sf.p != (void*)out[f]) {
        gs.x[f] = sf.p;
      } else {
        if (used == 2) throw std::runtime_error("dense_multi_2d: two staged sources at most");
        double* sp = stage[used++];
        hipLaunchKernelGGL((k_dense_in<double, Src<double, double>>), dim3(nblk(n1)), dim3(256), 0, cs, n1, 1, sf,
                           sp);
        gs.x[f] = sp;
      }
      gs.y[f] = out[f];
    }
    hipLaunchKernelGGL(k_gemv_multi, dim3(n1), dim3(256), 0, cs, n1, (const double*)g.mx[KCOS][di], gs);
    HIP_OK(hipGetLastError());
    c->stats.n_base += nv;
  }

  void first_order_single(const double* e1, const double* a1, double* ka1, double* ke1) {
    hos_ctx& C = *c;
    hipStream_t st = cs;
    const int n1 = C.n1, n2 = C.n2;
    if constexpr (kSplit) {
      if (dense && n2 == 1) {
        // 2-D dense grid: the FP64 transforms of one direction share one read of
        // the matrix (inverse {kth1 a1st, kth1 phis}, forward {phidphiz, phiz_add,
        // forcing}); the same fields and kernels as below and in surface_forcing_modes
        const Src<double, double> inv[2] = {src_mul<double, double>(a1, n1, n1, n2, MUL_KTH, 1, C.pix, C.piy),
                                            src_mul<double, double>(phis64, n1, n1, n2, MUL_KTH, 1, C.pix, C.piy)};
        double* const invo[2] = {w1, wlin};
        dense_multi_2d(INV, 2, inv, invo);
        hipLaunchKernelGGL(k_first_order, dim3(nblk(ns)), dim3(256), 0, st, ns, (const double*)w1,
                           (const double*)phiz1, (const double*)phit1, e1, (const double*)C.nu_d,
                           (double)C.p.iabsnb, ke1, tmp1);
        hipLaunchKernelGGL(k_forcing, dim3(nblk(ns)), dim3(256), 0, st, ns, (const double*)phiz_add,
                           (const double*)phit_add, (const double*)wlin, (const double*)C.nu_d, Zf64);
        const Src<double, double> fwd[3] = {src_plain<double, double>(tmp1, n1, n1, n2),
                                            src_plain<double, double>(phiz_add, n1, n1, n2),
                                            src_plain<double, double>(Zf64, n1, n1, n2)};
        double* const fwdo[3] = {ka1, phizh64, Zh64};
        dense_multi_2d(FWD, 3, fwd, fwdo);
        hipLaunchKernelGGL(k_narrow<T>, dim3(nblk(ns)), dim3(256), 0, st, ns, (const double*)phizh64, phizh);
        base(KCOS, KCOS, FWD, src_plain<T, double>(phix_add, n1, n1, n2), n1, n2, Dst<T>{phixh, n1, 0});
        return;
      }
    } else {
      if (dense && n2 == 1) {
        // FP64 path, 2-D dense grid: the forward transforms {phidphiz, phix_add,
        // phiz_add} share one read of the matrix (surface_forcing_modes's one included)
        base64(KCOS, KCOS, INV, src_mul<double, double>(a1, n1, n1, n2, MUL_KTH, 1, C.pix, C.piy), n1, n2,
               Dst<double>{w1, n1, 0});
        hipLaunchKernelGGL(k_first_order, dim3(nblk(ns)), dim3(256), 0, st, ns, (const double*)w1,
                           (const double*)phiz1, (const double*)phit1, e1, (const double*)C.nu_d,
                           (double)C.p.iabsnb, ke1, tmp1);
        const Src<double, double> fwd[3] = {src_plain<double, double>(tmp1, n1, n1, n2),
                                            src_plain<double, double>(phix_add, n1, n1, n2),
                                            src_plain<double, double>(phiz_add, n1, n1, n2)};
        double* const fwdo[3] = {ka1, reinterpret_cast<double*>(phixh), reinterpret_cast<double*>(phizh)};
        dense_multi_2d(FWD, 3, fwd, fwdo);
        return;
      }
    }
    base64(KCOS, KCOS, INV, src_mul<double, double>(a1, n1, n1, n2, MUL_KTH, 1, C.pix, C.piy), n1, n2,
           Dst<double>{w1, n1, 0});
    hipLaunchKernelGGL(k_first_order, dim3(nblk(ns)), dim3(256), 0, st, ns, (const double*)w1,
                       (const double*)phiz1, (const double*)phit1, e1, (const double*)C.nu_d,
                       (double)C.p.iabsnb, ke1, tmp1);
    base64(KCOS, KCOS, FWD, src_plain<double, double>(tmp1, n1, n1, n2), n1, n2, Dst<double>{ka1, n1, 0});
    // grad phi_add onto the extended grid (resol_HOS.f90:534-544)
    base(KCOS, KCOS, FWD, src_plain<T, double>(phix_add, n1, n1, n2), n1, n2, Dst<T>{phixh, n1, 0});
    if (n2 > 1)  // (a 2-D tank reads no y gradient)
      base(KCOS, KCOS, FWD, src_plain<T, double>(phiy_add, n1, n1, n2), n1, n2, Dst<T>{phiyh, n1, 0});
    surface_forcing_modes();
  }

  // The same on a dense 3-D base grid, the transforms of one direction and
  // precision batched (dense_multi): inverse {kth1 a1st, kth1 phis (mixed)},
  // forward FP64 {phidphiz, phiz_add, forcing (mixed)} or {phidphiz, grad phi_add}
  // (FP64), forward FP32 {phix_add, phiy_add} (mixed).
  // The same on a dense 3-D base grid in two parts.  first_order_gradients: the
  // modes of grad phi_add, which the free-surface products read after the join;
  // first_order_rest (enqueued on the wavemaker stream after the join's event):
  // the first-order free-surface conditions (ke1, ka1) and, on the mixed path, the
  // FP64 forcing modes Zh64, read only at the end of the right-hand side.
  void first_order_gradients() {
    const int n1 = c->n1, n2 = c->n2;
    if constexpr (kSplit) {
      Src<double, double> in[1] = {src_plain<double, double>(phiz_add, n1, n1, n2)};
      double* out[1] = {phizh64};
      dense_multi<double>(grid64(), FWD, 1, in, out);
      hipLaunchKernelGGL(k_narrow<T>, dim3(nblk(ns)), dim3(256), 0, cs, ns, (const double*)phizh64, phizh);
      Src<T, double> inT[2] = {src_plain<T, double>(phix_add, n1, n1, n2), src_plain<T, double>(phiy_add, n1, n1, n2)};
      T* outT[2] = {phixh, phiyh};
      dense_multi<T>(dT, FWD, 2, inT, outT);
    } else {
      Src<double, double> in[3] = {src_plain<double, double>(phix_add, n1, n1, n2),
                                   src_plain<double, double>(phiy_add, n1, n1, n2),
                                   src_plain<double, double>(phiz_add, n1, n1, n2)};
      double* out[3] = {(double*)phixh, (double*)phiyh, (double*)phizh};
      dense_multi<double>(grid64(), FWD, 3, in, out);
    }
  }
  void first_order_rest(const double* e1, const double* a1, double* ka1, double* ke1) {
    hos_ctx& C = *c;
    hipStream_t st = cs;
    const int n1 = C.n1, n2 = C.n2;
    {
      // (the FP64 path transforms the first field only)
      Src<double, double> in[2] = {src_mul<double, double>(a1, n1, n1, n2, MUL_KTH, 1, C.pix, C.piy),
                                   src_mul<double, double>(kSplit ? phis64 : a1, n1, n1, n2, MUL_KTH, 1, C.pix,
                                                           C.piy)};
      double* out[2] = {w1, wlin};
      dense_multi<double>(grid64(), INV, kSplit ? 2 : 1, in, out);
    }
    hipLaunchKernelGGL(k_first_order, dim3(nblk(ns)), dim3(256), 0, st, ns, (const double*)w1,
                       (const double*)phiz1, (const double*)phit1, e1, (const double*)C.nu_d,
                       (double)C.p.iabsnb, ke1, tmp1);
    if constexpr (kSplit) {
      hipLaunchKernelGGL(k_forcing, dim3(nblk(ns)), dim3(256), 0, st, ns, (const double*)phiz_add,
                         (const double*)phit_add, (const double*)wlin, (const double*)C.nu_d, Zf64);
      Src<double, double> in[2] = {src_plain<double, double>(Zf64, n1, n1, n2),
                                   src_plain<double, double>(tmp1, n1, n1, n2)};
      double* out[2] = {Zh64, ka1};
      dense_multi<double>(grid64(), FWD, 2, in, out);
    } else {
      Src<double, double> in[1] = {src_plain<double, double>(tmp1, n1, n1, n2)};
      double* out[1] = {ka1};
      dense_multi<double>(grid64(), FWD, 1, in, out);
    }
  }

  // modes of phiz_add in T (the extended-grid gradient products); on the mixed
  // path from the FP64 modes, which the free-surface condition uses directly, and
  // the FP64 modes of the first-order part -phit_add - nu (phiz_add + F2S(kth1 phis))
  void surface_forcing_modes() {
    const int n1 = c->n1, n2 = c->n2;
    if constexpr (kSplit) {
      base64(KCOS, KCOS, FWD, src_plain<double, double>(phiz_add, n1, n1, n2), n1, n2, Dst<double>{phizh64, n1, 0});
      hipLaunchKernelGGL(k_narrow<T>, dim3(nblk(ns)), dim3(256), 0, cs, ns, (const double*)phizh64, phizh);
      base64(KCOS, KCOS, INV, src_mul<double, double>(phis64, n1, n1, n2, MUL_KTH, 1, c->pix, c->piy), n1, n2,
             Dst<double>{wlin, n1, 0});
      hipLaunchKernelGGL(k_forcing, dim3(nblk(ns)), dim3(256), 0, cs, ns, (const double*)phiz_add,
                         (const double*)phit_add, (const double*)wlin, (const double*)c->nu_d, Zf64);
      base64(KCOS, KCOS, FWD, src_plain<double, double>(Zf64, n1, n1, n2), n1, n2, Dst<double>{Zh64, n1, 0});
    } else {
      base(KCOS, KCOS, FWD, src_plain<T, double>(phiz_add, n1, n1, n2), n1, n2, Dst<T>{phizh, n1, 0});
    }
  }

  // Stacked paddle-grid modes stk6 [xx, xxt, y, yt, z, zt] (kinds cos-cos x2,
  // cos-sin x2, sin-cos x2) to their values at the paddle nodes (IDCT_FFTW_add,
  // fourier_FFTW.f90:451-492) in out; when a_fs is given, the free-surface
  // modes a_fs, da_fs evaluated at x = 0 are added before the y transform.
  void paddle_fields(const double* a_fs, const double* da_fs, double* out) {
    hos_ctx& C = *c;
    const int n1 = C.n1, n2 = C.n2, n3 = C.n3, n3a = C.n3a;
    gemm<double>(cs, true, true, n3a, 4 * n2, n3a, Cz, n3a, stk6, n3a, stkz, n3a);
    gemm<double>(cs, true, true, n3a, 2 * n2, n3a, Sz, n3a, stk6 + 4 * na, n3a, stkz + 4 * na, n3a);
    if (a_fs)
      hipLaunchKernelGGL(k_fs_on_paddle, dim3(n3, n2), dim3(256), 0, cs, n1, n2, n3, n3a, C.delz, C.pix, C.piy,
                         a_fs, da_fs, stkz);
    if (n2 > 1) {
      const double* My[3] = {Cy, Sy, Cy};
      for (int g = 0; g < 3; ++g)
        hosmm::mm<double>(cs, false, true, n3a, n2, n2, stkz + 2 * g * na, n3a, (long)na, My[g], n2, 0,
                          out + 2 * g * na, n3a, (long)na, 2);
    } else {
      HIP_OK(hipMemcpyAsync(out, stkz, 6 * na * sizeof(double), hipMemcpyDeviceToDevice, cs));
    }
    HIP_OK(hipGetLastError());
  }

  // Third-order paddle condition and modes (solveHOS_SWEET_3 with wmk_3rdorder,
  // resol_wmkr.f90:338-437, :942-1115); the paddle motion X.., its derivatives
  // and the second-order derivatives q.ka2, q.kad2 of this stage are in place.
  void third_order_paddle(double al, double alt, const double* aad, const double* kad1, const S2& q) {
    hos_ctx& C = *c;
    const int n2 = C.n2, n3 = C.n3, n3a = C.n3a;
    hipLaunchKernelGGL(k_paddle_prep6_x3, dim3(nblk(na)), dim3(256), 0, cs, n2, n3a, C.piy, C.p.xlen, aad, kad1,
                       stk6);
    paddle_fields(nullptr, nullptr, Pw3);
    hipLaunchKernelGGL(k_paddle_prep6, dim3(nblk(na)), dim3(256), 0, cs, n2, n3a, C.piy, q.aad2,
                       (const double*)q.kad2, stk6);
    paddle_fields(q.a2, q.ka2, Pw);
    hipLaunchKernelGGL(k_paddle_bc3, dim3(nblk(n2, 64)), dim3(64), 0, cs, n2, n3, n3a, C.delz, al, alt,
                       (const double*)Pw3, (const double*)Pw, (const double*)Xp, (const double*)Xpt,
                       (const double*)Xdy, (const double*)Xdyt, (const double*)Xdz, (const double*)Xdzt, cont, f1);
    paddle_modes(1, q.kad3, (double*)nullptr);
  }

  // z = 0 wavemaker sums of the SET 0 (11 fields) or SET 1 (2 fields) kind.
  template <int SET>
  void phiadd_z0(const double* a, const double* da, Ptr11 outp) {
    hos_ctx& C = *c;
    // one thread per y-node pair (kPhi3Per pairs per thread beyond 256 threads)
    const int np = (C.n2 + 1) / 2;
    int nt = 64;
    while (nt < 256 && nt < np) nt *= 2;
    if (nt * kPhi3Per < np) throw std::runtime_error("n2 too large for the wavemaker kernel");
    hipLaunchKernelGGL((k_phiadd_z0<SET>), dim3(C.n1, nch), dim3(nt), 0, cs, C.n1, C.n2, C.n3a, C.p.xlen, C.piy,
                       C.dx, a, da, part2);
    hipLaunchKernelGGL(k_phiadd_reduce_n, dim3(nblk(ns)), dim3(256), 0, cs, C.n1, C.n2, nch, Z0Set<SET>::NF, C.dx,
                       (const double*)part2, outp);
    HIP_OK(hipGetLastError());
  }

  // First- and second-order wavemaker (i_wmk = 2): solveHOS_SWEET_1 with
  // d1st_for_2nd (resol_wmkr.f90:109-210, :645-793), solveHOS_SWEET_2
  // (:219-329), then the additional potential of a_add_1 + a_add_2 on the free
  // surface (resol_HOS.f90:498-502).
  void second_order_wavemaker(double t, const double* a1, const double* e1, const double* aad, double* ka1,
                              double* ke1, double* kad1, double* kadr, const S2& q) {
    hos_ctx& C = *c;
    hipStream_t st = cs;
    const int n1 = C.n1, n2 = C.n2, n3 = C.n3, n3a = C.n3a;
    const int ldb = N1b;
    double al, alt, altt;
    ramp_eval(C.p, t, al, alt, altt);
    C.last_alpha = al;
    C.last_alphat = alt;
    C.last_pos_top = C.pos_top(t);
    // --- SWEET 1: paddle condition and modes
    hipLaunchKernelGGL(k_pos, dim3(n2), dim3(256), 0, st, n2, n3, C.p.n_freq, t, (const double*)freq,
                       (const double*)pre, (const double*)pim, (const double*)gz, Xp, Xpt, Xptt);
    hipLaunchKernelGGL(k_paddle_rhs, dim3(nblk((size_t)n3a * n2)), dim3(256), 0, st, n2, n3, n3a, C.delz, (const double*)Xp,
                       (const double*)Xpt, (const double*)Xptt, altt, 2.0 * alt, al, altt, alt, 0.0, cont, f1, f2);
    paddle_modes(2, kad1, kadr);
    // phiadd_1st_SWEET at z = 0: the 11 fields (phit, phiz first)
    Ptr11 z1{};
    for (int f = 0; f < 11; ++f) z1.p[f] = zf[f];
    z1.p[8] = nullptr;  // phixx_add_1 feeds only the third-order wavemaker
    phiadd_z0<0>(aad, kad1, z1);
    // phidphiz and the first-order free-surface conditions
    base64(KCOS, KCOS, INV, src_mul<double, double>(a1, n1, n1, n2, MUL_KTH, 1, C.pix, C.piy), n1, n2,
           Dst<double>{w1, n1, 0});
    hipLaunchKernelGGL(k_first_order, dim3(nblk(ns)), dim3(256), 0, st, ns, (const double*)w1,
                       (const double*)phiz1, (const double*)phit1, e1, (const double*)C.nu_d,
                       (double)C.p.iabsnb, ke1, tmp1);
    base64(KCOS, KCOS, FWD, src_plain<double, double>(tmp1, n1, n1, n2), n1, n2, Dst<double>{ka1, n1, 0});
    // --- d1st_for_2nd: optional x/y filters of a1st and da1st (resol_wmkr.f90:673-681)
    const double* a1u = a1;
    if (C.filt_fs) {
      HIP_OK(hipMemcpyAsync(a1f, a1, ns * sizeof(double), hipMemcpyDeviceToDevice, st));
      hipLaunchKernelGGL(k_mulw, dim3(nblk(ns)), dim3(256), 0, st, ns, (const double*)C.filt_fs, a1f);
      hipLaunchKernelGGL(k_mulw, dim3(nblk(ns)), dim3(256), 0, st, ns, (const double*)C.filt_fs, ka1);
      a1u = a1f;
    }
    // first-order potential derivatives on the paddle: additional-potential part
    // (inverse paddle-grid transforms as dense products) plus the free-surface modes
    hipLaunchKernelGGL(k_paddle_prep6, dim3(nblk(na)), dim3(256), 0, st, n2, n3a, C.piy, aad, (const double*)kad1,
                       stk6);
    paddle_fields(a1u, ka1, Pw);
    // first-order free-surface fields
    hipLaunchKernelGGL(k_fs_modes8, dim3(nblk(ns)), dim3(256), 0, st, n1, n2, C.pix, C.piy, a1u, (const double*)ka1,
                       modes8);
    const int kxs[8] = {KSIN, KCOS, KSIN, KCOS, KCOS, KCOS, KSIN, KCOS};
    const int kys[8] = {KCOS, KSIN, KCOS, KSIN, KCOS, KCOS, KCOS, KSIN};
    for (int f = 0; f < 8; ++f) {
      if (n2 == 1 && kys[f] == KSIN) {  // ky = 0 in a 2-D tank
        HIP_OK(hipMemsetAsync(T8 + f * ns, 0, ns * sizeof(double), st));
        continue;
      }
      base64(kxs[f], kys[f], INV, src_plain<double, double>(modes8 + f * ns, n1, n1, n2), n1, n2,
             Dst<double>{T8 + f * ns, n1, 0});
    }
    Z0Fields zz{zf[2], zf[3], zf[4], zf[5], zf[6], zf[7], zf[9], zf[10]};
    hipLaunchKernelGGL(k_fs_fields, dim3(nblk(ns)), dim3(256), 0, st, ns, T8, zz, (const double*)ke1,
                       (const double*)C.nu_d, (const double*)nux, (const double*)nuy, (double)C.p.iabsnb);
    // --- SWEET 2: paddle condition, modes, z = 0 potential, free-surface conditions
    hipLaunchKernelGGL(k_dpos, dim3(n2), dim3(256), 0, st, n2, n3, C.p.n_freq, t, (const double*)freq,
                       (const double*)pre, (const double*)pim, (const double*)dyre, (const double*)dyim,
                       (const double*)gz, (const double*)dgz, Xdy, Xdyt, Xdz, Xdzt);
    hipLaunchKernelGGL(k_paddle_bc2, dim3(nblk(n2, 64)), dim3(64), 0, st, n2, n3, n3a, C.delz, al, alt,
                       (const double*)Pw, (const double*)Xp, (const double*)Xpt, (const double*)Xdy,
                       (const double*)Xdyt, (const double*)Xdz, (const double*)Xdzt, cont, f1);
    paddle_modes(1, q.kad2, (double*)nullptr);
    Ptr11 z2{};
    z2.p[0] = phit2;
    z2.p[1] = phiz2;
    phiadd_z0<1>(q.aad2, q.kad2, z2);
    base64(KCOS, KCOS, INV, src_mul<double, double>(q.a2, n1, n1, n2, MUL_KTH, 1, C.pix, C.piy), n1, n2,
           Dst<double>{w2, n1, 0});
    hipLaunchKernelGGL(k_second_fs, dim3(nblk(ns)), dim3(256), 0, st, ns, (const double*)w2, (const double*)phiz2,
                       (const double*)phit2, (const double*)T8, (const double*)ke1, e1, q.e2, (const double*)C.nu_d,
                       (double)C.p.iabsnb, q.ke2, tmp2);
    base64(KCOS, KCOS, FWD, src_plain<double, double>(tmp2, n1, n1, n2), n1, n2, Dst<double>{q.ka2, n1, 0});
    // --- SWEET 3: third-order paddle modes
    if (q.aad3) third_order_paddle(al, alt, aad, kad1, q);
    // --- additional potential of the total paddle modes on the free surface
    hipLaunchKernelGGL(k_sum_orders3, dim3(nblk(na)), dim3(256), 0, st, na, aad, q.aad2, q.aad3, kad1,
                       (const double*)q.kad2, (const double*)q.kad3, atot, dtot);
    phiadd(atot, dtot, eta_phys);
    // phit1 / phiz1 were overwritten by the surface sums of the totals; the first-order
    // conditions above have used the first-order values already.
    base(KCOS, KCOS, FWD, src_plain<T, double>(phix_add, n1, n1, n2), n1, n2, Dst<T>{phixh, n1, 0});
    if (n2 > 1)  // (a 2-D tank reads no y gradient)
      base(KCOS, KCOS, FWD, src_plain<T, double>(phiy_add, n1, n1, n2), n1, n2, Dst<T>{phiyh, n1, 0});
    surface_forcing_modes();
    {
      Batch2<T, Src<T, T>, Dst<T>> b;
      b.add(KCOS, KCOS, src_plain<T>(phixh, n1, n1, n2), N1b, N2b, Dst<T>{phixa, ldb, 0}, (*ctmpX)[0]);
      if (n2 > 1) b.add(KCOS, KCOS, src_plain<T>(phiyh, n1, n1, n2), N1b, N2b, Dst<T>{phiya, ldb, 0}, (*ctmpX)[1]);
      b.add(KCOS, KCOS, src_plain<T>(phizh, n1, n1, n2), N1b, N2b, Dst<T>{phiza, ldb, 0}, (*ctmpX)[2]);
      bigB(INV, b);
    }
  }

  // S2F_add of nfields paddle columns stored contiguously from f1, divided by
  // k_add tanh(k_add xlen): z transform Wz * F, y transform (.) * Wy.
  void paddle_modes(int nfields, double* out1, double* out2) {
    hos_ctx& C = *c;
    const int n2 = C.n2, n3a = C.n3a;
    gemm<double>(cs, true, true, n3a, nfields * n2, n3a, Wz, n3a, f1, n3a, fz1, n3a);
    const double* zin = fz1;
    if (n2 > 1) {
      hosmm::mm<double>(cs, false, true, n3a, n2, n2, fz1, n3a, (long)na, Wy, n2, 0, fz2, n3a, (long)na, nfields);
      zin = fz2;
    }
    hipLaunchKernelGGL(k_add_scale, dim3(nblk(na * nfields)), dim3(256), 0, cs, n2, n3a, nfields, C.p.xlen, C.piy,
                       (const double*)C.filt_z,
                       zin, out1, out2);
    HIP_OK(hipGetLastError());
  }

  void phiadd(const double* a, const double* da, const T* eta) {
    hos_ctx& C = *c;
    hipStream_t st = cs;
    const int n1 = C.n1, n2 = C.n2, n3a = C.n3a;
    if (n2 == 1) {
      hipLaunchKernelGGL((k_phiadd2d<T>), dim3(nblk(n1, 4)), dim3(256), 0, st, n1, n3a, C.p.xlen, C.dx, a, da, eta,
                         phi_add, phit_add, phix_add, phiy_add, phiz_add, phit1, phiz1);
    } else {
      // one thread per y-node pair (kPhi3Per pairs per thread beyond 256 threads)
      const int np = (n2 + 1) / 2;
      int nt = 64;
      while (nt < 256 && nt < np) nt *= 2;
      if (nt * kPhi3Per < np) throw std::runtime_error("n2 too large for the wavemaker kernel");
      hipLaunchKernelGGL((k_phiadd3d<T>), dim3(n1, nch), dim3(nt), 0, st, n1, n2, n3a, C.p.xlen, C.piy, C.dx, a, da,
                         eta, part);
      hipLaunchKernelGGL(k_phiadd_reduce, dim3(nblk(ns)), dim3(256), 0, st, n1, n2, nch, C.dx, (const double*)part,
                         phi_add, phit_add, phix_add, phiy_add, phiz_add, phit1, phiz1);
    }
    HIP_OK(hipGetLastError());
  }

  // `items` independent full inverse cos-cos transforms in one launch per pass
  // (items = 1: a lone transform; more: the batched launches of a right-hand
  // side), each item with its own source and destination
  void time_big(int reps, int items, double* out4) override {
    hos_ctx& C = *c;
    if (items < 1 || items > 4) throw std::runtime_error("timed batch of 1 to 4 transforms");
    T* xin[4] = {X, tmpX[0], tmpX[1], tmpX[2]};
    T* mid[4] = {N2b > 1 ? tmpB : Y, tmpX[3], tmpX[4], tmpX[5]};
    T* yout[4] = {Y, X, tmpX[0], tmpX[1]};
    hipEvent_t e0, e1, e2;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventCreate(&e2));
    LineBatch<Src<T, T>, Dst<T>> bx, by;
    bx.n = by.n = items;
    for (int i = 0; i < items; ++i) {
      bx.kind[i] = by.kind[i] = KCOS;
      bx.n_in[i] = bx.n_out[i] = N1b;
      bx.nlines[i] = N2b;
      bx.src[i] = src_plain<T>(xin[i], N1b, N1b, N2b);
      bx.dst[i] = Dst<T>{mid[i], N1b, 0};
      by.n_in[i] = by.n_out[i] = N2b;
      by.nlines[i] = N1b;
      by.src[i] = src_plain<T>(mid[i], N1b, N1b, N2b);
      by.dst[i] = Dst<T>{yout[i], N1b, 0};
    }
    HIP_OK(hipEventRecord(e0, C.st));
    for (int r = 0; r < reps; ++r) launch_x<T>(gb.px, N1b, INV, bx, C.st, gwsM);
    HIP_OK(hipEventRecord(e1, C.st));
    if (N2b > 1)
      for (int r = 0; r < reps; ++r) launch_y<T>(gb.py, gb.TX, N2b, INV, by, C.st);
    HIP_OK(hipEventRecord(e2, C.st));
    HIP_OK(hipEventSynchronize(e2));
    float mx = 0.f, my = 0.f;
    HIP_OK(hipEventElapsedTime(&mx, e0, e1));
    HIP_OK(hipEventElapsedTime(&my, e1, e2));
    out4[0] = mx / reps;
    out4[1] = 2.0 * items * (double)nb * sizeof(T);
    out4[2] = N2b > 1 ? my / reps : 0.0;
    out4[3] = N2b > 1 ? 2.0 * items * (double)nb * sizeof(T) : 0.0;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
  }

  void slope_max(double& sx, double& sy) override {
    hos_ctx& C = *c;
    const unsigned nbk = nblk(nb);
    // (k_absmax2 writes every block's entry of both partial buffers)
    double* pa = slope_part;
    double* pb = slope_part + nbk;
    hipLaunchKernelGGL(k_absmax2, dim3(nbk), dim3(256), 0, C.st, nb, (const void*)etax, (const void*)etay,
                       sizeof(T) == 4 ? 1 : 0, pa, pb);
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(256), 0, C.st, (int)nbk, (const double*)pa, C.red6);
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(256), 0, C.st, (int)nbk, (const double*)pb, C.red6 + 1);
    double h[2];
    HIP_OK(hipMemcpyAsync(h, C.red6, 2 * sizeof(double), hipMemcpyDeviceToHost, C.st));
    HIP_OK(hipStreamSynchronize(C.st));
    sx = h[0];
    sy = h[1];
  }
};

// ===========================================================================
// context life cycle and RK driver
// ===========================================================================
namespace {

// Makes the order-m engine the active one (built on first use).  Orders below
// mHOS stand for a context created with mHOS = m; the wavemaker coupling needs
// m >= 3 (or m = 1).
void select_order(hos_ctx* c, int m) {
  if (m < 1 || m > c->p.mHOS) throw std::runtime_error("order outside 1..mHOS");
  if (c->add_pot() && m == 2) throw std::runtime_error("the wavemaker coupling needs an order >= 3 (or 1)");
  if (!c->engines[m]) {
    if (c->p.precision == 0)
      c->engines[m] = std::make_unique<Engine<double>>(c, m);
    else
      c->engines[m] = std::make_unique<Engine<float>>(c, m);
  }
  c->eng = c->engines[m].get();
  c->M = m;
  int q1, q2;
  order_grid(c->p, m, q1, q2, c->N1b, c->N2b);
  c->stats.big_n1 = c->N1b;
  c->stats.big_n2 = c->N2b;
}

// Device tables of the dipole wavemaker: the components grouped by dipole depth (the
// image sums depend on the depth only), each group in the caller's order.
void dipoles_setup(hos_ctx* c, const hos_params& p) {
  const int nf = p.n_dip;
  if (nf <= 0 || !p.dip_freq || !p.dip_mux_re || !p.dip_mux_im || !p.dip_muz_re || !p.dip_muz_im || !p.dip_z)
    throw std::runtime_error("igeom = 3 needs the dipole components (n_dip > 0 and the dip_* arrays)");
  if (p.n_side < 0 || p.n_side > 64) throw std::runtime_error("n_side must be in 0..64");
  std::vector<double> zs;
  std::vector<int> cls(nf);
  for (int f = 0; f < nf; ++f) {
    const double v[5] = {p.dip_freq[f], p.dip_mux_re[f], p.dip_mux_im[f], p.dip_muz_re[f], p.dip_muz_im[f]};
    for (double x : v)
      if (!std::isfinite(x))
        throw std::runtime_error("dipole component " + std::to_string(f + 1) +
                                 " is not finite (the dipole transfer function of linear_wave.f90:316-327 "
                                 "divides by k^2 - (k tanh k)^2, zero in FP64 for 18.7 < k <= 20)");
    if (!(p.dip_z[f] > -1.0 && p.dip_z[f] < 0.0)) throw std::runtime_error("dipole depth must lie in (-1, 0)");
    size_t q = 0;
    while (q < zs.size() && zs[q] != p.dip_z[f]) ++q;
    if (q == zs.size()) {
      if ((int)zs.size() == kDipMaxDepth) throw std::runtime_error("more than 8 distinct dipole depths");
      zs.push_back(p.dip_z[f]);
    }
    cls[f] = (int)q;
  }
  DipSet& d = c->dips;
  d.nf = nf;
  d.nz = (int)zs.size();
  d.n_side = p.n_side;
  d.x_dip = p.x_dip;
  d.xlen = p.xlen;
  d.dx = c->dx;
  std::vector<double> w, m(4 * (size_t)nf);
  int o = 0;
  for (int q = 0; q < d.nz; ++q) {
    d.z[q] = zs[q];
    d.start[q] = o;
    for (int f = 0; f < nf; ++f) {
      if (cls[f] != q) continue;
      w.push_back(2.0 * M_PI * p.dip_freq[f]);
      m[o] = p.dip_mux_re[f];
      m[nf + o] = p.dip_mux_im[f];
      m[2 * nf + o] = p.dip_muz_re[f];
      m[3 * nf + o] = p.dip_muz_im[f];
      ++o;
    }
  }
  d.start[d.nz] = o;
  double* wd = c->ar.get<double>(nf);
  double* md = c->ar.get<double>(4 * (size_t)nf);
  d.cs = c->ar.get<double>(2 * (size_t)nf);
  HIP_OK(hipMemcpy(wd, w.data(), nf * sizeof(double), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(md, m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice));
  d.w = wd;
  d.m = md;
}

void check_params(const hos_params& p) {
  if (p.n1 < 4) throw std::runtime_error("n1 must be >= 4");
  if (p.n2 < 1 || p.n2 == 2 || p.n2 == 3) throw std::runtime_error("n2 must be 1 or >= 4");
  if (p.n3 < 5) throw std::runtime_error("n3 must be >= 5");
  if (p.mHOS < 1 || p.mHOS > 8) throw std::runtime_error("mHOS must be in 1..8");
  const int p2ok = (p.n2 == 1) ? 1 : p.mHOS;
  if (p.p1 != p.mHOS || (p.p2 != p2ok && !(p.n2 == 1 && p.p2 == p.mHOS)))
    throw std::runtime_error("supported dealiasing: p1 = mHOS, p2 = mHOS (3-D) or 1 (2-D)");
  if (p.i_wmk != 1 && p.i_wmk != 2 && p.i_wmk != 3 && (p.icase != 1))
    throw std::runtime_error("supported wavemaker order: i_wmk = 1, 2 or 3");
  // wmk_3rdorder filters its INTENT(IN) stage inputs in place (resol_wmkr.f90:1019-1030),
  // which the Fortran standard leaves undefined for the caller's arrays
  if (p.i_wmk == 3 && p.icase != 1 &&
      (filter_active(p.coeffilt[0]) || (p.n2 > 1 && filter_active(p.coeffilt[1])) || filter_active(p.coeffilt[2])))
    throw std::runtime_error("i_wmk = 3 is supported without filters (coeffilt = 1)");
  if (p.precision != 0 && p.precision != 1) throw std::runtime_error("precision must be 0 or 1");
  if (p.xlen <= 0.0 || (p.n2 > 1 && p.ylen <= 0.0)) throw std::runtime_error("tank size must be positive");
  if (p.n_freq < 0 || (p.n_freq > 0 && (!p.freq || !p.pos_ini_re || !p.pos_ini_im)))
    throw std::runtime_error("wavemaker spectrum arrays missing");
}

// Stage time t + h c with two roundings, as the reference computes it
// (runge_kutta.f90:186): a fused multiply-add can move a stage across a ramp
// boundary such as t = Tramp.
double stage_time(double t, double h, double cs) {
#pragma clang fp contract(off)
  const double hc = h * cs;
  return t + hc;
}

void rhs_eval(hos_ctx* c, double t, double dt, const double* G1, const double* G2, const double* a1,
              const double* e1, const double* aad, int s, const double* a2 = nullptr, const double* e2 = nullptr,
              const double* aad2 = nullptr, const double* aad3 = nullptr) {
  if (c->wmk2 && !(a2 && e2 && aad2)) throw std::runtime_error("i_wmk >= 2 needs the second-order state");
  if (c->wmk3 && !aad3) throw std::runtime_error("i_wmk = 3 needs the third-order paddle modes");
  EngineBase::S2 s2{a2, e2, aad2, c->ka2[s], c->ke2[s], c->kad2[s], c->wmk3 ? aad3 : nullptr,
                    c->wmk3 ? c->kad3[s] : nullptr};
  c->cur_stage = s;
  c->eng->rhs(t, dt, G1, G2, a1, e1, aad, c->kG1[s], c->kG2[s], c->ka1[s], c->ke1[s], c->kad1[s], c->kadr[s],
              c->wmk2 ? &s2 : nullptr);
}

// Enqueues one runge4_adapt2 attempt and the readback of its error maxima on
// the context's stream, without waiting (rk_finish completes it).
void rk_enqueue(hos_ctx* c, double t, double h) {
  const size_t ns = (size_t)c->n1 * c->n2, na = (size_t)c->n3a * c->n2;
  hipStream_t st = c->st;
  hipLaunchKernelGGL(k_step_begin, dim3(nblk(ns)), dim3(256), 0, st, c->n1, c->n2, c->pix, c->piy,
                     (const double*)c->eta_hat, (const double*)c->phis_hat, c->G1, c->G2);
  rhs_eval(c, t, h * c->C[0], c->G1, c->G2, c->a1st, c->eta1st, c->a1st_add, 0, c->a2nd, c->eta2nd, c->a2nd_add,
           c->a3rd_add);
  for (int jl = 1; jl < 5; ++jl) {
    AxpyBatch ab{};
    auto comb = [&](double* out, const double* base, double* const* karr, size_t n) {
      KSet& k = ab.ks[ab.n];
      k.nk = jl;
      for (int s = 0; s < jl; ++s) {
        k.k[s] = karr[s];
        k.c[s] = c->A[jl][s];
      }
      ab.out[ab.n] = out;
      ab.base[ab.n] = base;
      ab.len[ab.n] = n;
      ++ab.n;
    };
    comb(c->G1m, c->G1, c->kG1, ns);
    comb(c->G2m, c->G2, c->kG2, ns);
    comb(c->aadm, c->a1st_add, c->kad1, na);
    comb(c->a1m, c->a1st, c->ka1, ns);
    comb(c->e1m, c->eta1st, c->ke1, ns);
    if (c->wmk2) {
      comb(c->a2m, c->a2nd, c->ka2, ns);
      comb(c->e2m, c->eta2nd, c->ke2, ns);
      comb(c->aad2m, c->a2nd_add, c->kad2, na);
    }
    if (c->wmk3) comb(c->aad3m, c->a3rd_add, c->kad3, na);
    hipLaunchKernelGGL(k_axpyk_batch, dim3(nblk(std::max(ns, na)), ab.n), dim3(256), 0, st, h, ab,
                       (const double*)c->dsc_active);
    rhs_eval(c, stage_time(t, h, c->C[jl]), h * c->C[jl], c->G1m, c->G2m, c->a1m, c->e1m, c->aadm, jl, c->a2m,
             c->e2m, c->aad2m, c->aad3m);
  }
  // final combination and embedded error (runge_kutta.f90:225-279)
  auto kset = [&](double* const* karr, const double* w) {
    KSet k{};
    k.nk = 5;
    for (int s = 0; s < 5; ++s) {
      k.k[s] = karr[s];
      k.c[s] = w[s];
    }
    return k;
  };
  const unsigned nbs = nblk(ns), nba = nblk(na);
  double* Gn1 = c->G1m;  // reuse stage buffers for the new modal G
  double* Gn2 = c->G2m;
  // runge_kutta.f90:243-279: the new state, the embedded-error maxima and the
  // wavemaker mode combinations, three launches in all
  FinalBatch fb{};
  FinalCounts fc{};
  fb.stride = c->pmax_stride;
  auto fin = [&](double* out, const double* base, size_t len, double* const (&k)[5]) {
    const int f = fb.n++;
    fb.out[f] = out; fb.base[f] = base; fb.len[f] = len;
    fb.kb[f] = kset(k, c->B); fb.ke[f] = kset(k, c->E);
    fc.nb[f] = (int)nblk(len);
  };
  fin(Gn1, c->G1, ns, c->kG1);
  fin(Gn2, c->G2, ns, c->kG2);
  fin(c->a1st_add_t, c->a1st_add, na, c->kad1);
  // second- and third-order wavemaker modes and their error terms (runge_kutta.f90:249-256, :275-276)
  if (c->wmk2) fin(c->a2nd_add_t, c->a2nd_add, na, c->kad2);
  if (c->wmk3) fin(c->a3rd_add_t, c->a3rd_add, na, c->kad3);
  hipLaunchKernelGGL(k_final_err_batch, dim3(std::max(nbs, nba), fb.n), dim3(256), 0, st, h, fb, c->pmax_e,
                     c->pmax_x, (const double*)c->dsc_active);
  hipLaunchKernelGGL(k_reduce_max_batch, dim3(2 * fb.n), dim3(256), 0, st, c->pmax_stride, fc,
                     (const double*)c->pmax_e, (const double*)c->pmax_x, c->red6);
  AxpyBatch ab{};
  auto ax = [&](double* out, const double* base, size_t len, double* const (&k)[5]) {
    const int f = ab.n++;
    ab.out[f] = out; ab.base[f] = base; ab.len[f] = len; ab.ks[f] = kset(k, c->B);
  };
  ax(c->a_add_ramp_t, c->a_add_ramp, na, c->kadr);
  ax(c->a1st_t, c->a1st, ns, c->ka1);
  ax(c->eta1st_t, c->eta1st, ns, c->ke1);
  if (c->wmk2) {
    ax(c->a2nd_t, c->a2nd, ns, c->ka2);
    ax(c->eta2nd_t, c->eta2nd, ns, c->ke2);
  }
  hipLaunchKernelGGL(k_axpyk_batch, dim3(std::max(nbs, nba), ab.n), dim3(256), 0, st, h, ab,
                     (const double*)c->dsc_active);
  // GtoF back, deta, volume fix with the last stage's wavemaker globals
  hipLaunchKernelGGL(k_step_end, dim3(nbs), dim3(256), 0, st, c->n1, c->n2, c->pix, c->piy, h, (const double*)Gn1,
                     (const double*)Gn2, (const double*)c->kG1[4], (const double*)c->kG2[4],
                     (const double*)c->phis_hat, c->eta_hat_t,
                     c->phis_hat_t, c->deta_hat_t, c->partial, (const double*)c->dsc_active);
  hipLaunchKernelGGL(k_volfix<double>, dim3(1), dim3(256), 0, st, (int)nbs, (const double*)c->partial,
                     (double*)nullptr, c->eta_hat_t, c->volmode(), c->last_alpha, c->last_pos_top, c->p.xlen,
                     c->p.d_hinge, (const double*)c->dsc_active);
  const int nr = c->wmk3 ? 10 : (c->wmk2 ? 8 : 6);
  if (c->dsc_active) {
    // graph-captured attempt: the controller and the copy-back on the device
    hipLaunchKernelGGL(k_control, dim3(1), dim3(1), 0, st, c->dsc_active, (const double*)c->red6, nr / 2,
                       c->p.iramp, c->p.Tramp);
    CopyBatch cb{};
    auto cp = [&](double* dst, const double* src, size_t len) {
      cb.dst[cb.n] = dst;
      cb.src[cb.n] = src;
      cb.len[cb.n] = len;
      ++cb.n;
    };
    cp(c->eta_hat, c->eta_hat_t, ns);
    cp(c->phis_hat, c->phis_hat_t, ns);
    cp(c->a1st_add, c->a1st_add_t, na);
    cp(c->a_add_ramp, c->a_add_ramp_t, na);
    cp(c->a1st, c->a1st_t, ns);
    cp(c->eta1st, c->eta1st_t, ns);
    cp(c->deta_hat, c->deta_hat_t, ns);
    hipLaunchKernelGGL(k_commit_if, dim3(nblk(std::max(ns, na)), cb.n), dim3(256), 0, st,
                       (const double*)c->dsc_active, cb);
    hipLaunchKernelGGL(k_sub_if, dim3(nblk(na)), dim3(256), 0, st, na, c->a1st_add, (const double*)c->a_add_ramp,
                       (const double*)c->dsc_active);
  } else {
    HIP_OK(hipMemcpyAsync(c->red_host, c->red6, nr * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  c->have_trial = false;
}

// Waits for the enqueued attempt and forms the embedded error (runge_kutta.f90:272-279).
void rk_finish(hos_ctx* c, double* err) {
  HIP_OK(hipStreamSynchronize(c->st));
  const double* r = c->red_host;
  const double tiny = 2.220446049250313e-16;
  double e = r[0] / (r[1] + tiny);
  e = std::max(e, r[2] / (r[3] + tiny));
  e = std::max(e, r[4] / (r[5] + tiny));
  if (c->wmk2) e = std::max(e, r[6] / (r[7] + tiny));
  if (c->wmk3) e = std::max(e, r[8] / (r[9] + tiny));
  *err = e;
  c->have_trial = true;
}

void rk_attempt(hos_ctx* c, double t, double h, double* err) {
  rk_enqueue(c, t, h);
  rk_finish(c, err);
}

void commit(hos_ctx* c);

// The accept / reject controller of the inner loop (HOS-NWT.f90:518-563) for one
// finished attempt of length hl from t; updates t, h and the counters.
struct Advance {
  double t, h;
  int na, nr;
};
void control(hos_ctx* c, Advance& a, double hl, double err, double toler, double dt_stab, double delt_out) {
  const double tiny = 2.220446049250313e-16;
  const hos_params& p = c->p;
  double corr;
  if (err > toler && a.h > 0.001 * 0.95 * dt_stab) {
    corr = 0.9 * std::pow(err / toler, -(1.0 / 4.0));
    ++a.nr;
    if (std::fabs(p.Tramp - (a.t + hl)) <= tiny && p.Tramp > a.t) c->endramp = 0;
    c->have_trial = false;
  } else {
    commit(c);
    if (p.iramp == 1 && p.Tramp - (a.t + hl) <= tiny && p.Tramp > a.t && c->endramp != 1) {
      const size_t n = (size_t)c->n3a * c->n2;
      hipLaunchKernelGGL(k_sub, dim3(nblk(n)), dim3(256), 0, c->st, n, c->a1st_add, (const double*)c->a_add_ramp);
      c->endramp = 1;
    }
    a.t = a.t + hl;
    corr = (std::fabs(err) < tiny) ? 2.0 : 0.9 * std::pow(err / toler, -(1.0 / 4.0));
    ++a.na;
  }
  a.h = a.h * corr;
  a.h = std::min({a.h, 0.95 * dt_stab, delt_out});
  a.h = std::max({a.h, 0.0001 * 0.95 * dt_stab, 0.0001 * delt_out});
}

void commit(hos_ctx* c) {
  if (!c->have_trial) throw std::runtime_error("no trial state to commit");
  std::swap(c->eta_hat, c->eta_hat_t);
  std::swap(c->phis_hat, c->phis_hat_t);
  std::swap(c->a1st_add, c->a1st_add_t);
  std::swap(c->a_add_ramp, c->a_add_ramp_t);
  std::swap(c->a1st, c->a1st_t);
  std::swap(c->eta1st, c->eta1st_t);
  std::swap(c->deta_hat, c->deta_hat_t);
  if (c->wmk2) {
    std::swap(c->a2nd_add, c->a2nd_add_t);
    std::swap(c->a2nd, c->a2nd_t);
    std::swap(c->eta2nd, c->eta2nd_t);
  }
  if (c->wmk3) std::swap(c->a3rd_add, c->a3rd_add_t);
  c->have_trial = false;
}

template <typename F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 1;
  }
}

}  // namespace

// ===========================================================================
// graph-captured attempts with the controller on the device (hos_advance_graph)
// ===========================================================================
namespace {

std::vector<const void*> graph_signature(hos_ctx* c) {
  return {c->eng, c->eta_hat, c->phis_hat, c->a1st_add, c->a_add_ramp, c->a1st, c->eta1st, c->deta_hat,
          c->eta_hat_t, c->phis_hat_t, c->a1st_add_t, c->a_add_ramp_t, c->a1st_t, c->eta1st_t, c->deta_hat_t};
}

void graph_release(hos_ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_sig.clear();
}

void graph_prepare(hos_ctx* c) {
  if (t_pass_log) throw std::runtime_error("graph capture while a pass log is recording");
  if (!c->wmk || c->wmk2)
    throw std::runtime_error("graph-captured attempts need a wavemaker tank with i_wmk = 1");
  if (c->feat_mode == 2) throw std::runtime_error("graph-captured attempts record features in mode 0 or 1 only");
  if (!c->dsc) {
    HIP_OK(hipStreamSynchronize(c->st));
    c->dsc = c->ar.get<double>(DS_COUNT);
    const int nf = (int)c->freq.size();
    c->freq_d = c->ar.get<double>(std::max(nf, 1));
    c->pre0_d = c->ar.get<double>(std::max(nf, 1));
    c->pim0_d = c->ar.get<double>(std::max(nf, 1));
    if (nf) {
      HIP_OK(hipMemcpy(c->freq_d, c->freq.data(), nf * sizeof(double), hipMemcpyHostToDevice));
      HIP_OK(hipMemcpy(c->pre0_d, c->pre0.data(), nf * sizeof(double), hipMemcpyHostToDevice));
      HIP_OK(hipMemcpy(c->pim0_d, c->pim0.data(), nf * sizeof(double), hipMemcpyHostToDevice));
    }
    HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&c->dsc_host), DS_COUNT * sizeof(double), hipHostMallocDefault));
    HIP_OK(hipDeviceSynchronize());
  }
  if (c->graph_exec && c->graph_sig == graph_signature(c)) return;
  graph_release(c);
  // one eager attempt first, so that every library kernel of the attempt is loaded and
  // initialised outside the capture; its trial result and wavemaker globals are discarded
  const double la = c->last_alpha, lat = c->last_alphat, lp = c->last_pos_top;
  double err = 0.0;
  rk_enqueue(c, 0.0, 1e-3);
  rk_finish(c, &err);
  c->last_alpha = la;
  c->last_alphat = lat;
  c->last_pos_top = lp;
  c->have_trial = false;
  HIP_OK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  c->dsc_active = c->dsc;
  try {
    hipLaunchKernelGGL(k_stage_scalars, dim3(1), dim3(256), 0, c->st, c->dsc, c->p.iramp, c->p.Tramp, c->p.T_stop,
                       c->p.T_mono, (int)c->freq.size(), (const double*)c->freq_d, (const double*)c->pre0_d,
                       (const double*)c->pim0_d, c->gz_top);
    rk_enqueue(c, 0.0, 0.0);
  } catch (...) {
    c->dsc_active = nullptr;
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(c->st, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  c->dsc_active = nullptr;
  HIP_OK(hipStreamEndCapture(c->st, &c->graph));
  HIP_OK(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
  c->graph_sig = graph_signature(c);
}

// the controller's state to the device step scalars
void graph_upload(hos_ctx* c, double t, double t_next, double h, double toler, double dt_stab, double delt_out) {
  double* d = c->dsc_host;
  std::fill(d, d + DS_COUNT, 0.0);
  d[DS_T] = t;
  d[DS_H] = h;
  d[DS_TNEXT] = t_next;
  d[DS_TOLER] = toler;
  d[DS_STAB] = dt_stab;
  d[DS_DOUT] = delt_out;
  d[DS_ENDRAMP] = c->endramp;
  d[DS_GALPHA] = c->last_alpha;
  d[DS_GPOS] = c->last_pos_top;
  d[DS_GALPHAT] = c->last_alphat;
  HIP_OK(hipMemcpyAsync(c->dsc, d, DS_COUNT * sizeof(double), hipMemcpyHostToDevice, c->st));
}

void graph_readback(hos_ctx* c) {
  HIP_OK(hipMemcpyAsync(c->dsc_host, c->dsc, DS_COUNT * sizeof(double), hipMemcpyDeviceToHost, c->st));
}

// Attempts that cannot overshoot t_next: each advances at most min(0.95 dt_stab, delt_out).
int graph_round(double t, double t_next, double dt_stab, double delt_out) {
  const double hmax = std::min(0.95 * dt_stab, delt_out);
  const double n = std::floor((t_next - t) / hmax);
  return (int)std::max(1.0, std::min(n, 4096.0));
}

void graph_finish(hos_ctx* c, double& t, double& h, int& na, int& nr) {
  const double* d = c->dsc_host;
  t = d[DS_T];
  h = d[DS_H];
  na = (int)d[DS_NACC];
  nr = (int)d[DS_NREJ];
  c->endramp = (int)d[DS_ENDRAMP];
  c->last_alpha = d[DS_GALPHA];
  c->last_pos_top = d[DS_GPOS];
  c->last_alphat = d[DS_GALPHAT];
  c->have_trial = false;
}

}  // namespace

// ===========================================================================
// learned truncation order (hos_advance_learned)
// ===========================================================================
namespace {

// the committed RK state as commit() swaps it (current pointers, lengths)
std::vector<std::pair<double*, size_t>> state_fields(hos_ctx* c) {
  const size_t ns = (size_t)c->n1 * c->n2, na = (size_t)c->n3a * c->n2;
  std::vector<std::pair<double*, size_t>> f = {{c->eta_hat, ns},    {c->phis_hat, ns}, {c->a1st_add, na},
                                               {c->a_add_ramp, na}, {c->a1st, ns},     {c->eta1st, ns},
                                               {c->deta_hat, ns}};
  if (c->wmk2) {
    f.push_back({c->a2nd_add, na});
    f.push_back({c->a2nd, ns});
    f.push_back({c->eta2nd, ns});
  }
  if (c->wmk3) f.push_back({c->a3rd_add, na});
  return f;
}

void learn_checkpoint(hos_ctx* c, double t, double h) {
  hos_ctx::Learn& L = c->learn;
  const auto f = state_fields(c);
  for (size_t i = 0; i < f.size(); ++i)
    HIP_OK(hipMemcpyAsync(L.ck[i], f[i].first, f[i].second * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  L.t_ck = t;
  L.h_ck = h;
  L.alpha_ck = c->last_alpha;
  L.alphat_ck = c->last_alphat;
  L.pos_ck = c->last_pos_top;
  L.endramp_ck = c->endramp;
}

void learn_restore(hos_ctx* c, double& t, double& h) {
  hos_ctx::Learn& L = c->learn;
  const auto f = state_fields(c);
  for (size_t i = 0; i < f.size(); ++i)
    HIP_OK(hipMemcpyAsync(f[i].first, L.ck[i], f[i].second * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  t = L.t_ck;
  h = L.h_ck;
  c->last_alpha = L.alpha_ck;
  c->last_alphat = L.alphat_ck;
  c->last_pos_top = L.pos_ck;
  c->endramp = L.endramp_ck;
  c->have_trial = false;
}

void learn_prepare(hos_ctx* c, const hos_order_policy& pol) {
  hos_ctx::Learn& L = c->learn;
  if (L.ready) return;
  HIP_OK(hipStreamSynchronize(c->st));
  const int cur = c->M;
  for (int m = std::max(pol.m_min, 1); m <= c->p.mHOS; ++m) {
    if (c->add_pot() && m == 2) continue;
    select_order(c, m);  // every engine the selection may use, allocated now
  }
  select_order(c, cur);
  const size_t ns = (size_t)c->n1 * c->n2;
  for (const auto& f : state_fields(c)) L.ck.push_back(c->ar.get<double>(f.second));
  L.judge_eta = c->ar.get<double>(ns);
  L.dpart = c->ar.get<double>(2 * 256);
  L.dn = c->ar.get<double>(2);
  HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&L.rec_host), kFeatRec * sizeof(double), hipHostMallocDefault));
  HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&L.pred_host), 64 * 64 * sizeof(float), hipHostMallocDefault));
  HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&L.dn_host), 2 * sizeof(double), hipHostMallocDefault));
  hosmlp::workspace_alloc(L.ws, 64);
  HIP_OK(hipDeviceSynchronize());  // set-up fills of new engines run on the null stream
  L.m = 0;                          // the first attempt runs at mHOS
  L.floor = pol.m_min;
  L.since = 0;
  L.ready = true;
}

// eta drift rate sqrt(E_j / E_eta) of order j in a feature record
double rate_of(const double* rec, int j) {
  const double et = rec[kFeatEta];
  if (!(et > 0.0)) return 0.0;
  double s = 0.0;
  for (int b = 0; b < kFeatBands; ++b) s += rec[j * kFeatBands + b];
  return std::sqrt(s / et);
}

// the rate of the order after the record's last one: the network's band
// energies, or the geometric extrapolation r_m^2 / r_{m-1}
double next_rate(const hos_ctx::Learn& L, bool have_net) {
  const double* rec = L.rec_host;
  const int mr = (int)rec[kFeatM];
  if (have_net && mr >= 1 && mr <= kPredOrders) {
    const float* y = L.pred_host + (size_t)(mr - 1) * 64;
    double s = 0.0;
    for (int b = 0; b < kFeatBands; ++b) s += std::pow(10.0, 5.0 * (double)y[b] - 15.0);
    return std::sqrt(s);
  }
  const double rm = rate_of(rec, mr), rm1 = mr >= 2 ? rate_of(rec, mr - 1) : 0.0;
  if (rm1 > 0.0) return rm * rm / rm1;
  return rm > 0.0 ? std::numeric_limits<double>::infinity() : 0.0;
}

// the lowest order q >= max(m_min, floor) whose first omitted order's rate is
// below safety * rate_tol (actual rates up to the record's order, the
// estimate one beyond it; one order up when that is not enough)
int choose_order(hos_ctx* c, const hos_order_policy& pol, bool have_net, double* rate_used) {
  const hos_ctx::Learn& L = c->learn;
  const double* rec = L.rec_host;
  const int mr = (int)rec[kFeatM], Mx = c->p.mHOS;
  const double thr = pol.safety * pol.rate_tol;
  if (!(rec[kFeatEta] > 0.0)) {  // a surface at rest tells nothing about the orders: full order
    *rate_used = 0.0;
    return Mx;
  }
  int q = std::max(std::max(pol.m_min, L.floor), 1);
  double r = 0.0;
  while (q < Mx) {
    if (q + 1 <= mr)
      r = rate_of(rec, q + 1);
    else if (q + 1 == mr + 1)
      r = next_rate(L, have_net);
    else
      break;
    if (L.kappa * r <= thr) break;
    ++q;
  }
  if (c->add_pot() && q == 2) q = 3;
  q = std::min(q, Mx);
  *rate_used = (q < Mx) ? r : 0.0;
  return q;
}

// one attempt at order m, its first right-hand side's band energies and (with
// a network) the predictions, all copied to pinned host memory on the stream
void learn_enqueue(hos_ctx* c, const hos_mlp* net, double t, double hl, int m) {
  hos_ctx::Learn& L = c->learn;
  select_order(c, m);
  L.m_attempt = m;
  c->feat_mode = 1;
  rk_enqueue(c, t, hl);
  hipStream_t st = c->st;
  if (net) {
    hipLaunchKernelGGL(k_feat_rows, dim3(64 * kRowIn / 256), dim3(256), 0, st, (const double*)c->feat, m, L.ws.x);
    hosmlp::forward(net, 64, L.ws, st);
    HIP_OK(hipMemcpyAsync(L.pred_host, L.ws.y, 64 * 64 * sizeof(float), hipMemcpyDeviceToHost, st));
  }
  HIP_OK(hipMemcpyAsync(L.rec_host, c->feat, kFeatRec * sizeof(double), hipMemcpyDeviceToHost, st));
  L.stats.attempts[m]++;
}

struct LearnMember {
  hos_ctx* c;
  Advance a;
  double dt_stab, delt_out;
  bool final_attempt = false;
  double rate_sel = 0.0;  // the rate estimate behind the order of the attempt in flight
};

void learn_launch(LearnMember& mb, double t_next, const hos_mlp* net, const hos_order_policy& pol) {
  hos_ctx* c = mb.c;
  hos_ctx::Learn& L = c->learn;
  const double hl = (t_next - mb.a.t < mb.a.h) ? t_next - mb.a.t : mb.a.h;
  mb.final_attempt = !(mb.a.t + hl < t_next);
  int m = (L.m == 0) ? c->p.mHOS : L.m;
  m = std::max(m, std::max(pol.m_min, L.floor));
  if (c->add_pot() && m == 2) m = 3;
  m = std::min(m, c->p.mHOS);
  L.hl = hl;
  L.phase = 0;
  learn_enqueue(c, net, mb.a.t, hl, m);
}

// Completes the member's attempt in flight; returns true when the member has
// enqueued another attempt (a judge, or its next attempt before t_next).
bool learn_finish(LearnMember& mb, double t_next, double toler, const hos_mlp* net, const hos_order_policy& pol) {
  hos_ctx* c = mb.c;
  hos_ctx::Learn& L = c->learn;
  const int Mx = c->p.mHOS;
  const size_t ns = (size_t)c->n1 * c->n2;
  double err = 0.0;
  rk_finish(c, &err);
  const double hl = L.hl;
  if (L.phase == 0) {
    const int m = L.m_attempt;
    double rate_next = 0.0;
    const int m_next = choose_order(c, pol, net != nullptr, &rate_next);
    const bool accepts = !(err > toler && mb.a.h > 0.001 * 0.95 * mb.dt_stab);
    if (m < Mx && accepts && (L.since + 1 >= pol.judge_every || mb.final_attempt || m_next != m)) {
      // the judge: the same attempt at mHOS from the same state
      HIP_OK(hipMemcpyAsync(L.judge_eta, c->eta_hat_t, ns * sizeof(double), hipMemcpyDeviceToDevice, c->st));
      L.err_m = err;
      L.phase = 1;
      L.stats.judges++;
      const int judged = m;
      learn_enqueue(c, net, mb.a.t, hl, Mx);
      L.m_attempt = judged;  // (the order under judgment)
      return true;
    }
    const int na0 = mb.a.na;
    control(c, mb.a, hl, err, toler, mb.dt_stab, mb.delt_out);
    if (mb.a.na > na0) {
      if (m < Mx) {
        L.since++;
      } else if (L.since == 0 && ++L.full_run >= std::max(pol.judge_every, 1)) {
        // judge_every accepted full-order attempts: progress that needs no judge, and the
        // floor of an earlier reversion relaxes by one order
        learn_checkpoint(c, mb.a.t, mb.a.h);
        L.floor = std::max(pol.m_min, L.floor - 1);
        L.full_run = 0;
      }
    }
    L.m = m_next;
    mb.rate_sel = rate_next;
  } else {
    const int judged = L.m_attempt;
    hipLaunchKernelGGL(k_diff_norm, dim3(256), dim3(256), 0, c->st, ns, (const double*)L.judge_eta,
                       (const double*)c->eta_hat_t, L.dpart);
    hipLaunchKernelGGL(k_diff_reduce, dim3(1), dim3(256), 0, c->st, 256, (const double*)L.dpart, L.dn);
    HIP_OK(hipMemcpyAsync(L.dn_host, L.dn, 2 * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
    const double rate = L.dn_host[1] > 0.0 ? std::sqrt(L.dn_host[0] / L.dn_host[1]) / hl : 0.0;
    // calibration: up at once to the observed ratio, down by a fifth per judge
    if (mb.rate_sel > 0.0 && std::isfinite(mb.rate_sel))
      L.kappa = std::min(1e6, std::max(0.1, std::max(rate / mb.rate_sel, 0.8 * L.kappa)));
    if (rate > L.stats.max_judged_rate) {
      L.stats.max_judged_rate = rate;
      L.stats.max_pred_rate = mb.rate_sel;
    }
    if (rate > pol.rate_tol) {
      // reversion: back to the last passed judge, above the judged order
      learn_restore(c, mb.a.t, mb.a.h);
      L.floor = std::min(Mx, judged + 1);
      L.since = 0;
      L.m = L.floor;
      L.stats.reverts++;
    } else {
      const int na0 = mb.a.na;
      control(c, mb.a, hl, err, toler, mb.dt_stab, mb.delt_out);  // the mHOS attempt's decision
      if (mb.a.na > na0) {
        L.since = 0;
        learn_checkpoint(c, mb.a.t, mb.a.h);
        L.floor = std::max(pol.m_min, L.floor - 1);
      }
      double rate_next = 0.0;
      L.m = choose_order(c, pol, net != nullptr, &rate_next);  // from the full-order record
      mb.rate_sel = rate_next;
    }
  }
  if (mb.a.t < t_next) {
    learn_launch(mb, t_next, net, pol);
    return true;
  }
  return false;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
namespace {
// hos_rhs2 / hos_rhs3: one solveHOS2 with the higher-order wavemaker state
void rhs_higher(hos_ctx* c, double t, double delta_t, const double* G1, const double* G2, const double* a1st,
                const double* eta1st, const double* a_add_1, const double* a2nd, const double* eta2nd,
                const double* a_add_2, const double* a_add_3, double* dG1, double* dG2, double* da1st,
                double* deta1st, double* da_add_1, double* da_add_ramp, double* da2nd, double* deta2nd,
                double* da_add_2, double* da_add_3) {
  const size_t ns = (size_t)c->n1 * c->n2, na = (size_t)c->n3a * c->n2;
  hipStream_t st = c->st;
  HIP_OK(hipMemcpyAsync(c->G1m, G1, ns * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(c->G2m, G2, ns * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(c->a1m, a1st, ns * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(c->e1m, eta1st, ns * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(c->aadm, a_add_1, na * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(c->a2m, a2nd, ns * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(c->e2m, eta2nd, ns * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_OK(hipMemcpyAsync(c->aad2m, a_add_2, na * sizeof(double), hipMemcpyHostToDevice, st));
  if (c->wmk3) {
    if (a_add_3)
      HIP_OK(hipMemcpyAsync(c->aad3m, a_add_3, na * sizeof(double), hipMemcpyHostToDevice, st));
    else
      HIP_OK(hipMemsetAsync(c->aad3m, 0, na * sizeof(double), st));
  }
  rhs_eval(c, t, delta_t, c->G1m, c->G2m, c->a1m, c->e1m, c->aadm, 0, c->a2m, c->e2m, c->aad2m, c->aad3m);
  HIP_OK(hipMemcpyAsync(dG1, c->kG1[0], ns * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_OK(hipMemcpyAsync(dG2, c->kG2[0], ns * sizeof(double), hipMemcpyDeviceToHost, st));
  if (da1st) HIP_OK(hipMemcpyAsync(da1st, c->ka1[0], ns * sizeof(double), hipMemcpyDeviceToHost, st));
  if (deta1st) HIP_OK(hipMemcpyAsync(deta1st, c->ke1[0], ns * sizeof(double), hipMemcpyDeviceToHost, st));
  if (da_add_1) HIP_OK(hipMemcpyAsync(da_add_1, c->kad1[0], na * sizeof(double), hipMemcpyDeviceToHost, st));
  if (da_add_ramp) HIP_OK(hipMemcpyAsync(da_add_ramp, c->kadr[0], na * sizeof(double), hipMemcpyDeviceToHost, st));
  if (da2nd) HIP_OK(hipMemcpyAsync(da2nd, c->ka2[0], ns * sizeof(double), hipMemcpyDeviceToHost, st));
  if (deta2nd) HIP_OK(hipMemcpyAsync(deta2nd, c->ke2[0], ns * sizeof(double), hipMemcpyDeviceToHost, st));
  if (da_add_2) HIP_OK(hipMemcpyAsync(da_add_2, c->kad2[0], na * sizeof(double), hipMemcpyDeviceToHost, st));
  if (da_add_3 && c->wmk3)
    HIP_OK(hipMemcpyAsync(da_add_3, c->kad3[0], na * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_OK(hipStreamSynchronize(st));
}
}  // namespace

namespace hosnwt_err {
void set(const char* msg) { g_last_error = msg; }
}  // namespace hosnwt_err

extern "C" {

const char* hos_last_error(void) { return g_last_error.c_str(); }

int hos_create(const hos_params* pin, hos_ctx** out) {
  return guard([&] {
    if (!pin || !out) throw std::runtime_error("null argument");
    check_params(*pin);
    int ndev = 0;
    HIP_OK(hipGetDeviceCount(&ndev));
    if (ndev < 1) throw std::runtime_error("no GPU device");
    auto c = std::make_unique<hos_ctx>();
    HIP_OK(hipGetDevice(&c->device));
    c->p = *pin;
    const hos_params& p = c->p;
    c->n1 = p.n1;
    c->n2 = p.n2;
    c->n3 = p.n3;
    c->n3a = L_ADD * (p.n3 - 1);
    c->M = p.mHOS;
    {
      int q1, q2;
      order_grid(p, p.mHOS, q1, q2, c->N1b, c->N2b);
    }
    c->pix = M_PI / p.xlen;
    c->piy = p.n2 > 1 ? M_PI / p.ylen : 0.0;
    c->dx = p.xlen / (p.n1 - 1);
    c->dy = p.n2 > 1 ? p.ylen / (p.n2 - 1) : 0.0;
    c->delz = 1.0 / (p.n3 - 1);
    const bool wave_case = p.icase == 2 || p.icase == 3 || p.icase == 31 || p.icase == 32 || p.icase == 33 ||
                           p.icase == 4 || p.icase == 41 || p.icase == 42;
    c->wmk = wave_case && (p.igeom == 1 || p.igeom == 2);
    c->dip = wave_case && p.igeom == 3;
    c->stats.big_n1 = c->N1b;
    c->stats.big_n2 = c->N2b;
    HIP_OK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    c->gs64.build(p.n1, p.n2);
    const size_t ns = (size_t)p.n1 * p.n2, na = (size_t)c->n3a * p.n2;
    Arena& ar = c->ar;
    c->eta_hat = ar.get<double>(ns); c->phis_hat = ar.get<double>(ns);
    c->a1st_add = ar.get<double>(na); c->a_add_ramp = ar.get<double>(na);
    c->a1st = ar.get<double>(ns); c->eta1st = ar.get<double>(ns);
    c->eta_hat_t = ar.get<double>(ns); c->phis_hat_t = ar.get<double>(ns);
    c->a1st_add_t = ar.get<double>(na); c->a_add_ramp_t = ar.get<double>(na);
    c->a1st_t = ar.get<double>(ns); c->eta1st_t = ar.get<double>(ns);
    c->deta_hat = ar.get<double>(ns); c->deta_hat_t = ar.get<double>(ns);
    c->G1 = ar.get<double>(ns); c->G2 = ar.get<double>(ns); c->G1m = ar.get<double>(ns); c->G2m = ar.get<double>(ns);
    c->aadm = ar.get<double>(na); c->a1m = ar.get<double>(ns); c->e1m = ar.get<double>(ns);
    for (int s = 0; s < 5; ++s) {
      c->kG1[s] = ar.get<double>(ns); c->kG2[s] = ar.get<double>(ns);
      c->kad1[s] = ar.get<double>(na); c->kadr[s] = ar.get<double>(na);
      c->ka1[s] = ar.get<double>(ns); c->ke1[s] = ar.get<double>(ns);
    }
    c->wmk2 = c->wmk && p.i_wmk >= 2;
    c->wmk3 = c->wmk && p.i_wmk == 3;
    if (c->wmk2) {
      c->a2nd_add = ar.get<double>(na); c->a2nd = ar.get<double>(ns); c->eta2nd = ar.get<double>(ns);
      c->a2nd_add_t = ar.get<double>(na); c->a2nd_t = ar.get<double>(ns); c->eta2nd_t = ar.get<double>(ns);
      c->a2m = ar.get<double>(ns); c->e2m = ar.get<double>(ns); c->aad2m = ar.get<double>(na);
      for (int s = 0; s < 5; ++s) {
        c->ka2[s] = ar.get<double>(ns); c->ke2[s] = ar.get<double>(ns); c->kad2[s] = ar.get<double>(na);
      }
    }
    if (c->wmk3) {
      c->a3rd_add = ar.get<double>(na); c->a3rd_add_t = ar.get<double>(na); c->aad3m = ar.get<double>(na);
      for (int s = 0; s < 5; ++s) c->kad3[s] = ar.get<double>(na);
    }
    const size_t nbk = nblk(std::max(ns, na)) + 1;
    c->partial = ar.get<double>(nbk);
    c->pmax_stride = nbk;
    c->pmax_e = ar.get<double>(kMaxFinal * nbk);
    c->pmax_x = ar.get<double>(kMaxFinal * nbk);
    c->red6 = ar.get<double>(10);
    HIP_OK(hipHostMalloc(reinterpret_cast<void**>(&c->red_host), 16 * sizeof(double), hipHostMallocDefault));
    c->phys_tmp = ar.get<double>(ns);
    c->phi_add_last = ar.get<double>(ns);
    c->feat_cap = 512;
    c->feat = ar.get<double>((size_t)c->feat_cap * kFeatRec);
    c->feat_part = ar.get<double>((size_t)kFeatGrid * kFeatRec);
    c->nu_d = nullptr;
    if (p.nu) {
      c->nu_d = ar.get<double>(ns);
      HIP_OK(hipMemcpy(c->nu_d, p.nu, ns * sizeof(double), hipMemcpyHostToDevice));
      // a beach of x alone (the reference's front beach, initial_condition.f90:724-794)
      bool sep = true;
      for (int j = 1; j < p.n2 && sep; ++j)
        for (int i = 0; i < p.n1 && sep; ++i) sep = p.nu[i + (size_t)p.n1 * j] == p.nu[i];
      if (sep) c->h_nu_x.assign(p.nu, p.nu + p.n1);
    }
    c->p.nu = nullptr;
    if (filter_active(p.coeffilt[0]) || (p.n2 > 1 && filter_active(p.coeffilt[1]))) {
      const std::vector<double> w = filter_weights_2d(p.n1, p.n2, p.coeffilt[0], p.coeffilt[1]);
      c->filt_fs = ar.get<double>(ns);
      HIP_OK(hipMemcpy(c->filt_fs, w.data(), ns * sizeof(double), hipMemcpyHostToDevice));
    }
    if (filter_active(p.coeffilt[2])) {
      const std::vector<double> w = filter_weights(c->n3a, p.coeffilt[2]);
      c->filt_z = ar.get<double>(c->n3a);
      HIP_OK(hipMemcpy(c->filt_z, w.data(), c->n3a * sizeof(double), hipMemcpyHostToDevice));
    }
    c->freq.assign(p.freq, p.freq + p.n_freq);
    c->pre0.resize(p.n_freq);
    c->pim0.resize(p.n_freq);
    for (int f = 0; f < p.n_freq; ++f) {
      c->pre0[f] = p.pos_ini_re[(size_t)p.n2 * f];
      c->pim0[f] = p.pos_ini_im[(size_t)p.n2 * f];
    }
    c->gz_top = gzed(0.0, p.igeom, p.d_hinge);
    {
      const size_t nfy = (size_t)p.n2 * p.n_freq;
      if (p.n_freq) {
        c->h_pos_re.assign(p.pos_ini_re, p.pos_ini_re + nfy);
        c->h_pos_im.assign(p.pos_ini_im, p.pos_ini_im + nfy);
        if (p.dposdy_ini_re && p.dposdy_ini_im) {
          c->h_dy_re.assign(p.dposdy_ini_re, p.dposdy_ini_re + nfy);
          c->h_dy_im.assign(p.dposdy_ini_im, p.dposdy_ini_im + nfy);
        }
      }
      if (p.nu_x) c->h_nux.assign(p.nu_x, p.nu_x + ns);
      if (p.nu_y) c->h_nuy.assign(p.nu_y, p.nu_y + ns);
    }
    if (c->dip) dipoles_setup(c.get(), p);
    select_order(c.get(), p.mHOS);
    c->p.freq = nullptr;
    c->p.pos_ini_re = c->p.pos_ini_im = nullptr;
    c->p.dposdy_ini_re = c->p.dposdy_ini_im = c->p.nu_x = c->p.nu_y = nullptr;
    c->p.dip_freq = c->p.dip_mux_re = c->p.dip_mux_im = c->p.dip_muz_re = c->p.dip_muz_im = c->p.dip_z = nullptr;
    HIP_OK(hipDeviceSynchronize());
    *out = c.release();
  });
}

int hos_destroy(hos_ctx* c) {
  return guard([&] {
    if (!c) return;
    (void)hipStreamSynchronize(c->st);
    if (c->plog && t_pass_log == c->plog.get()) t_pass_log = nullptr;
    c->plog.reset();
    for (auto& e : c->engines) e.reset();
    c->eng = nullptr;
    c->gs64.release();
    hipStream_t st = c->st;
    delete c;
    (void)hipStreamDestroy(st);
  });
}

int hos_initialize(hos_ctx* c) {
  return guard([&] {
    const size_t ns = (size_t)c->n1 * c->n2, na = (size_t)c->n3a * c->n2;
    hipStream_t st = c->st;
    HIP_OK(hipMemsetAsync(c->a1st, 0, ns * sizeof(double), st));
    HIP_OK(hipMemsetAsync(c->eta1st, 0, ns * sizeof(double), st));
    HIP_OK(hipMemsetAsync(c->a1st_add, 0, na * sizeof(double), st));
    HIP_OK(hipMemsetAsync(c->a_add_ramp, 0, na * sizeof(double), st));
    HIP_OK(hipMemsetAsync(c->deta_hat, 0, ns * sizeof(double), st));
    // the free surface starts at rest (initial_condition.f90:329-348)
    HIP_OK(hipMemsetAsync(c->eta_hat, 0, ns * sizeof(double), st));
    HIP_OK(hipMemsetAsync(c->phis_hat, 0, ns * sizeof(double), st));
    c->have_trial = false;
    if (c->wmk2) {  // a2nd, eta2nd, a2nd_add start at zero (initial_condition.f90:370-373)
      HIP_OK(hipMemsetAsync(c->a2nd, 0, ns * sizeof(double), st));
      HIP_OK(hipMemsetAsync(c->eta2nd, 0, ns * sizeof(double), st));
      HIP_OK(hipMemsetAsync(c->a2nd_add, 0, na * sizeof(double), st));
    }
    if (c->wmk3) HIP_OK(hipMemsetAsync(c->a3rd_add, 0, na * sizeof(double), st));  // :376
    c->endramp = 0;
    if (!c->wmk) {
      c->last_alpha = 0.0;
      c->last_alphat = 0.0;
      c->last_pos_top = 0.0;
      if (c->dip) {
        // the reference's solveHOS at t = 0 (initial_condition.f90:467-478) sets the
        // dipole potential of the rest state; one right-hand side at t = 0 does the same
        hipLaunchKernelGGL(k_step_begin, dim3(nblk(ns)), dim3(256), 0, st, c->n1, c->n2, c->pix, c->piy,
                           (const double*)c->eta_hat, (const double*)c->phis_hat, c->G1, c->G2);
        rhs_eval(c, 0.0, 0.0, c->G1, c->G2, c->a1st, c->eta1st, c->a1st_add, 0, c->a2nd, c->eta2nd, c->a2nd_add,
                 c->a3rd_add);
      }
      HIP_OK(hipStreamSynchronize(st));
      return;
    }
    // a1st_add from -alpha X_t - alpha' X (initial_condition.f90:372-392)
    Engine<double>* e64 = dynamic_cast<Engine<double>*>(c->eng);
    Engine<float>* e32 = dynamic_cast<Engine<float>*>(c->eng);
    double al, alt, altt;
    ramp_eval(c->p, 0.0, al, alt, altt);
    auto run = [&](auto* e) {
      hipLaunchKernelGGL(k_pos, dim3(c->n2), dim3(256), 0, st, c->n2, c->n3, c->p.n_freq, 0.0, (const double*)e->freq,
                         (const double*)e->pre, (const double*)e->pim, (const double*)e->gz, e->Xp, e->Xpt, e->Xptt);
      hipLaunchKernelGGL(k_paddle_rhs, dim3(nblk((size_t)c->n3a * c->n2)), dim3(256), 0, st, c->n2, c->n3, c->n3a, c->delz,
                         (const double*)e->Xp, (const double*)e->Xpt, (const double*)e->Xptt, alt, al, 0.0, 0.0, 0.0,
                         0.0, e->cont, e->f1, (double*)nullptr);
      e->paddle_modes(1, c->a1st_add, (double*)nullptr);
    };
    if (e64) run(e64);
    else run(e32);
    HIP_OK(hipMemcpyAsync(c->a_add_ramp, c->a1st_add, na * sizeof(double), hipMemcpyDeviceToDevice, st));
    // one RHS at t = 0 sets phi_add (phiadd_SL with da1st_add) and the wavemaker globals
    hipLaunchKernelGGL(k_step_begin, dim3(nblk(ns)), dim3(256), 0, st, c->n1, c->n2, c->pix, c->piy,
                       (const double*)c->eta_hat, (const double*)c->phis_hat, c->G1, c->G2);
    rhs_eval(c, 0.0, 0.0, c->G1, c->G2, c->a1st, c->eta1st, c->a1st_add, 0, c->a2nd, c->eta2nd, c->a2nd_add,
             c->a3rd_add);
    HIP_OK(hipStreamSynchronize(st));
  });
}

int hos_set_state(hos_ctx* c, const double* eta, const double* phis, const double* a1st_add,
                  const double* a_add_ramp, const double* a1st, const double* eta1st) {
  return guard([&] {
    const size_t ns = (size_t)c->n1 * c->n2, na = (size_t)c->n3a * c->n2;
    hipStream_t st = c->st;
    // eta, phis physical -> modal (runge_kutta.f90:128-134)
    // staged through G1m: the non-dense base transform uses phys_tmp as its intermediate
    HIP_OK(hipMemcpyAsync(c->G1m, eta, ns * sizeof(double), hipMemcpyHostToDevice, st));
    c->eng->base_cc64(FWD, c->G1m, c->eta_hat);
    HIP_OK(hipMemcpyAsync(c->G1m, phis, ns * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
    c->eng->base_cc64(FWD, c->G1m, c->phis_hat);
    if (a1st_add) HIP_OK(hipMemcpyAsync(c->a1st_add, a1st_add, na * sizeof(double), hipMemcpyHostToDevice, st));
    if (a_add_ramp) HIP_OK(hipMemcpyAsync(c->a_add_ramp, a_add_ramp, na * sizeof(double), hipMemcpyHostToDevice, st));
    if (a1st) HIP_OK(hipMemcpyAsync(c->a1st, a1st, ns * sizeof(double), hipMemcpyHostToDevice, st));
    if (eta1st) HIP_OK(hipMemcpyAsync(c->eta1st, eta1st, ns * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
    c->have_trial = false;
  });
}

static void to_phys(hos_ctx* c, const double* modal, double* host) {
  const size_t ns = (size_t)c->n1 * c->n2;
  c->eng->base_cc64(INV, modal, c->G1m);
  HIP_OK(hipMemcpyAsync(host, c->G1m, ns * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIP_OK(hipStreamSynchronize(c->st));
}

int hos_get_state(hos_ctx* c, double* eta, double* phis, double* a1st_add, double* a_add_ramp, double* a1st,
                  double* eta1st) {
  return guard([&] {
    const size_t ns = (size_t)c->n1 * c->n2, na = (size_t)c->n3a * c->n2;
    if (eta) to_phys(c, c->eta_hat, eta);
    if (phis) to_phys(c, c->phis_hat, phis);
    if (a1st_add) HIP_OK(hipMemcpy(a1st_add, c->a1st_add, na * sizeof(double), hipMemcpyDeviceToHost));
    if (a_add_ramp) HIP_OK(hipMemcpy(a_add_ramp, c->a_add_ramp, na * sizeof(double), hipMemcpyDeviceToHost));
    if (a1st) HIP_OK(hipMemcpy(a1st, c->a1st, ns * sizeof(double), hipMemcpyDeviceToHost));
    if (eta1st) HIP_OK(hipMemcpy(eta1st, c->eta1st, ns * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int hos_set_state2(hos_ctx* c, const double* a2nd_add, const double* a2nd, const double* eta2nd) {
  return guard([&] {
    if (!c->wmk2) throw std::runtime_error("no second-order wavemaker state (i_wmk = 2, 3 only)");
    const size_t ns = (size_t)c->n1 * c->n2, na = (size_t)c->n3a * c->n2;
    if (a2nd_add) HIP_OK(hipMemcpyAsync(c->a2nd_add, a2nd_add, na * sizeof(double), hipMemcpyHostToDevice, c->st));
    if (a2nd) HIP_OK(hipMemcpyAsync(c->a2nd, a2nd, ns * sizeof(double), hipMemcpyHostToDevice, c->st));
    if (eta2nd) HIP_OK(hipMemcpyAsync(c->eta2nd, eta2nd, ns * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
    c->have_trial = false;
  });
}

int hos_get_state2(hos_ctx* c, double* a2nd_add, double* a2nd, double* eta2nd) {
  return guard([&] {
    if (!c->wmk2) throw std::runtime_error("no second-order wavemaker state (i_wmk = 2, 3 only)");
    const size_t ns = (size_t)c->n1 * c->n2, na = (size_t)c->n3a * c->n2;
    HIP_OK(hipStreamSynchronize(c->st));
    if (a2nd_add) HIP_OK(hipMemcpy(a2nd_add, c->a2nd_add, na * sizeof(double), hipMemcpyDeviceToHost));
    if (a2nd) HIP_OK(hipMemcpy(a2nd, c->a2nd, ns * sizeof(double), hipMemcpyDeviceToHost));
    if (eta2nd) HIP_OK(hipMemcpy(eta2nd, c->eta2nd, ns * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int hos_get_physical(hos_ctx* c, double* eta, double* phis, double* deta, double* phi_add) {
  return guard([&] {
    const size_t ns = (size_t)c->n1 * c->n2;
    if (eta) to_phys(c, c->eta_hat, eta);
    if (phis) to_phys(c, c->phis_hat, phis);
    if (deta) to_phys(c, c->deta_hat, deta);
    if (phi_add) {
      HIP_OK(hipMemcpyAsync(phi_add, c->phi_add_last, ns * sizeof(double), hipMemcpyDeviceToHost, c->st));
      HIP_OK(hipStreamSynchronize(c->st));
    }
  });
}

int hos_rk_step(hos_ctx* c, double t, double h, double* err) {
  return guard([&] { rk_attempt(c, t, h, err); });
}

int hos_commit(hos_ctx* c) {
  return guard([&] { commit(c); });
}

int hos_end_ramp(hos_ctx* c) {
  return guard([&] {
    const size_t na = (size_t)c->n3a * c->n2;
    hipLaunchKernelGGL(k_sub, dim3(nblk(na)), dim3(256), 0, c->st, na, c->a1st_add, (const double*)c->a_add_ramp);
    HIP_OK(hipStreamSynchronize(c->st));
  });
}

int hos_advance(hos_ctx* c, double* t_cur, double t_next, double* h_rk, double toler, double dt_stab,
                double delt_out, int* n_acc, int* n_rej) {
  return guard([&] {
    Advance a{*t_cur, *h_rk, 0, 0};
    while (a.t < t_next) {
      double hl = a.h;
      if (t_next - a.t < a.h) hl = t_next - a.t;
      double err = 0.0;
      rk_attempt(c, a.t, hl, &err);
      control(c, a, hl, err, toler, dt_stab, delt_out);
    }
    HIP_OK(hipStreamSynchronize(c->st));
    *t_cur = a.t;
    *h_rk = a.h;
    if (n_acc) *n_acc = a.na;
    if (n_rej) *n_rej = a.nr;
  });
}

int hos_advance_ensemble(int n, hos_ctx* const* ctx, double* t_cur, double t_next, double* h_rk, double toler,
                         const double* dt_stab, const double* delt_out, int* n_acc, int* n_rej) {
  return guard([&] {
    if (n < 0) throw std::runtime_error("negative ensemble size");
    std::vector<Advance> a(n);
    std::vector<double> hl(n, 0.0);
    for (int i = 0; i < n; ++i) {
      if (!ctx[i]) throw std::runtime_error("null context in the ensemble");
      for (int j = 0; j < i; ++j)
        if (ctx[j] == ctx[i]) throw std::runtime_error("a context appears twice in the ensemble");
      a[i] = Advance{t_cur[i], h_rk[i], 0, 0};
    }
    // Host threads: an attempt is a few hundred small launches, so a ragged
    // ensemble of small tanks is bound by the launch rate of one host thread.
    // Members are dealt to the threads longest first (cost ~ extended points x
    // M / dt_stab), each thread keeping all of its members' attempts in flight.
    // 4 threads measured best (2-D ensemble of 32: 286 / 356 / 347 attempts/s
    // with 1 / 4 / 8).
    int nth = 4;
    if (const char* e = std::getenv("HOSNWT_ENSEMBLE_THREADS")) nth = std::max(1, std::atoi(e));
    nth = std::max(1, std::min(nth, n));
    std::vector<int> order(n);
    std::vector<double> cost(n);
    for (int i = 0; i < n; ++i) {
      order[i] = i;
      cost[i] = (double)ctx[i]->N1b * ctx[i]->N2b * ctx[i]->M * std::max(0.0, t_next - a[i].t) / dt_stab[i];
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    std::vector<std::vector<int>> part(nth);
    std::vector<double> load(nth, 0.0);
    for (int i : order) {
      const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      part[k].push_back(i);
      load[k] += cost[i];
    }
    std::mutex em;
    std::string first_error;
    std::atomic<bool> failed{false};
    auto run = [&](const std::vector<int>& mine) {
      try {
        if (!mine.empty()) HIP_OK(hipSetDevice(ctx[mine[0]]->device));
        auto launch = [&](int i) {
          hl[i] = (t_next - a[i].t < a[i].h) ? t_next - a[i].t : a[i].h;
          rk_enqueue(ctx[i], a[i].t, hl[i]);
        };
        // every member's attempt is in flight on its own streams before the first
        // wait; a member's next attempt is enqueued as soon as its decision is taken
        for (int i : mine)
          if (a[i].t < t_next) launch(i);
        for (bool any = true; any && !failed.load();) {
          any = false;
          for (int i : mine) {
            if (!(a[i].t < t_next)) continue;
            double err = 0.0;
            rk_finish(ctx[i], &err);
            control(ctx[i], a[i], hl[i], err, toler, dt_stab[i], delt_out[i]);
            if (a[i].t < t_next) {
              launch(i);
              any = true;
            }
          }
        }
        for (int i : mine) HIP_OK(hipStreamSynchronize(ctx[i]->st));
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(em);
        if (first_error.empty()) first_error = e.what();
        failed = true;
      }
    };
    if (nth == 1) {
      run(part[0]);
    } else {
      std::vector<std::thread> th;
      for (int k = 0; k < nth; ++k) th.emplace_back(run, std::cref(part[k]));
      for (auto& t : th) t.join();
    }
    // the members' committed times, proposals and counts, also when one failed
    for (int i = 0; i < n; ++i) {
      t_cur[i] = a[i].t;
      h_rk[i] = a[i].h;
      if (n_acc) n_acc[i] = a[i].na;
      if (n_rej) n_rej[i] = a[i].nr;
    }
    if (!first_error.empty()) throw std::runtime_error(first_error);
  });
}

int hos_advance_graph(hos_ctx* c, double* t_cur, double t_next, double* h_rk, double toler, double dt_stab,
                      double delt_out, int* n_acc, int* n_rej) {
  return guard([&] {
    if (!c || !t_cur || !h_rk) throw std::runtime_error("null argument");
    graph_prepare(c);
    graph_upload(c, *t_cur, t_next, *h_rk, toler, dt_stab, delt_out);
    double t = *t_cur, h = *h_rk;
    int na = 0, nr = 0;
    while (t < t_next) {
      const int k = graph_round(t, t_next, dt_stab, delt_out);
      for (int r = 0; r < k; ++r) HIP_OK(hipGraphLaunch(c->graph_exec, c->st));
      graph_readback(c);
      HIP_OK(hipStreamSynchronize(c->st));
      graph_finish(c, t, h, na, nr);
    }
    *t_cur = t;
    *h_rk = h;
    if (n_acc) *n_acc = na;
    if (n_rej) *n_rej = nr;
  });
}

int hos_advance_ensemble_graph(int n, hos_ctx* const* ctx, double* t_cur, double t_next, double* h_rk,
                               double toler, const double* dt_stab, const double* delt_out, int* n_acc,
                               int* n_rej) {
  return guard([&] {
    if (n < 0) throw std::runtime_error("negative ensemble size");
    for (int i = 0; i < n; ++i) {
      if (!ctx[i]) throw std::runtime_error("null context in the ensemble");
      for (int j = 0; j < i; ++j)
        if (ctx[j] == ctx[i]) throw std::runtime_error("a context appears twice in the ensemble");
    }
    std::vector<double> t(t_cur, t_cur + n), h(h_rk, h_rk + n);
    std::vector<int> na(n, 0), nr(n, 0);
    for (int i = 0; i < n; ++i) {
      graph_prepare(ctx[i]);
      graph_upload(ctx[i], t[i], t_next, h[i], toler, dt_stab[i], delt_out[i]);
    }
    // rounds: every unfinished member's replays go onto its own stream, then one wait for all
    for (bool any = true; any;) {
      any = false;
      for (int i = 0; i < n; ++i) {
        if (!(t[i] < t_next)) continue;
        const int k = graph_round(t[i], t_next, dt_stab[i], delt_out[i]);
        for (int r = 0; r < k; ++r) HIP_OK(hipGraphLaunch(ctx[i]->graph_exec, ctx[i]->st));
        graph_readback(ctx[i]);
        any = true;
      }
      for (int i = 0; i < n; ++i) {
        if (!(t[i] < t_next)) continue;
        HIP_OK(hipStreamSynchronize(ctx[i]->st));
        graph_finish(ctx[i], t[i], h[i], na[i], nr[i]);
      }
    }
    for (int i = 0; i < n; ++i) {
      t_cur[i] = t[i];
      h_rk[i] = h[i];
      if (n_acc) n_acc[i] = na[i];
      if (n_rej) n_rej[i] = nr[i];
    }
  });
}

int hos_learned_last(hos_ctx* c, double* rec, float* pred) {
  return guard([&] {
    if (!c) throw std::runtime_error("null context");
    const hos_ctx::Learn& L = c->learn;
    if (!L.ready) throw std::runtime_error("no learned-order attempt on this context yet");
    HIP_OK(hipStreamSynchronize(c->st));
    if (rec) std::memcpy(rec, L.rec_host, kFeatRec * sizeof(double));
    if (pred) std::memcpy(pred, L.pred_host, 64 * 64 * sizeof(float));
  });
}

int hos_learned_reset(hos_ctx* c) {
  return guard([&] {
    if (!c) throw std::runtime_error("null context");
    hos_ctx::Learn& L = c->learn;
    L.m = 0;
    L.floor = 0;
    L.since = 0;
    L.phase = 0;
    L.full_run = 0;
    L.kappa = 1.0;
    L.stats = hos_learned_stats{};
  });
}

int hos_advance_learned(int n, hos_ctx* const* ctx, double* t_cur, double t_next, double* h_rk, double toler,
                        const double* dt_stab, const double* delt_out, const hos_mlp* net,
                        const hos_order_policy* policy, int* n_acc, int* n_rej, hos_learned_stats* stats) {
  return guard([&] {
    if (n < 0 || !policy) throw std::runtime_error("bad arguments");
    const hos_order_policy pol = *policy;
    if (pol.judge_every < 1 || !(pol.rate_tol > 0.0) || !(pol.safety > 0.0))
      throw std::runtime_error("policy: judge_every >= 1, rate_tol > 0, safety > 0");
    std::vector<LearnMember> mb(n);
    for (int i = 0; i < n; ++i) {
      if (!ctx[i]) throw std::runtime_error("null context in the ensemble");
      for (int j = 0; j < i; ++j)
        if (ctx[j] == ctx[i]) throw std::runtime_error("a context appears twice in the ensemble");
      if (pol.m_min < 1 || pol.m_min > ctx[i]->p.mHOS) throw std::runtime_error("policy: m_min outside 1..mHOS");
      mb[i] = LearnMember{ctx[i], Advance{t_cur[i], h_rk[i], 0, 0}, dt_stab[i], delt_out[i]};
    }
    // set-up (engines of every order, checkpoint buffers) before any thread starts
    for (int i = 0; i < n; ++i) {
      hos_ctx* c = ctx[i];
      const bool first = !c->learn.ready;
      learn_prepare(c, pol);
      if (first) c->learn.floor = pol.m_min;
      learn_checkpoint(c, t_cur[i], h_rk[i]);  // no reversion reaches back beyond this call
      c->learn.since = 0;
      HIP_OK(hipStreamSynchronize(c->st));
    }
    int nth = 4;
    if (const char* e = std::getenv("HOSNWT_ENSEMBLE_THREADS")) nth = std::max(1, std::atoi(e));
    nth = std::max(1, std::min(nth, n));
    std::vector<int> order(n);
    std::vector<double> cost(n);
    for (int i = 0; i < n; ++i) {
      order[i] = i;
      cost[i] = (double)ctx[i]->N1b * ctx[i]->N2b * ctx[i]->p.mHOS * std::max(0.0, t_next - t_cur[i]) / dt_stab[i];
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    std::vector<std::vector<int>> part(nth);
    std::vector<double> load(nth, 0.0);
    for (int i : order) {
      const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      part[k].push_back(i);
      load[k] += cost[i];
    }
    std::mutex em;
    std::string first_error;
    std::atomic<bool> failed{false};
    auto run = [&](const std::vector<int>& mine) {
      try {
        if (!mine.empty()) HIP_OK(hipSetDevice(ctx[mine[0]]->device));
        std::vector<char> busy(n, 0);
        for (int i : mine)
          if (mb[i].a.t < t_next) {
            learn_launch(mb[i], t_next, net, pol);
            busy[i] = 1;
          }
        for (bool any = true; any && !failed.load();) {
          any = false;
          for (int i : mine) {
            if (!busy[i]) continue;
            busy[i] = learn_finish(mb[i], t_next, toler, net, pol) ? 1 : 0;
            any = any || busy[i];
          }
        }
        for (int i : mine) {
          ctx[i]->feat_mode = 0;
          HIP_OK(hipStreamSynchronize(ctx[i]->st));
        }
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(em);
        if (first_error.empty()) first_error = e.what();
        failed = true;
      }
    };
    if (nth == 1) {
      run(part[0]);
    } else {
      std::vector<std::thread> th;
      for (int k = 0; k < nth; ++k) th.emplace_back(run, std::cref(part[k]));
      for (auto& t : th) t.join();
    }
    for (int i = 0; i < n; ++i) {
      t_cur[i] = mb[i].a.t;
      h_rk[i] = mb[i].a.h;
      if (n_acc) n_acc[i] = mb[i].a.na;
      if (n_rej) n_rej[i] = mb[i].a.nr;
      if (stats) stats[i] = ctx[i]->learn.stats;
    }
    if (!first_error.empty()) throw std::runtime_error(first_error);
  });
}

int hos_rhs(hos_ctx* c, double t, double delta_t, const double* G1, const double* G2, const double* a1st,
            const double* eta1st, const double* a_add_1, double* dG1, double* dG2, double* da1st, double* deta1st,
            double* da_add_1, double* da_add_ramp) {
  return guard([&] {
    const size_t ns = (size_t)c->n1 * c->n2, na = (size_t)c->n3a * c->n2;
    hipStream_t st = c->st;
    HIP_OK(hipMemcpyAsync(c->G1m, G1, ns * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(c->G2m, G2, ns * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(c->a1m, a1st, ns * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(c->e1m, eta1st, ns * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(c->aadm, a_add_1, na * sizeof(double), hipMemcpyHostToDevice, st));
    if (c->wmk2) {  // the higher-order state taken as zero
      HIP_OK(hipMemsetAsync(c->a2m, 0, ns * sizeof(double), st));
      HIP_OK(hipMemsetAsync(c->e2m, 0, ns * sizeof(double), st));
      HIP_OK(hipMemsetAsync(c->aad2m, 0, na * sizeof(double), st));
      if (c->wmk3) HIP_OK(hipMemsetAsync(c->aad3m, 0, na * sizeof(double), st));
      rhs_eval(c, t, delta_t, c->G1m, c->G2m, c->a1m, c->e1m, c->aadm, 0, c->a2m, c->e2m, c->aad2m, c->aad3m);
    } else {
      rhs_eval(c, t, delta_t, c->G1m, c->G2m, c->a1m, c->e1m, c->aadm, 0);
    }
    HIP_OK(hipMemcpyAsync(dG1, c->kG1[0], ns * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(dG2, c->kG2[0], ns * sizeof(double), hipMemcpyDeviceToHost, st));
    if (da1st) HIP_OK(hipMemcpyAsync(da1st, c->ka1[0], ns * sizeof(double), hipMemcpyDeviceToHost, st));
    if (deta1st) HIP_OK(hipMemcpyAsync(deta1st, c->ke1[0], ns * sizeof(double), hipMemcpyDeviceToHost, st));
    if (da_add_1) HIP_OK(hipMemcpyAsync(da_add_1, c->kad1[0], na * sizeof(double), hipMemcpyDeviceToHost, st));
    if (da_add_ramp) HIP_OK(hipMemcpyAsync(da_add_ramp, c->kadr[0], na * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
  });
}

int hos_rhs2(hos_ctx* c, double t, double delta_t, const double* G1, const double* G2, const double* a1st,
             const double* eta1st, const double* a_add_1, const double* a2nd, const double* eta2nd,
             const double* a_add_2, double* dG1, double* dG2, double* da1st, double* deta1st, double* da_add_1,
             double* da_add_ramp, double* da2nd, double* deta2nd, double* da_add_2) {
  return guard([&] {
    if (!c->wmk2) throw std::runtime_error("hos_rhs2 needs the second-order wavemaker (i_wmk >= 2)");
    rhs_higher(c, t, delta_t, G1, G2, a1st, eta1st, a_add_1, a2nd, eta2nd, a_add_2, nullptr, dG1, dG2, da1st,
               deta1st, da_add_1, da_add_ramp, da2nd, deta2nd, da_add_2, nullptr);
  });
}

int hos_rhs3(hos_ctx* c, double t, double delta_t, const double* G1, const double* G2, const double* a1st,
             const double* eta1st, const double* a_add_1, const double* a2nd, const double* eta2nd,
             const double* a_add_2, const double* a_add_3, double* dG1, double* dG2, double* da1st,
             double* deta1st, double* da_add_1, double* da_add_ramp, double* da2nd, double* deta2nd,
             double* da_add_2, double* da_add_3) {
  return guard([&] {
    if (!c->wmk3) throw std::runtime_error("hos_rhs3 needs the third-order wavemaker (i_wmk = 3)");
    rhs_higher(c, t, delta_t, G1, G2, a1st, eta1st, a_add_1, a2nd, eta2nd, a_add_2, a_add_3, dG1, dG2, da1st,
               deta1st, da_add_1, da_add_ramp, da2nd, deta2nd, da_add_2, da_add_3);
  });
}

int hos_set_state3(hos_ctx* c, const double* a3rd_add) {
  return guard([&] {
    if (!c->wmk3) throw std::runtime_error("no third-order wavemaker state (i_wmk = 3 only)");
    const size_t na = (size_t)c->n3a * c->n2;
    HIP_OK(hipMemcpyAsync(c->a3rd_add, a3rd_add, na * sizeof(double), hipMemcpyHostToDevice, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
    c->have_trial = false;
  });
}

int hos_get_state3(hos_ctx* c, double* a3rd_add) {
  return guard([&] {
    if (!c->wmk3) throw std::runtime_error("no third-order wavemaker state (i_wmk = 3 only)");
    const size_t na = (size_t)c->n3a * c->n2;
    HIP_OK(hipStreamSynchronize(c->st));
    HIP_OK(hipMemcpy(a3rd_add, c->a3rd_add, na * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int hos_get_wmk(hos_ctx* c, double* alpha, double* pos_top) {
  return guard([&] {
    *alpha = c->last_alpha;
    *pos_top = c->last_pos_top;
  });
}

int hos_get_wmk_motion(hos_ctx* c, double* alpha, double* alphat, double* x_top, double* v_top) {
  return guard([&] {
    *alpha = c->last_alpha;
    *alphat = c->last_alphat;
    const int n2 = c->n2;
    if (!c->wmk) {
      for (int i = 0; i < n2; ++i) x_top[i] = v_top[i] = 0.0;
      return;
    }
    std::vector<double> X(n2), Xt(n2);
    c->eng->paddle_top(X.data(), Xt.data());
    for (int i = 0; i < n2; ++i) {
      x_top[i] = c->last_alpha * X[i];
      v_top[i] = c->last_alphat * X[i] + c->last_alpha * Xt[i];
    }
  });
}

int hos_swense_modes(hos_ctx* c, double t, int meta2, double* specx, double* specy, double* specz, double* spect,
                     double* modes_fs, double* modes_fst, double* modes_add, double* modes_addt) {
  return guard([&] {
    if (!c->wmk) throw std::runtime_error("the SWENSE output is built for wavemaker tanks (igeom 1 or 2)");
    if (filter_active(c->p.coeffilt[0]) || (c->n2 > 1 && filter_active(c->p.coeffilt[1])) ||
        filter_active(c->p.coeffilt[2]))
      throw std::runtime_error("the SWENSE output is built without filters (coeffilt = 1)");
    if (meta2 < 1) throw std::runtime_error("meta2 must be >= 1");
    const size_t ns = (size_t)c->n1 * c->n2, na = (size_t)c->n3a * c->n2;
    hipStream_t st = c->st;
    // init_velocity_comp's right-hand side at the committed state and time t
    // (velocities.f90:571-577), into the stage-0 buffers
    hipLaunchKernelGGL(k_step_begin, dim3(nblk(ns)), dim3(256), 0, st, c->n1, c->n2, c->pix, c->piy,
                       (const double*)c->eta_hat, (const double*)c->phis_hat, c->G1, c->G2);
    rhs_eval(c, t, 0.0, c->G1, c->G2, c->a1st, c->eta1st, c->a1st_add, 0, c->a2nd, c->eta2nd, c->a2nd_add,
             c->a3rd_add);
    double* outs[5];
    c->eng->swense(meta2, outs);
    HIP_OK(hipStreamSynchronize(st));
    double* host[5] = {specx, specy, specz, spect, modes_fst};
    for (int f = 0; f < 5; ++f) HIP_OK(hipMemcpy(host[f], outs[f], ns * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(modes_fs, c->eta_hat, ns * sizeof(double), hipMemcpyDeviceToHost));
    // modesadd, modesaddt: the paddle modes of every order and their derivatives at the
    // committed state (velocities.f90:623-628)
    std::vector<double> h(na);
    const double* adds[3] = {c->a1st_add, c->a2nd_add, c->a3rd_add};
    const double* dadds[3] = {c->kad1[0], c->wmk2 ? c->kad2[0] : nullptr, c->wmk3 ? c->kad3[0] : nullptr};
    for (size_t i = 0; i < na; ++i) modes_add[i] = modes_addt[i] = 0.0;
    for (int q = 0; q < 3; ++q) {
      if (adds[q]) {
        HIP_OK(hipMemcpy(h.data(), adds[q], na * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < na; ++i) modes_add[i] += h[i];
      }
      if (dadds[q]) {
        HIP_OK(hipMemcpy(h.data(), dadds[q], na * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < na; ++i) modes_addt[i] += h[i];
      }
    }
  });
}

int hos_get_modes(hos_ctx* c, double* eta_modes, double* phis_modes) {
  return guard([&] {
    const size_t ns = (size_t)c->n1 * c->n2;
    HIP_OK(hipStreamSynchronize(c->st));
    if (eta_modes) HIP_OK(hipMemcpy(eta_modes, c->eta_hat, ns * sizeof(double), hipMemcpyDeviceToHost));
    if (phis_modes) HIP_OK(hipMemcpy(phis_modes, c->phis_hat, ns * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int hos_probes(hos_ctx* c, int np, const double* xp, const double* yp, double* eta_p) {
  return guard([&] {
    if (np < 0) throw std::runtime_error("negative probe count");
    if (np == 0) return;
    const size_t lds = (size_t)(c->n1 + c->n2) * sizeof(double);
    if (lds > 160 * 1024) throw std::runtime_error("grid too large for the probe kernel");
    std::vector<double> y0(np, 0.0);
    double *d = nullptr;
    HIP_OK(hipMalloc(&d, 3 * (size_t)np * sizeof(double)));
    struct Free {
      double* p;
      ~Free() { (void)hipFree(p); }
    } fr{d};
    HIP_OK(hipMemcpy(d, xp, np * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d + np, (c->n2 > 1 && yp) ? yp : y0.data(), np * sizeof(double), hipMemcpyHostToDevice));
    HIP_OK(hipStreamSynchronize(c->st));
    if (lds > 64 * 1024)
      HIP_OK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_probes),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_probes, dim3(np), dim3(256), lds, c->st, c->n1, c->n2, c->pix, c->piy, (const double*)d,
                       (const double*)(d + np), (const double*)c->eta_hat, d + 2 * np);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(eta_p, d + 2 * np, np * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HIP_OK(hipStreamSynchronize(c->st));
  });
}

int hos_get_slope_max(hos_ctx* c, double* sx, double* sy) {
  return guard([&] { c->eng->slope_max(*sx, *sy); });
}

int hos_check_slope(hos_ctx* c, double threshold, double coeff, int* fx, int* fy) {
  return guard([&] {
    double sx = 0.0, sy = 0.0;
    c->eng->slope_max(sx, sy);
    *fx = sx > threshold ? 1 : 0;
    *fy = (sy > threshold && c->n2 != 1) ? 1 : 0;
    if (!*fx && !*fy) return;
    // filtering_x then filtering_y of the committed eta and phis, i.e. their
    // modes times the filter weights
    const size_t ns = (size_t)c->n1 * c->n2;
    const std::vector<double> w = filter_weights_2d(c->n1, c->n2, *fx ? coeff : 1.0, *fy ? coeff : 1.0);
    HIP_OK(hipMemcpyAsync(c->phys_tmp, w.data(), ns * sizeof(double), hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL(k_mulw, dim3(nblk(ns)), dim3(256), 0, c->st, ns, (const double*)c->phys_tmp, c->eta_hat);
    hipLaunchKernelGGL(k_mulw, dim3(nblk(ns)), dim3(256), 0, c->st, ns, (const double*)c->phys_tmp, c->phis_hat);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->st));
  });
}

int hos_time_big_transform(hos_ctx* c, int reps, double* out4) {
  return guard([&] {
    if (reps < 1) throw std::runtime_error("reps must be >= 1");
    c->eng->time_big(reps, 1, out4);
  });
}

int hos_time_big_batch(hos_ctx* c, int reps, int items, double* out4) {
  return guard([&] {
    if (reps < 1) throw std::runtime_error("reps must be >= 1");
    c->eng->time_big(reps, items, out4);
  });
}

int hos_pass_log_begin(hos_ctx* c, int max_launches) {
  return guard([&] {
    if (max_launches < 1 || max_launches > (1 << 16)) throw std::runtime_error("pass log of 1 to 65536 launches");
    if (t_pass_log) throw std::runtime_error("a pass log is already recording on this host thread");
    auto lg = std::make_unique<PassLog>();
    lg->N[0] = c->N1b;
    lg->N[1] = c->N2b > 1 ? c->N2b : 0;
    lg->cap = (size_t)max_launches;
    lg->ev.resize(2 * lg->cap, nullptr);
    for (auto& e : lg->ev) HIP_OK(hipEventCreate(&e));
    lg->bytes.reserve(lg->cap);
    lg->pass.reserve(lg->cap);
    lg->items.reserve(lg->cap);
    c->plog = std::move(lg);
    t_pass_log = c->plog.get();
  });
}

int hos_pass_log_end(hos_ctx* c, double* out9) {
  return guard([&] {
    if (!c->plog) throw std::runtime_error("no pass log is recording");
    if (t_pass_log != c->plog.get()) throw std::runtime_error("pass log ended on another host thread");
    t_pass_log = nullptr;
    std::unique_ptr<PassLog> lg = std::move(c->plog);
    for (int k = 0; k < 9; ++k) out9[k] = 0.0;
    for (size_t i = 0; i < lg->bytes.size(); ++i) {
      HIP_OK(hipEventSynchronize(lg->ev[2 * i + 1]));
      float ms = 0.f;
      HIP_OK(hipEventElapsedTime(&ms, lg->ev[2 * i], lg->ev[2 * i + 1]));
      double* o = out9 + 4 * lg->pass[i];
      o[0] += 1.0;
      o[1] += lg->bytes[i];
      o[2] += ms;
      o[3] += lg->items[i];
    }
    out9[8] = lg->full ? 1.0 : 0.0;
  });
}

int hos_get_stats(hos_ctx* c, hos_stats* s) {
  return guard([&] { *s = c->stats; });
}

int hos_set_order(hos_ctx* c, int m) {
  return guard([&] {
    if (!c) throw std::runtime_error("null context");
    HIP_OK(hipStreamSynchronize(c->st));
    select_order(c, m);
    HIP_OK(hipDeviceSynchronize());  // a new engine's set-up fills run on the null stream
  });
}

int hos_get_order(hos_ctx* c, int* m) {
  return guard([&] {
    if (!c || !m) throw std::runtime_error("null argument");
    *m = c->M;
  });
}

int hos_record_features(hos_ctx* c, int mode) {
  return guard([&] {
    if (!c) throw std::runtime_error("null context");
    if (mode < 0 || mode > 2) throw std::runtime_error("feature mode must be 0, 1 or 2");
    HIP_OK(hipStreamSynchronize(c->st));
    c->feat_mode = mode;
    c->feat_n = 0;
  });
}

int hos_take_features(hos_ctx* c, int max_rec, double* out, int* n_out) {
  return guard([&] {
    if (!c || !n_out) throw std::runtime_error("null argument");
    HIP_OK(hipStreamSynchronize(c->st));
    const int avail = (c->feat_mode == 2) ? std::min(c->feat_n, c->feat_cap) : 1;
    const int n = std::max(0, std::min(avail, max_rec));
    // oldest first: ring slots (feat_n - avail .. feat_n - 1) mod cap
    const int first = (c->feat_mode == 2) ? c->feat_n - avail : 0;
    std::vector<double> ring((size_t)c->feat_cap * kFeatRec);
    HIP_OK(hipMemcpy(ring.data(), c->feat, ring.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int r = 0; r < n; ++r) {
      const int slot = (c->feat_mode == 2) ? (first + r) % c->feat_cap : 0;
      std::memcpy(out + (size_t)r * kFeatRec, ring.data() + (size_t)slot * kFeatRec, kFeatRec * sizeof(double));
    }
    if (c->feat_mode == 2) c->feat_n = 0;
    *n_out = n;
  });
}

int hos_synchronize(hos_ctx* c) {
  return guard([&] { HIP_OK(hipStreamSynchronize(c->st)); });
}

static int standalone(int n1, int n2, int kx, int ky, int prec, int dir, const double* in, double* out) {
  return guard([&] {
    if (n1 < 2 || n2 < 1) throw std::runtime_error("bad grid");
    const size_t n = (size_t)n1 * n2;
    hipStream_t st;
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    Arena ar;
    double* din = ar.get<double>(n);
    double* dout = ar.get<double>(n);
    HIP_OK(hipMemcpy(din, in, n * sizeof(double), hipMemcpyHostToDevice));
    if (prec == 0) {
      Grid2<double> g;
      g.build(n1, n2);
      g.check_lengths(true);
      double* tmp = ar.get<double>(n);
      double* gws = g.gline_ws_elems() ? ar.get<double>(g.gline_ws_elems()) : nullptr;
      HIP_OK(hipDeviceSynchronize());  // the arena's zero fills run on the null stream
      r2r2d<double>(g, kx, ky, dir, src_plain<double, double>(din, n1, n1, n2), n1, n2, Dst<double>{dout, n1, 0},
                    tmp, st, gws);
      HIP_OK(hipStreamSynchronize(st));
      g.release();
    } else {
      Grid2<float> g;
      g.build(n1, n2);
      g.check_lengths(true);
      float* tmp = ar.get<float>(n);
      float* fo = ar.get<float>(n);
      float* gws = g.gline_ws_elems() ? ar.get<float>(g.gline_ws_elems()) : nullptr;
      HIP_OK(hipDeviceSynchronize());  // the arena's zero fills run on the null stream
      r2r2d<float>(g, kx, ky, dir, src_plain<float, double>(din, n1, n1, n2), n1, n2, Dst<float>{fo, n1, 0}, tmp,
                   st, gws);
      hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk(n)), dim3(256), 0, st, n, (const float*)fo, dout);
      HIP_OK(hipStreamSynchronize(st));
      g.release();
    }
    HIP_OK(hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost));
    (void)hipStreamDestroy(st);
  });
}

int hos_space_2_fourier(int n1, int n2, int kind_x, int kind_y, int precision, const double* in, double* out) {
  return standalone(n1, n2, kind_x, kind_y, precision, FWD, in, out);
}

int hos_fourier_2_space(int n1, int n2, int kind_x, int kind_y, int precision, const double* in, double* out) {
  return standalone(n1, n2, kind_x, kind_y, precision, INV, in, out);
}

int hos_get_schedule(hos_ctx* c, int* s8, double* us2) {
  return guard([&] {
    if (!c || !s8 || !us2) throw std::runtime_error("null argument");
    c->eng->schedule(s8, us2);
  });
}

int hos_dense_product(int precision, int a_kmajor, int b_kmajor, int b_shared, int m, int n, int k, int batch, int tile,
                      const double* A, const double* B, double* C) {
  return guard([&] {
    if (m < 1 || n < 1 || k < 1 || batch < 1 || !A || !B || !C) throw std::runtime_error("bad dense product shape");
    if (precision != 0 && precision != 1) throw std::runtime_error("precision must be 0 or 1");
    const size_t na = (size_t)m * k, nb = (size_t)k * n, nc = (size_t)m * n;
    const size_t tb = b_shared ? nb : nb * batch;
    auto run = [&](auto zero) {
      using U = decltype(zero);
      std::vector<U> hA(na * batch), hB(tb), hC(nc * batch);
      for (size_t i = 0; i < hA.size(); ++i) hA[i] = (U)A[i];
      for (size_t i = 0; i < hB.size(); ++i) hB[i] = (U)B[i];
      U *dA = nullptr, *dB = nullptr, *dC = nullptr;
      HIP_OK(hipMalloc(&dA, hA.size() * sizeof(U)));
      HIP_OK(hipMalloc(&dB, hB.size() * sizeof(U)));
      HIP_OK(hipMalloc(&dC, hC.size() * sizeof(U)));
      std::unique_ptr<void, void (*)(void*)> fa(dA, [](void* q) { (void)hipFree(q); });
      std::unique_ptr<void, void (*)(void*)> fb(dB, [](void* q) { (void)hipFree(q); });
      std::unique_ptr<void, void (*)(void*)> fc(dC, [](void* q) { (void)hipFree(q); });
      HIP_OK(hipMemcpy(dA, hA.data(), hA.size() * sizeof(U), hipMemcpyHostToDevice));
      HIP_OK(hipMemcpy(dB, hB.data(), hB.size() * sizeof(U), hipMemcpyHostToDevice));
      hosmm::mm<U>(nullptr, a_kmajor != 0, b_kmajor != 0, m, n, k, dA, a_kmajor ? k : m, (long)na, dB, b_kmajor ? k : n,
                   b_shared ? 0L : (long)nb, dC,
                   m, (long)nc, batch, tile);
      HIP_OK(hipDeviceSynchronize());
      HIP_OK(hipMemcpy(hC.data(), dC, hC.size() * sizeof(U), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < hC.size(); ++i) C[i] = (double)hC[i];
    };
    if (precision == 0)
      run(0.0);
    else
      run(0.0f);
  });
}

}  // extern "C"
