// Dense products of the solver on the matrix cores (FP64 and FP32 MFMA).
//
// Every dense product the right-hand side needs has the form
//     C_b (m x n) = A_b (m x k) B_b (k x n),   b < batch,
// with C column-major and no accumulation into C.  These are the base-grid
// cos/sin transforms of non-smooth lengths (the matrix forms of
// fourier_FFTW.f90:552-653), the paddle-grid transforms (fourier_FFTW.f90:432-492)
// and the sine-to-cosine map of the mixed-parity products.  Each operand is
// stored either k-contiguous ("k-major": A[kk + lda i], B[kk + ldb j]) or not
// (A[i + lda kk], B[j + ldb kk]).  The solver stores its constant matrices
// k-major; the fields are k-major as the right operand of an x product and
// row-contiguous as the left operand of a y product.
//
// Layout of one launch: a workgroup of four waves owns a (16 TM) x (16 TN)
// tile of C; the waves split K into chunks of 16 (wave w takes chunks w, w+4,
// ...), each wave accumulating the whole tile with 16x16x4 MFMAs, and the four
// partial tiles are summed in LDS in a fixed order (results do not depend on
// timing).  Inside a chunk the k index of lane group g = lane / 16 at MFMA step
// s < 4 is 16 c + 4 g + s, the same permutation for both operands, so a lane
// of a k-major operand reads four consecutive values, one vector load when the
// host has checked alignment (VEC).
// MFMA operand layout (16x16x4, one block): lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; the accumulator's register r of lane l is C[row][l % 16]
// with row = 4 (l / 16) + r for the FP32 instruction and l / 16 + 4 r for the
// FP64 one.
//
// The loop issues the loads of chunk c + 4 before the MFMAs of chunk c into a
// second register set (two sets used in turn, no copies between them).  Loads
// are unconditional: rows / columns past the edge are clamped (their outputs
// are never stored) and k past the end reads the clamped last k, masked to zero
// on A where the values are used.  Every iteration handles two chunks, so the
// loop body is one basic block in which scheduling barriers keep the compiler
// from sinking the prefetch below the MFMAs.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>

namespace hosmm {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

constexpr int kS = 4;   // MFMA steps per chunk (k values per lane per chunk)
constexpr int kKW = 4;  // waves per workgroup (K split)

template <typename T>
struct Acc;
template <>
struct Acc<double> {
  using V = v4d;
  __device__ static V mfma(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int g, int r) { return g + 4 * r; }
  // four consecutive values from a 16-byte aligned address
  __device__ static void load4(const double* p, double (&o)[4]) {
    const v2d x = *reinterpret_cast<const v2d*>(p), y = *reinterpret_cast<const v2d*>(p + 2);
    o[0] = x[0];
    o[1] = x[1];
    o[2] = y[0];
    o[3] = y[1];
  }
};
template <>
struct Acc<float> {
  using V = v4f;
  __device__ static V mfma(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int g, int r) { return 4 * g + r; }
  __device__ static void load4(const float* p, float (&o)[4]) {
    const v4f x = *reinterpret_cast<const v4f*>(p);
    o[0] = x[0];
    o[1] = x[1];
    o[2] = x[2];
    o[3] = x[3];
  }
};

template <typename T, bool AK, bool BK, int TM, int TN, bool VEC>
__global__ void __launch_bounds__(64 * kKW) k_mm(int m, int n, int k, const T* __restrict__ A, int lda, long sa,
                                                 const T* __restrict__ B, int ldb, long sb, T* __restrict__ C, int ldc,
                                                 long sc, int rsc) {
  using V = typename Acc<T>::V;
  constexpr int RM = 16 * TM, RN = 16 * TN;
  __shared__ T red[kKW][RN][RM + 1];
  A += (size_t)blockIdx.z * sa;
  B += (size_t)blockIdx.z * sb;
  C += (size_t)blockIdx.z * sc;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r16 = lane & 15, g = lane >> 4;
  const int row0 = blockIdx.x * RM, col0 = blockIdx.y * RN;
  // this lane's operand row of A / column of B, clamped to a valid index
  const T* pa[TM];
  const T* pb[TN];
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const size_t r = (size_t)min(row0 + 16 * i + r16, m - 1);
    pa[i] = AK ? A + (size_t)lda * r : A + r;
  }
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const size_t cc = (size_t)min(col0 + 16 * j + r16, n - 1);
    pb[j] = BK ? B + (size_t)ldb * cc : B + cc;
  }
  V acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = V{};
  const int nc = (k + 4 * kS - 1) / (4 * kS);
  T a0[TM][kS], b0[TN][kS], a1[TM][kS], b1[TN][kS];
  auto load = [&](T (&ab)[TM][kS], T (&bb)[TN][kS], int c) {
    const int kb = kS * (4 * c + g);
    if constexpr (VEC) {  // k % 4 == 0: a vector is wholly inside or wholly past k
      const int kv = min(kb, k - kS);
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        if constexpr (AK) {
          Acc<T>::load4(pa[i] + kv, ab[i]);
        } else {
#pragma unroll
          for (int s = 0; s < kS; ++s) ab[i][s] = pa[i][(size_t)lda * min(kb + s, k - 1)];
        }
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        if constexpr (BK) {
          Acc<T>::load4(pb[j] + kv, bb[j]);
        } else {
#pragma unroll
          for (int s = 0; s < kS; ++s) bb[j][s] = pb[j][(size_t)ldb * min(kb + s, k - 1)];
        }
      }
    } else {
#pragma unroll
      for (int s = 0; s < kS; ++s) {
        const int kc = min(kb + s, k - 1);
#pragma unroll
        for (int i = 0; i < TM; ++i) ab[i][s] = AK ? pa[i][kc] : pa[i][(size_t)lda * kc];
#pragma unroll
        for (int j = 0; j < TN; ++j) bb[j][s] = BK ? pb[j][kc] : pb[j][(size_t)ldb * kc];
      }
    }
  };
  auto compute = [&](const T (&ab)[TM][kS], const T (&bb)[TN][kS], int c) {
#pragma unroll
    for (int s = 0; s < kS; ++s) {
      const T mk = (kS * (4 * c + g) + s < k) ? T(1) : T(0);  // the K tail
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const T av = ab[i][s] * mk;
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = Acc<T>::mfma(av, bb[j][s], acc[i][j]);
      }
    }
  };
  // chunks of this wave; a wave with an odd number computes one masked chunk more
  const int mine = w < nc ? (nc - w + kKW - 1) / kKW : 0;
  if (mine > 0) {
    load(a0, b0, w);
    for (int p = 0, c = w; p < (mine + 1) / 2; ++p, c += 2 * kKW) {
      load(a1, b1, c + kKW);
      __builtin_amdgcn_sched_barrier(0);
      compute(a0, b0, c);
      __builtin_amdgcn_sched_barrier(0);
      load(a0, b0, c + 2 * kKW);
      __builtin_amdgcn_sched_barrier(0);
      compute(a1, b1, c + kKW);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[w][16 * j + r16][16 * i + Acc<T>::row(g, r)] = acc[i][j][r];
  __syncthreads();
  for (int o = threadIdx.x; o < RM * RN; o += 64 * kKW) {
    const int rr = o % RM, cc = o / RM;
    T s = red[0][cc][rr];
#pragma unroll
    for (int q = 1; q < kKW; ++q) s += red[q][cc][rr];
    if (row0 + rr < m && col0 + cc < n) C[(size_t)(row0 + rr) * rsc + (size_t)ldc * (col0 + cc)] = s;
  }
}

// Tile of one launch: 32 x 32 when that still gives most CUs a workgroup,
// else 16 x 16 (small paddle-grid products).
inline int default_tile(int m, int n, int batch) {
  const long t32 = (long)((m + 31) / 32) * ((n + 31) / 32) * batch;
  return t32 >= 192 ? 32 : 16;
}

template <typename T, bool AK, bool BK, int TM, int TN>
void launch(hipStream_t st, bool vec, dim3 grid, int m, int n, int k, const T* A, int lda, long sa, const T* B, int ldb,
            long sb, T* C, int ldc, long sc, int rsc) {
  if (vec)
    hipLaunchKernelGGL((k_mm<T, AK, BK, TM, TN, true>), grid, dim3(64 * kKW), 0, st, m, n, k, A, lda, sa, B, ldb, sb,
                       C, ldc, sc, rsc);
  else
    hipLaunchKernelGGL((k_mm<T, AK, BK, TM, TN, false>), grid, dim3(64 * kKW), 0, st, m, n, k, A, lda, sa, B, ldb, sb,
                       C, ldc, sc, rsc);
}

template <typename T, bool AK, bool BK>
void launch_tile(hipStream_t st, int tile, bool vec, int m, int n, int k, int batch, const T* A, int lda, long sa,
                 const T* B, int ldb, long sb, T* C, int ldc, long sc, int rsc) {
  if (tile == 32)
    launch<T, AK, BK, 2, 2>(st, vec, dim3((m + 31) / 32, (n + 31) / 32, batch), m, n, k, A, lda, sa, B, ldb, sb, C,
                            ldc, sc, rsc);
  else
    launch<T, AK, BK, 1, 1>(st, vec, dim3((m + 15) / 16, (n + 15) / 16, batch), m, n, k, A, lda, sa, B, ldb, sb, C,
                            ldc, sc, rsc);
}

// C_b = A_b B_b for b < batch on stream st.  ak / bk: the operand is k-major
// (A[kk + lda i], B[kk + ldb j]); else A[i + lda kk], B[j + ldb kk].  Strides
// sa, sb, sc between the batch's operands (0: shared).  tile 0: the default.
// Row i of C_b lies at C_b + i * rsc (rsc = 2: every other row).
template <typename T>
void mm(hipStream_t st, bool ak, bool bk, int m, int n, int k, const T* A, int lda, long sa, const T* B, int ldb,
        long sb, T* C, int ldc, long sc, int batch = 1, int tile = 0, int rsc = 1) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  if (k < 1 || rsc < 1 || ldc < (long)(m - 1) * rsc + 1 || lda < (ak ? k : m) || ldb < (bk ? k : n))
    throw std::runtime_error("dense product: inconsistent operand shapes");
  if (batch > 65535) throw std::runtime_error("dense product: batch too large");
  if (tile == 0) tile = default_tile(m, n, batch);
  if (tile != 16 && tile != 32) throw std::runtime_error("dense product: tile must be 16 or 32");
  // vector loads of the k-major operands: four values, 16-byte aligned
  auto aligned = [](const void* p, long ld, long s) {
    return ((uintptr_t)p % 16 == 0) && ((ld * (long)sizeof(T)) % 16 == 0) && ((s * (long)sizeof(T)) % 16 == 0);
  };
  const bool vec = (k % kS == 0) && (!ak || aligned(A, lda, sa)) && (!bk || aligned(B, ldb, sb)) && (ak || bk);
  if (ak && bk)
    launch_tile<T, true, true>(st, tile, vec, m, n, k, batch, A, lda, sa, B, ldb, sb, C, ldc, sc, rsc);
  else if (!ak && bk)
    launch_tile<T, false, true>(st, tile, vec, m, n, k, batch, A, lda, sa, B, ldb, sb, C, ldc, sc, rsc);
  else if (ak && !bk)
    launch_tile<T, true, false>(st, tile, vec, m, n, k, batch, A, lda, sa, B, ldb, sb, C, ldc, sc, rsc);
  else
    launch_tile<T, false, false>(st, tile, false, m, n, k, batch, A, lda, sa, B, ldb, sb, C, ldc, sc, rsc);
  if (hipGetLastError() != hipSuccess) throw std::runtime_error("dense product: launch failed");
}

}  // namespace hosmm
