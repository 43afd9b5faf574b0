// Order-prediction MLP of the learned truncation order: 512 -> 256 -> 256 -> 256
// -> 64, ReLU hidden layers, trained and run on the GPU.
//
// Every product is one hand-written kernel on the gfx950 matrix cores:
// v_mfma_f32_32x32x16_bf16, BF16 operands, FP32 accumulation.  A workgroup of
// four waves computes a 64 x 64 tile of C = A Bt^T, each wave a 32 x 32 block;
// both operands are stored K-contiguous (A[M][K], Bt[N][K]) so a lane reads its
// eight K values as one 16-byte load straight from L2 (the matrices are a few
// hundred KB: no LDS staging).  The epilogues are fused: bias + ReLU + BF16
// store of the activation and of its transpose (forward), ReLU mask + BF16 store
// (backward data), FP32 store (weight gradients, output layer).
//
// Layouts (row-major): weights W_l [out][in] in FP32 (master, Adam) with BF16
// copies Wt_l = W_l (forward operand) and Wk_l = W_l^T (backward-data operand);
// activations H_l [rows][out] and H_l^T [out][rows] (the weight gradient
// dW_l = dZ_l^T X_l reduces over rows, so both its operands are the transposes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "mlp.h"

#define MLP_OK(expr)                                                                               \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at mlp.hip:" + \
                               std::to_string(__LINE__));                                          \
  } while (0)

namespace hosnwt_err {
void set(const char* msg);  // hos_last_error's message (hosnwt.hip)
}

using namespace hosmlp;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kDims[kLayers + 1] = {kIn, kWidth, kWidth, kWidth, kOut};

enum Epi { EPI_RELU = 0, EPI_LINEAR = 1, EPI_MASK = 2, EPI_STORE = 3 };

struct EpiArgs {
  const float* bias = nullptr;  // [N]             (RELU, LINEAR)
  __bf16* out = nullptr;        // [M][ldo] bf16   (RELU, MASK)
  int ldo = 0;
  __bf16* outT = nullptr;       // [N][ldt] bf16   (RELU, MASK; may be null)
  int ldt = 0;
  float* outf = nullptr;        // [M][ldf] fp32   (LINEAR, STORE)
  int ldf = 0;
  const __bf16* mask = nullptr; // [M][ldm]: keep where mask > 0 (MASK)
  int ldm = 0;
};

// C[M][N] = sum_k A[m][k] Bt[n][k]; M, N multiples of 64, K of 32; grid (N/64, M/64).
// Lane l of a wave holds rows (or columns) l % 32 and K values 8 (l / 32) .. +7 of
// each 16-deep slice; the accumulator's value q of lane l is element
// (8 (q / 4) + 4 (l / 32) + q % 4, l % 32) of the wave's 32 x 32 block.
template <int EPI>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const __bf16* __restrict__ A, int lda,
                                              const __bf16* __restrict__ Bt, int ldb, EpiArgs e) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tm = blockIdx.y * 64 + (wave >> 1) * 32;
  const int tn = blockIdx.x * 64 + (wave & 1) * 32;
  const int r = lane & 31, kh = (lane >> 5) * 8;
  const __bf16* ap = A + (size_t)(tm + r) * lda + kh;
  const __bf16* bp = Bt + (size_t)(tn + r) * ldb + kh;
  f32x16 acc = {};
  for (int k = 0; k < K; k += 32) {
    const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(ap + k);
    const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(bp + k);
    const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(ap + k + 16);
    const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(bp + k + 16);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc, 0, 0, 0);
  }
  const int col = tn + r;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int row = tm + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
    float v = acc[q];
    if constexpr (EPI == EPI_RELU || EPI == EPI_LINEAR) v += e.bias[col];
    if constexpr (EPI == EPI_RELU) v = v > 0.f ? v : 0.f;
    if constexpr (EPI == EPI_MASK) v = ((float)e.mask[(size_t)row * e.ldm + col] > 0.f) ? v : 0.f;
    if constexpr (EPI == EPI_RELU || EPI == EPI_MASK) {
      const __bf16 b = (__bf16)v;
      e.out[(size_t)row * e.ldo + col] = b;
      if (e.outT) e.outT[(size_t)col * e.ldt + row] = b;
    } else {
      e.outf[(size_t)row * e.ldf + col] = v;
    }
  }
}

// Mean squared error over the first nvalid rows and its gradient
// dY = 2 (Y - T) / (nvalid n), zero on padding rows; the loss sum goes to *loss.
__global__ void k_mse(int rows, int nvalid, int n, const float* __restrict__ Y, const float* __restrict__ T,
                      __bf16* __restrict__ dZ, __bf16* __restrict__ dZt, float* loss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float d2 = 0.f;
  if (i < rows * n) {
    const int row = i / n, col = i - row * n;
    const float d = (row < nvalid) ? Y[i] - T[i] : 0.f;
    d2 = d * d;
    const __bf16 g = (__bf16)(2.f * d / (float)(nvalid * n));
    dZ[i] = g;
    dZt[(size_t)col * rows + row] = g;
  }
  for (int o = 32; o > 0; o >>= 1) d2 += __shfl_xor(d2, o, 64);
  if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, d2);  // reported loss only
}

// g[o] = sum_rows dZt[o][row] (bias gradient), one wave per output.
__global__ void k_colsum(int n, int rows, const __bf16* __restrict__ dZt, float* __restrict__ g) {
  const int o = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (o >= n) return;
  float s = 0.f;
  for (int r = lane; r < rows; r += 64) s += (float)dZt[(size_t)o * rows + r];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) g[o] = s;
}

__global__ void k_adam(size_t n, float* __restrict__ P, const float* __restrict__ G, float* __restrict__ Mo,
                       float* __restrict__ Ve, float lr, float b1, float b2, float eps, float bc1, float bc2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float m = b1 * Mo[i] + (1.f - b1) * g;
  const float v = b2 * Ve[i] + (1.f - b2) * g * g;
  Mo[i] = m;
  Ve[i] = v;
  P[i] -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
}

// BF16 copies of a layer's weights: Wt[o][i] = W[o][i], Wk[i][o] = W[o][i].
__global__ void k_refresh(int out, int in, const float* __restrict__ W, __bf16* __restrict__ Wt,
                          __bf16* __restrict__ Wk) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= out * in) return;
  const int o = idx / in, i = idx - o * in;
  const __bf16 b = (__bf16)W[idx];
  Wt[idx] = b;
  Wk[(size_t)i * out + o] = b;
}

// Training batch by sample index: X[r] = data[idx[r]] and its transpose, T[r] the
// target; rows beyond nvalid are zero.
__global__ void k_gather(int rows, int nvalid, const int* __restrict__ idx, const __bf16* __restrict__ dX,
                         const float* __restrict__ dY, __bf16* __restrict__ X, __bf16* __restrict__ Xt,
                         float* __restrict__ T) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < rows * kIn) {
    const int r = i / kIn, c = i - r * kIn;
    const __bf16 v = (r < nvalid) ? dX[(size_t)idx[r] * kIn + c] : (__bf16)0.f;
    X[i] = v;
    Xt[(size_t)c * rows + r] = v;
  }
  if (i < rows * kOut) {
    const int r = i / kOut, c = i - r * kOut;
    T[i] = (r < nvalid) ? dY[(size_t)idx[r] * kOut + c] : 0.f;
  }
}

__global__ void k_to_bf16(size_t n, const float* __restrict__ x, __bf16* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (__bf16)x[i];
}

inline unsigned nblk(size_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// allocation-free inference (hos_mlp_forward_noalloc): the activations of up to
// kNoAllocRows rows live in one static device array per GPU
// ---------------------------------------------------------------------------
constexpr int kNoAllocRows = 256;
constexpr size_t kScratchX = (size_t)kNoAllocRows * kIn * sizeof(__bf16);      // input rows (bf16)
constexpr size_t kScratchH = (size_t)kNoAllocRows * kWidth * sizeof(__bf16);   // one hidden layer
constexpr size_t kScratchY = (size_t)kNoAllocRows * kOut * sizeof(float);      // padded outputs
constexpr size_t kScratchBytes = kScratchX + 3 * kScratchH + kScratchY;      // 704 KiB
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of `scratch` on the current device (hipGetSymbolAddress once per device;
// in host code the array's name is not a device address).
unsigned char* scratch_ptr() {
  static std::mutex mu;
  static std::map<int, unsigned char*> per_dev;
  int dev = 0;
  MLP_OK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto it = per_dev.find(dev);
  if (it != per_dev.end()) return it->second;
  void* p = nullptr;
  MLP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  per_dev[dev] = reinterpret_cast<unsigned char*>(p);
  return per_dev[dev];
}

// x (rows x kIn fp32) -> padded bf16 rows (zeros up to prows)
__global__ void k_rows_bf16(int rows, int prows, const float* __restrict__ x, __bf16* __restrict__ xb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= prows * kIn) return;
  xb[i] = (i < rows * kIn) ? (__bf16)x[i] : (__bf16)0.f;
}

template <int EPI>
void gemm(hipStream_t st, int M, int N, int K, const __bf16* A, int lda, const __bf16* Bt, int ldb,
          const EpiArgs& e) {
  if (M % 64 || N % 64 || K % 32) throw std::runtime_error("MLP GEMM: M, N must be multiples of 64, K of 32");
  hipLaunchKernelGGL(k_gemm<EPI>, dim3(N / 64, M / 64), dim3(256), 0, st, M, N, K, A, lda, Bt, ldb, e);
}

template <typename T>
T* dmalloc(size_t n) {
  void* p = nullptr;
  MLP_OK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  MLP_OK(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
  return reinterpret_cast<T*>(p);
}

template <typename F>
int guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    hosnwt_err::set(e.what());
    return 1;
  }
}

}  // namespace

struct hos_mlp {
  size_t off_w[kLayers], off_b[kLayers], nparam = 0;
  float *P = nullptr, *G = nullptr, *Mo = nullptr, *Ve = nullptr;
  __bf16 *Wt[kLayers] = {}, *Wk[kLayers] = {};
  int max_rows = 0;
  // training activations (max_rows rows)
  __bf16 *X = nullptr, *Xt = nullptr, *H[3] = {}, *Ht[3] = {}, *dZ[kLayers] = {}, *dZt[kLayers] = {};
  float *Y = nullptr, *T = nullptr, *loss = nullptr;
  int* idx = nullptr;
  float *xs = nullptr, *ts = nullptr;  // staging of host batches
  __bf16* xb = nullptr;
  // training set on the device
  __bf16* dX = nullptr;
  float* dY = nullptr;
  int n_data = 0;
  long step = 0;  // Adam steps taken
  hipStream_t st = nullptr;
  std::vector<void*> owned;

  template <typename U>
  U* get(size_t n) {
    U* p = dmalloc<U>(n);
    owned.push_back(p);
    return p;
  }
  ~hos_mlp() {
    if (st) (void)hipStreamSynchronize(st);
    for (void* p : owned) (void)hipFree(p);
    if (dX) (void)hipFree(dX);
    if (dY) (void)hipFree(dY);
    if (st) (void)hipStreamDestroy(st);
  }
  void refresh() {
    for (int l = 0; l < kLayers; ++l) {
      const int out = kDims[l + 1], in = kDims[l];
      hipLaunchKernelGGL(k_refresh, dim3(nblk((size_t)out * in)), dim3(256), 0, st, out, in,
                         (const float*)(P + off_w[l]), Wt[l], Wk[l]);
    }
    MLP_OK(hipGetLastError());
  }
  // forward over rows (X, Xt filled): hidden activations and their transposes, Y
  void forward_train(int rows) {
    const __bf16* in = X;
    for (int l = 0; l < 3; ++l) {
      EpiArgs e;
      e.bias = P + off_b[l];
      e.out = H[l];
      e.ldo = kWidth;
      e.outT = Ht[l];
      e.ldt = rows;
      gemm<EPI_RELU>(st, rows, kWidth, kDims[l], in, kDims[l], Wt[l], kDims[l], e);
      in = H[l];
    }
    EpiArgs e;
    e.bias = P + off_b[3];
    e.outf = Y;
    e.ldf = kOut;
    gemm<EPI_LINEAR>(st, rows, kOut, kWidth, H[2], kWidth, Wt[3], kWidth, e);
  }
  // loss gradient dZ[3] set: weight and bias gradients of every layer into G
  void backward(int rows) {
    for (int l = kLayers - 1; l >= 0; --l) {
      const int out = kDims[l + 1], in = kDims[l];
      const __bf16* inT = (l == 0) ? Xt : Ht[l - 1];
      EpiArgs e;
      e.outf = G + off_w[l];
      e.ldf = in;
      gemm<EPI_STORE>(st, out, in, rows, dZt[l], rows, inT, rows, e);
      hipLaunchKernelGGL(k_colsum, dim3(nblk(out, 4)), dim3(256), 0, st, out, rows, (const __bf16*)dZt[l],
                         G + off_b[l]);
      if (l > 0) {
        EpiArgs d;
        d.mask = H[l - 1];
        d.ldm = in;
        d.out = dZ[l - 1];
        d.ldo = in;
        d.outT = dZt[l - 1];
        d.ldt = rows;
        gemm<EPI_MASK>(st, rows, in, out, dZ[l], out, Wk[l], out, d);
      }
    }
    MLP_OK(hipGetLastError());
  }
  void loss_grad(int rows, int nvalid) {
    hipLaunchKernelGGL(k_mse, dim3(nblk((size_t)rows * kOut)), dim3(256), 0, st, rows, nvalid, kOut, (const float*)Y,
                       (const float*)T, dZ[3], dZt[3], loss);
  }
  void adam(float lr) {
    ++step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - std::pow(b1, (float)step), bc2 = 1.f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(nparam)), dim3(256), 0, st, nparam, P, (const float*)G, Mo, Ve, lr, b1, b2,
                       1e-8f, bc1, bc2);
    refresh();
  }
  // host rows -> X, Xt (bf16), T through the staging buffers
  void upload_batch(int rows, int nvalid, const float* x, const float* t) {
    std::vector<float> hx((size_t)rows * kIn, 0.f), ht((size_t)rows * kOut, 0.f);
    std::memcpy(hx.data(), x, (size_t)nvalid * kIn * sizeof(float));
    if (t) std::memcpy(ht.data(), t, (size_t)nvalid * kOut * sizeof(float));
    std::vector<int> id(rows);
    for (int r = 0; r < rows; ++r) id[r] = r;
    MLP_OK(hipMemcpyAsync(xs, hx.data(), hx.size() * sizeof(float), hipMemcpyHostToDevice, st));
    MLP_OK(hipMemcpyAsync(ts, ht.data(), ht.size() * sizeof(float), hipMemcpyHostToDevice, st));
    MLP_OK(hipMemcpyAsync(idx, id.data(), rows * sizeof(int), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(hx.size())), dim3(256), 0, st, hx.size(), (const float*)xs, xb);
    hipLaunchKernelGGL(k_gather, dim3(nblk((size_t)rows * kIn)), dim3(256), 0, st, rows, rows, (const int*)idx,
                       (const __bf16*)xb, (const float*)ts, X, Xt, T);
    MLP_OK(hipStreamSynchronize(st));
  }
  void check_rows(int rows) const {
    if (rows <= 0 || rows % kRowTile || rows > max_rows)
      throw std::runtime_error("MLP rows must be a positive multiple of 64 and at most the batch capacity");
  }
};

namespace hosmlp {

void workspace_alloc(Workspace& w, int max_rows) {
  if (max_rows <= 0 || max_rows % kRowTile) throw std::runtime_error("MLP workspace rows: multiple of 64");
  w.max_rows = max_rows;
  w.x = dmalloc<__bf16>((size_t)max_rows * kIn);
  for (int l = 0; l < 3; ++l) w.h[l] = dmalloc<__bf16>((size_t)max_rows * kWidth);
  w.y = dmalloc<float>((size_t)max_rows * kOut);
}

void workspace_free(Workspace& w) {
  (void)hipFree(w.x);
  for (int l = 0; l < 3; ++l) (void)hipFree(w.h[l]);
  (void)hipFree(w.y);
  w = Workspace{};
}

void forward(const hos_mlp* net, int rows, Workspace& w, hipStream_t st) {
  if (rows <= 0 || rows % kRowTile || rows > w.max_rows) throw std::runtime_error("MLP forward: bad row count");
  const __bf16* in = w.x;
  for (int l = 0; l < 3; ++l) {
    EpiArgs e;
    e.bias = net->P + net->off_b[l];
    e.out = w.h[l];
    e.ldo = kWidth;
    gemm<EPI_RELU>(st, rows, kWidth, kDims[l], in, kDims[l], net->Wt[l], kDims[l], e);
    in = w.h[l];
  }
  EpiArgs e;
  e.bias = net->P + net->off_b[3];
  e.outf = w.y;
  e.ldf = kOut;
  gemm<EPI_LINEAR>(st, rows, kOut, kWidth, w.h[2], kWidth, net->Wt[3], kWidth, e);
}

}  // namespace hosmlp

extern "C" {

int hos_mlp_info(int* n_in, int* width, int* n_out, long long* n_params) {
  return guard([&] {
    if (n_in) *n_in = kIn;
    if (width) *width = kWidth;
    if (n_out) *n_out = kOut;
    long long n = 0;
    for (int l = 0; l < kLayers; ++l) n += (long long)kDims[l + 1] * kDims[l] + kDims[l + 1];
    if (n_params) *n_params = n;
  });
}

int hos_mlp_create(int max_rows, hos_mlp** out) {
  return guard([&] {
    if (!out) throw std::runtime_error("null argument");
    if (max_rows <= 0 || max_rows % kRowTile) throw std::runtime_error("MLP batch capacity: multiple of 64");
    auto m = std::make_unique<hos_mlp>();
    MLP_OK(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
    size_t off = 0;
    for (int l = 0; l < kLayers; ++l) {
      m->off_w[l] = off;
      off += (size_t)kDims[l + 1] * kDims[l];
      m->off_b[l] = off;
      off += kDims[l + 1];
    }
    m->nparam = off;
    m->P = m->get<float>(off);
    m->G = m->get<float>(off);
    m->Mo = m->get<float>(off);
    m->Ve = m->get<float>(off);
    for (int l = 0; l < kLayers; ++l) {
      m->Wt[l] = m->get<__bf16>((size_t)kDims[l + 1] * kDims[l]);
      m->Wk[l] = m->get<__bf16>((size_t)kDims[l + 1] * kDims[l]);
    }
    m->max_rows = max_rows;
    m->X = m->get<__bf16>((size_t)max_rows * kIn);
    m->Xt = m->get<__bf16>((size_t)max_rows * kIn);
    for (int l = 0; l < 3; ++l) {
      m->H[l] = m->get<__bf16>((size_t)max_rows * kWidth);
      m->Ht[l] = m->get<__bf16>((size_t)max_rows * kWidth);
    }
    for (int l = 0; l < kLayers; ++l) {
      m->dZ[l] = m->get<__bf16>((size_t)max_rows * kDims[l + 1]);
      m->dZt[l] = m->get<__bf16>((size_t)max_rows * kDims[l + 1]);
    }
    m->Y = m->get<float>((size_t)max_rows * kOut);
    m->T = m->get<float>((size_t)max_rows * kOut);
    m->loss = m->get<float>(1);
    m->idx = m->get<int>(max_rows);
    m->xs = m->get<float>((size_t)max_rows * kIn);
    m->ts = m->get<float>((size_t)max_rows * kOut);
    m->xb = m->get<__bf16>((size_t)max_rows * kIn);
    MLP_OK(hipDeviceSynchronize());
    *out = m.release();
  });
}

int hos_mlp_destroy(hos_mlp* m) {
  return guard([&] { delete m; });
}

int hos_mlp_set_params(hos_mlp* m, const float* p) {
  return guard([&] {
    if (!m || !p) throw std::runtime_error("null argument");
    MLP_OK(hipMemcpyAsync(m->P, p, m->nparam * sizeof(float), hipMemcpyHostToDevice, m->st));
    MLP_OK(hipMemsetAsync(m->Mo, 0, m->nparam * sizeof(float), m->st));
    MLP_OK(hipMemsetAsync(m->Ve, 0, m->nparam * sizeof(float), m->st));
    m->step = 0;
    m->refresh();
    MLP_OK(hipStreamSynchronize(m->st));
  });
}

int hos_mlp_get_params(hos_mlp* m, float* p) {
  return guard([&] {
    if (!m || !p) throw std::runtime_error("null argument");
    MLP_OK(hipMemcpyAsync(p, m->P, m->nparam * sizeof(float), hipMemcpyDeviceToHost, m->st));
    MLP_OK(hipStreamSynchronize(m->st));
  });
}

int hos_mlp_forward(hos_mlp* m, int rows, const float* x, float* y) {
  return guard([&] {
    if (!m || !x || !y) throw std::runtime_error("null argument");
    const int padded = (rows + kRowTile - 1) / kRowTile * kRowTile;
    m->check_rows(padded);
    m->upload_batch(padded, rows, x, nullptr);
    m->forward_train(padded);
    MLP_OK(hipMemcpyAsync(y, m->Y, (size_t)rows * kOut * sizeof(float), hipMemcpyDeviceToHost, m->st));
    MLP_OK(hipStreamSynchronize(m->st));
  });
}

int hos_mlp_loss_grad(hos_mlp* m, int rows, const float* x, const float* t, double* loss, float* grad) {
  return guard([&] {
    if (!m || !x || !t) throw std::runtime_error("null argument");
    const int padded = (rows + kRowTile - 1) / kRowTile * kRowTile;
    m->check_rows(padded);
    m->upload_batch(padded, rows, x, t);
    m->forward_train(padded);
    MLP_OK(hipMemsetAsync(m->loss, 0, sizeof(float), m->st));
    m->loss_grad(padded, rows);
    m->backward(padded);
    float l = 0.f;
    MLP_OK(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->st));
    if (grad) MLP_OK(hipMemcpyAsync(grad, m->G, m->nparam * sizeof(float), hipMemcpyDeviceToHost, m->st));
    MLP_OK(hipStreamSynchronize(m->st));
    if (loss) *loss = (double)l / ((double)rows * kOut);
  });
}

// No counterpart in the reference: hos_mlp_forward on device arrays, allocating
// nothing (the activations use the static scratch array of this GPU).
int hos_mlp_forward_noalloc(const hos_mlp* m, int rows, const float* x_dev, float* y_dev, void* stream) {
  return guard([&] {
    if (!m || !x_dev || !y_dev) throw std::runtime_error("null argument");
    if (rows < 1 || rows > kNoAllocRows)
      throw std::runtime_error("hos_mlp_forward_noalloc: rows must be in 1.." + std::to_string(kNoAllocRows));
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    unsigned char* base = scratch_ptr();
    const int prows = (rows + kRowTile - 1) / kRowTile * kRowTile;
    __bf16* xb = reinterpret_cast<__bf16*>(base);
    __bf16* h[3];
    for (int l = 0; l < 3; ++l) h[l] = reinterpret_cast<__bf16*>(base + kScratchX + l * kScratchH);
    float* yp = reinterpret_cast<float*>(base + kScratchX + 3 * kScratchH);
    hipLaunchKernelGGL(k_rows_bf16, dim3(nblk((size_t)prows * kIn)), dim3(256), 0, st, rows, prows, x_dev, xb);
    const __bf16* in = xb;
    for (int l = 0; l < 3; ++l) {
      EpiArgs e;
      e.bias = m->P + m->off_b[l];
      e.out = h[l];
      e.ldo = kWidth;
      gemm<EPI_RELU>(st, prows, kWidth, kDims[l], in, kDims[l], m->Wt[l], kDims[l], e);
      in = h[l];
    }
    EpiArgs e;
    e.bias = m->P + m->off_b[3];
    e.outf = yp;
    e.ldf = kOut;
    gemm<EPI_LINEAR>(st, prows, kOut, kWidth, h[2], kWidth, m->Wt[3], kWidth, e);
    MLP_OK(hipMemcpyAsync(y_dev, yp, (size_t)rows * kOut * sizeof(float), hipMemcpyDeviceToDevice, st));
    MLP_OK(hipGetLastError());
  });
}

int hos_mlp_set_data(hos_mlp* m, int n, const float* x, const float* t) {
  return guard([&] {
    if (!m || (n > 0 && (!x || !t)) || n < 0) throw std::runtime_error("bad training set");
    MLP_OK(hipStreamSynchronize(m->st));
    if (m->dX) (void)hipFree(m->dX);
    if (m->dY) (void)hipFree(m->dY);
    m->dX = nullptr;
    m->dY = nullptr;
    m->n_data = 0;
    if (n == 0) return;
    m->dX = dmalloc<__bf16>((size_t)n * kIn);
    m->dY = dmalloc<float>((size_t)n * kOut);
    float* tmp = dmalloc<float>((size_t)n * kIn);
    MLP_OK(hipMemcpyAsync(tmp, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice, m->st));
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk((size_t)n * kIn)), dim3(256), 0, m->st, (size_t)n * kIn,
                       (const float*)tmp, m->dX);
    MLP_OK(hipMemcpyAsync(m->dY, t, (size_t)n * kOut * sizeof(float), hipMemcpyHostToDevice, m->st));
    MLP_OK(hipStreamSynchronize(m->st));
    (void)hipFree(tmp);
    m->n_data = n;
  });
}

// Adam over the device training set: `epochs` passes in a seeded random order,
// batches of `batch` rows (the last one padded); epoch_loss[e] = mean squared
// error over the epoch's batches; *ms = the GPU time of all steps.
int hos_mlp_train(hos_mlp* m, int epochs, int batch, float lr, unsigned long long seed, float* epoch_loss,
                  double* ms) {
  return guard([&] {
    if (!m || m->n_data <= 0) throw std::runtime_error("no training set");
    m->check_rows(batch);
    std::mt19937_64 rng(seed);
    std::vector<int> perm(m->n_data);
    int* didx = dmalloc<int>(m->n_data);
    hipEvent_t e0, e1;
    MLP_OK(hipEventCreate(&e0));
    MLP_OK(hipEventCreate(&e1));
    double tot_ms = 0.0;
    for (int ep = 0; ep < epochs; ++ep) {
      for (int i = 0; i < m->n_data; ++i) perm[i] = i;
      std::shuffle(perm.begin(), perm.end(), rng);
      MLP_OK(hipMemcpyAsync(didx, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, m->st));
      MLP_OK(hipMemsetAsync(m->loss, 0, sizeof(float), m->st));
      MLP_OK(hipEventRecord(e0, m->st));
      for (int b0 = 0; b0 < m->n_data; b0 += batch) {
        const int nv = std::min(batch, m->n_data - b0);
        hipLaunchKernelGGL(k_gather, dim3(nblk((size_t)batch * kIn)), dim3(256), 0, m->st, batch, nv,
                           (const int*)(didx + b0), (const __bf16*)m->dX, (const float*)m->dY, m->X, m->Xt, m->T);
        m->forward_train(batch);
        m->loss_grad(batch, nv);
        m->backward(batch);
        m->adam(lr);
      }
      MLP_OK(hipEventRecord(e1, m->st));
      float l = 0.f, t = 0.f;
      MLP_OK(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->st));
      MLP_OK(hipStreamSynchronize(m->st));
      MLP_OK(hipEventElapsedTime(&t, e0, e1));
      tot_ms += t;
      if (epoch_loss) epoch_loss[ep] = l / ((float)m->n_data * kOut);  // mean squared error of the epoch
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(didx);
    if (ms) *ms = tot_ms;
  });
}

}  // extern "C"
