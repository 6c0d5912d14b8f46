// uph_gnn.hip -- kernels of the learned acceleration surrogate (graph network).
//
// The network's dense layers run as hand-written BF16 MFMA GEMMs
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation) with fused bias/activation;
// weight gradients use the same GEMM over a transposed operand with a split-K
// reduction so every output tile gets many work-groups.  The graph is the SPH
// neighbour graph of the solver's cell list (radius = the search radius), kept
// in CSR order by receiver so every aggregation is a deterministic segment
// sum; the backward scatter to senders goes through a sender-ordered
// permutation instead of atomics.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <string.h>
#include <stdio.h>
#include <string>
#include <mutex>
#include <algorithm>

#include "../../include/uph.h"

typedef __bf16 bf16_t;
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

namespace {
std::string g_gnn_err;
std::mutex g_gnn_mu;

int fail(const char* what, hipError_t e) {
  char b[256];
  snprintf(b, sizeof b, "%s: %s", what, hipGetErrorString(e));
  std::lock_guard<std::mutex> lk(g_gnn_mu);
  g_gnn_err = b;
  return -1;
}
int fail(const char* what) {
  std::lock_guard<std::mutex> lk(g_gnn_mu);
  g_gnn_err = what;
  return -1;
}
}  // namespace

#define CHK(expr)                          \
  do {                                     \
    hipError_t e_ = (expr);                \
    if (e_ != hipSuccess) return fail(#expr, e_); \
  } while (0)

// ------------------------------------------------------------ buffers

extern "C" int uph_dev_alloc(void** p, size_t bytes) {
  *p = nullptr;
  CHK(hipMalloc(p, bytes ? bytes : 16));
  return 0;
}
extern "C" int uph_dev_free(void* p) {
  if (p) CHK(hipFree(p));
  return 0;
}
extern "C" int uph_dev_upload(void* dst, const void* src, size_t bytes) {
  CHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return 0;
}
extern "C" int uph_dev_download(void* dst, const void* src, size_t bytes) {
  CHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}
extern "C" int uph_dev_copy(void* dst, const void* src, size_t bytes) {
  CHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return 0;
}
extern "C" int uph_dev_zero(void* p, size_t bytes) {
  CHK(hipMemset(p, 0, bytes));
  return 0;
}
extern "C" int uph_dev_sync(void) {
  CHK(hipDeviceSynchronize());
  return 0;
}
extern "C" const char* uph_gnn_last_error(void) {
  std::lock_guard<std::mutex> lk(g_gnn_mu);
  return g_gnn_err.c_str();
}

// ------------------------------------------------------------ GEMM

#define GBM 64
#define GBN 64
#define GBK 32
#define GPAD 8

__device__ __forceinline__ float act_fwd(float x, int act) { return act == 1 ? fmaxf(x, 0.f) : x; }

// C[M x N] = act(op(A) . B + bias), op(A) = A (M x K, row stride lda) or, with TA,
// A^T where A is stored K x M (row stride lda).  B is K x N (row stride ldb).
// With split-K (gridDim.z > 1) each z-slice writes its FP32 partial product to
// part + z*M*N (no bias/activation); k_gemm_reduce finishes.
// With TB, B is given transposed: stored N x K (row stride ldb), as the weight
// copies W^T are, so it stages with vector loads and stores.
// Requires: N, lda, ldb, ldc multiples of 8; K a multiple of 8 unless TA.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const bf16_t* __restrict__ A, int lda,
                                              const bf16_t* __restrict__ B, int ldb, const float* __restrict__ bias,
                                              void* __restrict__ C, int ldc, int out_bf16, int act, int kchunk,
                                              float* __restrict__ part) {
  __shared__ bf16_t As[GBM][GBK + GPAD];
  __shared__ bf16_t Bs[GBN][GBK + GPAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
  const int kb = blockIdx.z * kchunk;
  const int ke = min(K, kb + kchunk);
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  v4f acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
  const bf16_t zero = (bf16_t)0.0f;
  for (int k0 = kb; k0 < ke; k0 += GBK) {
    // stage A (64 x 32) as As[m][k]
    if (!TA) {
      const int r = tid >> 2, k8 = (tid & 3) * 8;
      const int gm = m0 + r, gk = k0 + k8;
      if (gm < M && gk + 7 < ke) {
        *(v8bf*)&As[r][k8] = *(const v8bf*)(A + (size_t)gm * lda + gk);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) As[r][k8 + e] = (gm < M && gk + e < ke) ? A[(size_t)gm * lda + gk + e] : zero;
      }
    } else {
      const int kk = tid >> 3, m8 = (tid & 7) * 8;
      const int gk = k0 + kk, gm = m0 + m8;
      if (gk < ke && gm + 7 < M) {
        const v8bf v = *(const v8bf*)(A + (size_t)gk * lda + gm);
#pragma unroll
        for (int e = 0; e < 8; ++e) As[m8 + e][kk] = v[e];
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) As[m8 + e][kk] = (gk < ke && gm + e < M) ? A[(size_t)gk * lda + gm + e] : zero;
      }
    }
    // stage B (32 x 64) transposed as Bs[n][k]
    if (TB) {
      const int r = tid >> 2, k8 = (tid & 3) * 8;
      const int gn = n0 + r, gk = k0 + k8;
      if (gn < N && gk + 7 < ke) {
        *(v8bf*)&Bs[r][k8] = *(const v8bf*)(B + (size_t)gn * ldb + gk);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) Bs[r][k8 + e] = (gn < N && gk + e < ke) ? B[(size_t)gn * ldb + gk + e] : zero;
      }
    } else {
      const int kk = tid >> 3, n8 = (tid & 7) * 8;
      const int gk = k0 + kk, gn = n0 + n8;
      if (gk < ke && gn + 7 < N) {
        const v8bf v = *(const v8bf*)(B + (size_t)gk * ldb + gn);
#pragma unroll
        for (int e = 0; e < 8; ++e) Bs[n8 + e][kk] = v[e];
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) Bs[n8 + e][kk] = (gk < ke && gn + e < N) ? B[(size_t)gk * ldb + gn + e] : zero;
      }
    }
    __syncthreads();
    // one 16x16x32 MFMA per (mi, ni): lane holds row/col (lane % 16), k = (lane / 16) * 8 + 0..7
    v8bf a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      a[i] = *(const v8bf*)&As[wr + i * 16 + (lane & 15)][(lane >> 4) * 8];
      b[i] = *(const v8bf*)&Bs[wc + i * 16 + (lane & 15)][(lane >> 4) * 8];
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // D layout: lane holds column (lane % 16), rows (lane / 16) * 4 + 0..3
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wr + i * 16 + (lane >> 4) * 4 + r;
        const int gn = n0 + wc + j * 16 + (lane & 15);
        if (gm >= M || gn >= N) continue;
        float v = acc[i][j][r];
        if (gridDim.z > 1) {
          part[(size_t)blockIdx.z * M * N + (size_t)gm * N + gn] = v;
          continue;
        }
        if (bias) v += bias[gn];
        v = act_fwd(v, act);
        if (out_bf16)
          ((bf16_t*)C)[(size_t)gm * ldc + gn] = (bf16_t)v;
        else
          ((float*)C)[(size_t)gm * ldc + gn] = v;
      }
}

__global__ void k_gemm_reduce(int M, int N, int splits, const float* __restrict__ part, const float* __restrict__ bias,
                              void* __restrict__ C, int ldc, int out_bf16, int act, int accumulate) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)M * N) return;
  const int gm = (int)(i / N), gn = (int)(i % N);
  float v = 0.f;
  for (int z = 0; z < splits; ++z) v += part[(size_t)z * M * N + i];
  if (bias) v += bias[gn];
  v = act_fwd(v, act);
  if (out_bf16) {
    ((bf16_t*)C)[(size_t)gm * ldc + gn] = (bf16_t)v;
  } else {
    float* c = (float*)C + (size_t)gm * ldc + gn;
    *c = accumulate ? *c + v : v;
  }
}

extern "C" int uph_gemm_bf16(int M, int N, int K, const void* A, int lda, int trans, const void* B, int ldb,
                             const float* bias, void* C, int ldc, int out_bf16, int act, int splits, float* work,
                             int accumulate, void* stream) {
  if (M <= 0 || N <= 0 || K <= 0) return 0;
  const int trans_a = trans & 1, trans_b = (trans >> 1) & 1;
  if (trans & ~3) return fail("uph_gemm_bf16: trans takes bit 0 (A^T) and bit 1 (B given as N x K)");
  if (trans_b && (K & 7)) return fail("uph_gemm_bf16: K must be a multiple of 8 when B is given as N x K");
  if ((N & 7) || (lda & 7) || (ldb & 7) || (ldc & 7) || (!trans_a && (K & 7)))
    return fail("uph_gemm_bf16: N, K (unless trans_a) and leading dimensions must be multiples of 8");
  if (splits < 1) splits = 1;
  if ((splits > 1 || accumulate) && !work) return fail("uph_gemm_bf16: split-K or accumulate needs a workspace");
  if (accumulate && out_bf16) return fail("uph_gemm_bf16: accumulate needs an FP32 output");
  if (accumulate && splits == 1) splits = 1;
  hipStream_t st = (hipStream_t)stream;
  int kchunk = (K + splits - 1) / splits;
  kchunk = ((kchunk + GBK - 1) / GBK) * GBK;
  splits = (K + kchunk - 1) / kchunk;
  const bool via_work = splits > 1 || accumulate;
  dim3 grid((N + GBN - 1) / GBN, (M + GBM - 1) / GBM, via_work ? (splits > 1 ? splits : 1) : 1);
  // a single-slice accumulate still goes through the workspace
  if (via_work && splits == 1) grid.z = 1;
  const bool part_path = via_work;
#define UPH_GEMM(TA_, TB_)                                                                                        \
  hipLaunchKernelGGL((k_gemm<TA_, TB_>), grid, dim3(256), 0, st, M, N, K, (const bf16_t*)A, lda, (const bf16_t*)B, \
                     ldb, part_path ? nullptr : bias, part_path ? (void*)work : C, part_path ? N : ldc,             \
                     part_path ? 0 : out_bf16, part_path ? 0 : act, kchunk, work)
  if (trans_a && trans_b)
    UPH_GEMM(true, true);
  else if (trans_a)
    UPH_GEMM(true, false);
  else if (trans_b)
    UPH_GEMM(false, true);
  else
    UPH_GEMM(false, false);
#undef UPH_GEMM
  if (part_path) {
    const long mn = (long)M * N;
    hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, st, M, N, (int)grid.z, work,
                       bias, C, ldc, out_bf16, act, accumulate);
  }
  CHK(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------ elementwise

__global__ void k_f32_to_bf16(long n, const float* __restrict__ x, bf16_t* __restrict__ y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (bf16_t)x[i];
}
__global__ void k_bf16_to_f32(long n, const bf16_t* __restrict__ x, float* __restrict__ y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}
// dz = dy * (y > 0) (ReLU backward through its output), FP32 in, BF16 out
__global__ void k_relu_bwd(long n, const float* __restrict__ dy, const bf16_t* __restrict__ y,
                           bf16_t* __restrict__ dz) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dz[i] = (bf16_t)((float)y[i] > 0.f ? dy[i] : 0.f);
}
// y[r, 0:n] += x[r, 0:n] for FP32 rows (residual connections, gradient sums)
__global__ void k_add_f32(long n, const float* __restrict__ x, float* __restrict__ y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += x[i];
}
// column sums of an FP32 M x N matrix (bias gradients), one thread per column
__global__ void k_colsum(int M, int N, const float* __restrict__ x, int ldx, float* __restrict__ out, int accumulate) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float s = 0.f;
  for (int r = 0; r < M; ++r) s += x[(size_t)r * ldx + c];
  out[c] = accumulate ? out[c] + s : s;
}

static inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

extern "C" int uph_cast_f32_bf16(long n, const float* x, void* y, void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, n, x, (bf16_t*)y);
  CHK(hipGetLastError());
  return 0;
}
extern "C" int uph_cast_bf16_f32(long n, const void* x, float* y, void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_bf16_to_f32, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, n, (const bf16_t*)x, y);
  CHK(hipGetLastError());
  return 0;
}
extern "C" int uph_relu_bwd(long n, const float* dy, const void* y, void* dz, void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, n, dy, (const bf16_t*)y,
                     (bf16_t*)dz);
  CHK(hipGetLastError());
  return 0;
}
extern "C" int uph_add_f32(long n, const float* x, float* y, void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_add_f32, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, n, x, y);
  CHK(hipGetLastError());
  return 0;
}
extern "C" int uph_colsum_f32(int M, int N, const float* x, int ldx, float* out, int accumulate, void* stream) {
  if (N <= 0) return 0;
  hipLaunchKernelGGL(k_colsum, dim3(nblk(N)), dim3(256), 0, (hipStream_t)stream, M, N, x, ldx, out, accumulate);
  CHK(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------ graph ops

// out[i, :] = sum over e in [off[i], off[i+1]) of x[idx ? idx[e] : e, :]  (FP32 or BF16 in, FP32 out)
template <bool XBF>
__global__ void k_segsum(int nseg, int F, const int* __restrict__ off, const int* __restrict__ idx,
                         const void* __restrict__ x, int ldx, float* __restrict__ out, int ldo, int accumulate) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nseg * F) return;
  const int i = (int)(t / F), f = (int)(t % F);
  float s = 0.f;
  for (int e = off[i]; e < off[i + 1]; ++e) {
    const int r = idx ? idx[e] : e;
    s += XBF ? (float)((const bf16_t*)x)[(size_t)r * ldx + f] : ((const float*)x)[(size_t)r * ldx + f];
  }
  float* o = out + (size_t)i * ldo + f;
  *o = accumulate ? *o + s : s;
}

extern "C" int uph_segsum(int nseg, int F, const int* off, const int* idx, const void* x, int ldx, int x_bf16,
                          float* out, int ldo, int accumulate, void* stream) {
  if (nseg <= 0 || F <= 0) return 0;
  const long n = (long)nseg * F;
  if (x_bf16)
    hipLaunchKernelGGL(k_segsum<true>, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, nseg, F, off, idx, x, ldx,
                       out, ldo, accumulate);
  else
    hipLaunchKernelGGL(k_segsum<false>, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, nseg, F, off, idx, x, ldx,
                       out, ldo, accumulate);
  CHK(hipGetLastError());
  return 0;
}

// y[e, :] = x[rows[e], :]  (row gather; BF16 or FP32 by element size)
__global__ void k_gather_rows(int n, int F, const int* __restrict__ rows, const char* __restrict__ x, int ldx_bytes,
                              char* __restrict__ y, int ldy_bytes, int esize) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n * F) return;
  const int e = (int)(t / F), f = (int)(t % F);
  const char* s = x + (size_t)rows[e] * ldx_bytes + (size_t)f * esize;
  char* d = y + (size_t)e * ldy_bytes + (size_t)f * esize;
  if (esize == 2)
    *(uint16_t*)d = *(const uint16_t*)s;
  else
    *(uint32_t*)d = *(const uint32_t*)s;
}

extern "C" int uph_gather_rows(int n, int F, const int* rows, const void* x, int ldx, void* y, int ldy, int esize,
                               void* stream) {
  if (n <= 0 || F <= 0) return 0;
  if (esize != 2 && esize != 4) return fail("uph_gather_rows: element size must be 2 or 4");
  const long t = (long)n * F;
  hipLaunchKernelGGL(k_gather_rows, dim3(nblk(t)), dim3(256), 0, (hipStream_t)stream, n, F, rows, (const char*)x,
                     ldx * esize, (char*)y, ldy * esize, esize);
  CHK(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------ optimiser / loss

// Adam on FP32 master weights; writes the BF16 copy used by the GEMMs.
__global__ void k_adam(long n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, bf16_t* __restrict__ wb, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = (bf16_t)wi;
}

extern "C" int uph_adam(long n, float* w, const float* g, float* m, float* v, void* w_bf16, float lr, float b1,
                        float b2, float eps, int t, void* stream) {
  if (n <= 0) return 0;
  const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, n, w, g, m, v, (bf16_t*)w_bf16, lr,
                     b1, b2, eps, c1, c2);
  CHK(hipGetLastError());
  return 0;
}

// Masked MSE: d[i, c] = 2 (y - t) / count over rows with mask; per-block partial losses.
__global__ void k_mse(int M, int C, const float* __restrict__ y, int ldy, const float* __restrict__ t, int ldt,
                      const int* __restrict__ mask, float inv_count, float* __restrict__ dy, float* __restrict__ lpart) {
  __shared__ float red[256];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < (long)M * C) {
    const int r = (int)(i / C), c = (int)(i % C);
    const bool on = !mask || mask[r];
    const float d = on ? y[(size_t)r * ldy + c] - t[(size_t)r * ldt + c] : 0.f;
    l = d * d * inv_count;
    dy[(size_t)r * ldy + c] = 2.f * d * inv_count;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

extern "C" int uph_mse(int M, int C, const float* y, int ldy, const float* t, int ldt, const int* mask, int count,
                       float* dy, float* loss_part, void* stream) {
  if (M <= 0 || C <= 0 || count <= 0) return fail("uph_mse: empty input");
  const long n = (long)M * C;
  hipLaunchKernelGGL(k_mse, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, M, C, y, ldy, t, ldt, mask,
                     1.0f / (float)count, dy, loss_part);
  CHK(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------ surrogate features

// Node features (BF16, n x 16): columns 0-9 are the velocities at the last five
// macro steps divided by vscale (hist holds 5 x n x 2 doubles, newest first),
// 10 the wall flag (i >= n_recv), 11 the constant 1 that carries the first
// layer's bias, 12-15 zero.
__global__ void k_gns_node_feats(int n, int n_recv, const double* __restrict__ hist, float inv_vscale,
                                 bf16_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bf16_t* o = out + (size_t)i * 16;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    o[2 * k] = (bf16_t)((float)hist[((size_t)k * n + i) * 2] * inv_vscale);
    o[2 * k + 1] = (bf16_t)((float)hist[((size_t)k * n + i) * 2 + 1] * inv_vscale);
  }
  o[10] = (bf16_t)(i >= n_recv ? 1.0f : 0.0f);
  o[11] = (bf16_t)1.0f;
#pragma unroll
  for (int k = 12; k < 16; ++k) o[k] = (bf16_t)0.0f;
}

// Edge features (BF16, E x 8): (x_recv - x_send) / R, |r| / R, the constant 1, zeros.
__global__ void k_gns_edge_feats(long E, const int* __restrict__ send, const int* __restrict__ recv,
                                 const double* __restrict__ pos, double invR, bf16_t* __restrict__ out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int i = recv[e], j = send[e];
  const double dx = (pos[2 * i] - pos[2 * j]) * invR, dy = (pos[2 * i + 1] - pos[2 * j + 1]) * invR;
  bf16_t* o = out + (size_t)e * 8;
  o[0] = (bf16_t)(float)dx;
  o[1] = (bf16_t)(float)dy;
  o[2] = (bf16_t)(float)sqrt(dx * dx + dy * dy);
  o[3] = (bf16_t)1.0f;
#pragma unroll
  for (int k = 4; k < 8; ++k) o[k] = (bf16_t)0.0f;
}

// Semi-implicit Euler on the predicted acceleration (FP64 state):
// v += dt * ascale * a;  x += dt * v   for the receivers (fluid) only.
__global__ void k_gns_integrate(int n_recv, double* __restrict__ pos, double* __restrict__ vel,
                                const float* __restrict__ acc, int lda, double dt, double ascale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_recv) return;
  const double vx = vel[2 * i] + dt * ascale * (double)acc[(size_t)i * lda];
  const double vy = vel[2 * i + 1] + dt * ascale * (double)acc[(size_t)i * lda + 1];
  vel[2 * i] = vx;
  vel[2 * i + 1] = vy;
  pos[2 * i] += dt * vx;
  pos[2 * i + 1] += dt * vy;
}

// SPH summation density of the receivers over the radius graph (cubic spline,
// derivatives.c:44-66): rho_i = m_i W(0) + sum_j m_j W(r_ij).
__global__ void k_sph_density(int n_recv, const int* __restrict__ off, const int* __restrict__ send,
                              const double* __restrict__ pos, const double* __restrict__ mass, double h,
                              double* __restrict__ rho) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_recv) return;
  const double alpha = 10.0 / (7.0 * M_PI * h * h);
  double s = mass[i] * alpha;
  for (int e = off[i]; e < off[i + 1]; ++e) {
    const int j = send[e];
    const double dx = pos[2 * i] - pos[2 * j], dy = pos[2 * i + 1] - pos[2 * j + 1];
    const double q = sqrt(dx * dx + dy * dy) / h;
    double w = 0.0;
    if (q < 1.0)
      w = alpha * (1.0 - 1.5 * q * q + 0.75 * q * q * q);
    else if (q < 2.0)
      w = alpha * 0.25 * (2.0 - q) * (2.0 - q) * (2.0 - q);
    s += mass[j] * w;
  }
  rho[i] = s;
}

extern "C" int uph_gns_node_feats(int n, int n_recv, const double* hist, float inv_vscale, void* out, void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_gns_node_feats, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, n, n_recv, hist,
                     inv_vscale, (bf16_t*)out);
  CHK(hipGetLastError());
  return 0;
}
extern "C" int uph_gns_edge_feats(long E, const int* send, const int* recv, const double* pos, double invR, void* out,
                                  void* stream) {
  if (E <= 0) return 0;
  hipLaunchKernelGGL(k_gns_edge_feats, dim3(nblk(E)), dim3(256), 0, (hipStream_t)stream, E, send, recv, pos, invR,
                     (bf16_t*)out);
  CHK(hipGetLastError());
  return 0;
}
extern "C" int uph_gns_integrate(int n_recv, double* pos, double* vel, const float* acc, int lda, double dt,
                                 double ascale, void* stream) {
  if (n_recv <= 0) return 0;
  hipLaunchKernelGGL(k_gns_integrate, dim3(nblk(n_recv)), dim3(256), 0, (hipStream_t)stream, n_recv, pos, vel, acc,
                     lda, dt, ascale);
  CHK(hipGetLastError());
  return 0;
}
extern "C" int uph_sph_density(int n_recv, const int* off, const int* send, const double* pos, const double* mass,
                               double h, double* rho, void* stream) {
  if (n_recv <= 0) return 0;
  hipLaunchKernelGGL(k_sph_density, dim3(nblk(n_recv)), dim3(256), 0, (hipStream_t)stream, n_recv, off, send, pos,
                     mass, h, rho);
  CHK(hipGetLastError());
  return 0;
}

// W^T as BF16 (K x N FP32 -> N x K BF16), for the input-gradient GEMMs.
__global__ void k_transpose_bf16(int K, int N, const float* __restrict__ w, bf16_t* __restrict__ wt) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)K * N) return;
  const int k = (int)(t / N), n = (int)(t % N);
  wt[(size_t)n * K + k] = (bf16_t)w[t];
}
extern "C" int uph_transpose_bf16(int K, int N, const float* w, void* wt, void* stream) {
  if (K <= 0 || N <= 0) return 0;
  hipLaunchKernelGGL(k_transpose_bf16, dim3(nblk((long)K * N)), dim3(256), 0, (hipStream_t)stream, K, N, w,
                     (bf16_t*)wt);
  CHK(hipGetLastError());
  return 0;
}

// z[e, :] += a[ra[e], :] + b[rb[e], :]   (FP32 rows; the per-node projections of
// the edge MLP's first layer added to its per-edge part)
__global__ void k_gather_add2(long E, int F, const int* __restrict__ ra, const float* __restrict__ a,
                              const int* __restrict__ rb, const float* __restrict__ b, int ldab,
                              float* __restrict__ z, int ldz) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * F) return;
  const long e = t / F;
  const int f = (int)(t % F);
  z[e * ldz + f] += a[(size_t)ra[e] * ldab + f] + b[(size_t)rb[e] * ldab + f];
}
extern "C" int uph_gather_add2(long E, int F, const int* ra, const float* a, const int* rb, const float* b, int ldab,
                               float* z, int ldz, void* stream) {
  if (E <= 0 || F <= 0) return 0;
  hipLaunchKernelGGL(k_gather_add2, dim3(nblk(E * F)), dim3(256), 0, (hipStream_t)stream, E, F, ra, a, rb, b, ldab, z,
                     ldz);
  CHK(hipGetLastError());
  return 0;
}

// y = relu(z) as BF16 into a row of width ld (columns F..ld-1 untouched)
__global__ void k_relu_to_bf16(long M, int F, const float* __restrict__ z, int ldz, bf16_t* __restrict__ y, int ldy) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M * F) return;
  const long r = t / F;
  const int f = (int)(t % F);
  y[r * ldy + f] = (bf16_t)fmaxf(z[r * ldz + f], 0.f);
}
extern "C" int uph_relu_to_bf16(long M, int F, const float* z, int ldz, void* y, int ldy, void* stream) {
  if (M <= 0 || F <= 0) return 0;
  hipLaunchKernelGGL(k_relu_to_bf16, dim3(nblk(M * F)), dim3(256), 0, (hipStream_t)stream, M, F, z, ldz, (bf16_t*)y,
                     ldy);
  CHK(hipGetLastError());
  return 0;
}

// FP32 rows -> BF16 rows of width ldy (the latent streams' GEMM inputs)
__global__ void k_rows_to_bf16(long M, int F, const float* __restrict__ x, int ldx, bf16_t* __restrict__ y, int ldy) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M * F) return;
  const long r = t / F;
  const int f = (int)(t % F);
  y[r * ldy + f] = (bf16_t)x[r * ldx + f];
}
extern "C" int uph_rows_to_bf16(long M, int F, const float* x, int ldx, void* y, int ldy, void* stream) {
  if (M <= 0 || F <= 0) return 0;
  hipLaunchKernelGGL(k_rows_to_bf16, dim3(nblk(M * F)), dim3(256), 0, (hipStream_t)stream, M, F, x, ldx, (bf16_t*)y,
                     ldy);
  CHK(hipGetLastError());
  return 0;
}

// Fill column `col` of an M x ld BF16 matrix with `v` (the constant-1 bias feature)
__global__ void k_fill_col(long M, int ld, int col, float v, bf16_t* __restrict__ y) {
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < M) y[r * ld + col] = (bf16_t)v;
}
extern "C" int uph_fill_col_bf16(long M, int ld, int col, float v, void* y, void* stream) {
  if (M <= 0) return 0;
  hipLaunchKernelGGL(k_fill_col, dim3(nblk(M)), dim3(256), 0, (hipStream_t)stream, M, ld, col, v, (bf16_t*)y);
  CHK(hipGetLastError());
  return 0;
}

// dz[r, f] = dy[r, f] * (y[r, f] > 0) on the first F columns of strided rows
__global__ void k_relu_bwd_rows(long M, int F, const float* __restrict__ dy, int lddy, const bf16_t* __restrict__ y,
                                int ldy, bf16_t* __restrict__ dz, int ldz) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M * F) return;
  const long r = t / F;
  const int f = (int)(t % F);
  dz[r * ldz + f] = (bf16_t)((float)y[r * ldy + f] > 0.f ? dy[r * lddy + f] : 0.f);
}
extern "C" int uph_relu_bwd_rows(long M, int F, const float* dy, int lddy, const void* y, int ldy, void* dz, int ldz,
                                 void* stream) {
  if (M <= 0 || F <= 0) return 0;
  hipLaunchKernelGGL(k_relu_bwd_rows, dim3(nblk(M * F)), dim3(256), 0, (hipStream_t)stream, M, F, dy, lddy,
                     (const bf16_t*)y, ldy, (bf16_t*)dz, ldz);
  CHK(hipGetLastError());
  return 0;
}

// y[r, 0:F] += x[r, 0:F] on strided FP32 rows
__global__ void k_add_rows(long M, int F, const float* __restrict__ x, int ldx, float* __restrict__ y, int ldy) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= M * F) return;
  const long r = t / F;
  const int f = (int)(t % F);
  y[r * ldy + f] += x[r * ldx + f];
}
extern "C" int uph_add_rows_f32(long M, int F, const float* x, int ldx, float* y, int ldy, void* stream) {
  if (M <= 0 || F <= 0) return 0;
  hipLaunchKernelGGL(k_add_rows, dim3(nblk(M * F)), dim3(256), 0, (hipStream_t)stream, M, F, x, ldx, y, ldy);
  CHK(hipGetLastError());
  return 0;
}

// out = a x + b y over FP64 arrays
__global__ void k_lincomb_f64(long n, double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                              double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a * x[i] + b * y[i];
}
extern "C" int uph_lincomb_f64(long n, double a, const double* x, double b, const double* y, double* out,
                               void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_lincomb_f64, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, n, a, x, b, y, out);
  CHK(hipGetLastError());
  return 0;
}

// out = x * num / den over FP64 arrays
__global__ void k_scale_ratio_f64(long n, const double* __restrict__ x, const double* __restrict__ num,
                                  const double* __restrict__ den, double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = x[i] * (num[i] / den[i]);
}
extern "C" int uph_scale_ratio_f64(long n, const double* x, const double* num, const double* den, double* out,
                                   void* stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_scale_ratio_f64, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, n, x, num, den, out);
  CHK(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------ fused edge step
//
// One message-passing step on the edges, width 128, in one kernel:
//   z  = [e,1] A + zb[recv] + zc[send];  u = relu(z)            (GEMM 1 + gathers)
//   m  = [u,1] D                                                (GEMM 2)
//   e += m;  eb_out = bf16(e);  m_out = m                        (residual update)
// A tile of 192 edges stays in LDS between the two GEMMs (u never goes to HBM
// unless u1_out asks for it, as training does), the two weight matrices stay
// in LDS for the whole persistent loop, and the MFMA tiles are those of k_gemm
// (16x16x32 BF16, FP32 accumulation).  Per edge the HBM traffic is the BF16
// input row, the FP32 latent read and written, the BF16 output row and the
// FP32 message.
#define EW 128            // latent width
#define EKP 136           // row width of the BF16 inputs: latent, the constant 1, zeros
#define EKC 160           // K extent of the MFMA loop (5 chunks of 32); columns >= EKP are zero
#define ELD 168           // LDS row stride (bf16)
#define ETILE 192         // edges per tile: 12 wavefronts of 16 rows
#define ETHREADS 768

__global__ __launch_bounds__(ETHREADS) void k_edge_step(long E, const bf16_t* __restrict__ eb_in,
                                                        bf16_t* __restrict__ eb_out, float* __restrict__ e,
                                                        const float* __restrict__ zb, const float* __restrict__ zc,
                                                        const int* __restrict__ recv, const int* __restrict__ send,
                                                        const bf16_t* __restrict__ At, const bf16_t* __restrict__ Dt,
                                                        float* __restrict__ m_out, bf16_t* __restrict__ u1_out) {
  extern __shared__ __align__(16) unsigned char esm[];
  bf16_t* sA = (bf16_t*)esm;
  bf16_t* sD = sA + EW * ELD;
  bf16_t* sX = sD + EW * ELD;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const bf16_t zero = (bf16_t)0.0f;
  v8bf z8;
#pragma unroll
  for (int q = 0; q < 8; ++q) z8[q] = zero;
  // transposed weights (row n = output column n), zero past column EKP
  for (int idx = tid; idx < EW * (ELD / 8); idx += ETHREADS) {
    const int n = idx / (ELD / 8), k8 = (idx % (ELD / 8)) * 8;
    v8bf va = z8, vd = z8;
    if (k8 + 8 <= EKP) {
      va = *(const v8bf*)(At + (size_t)n * EKP + k8);
      vd = *(const v8bf*)(Dt + (size_t)n * EKP + k8);
    }
    *(v8bf*)(sA + n * ELD + k8) = va;
    *(v8bf*)(sD + n * ELD + k8) = vd;
  }
  for (int idx = tid; idx < ETILE * ((ELD - EKP) / 8); idx += ETHREADS) {
    const int r = idx / ((ELD - EKP) / 8), k8 = EKP + (idx % ((ELD - EKP) / 8)) * 8;
    *(v8bf*)(sX + r * ELD + k8) = z8;
  }
  __syncthreads();
  const long ntiles = (E + ETILE - 1) / ETILE;
  const int ar = 16 * w + (lane & 15), kq = (lane >> 4) * 8;
  for (long t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const long e0 = t * ETILE;
    for (int idx = tid; idx < ETILE * (EKP / 8); idx += ETHREADS) {
      const int r = idx / (EKP / 8), k8 = (idx % (EKP / 8)) * 8;
      const long er = e0 + r;
      *(v8bf*)(sX + r * ELD + k8) = er < E ? *(const v8bf*)(eb_in + er * EKP + k8) : z8;
    }
    __syncthreads();
    // the latent rows this lane will update are loaded first, so their HBM latency
    // overlaps both GEMMs
    float ev[8][4];
    int rcv[4], snd[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const long er = e0 + 16 * w + (lane >> 4) * 4 + i;
      const bool live = er < E;
      rcv[i] = live ? recv[er] : 0;
      snd[i] = live ? send[er] : 0;
#pragma unroll
      for (int j = 0; j < 8; ++j) ev[j][i] = live ? e[er * EW + 16 * j + (lane & 15)] : 0.f;
    }
    v4f acc[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < EKC; kc += 32) {
      const v8bf a = *(const v8bf*)(sX + ar * ELD + kc + kq);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const v8bf bb = *(const v8bf*)(sA + (16 * j + (lane & 15)) * ELD + kc + kq);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[j], 0, 0, 0);
      }
    }
    // D layout: lane holds column (lane % 16) of each 16x16 tile, rows (lane / 16) * 4 + i
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int col = 16 * j + (lane & 15);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 16 * w + (lane >> 4) * 4 + i;
        const long er = e0 + r;
        const float z = acc[j][i] + zb[(size_t)rcv[i] * EW + col] + zc[(size_t)snd[i] * EW + col];
        const bf16_t u = (bf16_t)fmaxf(z, 0.f);
        sX[r * ELD + col] = er < E ? u : zero;
        if (u1_out && er < E) u1_out[er * EKP + col] = u;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kc = 0; kc < EKC; kc += 32) {
      const v8bf a = *(const v8bf*)(sX + ar * ELD + kc + kq);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const v8bf bb = *(const v8bf*)(sD + (16 * j + (lane & 15)) * ELD + kc + kq);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[j], 0, 0, 0);
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int col = 16 * j + (lane & 15);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const long er = e0 + 16 * w + (lane >> 4) * 4 + i;
        if (er >= E) continue;
        const float mv = acc[j][i];
        const float en = ev[j][i] + mv;
        e[er * EW + col] = en;
        eb_out[er * EKP + col] = (bf16_t)en;
        m_out[er * EW + col] = mv;
      }
    }
    __syncthreads();
  }
}

extern "C" int uph_gns_edge_step(long E, const void* eb_in, void* eb_out, float* e, const float* zb, const float* zc,
                                 const int* recv, const int* send, const void* At, const void* Dt, float* m_out,
                                 void* u1_out, void* stream) {
  if (E <= 0) return 0;
  const size_t lds = (size_t)(2 * EW + ETILE) * ELD * sizeof(bf16_t);
  static int ready = 0;
  if (!ready) {
    int dev = 0, maxlds = 0;
    CHK(hipGetDevice(&dev));
    CHK(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if ((size_t)maxlds < lds) return fail("uph_gns_edge_step: the device has too little LDS per work-group");
    CHK(hipFuncSetAttribute((const void*)k_edge_step, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    ready = 1;
  }
  int dev = 0, ncu = 0;
  CHK(hipGetDevice(&dev));
  CHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const long ntiles = (E + ETILE - 1) / ETILE;
  const int grid = (int)std::min<long>(ntiles, (long)ncu);
  hipLaunchKernelGGL(k_edge_step, dim3(grid), dim3(ETHREADS), lds, (hipStream_t)stream, E, (const bf16_t*)eb_in,
                     (bf16_t*)eb_out, e, zb, zc, recv, send, (const bf16_t*)At, (const bf16_t*)Dt, m_out,
                     (bf16_t*)u1_out);
  CHK(hipGetLastError());
  return 0;
}
