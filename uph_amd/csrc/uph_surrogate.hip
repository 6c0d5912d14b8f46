// uph_surrogate.hip -- edge kernels of the rollout network (uph_amd/surrogate_net.py).
//
// The rollout network keeps no per-edge latent: message-passing step l computes,
// for every receiver i over its edges e = (i <- j),
//   s_i = sum_e relu(r_e A + zb_i + zc_j),   r_e = ((x_i - x_j)/R, |x_i - x_j|/R, 1),
// and applies the edge MLP's second layer after that sum ([s_i, deg_i] D, a GEMM over
// nodes).  zb = [h,1] B and zc = [h,1] C come from node GEMMs, so the only per-edge
// work is this gather-add-relu-sum, done here with the edge features formed from the
// FP64 positions on the fly.  One wavefront per receiver (lanes over features); the
// edge loop is wavefront-uniform.  The backward pass recomputes each edge's
// pre-activation: the receiver pass forms dzb and the partial products of dA, the
// sender pass (over the sender-ordered permutation) forms dzc, so every sum runs in
// a fixed order.  No counterpart in the reference.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>
#include <mutex>

#include "../../include/uph.h"

namespace {
std::mutex g_sn_mu;
std::string g_sn_err;
int sn_fail(const char* what, hipError_t e) {
  char b[256];
  snprintf(b, sizeof b, "%s: %s", what, hipGetErrorString(e));
  std::lock_guard<std::mutex> lk(g_sn_mu);
  g_sn_err = b;
  return -1;
}
int sn_fail(const char* what) {
  std::lock_guard<std::mutex> lk(g_sn_mu);
  g_sn_err = what;
  return -1;
}
}  // namespace

#define SN_CHK(expr)                                   \
  do {                                                 \
    hipError_t e_ = (expr);                            \
    if (e_ != hipSuccess) return sn_fail(#expr, e_);   \
  } while (0)

#define SN_WAVES 4                 // wavefronts per block
#define SN_MAXF 2                  // features per lane: widths up to 128
#define SN_PART_BLOCKS 256         // blocks of the dA receiver pass (fixed: partial sums)

extern "C" const char* uph_sn_last_error(void) {
  std::lock_guard<std::mutex> lk(g_sn_mu);
  return g_sn_err.c_str();
}

// The edge features of e = (i <- j): ((x_i - x_j)/R, |x_i - x_j|/R) in FP32.
__device__ __forceinline__ void sn_feat(double2 xi, double2 xj, double invR, float& fx, float& fy, float& fr) {
  const double dx = (xi.x - xj.x) * invR, dy = (xi.y - xj.y) * invR;
  fx = (float)dx;
  fy = (float)dy;
  fr = (float)sqrt(dx * dx + dy * dy);
}

// s (FP32, row stride lds >= W + 1): columns 0..W-1 the sum over the receiver's
// edges of relu(r_e A + zb_i + zc_j), column W the edge count, W+1..lds-1 zero.
// Rows of non-receivers (i >= n_recv) are zero.  A: FP32, rows 0..3 (row stride ldA).
__global__ __launch_bounds__(64 * SN_WAVES) void k_sn_edge_agg(int n, int n_recv, int W, const int* __restrict__ off,
                                                              const int* __restrict__ send,
                                                              const double2* __restrict__ pos, double invR,
                                                              const float* __restrict__ A, int ldA,
                                                              const float* __restrict__ zb, int ldzb,
                                                              const float* __restrict__ zc, int ldzc,
                                                              float* __restrict__ s, int lds) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * SN_WAVES + (threadIdx.x >> 6);
  const int nwaves = gridDim.x * SN_WAVES;
  float a0[SN_MAXF], a1[SN_MAXF], a2[SN_MAXF], a3[SN_MAXF];
#pragma unroll
  for (int t = 0; t < SN_MAXF; ++t) {
    const int f = lane + 64 * t;
    const bool on = f < W;
    a0[t] = on ? A[f] : 0.f;
    a1[t] = on ? A[ldA + f] : 0.f;
    a2[t] = on ? A[2 * ldA + f] : 0.f;
    a3[t] = on ? A[3 * ldA + f] : 0.f;
  }
  for (int i = wave; i < n; i += nwaves) {
    float acc[SN_MAXF] = {0.f, 0.f};
    int deg = 0;
    if (i < n_recv) {
      const int e0 = off[i], e1 = off[i + 1];
      deg = e1 - e0;
      float zbi[SN_MAXF];
#pragma unroll
      for (int t = 0; t < SN_MAXF; ++t) {
        const int f = lane + 64 * t;
        zbi[t] = f < W ? zb[(size_t)i * ldzb + f] + a3[t] : 0.f;
      }
      const double2 xi = pos[i];
      for (int e = e0; e < e1; ++e) {
        const int j = send[e];
        float fx, fy, fr;
        sn_feat(xi, pos[j], invR, fx, fy, fr);
#pragma unroll
        for (int t = 0; t < SN_MAXF; ++t) {
          const int f = lane + 64 * t;
          if (f < W) {
            const float pre = zbi[t] + zc[(size_t)j * ldzc + f] + a0[t] * fx + a1[t] * fy + a2[t] * fr;
            acc[t] += fmaxf(pre, 0.f);
          }
        }
      }
    }
#pragma unroll
    for (int t = 0; t < SN_MAXF; ++t) {
      const int f = lane + 64 * t;
      if (f < W) s[(size_t)i * lds + f] = acc[t];
    }
    for (int f = W + lane; f < lds; f += 64) s[(size_t)i * lds + f] = f == W ? (float)deg : 0.f;
  }
}

// Receiver pass of the backward: g_e = ds_i * [pre_e > 0]; dzb_i = sum_e g_e (FP32, row
// stride lddz, W columns); per block, the partial sums of g_e r_e (4 x W) over the
// receivers its wavefronts take, in a fixed order, into part[block][4][W].
__global__ __launch_bounds__(64 * SN_WAVES) void k_sn_edge_bwd_recv(
    int n_recv, int W, const int* __restrict__ off, const int* __restrict__ send, const double2* __restrict__ pos,
    double invR, const float* __restrict__ A, int ldA, const float* __restrict__ zb, int ldzb,
    const float* __restrict__ zc, int ldzc, const float* __restrict__ ds, int ldds, float* __restrict__ dzb,
    int lddz, float* __restrict__ part) {
  __shared__ float red[SN_WAVES][4][64 * SN_MAXF];
  const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
  const int wave = blockIdx.x * SN_WAVES + wl;
  const int nwaves = gridDim.x * SN_WAVES;
  float a0[SN_MAXF], a1[SN_MAXF], a2[SN_MAXF], a3[SN_MAXF];
  float pa[4][SN_MAXF];
#pragma unroll
  for (int t = 0; t < SN_MAXF; ++t) {
    const int f = lane + 64 * t;
    const bool on = f < W;
    a0[t] = on ? A[f] : 0.f;
    a1[t] = on ? A[ldA + f] : 0.f;
    a2[t] = on ? A[2 * ldA + f] : 0.f;
    a3[t] = on ? A[3 * ldA + f] : 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) pa[k][t] = 0.f;
  }
  for (int i = wave; i < n_recv; i += nwaves) {
    float zbi[SN_MAXF], dsi[SN_MAXF], acc[SN_MAXF] = {0.f, 0.f};
#pragma unroll
    for (int t = 0; t < SN_MAXF; ++t) {
      const int f = lane + 64 * t;
      zbi[t] = f < W ? zb[(size_t)i * ldzb + f] + a3[t] : 0.f;
      dsi[t] = f < W ? ds[(size_t)i * ldds + f] : 0.f;
    }
    const double2 xi = pos[i];
    for (int e = off[i]; e < off[i + 1]; ++e) {
      const int j = send[e];
      float fx, fy, fr;
      sn_feat(xi, pos[j], invR, fx, fy, fr);
#pragma unroll
      for (int t = 0; t < SN_MAXF; ++t) {
        const int f = lane + 64 * t;
        if (f < W) {
          const float pre = zbi[t] + zc[(size_t)j * ldzc + f] + a0[t] * fx + a1[t] * fy + a2[t] * fr;
          const float g = pre > 0.f ? dsi[t] : 0.f;
          acc[t] += g;
          pa[0][t] += g * fx;
          pa[1][t] += g * fy;
          pa[2][t] += g * fr;
          pa[3][t] += g;
        }
      }
    }
#pragma unroll
    for (int t = 0; t < SN_MAXF; ++t) {
      const int f = lane + 64 * t;
      if (f < W) dzb[(size_t)i * lddz + f] = acc[t];
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int t = 0; t < SN_MAXF; ++t) red[wl][k][lane + 64 * t] = pa[k][t];
  __syncthreads();
  // the block's partial: its wavefronts' sums added in wavefront order
  for (int x = threadIdx.x; x < 4 * W; x += 64 * SN_WAVES) {
    const int k = x / W, f = x % W;
    float v = 0.f;
    for (int w = 0; w < SN_WAVES; ++w) v += red[w][k][f];
    part[((size_t)blockIdx.x * 4 + k) * W + f] = v;
  }
}

// dA (rows 0..3, row stride ldA) += the sum of the blocks' partials, in block order.
__global__ void k_sn_reduce_dA(int nblocks, int W, const float* __restrict__ part, float* __restrict__ dA, int ldA) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= 4 * W) return;
  const int k = x / W, f = x % W;
  float v = 0.f;
  for (int b = 0; b < nblocks; ++b) v += part[((size_t)b * 4 + k) * W + f];
  dA[(size_t)k * ldA + f] += v;
}

// Sender pass: dzc_j = sum over the edges e = (i <- j) of ds_i * [pre_e > 0], the
// edges of j taken through the sender-ordered permutation (soff, sperm).
__global__ __launch_bounds__(64 * SN_WAVES) void k_sn_edge_bwd_send(
    int n, int W, const int* __restrict__ soff, const int* __restrict__ sperm, const int* __restrict__ recv,
    const double2* __restrict__ pos, double invR, const float* __restrict__ A, int ldA,
    const float* __restrict__ zb, int ldzb, const float* __restrict__ zc, int ldzc, const float* __restrict__ ds,
    int ldds, float* __restrict__ dzc, int lddz) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * SN_WAVES + (threadIdx.x >> 6);
  const int nwaves = gridDim.x * SN_WAVES;
  float a0[SN_MAXF], a1[SN_MAXF], a2[SN_MAXF], a3[SN_MAXF];
#pragma unroll
  for (int t = 0; t < SN_MAXF; ++t) {
    const int f = lane + 64 * t;
    const bool on = f < W;
    a0[t] = on ? A[f] : 0.f;
    a1[t] = on ? A[ldA + f] : 0.f;
    a2[t] = on ? A[2 * ldA + f] : 0.f;
    a3[t] = on ? A[3 * ldA + f] : 0.f;
  }
  for (int j = wave; j < n; j += nwaves) {
    float zcj[SN_MAXF], acc[SN_MAXF] = {0.f, 0.f};
#pragma unroll
    for (int t = 0; t < SN_MAXF; ++t) {
      const int f = lane + 64 * t;
      zcj[t] = f < W ? zc[(size_t)j * ldzc + f] + a3[t] : 0.f;
    }
    const double2 xj = pos[j];
    for (int k = soff[j]; k < soff[j + 1]; ++k) {
      const int i = recv[sperm[k]];
      float fx, fy, fr;
      sn_feat(pos[i], xj, invR, fx, fy, fr);
#pragma unroll
      for (int t = 0; t < SN_MAXF; ++t) {
        const int f = lane + 64 * t;
        if (f < W) {
          const float pre = zb[(size_t)i * ldzb + f] + zcj[t] + a0[t] * fx + a1[t] * fy + a2[t] * fr;
          if (pre > 0.f) acc[t] += ds[(size_t)i * ldds + f];
        }
      }
    }
#pragma unroll
    for (int t = 0; t < SN_MAXF; ++t) {
      const int f = lane + 64 * t;
      if (f < W) dzc[(size_t)j * lddz + f] = acc[t];
    }
  }
}

static inline int sn_blocks(long waves) {
  const long b = (waves + SN_WAVES - 1) / SN_WAVES;
  return (int)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

static bool sn_width_ok(int W) { return W > 0 && W <= 64 * SN_MAXF; }

extern "C" int uph_sn_edge_agg(int n, int n_recv, int W, const int* off, const int* send, const double* pos,
                               double invR, const float* A, int ldA, const float* zb, int ldzb, const float* zc,
                               int ldzc, float* s, int lds, void* stream) {
  if (!sn_width_ok(W) || lds < W + 1 || n_recv > n) return sn_fail("uph_sn_edge_agg: bad sizes");
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_sn_edge_agg, dim3(sn_blocks(n)), dim3(64 * SN_WAVES), 0, (hipStream_t)stream, n, n_recv, W,
                     off, send, (const double2*)pos, invR, A, ldA, zb, ldzb, zc, ldzc, s, lds);
  SN_CHK(hipGetLastError());
  return 0;
}

// part: SN_PART_BLOCKS * 4 * W floats of scratch.
extern "C" int uph_sn_edge_bwd(int n, int n_recv, int W, const int* off, const int* send, const int* soff,
                               const int* sperm, const int* recv, const double* pos, double invR, const float* A,
                               int ldA, const float* zb, int ldzb, const float* zc, int ldzc, const float* ds,
                               int ldds, float* dzb, float* dzc, int lddz, float* dA, float* part, void* stream) {
  if (!sn_width_ok(W) || n_recv > n) return sn_fail("uph_sn_edge_bwd: bad sizes");
  if (n <= 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (n_recv > 0) {
    hipLaunchKernelGGL(k_sn_edge_bwd_recv, dim3(SN_PART_BLOCKS), dim3(64 * SN_WAVES), 0, st, n_recv, W, off, send,
                       (const double2*)pos, invR, A, ldA, zb, ldzb, zc, ldzc, ds, ldds, dzb, lddz, part);
    hipLaunchKernelGGL(k_sn_reduce_dA, dim3((4 * W + 255) / 256), dim3(256), 0, st, SN_PART_BLOCKS, W, part, dA,
                       ldA);
  }
  // rows of dzb past the receivers are zero
  if (n > n_recv) SN_CHK(hipMemsetAsync(dzb + (size_t)n_recv * lddz, 0, (size_t)(n - n_recv) * lddz * sizeof(float), st));
  hipLaunchKernelGGL(k_sn_edge_bwd_send, dim3(sn_blocks(n)), dim3(64 * SN_WAVES), 0, st, n, W, soff, sperm, recv,
                     (const double2*)pos, invR, A, ldA, zb, ldzb, zc, ldzc, ds, ldds, dzc, lddz);
  SN_CHK(hipGetLastError());
  return 0;
}
