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

// The edge features of e = (i <- j): ((x_i - x_j)/R, |x_i - x_j|/R), the differences
// formed in FP64, the rest in FP32.
__device__ __forceinline__ void sn_feat(double2 xi, double2 xj, double invR, float& fx, float& fy, float& fr) {
  fx = (float)((xi.x - xj.x) * invR);
  fy = (float)((xi.y - xj.y) * invR);
  fr = sqrtf(fx * fx + fy * fy);
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

// ------------------------------------------------------------ fused rollout step
//
// Inference of the rollout network at width 128 in 1 + S launches: k_sn_node<SN_ENC>
// runs the node encoder and step 0's projections; k_sn_node<SN_STEP> (or SN_LAST) one
// message-passing step: the edge sum of every receiver, the node MLP, the residual
// update and the next step's projections (or the decoder).  Every node layer of a
// 16-node strip stays inside one wavefront: the strip's BF16 activations sit in a
// wavefront-private LDS tile (rows of 136) that feeds v_mfma_f32_16x16x32_bf16 as the
// A operand; the weights, transposed (row n = output column n, the bias at column
// 128), are read as B fragments from memory (L2-resident: 0.2 MB per step).  Biases
// start the accumulators, so every GEMM has K = 128; the edge layer's bias multiplies
// the edge count.  zc travels in BF16 (the gathered operand), zb and h in FP32.

#include <hip/hip_bf16.h>
typedef __bf16 sn_bf16;
typedef __bf16 sn_v8bf __attribute__((ext_vector_type(8)));
typedef float sn_v4f __attribute__((ext_vector_type(4)));

#define SN_W 128                   // width of the fused path
#define SN_KP 136                  // row stride of the transposed weights (bias at column 128)
#define SN_LD 136                  // LDS row stride of a strip tile (bf16)
#define SN_STRIP 16                // nodes per wavefront
#define SN_ENC 0
#define SN_STEP 1
#define SN_LAST 2

struct SnStep {
  int n, n_recv;
  const int* off;
  const int* send;
  const double2* pos;
  double invR;
  const float* A;        // this step's edge layer, FP32 rows 0..3 of 128
  const float* zb;       // n x 128 FP32
  const sn_bf16* zc;     // n x 128 BF16
  const sn_bf16* Dt;     // 128 x 136
  const sn_bf16* Et;     // 128 x 136
  const sn_bf16* Ft;     // 128 x 128
  const sn_bf16* Gt;     // 128 x 136
  float* h;              // n x 128 FP32, updated in place
  const sn_bf16* Bt;     // next step's projections, 128 x 136 each
  const sn_bf16* Ct;
  float* zb_out;
  sn_bf16* zc_out;
  const sn_bf16* P1t;    // decoder, 128 x 136 and 8 x 136
  const sn_bf16* P2t;
  float* out;            // n x 8 FP32
  const sn_bf16* xn;     // encoder input, n x 16 BF16
  const sn_bf16* L1t;    // 128 x 16
  const sn_bf16* L2t;    // 128 x 136
};

__device__ __forceinline__ void sn_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// acc[j] (16x16 tile j of a 16 x 128 strip) += sA (16 x 128, LDS) . W, W given transposed
// (wt: 128 x ldw, columns 0..127).  Lane layout of v_mfma_f32_16x16x32_bf16: A and B
// fragments hold row / column (lane % 16), k = (lane / 16) * 8 + 0..7; the result holds
// column (lane % 16), rows (lane / 16) * 4 + 0..3.
__device__ __forceinline__ void sn_strip_mm(const sn_bf16* sA, const sn_bf16* __restrict__ wt, int ldw, sn_v4f* acc,
                                            int lane) {
  const int ar = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
  for (int kc = 0; kc < SN_W; kc += 32) {
    const sn_v8bf a = *(const sn_v8bf*)(sA + ar * SN_LD + kc + kq);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const sn_v8bf b = *(const sn_v8bf*)(wt + (size_t)(16 * j + ar) * ldw + kc + kq);
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
    }
  }
}

// acc[j] = the bias of output column 16 j + lane % 16 (column 128 of wt), times scale[r].
__device__ __forceinline__ void sn_bias(const sn_bf16* __restrict__ wt, sn_v4f* acc, int lane, const float* scale) {
  const int ar = lane & 15;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const float bv = (float)wt[(size_t)(16 * j + ar) * SN_KP + SN_W];
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[j][r] = scale ? bv * scale[(lane >> 4) * 4 + r] : bv;
  }
}

// sX = bf16(acc) (relu first when RELU), in the strip's row-major layout.
template <bool RELU>
__device__ __forceinline__ void sn_store(sn_bf16* sX, const sn_v4f* acc, int lane) {
  const int ar = lane & 15;
#pragma unroll
  for (int j = 0; j < 8; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float v = RELU ? fmaxf(acc[j][r], 0.f) : acc[j][r];
      sX[((lane >> 4) * 4 + r) * SN_LD + 16 * j + ar] = (sn_bf16)v;
    }
}

// zb_out (FP32) and zc_out (BF16) of the strip from its latent tile sH.
__device__ __forceinline__ void sn_project(const SnStep& p, const sn_bf16* sH, int i0, int lane) {
  const int ar = lane & 15;
  sn_v4f acc[8];
  sn_bias(p.Bt, acc, lane, nullptr);
  sn_strip_mm(sH, p.Bt, SN_KP, acc, lane);
#pragma unroll
  for (int j = 0; j < 8; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + (lane >> 4) * 4 + r;
      if (i < p.n) p.zb_out[(size_t)i * SN_W + 16 * j + ar] = acc[j][r];
    }
  sn_bias(p.Ct, acc, lane, nullptr);
  sn_strip_mm(sH, p.Ct, SN_KP, acc, lane);
#pragma unroll
  for (int j = 0; j < 8; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + (lane >> 4) * 4 + r;
      if (i < p.n) p.zc_out[(size_t)i * SN_W + 16 * j + ar] = (sn_bf16)acc[j][r];
    }
}

// acc = h_old + acc (rows of the strip), written to p.h and, as BF16, to sH.
__device__ __forceinline__ void sn_residual(const SnStep& p, sn_bf16* sH, sn_v4f* acc, int i0, int lane) {
  const int ar = lane & 15;
#pragma unroll
  for (int j = 0; j < 8; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = (lane >> 4) * 4 + r, i = i0 + row;
      float v = 0.f;
      if (i < p.n) {
        float* hp = p.h + (size_t)i * SN_W + 16 * j + ar;
        v = *hp + acc[j][r];
        *hp = v;
      }
      sH[row * SN_LD + 16 * j + ar] = (sn_bf16)v;
    }
}

template <int MODE>
__global__ __launch_bounds__(64 * SN_WAVES) void k_sn_node(SnStep p) {
  __shared__ __align__(16) sn_bf16 tiles[SN_WAVES][2][SN_STRIP * SN_LD];
  __shared__ float degs[SN_WAVES][SN_STRIP];
  const int lane = threadIdx.x & 63, wl = threadIdx.x >> 6;
  const int strip = blockIdx.x * SN_WAVES + wl;
  const int i0 = __builtin_amdgcn_readfirstlane(strip * SN_STRIP);
  if (i0 >= p.n) return;   // a whole wavefront; nothing below synchronises the block
  sn_bf16* sH = tiles[wl][0];
  sn_bf16* sX = tiles[wl][1];
  const int ar = lane & 15, kq = (lane >> 4) * 8;
  sn_v4f acc[8];
  if (MODE == SN_ENC) {
    // h = [relu(x L1),1] L2: x is 16 wide, so the upper half of the K = 32 fragment is 0
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = (sn_v4f){0.f, 0.f, 0.f, 0.f};
    sn_v8bf a;
#pragma unroll
    for (int q = 0; q < 8; ++q) a[q] = (sn_bf16)0.0f;
    const int ir = i0 + ar;
    if (kq < 16 && ir < p.n) a = *(const sn_v8bf*)(p.xn + (size_t)ir * 16 + kq);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      sn_v8bf b;
#pragma unroll
      for (int q = 0; q < 8; ++q) b[q] = (sn_bf16)0.0f;
      if (kq < 16) b = *(const sn_v8bf*)(p.L1t + (size_t)(16 * j + ar) * 16 + kq);
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
    }
    sn_store<true>(sX, acc, lane);
    sn_wave_sync();
    sn_bias(p.L2t, acc, lane, nullptr);
    sn_strip_mm(sX, p.L2t, SN_KP, acc, lane);
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = (lane >> 4) * 4 + r, i = i0 + row;
        if (i < p.n) p.h[(size_t)i * SN_W + 16 * j + ar] = acc[j][r];
        sH[row * SN_LD + 16 * j + ar] = (sn_bf16)(i < p.n ? acc[j][r] : 0.f);
      }
    sn_wave_sync();
    sn_project(p, sH, i0, lane);
    return;
  }
  // --- the edge sums of the strip's receivers: four groups of 16 lanes take one node
  // each (four rounds), a lane holding features 8 (lane % 16) .. + 7, so four nodes'
  // edge chains are in flight per wavefront
  const int grp = lane >> 4, f = 8 * ar;
  float a0[8], a1[8], a2[8], a3[8];
  {
    const float4* Ap = (const float4*)p.A;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float4 u = Ap[(k * SN_W + f) / 4], v = Ap[(k * SN_W + f) / 4 + 1];
      float* dst = k == 0 ? a0 : k == 1 ? a1 : k == 2 ? a2 : a3;
      dst[0] = u.x; dst[1] = u.y; dst[2] = u.z; dst[3] = u.w;
      dst[4] = v.x; dst[5] = v.y; dst[6] = v.z; dst[7] = v.w;
    }
  }
  for (int q = 0; q < SN_STRIP / 4; ++q) {
    const int r = 4 * q + grp, i = i0 + r;
    float acc8[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc8[t] = 0.f;
    sn_v8bf hv;
#pragma unroll
    for (int t = 0; t < 8; ++t) hv[t] = (sn_bf16)0.0f;
    int deg = 0;
    if (i < p.n) {
      const float4* hp = (const float4*)(p.h + (size_t)i * SN_W + f);
      const float4 u = hp[0], v = hp[1];
      hv[0] = (sn_bf16)u.x; hv[1] = (sn_bf16)u.y; hv[2] = (sn_bf16)u.z; hv[3] = (sn_bf16)u.w;
      hv[4] = (sn_bf16)v.x; hv[5] = (sn_bf16)v.y; hv[6] = (sn_bf16)v.z; hv[7] = (sn_bf16)v.w;
    }
    if (i < p.n_recv) {
      const int e0 = p.off[i], e1 = p.off[i + 1];
      deg = e1 - e0;
      float zbi[8];
      {
        const float4* zp = (const float4*)(p.zb + (size_t)i * SN_W + f);
        const float4 u = zp[0], v = zp[1];
        zbi[0] = u.x; zbi[1] = u.y; zbi[2] = u.z; zbi[3] = u.w;
        zbi[4] = v.x; zbi[5] = v.y; zbi[6] = v.z; zbi[7] = v.w;
#pragma unroll
        for (int t = 0; t < 8; ++t) zbi[t] += a3[t];
      }
      const double2 xi = p.pos[i];
      // four edges per trip, their gathers issued together; the next trip's sender ids
      // are loaded before this trip's arithmetic (past the end: the receiver itself,
      // masked out)
      int jn[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) jn[u] = e0 + u < e1 ? p.send[e0 + u] : i;
      for (int e = e0; e < e1; e += 4) {
        double2 xj[4];
        sn_v8bf zc8[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          xj[u] = p.pos[jn[u]];
          zc8[u] = *(const sn_v8bf*)(p.zc + (size_t)jn[u] * SN_W + f);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) jn[u] = e + 4 + u < e1 ? p.send[e + 4 + u] : i;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (e + u >= e1) continue;
          float fx, fy, fr;
          sn_feat(xi, xj[u], p.invR, fx, fy, fr);
#pragma unroll
          for (int t = 0; t < 8; ++t)
            acc8[t] += fmaxf(zbi[t] + (float)zc8[u][t] + a0[t] * fx + a1[t] * fy + a2[t] * fr, 0.f);
        }
      }
    }
    sn_v8bf sv;
#pragma unroll
    for (int t = 0; t < 8; ++t) sv[t] = (sn_bf16)acc8[t];
    *(sn_v8bf*)(sX + r * SN_LD + f) = sv;
    *(sn_v8bf*)(sH + r * SN_LD + f) = hv;
    if (ar == 0) degs[wl][r] = (float)deg;
  }
  sn_wave_sync();
  // m = [s, deg] D
  sn_bias(p.Dt, acc, lane, degs[wl]);
  sn_strip_mm(sX, p.Dt, SN_KP, acc, lane);
  sn_wave_sync();
  sn_store<false>(sX, acc, lane);
  sn_wave_sync();
  // y = [h,1] E + m F;  u = relu(y)
  sn_bias(p.Et, acc, lane, nullptr);
  sn_strip_mm(sH, p.Et, SN_KP, acc, lane);
  sn_strip_mm(sX, p.Ft, SN_W, acc, lane);
  sn_wave_sync();
  sn_store<true>(sX, acc, lane);
  sn_wave_sync();
  // h += [u,1] G
  sn_bias(p.Gt, acc, lane, nullptr);
  sn_strip_mm(sX, p.Gt, SN_KP, acc, lane);
  sn_wave_sync();
  sn_residual(p, sH, acc, i0, lane);
  sn_wave_sync();
  if (MODE == SN_STEP) {
    sn_project(p, sH, i0, lane);
    return;
  }
  // decoder: out = [relu([h,1] P1),1] P2, 8 output columns
  sn_bias(p.P1t, acc, lane, nullptr);
  sn_strip_mm(sH, p.P1t, SN_KP, acc, lane);
  sn_store<true>(sX, acc, lane);
  sn_wave_sync();
  sn_v4f o = (sn_v4f){0.f, 0.f, 0.f, 0.f};
  const float bo = ar < 8 ? (float)p.P2t[(size_t)ar * SN_KP + SN_W] : 0.f;
#pragma unroll
  for (int r = 0; r < 4; ++r) o[r] = bo;
#pragma unroll
  for (int kc = 0; kc < SN_W; kc += 32) {
    const sn_v8bf a = *(const sn_v8bf*)(sX + ar * SN_LD + kc + kq);
    sn_v8bf b;
#pragma unroll
    for (int q = 0; q < 8; ++q) b[q] = (sn_bf16)0.0f;
    if (ar < 8) b = *(const sn_v8bf*)(p.P2t + (size_t)ar * SN_KP + kc + kq);
    o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, o, 0, 0, 0);
  }
  if (ar < 8) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = i0 + (lane >> 4) * 4 + r;
      if (i < p.n) p.out[(size_t)i * 8 + ar] = o[r];
    }
  }
}

// One launch of the fused path; mode: 0 encoder (+ step 0 projections), 1 a step with
// the next step's projections, 2 the last step with the decoder.  Arguments as in SnStep
// (unused ones may be NULL).
extern "C" int uph_sn_fused(int mode, int n, int n_recv, const int* off, const int* send, const double* pos,
                            double invR, const float* A, const float* zb, const void* zc, const void* Dt,
                            const void* Et, const void* Ft, const void* Gt, float* h, const void* Bt, const void* Ct,
                            float* zb_out, void* zc_out, const void* P1t, const void* P2t, float* out,
                            const void* xn, const void* L1t, const void* L2t, void* stream) {
  if (n <= 0) return 0;
  if (n_recv < 0 || n_recv > n || mode < 0 || mode > 2) return sn_fail("uph_sn_fused: bad arguments");
  SnStep p;
  p.n = n;
  p.n_recv = n_recv;
  p.off = off;
  p.send = send;
  p.pos = (const double2*)pos;
  p.invR = invR;
  p.A = A;
  p.zb = zb;
  p.zc = (const sn_bf16*)zc;
  p.Dt = (const sn_bf16*)Dt;
  p.Et = (const sn_bf16*)Et;
  p.Ft = (const sn_bf16*)Ft;
  p.Gt = (const sn_bf16*)Gt;
  p.h = h;
  p.Bt = (const sn_bf16*)Bt;
  p.Ct = (const sn_bf16*)Ct;
  p.zb_out = zb_out;
  p.zc_out = (sn_bf16*)zc_out;
  p.P1t = (const sn_bf16*)P1t;
  p.P2t = (const sn_bf16*)P2t;
  p.out = out;
  p.xn = (const sn_bf16*)xn;
  p.L1t = (const sn_bf16*)L1t;
  p.L2t = (const sn_bf16*)L2t;
  const bool need_proj = mode != 2, need_step = mode != 0, need_dec = mode == 2, need_enc = mode == 0;
  if ((need_proj && (!Bt || !Ct || !zb_out || !zc_out)) ||
      (need_step && (!off || !send || !pos || !A || !zb || !zc || !Dt || !Et || !Ft || !Gt)) ||
      (need_dec && (!P1t || !P2t || !out)) || (need_enc && (!xn || !L1t || !L2t)) || !h)
    return sn_fail("uph_sn_fused: a buffer the mode needs is NULL");
  const int strips = (n + SN_STRIP - 1) / SN_STRIP;
  const int blocks = (strips + SN_WAVES - 1) / SN_WAVES;
  hipStream_t st = (hipStream_t)stream;
  if (mode == 0)
    hipLaunchKernelGGL(k_sn_node<SN_ENC>, dim3(blocks), dim3(64 * SN_WAVES), 0, st, p);
  else if (mode == 1)
    hipLaunchKernelGGL(k_sn_node<SN_STEP>, dim3(blocks), dim3(64 * SN_WAVES), 0, st, p);
  else
    hipLaunchKernelGGL(k_sn_node<SN_LAST>, dim3(blocks), dim3(64 * SN_WAVES), 0, st, p);
  SN_CHK(hipGetLastError());
  return 0;
}
