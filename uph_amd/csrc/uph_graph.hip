// uph_graph.hip -- the radius-graph builder of the learned surrogate.
#include "uph_internal.h"

// Radius graph on device positions: receivers are the first n_recv particles.
__global__ void k_graph_count(int n, StepBufs b, const DevScene* __restrict__ sc, int n_recv, int* __restrict__ cnt,
                              int by_sender) {
  const int p = xcd_block() * BLK + threadIdx.x;
  if (p >= n) return;
  const DevScene& S = sc[0];
  const uint32_t idp = tag_id(b.tagA[p]);
  if (!by_sender && (int)idp >= n_recv) {
    cnt[idp] = 0;
    return;
  }
  const double2 pp = b.posA[p];
  int m = 0;
  for_candidates(S, pp, b.cellStart, [&](int q) {
    if (q == p) return;
    if (by_sender && (int)tag_id(b.tagA[q]) >= n_recv) return;
    if (dist2(pp, b.posA[q]) < S.cutoff) ++m;
  });
  cnt[idp] = m;
}

__global__ void k_graph_emit(int n, StepBufs b, const DevScene* __restrict__ sc, int n_recv,
                             const int* __restrict__ off, int* __restrict__ send, int* __restrict__ recv) {
  const int p = xcd_block() * BLK + threadIdx.x;
  if (p >= n) return;
  const DevScene& S = sc[0];
  const uint32_t idp = tag_id(b.tagA[p]);
  if ((int)idp >= n_recv) return;
  const double2 pp = b.posA[p];
  int o = off[idp];
  for_candidates(S, pp, b.cellStart, [&](int q) {
    if (q == p) return;
    if (dist2(pp, b.posA[q]) < S.cutoff) {
      send[o] = (int)tag_id(b.tagA[q]);
      recv[o] = (int)idp;
      ++o;
    }
  });
}

// For node j: the edges (i -> j) of every receiver i, in candidate order.
__global__ void k_graph_sender_perm(int n, StepBufs b, const DevScene* __restrict__ sc, int n_recv,
                                    const int* __restrict__ off, const int* __restrict__ send,
                                    const int* __restrict__ soff, int* __restrict__ sperm) {
  const int q = xcd_block() * BLK + threadIdx.x;
  if (q >= n) return;
  const DevScene& S = sc[0];
  const int j = (int)tag_id(b.tagA[q]);
  const double2 pq = b.posA[q];
  int o = soff[j];
  for_candidates(S, pq, b.cellStart, [&](int p) {
    if (p == q) return;
    const int i = (int)tag_id(b.tagA[p]);
    if (i >= n_recv || !(dist2(pq, b.posA[p]) < S.cutoff)) return;
    int e = off[i];
    while (e < off[i + 1] && send[e] != j) ++e;
    sperm[o++] = e;
  });
}

#define BBOX_BLOCKS 128

// Block b's bounding box {min x, min y, max x, max y} of its grid-stride share of the
// positions into part[4 b ..]; k_bbox_final folds the blocks' boxes into out[0..3].
__global__ __launch_bounds__(256) void k_bbox(int n, const double2* __restrict__ pos, double* __restrict__ part) {
  __shared__ double red[4][256];
  double lx = INFINITY, ly = INFINITY, hx = -INFINITY, hy = -INFINITY;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double2 p = pos[i];
    lx = fmin(lx, p.x); ly = fmin(ly, p.y); hx = fmax(hx, p.x); hy = fmax(hy, p.y);
  }
  red[0][threadIdx.x] = lx; red[1][threadIdx.x] = ly; red[2][threadIdx.x] = hx; red[3][threadIdx.x] = hy;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if (threadIdx.x < s2) {
      red[0][threadIdx.x] = fmin(red[0][threadIdx.x], red[0][threadIdx.x + s2]);
      red[1][threadIdx.x] = fmin(red[1][threadIdx.x], red[1][threadIdx.x + s2]);
      red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + s2]);
      red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + s2]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// One wavefront; nb <= 2 x 64.
__global__ void k_bbox_final(int nb, const double* __restrict__ part, double* __restrict__ out) {
  const int l = threadIdx.x;
  double v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = k < 2 ? INFINITY : -INFINITY;
    for (int b = l; b < nb; b += 64) v[k] = k < 2 ? fmin(v[k], part[4 * b + k]) : fmax(v[k], part[4 * b + k]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double w = __shfl_xor(v[k], o, 64);
      v[k] = k < 2 ? fmin(v[k], w) : fmax(v[k], w);
    }
  }
  if (l == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = v[k];
}

namespace {
struct GraphCtx {
  std::mutex mu;
  Err err;
  hipStream_t st = nullptr;
  int cap_n = 0, cap_cells = 0, cap_nb = 0;
  StepBufs b{};
  DevScene* dsc = nullptr;
  int *cnt = nullptr, *qpartial = nullptr;
  double* bbox = nullptr;
};
GraphCtx g_graph;

int graph_reserve(GraphCtx& c, int n, int ncells, int nb) {
  if (!c.st) HIPCHK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking), c.err);
  if (!c.dsc) HIPCHK(hipMalloc((void**)&c.dsc, sizeof(DevScene)), c.err);
  if (!c.bbox) HIPCHK(hipMalloc((void**)&c.bbox, 4 * sizeof(double)), c.err);
  if (n > c.cap_n || ncells > c.cap_cells || nb > c.cap_nb) {
    free_bufs(c.b);
    if (c.cnt) (void)hipFree(c.cnt);
    if (c.qpartial) (void)hipFree(c.qpartial);
    c.cnt = c.qpartial = nullptr;
    const int nn = std::max(n, c.cap_n), nc = std::max(ncells, c.cap_cells);
    const int nbb = std::max(std::max(nb, c.cap_nb), (nn + SCAN_TILE - 1) / SCAN_TILE + 1);
    if (alloc_bufs(c.b, nn, nc, nbb, c.err, true)) return -1;
    HIPCHK(hipMemset(c.b.cellCount, 0, (size_t)nc * sizeof(int)), c.err);
    HIPCHK(hipMalloc((void**)&c.cnt, ((size_t)nn + 1) * sizeof(int)), c.err);
    HIPCHK(hipMalloc((void**)&c.qpartial, (size_t)nbb * sizeof(int)), c.err);
    c.cap_n = nn;
    c.cap_cells = nc;
    c.cap_nb = nbb;
  }
  return 0;
}

__global__ void k_graph_tags(int n, int n_recv, const double2* __restrict__ src, double2* __restrict__ pos,
                             uint32_t* __restrict__ tag, int* __restrict__ sceneOf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pos[i] = src[i];
  tag[i] = ((uint32_t)(i < n_recv ? 1 : 2) << UPH_ID_BITS) | (uint32_t)i;
  sceneOf[i] = 0;
}
}  // namespace

extern "C" const char* uph_graph_last_error(void) { return g_graph.err.msg.c_str(); }

extern "C" int uph_graph_build(const double* pos, int n, int n_recv, double dist, int* off, int* send, int* recv,
                               int* soff, int* sperm, long capacity, long* n_edges) {
  GraphCtx& c = g_graph;
  std::lock_guard<std::mutex> lk(c.mu);
  c.err.msg.clear();
  *n_edges = 0;
  if (n <= 0 || n_recv < 0 || n_recv > n || !(dist > 0) || n >= (int)UPH_ID_MASK) {
    c.err.msg = "uph_graph_build: invalid sizes";
    return -1;
  }
  if (!c.st) HIPCHK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking), c.err);
  if (!c.bbox) HIPCHK(hipMalloc((void**)&c.bbox, 4 * (1 + BBOX_BLOCKS) * sizeof(double)), c.err);
  // the caller may have written pos on any stream
  HIPCHK(hipDeviceSynchronize(), c.err);
  {
    const int nbb = std::max(1, std::min(BBOX_BLOCKS, (n + 255) / 256));
    hipLaunchKernelGGL(k_bbox, dim3(nbb), dim3(256), 0, c.st, n, (const double2*)pos, c.bbox + 4);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, c.st, nbb, c.bbox + 4, c.bbox);
  }
  double bb[4];
  HIPCHK(hipMemcpyAsync(bb, c.bbox, sizeof bb, hipMemcpyDeviceToHost, c.st), c.err);
  HIPCHK(hipStreamSynchronize(c.st), c.err);
  if (!std::isfinite(bb[0]) || !std::isfinite(bb[1]) || !std::isfinite(bb[2]) || !std::isfinite(bb[3])) {
    c.err.msg = "uph_graph_build: non-finite position";
    return -1;
  }
  DevScene S;
  memset(&S, 0, sizeof S);
  S.cutoff = dist;
  double cell = sqrt(dist) * (1.0 + 1e-9);
  double ncx = floor((bb[2] - bb[0]) / cell) + 1, ncy = floor((bb[3] - bb[1]) / cell) + 1;
  while (ncx * ncy > 4.0 * n + 1024) {
    cell *= 2.0;
    ncx = floor((bb[2] - bb[0]) / cell) + 1;
    ncy = floor((bb[3] - bb[1]) / cell) + 1;
  }
  S.cell_inv = 1.0 / cell;
  S.xmin = bb[0];
  S.ymin = bb[1];
  S.ncx = (int)ncx;
  S.ncy = (int)ncy;
  S.np = n;
  S.nf = n_recv;
  S.pb_gid = n;
  S.K = 1 << 30;
  const int ncells = S.ncx * S.ncy;
  const int nb = (ncells + SCAN_TILE - 1) / SCAN_TILE;
  if (graph_reserve(c, n, ncells, nb)) return -1;
  hipStream_t st = c.st;
  HIPCHK(hipMemcpyAsync(c.dsc, &S, sizeof S, hipMemcpyHostToDevice, st), c.err);
  hipLaunchKernelGGL(k_graph_tags, dim3(grid_of(n)), dim3(BLK), 0, st, n, n_recv, (const double2*)pos, c.b.posB,
                     c.b.tagA, c.b.sceneOf);
  if (launch_sort(st, n, ncells, nb, c.b, c.dsc, 1, false, c.err)) return -1;
  const int qnb = (n + SCAN_TILE - 1) / SCAN_TILE;
  // receivers
  hipLaunchKernelGGL(k_graph_count, dim3(grid8_of(n)), dim3(BLK), 0, st, n, c.b, c.dsc, n_recv, c.cnt, 0);
  launch_scan(st, n, qnb, c.cnt, c.qpartial, off);
  int ne = 0;
  HIPCHK(hipMemcpyAsync(&ne, off + n, sizeof(int), hipMemcpyDeviceToHost, st), c.err);
  HIPCHK(hipStreamSynchronize(st), c.err);
  *n_edges = ne;
  if ((long)ne > capacity) {
    c.err.msg = "uph_graph_build: more edges than the capacity";
    return -2;
  }
  hipLaunchKernelGGL(k_graph_emit, dim3(grid8_of(n)), dim3(BLK), 0, st, n, c.b, c.dsc, n_recv, off, send, recv);
  // senders
  if (soff && sperm) {
    hipLaunchKernelGGL(k_graph_count, dim3(grid8_of(n)), dim3(BLK), 0, st, n, c.b, c.dsc, n_recv, c.cnt, 1);
    launch_scan(st, n, qnb, c.cnt, c.qpartial, soff);
    hipLaunchKernelGGL(k_graph_sender_perm, dim3(grid8_of(n)), dim3(BLK), 0, st, n, c.b, c.dsc, n_recv, off,
                       send, soff, sperm);
  }
  HIPCHK(hipGetLastError(), c.err);
  HIPCHK(hipStreamSynchronize(st), c.err);
  return 0;
}
