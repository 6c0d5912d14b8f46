// uph_search.hip -- SearchNeighbors (Search.cc:5-44) on the device: the kernels, the
// allocating entry points and the allocation-free one.
#include "uph_internal.h"

// ---------------------------------------------------------------- search

// Per query: the K-th key (when more than K are in range) and the number of
// pairs it emits.  A lane counts its own query in one pass; queries with more
// than K in range are left to the whole wavefront.
__global__ __launch_bounds__(BLK) void k_search_count(int nf_slots_unused, int n, StepBufs b, const DevScene* __restrict__ sc,
                                                      int* __restrict__ qcount) {
  __shared__ KthLds lds[BLK / 64];
  const int p = xcd_block() * BLK + threadIdx.x;
  const DevScene& S = sc[0];
  bool capped = false;
  if (p < n && tag_type(b.tagA[p]) == 1) {
    const uint32_t idp = tag_id(b.tagA[p]);
    const double2 pp = b.posA[p];
    int c = 0, m = 0;
    for_candidates(S, pp, b.cellStart, [&](int q) {
      if (q == p) return;
      const double d2 = dist2(pp, b.posA[q]);
      if (!(d2 < S.cutoff)) return;
      ++c;
      const uint32_t idq = tag_id(b.tagA[q]);
      if (idq > idp && (int)idq < S.pb_gid) ++m;
    });
    if (c > S.K) {
      capped = true;
    } else {
      b.kth[p] = KthRec{INFINITY, 0xffffffffu, 0};
      qcount[idp] = m;
    }
  }
  unsigned long long w = __ballot(capped);
  const int lane = threadIdx.x & 63;
  KthLds& L = lds[threadIdx.x >> 6];
#pragma unroll 1
  while (w) {
    const int l = __ffsll((long long)w) - 1;
    w &= w - 1ull;
    const int q = bcast(p, l);
    const uint32_t idq = tag_id(b.tagA[q]);
    int nin;
    double kd2;
    uint32_t kid;
    wave_kth(S, q, b.posA, b.tagA, b.cellStart, L, nin, kd2, kid, true);
    int m = 0;
    if (nin >= 0 && nin <= KTH_CAP) {
      // the listed keys this lane owns that the query keeps
      for (int j = lane; j < nin; j += 64) m += search_keeps(S, idq, L.d2[j], L.id[j], kd2, kid) ? 1 : 0;
      for (int o = 32; o > 0; o >>= 1) m += __shfl_xor(m, o, 64);
    } else if (lane == 0) {
      const double2 pq = b.posA[q];
      for_candidates(S, pq, b.cellStart, [&](int u) {
        if (u == q) return;
        const double d2 = dist2(pq, b.posA[u]);
        if (d2 < S.cutoff && search_keeps(S, idq, d2, tag_id(b.tagA[u]), kd2, kid)) ++m;
      });
    }
    if (lane == 0) {
      b.kth[q] = KthRec{kd2, kid, 0};
      qcount[idq] = m;
    }
    wave_lds_sync();
  }
}

// Writes the pairs of query p in (d2, id) order from qoff[id_p].  A lane writes
// its own query's pairs when they are few (each pair is ranked by a pass over
// the candidates); longer lists are ranked by the whole wavefront.
#define EMIT_LANE_MAX 16

__global__ __launch_bounds__(BLK) void k_search_emit(int n, StepBufs b, const DevScene* __restrict__ sc,
                                                     const int* __restrict__ qoff, int* __restrict__ ip,
                                                     int* __restrict__ jp, double* __restrict__ r2) {
  __shared__ KthLds lds[BLK / 64];
  const int p = xcd_block() * BLK + threadIdx.x;
  const DevScene& S = sc[0];
  bool wide = false;
  if (p < n && tag_type(b.tagA[p]) == 1) {
    const uint32_t idp = tag_id(b.tagA[p]);
    const int o = qoff[idp];
    if (qoff[idp + 1] - o > EMIT_LANE_MAX) {
      wide = true;
    } else {
      const double kd2 = b.kth[p].d2;
      const uint32_t kid = b.kth[p].id;
      const double2 pp = b.posA[p];
      for_candidates(S, pp, b.cellStart, [&](int q) {
        if (q == p) return;
        const double d2 = dist2(pp, b.posA[q]);
        if (!(d2 < S.cutoff)) return;
        const uint32_t idq = tag_id(b.tagA[q]);
        if (!search_keeps(S, idp, d2, idq, kd2, kid)) return;
        // rank among the kept candidates by (d2, id)
        int rnk = 0;
        for_candidates(S, pp, b.cellStart, [&](int u) {
          if (u == p) return;
          const double e2 = dist2(pp, b.posA[u]);
          if (!(e2 < S.cutoff)) return;
          const uint32_t idu = tag_id(b.tagA[u]);
          if (search_keeps(S, idp, e2, idu, kd2, kid) && key_lt(e2, idu, d2, idq)) ++rnk;
        });
        ip[o + rnk] = (int)idp;
        jp[o + rnk] = (int)idq;
        r2[o + rnk] = d2;
      });
    }
  }
  unsigned long long w = __ballot(wide);
  const int lane = threadIdx.x & 63;
  KthLds& L = lds[threadIdx.x >> 6];
#pragma unroll 1
  while (w) {
    const int l = __ffsll((long long)w) - 1;
    w &= w - 1ull;
    const int q = bcast(p, l);
    const uint32_t idq = tag_id(b.tagA[q]);
    const double kd2 = b.kth[q].d2;
    const uint32_t kid = b.kth[q].id;
    const int o = qoff[idq];
    const int nin = wave_gather(S, q, b.posA, b.tagA, b.cellStart, L);
    if (nin >= 0 && nin <= KTH_CAP) {
      // each kept key this lane owns, ranked among the kept keys
      for (int j0 = 0; j0 < nin; j0 += 64) {
        const int j = j0 + lane;
        const bool own = j < nin;
        const double dj = own ? L.d2[j] : 0.0;
        const uint32_t jd = own ? L.id[j] : 0u;
        const bool keep = own && search_keeps(S, idq, dj, jd, kd2, kid);
        int rnk = 0;
#pragma unroll 2
        for (int i = 0; i < nin; ++i) {
          const double di = L.d2[i];
          const uint32_t id = L.id[i];
          rnk += (search_keeps(S, idq, di, id, kd2, kid) && key_lt(di, id, dj, jd)) ? 1 : 0;
        }
        if (keep) {
          ip[o + rnk] = (int)idq;
          jp[o + rnk] = (int)jd;
          r2[o + rnk] = dj;
        }
      }
    } else if (lane == 0) {
      const double2 pq = b.posA[q];
      for_candidates(S, pq, b.cellStart, [&](int u) {
        if (u == q) return;
        const double d2 = dist2(pq, b.posA[u]);
        if (!(d2 < S.cutoff)) return;
        const uint32_t idu = tag_id(b.tagA[u]);
        if (!search_keeps(S, idq, d2, idu, kd2, kid)) return;
        int rnk = 0;
        for_candidates(S, pq, b.cellStart, [&](int v) {
          if (v == q) return;
          const double e2 = dist2(pq, b.posA[v]);
          if (!(e2 < S.cutoff)) return;
          const uint32_t idv = tag_id(b.tagA[v]);
          if (search_keeps(S, idq, e2, idv, kd2, kid) && key_lt(e2, idv, d2, idu)) ++rnk;
        });
        ip[o + rnk] = (int)idq;
        jp[o + rnk] = (int)idu;
        r2[o + rnk] = d2;
      });
    }
    wave_lds_sync();
  }
}

// ------------------------------------------------------ SearchNeighbors

namespace {
struct SearchCtx {
  std::mutex mu;
  Err err;
  hipStream_t st = nullptr;
  int cap_n = 0, cap_cells = 0, cap_nb = 0;
  long cap_pairs = 0;
  StepBufs b{};
  DevScene* dsc = nullptr;
  int *qcount = nullptr, *qoff = nullptr, *qpartial = nullptr;
  int *ip = nullptr, *jp = nullptr;
  double* r2 = nullptr;
  double2* hpos = nullptr;  // pinned staging
};
SearchCtx g_search;
// the last search call was uph_search_neighbors_noalloc: uph_search_last_error reports its message
std::atomic<bool> g_search_err_noalloc{false};

int search_reserve(SearchCtx& c, int n, int ncells, int nb) {
  if (!c.st) HIPCHK(hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking), c.err);
  if (!c.dsc) HIPCHK(hipMalloc((void**)&c.dsc, sizeof(DevScene)), c.err);
  if (n > c.cap_n || ncells > c.cap_cells || nb > c.cap_nb) {
    free_bufs(c.b);
    if (c.qcount) (void)hipFree(c.qcount);
    if (c.qoff) (void)hipFree(c.qoff);
    if (c.qpartial) (void)hipFree(c.qpartial);
    if (c.hpos) (void)hipHostFree(c.hpos);
    c.qcount = c.qoff = c.qpartial = nullptr;
    c.hpos = nullptr;
    const int nn = std::max(n, c.cap_n), nc = std::max(ncells, c.cap_cells);
    const int nbb = std::max(std::max(nb, c.cap_nb), (nn + SCAN_TILE - 1) / SCAN_TILE + 1);
    if (alloc_bufs(c.b, nn, nc, nbb, c.err, true)) return -1;
    HIPCHK(hipMemset(c.b.cellCount, 0, (size_t)nc * sizeof(int)), c.err);
    HIPCHK(hipMemset(c.b.partial, 0, (size_t)nbb * sizeof(int)), c.err);
    HIPCHK(hipMalloc((void**)&c.qcount, ((size_t)nn + 1) * sizeof(int)), c.err);
    HIPCHK(hipMalloc((void**)&c.qoff, ((size_t)nn + 1) * sizeof(int)), c.err);
    HIPCHK(hipMalloc((void**)&c.qpartial, (size_t)nbb * sizeof(int)), c.err);
    HIPCHK(hipHostMalloc((void**)&c.hpos, (size_t)nn * sizeof(double2), hipHostMallocDefault), c.err);
    c.cap_n = nn;
    c.cap_cells = nc;
    c.cap_nb = nbb;
  }
  return 0;
}

// The search's scene: one grid over the bounding box of all particles, at most
// 4 n + 1024 cells.  Returns false on a non-finite position.
bool search_grid(const double* const* rows, const double* flat, int nf, int n, int pb, int K, double dist,
                 DevScene& S) {
  double lx = INFINITY, ly = INFINITY, hx = -INFINITY, hy = -INFINITY;
  for (int i = 0; i < n; ++i) {
    const double x = rows ? rows[i][0] : flat[2 * i];
    const double y = rows ? rows[i][1] : flat[2 * i + 1];
    if (!std::isfinite(x) || !std::isfinite(y)) return false;
    lx = std::min(lx, x); hx = std::max(hx, x);
    ly = std::min(ly, y); hy = std::max(hy, y);
  }
  memset(&S, 0, sizeof S);
  S.cutoff = dist;
  const double cs = sqrt(dist) * (1.0 + 1e-9);
  S.cell_inv = 1.0 / cs;
  S.xmin = lx;
  S.ymin = ly;
  double ncx = floor((hx - lx) * S.cell_inv) + 1, ncy = floor((hy - ly) * S.cell_inv) + 1;
  // bound the grid: coarser cells keep the result exact, only slower
  double cell = cs;
  while (ncx * ncy > 4.0 * n + 1024) {
    cell *= 2.0;
    S.cell_inv = 1.0 / cell;
    ncx = floor((hx - lx) * S.cell_inv) + 1;
    ncy = floor((hy - ly) * S.cell_inv) + 1;
  }
  // a coarser grid must still hold every in-range pair inside a 3x3 stencil: cell >= sqrt(dist)
  S.ncx = (int)ncx;
  S.ncy = (int)ncy;
  S.coff = 0;
  S.poff = 0;
  S.np = n;
  S.nf = nf;
  S.pb_gid = pb;
  S.K = K;
  return true;
}

int search_run(const double* const* rows, const double* flat, int* i_pair, int* j_pair, double* rij2, int nf, int n,
               int pb, int K, double dist, long capacity) {
  SearchCtx& c = g_search;
  std::lock_guard<std::mutex> lk(c.mu);
  g_search_err_noalloc = false;
  c.err.msg.clear();
  if (nf <= 0 || n <= 0) return 0;
  if (nf > n || K <= 0 || !(dist > 0)) {
    c.err.msg = "SearchNeighbors: invalid sizes";
    return -1;
  }
  if (n >= (int)UPH_ID_MASK) {
    c.err.msg = "SearchNeighbors: too many particles";
    return -1;
  }
  DevScene S;
  if (!search_grid(rows, flat, nf, n, pb, K, dist, S)) {
    c.err.msg = "SearchNeighbors: non-finite position";
    return -1;
  }
  const int ncells = S.ncx * S.ncy;
  const int nb = (ncells + SCAN_TILE - 1) / SCAN_TILE;
  if (search_reserve(c, n, ncells, nb)) return -1;
  for (int i = 0; i < n; ++i)
    c.hpos[i] = rows ? make_double2(rows[i][0], rows[i][1]) : make_double2(flat[2 * i], flat[2 * i + 1]);
  std::vector<uint32_t> tag(n);
  for (int i = 0; i < n; ++i) tag[i] = ((uint32_t)(i < nf ? 1 : 2) << UPH_ID_BITS) | (uint32_t)i;
  hipStream_t st = c.st;
  HIPCHK(hipMemcpyAsync(c.b.posB, c.hpos, n * sizeof(double2), hipMemcpyHostToDevice, st), c.err);
  HIPCHK(hipMemcpyAsync(c.b.tagA, tag.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st), c.err);
  HIPCHK(hipMemsetAsync(c.b.sceneOf, 0, n * sizeof(int), st), c.err);
  HIPCHK(hipMemcpyAsync(c.dsc, &S, sizeof S, hipMemcpyHostToDevice, st), c.err);
  if (launch_sort(st, n, ncells, nb, c.b, c.dsc, 1, false, c.err)) return -1;
  hipLaunchKernelGGL(k_search_count, dim3(grid8_of(n)), dim3(BLK), 0, st, nf, n, c.b, c.dsc, c.qcount);
  // exclusive scan of per-query counts (nf entries) into qoff
  const int qnb = (nf + SCAN_TILE - 1) / SCAN_TILE;
  launch_scan(st, nf, qnb, c.qcount, c.qpartial, c.qoff);
  HIPCHK(hipGetLastError(), c.err);
  int niac = 0;
  HIPCHK(hipMemcpyAsync(&niac, c.qoff + nf, sizeof(int), hipMemcpyDeviceToHost, st), c.err);
  HIPCHK(hipStreamSynchronize(st), c.err);
  if (niac < 0) {
    c.err.msg = "SearchNeighbors: pair count overflow";
    return -1;
  }
  if ((long)niac > capacity) {
    c.err.msg = "SearchNeighbors: more pairs than the caller's capacity";
    return -2;
  }
  if (niac > c.cap_pairs) {
    if (c.ip) (void)hipFree(c.ip);
    if (c.jp) (void)hipFree(c.jp);
    if (c.r2) (void)hipFree(c.r2);
    c.ip = c.jp = nullptr;
    c.r2 = nullptr;
    c.cap_pairs = 0;
    HIPCHK(hipMalloc((void**)&c.ip, (size_t)niac * sizeof(int)), c.err);
    HIPCHK(hipMalloc((void**)&c.jp, (size_t)niac * sizeof(int)), c.err);
    HIPCHK(hipMalloc((void**)&c.r2, (size_t)niac * sizeof(double)), c.err);
    c.cap_pairs = niac;
  }
  if (niac > 0) {
    hipLaunchKernelGGL(k_search_emit, dim3(grid8_of(n)), dim3(BLK), 0, st, n, c.b, c.dsc, c.qoff, c.ip, c.jp, c.r2);
    HIPCHK(hipGetLastError(), c.err);
    HIPCHK(hipMemcpyAsync(i_pair, c.ip, (size_t)niac * sizeof(int), hipMemcpyDeviceToHost, st), c.err);
    HIPCHK(hipMemcpyAsync(j_pair, c.jp, (size_t)niac * sizeof(int), hipMemcpyDeviceToHost, st), c.err);
    HIPCHK(hipMemcpyAsync(rij2, c.r2, (size_t)niac * sizeof(double), hipMemcpyDeviceToHost, st), c.err);
    HIPCHK(hipStreamSynchronize(st), c.err);
  }
  return niac;
}
}  // namespace

extern "C" int SearchNeighbors(const double** Position, int* i_pair, int* j_pair, double* rij2,
                               const int NumberOfFluidParticles, const int NumberOfParticles, int ParticleBoundary,
                               const int NumberOfNeighbors, const double dist) {
  return search_run(Position, nullptr, i_pair, j_pair, rij2, NumberOfFluidParticles, NumberOfParticles,
                    ParticleBoundary, NumberOfNeighbors, dist, 0x7fffffffffffffffL);
}

extern "C" int uph_search_neighbors_cap(const double* pos_xy, int* i_pair, int* j_pair, double* rij2, int n_fluid,
                                        int n_particles, int particle_boundary, int n_neighbors, double dist,
                                        long capacity) {
  return search_run(nullptr, pos_xy, i_pair, j_pair, rij2, n_fluid, n_particles, particle_boundary, n_neighbors, dist,
                    capacity);
}

// ------------------------------------- SearchNeighbors without allocation
//
// uph_search_neighbors_noalloc: the search of uph_search_neighbors_cap with every
// device buffer (sort arrays, scene table, per-query counts, pair staging) carved out
// of one file-scope device array of fixed size, so a call allocates nothing.  The
// array's device address is looked up once per device with hipGetSymbolAddress and
// only that address is handed to kernels and copies.

namespace {

#define NOALLOC_MAX_PARTICLES 65536
#define NOALLOC_MAX_CELLS (4 * NOALLOC_MAX_PARTICLES + 1024)   // the search grid's bound
#define NOALLOC_MAX_PAIRS 1572864
#define NOALLOC_MAX_DEVICES 64

constexpr size_t na_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte offsets of the buffers inside the scratch array.
struct NaLayout {
  size_t dsc, posA, posB, tagA, tagB, keys, rank, tmpIdx, sceneOf, kth, cellStart, cellCount, partial, ctl, qcount,
      qoff, qpartial, ip, jp, r2, total;
};

constexpr NaLayout na_layout() {
  const size_t P = NOALLOC_MAX_PARTICLES, C = NOALLOC_MAX_CELLS;
  const size_t nb1 = (C + SCAN_TILE - 1) / SCAN_TILE, nb2 = (P + SCAN_TILE - 1) / SCAN_TILE + 1;
  const size_t NB = nb1 > nb2 ? nb1 : nb2;
  NaLayout L{};
  size_t o = 0;
  L.dsc = o; o += na_up(sizeof(DevScene));
  L.posA = o; o += na_up((P + POS_PAD) * sizeof(double2));
  L.posB = o; o += na_up(P * sizeof(double2));
  L.tagA = o; o += na_up(P * sizeof(uint32_t));
  L.tagB = o; o += na_up(P * sizeof(uint32_t));
  L.keys = o; o += na_up(P * sizeof(int));
  L.rank = o; o += na_up(P * sizeof(int));
  L.tmpIdx = o; o += na_up(P * sizeof(int));
  L.sceneOf = o; o += na_up(P * sizeof(int));
  L.kth = o; o += na_up(P * sizeof(KthRec));
  L.cellStart = o; o += na_up((C + 1) * sizeof(int));
  L.cellCount = o; o += na_up(C * sizeof(int));
  L.partial = o; o += na_up(NB * sizeof(int));
  L.ctl = o; o += na_up(CTL_WORDS * sizeof(int));
  L.qcount = o; o += na_up((P + 1) * sizeof(int));
  L.qoff = o; o += na_up((P + 1) * sizeof(int));
  L.qpartial = o; o += na_up(NB * sizeof(int));
  L.ip = o; o += na_up((size_t)NOALLOC_MAX_PAIRS * sizeof(int));
  L.jp = o; o += na_up((size_t)NOALLOC_MAX_PAIRS * sizeof(int));
  L.r2 = o; o += na_up((size_t)NOALLOC_MAX_PAIRS * sizeof(double));
  L.total = o;
  return L;
}

constexpr NaLayout NA = na_layout();

}  // namespace

// 32.5 MB; every device the library runs on holds its own copy until the process ends.
static __device__ __align__(256) unsigned char scratch[NA.total];

namespace {

std::mutex g_na_mu;
unsigned char* g_na_ptr[NOALLOC_MAX_DEVICES] = {};   // scratch's device address, per device
char g_na_err[256] = "";

// The buffers' zero-start state of one call: input tags (fluid below nf), scene 0 for
// every particle, empty cells, no pairs counted.
__global__ void k_na_clear(int n, int nf, int ncells, uint32_t* __restrict__ tag, int* __restrict__ sceneOf,
                           int* __restrict__ cellCount, int* __restrict__ qcount) {
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ncells || i <= nf || i < n; i += stride) {
    if (i < n) {
      tag[i] = ((uint32_t)(i < nf ? 1 : 2) << UPH_ID_BITS) | (uint32_t)i;
      sceneOf[i] = 0;
    }
    if (i < ncells) cellCount[i] = 0;
    if (i <= nf) qcount[i] = 0;
  }
}

int na_fail(int rc, const char* msg) {
  snprintf(g_na_err, sizeof g_na_err, "uph_search_neighbors_noalloc: %s", msg);
  return rc;
}

#define NACHK(x)                                              \
  do {                                                        \
    hipError_t e_ = (x);                                      \
    if (e_ != hipSuccess) return na_fail(-1, hipGetErrorString(e_)); \
  } while (0)

}  // namespace

extern "C" int uph_search_neighbors_noalloc(const double* pos_xy, int* i_pair, int* j_pair, double* rij2,
                                            int n_fluid, int n_particles, int particle_boundary, int n_neighbors,
                                            double dist, long capacity) {
  std::lock_guard<std::mutex> lk(g_na_mu);
  g_search_err_noalloc = true;
  g_na_err[0] = 0;
  const int nf = n_fluid, n = n_particles;
  if (nf <= 0 || n <= 0) return 0;
  if (nf > n || n_neighbors <= 0 || !(dist > 0)) return na_fail(-1, "invalid sizes");
  if (n > NOALLOC_MAX_PARTICLES) {
    char m[160];
    snprintf(m, sizeof m, "%d particles exceed this entry point's fixed maximum of %d", n, NOALLOC_MAX_PARTICLES);
    return na_fail(-3, m);
  }
  DevScene S;
  if (!search_grid(nullptr, pos_xy, nf, n, particle_boundary, n_neighbors, dist, S))
    return na_fail(-1, "non-finite position");
  const int ncells = S.ncx * S.ncy;
  if (ncells > NOALLOC_MAX_CELLS) return na_fail(-3, "search grid beyond the fixed maximum of cells");
  const int nb = (ncells + SCAN_TILE - 1) / SCAN_TILE;
  // set-up, once per device: the device address of this device's copy of scratch
  int dev = 0;
  NACHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return na_fail(-1, "device index beyond the supported range");
  if (!g_na_ptr[dev]) {
    void* ptr = nullptr;
    NACHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_na_ptr[dev] = static_cast<unsigned char*>(ptr);
  }
  unsigned char* base = g_na_ptr[dev];
  StepBufs b;
  memset(&b, 0, sizeof b);
  b.posA = reinterpret_cast<double2*>(base + NA.posA);
  b.posB = reinterpret_cast<double2*>(base + NA.posB);
  b.tagA = reinterpret_cast<uint32_t*>(base + NA.tagA);
  b.tagB = reinterpret_cast<uint32_t*>(base + NA.tagB);
  b.keys = reinterpret_cast<int*>(base + NA.keys);
  b.rank = reinterpret_cast<int*>(base + NA.rank);
  b.tmpIdx = reinterpret_cast<int*>(base + NA.tmpIdx);
  b.sceneOf = reinterpret_cast<int*>(base + NA.sceneOf);
  b.kth = reinterpret_cast<KthRec*>(base + NA.kth);
  b.cellStart = reinterpret_cast<int*>(base + NA.cellStart);
  b.cellCount = reinterpret_cast<int*>(base + NA.cellCount);
  b.partial = reinterpret_cast<int*>(base + NA.partial);
  b.ctl = reinterpret_cast<int*>(base + NA.ctl);
  b.it = b.ctl + C_IT;
  DevScene* dsc = reinterpret_cast<DevScene*>(base + NA.dsc);
  int* qcount = reinterpret_cast<int*>(base + NA.qcount);
  int* qoff = reinterpret_cast<int*>(base + NA.qoff);
  int* qpartial = reinterpret_cast<int*>(base + NA.qpartial);
  int* ip = reinterpret_cast<int*>(base + NA.ip);
  int* jp = reinterpret_cast<int*>(base + NA.jp);
  double* r2 = reinterpret_cast<double*>(base + NA.r2);
  hipStream_t st = nullptr;   // the null stream: the copies below are synchronous
  NACHK(hipMemcpy(dsc, &S, sizeof S, hipMemcpyHostToDevice));
  NACHK(hipMemcpy(b.posB, pos_xy, (size_t)n * sizeof(double2), hipMemcpyHostToDevice));
  const int span = std::max(std::max(ncells, nf + 1), n);
  hipLaunchKernelGGL(k_na_clear, dim3(std::min(grid_of(span), 1024)), dim3(BLK), 0, st, n, nf, ncells, b.tagA,
                     b.sceneOf, b.cellCount, qcount);
  Err err;
  if (launch_sort(st, n, ncells, nb, b, dsc, 1, false, err)) return na_fail(-1, "sort launch failed");
  hipLaunchKernelGGL(k_search_count, dim3(grid8_of(n)), dim3(BLK), 0, st, nf, n, b, dsc, qcount);
  const int qnb = (nf + SCAN_TILE - 1) / SCAN_TILE;
  launch_scan(st, nf, qnb, qcount, qpartial, qoff);
  NACHK(hipGetLastError());
  int niac = 0;
  NACHK(hipMemcpy(&niac, qoff + nf, sizeof(int), hipMemcpyDeviceToHost));
  if (niac < 0) return na_fail(-1, "pair count overflow");
  if ((long)niac > capacity) return na_fail(-2, "more pairs than the caller's capacity");
  if (niac > NOALLOC_MAX_PAIRS) {
    char m[160];
    snprintf(m, sizeof m, "%d pairs exceed this entry point's fixed maximum of %d", niac, NOALLOC_MAX_PAIRS);
    return na_fail(-3, m);
  }
  if (niac > 0) {
    hipLaunchKernelGGL(k_search_emit, dim3(grid8_of(n)), dim3(BLK), 0, st, n, b, dsc, qoff, ip, jp, r2);
    NACHK(hipGetLastError());
    NACHK(hipMemcpy(i_pair, ip, (size_t)niac * sizeof(int), hipMemcpyDeviceToHost));
    NACHK(hipMemcpy(j_pair, jp, (size_t)niac * sizeof(int), hipMemcpyDeviceToHost));
    NACHK(hipMemcpy(rij2, r2, (size_t)niac * sizeof(double), hipMemcpyDeviceToHost));
  }
  return niac;
}

extern "C" const char* uph_search_last_error(void) {
  return g_search_err_noalloc ? g_na_err : g_search.err.msg.c_str();
}
