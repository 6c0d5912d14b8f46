// uph_step.hip -- the step's kernels (sort, reorder, mirror rows, pair passes, the
// persistent form) and the handle API that launches them.
#include "uph_internal.h"

// ------------------------------------------------------------------ sort

__global__ __launch_bounds__(BLK) void k_bin(int n, const double2* __restrict__ pos, const int* __restrict__ sceneOf,
                                             const DevScene* __restrict__ sc, const uint32_t* __restrict__ tagA,
                                             uint32_t* __restrict__ tagB, int* __restrict__ keys,
                                             int* __restrict__ rank, int* __restrict__ count) {
  bin_block(blockIdx.x, n, pos, sceneOf, sc, tagA, tagB, keys, rank, count);
}

// With `gate` (the first step of a lazy handle's call), the scan, scatter and bin
// kernels return at once unless the last call held a re-sort over (*gate != 0).
__global__ __launch_bounds__(SCAN_BLK) void k_scan_reduce(int m, const int* __restrict__ in, int* __restrict__ partial,
                                                          int* __restrict__ step_ctl, int* __restrict__ sstale,
                                                          int nscenes, int view, const int* __restrict__ gate) {
  __shared__ int lds[SCAN_BLK / 64 + 1];
  if (gate && *gate == 0) return;
  if (step_ctl && blockIdx.x == 0 && threadIdx.x == 0) open_step(step_ctl, sstale, nscenes, view);
  scan_reduce_tile(blockIdx.x, m, in, partial, lds);
}

__global__ __launch_bounds__(SCAN_BLK) void k_scan_down(int m, int* __restrict__ count, const int* __restrict__ partial,
                                                        int* __restrict__ start, const int* __restrict__ gate) {
  __shared__ int tile[SCAN_TILE + SCAN_TILE / 16];
  __shared__ int lds[SCAN_BLK / 64 + 1];
  if (gate && *gate == 0) return;
  scan_down_tile(blockIdx.x, m, count, partial, start, tile, lds);
}

__global__ __launch_bounds__(BLK) void k_scatter(int n, const int* __restrict__ keys, const int* __restrict__ rank,
                                                 const int* __restrict__ cellStart, int* __restrict__ tmpIdx,
                                                 const int* __restrict__ gate) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n || (gate && *gate == 0)) return;
  tmpIdx[cellStart[keys[i]] + rank[i]] = i;
}

template <bool MIRRORS>
__global__ __launch_bounds__(BLK) void k_reorder(int n, StepBufs b, const DevScene* __restrict__ sc, int search_only) {
  reorder_slot<MIRRORS>(blockIdx.x * BLK + threadIdx.x, n, b, sc, search_only);
}

// The re-sort a lazy handle's last call held over (C_PEND), at the first step of the
// next call: the stale scenes binned again, the others kept in their cells, the B set
// of the step's view gathered into its A set.
__global__ __launch_bounds__(BLK) void k_bin_held(int n, StepBufs b, const DevScene* __restrict__ sc) {
  if (b.ctl[C_PEND] == 0) return;
  const StepBufs v = lazy_view(b);
  // the holding step's flags: every scene after a capped pass, else the stale ones
  const int par = (b.ctl[C_IT] - 1) & 1;
  const int* keep = b.ctl[C_LOVF + par] != 0 ? nullptr : b.sstale + par * b.nscenes;
  bin_block(blockIdx.x, n, v.posB, b.sceneOf, sc, nullptr, nullptr, b.keys, b.rank, b.cellCount, keep, b.cellOf);
}

__global__ __launch_bounds__(BLK) void k_permute_held(int n, StepBufs b, const DevScene* __restrict__ sc, int ref) {
  if (b.ctl[C_PEND] == 0) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&b.ctl[C_NRES], 1);
  permute_slot(blockIdx.x * BLK + threadIdx.x, n, lazy_view(b), sc, ref != 0);
}

// Popcount of each word of section-test bits (the scan's input); the words move to
// the step's copy and are cleared for the next step's sort.
__global__ __launch_bounds__(BLK) void k_ghost_count(StepBufs b) {
  const int w = blockIdx.x * BLK + threadIdx.x;
  if (w >= b.wtot) return;
  const uint32_t v = b.gbits[w];
  b.gbits2[w] = v;
  b.gbits[w] = 0u;
  b.gwcnt[w] = __popc(v);
}

// The listed rows' mirror rows (a grid-stride walk over the list).
__global__ __launch_bounds__(BLK) void k_ghost_emit(StepBufs b, const DevScene* __restrict__ sc) {
  const int nl = *b.gcandN;
  for (int c = blockIdx.x * BLK + threadIdx.x; c < nl; c += gridDim.x * BLK) emit_mirrors(b, sc, b.gcand[c]);
}

// Cell and in-cell rank of each mirror row (order within a cell is free: the rows
// only enter counts and key ranks).
__global__ __launch_bounds__(BLK) void k_ghost_bin(StepBufs b, const DevScene* __restrict__ sc) {
  const int j = blockIdx.x * BLK + threadIdx.x;
  if (j == 0) *b.gcandN = 0;   // the list is read; the next sort starts it again
  if (j >= b.gtot || j >= b.gwstart[b.wtot]) return;
  const DevScene& S = sc[b.gscene[j]];
  const double2 g = b.gpos[j];
  int cx, cy;
  cell_of(S, g.x, g.y, cx, cy);
  const int key = S.coff + cy * S.ncx + cx;
  b.gkey[j] = key;
  b.grank[j] = atomicAdd(&b.gcellCount[key], 1);
}

__global__ __launch_bounds__(BLK) void k_ghost_scatter(StepBufs b) {
  const int j = blockIdx.x * BLK + threadIdx.x;
  if (j >= b.gtot || j >= b.gwstart[b.wtot]) return;
  const int t = b.gcellStart[b.gkey[j]] + b.grank[j];
  b.gposS[t] = b.gpos[j];
  b.gidS[t] = b.gid[j];
}

// ------------------------------------------------------------ persistent step
//
// k_persist runs n_steps whole steps in one cooperative launch: the phases of the
// launched step (bin, scan totals, scan, scatter, reorder, compacted pair pass with the
// mixed-scene wavefronts, capped pass) as grid-stride loops over the same blocks the
// separate kernels would take, separated by grid-wide barriers (7 per step).  For
// handles whose queries are not expected to be capped and that have no mirror rows; a
// capped step (a query with more than K in range, rare there) takes the direct capped
// loop, which sums in the same slot order as the bit walk of k_pair<true>.

// bar (BAR_WORDS words): the arrivals of each XCD's blocks (bar[BAR_GROUP + 16 x], one
// cache line each), the arrivals of the XCD groups (bar[BAR_TOP]), the generation
// (bar[BAR_GEN]) and the broken flag (bar[BAR_BROKEN]).  Two levels, so no counter takes
// more than G / 8 atomics per barrier (one counter for every block serialised about
// 80 ns per arrival: 85 us a barrier with 1024 blocks).  A block that polls the
// generation PERSIST_SPIN times without a change gives up, sets the broken flag and goes
// on; every later phase then skips its work, so the grid drains and the host reports
// the failure.
#define PERSIST_SPIN (1u << 24)
#define BAR_GROUP 0
#define BAR_TOP 128
#define BAR_GEN 144
#define BAR_BROKEN 160
#define BAR_WORDS 176

__device__ __forceinline__ bool grid_sync(unsigned* bar, unsigned nblocks) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    int good = 1;
    const unsigned g = __hip_atomic_load(&bar[BAR_GEN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned x = blockIdx.x & 7;
    bool bumped = false;
    if (atomicAdd(&bar[BAR_GROUP + 16 * x], 1u) == (nblocks >> 3) - 1) {   // the last of its XCD group
      atomicExch(&bar[BAR_GROUP + 16 * x], 0u);
      __threadfence();                        // the reset lands before the generation can move
      if (atomicAdd(&bar[BAR_TOP], 1u) == 7u) {                             // the last group
        atomicExch(&bar[BAR_TOP], 0u);
        __threadfence();
        atomicAdd(&bar[BAR_GEN], 1u);
        bumped = true;
      }
    }
    if (!bumped) {
      unsigned spins = 0;
      while (__hip_atomic_load(&bar[BAR_GEN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (++spins > PERSIST_SPIN) {
          atomicOr(&bar[BAR_BROKEN], 1u);
          good = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __threadfence();
    if (__hip_atomic_load(&bar[BAR_BROKEN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) good = 0;
    ok = good;
  }
  __syncthreads();
  return ok != 0;
}

// The virtual blocks [0, V) of a phase taken by this block: each XCD (blocks are dealt
// to the 8 XCDs round robin; the grid is a multiple of 8) takes one contiguous run, as
// xcd_block() arranges for the separate kernels.  The trip count is uniform per block.
template <class F>
__device__ __forceinline__ void for_vblocks(int V, F&& f) {
  const int per = gridDim.x >> 3;
  const int x = blockIdx.x & 7, li = blockIdx.x >> 3;
  const int chunk = (V + 7) >> 3;
  const int v0 = x * chunk, v1 = min(V, v0 + chunk);
  for (int v = v0 + li; v < v1; v += per) f(v);
}

struct PersistArgs {
  int n, ncells, nb, nmixed, n_steps;
  StepBufs b;
  const DevScene* sc;
  const int* mixed;
  unsigned* bar;
};

// The arguments sit in device memory and each phase reloads the fields it uses from an
// opaque copy of the pointer, so no phase keeps the others' pointers live (loop-
// invariant, the ~100 pointers would otherwise be hoisted and spill into VGPRs).
__device__ __forceinline__ const PersistArgs& phase_args(const PersistArgs* pa) {
  asm volatile("" : "+s"(pa));
  return *pa;
}

// Four wavefronts per SIMD (128 VGPRs): inside the loop over a phase's blocks the pair
// pass needs about 133 VGPRs (91 as a kernel of its own).
template <int KT, int MODE, bool ONE>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4))) void k_persist(
    const PersistArgs* __restrict__ pa) {
  __shared__ __align__(16) unsigned char sm[sizeof(KthLds) * (BLK / 64)];
  static_assert(sizeof(KthLds) * (BLK / 64) >= sizeof(int) * (SCAN_TILE + SCAN_TILE / 16 + SCAN_BLK / 64 + 1),
                "the scan tile fits the pair pass's LDS");
  int* tile = reinterpret_cast<int*>(sm);
  int* slds = tile + SCAN_TILE + SCAN_TILE / 16;
  KthLds* klds = reinterpret_cast<KthLds*>(sm);
  const unsigned G = gridDim.x;
  const int n_steps = pa->n_steps;
  unsigned* bar = pa->bar;
  for (int s = 0; s < n_steps; ++s) {
    {
      const PersistArgs& a = phase_args(pa);
      const StepBufs& b = a.b;
      const int nbl = (a.n + BLK - 1) / BLK;
      for_vblocks(nbl, [&](int v) {
        bin_block(v, a.n, b.posB, b.sceneOf, a.sc, b.tagA, b.tagB, b.keys, b.rank, b.cellCount);
      });
    }
    if (!grid_sync(bar, G)) break;
    {
      const PersistArgs& a = phase_args(pa);
      if (blockIdx.x == 0 && threadIdx.x == 0) open_step(a.b.ctl);
      for_vblocks(a.nb, [&](int t) { scan_reduce_tile(t, a.ncells, a.b.cellCount, a.b.partial, slds); });
    }
    if (!grid_sync(bar, G)) break;
    {
      const PersistArgs& a = phase_args(pa);
      for_vblocks(a.nb, [&](int t) {
        scan_down_tile(t, a.ncells, a.b.cellCount, a.b.partial, a.b.cellStart, tile, slds);
      });
    }
    if (!grid_sync(bar, G)) break;
    {
      const PersistArgs& a = phase_args(pa);
      const StepBufs& b = a.b;
      const int nbl = (a.n + BLK - 1) / BLK;
      for_vblocks(nbl, [&](int v) {
        const int i = v * BLK + threadIdx.x;
        if (i < a.n) b.tmpIdx[b.cellStart[b.keys[i]] + b.rank[i]] = i;
      });
    }
    if (!grid_sync(bar, G)) break;
    {
      const PersistArgs& a = phase_args(pa);
      const int nbl = (a.n + BLK - 1) / BLK;
      for_vblocks(nbl, [&](int v) { reorder_slot<false>(v * BLK + threadIdx.x, a.n, a.b, a.sc, 0); });
    }
    if (!grid_sync(bar, G)) break;
    {
      const PersistArgs& a = phase_args(pa);
      const int nbl = (a.n + BLK - 1) / BLK;
      for_vblocks(nbl, [&](int v) {
        compact_query<KT, MODE, true, ONE>(a.n, v * BLK + threadIdx.x, a.b, a.sc, klds[threadIdx.x >> 6]);
      });
      // the wavefronts holding two scenes' slots, one per wavefront with the scene per lane
      const int wid = blockIdx.x * (BLK / 64) + (threadIdx.x >> 6), nw = G * (BLK / 64);
      for (int w = wid; w < a.nmixed; w += nw)
        compact_query<KT, MODE, false>(a.n, a.mixed[w] * 64 + (int)(threadIdx.x & 63), a.b, a.sc,
                                       klds[threadIdx.x >> 6]);
    }
    if (!grid_sync(bar, G)) break;
    {
      const PersistArgs& a = phase_args(pa);
      const StepBufs& b = a.b;
      const int ovf = __hip_atomic_load(&b.ctl[C_OVF], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const int broken = __hip_atomic_load(&b.ctl[C_BROKEN], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (ovf != 0 && broken == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&b.ctl[C_NOVF], 1);
        const int nbl = (a.n + BLK - 1) / BLK;
        for_vblocks(nbl, [&](int v) {
          const int p = v * BLK + threadIdx.x;
          if (p < a.n) (void)pair_direct<true, KT, MODE>(p, b, a.sc);
        });
      }
    }
    if (!grid_sync(bar, G)) break;
  }
}

// ----------------------------------------------------------- lazy re-sort
//
// A lazy handle's step is its pair kernel and five short kernels that decide on the
// device whether there is work and return at once when there is none: k_post (the
// capped pass over every slot when a query had more than K in range, as k_pair<true>;
// then, for a re-sort, the binning: the stale scenes binned again and the others kept in
// their cells, or every scene when the capped pass ran, since its stale flags are not
// final until it ends), the two scan kernels, the scatter, and k_close (a plain gather of
// the step's B set into its A set, the pair epilogue having predicted the mirrors; and the
// next step's control words).  Without a re-sort the next step reads the B set as its A
// set (the view flips).  The last step of a call (hold) leaves a needed re-sort to the
// next call's first step (C_PEND), so after a call the A set still holds the last step's
// input (the state getters, approxVorticity).  None of these kernels waits on another
// block: the re-sort's phases are separated by kernel boundaries, whatever else runs on
// the GPU.  They read the step's counters from the copies the pair kernel left; k_close,
// the last, writes the next step's, and clears the next step's parity flags.

// Whether the step re-sorts: a query overflowed or a scene went stale, and the step is
// not the last of its call.
__device__ __forceinline__ bool lazy_resort(const int* ctl, int hold) {
  const int par = ctl[C_ITC] & 1;
  return !hold && (ctl[C_LOVF + par] != 0 || ctl[C_LSTALE + par] != 0);
}

template <int KT, int MODE, bool ONE>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4))) void k_post(int n, int hold, StepBufs b0,
                                                                                      const DevScene* __restrict__ sc) {
  int* ctl = b0.ctl;
  const int par = ctl[C_ITC] & 1;
  const bool ovf = ctl[C_LOVF + par] != 0;
  const bool capped = ovf && ctl[C_BROKEN] == 0;
  const bool resort = lazy_resort(ctl, hold);
  if (!capped && !resort) return;
  if (capped && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&ctl[C_NOVF], 1);
  StepBufs b = (ctl[C_VC] & 1) ? swapped(b0) : b0;
  b.it = ctl + C_ITC;   // {itime, stamp} of the step, as the pair kernel saw them
  const int* keep = ovf ? nullptr : b.sstale + par * b.nscenes;
  const int nbl = (n + BLK - 1) / BLK;
  // the trip count is the same for the whole block (bin_block works by wavefronts)
  for (int v = xcd_block(); v < nbl; v += gridDim.x) {
    const int p = v * BLK + threadIdx.x;
    if (capped && p < n) {
      const bool walked = tag_type(b.tagA[p]) == 1 && exact_pairs<KT, MODE, ONE, true>(p, b, sc);
      if (!walked) (void)pair_direct<true, KT, MODE, ONE, true>(p, b, sc);
    }
    // the slot's new position came from this thread (or the pair kernel)
    if (resort) bin_block(v, n, b.posB, b.sceneOf, sc, nullptr, nullptr, b.keys, b.rank, b.cellCount, keep, b.cellOf);
  }
}

__global__ __launch_bounds__(SCAN_BLK) void k_lscan_reduce(int m, const int* __restrict__ in,
                                                           int* __restrict__ partial, const int* __restrict__ ctl,
                                                           int hold) {
  __shared__ int lds[SCAN_BLK / 64 + 1];
  if (!lazy_resort(ctl, hold)) return;
  scan_reduce_tile(blockIdx.x, m, in, partial, lds);
}

__global__ __launch_bounds__(SCAN_BLK) void k_lscan_down(int m, int* __restrict__ count,
                                                         const int* __restrict__ partial, int* __restrict__ start,
                                                         const int* __restrict__ ctl, int hold) {
  __shared__ int tile[SCAN_TILE + SCAN_TILE / 16];
  __shared__ int lds[SCAN_BLK / 64 + 1];
  if (!lazy_resort(ctl, hold)) return;
  scan_down_tile(blockIdx.x, m, count, partial, start, tile, lds);
}

__global__ __launch_bounds__(BLK) void k_lscatter(int n, const int* __restrict__ keys, const int* __restrict__ rank,
                                                  const int* __restrict__ cellStart, int* __restrict__ tmpIdx,
                                                  const int* __restrict__ ctl, int hold) {
  if (!lazy_resort(ctl, hold)) return;
  for (int i = blockIdx.x * BLK + threadIdx.x; i < n; i += gridDim.x * BLK) tmpIdx[cellStart[keys[i]] + rank[i]] = i;
}

__global__ __launch_bounds__(BLK) void k_close(int n, int hold, StepBufs b0, const DevScene* __restrict__ sc) {
  int* ctl = b0.ctl;
  const int itime = ctl[C_ITC];
  const int par = itime & 1;
  const int vc = ctl[C_VC] & 1;
  const bool resort = lazy_resort(ctl, hold);
  if (resort) {
    const StepBufs b = vc ? swapped(b0) : b0;
    const bool ref = b.lazy == 2;
    for (int d = blockIdx.x * BLK + threadIdx.x; d < n; d += gridDim.x * BLK) permute_slot(d, n, b, sc, ref);
  }
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      const bool want = ctl[C_LOVF + par] != 0 || ctl[C_LSTALE + par] != 0;
      ctl[C_IT] = itime + 1;
      ctl[C_STAMP] = ctl[C_STC] + 1;
      ctl[C_V] = resort ? vc : vc ^ 1;
      ctl[C_PEND] = (want && hold) ? 1 : 0;
      ctl[C_LOVF + (par ^ 1)] = 0;
      ctl[C_LSTALE + (par ^ 1)] = 0;
      if (resort) ctl[C_NRES] += 1;
    }
    for (int s = threadIdx.x; s < b0.nscenes; s += BLK) b0.sstale[(par ^ 1) * b0.nscenes + s] = 0;
  }
}

void fill_scene(DevScene& S, const uph_scene& u, double cell_factor) {
  memset(&S, 0, sizeof S);
  S.cutoff = u.h * u.h * 4;
  // a cell is at least the search radius, so the 3x3 stencil holds every pair
  const double cs = sqrt(S.cutoff) * (1.0 + 1e-9) * (cell_factor > 1.0 ? cell_factor : 1.0);
  S.cell_inv = 1.0 / cs;
  // the margin (in cells) of a lazy re-sort: 1 - 2h / cell, less a rounding guard
  S.skin_m = std::max(0.0, 1.0 - sqrt(S.cutoff) * S.cell_inv - 1e-9);
  S.xmin = u.box[0];
  S.ymin = u.box[1];
  S.ncx = (int)floor((u.box[2] - u.box[0]) * S.cell_inv) + 1;
  S.ncy = (int)floor((u.box[3] - u.box[1]) * S.cell_inv) + 1;
  S.nf = u.n_fluid;
  S.np = u.n_fluid + u.n_fixed + u.n_moving;
  S.K = u.K;
  S.gravity = u.gravity;
  S.motion = u.motion;
  S.h = u.h;
  S.dt = u.dt;
  S.rho0 = u.rho0;
  S.gamma = u.gamma;
  S.gamma_int = (u.gamma == floor(u.gamma) && u.gamma >= 1 && u.gamma <= 16) ? (int)u.gamma : 0;
  S.B = u.B;
  S.c0 = u.c0;
  S.g = u.g;
  S.r0 = 3.0 * u.dx;
  S.m_amp = u.m_amp;
  S.m_omega = u.m_omega;
  S.fh = (float)u.h;
  S.finvh = (float)(1.0 / u.h);
  S.fwalpha = (float)(10.0 / (7. * M_PI * u.h * u.h));
  S.fav = (float)u.alpha_av;
  S.fc0 = (float)u.c0;
  S.fk = (float)u.k_spring;
  S.fcd = (float)u.c_damp;
  S.fr0 = (float)(3.0 * u.dx);
  S.frho0 = (float)u.rho0;
  S.fk7 = (float)(-2145.0 / (192512.0 * M_PI * u.h * u.h * u.h * u.h));
  S.fw7 = (float)(2145.0 / (1347584.0 * M_PI * u.h * u.h));
  S.kernel = u.kernel;
  S.mode = u.mode;
  S.eps = u.eps;
  S.stress2 = u.mode == UPH_MODE_REF && u.stress_type == 2;
  S.lid_u = u.lid_u;
  S.t_stop = u.lid_u != 0.0 ? 1.5 * (u.depth - 15.0 * u.dx) / u.lid_u : 0.0;   // inputbp.c:227
  S.flat_bottom = u.flat_bottom;
  S.fK0 = (float)(u.rho0 * u.c0 * u.c0);
  S.fgamma = (float)u.gamma;
  S.fgk = (float)(u.gamma * u.kappa);
  S.fkap = (float)u.kappa;
  S.fmu = (float)u.mu;
  S.lbp0 = u.lbp;
  S.rbp0 = u.rbp;
  S.depth = u.depth;
  S.dx = u.dx;
}

// kmask (96 bytes per slot) is written and read only by the per-lane K-th search of
// handles whose queries are expected to be capped; other handles leave it null.
int alloc_bufs(StepBufs& b, int n, int ncells, int nblocks_scan, Err& err, bool search_only, bool capped) {
  memset(&b, 0, sizeof b);
  if (dalloc(&b.posA, (size_t)n + POS_PAD, err) || dalloc(&b.posB, (size_t)n + POS_PAD, err) || dalloc(&b.tagA, n, err) ||
      dalloc(&b.tagB, n, err) || dalloc(&b.keys, n, err) || dalloc(&b.rank, n, err) ||
      dalloc(&b.tmpIdx, n, err) || dalloc(&b.cellStart, (size_t)ncells + 1, err) ||
      dalloc(&b.cellCount, ncells, err) || dalloc(&b.partial, nblocks_scan, err) ||
      dalloc(&b.sceneOf, n, err) || dalloc(&b.kth, n, err) ||
      dalloc(&b.ctl, CTL_WORDS, err))
    return -1;
  b.it = b.ctl + C_IT;
  if (hipMemset(b.ctl, 0, CTL_WORDS * sizeof(int)) != hipSuccess ||
      hipMemset(b.posA + n, 0, POS_PAD * sizeof(double2)) != hipSuccess ||
      hipMemset(b.posB + n, 0, POS_PAD * sizeof(double2)) != hipSuccess) {
    err.msg = "step buffers: init failed";
    return -1;
  }
  if (search_only) return 0;
  if (dalloc(&b.velA, n, err) || dalloc(&b.velB, n, err) || dalloc(&b.rhoA, n, err) ||
      dalloc(&b.rhoB, n, err) || dalloc(&b.rtB, n, err) ||
      (capped && dalloc(&b.kmask, (size_t)n * 12, err)) ||
      dalloc(&b.mA, n, err) || dalloc(&b.mB, n, err) || dalloc(&b.deckPos, n, err) ||
      dalloc(&b.deckVel, n, err) || dalloc(&b.deckLvel, n, err) || dalloc(&b.massD, n, err) ||
      dalloc(&b.presA, n, err) || dalloc(&b.presB, n, err) || dalloc(&b.deckP, n, err) ||
      dalloc(&b.rt2B, n, err) || dalloc(&b.dvB, n, err) || dalloc(&b.mC, n, err) ||
      dalloc(&b.deckRl, n, err) || dalloc(&b.deckPl, n, err) || dalloc(&b.wallKind, n, err) ||
      dalloc(&b.wallRef, n, err))
    return -1;
  b.kmask_n = capped ? n : 0;
  return 0;
}

// The second state set, the binned cells and the stale flags of a lazy re-sort
// handle (see StepBufs), zeroed.
static int alloc_lazy_bufs(StepBufs& b, int n, int nscenes, Err& err) {
  if (dalloc(&b.mA2, n, err) || dalloc(&b.mB2, n, err) || dalloc(&b.rtA, n, err) || dalloc(&b.rt2A, n, err) ||
      dalloc(&b.mC2, n, err) || dalloc(&b.dvA, n, err) || dalloc(&b.cellOf, n, err) ||
      dalloc(&b.sstale, 2 * (size_t)nscenes, err))
    return -1;
  if (hipMemset(b.rtA, 0, (size_t)n * sizeof(float4)) != hipSuccess ||
      hipMemset(b.rt2A, 0, (size_t)n * sizeof(float4)) != hipSuccess ||
      hipMemset(b.dvA, 0, (size_t)n * sizeof(float2)) != hipSuccess ||
      hipMemset(b.sstale, 0, 2 * (size_t)nscenes * sizeof(int)) != hipSuccess) {
    err.msg = "lazy buffers: init failed";
    return -1;
  }
  b.lazy = 1;
  b.nscenes = nscenes;
  return 0;
}

// Mirror-row buffers for at most gtot rows and wtot words of test bits (see
// StepBufs), zeroed.
static int alloc_ghost_bufs(StepBufs& b, int n, int ncells, int gtot, int wtot, Err& err) {
  b.gtot = gtot;
  b.wtot = wtot;
  const size_t g = (size_t)gtot, w = (size_t)wtot;
  const size_t wnb = (w + SCAN_TILE - 1) / SCAN_TILE;
  if (dalloc(&b.gbits, w, err) || dalloc(&b.gbits2, w, err) || dalloc(&b.gwcnt, w, err) ||
      dalloc(&b.gwstart, w + 1, err) || dalloc(&b.gpart, wnb, err) || dalloc(&b.gcand, (size_t)n, err) ||
      dalloc(&b.gcandN, 1, err) || dalloc(&b.gpos, g, err) || dalloc(&b.gid, g, err) || dalloc(&b.gscene, g, err) || dalloc(&b.gkey, g, err) ||
      dalloc(&b.grank, g, err) || dalloc(&b.gcellCount, (size_t)ncells, err) ||
      dalloc(&b.gcellStart, (size_t)ncells + 1, err) || dalloc(&b.gposS, g, err) || dalloc(&b.gidS, g, err) ||
      dalloc(&b.gvel, g, err) || dalloc(&b.glvel, g, err) || dalloc(&b.grho, g, err) || dalloc(&b.grl, g, err) ||
      dalloc(&b.gP, g, err) || dalloc(&b.gPl, g, err) || dalloc(&b.gmass, g, err) || dalloc(&b.predV, n, err) ||
      dalloc(&b.predVl, n, err) || dalloc(&b.predRho, n, err) || dalloc(&b.predRl, n, err) ||
      dalloc(&b.predP, n, err) || dalloc(&b.predPl, n, err))
    return -1;
  if (hipMemset(b.gbits, 0, w * sizeof(uint32_t)) != hipSuccess ||
      hipMemset(b.gwstart, 0, (w + 1) * sizeof(int)) != hipSuccess ||
      hipMemset(b.gcandN, 0, sizeof(int)) != hipSuccess ||
      hipMemset(b.gcellCount, 0, (size_t)ncells * sizeof(int)) != hipSuccess ||
      hipMemset(b.gcellStart, 0, ((size_t)ncells + 1) * sizeof(int)) != hipSuccess) {
    err.msg = "mirror buffers: init failed";
    return -1;
  }
  return 0;
}

void free_bufs(StepBufs& b) {
  void* ptrs[] = {b.mA2, b.mB2, b.rtA, b.rt2A, b.mC2, b.dvA, b.cellOf, b.sstale,
                  b.posA, b.velA, b.rhoA, b.tagA, b.mA, b.mB, b.posB, b.velB, b.rhoB,
                  b.rtB, b.tagB, b.keys, b.rank, b.tmpIdx, b.cellStart, b.cellCount, b.partial,
                  b.sceneOf, b.kth, b.kmask, b.deckPos, b.deckVel, b.deckLvel, b.massD, b.ctl,
                  b.dbgRates, b.presA, b.presB, b.deckP, b.rt2B, b.dvB, b.mC, b.deckRl, b.deckPl,
                  b.wallKind, b.wallRef, b.gbits, b.gbits2, b.gwcnt, b.gwstart, b.gpart, b.gcand, b.gcandN, b.gpos, b.gid, b.gscene, b.gkey,
                  b.grank, b.gcellCount, b.gcellStart, b.gposS, b.gidS, b.gvel, b.glvel, b.grho, b.grl,
                  b.gP, b.gPl, b.gmass, b.predV, b.predVl, b.predRho, b.predRl, b.predP, b.predPl, b.sst};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  memset(&b, 0, sizeof b);
}

// Exclusive scan of count[0, m) into start[0, m] over nb tiles (partial: nb ints),
// zeroing count.
int launch_scan(hipStream_t st, int m, int nb, int* count, int* partial, int* start) {
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_BLK), 0, st, m, count, partial, (int*)nullptr, (int*)nullptr,
                     0, 0, (const int*)nullptr);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(SCAN_BLK), 0, st, m, count, partial, start, (const int*)nullptr);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// The sort kernels: bin, scan, scatter, reorder, gathering the B set of buffer view
// `view` into its A set (step_ctl: the first kernel opens the step; for a lazy handle
// also its parity flags, and the view becomes the pair kernel's).
int launch_sort(hipStream_t st, int n, int ncells, int nb, StepBufs& b0, const DevScene* dsc, int search_only,
                bool step_ctl, Err& err, int view) {
  const StepBufs b = view ? swapped(b0) : b0;
  hipLaunchKernelGGL(k_bin, dim3(grid_of(n)), dim3(BLK), 0, st, n, b.posB, b.sceneOf, dsc, b.tagA, b.tagB, b.keys,
                     b.rank, b.cellCount);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_BLK), 0, st, ncells, b.cellCount, b.partial,
                     step_ctl ? b.ctl : (int*)nullptr, (step_ctl && b.lazy) ? b.sstale : (int*)nullptr, b.nscenes,
                     view, (const int*)nullptr);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(SCAN_BLK), 0, st, ncells, b.cellCount, b.partial, b.cellStart,
                     (const int*)nullptr);
  hipLaunchKernelGGL(k_scatter, dim3(grid_of(n)), dim3(BLK), 0, st, n, b.keys, b.rank, b.cellStart, b.tmpIdx,
                     (const int*)nullptr);
  if (b.gtot > 0 && !search_only)
    hipLaunchKernelGGL(k_reorder<true>, dim3(grid_of(n)), dim3(BLK), 0, st, n, b, dsc, search_only);
  else
    hipLaunchKernelGGL(k_reorder<false>, dim3(grid_of(n)), dim3(BLK), 0, st, n, b, dsc, search_only);
  HIPCHK(hipGetLastError(), err);
  return 0;
}

// The re-sort a lazy handle's last call held over, at the first step of a call: each
// kernel returns at once unless C_PEND is set.
static int launch_held_sort(uph_handle* h, hipStream_t st) {
  StepBufs& b = h->b;
  const int* gate = b.ctl + C_PEND;
  const int n = h->n, nb = h->cs.nb;
  hipLaunchKernelGGL(k_bin_held, dim3(grid_of(n)), dim3(BLK), 0, st, n, b, h->dsc);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_BLK), 0, st, h->ncells, b.cellCount, b.partial,
                     (int*)nullptr, (int*)nullptr, 0, 0, gate);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(SCAN_BLK), 0, st, h->ncells, b.cellCount, b.partial, b.cellStart,
                     gate);
  hipLaunchKernelGGL(k_scatter, dim3(grid_of(n)), dim3(BLK), 0, st, n, b.keys, b.rank, b.cellStart, b.tmpIdx, gate);
  hipLaunchKernelGGL(k_permute_held, dim3(grid_of(n)), dim3(BLK), 0, st, n, b, (const DevScene*)h->dsc,
                     (int)(h->mode == UPH_MODE_REF));
  HIPCHK(hipGetLastError(), h->err);
  return 0;
}

// The kernels after a lazy step's pair pass (hold: the last step of a call).
template <int KT, int MODE, bool ONE>
static int launch_post_kt(uph_handle* h, hipStream_t st, int hold) {
  StepBufs& b = h->b;
  const int n = h->n, nb = h->cs.nb, g = std::min(grid8_of(n), FIX_GRID);
  hipLaunchKernelGGL((k_post<KT, MODE, ONE>), dim3(g), dim3(BLK), 0, st, n, hold, b, h->dsc);
  hipLaunchKernelGGL(k_lscan_reduce, dim3(nb), dim3(SCAN_BLK), 0, st, h->ncells, b.cellCount, b.partial,
                     (const int*)b.ctl, hold);
  hipLaunchKernelGGL(k_lscan_down, dim3(nb), dim3(SCAN_BLK), 0, st, h->ncells, b.cellCount, b.partial, b.cellStart,
                     (const int*)b.ctl, hold);
  hipLaunchKernelGGL(k_lscatter, dim3(g), dim3(BLK), 0, st, n, b.keys, b.rank, b.cellStart, b.tmpIdx,
                     (const int*)b.ctl, hold);
  hipLaunchKernelGGL(k_close, dim3(g), dim3(BLK), 0, st, n, hold, b, (const DevScene*)h->dsc);
  HIPCHK(hipGetLastError(), h->err);
  return 0;
}

template <int KT, int MODE>
static int launch_post(uph_handle* h, hipStream_t st, int hold) {
  return h->nscenes == 1 ? launch_post_kt<KT, MODE, true>(h, st, hold) : launch_post_kt<KT, MODE, false>(h, st, hold);
}

// The mirror rows of this step and their cell list (scenes with a virtualType that
// has sections only; the sort has set the test bits and listed the rows): counts,
// scan, rows, bin, scan, scatter.
static int launch_ghosts(uph_handle* h, hipStream_t st) {
  StepBufs& b = h->b;
  const int wnb = (b.wtot + SCAN_TILE - 1) / SCAN_TILE;
  hipLaunchKernelGGL(k_ghost_count, dim3(grid_of(b.wtot)), dim3(BLK), 0, st, b);
  if (launch_scan(st, b.wtot, wnb, b.gwcnt, b.gpart, b.gwstart)) return -1;
  hipLaunchKernelGGL(k_ghost_emit, dim3(std::min(grid_of(h->gfluid), 256)), dim3(BLK), 0, st, b, h->dsc);
  hipLaunchKernelGGL(k_ghost_bin, dim3(grid_of(b.gtot)), dim3(BLK), 0, st, b, h->dsc);
  if (launch_scan(st, h->ncells, h->cs.nb, b.gcellCount, b.partial, b.gcellStart)) return -1;
  hipLaunchKernelGGL(k_ghost_scatter, dim3(grid_of(b.gtot)), dim3(BLK), 0, st, b);
  HIPCHK(hipGetLastError(), h->err);
  return 0;
}

template <int KT, int MODE>
static int launch_step_kt(uph_handle* h, int k, int n_steps, hipEvent_t pa, hipEvent_t pb, hipEvent_t pc) {
  hipStream_t st = h->stream;
  if (!h->b.lazy) {
    if (launch_sort(st, h->n, h->ncells, h->cs.nb, h->b, h->dsc, 0, true, h->err)) return -1;
    if (h->b.gtot > 0 && launch_ghosts(h, st)) return -1;
  } else if (h->need_sort) {
    if (launch_sort(st, h->n, h->ncells, h->cs.nb, h->b, h->dsc, 0, true, h->err, h->sort_view)) return -1;
    h->need_sort = false;
  } else if (k == 0) {
    // the held re-sort's kernels return at once when none is pending; with the stream
    // idle the flag is read here instead of launching them
    int pend = 1;
    if (hipStreamQuery(st) == hipSuccess &&
        hipMemcpy(&pend, h->b.ctl + C_PEND, sizeof pend, hipMemcpyDeviceToHost) != hipSuccess)
      pend = 1;
    if (pend && launch_held_sort(h, st)) return -1;
  }
  if (pa) HIPCHK(hipEventRecord(pa, st), h->err);
  if constexpr (MODE == UPH_MODE_REF2) {
    if (launch_stress(h, st)) return -1;
  }
  if (launch_pair_pass(h, st)) return -1;
  if (pb) HIPCHK(hipEventRecord(pb, st), h->err);
  if constexpr (MODE == UPH_MODE_REF2) {
    if (launch_capped(h, st)) return -1;   // such a handle sorts every step
  } else {
    if (h->b.lazy ? launch_post<KT, MODE>(h, st, (int)(k == n_steps - 1)) : launch_capped(h, st)) return -1;
  }
  if (pc) HIPCHK(hipEventRecord(pc, st), h->err);
  return 0;
}

// Step k of a call of n_steps steps.  pa, pb, pc (may be null): events recorded after
// the sort, after the pair pass and after the capped pass (k_post for a lazy handle).
int launch_call_step(uph_handle* h, int k, int n_steps, hipEvent_t pa, hipEvent_t pb, hipEvent_t pc) {
  int rc;
  if (h->stress2)
    rc = h->kernel == UPH_KERNEL_POLY7 ? launch_step_kt<UPH_KERNEL_POLY7, UPH_MODE_REF2>(h, k, n_steps, pa, pb, pc)
                                       : launch_step_kt<UPH_KERNEL_CUBIC, UPH_MODE_REF2>(h, k, n_steps, pa, pb, pc);
  else if (h->mode == UPH_MODE_REF)
    rc = h->kernel == UPH_KERNEL_POLY7 ? launch_step_kt<UPH_KERNEL_POLY7, UPH_MODE_REF>(h, k, n_steps, pa, pb, pc)
                                       : launch_step_kt<UPH_KERNEL_CUBIC, UPH_MODE_REF>(h, k, n_steps, pa, pb, pc);
  else
    rc = h->kernel == UPH_KERNEL_POLY7 ? launch_step_kt<UPH_KERNEL_POLY7, UPH_MODE_CONFIG>(h, k, n_steps, pa, pb, pc)
                                       : launch_step_kt<UPH_KERNEL_CUBIC, UPH_MODE_CONFIG>(h, k, n_steps, pa, pb, pc);
  if (rc != 0) return -1;
  HIPCHK(hipGetLastError(), h->err);
  return 0;
}

// One step of a sort-every-step handle (the hipGraph capture, the derivatives() seam).
int launch_step(uph_handle* h, hipEvent_t pa, hipEvent_t pb, hipEvent_t pc) {
  return launch_call_step(h, 0, 1, pa, pb, pc);
}

int read_bufs(uph_t* h, StepBufs& out) {
  out = h->b;
  if (!h->b.lazy) return 0;
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  int v = 0;
  HIPCHK(hipMemcpy(&v, h->b.ctl + C_VC, sizeof v, hipMemcpyDeviceToHost), h->err);
  if (v & 1) out = swapped(h->b);
  out.it = h->b.ctl + C_ITC;   // the last step's itime and K-th key stamp
  return 0;
}

int lazy_reset(uph_t* h) {
  if (!h->b.lazy) return 0;
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  int v = 0;
  HIPCHK(hipMemcpy(&v, h->b.ctl + C_VC, sizeof v, hipMemcpyDeviceToHost), h->err);
  h->sort_view = v & 1;
  h->need_sort = true;
  // k_post left the next step's itime; the sort that opens that step counts it again
  HIPCHK(hipMemcpy(h->b.ctl + C_IT, &h->itime, sizeof(int), hipMemcpyHostToDevice), h->err);
  return 0;
}

static std::string g_create_err;

extern "C" int uph_create(const uph_scene* scenes, int n_scenes, const uph_opts* opts, uph_t** out) {
  *out = nullptr;
  if (!scenes || n_scenes <= 0) {
    g_create_err = "uph_create: no scenes";
    return -1;
  }
  uph_handle* h = new uph_handle();
  if (opts) h->opts = *opts;
  if (h->opts.pair_block != 128 && h->opts.pair_block != 512) h->opts.pair_block = 256;
  if (!(h->opts.cell_factor >= 1.0) || h->opts.cell_factor > 4.0) h->opts.cell_factor = 1.0;
  h->nscenes = n_scenes;
  h->hs.resize(n_scenes);
  // lazy re-sort where it applies: no scene with mirror rows (the sort evaluates their
  // section tests every step), none expected to be capped, no hipGraph (the step's
  // launches depend on the call); its cells grow by the skin.  By default (0) also only
  // for batches of at most LAZY_MAX_SCENES scenes: a batch re-sorts at every step where
  // one of its scenes has gone stale, so a large batch of moving scenes re-sorts nearly
  // every step, and a lazy re-sort step costs more than an every-step sort; -1 asks for
  // it whatever the batch size.
  bool any_mirrors = false, any_capped = false;
  for (int s = 0; s < n_scenes; ++s) {
    const uph_scene& u = scenes[s];
    int sec[5];
    if (u.n_fluid > 0 && mirror_sections(u.vtype, sec) > 0) any_mirrors = true;
    if (!(4.0 * M_PI * (u.h / u.dx) * (u.h / u.dx) <= 0.75 * u.K)) any_capped = true;
  }
  // stressType 2: the same for every scene; its stress sums take every neighbour in
  // range, so no mirror rows (which rank among the K nearest) and no capped queries; it
  // sorts every step
  const bool stress2 = scenes[0].mode == UPH_MODE_REF && scenes[0].stress_type == 2;
  for (int s = 0; s < n_scenes; ++s)
    if ((scenes[s].mode == UPH_MODE_REF && scenes[s].stress_type == 2) != stress2) {
      g_create_err = "uph_create: scene " + std::to_string(s) + " has not the stressType of scene 0";
      delete h;
      return -1;
    }
  if (stress2 && (any_mirrors || any_capped)) {
    g_create_err = "uph_create: stressType 2 needs scenes without mirror particles whose queries are not "
                   "expected to be capped (4 pi (h/dx)^2 at most 3/4 of K)";
    delete h;
    return -1;
  }
  h->stress2 = stress2;
  bool lazy = h->opts.sort_every_step <= 0 && !h->opts.use_graph && !any_mirrors && !any_capped && !stress2 &&
              (h->opts.sort_every_step < 0 || n_scenes <= LAZY_MAX_SCENES);
  // default growth: 2 % of the cell for one scene, 5 % for a batch (a batch re-sorts when
  // any of its scenes needs it, so its cells carry more slack)
  if (!(h->opts.skin > 0.0)) h->opts.skin = n_scenes == 1 ? 0.02 : 0.05;
  if (h->opts.skin > 1.0) h->opts.skin = 1.0;
  const double cell_factor = lazy ? h->opts.cell_factor * (1.0 + h->opts.skin) : h->opts.cell_factor;
  long n = 0, nc = 0, gtot = 0, wtot = 0, gfluid = 0;
  bool kth_all = !stress2 && h->opts.pair_variant == 0;
  for (int s = 0; s < n_scenes; ++s) {
    const uph_scene& u = scenes[s];
    const int np = u.n_fluid + u.n_fixed + u.n_moving;
    if (u.n_fluid < 0 || u.n_fixed < 0 || u.n_moving < 0 || np <= 0 || !(u.h > 0) || !(u.dt > 0) ||
        u.K <= 0) {
      g_create_err = "uph_create: scene " + std::to_string(s) + " has invalid sizes or constants";
      delete h;
      return -1;
    }
    if ((u.kernel != UPH_KERNEL_CUBIC && u.kernel != UPH_KERNEL_POLY7) || u.kernel != scenes[0].kernel) {
      g_create_err = "uph_create: scene " + std::to_string(s) +
                     " has an unknown smoothing kernel or not the kernel of scene 0";
      delete h;
      return -1;
    }
    if (u.motion < UPH_MOTION_NONE || u.motion > UPH_MOTION_SLOSH ||
        (u.motion == UPH_MOTION_GATE && !(u.lid_u != 0.0))) {
      g_create_err = "uph_create: scene " + std::to_string(s) + " has an unknown motion (or a gate with no speed)";
      delete h;
      return -1;
    }
    if ((u.mode != UPH_MODE_CONFIG && u.mode != UPH_MODE_REF) || u.mode != scenes[0].mode) {
      g_create_err = "uph_create: scene " + std::to_string(s) +
                     " has an unknown scheme (mode) or not the scheme of scene 0";
      delete h;
      return -1;
    }
    for (int i = u.n_fluid; i < np; ++i)
      if (!(u.rho[i] > 0)) {
        g_create_err = "uph_create: wall densities must be positive";
        delete h;
        return -1;
      }
    fill_scene(h->hs[s], u, cell_factor);
    DevScene& S = h->hs[s];
    S.sidx = s;
    if (S.ncx > 65535 || S.ncy > 65535) lazy = false;   // cellOf packs a cell's coordinates in 16 bits each
    S.poff = (int)n;
    S.coff = (int)nc;
    S.pb_gid = (int)n + (u.n_fluid + u.n_fixed + u.n_moving - 1);
    // mirror rows: one test per section and fluid row, ids after the scene's rows
    if (4.0 * M_PI * (u.h / u.dx) * (u.h / u.dx) > 0.75 * u.K) h->capped = true;
    S.nsec = u.n_fluid > 0 ? mirror_sections(u.vtype, S.sec) : 0;
    if (!(4.0 * M_PI * (u.h / u.dx) * (u.h / u.dx) > u.K) || S.nsec > 0) kth_all = false;
    S.nwords = (u.n_fluid + 31) / 32;
    S.gw = (int)wtot;
    S.gbase = (int)n + np;
    gtot += (long)S.nsec * u.n_fluid;
    wtot += (long)S.nsec * S.nwords;
    if (S.nsec > 0) gfluid += u.n_fluid;
    if ((long)S.gbase + (long)S.nsec * u.n_fluid >= (long)UPH_ID_MASK) {
      g_create_err = "uph_create: too many particles and mirror rows";
      delete h;
      return -1;
    }
    n += np;
    nc += (long)S.ncx * S.ncy;
    h->fluid_n.push_back(u.n_fluid);
    h->tot_n.push_back(np);
  }
  if (n >= (long)UPH_ID_MASK || nc >= (1L << 30) || gtot >= (1L << 30)) {
    g_create_err = "uph_create: too many particles or cells";
    delete h;
    return -1;
  }
  h->n = (int)n;
  h->kth_all = kth_all && h->capped && !lazy;
  h->kernel = scenes[0].kernel;
  h->mode = scenes[0].mode;
  h->ncells = (int)nc;
  h->cs.m = (int)nc;
  h->cs.nb = (int)((nc + SCAN_TILE - 1) / SCAN_TILE);
  Err& err = h->err;
  auto fail = [&]() {
    g_create_err = err.msg;
    free_bufs(h->b);
    if (h->dsc) (void)hipFree(h->dsc);
    if (h->mixedWaves) (void)hipFree(h->mixedWaves);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return -1;
  };
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    err.msg = "hipStreamCreate failed";
    return fail();
  }
  if (alloc_bufs(h->b, h->n, h->ncells, h->cs.nb, err, false, h->capped)) return fail();
  if (h->stress2 && dalloc(&h->b.sst, h->n, err)) return fail();
  if (lazy) {
    if (alloc_lazy_bufs(h->b, h->n, n_scenes, err)) return fail();
    if (h->mode == UPH_MODE_REF) h->b.lazy = 2;   // the re-sort's gather carries DynamicPressure and mC
  }
  h->gfluid = (int)gfluid;
  if (gtot > 0 && alloc_ghost_bufs(h->b, h->n, h->ncells, (int)gtot, (int)wtot, err)) return fail();
  if (h->opts.keep_rates && dalloc(&h->b.dbgRates, (h->mode == UPH_MODE_REF ? 4 : 2) * (size_t)h->n, err))
    return fail();
  if (dalloc(&h->dsc, n_scenes, err)) return fail();
  // host staging in original order (global index = slot at start)
  std::vector<double2> pos(n), vel(n), lv(n);
  std::vector<double> rho(n), massv(n), pres(n), rl0(n), wref(n, 0.0);
  std::vector<int> wkind(n, 0);
  std::vector<float4> rt(n);
  std::vector<uint32_t> tag(n);
  std::vector<int> sceneOf(n);
  for (int s = 0; s < n_scenes; ++s) {
    const uph_scene& u = scenes[s];
    const DevScene& S = h->hs[s];
    int halted = 0;
    for (int i = 0; i < S.np; ++i) {
      const long g = S.poff + i;
      pos[g] = make_double2(u.pos[2 * i], u.pos[2 * i + 1]);
      vel[g] = make_double2(u.vel[2 * i], u.vel[2 * i + 1]);
      lv[g] = make_double2(u.lvel[2 * i], u.lvel[2 * i + 1]);
      rho[g] = u.rho[i];
      pres[g] = u.pres ? u.pres[i] : 0.0;
      rl0[g] = u.rho0;
      rt[g] = make_float4(0.f, 0.f, 0.f, (float)u.mass[i]);
      massv[g] = u.mass[i];
      int ty = u.itype[i];
      if (i >= u.n_fluid) wkind[g] = wall_class(u.motion, i, u.n_fluid, u.n_fixed, ty, u.dx,
                                                u.gate_y ? u.gate_y[i] : 0.0, &wref[g], &halted);
      // walls keep their itype (2, 3, 5 take the spring force, derivatives.c:2406);
      // a wall of itype 1 (the sloshing tank's bottom) becomes 4, a wall without it
      if (i < u.n_fluid) ty = 1;
      else if (ty == 1) ty = 4;
      else if (ty < 0 || ty > 7) ty = 2;
      tag[g] = ((uint32_t)ty << UPH_ID_BITS) | (uint32_t)g;
      sceneOf[g] = s;
    }
    h->hs[s].slosh_halt = halted;   // uploaded with the scene table below
  }
  // wavefronts of the compacted pair kernel whose 64 slots (the last slot standing in
  // for those past the end) belong to two scenes: scenes hold contiguous slot ranges,
  // so the first and last slot decide
  {
    std::vector<int> mixed;
    for (long w = 0; 64 * w < n; ++w) {
      const long a = 64 * w, e = std::min(64 * w + 63, n - 1);
      if (sceneOf[a] != sceneOf[e]) mixed.push_back((int)w);
    }
    h->nmixed = (int)mixed.size();
    if (h->nmixed > 0 &&
        (dalloc(&h->mixedWaves, mixed.size(), err) ||
         hipMemcpy(h->mixedWaves, mixed.data(), mixed.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)) {
      if (err.msg.empty()) err.msg = "uph_create: upload failed";
      return fail();
    }
  }
  if (hipMemcpy(h->b.deckPos, pos.data(), n * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) {
    err.msg = "uph_create: upload failed";
    return fail();
  }
  h->deck_pos = pos;
  h->wall_kind = wkind;
  h->wall_ref = wref;
  // boundary particles start at their step-1 positions (derivatives is called with t = dt)
  for (int s = 0; s < n_scenes; ++s) {
    const DevScene& S = h->hs[s];
    if (S.motion == 0) continue;
    for (int i = S.nf; i < S.np; ++i) {
      const long g = S.poff + i;
      double lvx;
      const double2 xd = pos[g];
      (void)wall_kin(S, wkind[g], wref[g], 1 * S.dt, xd, pos[g], vel[g], &lvx);
    }
  }
  std::vector<double2> vdeck(n);
  for (int s = 0; s < n_scenes; ++s) {
    const uph_scene& u = scenes[s];
    for (int i = 0; i < h->hs[s].np; ++i) vdeck[h->hs[s].poff + i] = make_double2(u.vel[2 * i], u.vel[2 * i + 1]);
  }
  h->deck_vel = vdeck;
  int ctl[CTL_WORDS] = {0};
  if (hipMemcpy(h->b.deckVel, vdeck.data(), n * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->b.deckLvel, lv.data(), n * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->b.massD, massv.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->b.presB, pres.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->b.deckP, pres.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->b.deckRl, rl0.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->b.wallKind, wkind.data(), n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->b.wallRef, wref.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(h->b.deckPl, 0, n * sizeof(double)) != hipSuccess ||
      hipMemcpy(h->b.posB, pos.data(), n * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->b.velB, vel.data(), n * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->b.rhoB, rho.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->b.rtB, rt.data(), n * sizeof(float4), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->b.tagA, tag.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->b.sceneOf, sceneOf.data(), n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->dsc, h->hs.data(), n_scenes * sizeof(DevScene), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->b.ctl, ctl, sizeof ctl, hipMemcpyHostToDevice) != hipSuccess) {
    err.msg = "uph_create: upload failed";
    return fail();
  }
  // cell counts start at zero; every scan leaves them zero again
  if (hipMemset(h->b.cellCount, 0, (size_t)h->ncells * sizeof(int)) != hipSuccess ||
      hipMemset(h->b.posA + h->n, 0, POS_PAD * sizeof(double2)) != hipSuccess ||
      hipMemset(h->b.kth, 0, (size_t)h->n * sizeof(KthRec)) != hipSuccess ||
      hipMemset(h->b.rt2B, 0, (size_t)h->n * sizeof(float4)) != hipSuccess ||
      hipMemset(h->b.dvB, 0, (size_t)h->n * sizeof(float2)) != hipSuccess ||
      hipMemset(h->b.partial, 0, (size_t)h->cs.nb * sizeof(int)) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess) {
    err.msg = "uph_create: init failed";
    return fail();
  }
  if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
    err.msg = "hipEventCreate failed";
    return fail();
  }
  *out = h;
  return 0;
}

extern "C" void uph_destroy(uph_t* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->diag) (void)hipFree(h->diag);
  if (h->vort) (void)hipFree(h->vort);
  if (h->bar) (void)hipFree(h->bar);
  if (h->pargs) (void)hipFree(h->pargs);
  free_bufs(h->b);
  if (h->dsc) (void)hipFree(h->dsc);
  if (h->mixedWaves) (void)hipFree(h->mixedWaves);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

extern "C" int uph_step(uph_t* h, int n_steps) {
  if (!h) return -1;
  if (n_steps < 0) {
    h->err.msg = "uph_step: negative step count";
    return -1;
  }
  if (h->opts.use_graph && n_steps > 0) {
    if (!h->gexec) {
      HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal), h->err);
      int rc = launch_step(h, nullptr, nullptr);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(h->stream, &g);
      if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (!rc) h->err.msg = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
        return -1;
      }
      h->graph = g;
      HIPCHK(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0), h->err);
    }
    for (int k = 0; k < n_steps; ++k) HIPCHK(hipGraphLaunch(h->gexec, h->stream), h->err);
  } else {
    for (int k = 0; k < n_steps; ++k)
      if (launch_call_step(h, k, n_steps, nullptr, nullptr, nullptr)) return -1;
  }
  h->itime += n_steps;
  return 0;
}

template <int KT, int MODE, bool ONE>
static int persist_launch(uph_t* h, int n_steps) {
  const void* fn = (const void*)k_persist<KT, MODE, ONE>;
  int dev = 0, ncu = 0, bpc = 0;
  HIPCHK(hipGetDevice(&dev), h->err);
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), h->err);
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, fn, BLK, 0), h->err);
  int grid = (bpc * ncu) & ~7;
  if (grid < 8) {
    h->err.msg = "uph_step_persistent: the kernel does not fit 8 co-resident blocks";
    return -1;
  }
  // No more blocks than the widest phase has virtual blocks: a barrier's cost grows with
  // the blocks that arrive at it, and a small scene needs only a few dozen.
  const int widest = std::max((h->n + BLK - 1) / BLK, h->cs.nb);
  grid = std::min(grid, std::max(8, (widest + 7) & ~7));
  if (!h->pargs) HIPCHK(hipMalloc((void**)&h->pargs, sizeof(PersistArgs)), h->err);
  PersistArgs a;
  a.n = h->n;
  a.ncells = h->ncells;
  a.nb = h->cs.nb;
  a.nmixed = h->nmixed;
  a.n_steps = n_steps;
  a.b = h->b;
  a.sc = h->dsc;
  a.mixed = h->mixedWaves;
  a.bar = h->bar;
  HIPCHK(hipMemcpyAsync(h->pargs, &a, sizeof a, hipMemcpyHostToDevice, h->stream), h->err);
  const PersistArgs* dp = h->pargs;
  void* args[] = {(void*)&dp};
  HIPCHK(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(BLK), args, 0, h->stream), h->err);
  // the host copy must outlive the asynchronous upload
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  return 0;
}

extern "C" int uph_step_persistent(uph_t* h, int n_steps) {
  if (!h) return -1;
  if (n_steps < 0) {
    h->err.msg = "uph_step_persistent: negative step count";
    return -1;
  }
  if (h->b.gtot > 0 || h->capped || h->opts.pair_variant != 0 || h->b.lazy || h->stress2) {
    h->err.msg = "uph_step_persistent: for handles without mirror rows, not expected to be capped, with the "
                 "compacted pair kernel, created to re-sort every step, without stressType 2 (use uph_step)";
    return -1;
  }
  if (n_steps == 0) return 0;
  if (!h->bar) {
    HIPCHK(hipMalloc((void**)&h->bar, BAR_WORDS * sizeof(unsigned)), h->err);
    HIPCHK(hipMemset(h->bar, 0, BAR_WORDS * sizeof(unsigned)), h->err);
  }
  int rc;
  const bool one = h->nscenes == 1;
  if (h->mode == UPH_MODE_REF)
    rc = h->kernel == UPH_KERNEL_POLY7
             ? (one ? persist_launch<UPH_KERNEL_POLY7, UPH_MODE_REF, true>(h, n_steps)
                    : persist_launch<UPH_KERNEL_POLY7, UPH_MODE_REF, false>(h, n_steps))
             : (one ? persist_launch<UPH_KERNEL_CUBIC, UPH_MODE_REF, true>(h, n_steps)
                    : persist_launch<UPH_KERNEL_CUBIC, UPH_MODE_REF, false>(h, n_steps));
  else
    rc = h->kernel == UPH_KERNEL_POLY7
             ? (one ? persist_launch<UPH_KERNEL_POLY7, UPH_MODE_CONFIG, true>(h, n_steps)
                    : persist_launch<UPH_KERNEL_POLY7, UPH_MODE_CONFIG, false>(h, n_steps))
             : (one ? persist_launch<UPH_KERNEL_CUBIC, UPH_MODE_CONFIG, true>(h, n_steps)
                    : persist_launch<UPH_KERNEL_CUBIC, UPH_MODE_CONFIG, false>(h, n_steps));
  if (rc) return -1;
  h->itime += n_steps;
  return 0;
}

extern "C" int uph_sync(uph_t* h) {
  if (!h) return -1;
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  if (h->bar) {
    unsigned broken = 0;
    HIPCHK(hipMemcpy(&broken, h->bar + BAR_BROKEN, sizeof broken, hipMemcpyDeviceToHost), h->err);
    if (broken) {
      h->err.msg = "uph_sync: a grid barrier of uph_step_persistent gave up waiting; the steps' results are not "
                   "valid";
      return -1;
    }
  }
  int ctl[CTL_WORDS];
  HIPCHK(hipMemcpy(ctl, h->b.ctl, sizeof ctl, hipMemcpyDeviceToHost), h->err);
  if (ctl[C_SCAP]) {
    h->err.msg = "uph_sync: with stressType 2 a fluid particle had more than K neighbours in range; the viscous "
                 "stresses take every neighbour in range, so the steps' results are not valid";
    return -1;
  }
  if (ctl[C_BROKEN]) {
    h->err.msg = "uph_sync: a particle had more than " + std::to_string(CAND_MAX) +
                 " candidates in its 3x3 cells: the state has blown up (particles escaped or collapsed "
                 "into one cell); the step's results are not valid";
    return -1;
  }
  return ctl[C_NOVF];
}

extern "C" int uph_itime(const uph_t* h) { return h ? h->itime : -1; }

extern "C" double uph_skin(const uph_t* h) {
  if (!h || !h->b.lazy) return 0.0;
  return h->opts.skin;
}

extern "C" int uph_resorts(uph_t* h) {
  if (!h) return -1;
  if (!h->b.lazy) return -1;
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  int v = 0;
  HIPCHK(hipMemcpy(&v, h->b.ctl + C_NRES, sizeof v, hipMemcpyDeviceToHost), h->err);
  return v;
}

extern "C" int uph_time_steps(uph_t* h, int n_steps, double* ms_total, double* ms_pair) {
  if (!h || n_steps <= 0) return -1;
  std::vector<hipEvent_t> ev(2 * n_steps);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e), h->err);
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  HIPCHK(hipEventRecord(h->ev0, h->stream), h->err);
  for (int k = 0; k < n_steps; ++k)
    if (launch_call_step(h, k, n_steps, ev[2 * k], ev[2 * k + 1], nullptr)) return -1;
  HIPCHK(hipEventRecord(h->ev1, h->stream), h->err);
  HIPCHK(hipEventSynchronize(h->ev1), h->err);
  float tot = 0.f;
  HIPCHK(hipEventElapsedTime(&tot, h->ev0, h->ev1), h->err);
  double pair = 0.0;
  for (int k = 0; k < n_steps; ++k) {
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]), h->err);
    pair += t;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  h->itime += n_steps;
  if (ms_total) *ms_total = tot;
  if (ms_pair) *ms_pair = pair / n_steps;
  return 0;
}

extern "C" int uph_time_phases(uph_t* h, int n_steps, double* ms_total, double* ms_sort, double* ms_pair,
                               double* ms_exact) {
  if (!h || n_steps <= 0) return -1;
  // per step: start, after the sort, after the compacted pass, after the capped pass;
  // none when no phase is asked for (events between kernels cost a few us per step)
  const bool split = ms_sort || ms_pair || ms_exact;
  std::vector<hipEvent_t> ev(split ? 4 * n_steps : 0, nullptr);
  int rc = 0;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      h->err.msg = "uph_time_phases: hipEventCreate failed";
      rc = -1;
      break;
    }
  if (rc == 0 && hipStreamSynchronize(h->stream) != hipSuccess) rc = -1;
  if (rc == 0 && hipEventRecord(h->ev0, h->stream) != hipSuccess) rc = -1;
  for (int k = 0; rc == 0 && k < n_steps; ++k) {
    if (!split) {
      if (launch_call_step(h, k, n_steps, nullptr, nullptr, nullptr)) rc = -1;
    } else if (hipEventRecord(ev[4 * k], h->stream) != hipSuccess ||
               launch_call_step(h, k, n_steps, ev[4 * k + 1], ev[4 * k + 2], ev[4 * k + 3])) {
      rc = -1;
    }
  }
  if (rc == 0 && (hipEventRecord(h->ev1, h->stream) != hipSuccess || hipEventSynchronize(h->ev1) != hipSuccess))
    rc = -1;
  double sum[4] = {0.0, 0.0, 0.0, 0.0};
  if (rc == 0) {
    float tot = 0.f;
    (void)hipEventElapsedTime(&tot, h->ev0, h->ev1);
    sum[0] = tot;
    for (int k = 0; split && k < n_steps; ++k)
      for (int j = 0; j < 3; ++j) {
        float t = 0.f;
        (void)hipEventElapsedTime(&t, ev[4 * k + j], ev[4 * k + j + 1]);
        sum[j + 1] += t;
      }
    h->itime += n_steps;
  } else if (h->err.msg.empty()) {
    h->err.msg = "uph_time_phases: a HIP call failed";
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (rc) return -1;
  if (ms_total) *ms_total = sum[0];
  if (ms_sort) *ms_sort = sum[1] / n_steps;
  if (ms_pair) *ms_pair = sum[2] / n_steps;
  if (ms_exact) *ms_exact = sum[3] / n_steps;
  return 0;
}

extern "C" int uph_ghost_capacity(uph_t* h, int scene) {
  if (!h || scene < 0 || scene >= h->nscenes) return -1;
  return h->hs[scene].nsec * h->hs[scene].nf;
}

extern "C" int uph_get_ghosts(uph_t* h, int scene, int cap, double* pos, double* vel, double* lvel, double* rho,
                              double* rholoc, double* P, double* Ploc, double* mass) {
  if (!h) return -1;
  if (scene < 0 || scene >= h->nscenes) {
    h->err.msg = "uph_get_ghosts: scene out of range";
    return -1;
  }
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  const DevScene& S = h->hs[scene];
  if (S.nsec == 0) return 0;
  int se[2];
  HIPCHK(hipMemcpy(&se[0], h->b.gwstart + S.gw, sizeof(int), hipMemcpyDeviceToHost), h->err);
  HIPCHK(hipMemcpy(&se[1], h->b.gwstart + S.gw + S.nsec * S.nwords, sizeof(int), hipMemcpyDeviceToHost), h->err);
  const int cnt = se[1] - se[0];
  if (cnt < 0 || cnt > cap) {
    h->err.msg = "uph_get_ghosts: " + std::to_string(cnt) + " mirror rows, capacity " + std::to_string(cap);
    return -1;
  }
  auto get = [&](double* dst, const void* src, size_t width) -> int {
    if (!dst || cnt == 0) return 0;
    return hipMemcpy(dst, (const char*)src + (size_t)se[0] * width, (size_t)cnt * width, hipMemcpyDeviceToHost) ==
                   hipSuccess
               ? 0
               : -1;
  };
  if (get(pos, h->b.gpos, sizeof(double2)) || get(vel, h->b.gvel, sizeof(double2)) ||
      get(lvel, h->b.glvel, sizeof(double2)) || get(rho, h->b.grho, sizeof(double)) ||
      get(rholoc, h->b.grl, sizeof(double)) || get(P, h->b.gP, sizeof(double)) ||
      get(Ploc, h->b.gPl, sizeof(double)) || get(mass, h->b.gmass, sizeof(double))) {
    h->err.msg = "uph_get_ghosts: download failed";
    return -1;
  }
  return cnt;
}

extern "C" int uph_get_state(uph_t* h, int scene, double* pos, double* vel, double* rho) {
  if (!h) return -1;
  if (scene < 0 || scene >= h->nscenes) {
    h->err.msg = "uph_get_state: scene out of range";
    return -1;
  }
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  const DevScene& S = h->hs[scene];
  std::vector<double2> p(S.np), v(S.np);
  std::vector<double> r(S.np);
  std::vector<uint32_t> t(S.np);
  StepBufs rb;
  if (read_bufs(h, rb)) return -1;
  HIPCHK(hipMemcpy(p.data(), rb.posB + S.poff, S.np * sizeof(double2), hipMemcpyDeviceToHost), h->err);
  HIPCHK(hipMemcpy(v.data(), rb.velB + S.poff, S.np * sizeof(double2), hipMemcpyDeviceToHost), h->err);
  HIPCHK(hipMemcpy(r.data(), rb.rhoB + S.poff, S.np * sizeof(double), hipMemcpyDeviceToHost), h->err);
  HIPCHK(hipMemcpy(t.data(), rb.tagA + S.poff, S.np * sizeof(uint32_t), hipMemcpyDeviceToHost), h->err);
  for (int k = 0; k < S.np; ++k) {
    const int i = (int)((t[k] & UPH_ID_MASK) - (uint32_t)S.poff);
    if (i < 0 || i >= S.np) {
      h->err.msg = "uph_get_state: corrupt particle tags";
      return -1;
    }
    if (pos) { pos[2 * i] = p[k].x; pos[2 * i + 1] = p[k].y; }
    if (vel) { vel[2 * i] = v[k].x; vel[2 * i + 1] = v[k].y; }
    if (rho) rho[i] = r[k];
  }
  // the device holds the boundary rows at the next step's time; the reference's
  // state after step n has them where derivatives(t = n dt) put them (inputbp.c:192-283),
  // and before the first step at the deck
  if (S.motion != 0 && (pos || vel)) {
    const bool at_deck = h->itime == 0 && S.t0 == 0.0;
    const double t = S.t0 + h->itime * S.dt;
    for (int i = S.nf; i < S.np; ++i) {
      const long g = S.poff + i;
      double2 x = h->deck_pos[g], w = h->deck_vel[g];
      double lvx;
      if (!at_deck) (void)wall_kin(S, h->wall_kind[g], h->wall_ref[g], t, h->deck_pos[g], x, w, &lvx);
      if (pos) { pos[2 * i] = x.x; pos[2 * i + 1] = x.y; }
      if (vel) { vel[2 * i] = w.x; vel[2 * i + 1] = w.y; }
    }
  }
  return 0;
}

// Rates of one scene's fluid rows in reference order; the acceleration in FP64.
int get_rates_f64(uph_t* h, int scene, float* R, float* Rt, double* a, float* at) {
  if (!h) return -1;
  if (!h->b.dbgRates) {
    h->err.msg = "uph_get_rates: handle created without keep_rates";
    return -1;
  }
  if (scene < 0 || scene >= h->nscenes) {
    h->err.msg = "uph_get_rates: scene out of range";
    return -1;
  }
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  const DevScene& S = h->hs[scene];
  const int st = h->mode == UPH_MODE_REF ? 4 : 2;   // float4s per slot
  std::vector<float4> d(st * (size_t)S.np);
  std::vector<uint32_t> t(S.np);
  StepBufs rb;
  if (read_bufs(h, rb)) return -1;
  HIPCHK(hipMemcpy(d.data(), h->b.dbgRates + st * (size_t)S.poff, st * S.np * sizeof(float4),
                   hipMemcpyDeviceToHost),
         h->err);
  HIPCHK(hipMemcpy(t.data(), rb.tagA + S.poff, S.np * sizeof(uint32_t), hipMemcpyDeviceToHost), h->err);
  for (int k = 0; k < S.np; ++k) {
    const int i = (int)((t[k] & UPH_ID_MASK) - (uint32_t)S.poff);
    if (i < 0 || i >= S.np) continue;
    if (i >= S.nf) continue;
    if (R) R[i] = d[st * k].x;
    if (Rt) Rt[i] = d[st * k].y;
    if (at) { at[2 * i] = d[st * k].z; at[2 * i + 1] = d[st * k].w; }
    if (a) {
      double2 ad;
      memcpy(&ad, &d[st * k + 1], sizeof ad);
      a[2 * i] = ad.x;
      a[2 * i + 1] = ad.y;
    }
  }
  return 0;
}

extern "C" int uph_get_ref_rates(uph_t* h, int scene, float* RP, float* RPt, float* drho, float* dP, float* dv) {
  if (!h) return -1;
  if (h->mode != UPH_MODE_REF || !h->b.dbgRates) {
    h->err.msg = "uph_get_ref_rates: handle not in UPH_MODE_REF or created without keep_rates";
    return -1;
  }
  if (scene < 0 || scene >= h->nscenes) {
    h->err.msg = "uph_get_ref_rates: scene out of range";
    return -1;
  }
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  const DevScene& S = h->hs[scene];
  std::vector<float4> d(4 * (size_t)S.np);
  std::vector<uint32_t> t(S.np);
  StepBufs rb;
  if (read_bufs(h, rb)) return -1;
  HIPCHK(hipMemcpy(d.data(), h->b.dbgRates + 4 * (size_t)S.poff, 4 * S.np * sizeof(float4), hipMemcpyDeviceToHost),
         h->err);
  HIPCHK(hipMemcpy(t.data(), rb.tagA + S.poff, S.np * sizeof(uint32_t), hipMemcpyDeviceToHost), h->err);
  for (int k = 0; k < S.np; ++k) {
    const int i = (int)((t[k] & UPH_ID_MASK) - (uint32_t)S.poff);
    if (i < 0 || i >= S.nf) continue;
    const float4 e = d[4 * k + 2], f = d[4 * k + 3];
    if (RP) RP[i] = e.x;
    if (RPt) RPt[i] = e.y;
    if (drho) drho[i] = e.z;
    if (dP) dP[i] = e.w;
    if (dv) { dv[2 * i] = f.x; dv[2 * i + 1] = f.y; }
  }
  return 0;
}

extern "C" int uph_get_pressure(uph_t* h, int scene, double* P) {
  if (!h || !P) return -1;
  if (scene < 0 || scene >= h->nscenes) {
    h->err.msg = "uph_get_pressure: scene out of range";
    return -1;
  }
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  const DevScene& S = h->hs[scene];
  std::vector<double> v(S.np);
  std::vector<uint32_t> t(S.np);
  StepBufs rb;
  if (read_bufs(h, rb)) return -1;
  HIPCHK(hipMemcpy(v.data(), (h->mode == UPH_MODE_REF ? rb.presB : rb.rhoB) + S.poff, S.np * sizeof(double),
                   hipMemcpyDeviceToHost),
         h->err);
  HIPCHK(hipMemcpy(t.data(), rb.tagA + S.poff, S.np * sizeof(uint32_t), hipMemcpyDeviceToHost), h->err);
  for (int k = 0; k < S.np; ++k) {
    const int i = (int)((t[k] & UPH_ID_MASK) - (uint32_t)S.poff);
    if (i < 0 || i >= S.np) {
      h->err.msg = "uph_get_pressure: corrupt particle tags";
      return -1;
    }
    if (h->mode == UPH_MODE_REF) {
      P[i] = v[k];
    } else {   // EquationOfState (derivatives.c:1674-1682) of the density
      const double x = v[k] / S.rho0;
      P[i] = S.B * (pow(x, S.gamma) - 1.0);
    }
  }
  return 0;
}

extern "C" int uph_get_rates(uph_t* h, int scene, float* R, float* Rt, float* a, float* at) {
  std::vector<double> ad;
  if (a && h && scene >= 0 && scene < h->nscenes) ad.resize(2 * (size_t)h->hs[scene].nf);
  if (get_rates_f64(h, scene, R, Rt, a ? ad.data() : nullptr, at)) return -1;
  for (size_t k = 0; k < ad.size(); ++k) a[k] = (float)ad[k];
  return 0;
}

extern "C" const char* uph_last_error(const uph_t* h) { return h ? h->err.msg.c_str() : g_create_err.c_str(); }
