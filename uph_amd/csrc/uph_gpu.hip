// uph_gpu.hip -- kernels and C ABI of the MI355X SPH hot path.
//
// One step of sph.cc:90-168 (the predictor, derivatives() of
// derivatives.c:3137 in the scheme's configuration, the corrector) is, for all
// scenes of a batch at once:
//   k_bin         cell key of every particle, rank inside its cell
//                 (wavefront-aggregated atomics)
//   k_scan_reduce per-tile cell totals; opens the step
//   k_scan_down   exclusive scan of cell counts -> cell start offsets (each tile
//                 adds up the totals of the tiles before it)
//   k_scatter     provisional sorted slot of every particle
//   k_reorder     final slot by (cell, original index), gather of the FP64
//                 state, predictor and FP64 Tait pressure -> FP32 neighbour mirror
//   k_pair_compact fused neighbour search + W/gradW + continuity (both forms) +
//                 pressure gradient + artificial viscosity + gravity + wall
//                 springs, FP32 compensated sums, FP64 leapfrog corrector; a
//                 query with more than K neighbours in range records its K-th
//                 key and flags the step
//   k_pair<exact> returns at once unless the step was flagged; then applies the
//                 capped (K-nearest) neighbour sets of Search.cc:27-40.
#include "uph_kernels.h"
#include "../../include/uph.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include <mutex>
#include <atomic>
#include <algorithm>
#include <cmath>
#include <type_traits>

#define BLK 256
#define SCAN_ITEMS 4
#define SCAN_BLK 256
#define SCAN_TILE (SCAN_ITEMS * SCAN_BLK)
#define POS_PAD 128    // zeroed slots past the end of posA (read, never used, by k_pair_compact)
// Blocks of the capped-path launch: when it has work it walks every slot, so it
// fills the chip (4 blocks of 256 per CU on 256 CUs); when it has none it returns
// at once (measured no slower than a 512-block launch).
#define FIX_GRID 1024
// Candidates (slots of a query's 3x3 cells, or of one cell) beyond which a step
// gives up on the query and flags the state as broken: no physical particle distribution comes
// near it (a smoothing length of 6 lattice spacings with 4-fold coarsened cells
// gives about 21,000), while a run that has blown up piles its escaped particles
// into the clamped edge cells, where an exact search would cost O(N) per query.
#define CAND_MAX 65536

// ------------------------------------------------------------------ sort

// One block of the binning: slots blk * BLK + [0, BLK).
__device__ __forceinline__ void bin_block(int blk, int n, const double2* __restrict__ pos,
                                          const int* __restrict__ sceneOf, const DevScene* __restrict__ sc,
                                          const uint32_t* __restrict__ tagA, uint32_t* __restrict__ tagB,
                                          int* __restrict__ keys, int* __restrict__ rank, int* __restrict__ count) {
  const int i = blk * BLK + threadIdx.x;
  const bool valid = i < n;
  int key = -1;
  if (valid) {
    const DevScene& S = sc[sceneOf[i]];
    const double2 p = pos[i];
    int cx, cy;
    cell_of(S, p.x, p.y, cx, cy);
    key = S.coff + cy * S.ncx + cx;
    keys[i] = key;
    tagB[i] = tagA[i];
  }
  // Slots arrive sorted by the previous step's cells, so equal keys come in runs
  // of consecutive lanes: one atomic per run reserves the run's ranks.
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1, 64);
  const bool head = !valid || lane == 0 || prev != key;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = heads & ((2ull << lane) - 1ull);   // bits 0..lane
  const int rstart = 63 - __clzll(upto);
  const unsigned long long above = lane == 63 ? 0ull : (heads & (~0ull << (lane + 1)));
  const int rend = above ? __ffsll(above) - 1 : 64;
  int base = 0;
  if (valid && lane == rstart) base = atomicAdd(&count[key], rend - rstart);
  base = __shfl(base, rstart, 64);
  if (valid) rank[i] = base + (lane - rstart);
}

__global__ __launch_bounds__(BLK) void k_bin(int n, const double2* __restrict__ pos, const int* __restrict__ sceneOf,
                                             const DevScene* __restrict__ sc, const uint32_t* __restrict__ tagA,
                                             uint32_t* __restrict__ tagB, int* __restrict__ keys,
                                             int* __restrict__ rank, int* __restrict__ count) {
  bin_block(blockIdx.x, n, pos, sceneOf, sc, tagA, tagB, keys, rank, count);
}

// Block-wide exclusive scan of SCAN_TILE ints; returns the tile total.
__device__ int block_scan_tile(int* vals /* SCAN_ITEMS per thread, in registers via array */, int* lds) {
  const int tid = threadIdx.x;
  int sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int v = vals[k];
    vals[k] = sum;
    sum += v;
  }
  // wavefront inclusive scan of thread sums
  const int lane = tid & 63, wid = tid >> 6;
  int x = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[wid] = x;
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int w = 0; w < SCAN_BLK / 64; ++w) {
      const int t = lds[w];
      lds[w] = acc;
      acc += t;
    }
    lds[SCAN_BLK / 64] = acc;
  }
  __syncthreads();
  const int excl = x - sum + lds[wid];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) vals[k] += excl;
  const int total = lds[SCAN_BLK / 64];
  __syncthreads();
  return total;
}

// Per-tile totals of `in`.  With `step_ctl` it also opens a step: itime += 1,
// the overflow flag cleared, the K-th key stamp advanced.
// Opens a step: itime += 1, the overflow flag cleared, the K-th key stamp advanced.
__device__ __forceinline__ void open_step(int* step_ctl) {
  step_ctl[0] += 1;   // itime of the step being computed
  step_ctl[1] = 0;    // overflow flag of this step
  step_ctl[3] += 1;   // stamp of this step's K-th keys
}

// The total of tile `tile` into partial[tile] (lds: SCAN_BLK / 64 + 1 ints).
__device__ __forceinline__ void scan_reduce_tile(int tile, int m, const int* __restrict__ in, int* __restrict__ partial,
                                                 int* lds) {
  const long base = (long)tile * SCAN_TILE + threadIdx.x;
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const long i = base + (long)k * SCAN_BLK;
    if (i < m) s += in[i];
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if (lane == 0) lds[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < SCAN_BLK / 64; ++w) t += lds[w];
    partial[tile] = t;
  }
  __syncthreads();
}

__global__ __launch_bounds__(SCAN_BLK) void k_scan_reduce(int m, const int* __restrict__ in, int* __restrict__ partial,
                                                          int* __restrict__ step_ctl) {
  __shared__ int lds[SCAN_BLK / 64 + 1];
  if (step_ctl && blockIdx.x == 0 && threadIdx.x == 0) open_step(step_ctl);
  scan_reduce_tile(blockIdx.x, m, in, partial, lds);
}

// Exclusive scan of `count` into `start` (m+1 entries), zeroing `count`.  The
// tile goes through LDS (padded by one word per 16) so global accesses are
// coalesced and each thread scans 16 consecutive entries without bank conflicts.
__device__ __forceinline__ int tile_ix(int e) { return e + (e >> 4); }

// Tile `tb` of the exclusive scan (tile: SCAN_TILE + SCAN_TILE / 16 ints of LDS, lds:
// SCAN_BLK / 64 + 1 ints).
__device__ __forceinline__ void scan_down_tile(int tb, int m, int* __restrict__ count, const int* __restrict__ partial,
                                               int* __restrict__ start, int* tile, int* lds) {
  const long base = (long)tb * SCAN_TILE;
  const int tid = threadIdx.x;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int e = k * SCAN_BLK + tid;
    const long i = base + e;
    int v = 0;
    if (i < m) {
      v = count[i];
      count[i] = 0;
    }
    tile[tile_ix(e)] = v;
  }
  __syncthreads();
  int v[SCAN_ITEMS];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) v[k] = tile[tile_ix(tid * SCAN_ITEMS + k)];
  block_scan_tile(v, lds);
  // offset of this tile: the sum of the preceding tiles' totals
  int pre = 0;
  for (int t = tid; t < tb; t += SCAN_BLK) pre += partial[t];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) pre += __shfl_down(pre, o, 64);
  if ((tid & 63) == 0) lds[tid >> 6] = pre;
  __syncthreads();
  int off = 0;
#pragma unroll
  for (int w = 0; w < SCAN_BLK / 64; ++w) off += lds[w];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int e = tid * SCAN_ITEMS + k;
    const int orig = tile[tile_ix(e)];
    const int ex = v[k] + off;
    if (base + e == (long)m - 1) start[m] = ex + orig;
    tile[tile_ix(e)] = ex;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int e = k * SCAN_BLK + tid;
    const long i = base + e;
    if (i < m) start[i] = tile[tile_ix(e)];
  }
  __syncthreads();
}

__global__ __launch_bounds__(SCAN_BLK) void k_scan_down(int m, int* __restrict__ count, const int* __restrict__ partial,
                                                        int* __restrict__ start) {
  __shared__ int tile[SCAN_TILE + SCAN_TILE / 16];
  __shared__ int lds[SCAN_BLK / 64 + 1];
  scan_down_tile(blockIdx.x, m, count, partial, start, tile, lds);
}

__global__ __launch_bounds__(BLK) void k_scatter(int n, const int* __restrict__ keys, const int* __restrict__ rank,
                                                 const int* __restrict__ cellStart, int* __restrict__ tmpIdx) {
  const int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  tmpIdx[cellStart[keys[i]] + rank[i]] = i;
}

// The K-th smallest (d2, id) key of a capped query and the step stamp it is valid
// for, read by the pair admission in one 16-byte load.
struct __align__(16) KthRec {
  double d2;
  uint32_t id;
  int step;
};

struct StepBufs {
  // A: sorted input of the pair kernel
  double2 *posA, *velA;
  double* rhoA;
  uint32_t* tagA;
  float4 *mA, *mB;
  // B: pair-kernel output, same slot order as A
  double2 *posB, *velB;
  double* rhoB;
  float4* rtB;
  uint32_t* tagB;
  // sort scratch
  int *keys, *rank, *tmpIdx, *cellStart, *cellCount, *partial;
  int* sceneOf;
  // capped search: K-th key of overflowing queries, valid when its step equals the
  // step stamp ctl[3] (a counter that no restart resets, so a stamp never repeats),
  // or its negation when the per-lane search has also left the query's in-range
  // candidates in kmask (six 64-bit words: row r's slots rb[r] + [0, 64) in word 2r,
  // + [64, 128) in word 2r + 1; word u of slot q at kmask[u * kmask_n + q], so that a
  // wavefront's stores and loads of one word are contiguous)
  KthRec* kth;
  unsigned long long* kmask;
  int kmask_n;
  // per original index (global id)
  double2 *deckPos, *deckVel, *deckLvel;
  double* massD;   // FP64 mass per original index (diagnostics)
  int* wallKind;   // motion kind of each boundary particle (wall_class), per original index
  double* wallRef; // its gate height or sloshing-bottom offset
  // UPH_MODE_REF: DynamicPressure (A: at the start of the step, B: after it), its
  // deck value per original index, the predicting pressure rate and the density and
  // pressure fluctuations of the last step {dpdt_approx, drho, dP, 0}, the velocity
  // fluctuation, and the neighbour mirror {localDensity, localPressure}
  double *presA, *presB, *deckP;
  double *deckRl, *deckPl;   // localDensity, localPressure of the first step, per original index
  float4* rt2B;
  float2 *dvB, *mC;
  // control: [0] itime, [1] overflow flag of the step, [2] overflowed steps, [3] step stamp,
  // [4] broken state (a query with more than CAND_MAX candidates; sticky)
  int* ctl;
  float4* dbgRates;   // keep_rates: per slot {R, R~, a~x, a~y} (FP32), {ax, ay} (FP64);
                      // UPH_MODE_REF adds {dpdt, dpdt_approx, drho, dP}, {dvx, dvy, 0, 0}
  // mirror ("virtual") particles, allocated when some scene has them (gtot > 0, the
  // most mirror rows a step can make; wtot words of section-test bits): the bits the
  // sort sets (cleared again by k_ghost_count), the step's copy of them, their
  // popcounts and exclusive scan (gwstart[w] + the set bits below a row's in word w is
  // its mirror index; gwstart[wtot] is the step's total) with the scan's partials; the
  // slots of the fluid rows that have mirror rows and their number; per mirror index
  // its position, key id, scene, cell and in-cell rank; the mirror rows sorted by cell
  // (gcellStart per cell, ncells + 1); the fields derivatives.c copies into a mirror
  // row; the predicted fields of the listed rows, per A slot, from which they are copied
  int gtot, wtot;
  uint32_t *gbits, *gbits2;
  int *gwcnt, *gwstart, *gpart, *gcand, *gcandN;
  double2* gpos;
  uint32_t* gid;
  int *gscene, *gkey, *grank, *gcellCount, *gcellStart;
  double2* gposS;
  uint32_t* gidS;
  double2 *gvel, *glvel;
  double *grho, *grl, *gP, *gPl, *gmass;
  double2 *predV, *predVl;
  double *predRho, *predRl, *predP, *predPl;
};

// The section tests of fluid row `id` (global) at position x in this step
// (derivatives.c:3193-4192): each passing test sets the row's bit in its section's
// words; a row with any mirror row is listed (slot d).  Returns whether it was.
__device__ __forceinline__ bool mark_mirrors(const DevScene& S, const StepBufs& b, int itime, uint32_t id, double2 x,
                                             int d) {
  double lbp, rbp;
  boundary_lines(S, S.t0 + itime * S.dt, lbp, rbp);
  const int i = (int)id - S.poff;
  bool any = false;
  for (int k = 0; k < S.nsec; ++k) {
    double2 g, v = make_double2(0.0, 0.0), lv = v;
    if (mirror_row(S.sec[k], lbp, rbp, S.flat_bottom, S.depth, S.dx, S.h, x, g, v, lv)) {
      atomicOr(&b.gbits[S.gw + k * S.nwords + (i >> 5)], 1u << (i & 31));
      any = true;
    }
  }
  // one atomic per wavefront for the list slots of its listed rows (the lanes here
  // are the wavefront's fluid rows of scenes with mirror rows)
  const unsigned long long m = __ballot(any);
  if (any) {
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(b.gcandN, __popcll(m));
    base = __shfl(base, leader, 64);
    const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    b.gcand[base + below] = d;
  }
  return any;
}

// Final slot = cell start + number of same-cell particles with a smaller
// original index (deterministic order); gather the state and build the
// predictor mirror (sph.cc:93-126: rho += dt/2 Rt, v += dt/2 at,
// vloc = v + dt/2 at; Tait pressure of the predicted density).  UPH_MODE_REF
// also predicts DynamicPressure (P += dt/2 dpdt_approx) and adds the last step's
// fluctuations to the local fields: localDensity = rho + drho, localPressure =
// P + dP, localVelocity = v + dt/2 at + eps dv; at the first step they are
// deckRl, deckPl (rho0 and 0 from uph_create, input.c:109-110) and the deck's
// localVelocity.
// MIRRORS: some scene has mirror rows (the section tests run here); without them the
// kernel is built without that code and keeps its smaller register allocation.
template <bool MIRRORS>
__device__ __forceinline__ void reorder_slot(int d, int n, const StepBufs& b, const DevScene* __restrict__ sc,
                                             int search_only) {
  if (d >= n) return;
  const int i = b.tmpIdx[d];
  const int c = b.keys[i];
  const int c0 = b.cellStart[c], c1 = b.cellStart[c + 1];
  const uint32_t tg = b.tagB[i];
  const uint32_t my = tag_id(tg);
  int dst = d;   // the provisional slot: a cell beyond CAND_MAX marks a broken state
  if (c1 - c0 <= CAND_MAX) {
    int r = 0;
    for (int t = c0; t < c1; ++t) r += (tag_id(b.tagB[b.tmpIdx[t]]) < my) ? 1 : 0;
    dst = c0 + r;
  } else if (!search_only) {
    atomicOr(&b.ctl[4], 1);
  }
  const double2 p = b.posB[i];
  b.posA[dst] = p;
  b.tagA[dst] = tg;
  if (search_only) return;
  const DevScene& S = sc[b.sceneOf[i]];
  const double2 v = b.velB[i];
  const double rho = b.rhoB[i];
  const float4 rt = b.rtB[i];
  b.velA[dst] = v;
  b.rhoA[dst] = rho;
  const int itime = b.ctl[0];
  const double dt = S.dt;
  double vx = v.x, vy = v.y, lx, ly, rp = rho;
  if (tag_type(tg) == 1) {
    if (itime != 1) {
      rp = rho + 0.5 * dt * (double)rt.x;
      vx = v.x + 0.5 * dt * (double)rt.y;
      vy = v.y + 0.5 * dt * (double)rt.z;
      lx = vx + 0.5 * dt * (double)rt.y;
      ly = vy + 0.5 * dt * (double)rt.z;
    } else {
      const double2 l0 = b.deckLvel[my];
      lx = l0.x;
      ly = l0.y;
    }
  } else {
    const double2 l0 = b.deckLvel[my];
    lx = l0.x;
    ly = l0.y;
    if (S.motion != 0) {
      double2 xx = p, vv = v;
      double lvx;
      if (wall_kin(S, b.wallKind[my], b.wallRef[my], S.t0 + itime * dt, b.deckPos[my], xx, vv, &lvx)) lx = lvx;
    }
  }
  if (S.mode == 1) {
    // walls: DynamicPressure from the deck, localDensity rho0, localPressure 0
    // (inputbp.c:117-150, 289-290)
    double P = tag_type(tg) == 1 ? b.presB[i] : b.deckP[my];
    double Pp = P, rl = S.rho0, pl = 0.0;
    if (tag_type(tg) == 1 && itime == 1) {
      rl = b.deckRl[my];
      pl = b.deckPl[my];
    } else if (tag_type(tg) == 1) {
      const float4 r2 = b.rt2B[i];
      const float2 dv = b.dvB[i];
      Pp = P + 0.5 * dt * (double)r2.x;
      rl = rp + (double)r2.y;
      pl = Pp + (double)r2.z;
      lx += S.eps * (double)dv.x;
      ly += S.eps * (double)dv.y;
    }
    // densities enter the mirror as deviations from rho0 (formed in FP64), so the
    // differences the fluctuations and diffusion terms take keep FP32's relative
    // precision instead of losing the leading digits of rho0
    b.presA[dst] = P;
    b.mA[dst] = make_float4((float)vx, (float)vy, (float)lx, (float)ly);
    b.mB[dst] = make_float4((float)(rp - S.rho0), (float)Pp, rt.w, __uint_as_float(tg));
    b.mC[dst] = make_float2((float)(rl - S.rho0), (float)pl);
    if (MIRRORS && S.nsec > 0 && tag_type(tg) == 1 && mark_mirrors(S, b, itime, my, p, dst)) {
      b.predV[dst] = make_double2(vx, vy);   // the fields this row's mirror rows copy
      b.predVl[dst] = make_double2(lx, ly);
      b.predRho[dst] = rp;
      b.predRl[dst] = rl;
      b.predP[dst] = Pp;
      b.predPl[dst] = pl;
    }
    return;
  }
  const double P = tait(S, rp);
  const double pr = P / (rp * rp);
  b.mA[dst] = make_float4((float)vx, (float)vy, (float)lx, (float)ly);
  b.mB[dst] = make_float4((float)rp, (float)pr, rt.w, __uint_as_float(tg));
  if (MIRRORS && S.nsec > 0 && tag_type(tg) == 1 && mark_mirrors(S, b, itime, my, p, dst)) {
    b.predV[dst] = make_double2(vx, vy);
    b.predVl[dst] = make_double2(lx, ly);
    b.predRho[dst] = rp;
    b.predRl[dst] = rp;   // config mode keeps localDensity equal to density
    b.predP[dst] = P;
    b.predPl[dst] = 0.0;
  }
}

template <bool MIRRORS>
__global__ __launch_bounds__(BLK) void k_reorder(int n, StepBufs b, const DevScene* __restrict__ sc, int search_only) {
  reorder_slot<MIRRORS>(blockIdx.x * BLK + threadIdx.x, n, b, sc, search_only);
}

// ------------------------------------------------------ mirror particles
//
// derivatives.c:3193-4192 appends, after the particle rows, the mirror image of
// every fluid row that passes a section's test, section by section and in index
// order; the search ranks these rows among the K nearest (derivatives.c:1069-1085)
// and the pair filter drops them (Search.cc:33).  Here they are rebuilt each step:
// the sort (k_reorder) sets one bit per passing test, in words per section and 32
// fluid rows, and lists the rows with any; a scan over the words' popcounts gives
// each set bit its mirror index in the reference's order; the listed rows write
// their mirror rows by index; the mirror rows get a cell list of their own, which
// the pair kernels read only to count candidates in range and to rank keys.

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

// The mirror rows of the listed fluid rows, at their indices in the reference's
// order: section k's rows come after those of sections < k, and within a section
// in fluid index order (the set bits below the row's).
__device__ void emit_mirrors(const StepBufs& b, const DevScene* __restrict__ sc, int d) {
  const uint32_t tg = b.tagA[d];
  const int s = b.sceneOf[d];
  const DevScene& S = sc[s];
  const int i = (int)tag_id(tg) - S.poff;
  double lbp, rbp;
  boundary_lines(S, S.t0 + b.ctl[0] * S.dt, lbp, rbp);
  const double2 x = b.posA[d];
  const int first = b.gwstart[S.gw];
  const uint32_t below = (1u << (i & 31)) - 1u;
  for (int k = 0; k < S.nsec; ++k) {
    const int w = S.gw + k * S.nwords + (i >> 5);
    const uint32_t bits = b.gbits2[w];
    if (!((bits >> (i & 31)) & 1u)) continue;
    const int j = b.gwstart[w] + __popc(bits & below);
    double2 g = x, v = b.predV[d], lv = b.predVl[d];
    (void)mirror_row(S.sec[k], lbp, rbp, S.flat_bottom, S.depth, S.dx, S.h, x, g, v, lv);
    b.gpos[j] = g;
    b.gid[j] = (uint32_t)(S.gbase + (j - first));
    b.gscene[j] = s;
    b.gvel[j] = v;
    b.glvel[j] = lv;
    b.grho[j] = b.predRho[d];
    b.grl[j] = b.predRl[d];
    b.gP[j] = b.predP[d];
    b.gPl[j] = b.predPl[d];
    b.gmass[j] = b.massD[tag_id(tg)];
  }
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

// ------------------------------------------------------------- pair kernel

// Blocks are spread over the 8 XCDs round-robin by the dispatcher; remap so
// each XCD works on one contiguous run of slots (shared neighbour rows stay in
// that XCD's L2).  Grid size is a multiple of 8.
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x;
  const int per = nb >> 3;
  const int b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}

// Count of in-range candidates (d2 < cutoff, not itself) of slot p and, when
// it exceeds K, the K-th smallest (d2, id) key: the neighbour set of
// Search.cc:27-40 is then the in-range candidates with key <= that key.
__device__ void kth_key(const DevScene& S, int p, const double2* __restrict__ pos, const uint32_t* __restrict__ tag,
                        const int* __restrict__ cellStart, int& cnt, double& kd2, uint32_t& kid,
                        const StepBufs* gb = nullptr) {
  const double2 pp = pos[p];
  const double cut = S.cutoff;
  int c = 0;
  for_candidates(S, pp, cellStart, [&](int q) {
    if (q != p && dist2(pp, pos[q]) < cut) ++c;
  });
  if (gb)   // mirror rows rank too
    for_candidates(S, pp, gb->gcellStart, [&](int t) {
      if (dist2(pp, gb->gposS[t]) < cut) ++c;
    });
  cnt = c;
  kd2 = INFINITY;
  kid = 0xffffffffu;
  if (c <= S.K) return;
  // K-th smallest by repeated minimum above the previous threshold
  double td = -1.0;
  uint32_t ti = 0;
  for (int k = 0; k < S.K; ++k) {
    double bd = INFINITY;
    uint32_t bi = 0xffffffffu;
    auto take = [&](double d2, uint32_t id) {
      const bool above = d2 > td || (d2 == td && id > ti);
      if (above && (d2 < bd || (d2 == bd && id < bi))) {
        bd = d2;
        bi = id;
      }
    };
    for_candidates(S, pp, cellStart, [&](int q) {
      if (q == p) return;
      const double d2 = dist2(pp, pos[q]);
      if (d2 < cut) take(d2, tag_id(tag[q]));
    });
    if (gb)
      for_candidates(S, pp, gb->gcellStart, [&](int t) {
        const double d2 = dist2(pp, gb->gposS[t]);
        if (d2 < cut) take(d2, gb->gidS[t]);
      });
    td = bd;
    ti = bi;
  }
  kd2 = td;
  kid = ti;
}

// ------------------------------------------- K-th key, one wavefront per query
//
// A query with more than K candidates in range needs its K-th smallest (d2, id)
// key.  The 64 lanes of a wavefront find it together: the candidate slots are
// spread over the lanes (consecutive slots, so the loads coalesce), the keys in
// range are compacted into a per-wavefront LDS list, and each lane counts, for
// the keys it owns, how many listed keys are smaller; the key of rank K - 1 is
// the K-th.  Every candidate is read once, where a lane working alone re-reads
// all of them once per rank.

#define KTH_CAP 256   // keys in range one wavefront can list; beyond it one lane searches alone
#ifndef GATHER_CHUNKS
#define GATHER_CHUNKS 6   // 64-slot chunks of candidates loaded per pass
#endif

#define LANE_BUCKETS 16   // d2 buckets of the per-lane K-th key search
#define LANE_CAP 12       // candidates the bucket holding rank K - 1 may have

// A wavefront's listed keys for the wavefront K-th key search (3 KB).
struct KthLds {
  double d2[KTH_CAP];
  uint32_t id[KTH_CAP];
};
// The per-lane search's bucket counts (two 16-bit counts a word) and the slots of
// each lane's selected bucket, lane-major (5 KB), sharing a wavefront's LDS with its
// key list.  Only handles whose queries are expected to be capped build it (a
// larger per-block LDS made the uncapped kernel 2 % slower); packing it into the
// list's 3 KB (8-bit counts, 16-bit row indices) measured slower on capped steps
// (0.80 against 0.70 ms at 1 M).
struct LaneSel {
  uint32_t cnt[LANE_BUCKETS / 2][64];
  uint32_t slot[LANE_CAP][64];
};
union KthLaneLds {
  KthLds k;
  LaneSel l;
};

// LDS written by some lanes of a wavefront and then read by others: the
// wavefront's LDS operations complete in issue order; this keeps the compiler
// from moving them across.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lanes_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// (d2, id) < (e2, jd)
__device__ __forceinline__ bool key_lt(double d2, uint32_t id, double e2, uint32_t jd) {
  return d2 < e2 || (d2 == e2 && id < jd);
}

// The three candidate row ranges of a position: slots [rb[r], rb[r] + rl[r]).
__device__ __forceinline__ void candidate_rows(const DevScene& S, double2 x, const int* __restrict__ cellStart,
                                               int* rb, int* rl) {
  int cx, cy;
  cell_of(S, x.x, x.y, cx, cy);
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx < S.ncx - 1 ? cx + 1 : S.ncx - 1;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int row = cy - 1 + r;
    rb[r] = 0;
    rl[r] = 0;
    if (row >= 0 && row < S.ncy) {
      const int base = S.coff + row * S.ncx;
      rb[r] = cellStart[base + x0];
      rl[r] = cellStart[base + x1 + 1] - rb[r];
    }
  }
}

// Lists the in-range candidates (d2 < cutoff, not q) of slot q at xq, row ranges
// rb/rl, in L; all arguments are the same on every lane.  Returns how many are in
// range (only the first KTH_CAP are listed), or -1 for more than CAND_MAX
// candidates (nothing listed).
__device__ int wave_list(const DevScene& S, int q, double2 xq, const int* rb, const int* rl,
                         const double2* __restrict__ pos, const uint32_t* __restrict__ tag, KthLds& L) {
  const int lane = threadIdx.x & 63;
  const int total = rl[0] + rl[1] + rl[2];
  if (total > CAND_MAX) return -1;
  int n = 0;
  // up to 6 chunks of 64 candidates per pass, all loads issued before any is used
  // (a chunk past the end is skipped; slots past the end of the last chunk load
  // the query's own row, unused)
#pragma unroll 1
  for (int c0 = 0; c0 < total; c0 += 64 * GATHER_CHUNKS) {
    double2 xs[GATHER_CHUNKS];
    uint32_t ts[GATHER_CHUNKS];
    int sl[GATHER_CHUNKS];
#pragma unroll
    for (int u = 0; u < GATHER_CHUNKS; ++u) {
      if (c0 + 64 * u >= total) break;
      const int c = c0 + 64 * u + lane;
      sl[u] = c >= total ? q
              : c < rl[0] ? rb[0] + c
              : c < rl[0] + rl[1] ? rb[1] + (c - rl[0]) : rb[2] + (c - rl[0] - rl[1]);
      xs[u] = pos[sl[u]];
      ts[u] = tag[sl[u]];
    }
#pragma unroll
    for (int u = 0; u < GATHER_CHUNKS; ++u) {
      if (c0 + 64 * u >= total) break;
      const double d2 = dist2(xq, xs[u]);
      const bool in = sl[u] != q && d2 < S.cutoff;
      const unsigned long long m = __ballot(in);
      const int at = n + lanes_below(m);
      if (in && at < KTH_CAP) {
        L.d2[at] = d2;
        L.id[at] = tag_id(ts[u]);
      }
      n += __popcll(m);
    }
  }
  wave_lds_sync();
  return n;
}

// wave_list for slot q, its position and rows read here.
__device__ int wave_gather(const DevScene& S, int q, const double2* __restrict__ pos,
                           const uint32_t* __restrict__ tag, const int* __restrict__ cellStart, KthLds& L) {
  const double2 xq = pos[q];
  int rb[3], rl[3];
  candidate_rows(S, xq, cellStart, rb, rl);
  return wave_list(S, q, xq, rb, rl, pos, tag, L);
}

// Mirror rows in range (d2 < cutoff) of a query at x; CAND_MAX + 1 when its cells
// hold more than CAND_MAX of them.
__device__ int ghost_count(const DevScene& S, double2 x, const StepBufs& b) {
  int rb[3], rl[3];
  candidate_rows(S, x, b.gcellStart, rb, rl);
  if (rl[0] + rl[1] + rl[2] > CAND_MAX) return CAND_MAX + 1;
  int c = 0;
#pragma unroll 1
  for (int r = 0; r < 3; ++r)
#pragma unroll 1
    for (int t = rb[r]; t < rb[r] + rl[r]; ++t) c += dist2(x, b.gposS[t]) < S.cutoff ? 1 : 0;
  return c;
}

// Appends the mirror rows in range of a query at xq (the same on every lane) to the
// wavefront's list L after its first n entries: the new count (entries past KTH_CAP
// counted, not listed), or -1 beyond CAND_MAX candidates.
__device__ int wave_list_ghosts(const DevScene& S, double2 xq, int n, const StepBufs& b, KthLds& L) {
  const int lane = threadIdx.x & 63;
  int rb[3], rl[3];
  candidate_rows(S, xq, b.gcellStart, rb, rl);
  const int total = rl[0] + rl[1] + rl[2];
  if (total > CAND_MAX) return -1;
#pragma unroll 1
  for (int c0 = 0; c0 < total; c0 += 64) {
    const int c = c0 + lane;
    bool in = false;
    double d2 = 0.0;
    uint32_t id = 0u;
    if (c < total) {
      const int t = c < rl[0] ? rb[0] + c : c < rl[0] + rl[1] ? rb[1] + (c - rl[0]) : rb[2] + (c - rl[0] - rl[1]);
      d2 = dist2(xq, b.gposS[t]);
      in = d2 < S.cutoff;
      id = b.gidS[t];
    }
    const unsigned long long m = __ballot(in);
    const int at = n + lanes_below(m);
    if (in && at < KTH_CAP) {
      L.d2[at] = d2;
      L.id[at] = id;
    }
    n += __popcll(m);
  }
  wave_lds_sync();
  return n;
}

// 1 if (d2, id) < (e2, jd), else 0, for d2, e2 >= 0 given by their bit patterns
// (which order as integers): the borrow out of the 96-bit difference.
__device__ __forceinline__ int key_lt_bits(uint64_t d2, uint32_t id, uint64_t e2, uint32_t jd) {
  const int64_t t = (int64_t)d2 - (int64_t)e2 - (int64_t)(id < jd ? 1 : 0);
  return (int)((uint64_t)t >> 63);
}

// Value of lane l (the same l on every lane) on every lane.
__device__ __forceinline__ int bcast(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ uint64_t bcast64(uint64_t v, int l) {
  return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(v >> 32), l) << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
}
__device__ __forceinline__ double bcastd(double v, int l) {
  return __longlong_as_double((long long)bcast64((uint64_t)__double_as_longlong(v), l));
}

// The key of rank K - 1 among the n <= 64 NS listed keys (n > K), on every lane,
// by quickselect over the wavefront: a lane owns keys lane + 64 s, s < NS; each
// round takes one still-possible key as the pivot, counts the keys below it (one
// comparison per owned key and a ballot), and keeps the side that holds rank
// K - 1.  Every round removes at least the pivot, so it ends within n rounds;
// the pivot position is drawn from a fixed sequence, so the result and the work
// are the same on every run.
template <int NS>
__device__ void wave_kth_listed(const KthLds& L, int n, int K, uint32_t seed, double& kd2, uint32_t& kid) {
  const int lane = threadIdx.x & 63;
  uint64_t kd[NS];
  uint32_t ki[NS];
  bool act[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int j = lane + 64 * s;
    act[s] = j < n;
    kd[s] = act[s] ? (uint64_t)__double_as_longlong(L.d2[j]) : 0ull;
    ki[s] = act[s] ? L.id[j] : 0u;
  }
  int below = 0;   // keys ranked below every key still possible
  uint32_t rnd = seed * 2654435761u + 1u;
#pragma unroll 1
  for (int round = 0; round < n; ++round) {
    unsigned long long am[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) am[s] = __ballot(act[s]);
    // pivot: the first possible key at or after a drawn (slot, lane) position
    rnd = rnd * 1664525u + 1013904223u;
    const int s0 = (int)(rnd >> 28) & (NS - 1), l0 = (int)(rnd >> 22) & 63;
    int ps = -1, pl = 0;
#pragma unroll
    for (int t = 0; t < NS; ++t) {
      const int s = (s0 + t) & (NS - 1);
      unsigned long long msk = 0ull;
#pragma unroll
      for (int u = 0; u < NS; ++u)
        if (u == s) msk = am[u];
      if (ps < 0 && msk) {
        const unsigned long long r = l0 ? ((msk >> l0) | (msk << (64 - l0))) : msk;
        pl = (__ffsll((long long)r) - 1 + l0) & 63;
        ps = s;
      }
    }
    uint64_t pd = 0ull;
    uint32_t pi = 0u;
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (s == ps) {
        pd = bcast64(kd[s], pl);
        pi = (uint32_t)bcast((int)ki[s], pl);
      }
    bool lt[NS];
    int c = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      lt[s] = act[s] && key_lt_bits(kd[s], ki[s], pd, pi);
      c += __popcll(__ballot(lt[s]));
    }
    const int r = below + c;   // rank of the pivot
    if (r == K - 1) {
      kd2 = __longlong_as_double((long long)pd);
      kid = pi;
      break;
    }
    if (r > K - 1) {
#pragma unroll
      for (int s = 0; s < NS; ++s) act[s] = lt[s];
    } else {
      below = r + 1;
#pragma unroll
      for (int s = 0; s < NS; ++s) act[s] = act[s] && !lt[s] && ki[s] != pi;
    }
  }
  wave_lds_sync();
}

// K-th key of slot q on every lane (+inf when at most K are in range), from its
// n in-range candidates listed in L.  n = -1 (more than CAND_MAX candidates):
// +inf, unless `exact`, when one lane searches the candidates alone.
__device__ void wave_kth_of_list(const DevScene& S, int q, const double2* __restrict__ pos,
                                 const uint32_t* __restrict__ tag, const int* __restrict__ cellStart, KthLds& L,
                                 int n, double& kd2, uint32_t& kid, bool exact = false,
                                 const StepBufs* gb = nullptr) {
  kd2 = INFINITY;
  kid = 0xffffffffu;
  if (n < 0 && !exact) return;
  if (n >= 0 && n <= S.K) return;
  if (n >= 0 && n <= KTH_CAP) {
    // as few owned-key slots as the list needs (the count is the same on every lane)
    if (n <= 128)
      wave_kth_listed<2>(L, n, S.K, (uint32_t)q, kd2, kid);
    else
      wave_kth_listed<KTH_CAP / 64>(L, n, S.K, (uint32_t)q, kd2, kid);
    return;
  }
  // more in range than the list holds: one lane searches alone
  if ((threadIdx.x & 63) == 0) {
    int c;
    kth_key(S, q, pos, tag, cellStart, c, kd2, kid, gb);
  }
  kd2 = bcastd(kd2, 0);
  kid = (uint32_t)bcast((int)kid, 0);
}

// K-th key of slot q on every lane; n: the number in range (-1: more than
// CAND_MAX candidates, see wave_kth_of_list for `exact`).
__device__ void wave_kth(const DevScene& S, int q, const double2* __restrict__ pos, const uint32_t* __restrict__ tag,
                         const int* __restrict__ cellStart, KthLds& L, int& n, double& kd2, uint32_t& kid,
                         bool exact = false) {
  n = wave_gather(S, q, pos, tag, cellStart, L);
  wave_kth_of_list(S, q, pos, tag, cellStart, L, n, kd2, kid, exact);
}

// K-th key of query q (slot) at xq, candidate rows rb/rl, found by this lane alone:
// a histogram of the in-range d2 over LANE_BUCKETS equal buckets of [0, cutoff)
// finds the bucket that holds rank K - 1 (bucketing is monotone in d2, so every key
// of a lower bucket is smaller); its slots are collected and the rank taken among
// them by full (d2, id) keys.  Returns false, leaving the query to the wavefront
// search, when that bucket holds more than LANE_CAP or a row is longer than 128
// slots.  +inf when at most K are in range.
__device__ bool lane_kth(const DevScene& S, int q, double2 xq, const int* rb, const int* rl,
                         const double2* __restrict__ pos, const uint32_t* __restrict__ tag, LaneSel& L,
                         double& kd2, uint32_t& kid, unsigned long long* __restrict__ mask_out,
                         int mask_stride) {
  const int lane = threadIdx.x & 63;
  kd2 = INFINITY;
  kid = 0xffffffffu;
  // rows of up to 128 slots: the in-range candidates are kept as bits, two 64-bit
  // words per row, so the collecting pass visits only them (16-bit counts hold every
  // bucket: at most 384 are in range)
  if (rl[0] > 128 || rl[1] > 128 || rl[2] > 128) return false;
  const double cut = S.cutoff;
  const double scale = (double)LANE_BUCKETS / cut;
#pragma unroll
  for (int w = 0; w < LANE_BUCKETS / 2; ++w) L.cnt[w][lane] = 0u;
  int n = 0;
  unsigned long long mk[6];   // row r: slots rb[r] + [0, 64) in mk[2r], + [64, 128) in mk[2r + 1]
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int w = 0; w < 2; ++w) {
      const int k0 = 64 * w;
      const int k1 = rl[r] < k0 + 64 ? rl[r] : k0 + 64;
      unsigned long long mm = 0ull;
      // four candidates per trip, their loads issued together (a slot past the row's
      // end repeats its last one and is not counted)
#pragma unroll 1
      for (int k = k0; k < k1; k += 4) {
        double2 x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = pos[rb[r] + (k + u < k1 ? k + u : k1 - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const double d2 = dist2(xq, x[u]);
          if (k + u >= k1 || rb[r] + k + u == q || !(d2 < cut)) continue;
          const int bk = min((int)(d2 * scale), LANE_BUCKETS - 1);
          // an LDS atomic whose result is unused (ds_add_u32): the next candidate
          // does not wait for this count to be read back, as a read-modify-write would
          atomicAdd(&L.cnt[bk >> 1][lane], (bk & 1) ? 0x10000u : 1u);
          ++n;
          mm |= 1ull << (k + u - k0);
        }
      }
      mk[2 * r + w] = mm;
    }
#pragma unroll
  for (int u = 0; u < 6; ++u) mask_out[(size_t)u * mask_stride] = mk[u];
  if (n <= S.K) return true;
  int below = 0, kb = 0, m = 0;
#pragma unroll 1
  for (kb = 0; kb < LANE_BUCKETS; ++kb) {
    const uint32_t w = L.cnt[kb >> 1][lane];
    m = (int)((kb & 1) ? (w >> 16) : (w & 0xffffu));
    if (below + m >= S.K) break;
    below += m;
  }
  if (m > LANE_CAP) return false;
  int j = 0;
#pragma unroll
  for (int u = 0; u < 6; ++u) {
    const int base = rb[u >> 1] + 64 * (u & 1);
    unsigned long long mm = mk[u];
#pragma unroll 1
    while (mm) {
      const int t = base + __ffsll((long long)mm) - 1;
      mm &= mm - 1ull;
      const double d2 = dist2(xq, pos[t]);
      if (min((int)(d2 * scale), LANE_BUCKETS - 1) == kb) L.slot[j++][lane] = (uint32_t)t;
    }
  }
  // rank K - 1 - below of the bucket: the minimum above the previous one, that many times
  double td = -1.0;
  uint32_t ti = 0u;
#pragma unroll 1
  for (int rr = 0; rr <= S.K - 1 - below; ++rr) {
    double bd = INFINITY;
    uint32_t bi = 0xffffffffu;
    // four members per trip, their loads issued together (a position past the
    // bucket's end repeats its last member and is not compared)
#pragma unroll 1
    for (int i = 0; i < m; i += 4) {
      double2 x[4];
      uint32_t g[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t t = L.slot[min(i + u, m - 1)][lane];
        x[u] = pos[t];
        g[u] = tag[t];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double d2 = dist2(xq, x[u]);
        const uint32_t id = tag_id(g[u]);
        const bool above = d2 > td || (d2 == td && id > ti);
        if (i + u < m && above && key_lt(d2, id, bd, bi)) {
          bd = d2;
          bi = id;
        }
      }
    }
    td = bd;
    ti = bi;
  }
  kd2 = td;
  kid = ti;
  return true;
}

// The overflowing queries of this wavefront (more than K candidates in range,
// or a row range the compacted kernel cannot hold; ovf set on their lanes): each
// gets its K-th key, stamped with the step, and the step's flag makes the exact
// kernel run.  Every lane of the wavefront calls this.  With ROWS, each lane
// passes its own scene index, position and candidate rows, and the wavefront takes
// them from the query's lane instead of reading them again.
template <bool ROWS>
__device__ void wave_overflow(const StepBufs& b, const DevScene* __restrict__ sc, int p, bool ovf, KthLds& L,
                              int sidx = 0, double2 xp = double2{}, const int* rbp = nullptr,
                              const int* rlp = nullptr, LaneSel* ls = nullptr) {
  unsigned long long m = __ballot(ovf);
  if (m == 0ull) return;
  const bool lead = (threadIdx.x & 63) == 0;
  if (lead) atomicOr(&b.ctl[1], 1);
  // many overflowing lanes (the reference's smoothing length caps every query):
  // each finds its own key at once; what it leaves goes to the wavefront search
  if (ROWS && ls != nullptr && b.kmask != nullptr && __popcll(m) >= 8) {
    bool done = false;
    if (ovf) {
      const DevScene& S = sc[sidx];
      double kd2;
      uint32_t kid;
      if (S.nsec == 0 && lane_kth(S, p, xp, rbp, rlp, b.posA, b.tagA, *ls, kd2, kid, b.kmask + p, b.kmask_n)) {
        b.kth[p] = KthRec{kd2, kid, -b.ctl[3]};   // negated: kmask holds its in-range candidates
        done = true;
      }
    }
    wave_lds_sync();   // the wavefront search below writes the same LDS
    m = __ballot(ovf && !done);
  }
#pragma unroll 1
  while (m) {
    const int l = __ffsll((long long)m) - 1;
    m &= m - 1ull;
    const int q = bcast(p, l);
    int n;
    double kd2;
    uint32_t kid;
    if (ROWS) {
      const DevScene& S = sc[bcast(sidx, l)];
      const double2 xq = make_double2(bcastd(xp.x, l), bcastd(xp.y, l));
      int rb[3], rl[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        rb[r] = bcast(rbp[r], l);
        rl[r] = bcast(rlp[r], l);
      }
      n = wave_list(S, q, xq, rb, rl, b.posA, b.tagA, L);
      if (S.nsec > 0 && n >= 0) n = wave_list_ghosts(S, xq, n, b, L);
      wave_kth_of_list(S, q, b.posA, b.tagA, b.cellStart, L, n, kd2, kid, false, S.nsec > 0 ? &b : nullptr);
    } else {
      const DevScene& S = sc[b.sceneOf[q]];
      const double2 xq = b.posA[q];
      n = wave_gather(S, q, b.posA, b.tagA, b.cellStart, L);
      if (S.nsec > 0 && n >= 0) n = wave_list_ghosts(S, xq, n, b, L);
      wave_kth_of_list(S, q, b.posA, b.tagA, b.cellStart, L, n, kd2, kid, false, S.nsec > 0 ? &b : nullptr);
    }
    if (lead) {
      b.kth[q] = KthRec{kd2, kid, b.ctl[3]};
      if (n < 0) atomicOr(&b.ctl[4], 1);
    }
  }
}

// K-th key of slot q in this step, or +inf when q did not overflow.
__device__ __forceinline__ void kth_of(const StepBufs& b, int q, double& kd2, uint32_t& kid) {
  const KthRec k = b.kth[q];
  if (k.step == b.ctl[3] || k.step == -b.ctl[3]) {
    kd2 = k.d2;
    kid = k.id;
  } else {
    kd2 = INFINITY;
    kid = 0xffffffffu;
  }
}

// Kernel factors and pair admission of one fluid particle p, shared by both schemes.
struct PairBase {
  uint32_t idp;
  bool pb_ok;
  int pb;               // ParticleBoundary as a global index
  float invh, h2eps, cav;
  float fa, fb, fw;     // cubic-spline gradient factor: fa + fb q (q < 1), fw (2 - q)^2 / r (1 <= q < 2)
  float k7, invh2;      // PolyOrderSeven gradient factor: k7 (1 - u/4)^7 P(u), u = r^2 / h^2
  float wa, w7;         // W factors: cubic alpha = 10/(7 pi h^2), PolyOrderSeven 2145/(1347584 pi h^2)

  __device__ __forceinline__ void init_base(const DevScene& S, uint32_t tg) {
    idp = tag_id(tg);
    pb_ok = (int)idp < S.pb_gid;
    const float h = S.fh, walpha = S.fwalpha;
    invh = S.finvh;
    h2eps = 0.01f * h * h;
    cav = -2.0f * S.fav * S.fc0 * h;
    fa = -3.0f * walpha * invh * invh;
    fb = 2.25f * walpha * invh * invh;
    fw = -0.75f * walpha * invh;
    k7 = S.fk7;
    invh2 = invh * invh;
    wa = walpha;
    w7 = S.fw7;
    pb = S.pb_gid;
  }

  // grad W = f (x_p - x_q) for the kernel KT at |x_p - x_q|^2 = r2, 1/|.| = rinv
  template <int KT>
  __device__ __forceinline__ float grad_factor(float r2, float r, float rinv) const {
#pragma clang fp contract(on)
    if (KT == UPH_KERNEL_POLY7) {
      // PolyOrderSevenDeconvolutionKernel (derivatives.c:786-823); its correction
      // terms carry the integer quotients -131/329 and 150882/1840097, both 0, so
      // grad W = -a0 (1 - u/4)^7 (2048 - 3584 u + 3712 u^2 - 2912 u^3 + 1912 u^4
      // - 1106 u^5 + 305 u^6) dx / h^2 with u = q^2, evaluated by Horner in u
      const float u = r2 * invh2;
      const float s = 1.0f - 0.25f * u;
      const float s2 = s * s;
      const float s7 = s2 * s2 * s2 * s;
      float pu = 305.0f;
      pu = pu * u - 1106.0f;
      pu = pu * u + 1912.0f;
      pu = pu * u - 2912.0f;
      pu = pu * u + 3712.0f;
      pu = pu * u - 3584.0f;
      pu = pu * u + 2048.0f;
      return k7 * s7 * pu;
    }
    // cubic spline gradient factor (derivatives.c:44-66); both branches evaluated
    // and selected (lanes of a wavefront take both anyway)
    const float qq = r * invh;
    const float t = 2.0f - qq;
    const float f_far = fw * t * t * rinv;
    const float f_near = fa + fb * qq;
    return qq < 1.0f ? f_near : f_far;
  }

  // W at |x_p - x_q|^2 = r2, |.| = r for the kernel KT
  template <int KT>
  __device__ __forceinline__ float w_value(float r2, float r) const {
#pragma clang fp contract(on)
    if (KT == UPH_KERNEL_POLY7) {
      // PolyOrderSevenConvolutionKernel (derivatives.c:767-784): a (1 - u/4)^8
      // (2048 - 3072 u + 2816 u^2 - 1984 u^3 + 1184 u^4 - 628 u^5 + 305 u^6)
      const float u = r2 * invh2;
      const float s = 1.0f - 0.25f * u;
      const float s2 = s * s, s4 = s2 * s2;
      float pu = 305.0f;
      pu = pu * u - 628.0f;
      pu = pu * u + 1184.0f;
      pu = pu * u - 1984.0f;
      pu = pu * u + 2816.0f;
      pu = pu * u - 3072.0f;
      pu = pu * u + 2048.0f;
      return w7 * (s4 * s4) * pu;
    }
    // cubic spline W (derivatives.c:44-66)
    const float qq = r * invh;
    const float t = 2.0f - qq;
    const float w_far = 0.25f * wa * t * t * t;
    const float w_near = wa * (1.0f + qq * qq * (-1.5f + 0.75f * qq));
    return qq < 1.0f ? w_near : w_far;
  }

  // Admission of an in-range pair by the rules of Search.cc:33 (original index
  // order, ParticleBoundary bound) and, when EXACT, the K-nearest cap.
  template <bool EXACT>
  __device__ __forceinline__ bool admit(const DevScene& S, uint32_t idq, double d2, double kpd2, uint32_t kpid,
                                        const StepBufs& b, int q) const {
    if (!EXACT) return idq > idp ? (int)idq < pb : pb_ok;   // a select, no branch
    if (idq > idp) {
      if ((int)idq >= S.pb_gid) return false;
      if (EXACT && !key_le(d2, idq, kpd2, kpid)) return false;
    } else {
      if (!pb_ok) return false;
      if (EXACT) {
        double kd2;
        uint32_t kid;
        kth_of(b, q, kd2, kid);
        if (!key_le(d2, idp, kd2, kid)) return false;
      }
    }
    return true;
  }

};

// Config-mode sums of one fluid particle p.
struct PairCtx : PairBase {
  float vpx, vpy, lpx, lpy, rhop, prp, mp, inv_mp;
  CSum sR, sRt, sAx, sAy;   // sAx/sAy: the predicting acceleration (dvdt_approx)
  float dAx = 0.f, dAy = 0.f;   // dvdt - dvdt_approx: wall pairs with localDensity != density

  __device__ __forceinline__ void init(const DevScene& S, uint32_t tg, float4 ap, float4 bp, float2) {
    init_base(S, tg);
    vpx = ap.x; vpy = ap.y; lpx = ap.z; lpy = ap.w;
    rhop = bp.x; prp = bp.y; mp = bp.z;
    inv_mp = 1.0f / mp;
  }

  // One admitted pair: x_p - x_q = (ex, ey), |.|^2 = d2 (FP64 geometry, FP32 sums).
  // FP32 a * b + c within one expression contracts to FMA here, and only there: a
  // fusion across statements depends on the code this body is inlined into, and the
  // kernels that inline it (compacted, direct, capped bit walk) must round alike.
  // 1/r and 1/x use v_rsq_f32 / v_rcp_f32.
  template <int KT>
  __device__ __forceinline__ void add(const DevScene& S, double ex, double ey, uint32_t tq, float4 aq, float4 bq,
                                      float2) {
#pragma clang fp contract(on)
    const bool p_is_i = tag_id(tq) > idp;   // pair (i=p, j=q) or (i=q, j=p)
    const float dx = (float)ex, dy = (float)ey;
    const float r2 = dx * dx + dy * dy;
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    const float f = grad_factor<KT>(r2, r, rinv);
    const float gx = f * dx, gy = f * dy;
    const float mq = bq.z;
    const uint32_t tyq = tag_type(tq);
    // m_q times localDensity/density of q: the factor is 1 for fluid (the leapfrog
    // keeps them equal) and rho0/rho for walls (inputbp.c:288 resets their
    // localDensity to rho0).  PressureGradientForce weighs P_p by that factor
    // (derivatives.c:1129-1131), the predicting form (derivatives.c:1353-1355) does
    // not: the difference is kept apart in dA.
    float mql = mq;
    if (tyq != 1) {
      const float lam = S.frho0 * __builtin_amdgcn_rcpf(bq.x);
      mql = mq * lam;
      const float w = -mq * prp * (lam - 1.0f);
      dAx += w * gx;
      dAy += w * gy;
    }
    // continuity (derivatives.c:1520-1567), symmetric in the pair order
    const float cont = (vpx - aq.z) * gx + (vpy - aq.w) * gy;
    sR.add(mql * cont);
    // predicting continuity (derivatives.c:1217-1252), depends on the pair order
    const float alt = (lpx - aq.x) * gx + (lpy - aq.y) * gy;
    float rt_q = mq * cont, rt_p = mp * alt;
    // both forms are live in every wavefront; keep them as values so the choice is a
    // select rather than two masked branches
    asm volatile("" : "+v"(rt_q), "+v"(rt_p));
    sRt.add(p_is_i ? rt_q : rt_p);
    // pressure gradient (derivatives.c:1109-1163) + Monaghan viscosity
    // -alpha c0 mu / rho_bar with mu = h v.r / (r^2 + 0.01 h^2), rho_bar = (rho_p + rho_q) / 2,
    // through one reciprocal of the product of the two denominators
    const float dvx = vpx - aq.x, dvy = vpy - aq.y;
    const float vr = dvx * dx + dvy * dy;
    float pij = prp + bq.y;
    if (vr < 0.0f) pij += cav * vr * __builtin_amdgcn_rcpf((r2 + h2eps) * (rhop + bq.x));
    float ax = -mq * pij * gx;
    float ay = -mq * pij * gy;
    // wall spring (derivatives.c:2396-2449).  The damping part does not vanish at
    // r0, so the range test is the reference's FP64 sqrt(d2) < r0.
    if (p_is_i && (tyq == 2 || tyq == 3 || tyq == 5) &&
        sqrt(__dadd_rn(__dmul_rn(ex, ex), __dmul_rn(ey, ey))) < S.r0) {
      const float kr = S.fk * (S.fr0 - r) * rinv * inv_mp;
      const float cr = S.fcd * vr * (rinv * rinv) * inv_mp;
      ax -= kr * dx + cr * dx;
      ay -= kr * dy + cr * dy;
    }
    sAx.add(ax);
    sAy.add(ay);
  }
};

// Ref-mode sums of one fluid particle p: the reference's run-phase rates in gather
// form (every neighbour q of p; terms whose reference form depends on which of p, q
// is the pair's first index branch on the original ids, as a5 does).
struct RefCtx : PairBase {
  float vpx, vpy, lpx, lpy;   // predicted velocity, localVelocity
  float rhop, Pp, mp, rlp;    // predicted density, DynamicPressure, mass, localDensity
  float drp, drlp, r0;        // density and localDensity minus rho0; rho0
  float inv_mp, irp, irp2, Kp, gk, kap, mu;
  CSum sR, sRt;               // ddensitydt, ddensitydt_approx
  CSum sL, sC2, sLt;          // dpdt diffusion part; sum m_q (v_p - v_q).grad W; dpdt_approx diffusion part
  CSum sAx, sAy, sTx, sTy;    // dvdt, dvdt_approx
  float dr = 0.f, dP = 0.f, dvx = 0.f, dvy = 0.f;   // fluctuations

  __device__ __forceinline__ void init(const DevScene& S, uint32_t tg, float4 ap, float4 bp, float2 cp) {
    init_base(S, tg);
    vpx = ap.x; vpy = ap.y; lpx = ap.z; lpy = ap.w;
    r0 = S.frho0;
    drp = bp.x; Pp = bp.y; mp = bp.z;
    drlp = cp.x;
    rhop = r0 + drp;
    rlp = r0 + drlp;
    inv_mp = 1.0f / mp;
    irp = 1.0f / rhop;
    irp2 = irp * irp;
    Kp = S.fK0 + S.fgamma * Pp;
    gk = S.fgk;
    kap = S.fkap;
    mu = S.fmu;
  }

  template <int KT>
  __device__ __forceinline__ void add(const DevScene& S, double ex, double ey, uint32_t tq, float4 aq, float4 bq,
                                      float2 cq) {
    // no contraction: the compacted and the direct kernel inline this body into
    // different code, and only unfused arithmetic rounds the same in both
#pragma clang fp contract(off)
    const bool p_is_i = tag_id(tq) > idp;   // pair (i=p, j=q) or (i=q, j=p)
    const float dx = (float)ex, dy = (float)ey;
    const float r2 = dx * dx + dy * dy;
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * rinv;
    const float f = grad_factor<KT>(r2, r, rinv);
    const float gx = f * dx, gy = f * dy;   // grad_p W_pq
    const float w = w_value<KT>(r2, r);
    const float drq = bq.x, Pq = bq.y, mq = bq.z;
    const float drlq = cq.x, Plq = cq.y;
    const float rhoq = r0 + drq, rlq = r0 + drlq;
    const float irq = __builtin_amdgcn_rcpf(rhoq);
    const float Vq = mq * irq;
    const uint32_t tyq = tag_type(tq);
    // DensityEquation (derivatives.c:1520-1567)
    const float cont = (vpx - aq.z) * gx + (vpy - aq.w) * gy;
    sR.add(mq * rlq * irq * cont);
    // PredictingDensityEquation (derivatives.c:1217-1252), depends on the pair order
    const float dvx_ = vpx - aq.x, dvy_ = vpy - aq.y;
    const float std_ = dvx_ * gx + dvy_ * gy;          // (v_p - v_q).grad W
    const float alt = (lpx - aq.x) * gx + (lpy - aq.y) * gy;
    float rt_q = mq * cont, rt_p = mp * alt;
    asm volatile("" : "+v"(rt_q), "+v"(rt_p));
    sRt.add(p_is_i ? rt_q : rt_p);
    // DynamicPressureEquation (:1569-1672): its compression term is K_p/rho_p times
    // the continuity sum (added at the end); the diffusion terms of P and rho
    const float ir2 = __builtin_amdgcn_rcpf(r2 + h2eps);
    const float nu = (dx * gx + dy * gy) * ir2;
    const float Kq = S.fK0 + S.fgamma * Pq;
    const float ep = 4.0f * Kp * __builtin_amdgcn_rcpf(Kp + Kq);
    const float rb = (rhop + rhoq) * irp * irq;
    sL.add(Vq * (gk * (Pp - Plq) * ep - kap * Kp * (drp - drlq) * rb) * nu);
    // PredictingDynamicPressureEquation (:1254-1331): the reference's volume of the
    // second particle of a pair is mass[j]/density[i]
    sC2.add(mq * std_);
    const float Vt = p_is_i ? Vq : mp * irq;
    sLt.add(Vt * (gk * (Pp - Pq) * ep - kap * Kp * (drp - drq) * rb) * nu);
    // Local{Density,Pressure}Fluctuation (:2100-2171); LocalVelocityFluctuation
    // (:2060-2098) adds both of its terms to the pair's first particle only
    dr += Vq * (drq - drp) * w;
    dP += Vq * (Pq - Pp) * w;
    if (p_is_i) {
      // m_p/rholoc_q - m_q/rholoc_p over the common denominator, its numerator
      // from the deviations
      const float num = r0 * (mp - mq) + (mp * drlp - mq * drlq);
      const float cv = num * __builtin_amdgcn_rcpf(rlp * rlq) * w;
      dvx += cv * dvx_;
      dvy += cv * dvy_;
    }
    // PressureGradientForce (:1109-1163) with DynamicPressure, its predicting form
    // (:1333-1377), and Monaghan viscosity when alpha_av > 0
    const float vr = dvx_ * dx + dvy_ * dy;
    float pij = Pp * rlq * irp2 * irq + Pq * rlp * irq * irq * irp;
    float ptij = Pp * irp2 + Pq * irq * irq;
    if (vr < 0.0f) {
      const float av = cav * vr * __builtin_amdgcn_rcpf((r2 + h2eps) * (rhop + rhoq));
      pij += av;
      ptij += av;
    }
    float ax = -mq * pij * gx, ay = -mq * pij * gy;
    float tx = -mq * ptij * gx, ty = -mq * ptij * gy;
    // ViscousForcesSingleStep (:1770-1839) and its predicting form (:1448-1515):
    // X = x_q - x_p, D = v_p - v_q, sigma = mu_eff (X D deviatoric) / (r^2 + 0.01 h^2)
    {
      const float X0 = -dx, X1 = -dy;
      const float eff = 2.0f * mu * rhop * rhoq * __builtin_amdgcn_rcpf(rhop + rhoq);
      const float s00 = eff * ir2 * (X0 * dvx_ - X1 * dvy_);
      const float s01 = eff * ir2 * (X0 * dvy_ + X1 * dvx_);
      const float sgx = s00 * gx + s01 * gy, sgy = s01 * gx - s00 * gy;
      // rholoc_q/(rho_p^2 rho_q) - rholoc_p/(rho_q^2 rho_p) = (rholoc_q rho_q -
      // rholoc_p rho_p) / (rho_p rho_q)^2, the numerator from the deviations
      const float num = r0 * ((drlq + drq) - (drlp + drp)) + (drlq * drq - drlp * drp);
      const float dA = num * irp2 * irq * irq;
      const float sg = p_is_i ? dA : -dA;
      ax += mq * sg * sgx;
      ay += mq * sg * sgy;
      const float st = mq * (irp2 + irq * irq);
      tx += st * sgx;
      ty += st * sgy;
    }
    // wall spring (derivatives.c:2396-2449, 1392-1446), into both accelerations
    if (p_is_i && (tyq == 2 || tyq == 3 || tyq == 5) &&
        sqrt(__dadd_rn(__dmul_rn(ex, ex), __dmul_rn(ey, ey))) < S.r0) {
      const float kr = S.fk * (S.fr0 - r) * rinv * inv_mp;
      const float cr = S.fcd * vr * (rinv * rinv) * inv_mp;
      const float sx = kr * dx + cr * dx, sy = kr * dy + cr * dy;
      ax -= sx;
      ay -= sy;
      tx -= sx;
      ty -= sy;
    }
    sAx.add(ax);
    sAy.add(ay);
    sTx.add(tx);
    sTy.add(ty);
  }
};

// The pair context of a scheme.
template <int MODE>
using CtxOf = typename std::conditional<MODE == UPH_MODE_REF, RefCtx, PairCtx>::type;

// The {localDensity, localPressure} mirror of slot q, read by the ref-mode scheme only.
template <int MODE>
__device__ __forceinline__ float2 mirror_c(const StepBufs& b, uint32_t q) {
  return MODE == UPH_MODE_REF ? ld_off(b.mC, q) : make_float2(0.f, 0.f);
}

// Boundary slot: wall state for the next step (deck + kinematics at t = (itime+1) dt).
__device__ __forceinline__ void wall_update(const DevScene& S, const StepBufs& b, int p, uint32_t tg, int itime,
                                            float mass) {
  const uint32_t id = tag_id(tg);
  double2 x = b.deckPos[id], v = b.deckVel[id];
  if (S.motion != 0) {
    double lvx;
    (void)wall_kin(S, b.wallKind[id], b.wallRef[id], S.t0 + (itime + 1) * S.dt, b.deckPos[id], x, v, &lvx);
  }
  b.posB[p] = x;
  b.velB[p] = v;
  b.rhoB[p] = b.rhoA[p];
  b.rtB[p] = make_float4(0.f, 0.f, 0.f, mass);
  if (S.mode == 1) b.presB[p] = b.deckP[id];
}

// Gravity and the FP64 leapfrog corrector (sph.cc:137-168) for fluid slot p.
__device__ __forceinline__ void correct(const DevScene& S, const StepBufs& b, int p, int itime, double2 xp,
                                        double2 vn, float mass, const PairCtx& c) {
  const double dt = S.dt;
  const float R = c.sR.val(), Rt = c.sRt.val();
  double atx = (double)c.sAx.val(), aty = (double)c.sAy.val();
  if (S.gravity == 1) aty -= S.g;    // BodyForce / PredictingBodyForce (derivatives.c:1167-1178, 1379-1390)
  const double ax = atx + (double)c.dAx, ay = aty + (double)c.dAy;
  const double rho_n = b.rhoA[p];
  double rho1, v1x, v1y;
  if (itime == 1) {
    rho1 = rho_n + 0.5 * dt * (double)R;
    v1x = vn.x + 0.5 * dt * ax;
    v1y = vn.y + 0.5 * dt * ay;
  } else {
    rho1 = rho_n + dt * (double)R;
    v1x = vn.x + dt * ax;
    v1y = vn.y + dt * ay;
  }
  b.posB[p] = make_double2(xp.x + dt * v1x, xp.y + dt * v1y);
  b.velB[p] = make_double2(v1x, v1y);
  b.rhoB[p] = rho1;
  b.rtB[p] = make_float4(Rt, (float)atx, (float)aty, mass);
  if (b.dbgRates) {
    // {R, R~, a~ (as the predictor reads them)} then the FP64 acceleration the
    // corrector used, in the same 32 bytes
    b.dbgRates[2 * p] = make_float4(R, Rt, (float)atx, (float)aty);
    *reinterpret_cast<double2*>(b.dbgRates + 2 * p + 1) = make_double2(ax, ay);
  }
}

// UPH_MODE_REF: gravity, the sponge layer and the FP64 leapfrog corrector
// (sph.cc:137-168) with DynamicPressure integrated like the density.
__device__ __forceinline__ void correct(const DevScene& S, const StepBufs& b, int p, int itime, double2 xp,
                                        double2 vn, float mass, const RefCtx& c) {
  const double dt = S.dt;
  const float R = c.sR.val(), Rt = c.sRt.val();
  const float RP = c.Kp * c.irp * R + c.sL.val();      // compression + diffusion
  const float RPt = c.Kp * c.irp * c.sC2.val() + c.sLt.val();
  double ax = (double)c.sAx.val(), ay = (double)c.sAy.val();
  double atx = (double)c.sTx.val(), aty = (double)c.sTy.val();
  if (S.gravity == 1) {   // BodyForce / PredictingBodyForce (derivatives.c:1167-1178, 1379-1390)
    ay -= S.g;
    aty -= S.g;
  }
  // SpongeLayerWaveReflectionAbsorption (derivatives.c:2827-2838): spongeFactor0 = 0.2
  // beyond spongePosition0 = 7 (System.c:333-335) by the particle's initial x
  // (input.c:137-143), for the reference piston only
  if (S.motion == 1 && b.deckPos[tag_id(b.tagA[p])].x >= 7.0) ax -= 0.2 * (double)c.vpx;
  const double rho_n = b.rhoA[p], P_n = b.presA[p];
  const double hdt = itime == 1 ? 0.5 * dt : dt;
  const double rho1 = rho_n + hdt * (double)R;
  const double P1 = P_n + hdt * (double)RP;
  const double v1x = vn.x + hdt * ax, v1y = vn.y + hdt * ay;
  b.posB[p] = make_double2(xp.x + dt * v1x, xp.y + dt * v1y);
  b.velB[p] = make_double2(v1x, v1y);
  b.rhoB[p] = rho1;
  b.presB[p] = P1;
  b.rtB[p] = make_float4(Rt, (float)atx, (float)aty, mass);
  b.rt2B[p] = make_float4(RPt, c.dr, c.dP, 0.f);
  b.dvB[p] = make_float2(c.dvx, c.dvy);
  if (b.dbgRates) {
    b.dbgRates[4 * p] = make_float4(R, Rt, (float)atx, (float)aty);
    *reinterpret_cast<double2*>(b.dbgRates + 4 * p + 1) = make_double2(ax, ay);
    b.dbgRates[4 * p + 2] = make_float4(RP, RPt, c.dr, c.dP);
    b.dbgRates[4 * p + 3] = make_float4(c.dvx, c.dvy, 0.f, 0.f);
  }
}

// Direct form: one pass over the candidates, physics inside the distance test.
// EXACT applies the K-nearest cap (run only when some query overflowed).
// Returns true for a query with more than K candidates in range when !EXACT
// (nothing written; its K-th key is left to wave_overflow).
template <bool EXACT, int KT, int MODE>
__device__ __forceinline__ bool pair_direct(int p, const StepBufs& b, const DevScene* __restrict__ sc) {
  const uint32_t tg = b.tagA[p];
  const DevScene& S = sc[b.sceneOf[p]];
  const int itime = b.ctl[0];
  const double2 xp = b.posA[p];
  const double2 vn = b.velA[p];
  const float mass = b.mB[p].z;
  if (tag_type(tg) != 1) {
    wall_update(S, b, p, tg, itime, mass);
    return false;
  }
  {
    int rb[3], rl[3];
    candidate_rows(S, xp, b.cellStart, rb, rl);
    if (rl[0] + rl[1] + rl[2] > CAND_MAX) {   // a broken state: flag it, skip the query
      atomicOr(&b.ctl[4], 1);
      return false;
    }
  }
  CtxOf<MODE> c;
  c.init(S, tg, b.mA[p], b.mB[p], mirror_c<MODE>(b, (uint32_t)p));
  const double cut = S.cutoff;
  double kpd2 = INFINITY;
  uint32_t kpid = 0xffffffffu;
  if (EXACT) kth_of(b, p, kpd2, kpid);
  int cnt = 0;
  for_candidates(S, xp, b.cellStart, [&](int q) {
    if (q == p) return;
    const double2 xq = b.posA[q];
    const double ex = xp.x - xq.x, ey = xp.y - xq.y;
    const double d2 = __dadd_rn(__dmul_rn(ex, ex), __dmul_rn(ey, ey));
    if (!(d2 < cut)) return;
    ++cnt;
    const float4 bq = b.mB[q];
    const uint32_t tq = __float_as_uint(bq.w);
    if (!c.template admit<EXACT>(S, tag_id(tq), d2, kpd2, kpid, b, q)) return;
    c.template add<KT>(S, ex, ey, tq, b.mA[q], bq, mirror_c<MODE>(b, (uint32_t)q));
  });
  if (!EXACT && S.nsec > 0 && cnt <= S.K) cnt += ghost_count(S, xp, b);
  if (!EXACT && cnt > S.K) return true;
  correct(S, b, p, itime, xp, vn, mass, c);
  return false;
}

// The pair sums of fluid particle p over the set bits of m[0..2] (slot bases
// rb[0..2]), in that order, into c.
template <int KT, int MODE>
__device__ __forceinline__ void walk_pairs(const DevScene& S, const StepBufs& b, int p, double2 xp, CtxOf<MODE>& c,
                                           const int* rb, const unsigned long long* m) {
  // the set bits of the three rows as one stream, advanced without branches;
  // -1 when exhausted
  int rr = 0;
  unsigned long long cm = m[0];
  int cb = rb[0];
  auto next = [&]() -> int {
    const bool a0 = (cm == 0ull) & (rr == 0);
    cm = a0 ? m[1] : cm;
    cb = a0 ? rb[1] : cb;
    rr = a0 ? 1 : rr;
    const bool a1 = (cm == 0ull) & (rr == 1);
    cm = a1 ? m[2] : cm;
    cb = a1 ? rb[2] : cb;
    rr = a1 ? 2 : rr;
    const bool none = cm == 0ull;
    const int k = __ffsll((long long)cm) - 1;
    cm &= cm - 1ull;
    return none ? -1 : cb + k;
  };
  // two register sets in ping-pong: the next pair's loads are in flight while the
  // current pair's physics runs (an exhausted stream loads slot p, unused)
  struct Nb {
    double2 x;
    float4 a, b;
    float2 c;
  };
  auto load = [&](Nb& o, int q) {
    const uint32_t s = (uint32_t)(q >= 0 ? q : p);
    o.x = ld_off(b.posA, s);
    o.a = ld_off(b.mA, s);
    o.b = ld_off(b.mB, s);
    o.c = mirror_c<MODE>(b, s);
  };
  auto pair = [&](const Nb& o) {
    const uint32_t tq = __float_as_uint(o.b.w);
    if (c.template admit<false>(S, tag_id(tq), 0.0, 0.0, 0u, b, 0))
      c.template add<KT>(S, xp.x - o.x.x, xp.y - o.x.y, tq, o.a, o.b, o.c);
  };
  Nb n0, n1;
  int q0 = next();
  load(n0, q0);
#pragma unroll 1
  while (q0 >= 0) {
    const int q1 = next();
    load(n1, q1);
    pair(n0);
    if (q1 < 0) break;
    q0 = next();
    load(n0, q0);
    pair(n1);
  }
}

// Pass 2 of k_pair_compact for one fluid particle: the pair sums over the set bits
// of m[0..2] (row starts rb[0..2]) and the corrector.
template <int KT, int MODE>
__device__ __forceinline__ void compact_pairs(const DevScene& S, const StepBufs& b, int p, uint32_t tg, int itime,
                                              double2 xp, double2 vn, float mass, const int* rb,
                                              const unsigned long long* m) {
  CtxOf<MODE> c;
  c.init(S, tg, b.mA[p], b.mB[p], mirror_c<MODE>(b, (uint32_t)p));
  walk_pairs<KT, MODE>(S, b, p, xp, c, rb, m);
  correct(S, b, p, itime, xp, vn, mass, c);
}

// The capped path's pairs of fluid slot p by the same bit walk: pass 1 marks, in two
// 64-bit masks per candidate row (rows of up to 128 slots), the candidates in range
// that the K cap keeps; pass 2 walks them row by row, in slot order as the direct
// loop sums them.
// Returns false, leaving p to the direct loop, when a row is longer than 128 slots.
template <int KT, int MODE>
__device__ __forceinline__ bool exact_pairs(int p, const StepBufs& b, const DevScene* __restrict__ sc) {
  const uint32_t tg = b.tagA[p];
  const DevScene& S = sc[b.sceneOf[p]];
  const double2 xp = b.posA[p];
  int rb[3], rl[3];
  candidate_rows(S, xp, b.cellStart, rb, rl);
  if (rl[0] > 128 || rl[1] > 128 || rl[2] > 128) return false;
  const int itime = b.ctl[0];
  const double2 vn = b.velA[p];
  const float mass = b.mB[p].z;
  CtxOf<MODE> c;
  c.init(S, tg, b.mA[p], b.mB[p], mirror_c<MODE>(b, (uint32_t)p));
  double kpd2;
  uint32_t kpid;
  kth_of(b, p, kpd2, kpid);
  const double cut = S.cutoff;
  const uint32_t idp = tag_id(tg);
  const bool pb_ok = (int)idp < S.pb_gid;
  const int stamp = b.ctl[3];
  const bool have_mask = b.kmask != nullptr && b.kth[p].step == -stamp;
  // Row r's admitted candidates as bits of m[2r] (slots rb[r] + [0, 64)) and m[2r + 1]
  // (slots rb[r] + [64, 128)), by the rules of Search.cc:27-40 (admit<true>).  Pass 1a
  // tests distances and decides the candidates q with a larger original index, whose
  // admission needs only p's K-th key (q ranks within p's K nearest, q below
  // ParticleBoundary); it marks as pending those with a smaller one, whose admission
  // needs q's K-th key (p ranks within q's K nearest).  Pass 1b reads those keys, two
  // candidates in flight.
  unsigned long long m[6];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double2* rp = b.posA + rb[r];
    const uint32_t* tp = b.tagA + rb[r];
#pragma unroll
    for (int w = 0; w < 2; ++w) {
      const int k0 = 64 * w;
      const int k1 = rl[r] < k0 + 64 ? rl[r] : k0 + 64;
      unsigned long long am = 0ull, pm = 0ull;
      if (have_mask) {
        // the per-lane K-th search left this word's in-range candidates: four per trip
        unsigned long long mm = b.kmask[(size_t)(2 * r + w) * b.kmask_n + p];
#pragma unroll 1
        while (mm) {
          int kk[4];
          double2 x[4];
          uint32_t g[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            kk[u] = mm ? __ffsll((long long)mm) - 1 : -1;
            mm &= mm - 1ull;
            const int qq = kk[u] >= 0 ? rb[r] + k0 + kk[u] : p;
            x[u] = b.posA[qq];
            g[u] = b.tagA[qq];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const double d2 = dist2(xp, x[u]);
            const uint32_t idq = tag_id(g[u]);
            const bool in = kk[u] >= 0;
            const bool adm = in && idq > idp && (int)idq < S.pb_gid && key_le(d2, idq, kpd2, kpid);
            const bool pen = in && idq < idp && pb_ok;
            const int sh = kk[u] & 63;
            am |= (unsigned long long)adm << sh;
            pm |= (unsigned long long)pen << sh;
          }
        }
      } else {
      // four candidates per trip, their loads issued together (a slot past the row's
      // end repeats its last one and sets no bit)
#pragma unroll 1
      for (int k = k0; k < k1; k += 4) {
        double2 x[4];
        uint32_t g[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int kk = k + u < k1 ? k + u : k1 - 1;
          x[u] = rp[kk];
          g[u] = tp[kk];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const double d2 = dist2(xp, x[u]);
          const uint32_t idq = tag_id(g[u]);
          const bool in = k + u < k1 && d2 < cut;
          const bool adm = in && idq > idp && (int)idq < S.pb_gid && key_le(d2, idq, kpd2, kpid);
          const bool pen = in && idq < idp && pb_ok;   // the particle itself is neither
          const int sh = (k + u - k0) & 63;   // past the row's end both bits are 0
          am |= (unsigned long long)adm << sh;
          pm |= (unsigned long long)pen << sh;
        }
      }
      }
      const int base = rb[r] + k0;
      int ka = pm ? __ffsll((long long)pm) - 1 : -1;
      pm &= pm - 1ull;
      int qa = ka >= 0 ? base + ka : p;
      KthRec ra = b.kth[qa];
      double2 xa = b.posA[qa];
#pragma unroll 1
      while (ka >= 0) {
        const int kb = pm ? __ffsll((long long)pm) - 1 : -1;
        pm &= pm - 1ull;
        const int qb = kb >= 0 ? base + kb : p;
        const KthRec rk = b.kth[qb];
        const double2 xb = b.posA[qb];
        if ((ra.step != stamp && ra.step != -stamp) || key_le(dist2(xp, xa), idp, ra.d2, ra.id))
          am |= 1ull << ka;
        ka = kb;
        ra = rk;
        xa = xb;
      }
      m[2 * r + w] = am;
    }
  }
  const int rb0[3] = {rb[0], rb[0] + 64, rb[1]};
  const unsigned long long m0[3] = {m[0], m[1], m[2]};
  walk_pairs<KT, MODE>(S, b, p, xp, c, rb0, m0);
  const int rb1[3] = {rb[1] + 64, rb[2], rb[2] + 64};
  const unsigned long long m1[3] = {m[3], m[4], m[5]};
  walk_pairs<KT, MODE>(S, b, p, xp, c, rb1, m1);
  correct(S, b, p, itime, xp, vn, mass, c);
  return true;
}

// EXACT (the capped path) runs as a small grid-stride launch that returns at
// once unless the step's overflow flag is set.
template <bool EXACT, int KT, int MODE>
// Four wavefronts per SIMD (128 VGPRs): the configuration scheme needs 129, the
// reference scheme 179, which at this target spills 168 bytes per lane and still runs
// the capped step faster than with three wavefronts (no spill) or five (392 bytes).
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4)))
void k_pair(int n, StepBufs b, const DevScene* __restrict__ sc) {
  if (EXACT && (b.ctl[1] == 0 || b.ctl[4] != 0)) return;
  if (EXACT && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&b.ctl[2], 1);
  __shared__ KthLds lds[EXACT ? 1 : BLK / 64];
  const int nlog = (n + BLK - 1) / BLK;
  // the trip count is the same for the whole block, so every wavefront reaches
  // wave_overflow with all of its lanes
  for (int lb = xcd_block(); lb < nlog; lb += gridDim.x) {
    const int p = lb * BLK + threadIdx.x;
    bool ovf = false;
    if (p < n) {
      // the capped step's fluid rows take the bit walk when their rows fit it
      const bool walked = EXACT && tag_type(b.tagA[p]) == 1 && exact_pairs<KT, MODE>(p, b, sc);
      if (!walked) ovf = pair_direct<EXACT, KT, MODE>(p, b, sc);
    }
    if (!EXACT) wave_overflow<false>(b, sc, p, ovf, lds[threadIdx.x >> 6]);
  }
}

// Compacted form (default).  Pass 1 runs only the FP64 distance test over the
// three candidate row ranges, in a branch-free loop whose counter is uniform
// across the wavefront (each lane walks its own row range by pointer), and
// records in-range candidates as bits of one 64-bit mask per row.  Pass 2 walks
// the set bits, so every lane spends its physics on in-range pairs only, with the
// next pair's loads issued before the current pair's math.  A row range longer
// than 64 slots, or more than K candidates in range, raises the step's overflow
// flag and the exact direct kernel recomputes the step.  UNI: the wavefront takes
// its scene from its first lane, so the scene constants stay in scalar registers
// (101 VGPRs against 109: five wavefronts per SIMD instead of four); a wavefront
// whose slots belong to two scenes returns at once and k_pair_mixed runs it with
// the scene per lane.  ONE: the handle holds one scene (index 0, no check).  Every
// lane of the wavefront calls this.
template <int KT, int MODE, bool UNI, bool ONE = false>
__device__ __forceinline__ void compact_query(int n, int p, const StepBufs& b, const DevScene* __restrict__ sc,
                                              KthLds& lds, LaneSel* ls = nullptr) {
  const bool live = p < n;
  const int pp = live ? p : n - 1;
  const uint32_t tg = b.tagA[pp];
  const bool fluid = live && tag_type(tg) == 1;
  const int slane = ONE ? 0 : b.sceneOf[pp];
  const int sidx = ONE ? 0 : UNI ? __builtin_amdgcn_readfirstlane(slane) : slane;
  if (!ONE && UNI && __ballot(slane != sidx) != 0ull) return;   // a wavefront of two scenes
  const DevScene& S = sc[sidx];
  const int itime = b.ctl[0];
  const double2 xp = b.posA[pp];
  const double cut = S.cutoff;
  int cx, cy;
  cell_of(S, xp.x, xp.y, cx, cy);
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx < S.ncx - 1 ? cx + 1 : S.ncx - 1;
  int rb[3], rl[3], rlf[3];   // rlf: the row lengths before the 64-slot limit
  unsigned long long m[3];
  bool bad = false;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int row = cy - 1 + r;
    rb[r] = 0;
    rl[r] = 0;
    m[r] = 0ull;
    if (fluid && row >= 0 && row < S.ncy) {
      const int base = S.coff + row * S.ncx;
      rb[r] = b.cellStart[base + x0];
      rl[r] = b.cellStart[base + x1 + 1] - rb[r];
    }
    rlf[r] = rl[r];
    if (rl[r] > 64) {
      bad = true;
      rl[r] = 0;
    }
    // the trip count is the wave's maximum row length (uniform, in an SGPR); slots
    // past this lane's row end are tested too and their bits cleared below
    const int lmax = __builtin_amdgcn_readfirstlane(wave_max(rl[r]));
    const double2* rp = b.posA + rb[r];
    uint32_t mw[2] = {0u, 0u};
#pragma unroll
    for (int hw = 0; hw < 2; ++hw) {
      const int k0 = 32 * hw;
      const int kend = hw == 0 ? (lmax < 32 ? lmax : 32) : lmax;
      // candidate k0 + j enters as the low bit and is shifted up by each later one:
      // one add-with-carry per candidate; bit-reversed into place after the loop
      uint32_t bits = 0u;
#pragma unroll 1
      for (int k = k0; k < kend; k += 4) {
        double2 xs[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) xs[u] = rp[u];   // posA has POS_PAD zeroed slots past n
        rp += 4;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const double ex = xp.x - xs[u].x, ey = xp.y - xs[u].y;
          const double d2 = __dadd_rn(__dmul_rn(ex, ex), __dmul_rn(ey, ey));
          bits = bits + bits + (d2 < cut ? 1u : 0u);
        }
      }
      const int nb = (kend - k0 + 3) & ~3;   // candidates tested in this half (uniform)
      if (nb > 0) mw[hw] = __builtin_bitreverse32(bits) >> (32 - nb);
    }
    const int l = rl[r];
    mw[0] &= l >= 32 ? 0xffffffffu : ((1u << l) - 1u);
    mw[1] &= l >= 64 ? 0xffffffffu : (l > 32 ? ((1u << (l - 32)) - 1u) : 0u);
    m[r] = ((unsigned long long)mw[1] << 32) | mw[0];
  }
  // every lane reaches wave_overflow: the work of a lane is a call of this
  bool ovf = false;
  auto run = [&]() {
    if (!live) return;
    const double2 vn = b.velA[p];
    const float mass = b.mB[p].z;
    if (!fluid) {
      wall_update(S, b, p, tg, itime, mass);
      return;
    }
    // drop the particle itself (always in range, in its own row)
    const int self = p - rb[1];
    if (self >= 0 && self < 64) m[1] &= ~(1ull << self);
    int cnt = __popcll(m[0]) + __popcll(m[1]) + __popcll(m[2]);
    if (S.nsec > 0 && !bad && cnt <= S.K) cnt += ghost_count(S, xp, b);   // mirror rows rank too
    if (bad || cnt > S.K) {
      ovf = true;   // the exact kernel recomputes this step
      return;
    }
    compact_pairs<KT, MODE>(S, b, p, tg, itime, xp, vn, mass, rb, m);
  };
  run();
  wave_overflow<true>(b, sc, p, ovf, lds, sidx, xp, rb, rlf, ls);
}

// LANE: the handle's queries are expected to be capped, so the kernel carries the
// per-lane K-th key search and its LDS.
template <int PB, int KT, int MODE, bool ONE, bool LANE>
// Occupancy targets, each measured faster than the compiler's own choice: the
// reference scheme's sums need about 133 VGPRs, capped at 128 (4 wavefronts per SIMD
// instead of 3, a few spilled registers); with the per-lane K-th search (LANE) the
// configuration scheme runs 6 wavefronts per SIMD (80 VGPRs) instead of 5.  The
// default instance (91 VGPRs, 5 wavefronts) measured slower at 6.
__global__ __launch_bounds__(PB)
__attribute__((amdgpu_waves_per_eu(MODE == UPH_MODE_REF ? 4 : LANE ? 6 : 1)))
void k_pair_compact(int n, StepBufs b, const DevScene* __restrict__ sc) {
  const int p = xcd_block() * PB + threadIdx.x;
  if constexpr (LANE) {
    __shared__ KthLaneLds lds[PB / 64];
    compact_query<KT, MODE, true, ONE>(n, p, b, sc, lds[threadIdx.x >> 6].k, &lds[threadIdx.x >> 6].l);
  } else {
    __shared__ KthLds lds[PB / 64];
    compact_query<KT, MODE, true, ONE>(n, p, b, sc, lds[threadIdx.x >> 6]);
  }
}

// The wavefronts whose slots belong to two scenes (listed by uph_create), with the
// scene taken per lane: one wavefront per block.
template <int KT, int MODE>
__global__ __launch_bounds__(64) void k_pair_mixed(int n, StepBufs b, const DevScene* __restrict__ sc,
                                                   const int* __restrict__ waves) {
  __shared__ KthLds lds[1];
  compact_query<KT, MODE, false>(n, waves[blockIdx.x] * 64 + (int)threadIdx.x, b, sc, lds[0]);
}

// ---------------------------------------------------------------- search

// Pairs of query p that SearchNeighbors emits (Search.cc:27-40): in range,
// j > i, j < ParticleBoundary, key <= the K-th key (kd2, kid).
__device__ __forceinline__ bool search_keeps(const DevScene& S, uint32_t idp, double d2, uint32_t idq, double kd2,
                                             uint32_t kid) {
  return idq > idp && (int)idq < S.pb_gid && key_le(d2, idq, kd2, kid);
}

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
      const int ovf = __hip_atomic_load(&b.ctl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const int broken = __hip_atomic_load(&b.ctl[4], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (ovf != 0 && broken == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&b.ctl[2], 1);
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

// ================================================================== host

namespace {

struct Err {
  std::string msg;
};

#define HIPCHK(expr, errobj)                                                              \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      char b_[512];                                                                       \
      snprintf(b_, sizeof b_, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
               __LINE__);                                                                 \
      (errobj).msg = b_;                                                                  \
      return -1;                                                                          \
    }                                                                                     \
  } while (0)

template <class T>
int dalloc(T** p, size_t count, Err& err) {
  *p = nullptr;
  if (count == 0) count = 1;
  HIPCHK(hipMalloc((void**)p, count * sizeof(T)), err);
  return 0;
}

int grid_of(int n) { return (n + BLK - 1) / BLK; }
int grid8_of(int n) { return ((grid_of(n) + 7) / 8) * 8; }

struct CellScan {
  int m = 0, nb = 0;
};

}  // namespace

struct uph_handle {
  Err err;
  hipStream_t stream = nullptr;
  int n = 0, ncells = 0, nscenes = 0;
  int itime = 0;
  int kernel = UPH_KERNEL_CUBIC;   // smoothing kernel, the same for every scene
  int mode = UPH_MODE_CONFIG;      // scheme, the same for every scene
  std::vector<DevScene> hs;
  std::vector<int> fluid_n, tot_n;
  DevScene* dsc = nullptr;
  StepBufs b{};
  CellScan cs;
  uph_opts opts{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double* diag = nullptr;   // per-block partial sums of uph_diagnostics
  int diag_cap = 0;
  double* vort = nullptr;   // uph_vorticity's output, n_fluid of the largest scene asked for
  int vort_cap = 0;
  unsigned* bar = nullptr;  // grid barrier of uph_step_persistent: arrivals, generation, broken
  struct PersistArgs* pargs = nullptr;   // its arguments, in device memory
  // boundary rows' deck and kinematics (host copies, by global index), from which the
  // state getters place them at the last step's time
  std::vector<double2> deck_pos, deck_vel;
  std::vector<int> wall_kind;
  std::vector<double> wall_ref;
  int gfluid = 0;   // fluid rows of the scenes with mirror rows (the most rows the sort lists)
  // the wavefronts (64 consecutive slots) that hold two scenes' slots: the compacted pair
  // kernel leaves them to k_pair_mixed
  int nmixed = 0;
  int* mixedWaves = nullptr;
  // some scene's queries are expected to be capped (4 pi (h/dx)^2 in range, a lattice's
  // count, above 3/4 of K): the pair kernel carries the per-lane K-th key search
  bool capped = false;
};

static void fill_scene(DevScene& S, const uph_scene& u, double cell_factor) {
  memset(&S, 0, sizeof S);
  S.cutoff = u.h * u.h * 4;
  // a cell is at least the search radius, so the 3x3 stencil holds every pair
  const double cs = sqrt(S.cutoff) * (1.0 + 1e-9) * (cell_factor > 1.0 ? cell_factor : 1.0);
  S.cell_inv = 1.0 / cs;
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

// kmask (48 bytes per slot) is written and read only by the per-lane K-th search of
// handles whose queries are expected to be capped; other handles leave it null.
static int alloc_bufs(StepBufs& b, int n, int ncells, int nblocks_scan, Err& err, bool search_only,
                      bool capped = false) {
  memset(&b, 0, sizeof b);
  if (dalloc(&b.posA, (size_t)n + POS_PAD, err) || dalloc(&b.posB, n, err) || dalloc(&b.tagA, n, err) ||
      dalloc(&b.tagB, n, err) || dalloc(&b.keys, n, err) || dalloc(&b.rank, n, err) ||
      dalloc(&b.tmpIdx, n, err) || dalloc(&b.cellStart, (size_t)ncells + 1, err) ||
      dalloc(&b.cellCount, ncells, err) || dalloc(&b.partial, nblocks_scan, err) ||
      dalloc(&b.sceneOf, n, err) || dalloc(&b.kth, n, err) ||
      dalloc(&b.ctl, 8, err))
    return -1;
  if (search_only) return 0;
  if (dalloc(&b.velA, n, err) || dalloc(&b.velB, n, err) || dalloc(&b.rhoA, n, err) ||
      dalloc(&b.rhoB, n, err) || dalloc(&b.rtB, n, err) ||
      (capped && dalloc(&b.kmask, (size_t)n * 6, err)) ||
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

static void free_bufs(StepBufs& b) {
  void* ptrs[] = {b.posA, b.velA, b.rhoA, b.tagA, b.mA, b.mB, b.posB, b.velB, b.rhoB,
                  b.rtB, b.tagB, b.keys, b.rank, b.tmpIdx, b.cellStart, b.cellCount, b.partial,
                  b.sceneOf, b.kth, b.kmask, b.deckPos, b.deckVel, b.deckLvel, b.massD, b.ctl,
                  b.dbgRates, b.presA, b.presB, b.deckP, b.rt2B, b.dvB, b.mC, b.deckRl, b.deckPl,
                  b.wallKind, b.wallRef, b.gbits, b.gbits2, b.gwcnt, b.gwstart, b.gpart, b.gcand, b.gcandN, b.gpos, b.gid, b.gscene, b.gkey,
                  b.grank, b.gcellCount, b.gcellStart, b.gposS, b.gidS, b.gvel, b.glvel, b.grho, b.grl,
                  b.gP, b.gPl, b.gmass, b.predV, b.predVl, b.predRho, b.predRl, b.predP, b.predPl};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  memset(&b, 0, sizeof b);
}

// The sort kernels: bin, scan, scatter, reorder.
static int launch_sort(hipStream_t st, int n, int ncells, int nb, StepBufs& b, const DevScene* dsc, int search_only,
                       bool step_ctl, Err& err) {
  hipLaunchKernelGGL(k_bin, dim3(grid_of(n)), dim3(BLK), 0, st, n, b.posB, b.sceneOf, dsc, b.tagA, b.tagB, b.keys,
                     b.rank, b.cellCount);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_BLK), 0, st, ncells, b.cellCount, b.partial,
                     step_ctl ? b.ctl : (int*)nullptr);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(SCAN_BLK), 0, st, ncells, b.cellCount, b.partial, b.cellStart);
  hipLaunchKernelGGL(k_scatter, dim3(grid_of(n)), dim3(BLK), 0, st, n, b.keys, b.rank, b.cellStart, b.tmpIdx);
  if (b.gtot > 0 && !search_only)
    hipLaunchKernelGGL(k_reorder<true>, dim3(grid_of(n)), dim3(BLK), 0, st, n, b, dsc, search_only);
  else
    hipLaunchKernelGGL(k_reorder<false>, dim3(grid_of(n)), dim3(BLK), 0, st, n, b, dsc, search_only);
  HIPCHK(hipGetLastError(), err);
  return 0;
}

// The pair kernels of one smoothing kernel KT and scheme MODE: the selected pass,
// then the exact capped fixup (which returns at once unless the step overflowed).
template <int KT, int MODE>
static int launch_pairs(uph_handle* h, hipStream_t st, hipEvent_t pb, hipEvent_t pc) {
  if (h->opts.pair_variant == 1) {
    hipLaunchKernelGGL((k_pair<false, KT, MODE>), dim3(grid8_of(h->n)), dim3(BLK), 0, st, h->n, h->b, h->dsc);
  } else {
    const int pbk = h->opts.pair_block;
    const int g = (((h->n + pbk - 1) / pbk + 7) / 8) * 8;
    const bool one = h->nscenes == 1, lane = h->capped;
#define UPH_COMPACT(PBV, ONEV, LANEV)                                                                        \
  hipLaunchKernelGGL((k_pair_compact<PBV, KT, MODE, ONEV, LANEV>), dim3(g), dim3(PBV), 0, st, h->n, h->b, h->dsc)
#define UPH_COMPACT_PB(ONEV, LANEV)              \
  do {                                           \
    if (pbk == 128) UPH_COMPACT(128, ONEV, LANEV); \
    else if (pbk == 512) UPH_COMPACT(512, ONEV, LANEV); \
    else UPH_COMPACT(256, ONEV, LANEV);          \
  } while (0)
    if (one && lane) UPH_COMPACT_PB(true, true);
    else if (one) UPH_COMPACT_PB(true, false);
    else if (lane) UPH_COMPACT_PB(false, true);
    else UPH_COMPACT_PB(false, false);
#undef UPH_COMPACT_PB
#undef UPH_COMPACT
    if (h->nmixed > 0)
      hipLaunchKernelGGL((k_pair_mixed<KT, MODE>), dim3(h->nmixed), dim3(64), 0, st, h->n, h->b, h->dsc,
                         h->mixedWaves);
  }
  if (pb) HIPCHK(hipEventRecord(pb, st), h->err);
  const int gfix = std::min(grid8_of(h->n), FIX_GRID);
  hipLaunchKernelGGL((k_pair<true, KT, MODE>), dim3(gfix), dim3(BLK), 0, st, h->n, h->b, h->dsc);
  if (pc) HIPCHK(hipEventRecord(pc, st), h->err);
  return 0;
}

// The mirror rows of this step and their cell list (scenes with a virtualType that
// has sections only; the sort has set the test bits and listed the rows): counts,
// scan, rows, bin, scan, scatter.
static int launch_ghosts(uph_handle* h, hipStream_t st) {
  StepBufs& b = h->b;
  const int wnb = (b.wtot + SCAN_TILE - 1) / SCAN_TILE;
  hipLaunchKernelGGL(k_ghost_count, dim3(grid_of(b.wtot)), dim3(BLK), 0, st, b);
  hipLaunchKernelGGL(k_scan_reduce, dim3(wnb), dim3(SCAN_BLK), 0, st, b.wtot, b.gwcnt, b.gpart, (int*)nullptr);
  hipLaunchKernelGGL(k_scan_down, dim3(wnb), dim3(SCAN_BLK), 0, st, b.wtot, b.gwcnt, b.gpart, b.gwstart);
  hipLaunchKernelGGL(k_ghost_emit, dim3(std::min(grid_of(h->gfluid), 256)), dim3(BLK), 0, st, b, h->dsc);
  hipLaunchKernelGGL(k_ghost_bin, dim3(grid_of(b.gtot)), dim3(BLK), 0, st, b, h->dsc);
  hipLaunchKernelGGL(k_scan_reduce, dim3(h->cs.nb), dim3(SCAN_BLK), 0, st, h->ncells, b.gcellCount, b.partial,
                     (int*)nullptr);
  hipLaunchKernelGGL(k_scan_down, dim3(h->cs.nb), dim3(SCAN_BLK), 0, st, h->ncells, b.gcellCount, b.partial,
                     b.gcellStart);
  hipLaunchKernelGGL(k_ghost_scatter, dim3(grid_of(b.gtot)), dim3(BLK), 0, st, b);
  HIPCHK(hipGetLastError(), h->err);
  return 0;
}

// pa, pb, pc (may be null): events recorded after the sort, after the compacted pair
// pass and after the capped pass.
static int launch_step(uph_handle* h, hipEvent_t pa, hipEvent_t pb, hipEvent_t pc = nullptr) {
  hipStream_t st = h->stream;
  if (launch_sort(st, h->n, h->ncells, h->cs.nb, h->b, h->dsc, 0, true, h->err)) return -1;
  if (h->b.gtot > 0 && launch_ghosts(h, st)) return -1;
  if (pa) HIPCHK(hipEventRecord(pa, st), h->err);
  int rc;
  if (h->mode == UPH_MODE_REF)
    rc = h->kernel == UPH_KERNEL_POLY7 ? launch_pairs<UPH_KERNEL_POLY7, UPH_MODE_REF>(h, st, pb, pc)
                                       : launch_pairs<UPH_KERNEL_CUBIC, UPH_MODE_REF>(h, st, pb, pc);
  else
    rc = h->kernel == UPH_KERNEL_POLY7 ? launch_pairs<UPH_KERNEL_POLY7, UPH_MODE_CONFIG>(h, st, pb, pc)
                                       : launch_pairs<UPH_KERNEL_CUBIC, UPH_MODE_CONFIG>(h, st, pb, pc);
  if (rc != 0) return -1;
  HIPCHK(hipGetLastError(), h->err);
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
  long n = 0, nc = 0, gtot = 0, wtot = 0, gfluid = 0;
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
    fill_scene(h->hs[s], u, h->opts.cell_factor);
    DevScene& S = h->hs[s];
    S.poff = (int)n;
    S.coff = (int)nc;
    S.pb_gid = (int)n + (u.n_fluid + u.n_fixed + u.n_moving - 1);
    // mirror rows: one test per section and fluid row, ids after the scene's rows
    if (4.0 * M_PI * (u.h / u.dx) * (u.h / u.dx) > 0.75 * u.K) h->capped = true;
    S.nsec = u.n_fluid > 0 ? mirror_sections(u.vtype, S.sec) : 0;
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
  int ctl[8] = {0, 0, 0, 0, 0, 0, 0, 0};
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
      if (launch_step(h, nullptr, nullptr)) return -1;
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
  if (h->b.gtot > 0 || h->capped || h->opts.pair_variant != 0) {
    h->err.msg = "uph_step_persistent: for handles without mirror rows, not expected to be capped, with the "
                 "compacted pair kernel (use uph_step)";
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
  int ctl[8];
  HIPCHK(hipMemcpy(ctl, h->b.ctl, sizeof ctl, hipMemcpyDeviceToHost), h->err);
  if (ctl[4]) {
    h->err.msg = "uph_sync: a particle had more than " + std::to_string(CAND_MAX) +
                 " candidates in its 3x3 cells: the state has blown up (particles escaped or collapsed "
                 "into one cell); the step's results are not valid";
    return -1;
  }
  return ctl[2];
}

extern "C" int uph_itime(const uph_t* h) { return h ? h->itime : -1; }

extern "C" int uph_time_steps(uph_t* h, int n_steps, double* ms_total, double* ms_pair) {
  if (!h || n_steps <= 0) return -1;
  std::vector<hipEvent_t> ev(2 * n_steps);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e), h->err);
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  HIPCHK(hipEventRecord(h->ev0, h->stream), h->err);
  for (int k = 0; k < n_steps; ++k)
    if (launch_step(h, ev[2 * k], ev[2 * k + 1])) return -1;
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
      if (launch_step(h, nullptr, nullptr, nullptr)) rc = -1;
    } else if (hipEventRecord(ev[4 * k], h->stream) != hipSuccess ||
               launch_step(h, ev[4 * k + 1], ev[4 * k + 2], ev[4 * k + 3])) {
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
  HIPCHK(hipMemcpy(p.data(), h->b.posB + S.poff, S.np * sizeof(double2), hipMemcpyDeviceToHost), h->err);
  HIPCHK(hipMemcpy(v.data(), h->b.velB + S.poff, S.np * sizeof(double2), hipMemcpyDeviceToHost), h->err);
  HIPCHK(hipMemcpy(r.data(), h->b.rhoB + S.poff, S.np * sizeof(double), hipMemcpyDeviceToHost), h->err);
  HIPCHK(hipMemcpy(t.data(), h->b.tagA + S.poff, S.np * sizeof(uint32_t), hipMemcpyDeviceToHost), h->err);
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
static int get_rates_f64(uph_t* h, int scene, float* R, float* Rt, double* a, float* at) {
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
  HIPCHK(hipMemcpy(d.data(), h->b.dbgRates + st * (size_t)S.poff, st * S.np * sizeof(float4),
                   hipMemcpyDeviceToHost),
         h->err);
  HIPCHK(hipMemcpy(t.data(), h->b.tagA + S.poff, S.np * sizeof(uint32_t), hipMemcpyDeviceToHost), h->err);
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
  HIPCHK(hipMemcpy(d.data(), h->b.dbgRates + 4 * (size_t)S.poff, 4 * S.np * sizeof(float4), hipMemcpyDeviceToHost),
         h->err);
  HIPCHK(hipMemcpy(t.data(), h->b.tagA + S.poff, S.np * sizeof(uint32_t), hipMemcpyDeviceToHost), h->err);
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
  HIPCHK(hipMemcpy(v.data(), (h->mode == UPH_MODE_REF ? h->b.presB : h->b.rhoB) + S.poff, S.np * sizeof(double),
                   hipMemcpyDeviceToHost),
         h->err);
  HIPCHK(hipMemcpy(t.data(), h->b.tagA + S.poff, S.np * sizeof(uint32_t), hipMemcpyDeviceToHost), h->err);
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

// ------------------------------------------------------ diagnostics
//
// The reductions sph.cc's output reports, over one scene's current fluid state:
// TotalKineticEnergy (derivatives.c:2254-2263), NeighborsToPressurePoint
// (:2290-2338, pressure from the Tait equation of state) and
// DensityConsistencyFunctions (:3090-3117).  Per-block partial sums in FP64 are
// added on the host in block order, so the result does not depend on timing.
#define DIAG_Q 9
#define DIAG_BLK 256

__global__ __launch_bounds__(DIAG_BLK) void k_diag(int poff, int np, const uint32_t* __restrict__ tag,
                                                    const double2* __restrict__ pos, const double2* __restrict__ vel,
                                                    const double* __restrict__ rho, const double* __restrict__ massD,
                                                    const double* __restrict__ pres,
                                                    const DevScene* __restrict__ sc, int scene, double px, double py,
                                                    double px2, double py2, double inv_nf,
                                                    double* __restrict__ part) {
  __shared__ double red[DIAG_Q][DIAG_BLK];
  const int k = blockIdx.x * DIAG_BLK + threadIdx.x;
  double q[DIAG_Q];
#pragma unroll
  for (int j = 0; j < DIAG_Q; ++j) q[j] = 0.0;
  if (k < np) {
    const int s = poff + k;
    const uint32_t tg = tag[s];
    if (tag_type(tg) == 1) {
      const DevScene& S = sc[scene];
      const double m = massD[tag_id(tg)];
      const double2 x = pos[s], v = vel[s];
      const double rh = rho[s];
      q[0] = 0.5 * m * (v.x * v.x + v.y * v.y);
      q[1] = m * S.g * x.y;
      const double dx = x.x - px, dy = x.y - py;
      const double r = sqrt(dx * dx + dy * dy), hh = S.h;
      // DynamicPressure: the Tait pressure of the density in config mode, the
      // integrated state in ref mode
      const double P = pres ? pres[s] : tait(S, rh);
      if (r <= 3.0 * hh && P > 0.0) {
        const double coeff1 = 0.5 - 5.0 * exp(-9.0);
        const double w = (exp(-r * r / (hh * hh)) - exp(-9.0)) * m / (2.0 * M_PI * hh * hh * coeff1 * rh);
        q[2] = P * w;
        q[3] = w;
        q[4] = 1.0;
      }
      q[5] = rh * inv_nf;   // localDensity / Nf: localDensity equals density for fluid
      q[6] = rh * inv_nf;
      // NeighborsToPressurePointTwo (derivatives.c:2340-2367): the positive
      // pressures strictly within 2h of the second probe, averaged by
      // EnsembleAveragedPointPressureTwo (derivatives.c:2382-2393)
      const double ex = x.x - px2, ey = x.y - py2;
      if (sqrt(ex * ex + ey * ey) < 2.0 * hh && P > 0.0) {
        q[7] = P;
        q[8] = 1.0;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < DIAG_Q; ++j) red[j][threadIdx.x] = q[j];
  __syncthreads();
  for (int st = DIAG_BLK / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st)
#pragma unroll
      for (int j = 0; j < DIAG_Q; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < DIAG_Q) part[(size_t)blockIdx.x * DIAG_Q + threadIdx.x] = red[threadIdx.x][0];
}

// Sums of k_diag over one scene's fluid rows (t: DIAG_Q values).
static int diag_sums(uph_t* h, int scene, double px, double py, double px2, double py2, double* t) {
  if (!h) return -1;
  if (scene < 0 || scene >= h->nscenes) {
    h->err.msg = "uph_diagnostics: scene out of range";
    return -1;
  }
  const DevScene& S = h->hs[scene];
  const int nb = (S.np + DIAG_BLK - 1) / DIAG_BLK;
  if (nb > h->diag_cap) {
    if (h->diag) HIPCHK(hipFree(h->diag), h->err);
    h->diag = nullptr;
    HIPCHK(hipMalloc((void**)&h->diag, (size_t)nb * DIAG_Q * sizeof(double)), h->err);
    h->diag_cap = nb;
  }
  hipLaunchKernelGGL(k_diag, dim3(nb), dim3(DIAG_BLK), 0, h->stream, S.poff, S.np, h->b.tagA, h->b.posB,
                     h->b.velB, h->b.rhoB, h->b.massD, h->mode == UPH_MODE_REF ? h->b.presB : nullptr, h->dsc,
                     scene, px, py, px2, py2,
                     S.nf > 0 ? 1.0 / S.nf : 0.0, h->diag);
  HIPCHK(hipGetLastError(), h->err);
  std::vector<double> part((size_t)nb * DIAG_Q);
  HIPCHK(hipMemcpyAsync(part.data(), h->diag, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream),
         h->err);
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  for (int j = 0; j < DIAG_Q; ++j) t[j] = 0.0;
  for (int b = 0; b < nb; ++b)
    for (int j = 0; j < DIAG_Q; ++j) t[j] += part[(size_t)b * DIAG_Q + j];
  return 0;
}

extern "C" int uph_probe_two(uph_t* h, int scene, double x, double y, double* pressure, int* neighbors) {
  double t[DIAG_Q];
  if (diag_sums(h, scene, x, y, x, y, t)) return -1;
  const int n = (int)t[8];
  if (neighbors) *neighbors = n;
  if (pressure) *pressure = n > 0 ? t[7] / n : 0.0;
  return 0;
}

extern "C" int uph_diagnostics(uph_t* h, int scene, double probe_x, double probe_y, uph_diag* out) {
  if (!out) return -1;
  double t[DIAG_Q];
  if (diag_sums(h, scene, probe_x, probe_y, probe_x, probe_y, t)) return -1;
  memset(out, 0, sizeof *out);
  out->kinetic_energy = t[0];
  out->potential_energy = t[1];
  out->probe_neighbors = (int)t[4];
  out->probe_norm = out->probe_neighbors ? t[3] : 1.0;
  out->probe_pressure = out->probe_neighbors ? t[2] / t[3] : 0.0;
  // derivatives.c:3096-3099 accumulates outside its reduction and then overwrites
  // the result with the reduction's (empty) total: the reported value is 0
  out->consistency_one = 0.0;
  out->consistency_two = t[5];
  out->consistency_three = t[6];
  out->density_uncertainty = out->consistency_one - t[5] * t[5] / t[6];
  return 0;
}

// ApproximateVortcity (derivatives.c:2840-2885) of the last step's derivatives()
// evaluation, in gather form: for fluid p, the sum over its pair partners q of
// (m_q / rho_q) (dv_y dx_x - dv_x dx_y) W / (r^2 + 0.01 h^2) with dx = x_p - x_q,
// dv = v_p - v_q, on the state that evaluation saw (positions x_n, predicted
// velocities and densities: the pair kernel's inputs, still in the A buffers).  The
// pair term is symmetric in p and q (both differences change sign), so the gather
// equals the reference's scatter to both particles of each pair.  The pair set is
// the step's, K cap included.
template <int KT>
__global__ __launch_bounds__(BLK) void k_vort(int poff, int np, StepBufs b, const DevScene* __restrict__ sc,
                                              int scene, double* __restrict__ out) {
  const int k = blockIdx.x * BLK + threadIdx.x;
  if (k >= np) return;
  const int p = poff + k;
  const uint32_t tg = b.tagA[p];
  if (tag_type(tg) != 1) return;
  const DevScene& S = sc[scene];
  PairBase c;
  c.init_base(S, tg);
  double kpd2;
  uint32_t kpid;
  kth_of(b, p, kpd2, kpid);
  const double2 xp = b.posA[p];
  const float4 ap = b.mA[p];
  const double rho_off = S.mode == UPH_MODE_REF ? S.rho0 : 0.0;   // ref mode's mirror holds rho - rho0
  const double eta2 = 0.01 * S.h * S.h;
  double acc = 0.0;
  for_candidates(S, xp, b.cellStart, [&](int q) {
    if (q == p) return;
    const double2 xq = b.posA[q];
    const double ex = xp.x - xq.x, ey = xp.y - xq.y;
    const double d2 = __dadd_rn(__dmul_rn(ex, ex), __dmul_rn(ey, ey));
    if (!(d2 < S.cutoff)) return;
    const float4 bq = b.mB[q];
    const uint32_t tq = __float_as_uint(bq.w);
    if (!c.template admit<true>(S, tag_id(tq), d2, kpd2, kpid, b, q)) return;
    const float4 aq = b.mA[q];
    const float r2 = (float)d2;
    const double w = (double)c.template w_value<KT>(r2, sqrtf(r2));
    const double vol = (double)bq.z / ((double)bq.x + rho_off);
    const double dvx = (double)ap.x - (double)aq.x, dvy = (double)ap.y - (double)aq.y;
    acc += vol * (dvy * ex - dvx * ey) * w / (d2 + eta2);
  });
  out[tag_id(tg) - poff] = acc;
}

// approxVorticity of the last evaluation on the handle's stream, copied to host `out`.
static int vort_run(uph_t* h, int scene, double* out) {
  const DevScene& S = h->hs[scene];
  if (S.nf > h->vort_cap) {
    if (h->vort) HIPCHK(hipFree(h->vort), h->err);
    h->vort = nullptr;
    HIPCHK(hipMalloc((void**)&h->vort, (size_t)S.nf * sizeof(double)), h->err);
    h->vort_cap = S.nf;
  }
  if (h->kernel == UPH_KERNEL_POLY7)
    hipLaunchKernelGGL(k_vort<UPH_KERNEL_POLY7>, dim3(grid_of(S.np)), dim3(BLK), 0, h->stream, S.poff, S.np, h->b,
                       h->dsc, scene, h->vort);
  else
    hipLaunchKernelGGL(k_vort<UPH_KERNEL_CUBIC>, dim3(grid_of(S.np)), dim3(BLK), 0, h->stream, S.poff, S.np, h->b,
                       h->dsc, scene, h->vort);
  HIPCHK(hipGetLastError(), h->err);
  HIPCHK(hipMemcpyAsync(out, h->vort, (size_t)S.nf * sizeof(double), hipMemcpyDeviceToHost, h->stream), h->err);
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  return 0;
}

extern "C" int uph_vorticity(uph_t* h, int scene, double* out) {
  if (!h || !out) return -1;
  if (scene < 0 || scene >= h->nscenes) {
    h->err.msg = "uph_vorticity: scene out of range";
    return -1;
  }
  if (h->itime == 0) {
    h->err.msg = "uph_vorticity: no step has run";
    return -1;
  }
  return vort_run(h, scene, out);
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
  hipLaunchKernelGGL(k_scan_reduce, dim3(qnb), dim3(SCAN_BLK), 0, st, nf, c.qcount, c.qpartial, (int*)nullptr);
  hipLaunchKernelGGL(k_scan_down, dim3(qnb), dim3(SCAN_BLK), 0, st, nf, c.qcount, c.qpartial, c.qoff);
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

// ------------------------------------------------- surrogate plumbing

// Copy of a scene's current state in the reference's order (device to device).
__global__ void k_state_out(int poff, int np, const uint32_t* __restrict__ tag, const double2* __restrict__ pos,
                            const double2* __restrict__ vel, const double* __restrict__ rho, double2* __restrict__ opos,
                            double2* __restrict__ ovel, double* __restrict__ orho) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int s = poff + k;
  const int i = (int)tag_id(tag[s]) - poff;
  if (i < 0 || i >= np) return;
  if (opos) opos[i] = pos[s];
  if (ovel) ovel[i] = vel[s];
  if (orho) orho[i] = rho[s];
}

// Load fluid state from device arrays in the reference's order; on restart the
// next step takes the first-step form of the leapfrog (sph.cc:154-167) with
// localVelocity = velocity.
__global__ void k_state_in(int poff, int np, int nf, const uint32_t* __restrict__ tag, double2* __restrict__ pos,
                           double2* __restrict__ vel, double* __restrict__ rho, const double2* __restrict__ ipos,
                           const double2* __restrict__ ivel, const double* __restrict__ irho,
                           const double2* __restrict__ ilvel, double2* __restrict__ deckLvel, int restart) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int s = poff + k;
  const uint32_t g = tag_id(tag[s]);
  const int i = (int)g - poff;
  if (i < 0 || i >= nf) return;   // walls keep their own state
  if (ipos) pos[s] = ipos[i];
  if (ivel) vel[s] = ivel[i];
  if (irho) rho[s] = irho[i];
  if (restart && ilvel) deckLvel[g] = ilvel[i];
  else if (restart && ivel) deckLvel[g] = ivel[i];
}

// UPH_MODE_REF fluid rows from device arrays in the reference's order:
// DynamicPressure into the state, localDensity and localPressure into the fields
// the first-step form of the next step reads.
__global__ void k_ref_in(int poff, int np, int nf, const uint32_t* __restrict__ tag, double* __restrict__ pres,
                         double* __restrict__ deckRl, double* __restrict__ deckPl, const double* __restrict__ iP,
                         const double* __restrict__ irl, const double* __restrict__ ipl) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int s = poff + k;
  const uint32_t g = tag_id(tag[s]);
  const int i = (int)g - poff;
  if (i < 0 || i >= nf) return;
  pres[s] = iP[i];
  deckRl[g] = irl[i];
  deckPl[g] = ipl[i];
}

// Boundary slots of one scene at time t: the deck state plus the wall kinematics.
__global__ void k_walls_at(int poff, int np, int nf, const uint32_t* __restrict__ tag, double2* __restrict__ pos,
                           double2* __restrict__ vel, const double2* __restrict__ deckPos,
                           const double2* __restrict__ deckVel, const int* __restrict__ wallKind,
                           const double* __restrict__ wallRef, const DevScene* __restrict__ sc, int scene, double t) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const int s = poff + k;
  const uint32_t tg = tag[s];
  const uint32_t g = tag_id(tg);
  const int i = (int)g - poff;
  if (i < nf || i >= np) return;
  const DevScene& S = sc[scene];
  double2 x = deckPos[g], v = deckVel[g];
  if (S.motion != 0) {
    double lvx;
    (void)wall_kin(S, wallKind[g], wallRef[g], t, deckPos[g], x, v, &lvx);
  }
  pos[s] = x;
  vel[s] = v;
}

// Boundary rows of a state in reference order at time t, or at the deck (at_deck).
__global__ void k_walls_out(int poff, int np, int nf, const double2* __restrict__ deckPos,
                            const double2* __restrict__ deckVel, const int* __restrict__ wallKind,
                            const double* __restrict__ wallRef, const DevScene* __restrict__ sc, int scene, double t,
                            int at_deck, double2* __restrict__ pos, double2* __restrict__ vel) {
  const int i = nf + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= np) return;
  const int g = poff + i;
  double2 x = deckPos[g], v = deckVel[g];
  const DevScene& S = sc[scene];
  if (!at_deck) {
    double lvx;
    (void)wall_kin(S, wallKind[g], wallRef[g], t, deckPos[g], x, v, &lvx);
  }
  if (pos) pos[i] = x;
  if (vel) vel[i] = v;
}

extern "C" int uph_state_to_device(uph_t* h, int scene, double* pos, double* vel, double* rho) {
  if (!h) return -1;
  if (scene < 0 || scene >= h->nscenes) {
    h->err.msg = "uph_state_to_device: scene out of range";
    return -1;
  }
  const DevScene& S = h->hs[scene];
  // the caller's work on pos/vel/rho, on any stream, completes before they are overwritten
  HIPCHK(hipDeviceSynchronize(), h->err);
  hipLaunchKernelGGL(k_state_out, dim3(grid_of(S.np)), dim3(BLK), 0, h->stream, S.poff, S.np, h->b.tagA, h->b.posB,
                     h->b.velB, h->b.rhoB, (double2*)pos, (double2*)vel, rho);
  HIPCHK(hipGetLastError(), h->err);
  if (S.motion != 0 && S.np > S.nf) {   // boundary rows at the last step's time, as uph_get_state
    hipLaunchKernelGGL(k_walls_out, dim3(grid_of(S.np - S.nf)), dim3(BLK), 0, h->stream, S.poff, S.np, S.nf,
                       h->b.deckPos, h->b.deckVel, h->b.wallKind, h->b.wallRef, h->dsc, scene,
                       S.t0 + h->itime * S.dt, (int)(h->itime == 0 && S.t0 == 0.0), (double2*)pos, (double2*)vel);
    HIPCHK(hipGetLastError(), h->err);
  }
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  return 0;
}

extern "C" int uph_state_from_device(uph_t* h, int scene, const double* pos, const double* vel, const double* rho,
                                     int restart) {
  if (!h) return -1;
  if (scene < 0 || scene >= h->nscenes) {
    h->err.msg = "uph_state_from_device: scene out of range";
    return -1;
  }
  const DevScene& S = h->hs[scene];
  // the caller may have written pos/vel/rho on any stream
  HIPCHK(hipDeviceSynchronize(), h->err);
  hipLaunchKernelGGL(k_state_in, dim3(grid_of(S.np)), dim3(BLK), 0, h->stream, S.poff, S.np, S.nf, h->b.tagA,
                     h->b.posB, h->b.velB, h->b.rhoB, (const double2*)pos, (const double2*)vel, rho,
                     (const double2*)nullptr, h->b.deckLvel, restart);
  HIPCHK(hipGetLastError(), h->err);
  if (restart) {
    // the step counter restarts; the clock of the wall kinematics continues
    for (auto& T : h->hs) T.t0 += h->itime * T.dt;
    HIPCHK(hipMemcpyAsync(h->dsc, h->hs.data(), h->hs.size() * sizeof(DevScene), hipMemcpyHostToDevice, h->stream),
           h->err);
    const int zero = 0;
    HIPCHK(hipMemcpyAsync(h->b.ctl, &zero, sizeof(int), hipMemcpyHostToDevice, h->stream), h->err);
    h->itime = 0;
  }
  HIPCHK(hipStreamSynchronize(h->stream), h->err);
  return 0;
}

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
  hipLaunchKernelGGL(k_scan_reduce, dim3(qnb), dim3(SCAN_BLK), 0, st, n, c.cnt, c.qpartial, (int*)nullptr);
  hipLaunchKernelGGL(k_scan_down, dim3(qnb), dim3(SCAN_BLK), 0, st, n, c.cnt, c.qpartial, off);
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
    hipLaunchKernelGGL(k_scan_reduce, dim3(qnb), dim3(SCAN_BLK), 0, st, n, c.cnt, c.qpartial, (int*)nullptr);
    hipLaunchKernelGGL(k_scan_down, dim3(qnb), dim3(SCAN_BLK), 0, st, n, c.cnt, c.qpartial, soff);
    hipLaunchKernelGGL(k_graph_sender_perm, dim3(grid8_of(n)), dim3(BLK), 0, st, n, c.b, c.dsc, n_recv, off,
                       send, soff, sperm);
  }
  HIPCHK(hipGetLastError(), c.err);
  HIPCHK(hipStreamSynchronize(st), c.err);
  return 0;
}

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
  L.ctl = o; o += na_up(8 * sizeof(int));
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
  hipLaunchKernelGGL(k_scan_reduce, dim3(qnb), dim3(SCAN_BLK), 0, st, nf, qcount, qpartial, (int*)nullptr);
  hipLaunchKernelGGL(k_scan_down, dim3(qnb), dim3(SCAN_BLK), 0, st, nf, qcount, qpartial, qoff);
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

// ------------------------------------------------------ derivatives() compat
//
// The reference's rate routine (derivatives.c:3137) over its System object:
// the fluid rows go to a device context (one scene), one step of the first-step
// form runs with the caller's localVelocity, and the pair kernel's rates come
// back into the System.  The boundary rows are reset to their state at the
// first call (inputbp.c re-reads them from the decks every step) and moved by the
// kinematics of inputbp.c:192-283; the mirror rows of the System's virtualType
// (derivatives.c:3193-4192) are written after them.

#include "../../include/uph_system.h"

namespace {
struct CompatCtx {
  std::mutex mu;
  std::string err;
  uph_t* h = nullptr;
  const uph_system* sys = nullptr;
  int nf = -1, nfix = -1, nmov = -1;
  double hsml = 0.0, dt = 0.0;
  std::vector<double> key;                         // constants the context was built with
  std::vector<double2> deck_x, deck_v, deck_lv;   // boundary rows at the first call
  std::vector<double> deck_rho, deck_p;
  double2* dstage = nullptr;                       // pos, vel, lvel (nf each), then rho (nf doubles)
  long stage_n = -1;
  std::vector<double2> hpos, hvel, hlv;
  std::vector<double> hrho;
  std::vector<float> R, Rt, at;
  std::vector<double> a;
  std::vector<int> wkind;      // boundary rows' motion kind and reference (wall_class)
  std::vector<double> wref;
  // scheme of the seam: config mode (default) or the reference's run-phase scheme
  int mode = UPH_MODE_CONFIG, kernel = UPH_KERNEL_CUBIC;
  bool scheme_set = false;
  std::vector<double> hP, hrl, hpl;
  std::vector<float> RP, RPt, fr, fP, fv;
};
CompatCtx g_compat;

// The seam's scheme: set by uph_derivatives_scheme, else by the environment
// variable UPH_DERIVATIVES_SCHEME ("config", "ref" = the reference's scheme with
// its PolyOrderSeven kernel, "ref-cubic"), else config mode.
void compat_scheme_from_env() {
  CompatCtx& c = g_compat;
  if (c.scheme_set) return;
  c.scheme_set = true;
  const char* e = getenv("UPH_DERIVATIVES_SCHEME");
  if (!e) return;
  if (!strcmp(e, "ref")) {
    c.mode = UPH_MODE_REF;
    c.kernel = UPH_KERNEL_POLY7;
  } else if (!strcmp(e, "ref-cubic")) {
    c.mode = UPH_MODE_REF;
    c.kernel = UPH_KERNEL_CUBIC;
  }
}

int compat_fail(const std::string& m) {
  g_compat.err = m;
  return -1;
}

double compat_tait(const uph_system* s, double rho) {
  return s->CompressionFactor * (pow(rho / s->density0, s->AdiabaticConstant) - 1.0);
}

// The motion of the System's moving rows (inputbp.c:192-283): movingBoundary 1,
// 3, 2, else virtualType 5, else none.
int compat_motion(const uph_system* s) {
  if (s->movingBoundary == 1) return UPH_MOTION_PISTON;
  if (s->movingBoundary == 3) return UPH_MOTION_LID;
  if (s->movingBoundary == 2) return UPH_MOTION_GATE;
  if (s->virtualType == 5) return UPH_MOTION_SLOSH;
  return UPH_MOTION_NONE;
}

void compat_release() {
  CompatCtx& c = g_compat;
  if (c.h) uph_destroy(c.h);
  if (c.dstage) (void)hipFree(c.dstage);
  c.h = nullptr;
  c.dstage = nullptr;
  c.stage_n = -1;
  c.sys = nullptr;
}

// Everything the device context depends on besides the particle rows.
std::vector<double> compat_key(const uph_system* s) {
  const CompatCtx& c = g_compat;
  return {(double)s->NumberOfFluidParticles, (double)s->NumberOfFixedBoundaryParticles,
          (double)s->NumberOfMovingBoundaryParticles, s->hsml[0], s->timeStep, s->dx, s->density0,
          s->AdiabaticConstant, s->SpeedOfSound, s->CompressionFactor, s->Gravity, (double)s->gravityFlag,
          s->alpha, s->springConstant0, s->dampingConstant0, (double)s->HowManyNeighbors,
          (double)s->movingBoundary, s->stroke, s->omega, (double)c.mode, (double)c.kernel,
          s->thermalDiffusivity0, s->viscosity0, s->epsilon, (double)s->virtualType, s->lidVelocity,
          s->LocalDepth, s->FlatBottomLength,
          // the boundary positions the mirror sections read where no motion sets them
          (s->movingBoundary == 2 || s->movingBoundary == 3) ? s->LeftBoundaryPosition : 0.0,
          compat_motion(s) != UPH_MOTION_SLOSH ? s->RightBoundaryPosition : 0.0};
}

int compat_build(const uph_system* s) {
  CompatCtx& c = g_compat;
  compat_release();
  const int nf = s->NumberOfFluidParticles, nfix = s->NumberOfFixedBoundaryParticles;
  const int nmov = s->NumberOfMovingBoundaryParticles, np = nf + nfix + nmov;
  std::vector<double> pos(2 * (size_t)np), vel(2 * (size_t)np), lv(2 * (size_t)np), rho(np), mass(np);
  std::vector<int> ity(np);
  double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
  for (int i = 0; i < np; ++i) {
    for (int d = 0; d < 2; ++d) {
      pos[2 * i + d] = s->Position[i][d];
      vel[2 * i + d] = s->Velocity[i][d];
      lv[2 * i + d] = s->localVelocity[i][d];
      lo[d] = std::min(lo[d], pos[2 * i + d]);
      hi[d] = std::max(hi[d], pos[2 * i + d]);
    }
    rho[i] = s->density[i];
    mass[i] = s->mass[i];
    ity[i] = s->itype[i];
  }
  if (!std::isfinite(lo[0]) || !std::isfinite(lo[1]) || !std::isfinite(hi[0]) || !std::isfinite(hi[1]))
    return compat_fail("derivatives: non-finite positions");
  uph_scene u;
  memset(&u, 0, sizeof u);
  u.n_fluid = nf;
  u.n_fixed = nfix;
  u.n_moving = nmov;
  u.pos = pos.data();
  u.vel = vel.data();
  u.rho = rho.data();
  u.mass = mass.data();
  u.itype = ity.data();
  u.lvel = lv.data();
  u.dx = s->dx;
  u.h = s->hsml[0];
  u.rho0 = s->density0;
  u.gamma = s->AdiabaticConstant;
  u.c0 = s->SpeedOfSound;
  u.B = s->CompressionFactor;
  u.g = s->Gravity;
  u.gravity = s->gravityFlag;
  u.alpha_av = s->alpha;
  u.k_spring = s->springConstant0;
  u.c_damp = s->dampingConstant0;
  u.dt = s->timeStep;
  u.K = s->HowManyNeighbors > 0 ? s->HowManyNeighbors : 40;
  u.motion = compat_motion(s);
  u.m_amp = s->stroke;
  u.m_omega = s->omega;
  u.lid_u = s->lidVelocity;
  u.depth = s->LocalDepth;
  u.flat_bottom = s->FlatBottomLength;
  u.gate_y = s->tempPositionGate;
  u.kernel = c.kernel;
  u.mode = c.mode;
  u.kappa = s->thermalDiffusivity0;
  u.mu = s->viscosity0;
  u.eps = s->epsilon;
  u.vtype = s->virtualType;
  u.lbp = s->LeftBoundaryPosition;
  u.rbp = s->RightBoundaryPosition;
  std::vector<double> pres(np, 0.0);
  if (s->DynamicPressure)
    for (int i = 0; i < np; ++i) pres[i] = s->DynamicPressure[i];
  u.pres = pres.data();
  // cell grid: the particles' box with a 10 h margin, twice the height above
  const double mg = 10.0 * u.h;
  u.box[0] = lo[0] - mg;
  u.box[1] = lo[1] - mg;
  u.box[2] = hi[0] + mg;
  u.box[3] = hi[1] + 2.0 * (hi[1] - lo[1]) + mg;
  uph_opts o;
  memset(&o, 0, sizeof o);
  o.keep_rates = 1;
  o.pair_block = 256;
  o.cell_factor = 1.0;
  if (uph_create(&u, 1, &o, &c.h) != 0) return compat_fail(std::string("derivatives: ") + uph_last_error(nullptr));
  // staging: pos, vel, lvel (nf double2 each), then rho, P, localDensity, localPressure (nf doubles each)
  if (hipMalloc((void**)&c.dstage, ((size_t)3 * nf + 2 * (size_t)nf + 1) * sizeof(double2)) != hipSuccess) {
    compat_release();
    return compat_fail("derivatives: device staging allocation failed");
  }
  c.stage_n = nf;
  c.sys = s;
  c.nf = nf;
  c.nfix = nfix;
  c.nmov = nmov;
  c.hsml = u.h;
  c.dt = u.dt;
  c.key = compat_key(s);
  const int nb = nfix + nmov;
  c.wkind.assign(nb, 0);
  c.wref.assign(nb, 0.0);
  int halted = 0;
  for (int k = 0; k < nb; ++k)
    c.wkind[k] = wall_class(u.motion, nf + k, nf, nfix, s->itype[nf + k], s->dx,
                            s->tempPositionGate ? s->tempPositionGate[nf + k] : 0.0, &c.wref[k], &halted);
  c.deck_x.resize(nb);
  c.deck_v.resize(nb);
  c.deck_lv.resize(nb);
  c.deck_rho.resize(nb);
  c.deck_p.resize(nb);
  for (int k = 0; k < nb; ++k) {
    const int i = nf + k;
    c.deck_x[k] = make_double2(s->Position[i][0], s->Position[i][1]);
    c.deck_v[k] = make_double2(s->Velocity[i][0], s->Velocity[i][1]);
    c.deck_lv[k] = make_double2(s->localVelocity[i][0], s->localVelocity[i][1]);
    c.deck_rho[k] = s->density[i];
    c.deck_p[k] = s->DynamicPressure ? s->DynamicPressure[i] : 0.0;
  }
  c.hpos.resize(nf);
  c.hvel.resize(nf);
  c.hlv.resize(nf);
  c.hrho.resize(nf);
  c.R.resize(nf);
  c.Rt.resize(nf);
  c.a.resize(2 * (size_t)nf);
  c.at.resize(2 * (size_t)nf);
  c.hP.resize(nf);
  c.hrl.resize(nf);
  c.hpl.resize(nf);
  c.RP.resize(nf);
  c.RPt.resize(nf);
  c.fr.resize(nf);
  c.fP.resize(nf);
  c.fv.resize(2 * (size_t)nf);
  return 0;
}
}  // namespace

extern "C" int uph_derivatives(uph_system* s, double t) {
  CompatCtx& c = g_compat;
  std::lock_guard<std::mutex> lk(c.mu);
  c.err.clear();
  if (!s) return compat_fail("derivatives: null System");
  // (NumberOfVirtualParticles is an output: each call makes the mirror rows anew)
  if (s->NumberOfSolidParticles != 0 || s->NumberOfFrozenParticles != 0)
    return compat_fail("derivatives: solid and frozen particles are not supported");
  if (s->flagInitSimul != 1)   // derivatives.c:4197: 0 is the initialization phase
    return compat_fail("derivatives: flagInitSimul must be 1 (the run phase; the initialization phase is not supported)");
  if (s->dim != 2) return compat_fail("derivatives: only dim = 2 is supported");
  const int nf = s->NumberOfFluidParticles, nfix = s->NumberOfFixedBoundaryParticles;
  const int nmov = s->NumberOfMovingBoundaryParticles;
  if (nf < 0 || nfix < 0 || nmov < 0 || nf + nfix + nmov <= 0)
    return compat_fail("derivatives: invalid particle counts");
  if (!s->Position || !s->Velocity || !s->localVelocity || !s->density || !s->mass || !s->itype || !s->hsml ||
      !s->ddensitydt || !s->ddensitydt_approx || !s->dvdt || !s->dvdt_approx)
    return compat_fail("derivatives: a required System array is not allocated");
  if (!(s->timeStep > 0) || !(s->hsml[0] > 0) || !(s->density0 > 0))
    return compat_fail("derivatives: timeStep, hsml and density0 must be positive");
  compat_scheme_from_env();
  const bool ref = c.mode == UPH_MODE_REF;
  if (ref && (!s->DynamicPressure || !s->localDensity || !s->localPressure || !s->dpdt || !s->dpdt_approx))
    return compat_fail("derivatives: the reference scheme needs DynamicPressure, localDensity, localPressure, "
                       "dpdt and dpdt_approx");
  if (!c.h || c.sys != s || c.key != compat_key(s)) {
    if (compat_build(s)) return -1;
  }
  uph_handle* h = c.h;
  DevScene& S = h->hs[0];
  // fluid rows in, with the caller's localVelocity; the step counter restarts so
  // that the step takes the first-step form, at time t
  for (int i = 0; i < nf; ++i) {
    c.hpos[i] = make_double2(s->Position[i][0], s->Position[i][1]);
    c.hvel[i] = make_double2(s->Velocity[i][0], s->Velocity[i][1]);
    c.hlv[i] = make_double2(s->localVelocity[i][0], s->localVelocity[i][1]);
    c.hrho[i] = s->density[i];
  }
  double2* dpos = c.dstage;
  double2* dvel = c.dstage + nf;
  double2* dlv = c.dstage + 2 * (size_t)nf;
  double* drho = (double*)(c.dstage + 3 * (size_t)nf);
  Err& err = h->err;
  auto chk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess) {
      c.err = std::string("derivatives: ") + what + ": " + hipGetErrorString(e);
      return true;
    }
    return false;
  };
  if (nf > 0 && (chk(hipMemcpy(dpos, c.hpos.data(), nf * sizeof(double2), hipMemcpyHostToDevice), "upload") ||
                 chk(hipMemcpy(dvel, c.hvel.data(), nf * sizeof(double2), hipMemcpyHostToDevice), "upload") ||
                 chk(hipMemcpy(dlv, c.hlv.data(), nf * sizeof(double2), hipMemcpyHostToDevice), "upload") ||
                 chk(hipMemcpy(drho, c.hrho.data(), nf * sizeof(double), hipMemcpyHostToDevice), "upload")))
    return -1;
  double* dP = drho + nf;
  double* drl = dP + nf;
  double* dpl = drl + nf;
  if (ref && nf > 0) {
    // the reference's state: DynamicPressure and the local fields the host loop built
    for (int i = 0; i < nf; ++i) {
      c.hP[i] = s->DynamicPressure[i];
      c.hrl[i] = s->localDensity[i];
      c.hpl[i] = s->localPressure[i];
    }
    if (chk(hipMemcpy(dP, c.hP.data(), nf * sizeof(double), hipMemcpyHostToDevice), "upload") ||
        chk(hipMemcpy(drl, c.hrl.data(), nf * sizeof(double), hipMemcpyHostToDevice), "upload") ||
        chk(hipMemcpy(dpl, c.hpl.data(), nf * sizeof(double), hipMemcpyHostToDevice), "upload"))
      return -1;
  }
  S.t0 = t - S.dt;
  const int zero = 0;
  if (chk(hipMemcpy(h->dsc, h->hs.data(), sizeof(DevScene), hipMemcpyHostToDevice), "scene upload") ||
      chk(hipMemcpy(h->b.ctl, &zero, sizeof(int), hipMemcpyHostToDevice), "control upload"))
    return -1;
  h->itime = 0;
  hipLaunchKernelGGL(k_state_in, dim3(grid_of(S.np)), dim3(BLK), 0, h->stream, S.poff, S.np, S.nf, h->b.tagA,
                     h->b.posB, h->b.velB, h->b.rhoB, (const double2*)dpos, (const double2*)dvel,
                     (const double*)drho, (const double2*)dlv, h->b.deckLvel, 1);
  hipLaunchKernelGGL(k_walls_at, dim3(grid_of(S.np)), dim3(BLK), 0, h->stream, S.poff, S.np, S.nf, h->b.tagA,
                     h->b.posB, h->b.velB, h->b.deckPos, h->b.deckVel, h->b.wallKind, h->b.wallRef, h->dsc, 0, t);
  if (ref)
    hipLaunchKernelGGL(k_ref_in, dim3(grid_of(S.np)), dim3(BLK), 0, h->stream, S.poff, S.np, S.nf, h->b.tagA,
                       h->b.presB, h->b.deckRl, h->b.deckPl, (const double*)dP, (const double*)drl,
                       (const double*)dpl);
  if (chk(hipGetLastError(), "launch")) return -1;
  // the diagnostics derivatives() leaves in the System (derivatives.c:4317-4324), of
  // the state it was called with
  uph_diag dg;
  if (uph_diagnostics(h, 0, s->pressureProbeX, s->pressureProbeY, &dg) != 0)
    return compat_fail("derivatives: " + err.msg);
  s->totalKineticEnergy = dg.kinetic_energy;
  s->potentialEnergy = dg.potential_energy;
  s->pointNeighbors = dg.probe_neighbors;
  s->pressureNorm = dg.probe_norm;
  if (dg.probe_neighbors) s->pointPressure = dg.probe_pressure;
  s->consistencyDensityOne = dg.consistency_one;
  s->consistencyDensityTwo = dg.consistency_two;
  s->consistencyDensityThree = dg.consistency_three;
  s->densityUncertainty = dg.density_uncertainty;
  // the second probe at pressureProbeXX/YY (derivatives.c:4322-4323)
  if (uph_probe_two(h, 0, s->pressureProbeXX, s->pressureProbeYY, &s->pointPressureTwo,
                    &s->pointNeighborsTwo) != 0)
    return compat_fail("derivatives: " + err.msg);
  if (launch_step(h, nullptr, nullptr)) {
    if (c.err.empty()) c.err = "derivatives: " + err.msg;
    return -1;
  }
  // approxVorticity of this evaluation (derivatives.c:4251), where the System has it
  if (s->approxVorticity && vort_run(h, 0, s->approxVorticity) != 0)
    return compat_fail("derivatives: " + err.msg);
  if (get_rates_f64(h, 0, c.R.data(), c.Rt.data(), c.a.data(), c.at.data()) != 0)
    return compat_fail("derivatives: " + err.msg);
  if (ref && uph_get_ref_rates(h, 0, c.RP.data(), c.RPt.data(), c.fr.data(), c.fP.data(), c.fv.data()) != 0)
    return compat_fail("derivatives: " + err.msg);
  if (uph_sync(h) < 0) return compat_fail("derivatives: " + err.msg);
  // rates and the equation of state for the fluid rows
  const double B = s->CompressionFactor, gm = s->AdiabaticConstant, r0 = s->density0;
  for (int i = 0; i < nf; ++i) {
    s->ddensitydt[i] = c.R[i];
    s->ddensitydt_approx[i] = c.Rt[i];
    s->dvdt[i][0] = c.a[2 * i];
    s->dvdt[i][1] = c.a[2 * i + 1];
    s->dvdt_approx[i][0] = c.at[2 * i];
    s->dvdt_approx[i][1] = c.at[2 * i + 1];
    const double rho = s->density[i];
    const double P = compat_tait(s, rho);
    if (s->Pressure) s->Pressure[i] = P;   // EquationOfState (derivatives.c:1674-1682)
    if (ref) {
      // DynamicPressure is the caller's state; the pressure rates and the
      // fluctuation fields come from the device
      s->dpdt[i] = c.RP[i];
      s->dpdt_approx[i] = c.RPt[i];
      if (s->densityFluctuation) s->densityFluctuation[i] = c.fr[i];
      if (s->pressureFluctuation) s->pressureFluctuation[i] = c.fP[i];
      if (s->velocityFluctuation) {
        s->velocityFluctuation[i][0] = c.fv[2 * i];
        s->velocityFluctuation[i][1] = c.fv[2 * i + 1];
      }
      continue;
    }
    if (s->DynamicPressure) s->DynamicPressure[i] = P;
    const double dPdrho = B * gm / r0 * pow(rho / r0, gm - 1.0);
    if (s->dpdt) s->dpdt[i] = dPdrho * c.R[i];
    if (s->dpdt_approx) s->dpdt_approx[i] = dPdrho * c.Rt[i];
    if (s->densityFluctuation) s->densityFluctuation[i] = 0.0;
    if (s->pressureFluctuation) s->pressureFluctuation[i] = 0.0;
    if (s->velocityFluctuation) s->velocityFluctuation[i][0] = s->velocityFluctuation[i][1] = 0.0;
  }
  // boundary rows: the deck at time t (inputbp.c:117-150, 194-207, 288)
  for (int k = 0; k < nfix + nmov; ++k) {
    const int i = nf + k;
    double2 x = c.deck_x[k], v = c.deck_v[k], lv = c.deck_lv[k];
    double lvx;
    if (wall_kin(h->hs[0], c.wkind[k], c.wref[k], t, c.deck_x[k], x, v, &lvx)) lv.x = lvx;
    s->Position[i][0] = x.x;
    s->Position[i][1] = x.y;
    s->Velocity[i][0] = v.x;
    s->Velocity[i][1] = v.y;
    s->localVelocity[i][0] = lv.x;
    s->localVelocity[i][1] = lv.y;
    s->density[i] = c.deck_rho[k];
    if (s->DynamicPressure) s->DynamicPressure[i] = c.deck_p[k];
    if (s->Pressure) s->Pressure[i] = compat_tait(s, c.deck_rho[k]);
    if (s->localDensity) s->localDensity[i] = r0;
    if (s->localPressure) s->localPressure[i] = 0.0;   // inputbp.c:290
  }
  // mirror rows after the particle rows, with the fields derivatives.c:3193-4192
  // copies from their sources, and their number (derivatives.c:4192)
  s->NumberOfVirtualParticles = 0;
  if (S.nsec > 0) {
    const int cap = S.nsec * nf;
    std::vector<double> gx(2 * (size_t)cap), gv(2 * (size_t)cap), gl(2 * (size_t)cap), gr(cap), grl(cap), gp(cap),
        gpl(cap), gm(cap);
    const int ng = uph_get_ghosts(h, 0, cap, gx.data(), gv.data(), gl.data(), gr.data(), grl.data(), gp.data(),
                                  gpl.data(), gm.data());
    if (ng < 0) return compat_fail("derivatives: " + err.msg);
    const int first = nf + nfix + nmov;
    if ((long)first + ng > (long)s->MaxNumberOfParticles)
      return compat_fail("derivatives: the mirror rows need " + std::to_string(first + ng) +
                         " rows, the System's MaxNumberOfParticles is " + std::to_string(s->MaxNumberOfParticles));
    for (int k = 0; k < ng; ++k) {
      const int i = first + k;
      s->Position[i][0] = gx[2 * k];
      s->Position[i][1] = gx[2 * k + 1];
      s->Velocity[i][0] = gv[2 * k];
      s->Velocity[i][1] = gv[2 * k + 1];
      s->localVelocity[i][0] = gl[2 * k];
      s->localVelocity[i][1] = gl[2 * k + 1];
      s->mass[i] = gm[k];
      s->density[i] = gr[k];
      s->hsml[i] = s->hsml[0];
      s->itype[i] = 4;
      if (s->DynamicPressure) s->DynamicPressure[i] = gp[k];
      if (s->localDensity) s->localDensity[i] = grl[k];
      if (s->localPressure) s->localPressure[i] = gpl[k];
    }
    s->NumberOfVirtualParticles = ng;
  }
  return 0;
}

extern "C" int uph_derivatives_scheme(int mode, int kernel) {
  CompatCtx& c = g_compat;
  std::lock_guard<std::mutex> lk(c.mu);
  if ((mode != UPH_MODE_CONFIG && mode != UPH_MODE_REF) || (kernel != UPH_KERNEL_CUBIC && kernel != UPH_KERNEL_POLY7)) {
    c.err = "uph_derivatives_scheme: unknown mode or kernel";
    return -1;
  }
  c.mode = mode;
  c.kernel = kernel;
  c.scheme_set = true;
  return 0;
}

extern "C" void derivatives(uph_system* s, double t) {
  if (uph_derivatives(s, t) == 0) return;
  fprintf(stderr, "%s\n", uph_derivatives_last_error());
  if (!s || !s->ddensitydt || !s->dvdt) return;
  for (int i = 0; i < s->NumberOfFluidParticles; ++i) {
    s->ddensitydt[i] = NAN;
    s->dvdt[i][0] = s->dvdt[i][1] = NAN;
    if (s->ddensitydt_approx) s->ddensitydt_approx[i] = NAN;
    if (s->dvdt_approx) s->dvdt_approx[i][0] = s->dvdt_approx[i][1] = NAN;
  }
}

extern "C" const char* uph_derivatives_last_error(void) { return g_compat.err.c_str(); }
