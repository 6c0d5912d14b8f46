// uph_internal.h -- device code and host internals shared by the translation units
// of libuph_gpu.so.
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
#pragma once
#include "uph_kernels.h"
#include "../../include/uph.h"

#include <limits.h>
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
// the most scenes a handle re-sorts lazily by default (uph_opts.sort_every_step = 0)
#define LAZY_MAX_SCENES 16
// Candidates (slots of a query's 3x3 cells, or of one cell) beyond which a step
// gives up on the query and flags the state as broken: no physical particle distribution comes
// near it (a smoothing length of 6 lattice spacings with 4-fold coarsened cells
// gives about 21,000), while a run that has blown up piles its escaped particles
// into the clamped edge cells, where an exact search would cost O(N) per query.
#define CAND_MAX 65536

// Control words (StepBufs::ctl).  A word that a kernel reads is never written by
// another block of the same kernel: the lazy step's pair kernel copies the step
// counter, the stamp and its view for k_post, which writes the next step's values,
// and its flags are kept per itime parity (a step sets its own, k_post clears the
// next step's).
#define C_IT 0       // itime of the step being computed
#define C_STAMP 1    // stamp of that step's K-th keys (never reset, so a stamp never repeats)
#define C_NOVF 2     // steps whose capped pass ran (a query had more than K in range)
#define C_OVF 3      // overflow flag of the step (sort-every-step handles)
#define C_BROKEN 4   // broken state (a query with more than CAND_MAX candidates; sticky)
#define C_ITC 5      // lazy: itime and stamp of the step, copied by the pair kernel for k_post
#define C_STC 6
#define C_V 7        // lazy: buffer view of the next pair kernel (1: the A and B sets exchanged)
#define C_PEND 8     // lazy: a re-sort that the last step of a call held over to the next call
#define C_VC 9       // lazy: the view the pair kernel ran with, for k_post
#define C_LOVF 10    // lazy: overflow flag by itime parity (2 words)
#define C_LSTALE 12  // lazy: some scene's cell list went stale, by itime parity (2 words)
#define C_NRES 14    // lazy: re-sorts done (a count for reports)
#define C_SCAP 15    // stressType 2: a query had more than K neighbours in range (sticky)
#define CTL_WORDS 16

// ------------------------------------------------------------------ sort

// One block of the binning: slots blk * BLK + [0, BLK).  With `keep` (a lazy
// re-sort), a slot of a scene that is not stale (keep[scene] == 0) keeps the cell it
// was binned into (cellOf), so that scene's slot order does not change.  tagB (when
// given) receives a copy of tagA for the reorder.
__device__ __forceinline__ void bin_block(int blk, int n, const double2* __restrict__ pos,
                                          const int* __restrict__ sceneOf, const DevScene* __restrict__ sc,
                                          const uint32_t* __restrict__ tagA, uint32_t* __restrict__ tagB,
                                          int* __restrict__ keys, int* __restrict__ rank, int* __restrict__ count,
                                          const int* __restrict__ keep = nullptr,
                                          const int* __restrict__ cellOf = nullptr) {
  const int i = blk * BLK + threadIdx.x;
  const bool valid = i < n;
  int key = -1;
  if (valid) {
    const int s = sceneOf[i];
    const DevScene& S = sc[s];
    if (keep && keep[s] == 0) {
      const int pk = cellOf[i];
      key = S.coff + (pk >> 16) * S.ncx + (pk & 0xffff);
    } else {
      const double2 p = pos[i];
      int cx, cy;
      cell_of(S, p.x, p.y, cx, cy);
      key = S.coff + cy * S.ncx + cx;
    }
    keys[i] = key;
    if (tagB) tagB[i] = tagA[i];
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

// Opens a step: itime += 1, the overflow flag cleared, the K-th key stamp advanced.
// A lazy handle's first step after its state was set (sstale given) also clears that
// step's parity flags and takes the buffer view its sort gathers into.
__device__ __forceinline__ void open_step(int* ctl, int* sstale = nullptr, int nscenes = 0, int view = 0) {
  const int it = ctl[C_IT] + 1;
  ctl[C_IT] = it;            // itime of the step being computed
  ctl[C_OVF] = 0;            // overflow flag of this step
  ctl[C_STAMP] += 1;         // stamp of this step's K-th keys
  if (sstale) {
    ctl[C_NRES] += 1;
    ctl[C_LOVF + (it & 1)] = 0;
    ctl[C_LSTALE + (it & 1)] = 0;
    ctl[C_V] = view;
    ctl[C_PEND] = 0;
    for (int s = 0; s < nscenes; ++s) sstale[(it & 1) * nscenes + s] = 0;
  }
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
  // Lazy re-sort handles (lazy = 1; 2 in the reference scheme, whose re-sort also
  // gathers DynamicPressure and mC) keep two full state sets and exchange their roles
  // between steps: the pair kernel's epilogue also writes the next step's predicted
  // neighbour mirrors (mA2, mB2, mC2 on the B side), the tag and the rates, so that a
  // step without a re-sort reads the B set as its A set.  rtA, rt2A, dvA are the A
  // side's rates (written when the roles are exchanged).  cellOf: the cell (global
  // key) each slot was binned into at the last sort.  sstale: per scene and itime
  // parity, whether one of its particles left its binned cell by more than the margin.
  // `it` points at {itime, stamp} of the step (the pair of words C_IT, C_STAMP, or their
  // copies C_ITC, C_STC inside k_post).  cellOf holds the binned cell's coordinates
  // packed as cx | cy << 16 (cell_pack).
  float4 *mA2, *mB2, *rtA, *rt2A;
  float2 *mC2, *dvA;
  // stressType 2 handles: each slot's viscous stress {xx, xy (= yx), yy, 0} of the
  // step (k_stress), read by the pair kernel for the slot and its neighbours
  float4* sst;
  int* cellOf;
  int* sstale;
  int* it;
  int lazy, nscenes;
  int view;   // 1 in the view with the A and B sets exchanged (swapped)
  // sort scratch
  int *keys, *rank, *tmpIdx, *cellStart, *cellCount, *partial;
  int* sceneOf;
  // capped search: K-th key of overflowing queries, valid when its step equals the
  // step stamp ctl[3] (a counter that no restart resets, so a stamp never repeats),
  // or its negation when the per-lane search has also left the query's pair
  // partners in kmask (twelve 64-bit words: row r's slots rb[r] + [0, 64) in word 2r,
  // + [64, 128) in word 2r + 1, for the admitted candidates above the query; words
  // 6-11 the same for its candidates below, pending their own K-th keys; word u of
  // slot q at kmask[u * kmask_n + q], so that a wavefront's stores and loads of one
  // word are contiguous)
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
  // control words, C_* (CTL_WORDS)
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

// A cell's coordinates in one int (cellOf), and from its global key in scene S.
__host__ __device__ __forceinline__ int cell_pack(int cx, int cy) { return cx | (cy << 16); }
__device__ __forceinline__ int cell_pack_key(const DevScene& S, int key) {
  const int r = key - S.coff;
  return cell_pack(r % S.ncx, r / S.ncx);
}

// The buffers with the roles of the A and B state sets exchanged (a lazy handle's
// view 1).
__host__ __device__ __forceinline__ StepBufs swapped(const StepBufs& b) {
  StepBufs s = b;
  s.posA = b.posB;
  s.posB = b.posA;
  s.velA = b.velB;
  s.velB = b.velA;
  s.rhoA = b.rhoB;
  s.rhoB = b.rhoA;
  s.tagA = b.tagB;
  s.tagB = b.tagA;
  s.mA = b.mA2;
  s.mA2 = b.mA;
  s.mB = b.mB2;
  s.mB2 = b.mB;
  s.mC = b.mC2;
  s.mC2 = b.mC;
  s.presA = b.presB;
  s.presB = b.presA;
  s.rtB = b.rtA;
  s.rtA = b.rtB;
  s.rt2B = b.rt2A;
  s.rt2A = b.rt2B;
  s.dvB = b.dvA;
  s.dvA = b.dvB;
  s.view = b.view ^ 1;
  return s;
}

// The buffers of a lazy handle's step as the pair kernel and the sort see them: the
// view k_post chose (C_V), exchanged once more when the last call held a re-sort over
// (C_PEND: the gated sort of this call's first step then gathers the state back).
__device__ __forceinline__ StepBufs lazy_view(const StepBufs& b) {
  if (!b.lazy) return b;
  const int v = (b.ctl[C_V] ^ b.ctl[C_PEND]) & 1;
  return v ? swapped(b) : b;
}

// A lazy step's pair kernel: the buffers of the step's view (the A and B sets exchange
// roles from step to step unless a re-sort gathered the state back), and block 0
// leaves the step's itime, stamp and view for the kernels after it (which read only
// these copies, so the last of them may write the next step's values while its other
// blocks still run).
__device__ __forceinline__ StepBufs pair_view(const StepBufs& b) {
  int* ctl = b.ctl;
  const int v = (ctl[C_V] ^ ctl[C_PEND]) & 1;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl[C_ITC] = ctl[C_IT];
    ctl[C_STC] = ctl[C_STAMP];
    ctl[C_VC] = v;
  }
  return v ? swapped(b) : b;
}

// The overflow flag of the step that b.it describes.
__device__ __forceinline__ int* ovf_word(const StepBufs& b) {
  return b.lazy ? &b.ctl[C_LOVF + (b.it[0] & 1)] : &b.ctl[C_OVF];
}

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

// The predictor of sph.cc:93-126 for one slot at the start of step itime, from
// its state after the last step (position p, velocity v, density rho; UPH_MODE_REF:
// DynamicPressure Pst) and the last evaluation's predicting rates rt = {R~, a~, mass}
// (UPH_MODE_REF also r2 = {dpdt_approx, drho, dP} and the velocity fluctuation dv):
// rho += dt/2 R~, v += dt/2 a~, vloc = v + dt/2 a~; Tait pressure of the predicted
// density.  UPH_MODE_REF also predicts DynamicPressure (P += dt/2 dpdt_approx) and
// adds the last step's fluctuations to the local fields: localDensity = rho + drho,
// localPressure = P + dP, localVelocity = v + dt/2 a~ + eps dv; at the first step they
// are deckRl, deckPl (rho0 and 0 from uph_create, input.c:109-110) and the deck's
// localVelocity.  Boundary slots keep the deck's localVelocity (the kinematics' one
// for moving walls), DynamicPressure from the deck, localDensity rho0 and
// localPressure 0 (inputbp.c:117-150, 289-290).  The sort and a lazy handle's pair
// epilogue both call this, so their mirrors agree bit for bit.
struct Pred {
  double vx, vy, lx, ly;   // predicted velocity, localVelocity
  double rp;               // predicted density
  double P, Pp, rl, pl;    // P: Tait pressure of rp (config) or DynamicPressure at the step's start (ref);
                           // Pp: the predicted pressure; rl, pl: localDensity, localPressure
};

__device__ __forceinline__ Pred predict(const DevScene& S, const StepBufs& b, uint32_t tg, int itime, double2 p,
                                        double2 v, double rho, float4 rt, double Pst, float4 r2, float2 dv) {
  const uint32_t my = tag_id(tg);
  const double dt = S.dt;
  Pred o;
  o.vx = v.x;
  o.vy = v.y;
  o.rp = rho;
  if (tag_type(tg) == 1) {
    if (itime != 1) {
      o.rp = rho + 0.5 * dt * (double)rt.x;
      o.vx = v.x + 0.5 * dt * (double)rt.y;
      o.vy = v.y + 0.5 * dt * (double)rt.z;
      o.lx = o.vx + 0.5 * dt * (double)rt.y;
      o.ly = o.vy + 0.5 * dt * (double)rt.z;
    } else {
      const double2 l0 = b.deckLvel[my];
      o.lx = l0.x;
      o.ly = l0.y;
    }
  } else {
    const double2 l0 = b.deckLvel[my];
    o.lx = l0.x;
    o.ly = l0.y;
    if (S.motion != 0) {
      double2 xx = p, vv = v;
      double lvx;
      if (wall_kin(S, b.wallKind[my], b.wallRef[my], S.t0 + itime * dt, b.deckPos[my], xx, vv, &lvx)) o.lx = lvx;
    }
  }
  if (S.mode == 1) {
    o.P = tag_type(tg) == 1 ? Pst : b.deckP[my];
    o.Pp = o.P;
    o.rl = S.rho0;
    o.pl = 0.0;
    if (tag_type(tg) == 1 && itime == 1) {
      o.rl = b.deckRl[my];
      o.pl = b.deckPl[my];
    } else if (tag_type(tg) == 1) {
      o.Pp = o.P + 0.5 * dt * (double)r2.x;
      o.rl = o.rp + (double)r2.y;
      o.pl = o.Pp + (double)r2.z;
      o.lx += S.eps * (double)dv.x;
      o.ly += S.eps * (double)dv.y;
    }
  } else {
    o.P = tait(S, o.rp);
    o.Pp = o.P;
    o.rl = o.rp;   // config mode keeps localDensity equal to density
    o.pl = 0.0;
  }
  return o;
}

// predict() of a fluid slot at a step after the first (the lazy pair epilogue's next
// step): the same expressions, without the first step's deck fields.
__device__ __forceinline__ Pred predict_fluid_next(const DevScene& S, double2 v, double rho, float4 rt, double Pst,
                                                   float4 r2, float2 dv) {
  const double dt = S.dt;
  Pred o;
  o.rp = rho + 0.5 * dt * (double)rt.x;
  o.vx = v.x + 0.5 * dt * (double)rt.y;
  o.vy = v.y + 0.5 * dt * (double)rt.z;
  o.lx = o.vx + 0.5 * dt * (double)rt.y;
  o.ly = o.vy + 0.5 * dt * (double)rt.z;
  if (S.mode == 1) {
    o.P = Pst;
    o.Pp = o.P + 0.5 * dt * (double)r2.x;
    o.rl = o.rp + (double)r2.y;
    o.pl = o.Pp + (double)r2.z;
    o.lx += S.eps * (double)dv.x;
    o.ly += S.eps * (double)dv.y;
  } else {
    o.P = tait(S, o.rp);
    o.Pp = o.P;
    o.rl = o.rp;
    o.pl = 0.0;
  }
  return o;
}

// The FP32 neighbour mirrors of a predicted slot: {v, vloc} and {rho, P/rho^2, m, tag}
// (config) or {rho - rho0, P, m, tag} and {localDensity - rho0, localPressure} (ref).
// In ref mode densities enter the mirror as deviations from rho0 (formed in FP64), so
// the differences the fluctuations and diffusion terms take keep FP32's relative
// precision instead of losing the leading digits of rho0.
__device__ __forceinline__ void mirror_of(const DevScene& S, const Pred& o, float mass, uint32_t tg, float4& ma,
                                          float4& mb, float2& mc) {
  ma = make_float4((float)o.vx, (float)o.vy, (float)o.lx, (float)o.ly);
  if (S.mode == 1) {
    mb = make_float4((float)(o.rp - S.rho0), (float)o.Pp, mass, __uint_as_float(tg));
    mc = make_float2((float)(o.rl - S.rho0), (float)o.pl);
  } else {
    mb = make_float4((float)o.rp, (float)(o.P / (o.rp * o.rp)), mass, __uint_as_float(tg));
    mc = make_float2(0.f, 0.f);
  }
}

// Final slot = cell start + number of same-cell particles with a smaller
// original index (deterministic order); gather the state and build the
// predictor mirror of the step (predict, mirror_of); the slot's cell goes to cellOf.
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
    atomicOr(&b.ctl[C_BROKEN], 1);
  }
  const double2 p = b.posB[i];
  b.posA[dst] = p;
  b.tagA[dst] = tg;
  if (search_only) return;
  const DevScene& S = sc[b.sceneOf[i]];
  if (b.cellOf) b.cellOf[dst] = cell_pack_key(S, c);
  const double2 v = b.velB[i];
  const double rho = b.rhoB[i];
  const float4 rt = b.rtB[i];
  b.velA[dst] = v;
  b.rhoA[dst] = rho;
  const int itime = b.it[0];
  const bool ref = S.mode == 1;
  const Pred o = predict(S, b, tg, itime, p, v, rho, rt, ref ? b.presB[i] : 0.0,
                         ref ? b.rt2B[i] : make_float4(0.f, 0.f, 0.f, 0.f),
                         ref ? b.dvB[i] : make_float2(0.f, 0.f));
  float4 ma, mb;
  float2 mc;
  mirror_of(S, o, rt.w, tg, ma, mb, mc);
  b.mA[dst] = ma;
  b.mB[dst] = mb;
  if (ref) {
    b.presA[dst] = o.P;
    b.mC[dst] = mc;
  }
  if (MIRRORS && S.nsec > 0 && tag_type(tg) == 1 && mark_mirrors(S, b, itime, my, p, dst)) {
    b.predV[dst] = make_double2(o.vx, o.vy);   // the fields this row's mirror rows copy
    b.predVl[dst] = make_double2(o.lx, o.ly);
    b.predRho[dst] = o.rp;
    b.predRl[dst] = o.rl;
    b.predP[dst] = o.Pp;
    b.predPl[dst] = o.pl;
  }
}

// A lazy re-sort's reorder: the slot's final position as reorder_slot finds it, and
// a plain gather of the B set into the A set (the pair epilogue has already predicted
// the mirrors); the slot's cell goes to cellOf.
__device__ __forceinline__ void permute_slot(int d, int n, const StepBufs& b, const DevScene* __restrict__ sc,
                                             bool ref) {
  if (d >= n) return;
  const int i = b.tmpIdx[d];
  const int c = b.keys[i];
  const int c0 = b.cellStart[c], c1 = b.cellStart[c + 1];
  const uint32_t tg = b.tagB[i];
  const uint32_t my = tag_id(tg);
  int dst = d;
  if (c1 - c0 <= CAND_MAX) {
    int r = 0;
    for (int t = c0; t < c1; ++t) r += (tag_id(b.tagB[b.tmpIdx[t]]) < my) ? 1 : 0;
    dst = c0 + r;
  } else {
    atomicOr(&b.ctl[C_BROKEN], 1);
  }
  b.posA[dst] = b.posB[i];
  b.tagA[dst] = tg;
  b.velA[dst] = b.velB[i];
  b.rhoA[dst] = b.rhoB[i];
  b.mA[dst] = b.mA2[i];
  b.mB[dst] = b.mB2[i];
  if (ref) {
    b.presA[dst] = b.presB[i];
    b.mC[dst] = b.mC2[i];
  }
  b.cellOf[dst] = cell_pack_key(sc[b.sceneOf[i]], c);
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

// The mirror rows of the listed fluid rows, at their indices in the reference's
// order: section k's rows come after those of sections < k, and within a section
// in fluid index order (the set bits below the row's).
__device__ void emit_mirrors(const StepBufs& b, const DevScene* __restrict__ sc, int d) {
  const uint32_t tg = b.tagA[d];
  const int s = b.sceneOf[d];
  const DevScene& S = sc[s];
  const int i = (int)tag_id(tg) - S.poff;
  double lbp, rbp;
  boundary_lines(S, S.t0 + b.it[0] * S.dt, lbp, rbp);
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
// Where the per-lane search reads a candidate's position, given its row r and slot s:
// posA itself, or a copy of the block's candidate rows in LDS (k_kth), row r's slots
// [lo_r, hi_r) at so[r] + s.
// slot(s): the same for a slot of any of the three rows (rows lie in slot order; hi0 and
// hi1 end rows 0 and 1, an absent row being empty).
struct GlobalPos {
  const double2* __restrict__ p;
  __device__ __forceinline__ double2 operator()(int, int s) const { return p[s]; }
  __device__ __forceinline__ double2 slot(int s) const { return p[s]; }
};
struct StagedPos {
  const double2* sp;
  int so[3];
  int hi0, hi1;
  __device__ __forceinline__ double2 operator()(int r, int s) const { return sp[so[r] + s]; }
  __device__ __forceinline__ double2 slot(int s) const {
    return sp[s + (s < hi0 ? so[0] : s < hi1 ? so[1] : so[2])];
  }
};

// a histogram of the in-range d2 over LANE_BUCKETS equal buckets of [0, cutoff)
// finds the bucket that holds rank K - 1 (bucketing is monotone in d2, so every key
// of a lower bucket is smaller); its slots are collected and the rank taken among
// them by full (d2, id) keys.  Returns false, leaving the query to the wavefront
// search, when that bucket holds more than LANE_CAP or a row is longer than 128
// slots.  +inf when at most K are in range.
// With the K-th key known, the query's pair partners by the rules of
// Search.cc:27-40 go to mask_out in two sets of six words (row r's slots
// rb[r] + [0, 64) in word 2r, + [64, 128) in word 2r + 1): words 0-5 the candidates
// above q (by original index, below ParticleBoundary) that rank within q's K
// nearest, admitted whatever their own lists hold; words 6-11 the candidates below
// q, whose admission needs their own K-th key (the capped pass reads it).
// Candidate positions are read through `at`.
template <class PosAt>
__device__ bool lane_kth(const DevScene& S, int q, double2 xq, const int* rb, const int* rl,
                         const double2* __restrict__ pos, const uint32_t* __restrict__ tag, LaneSel& L,
                         double& kd2, uint32_t& kid, unsigned long long* __restrict__ mask_out,
                         int mask_stride, const PosAt& at) {
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
        for (int u = 0; u < 4; ++u) x[u] = at(r, rb[r] + (k + u < k1 ? k + u : k1 - 1));
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
  // the bucket holding rank K - 1; none (LANE_BUCKETS) when every key in range ranks
  // within K
  int below = 0, kb = LANE_BUCKETS, m = 0;
  if (n > S.K) {
#pragma unroll 1
    for (kb = 0; kb < LANE_BUCKETS; ++kb) {
      const uint32_t w = L.cnt[kb >> 1][lane];
      m = (int)((kb & 1) ? (w >> 16) : (w & 0xffffu));
      if (below + m >= S.K) break;
      below += m;
    }
    if (m > LANE_CAP) return false;
  }
  const uint32_t idq = tag_id(tag[q]);
  const bool pb_ok = (int)idq < S.pb_gid;
  // one pass over the in-range candidates, four per trip with their loads issued
  // together: a lower bucket ranks within K, a higher one does not; the selected
  // bucket's slots are collected for the rank below
  unsigned long long am[6];
  int j = 0;
#pragma unroll
  for (int u = 0; u < 6; ++u) {
    const int base = rb[u >> 1] + 64 * (u & 1);
    unsigned long long mm = mk[u], aw = 0ull, pw = 0ull;
#pragma unroll 1
    while (mm) {
      int kk[4];
      double2 x[4];
      uint32_t g[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        kk[v] = mm ? __ffsll((long long)mm) - 1 : -1;
        mm &= mm - 1ull;
        // an exhausted word reads its row's first slot (the query's own, for tag)
        x[v] = at(u >> 1, kk[v] >= 0 ? base + kk[v] : rb[u >> 1]);
        g[v] = tag[kk[v] >= 0 ? base + kk[v] : q];
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const bool in = kk[v] >= 0;
        const double d2 = dist2(xq, x[v]);
        const int bk = min((int)(d2 * scale), LANE_BUCKETS - 1);
        const uint32_t id = tag_id(g[v]);
        if (in && bk == kb) L.slot[j++][lane] = (uint32_t)(base + kk[v]);
        const int sh = kk[v] & 63;
        aw |= (unsigned long long)(in && bk < kb && id > idq && (int)id < S.pb_gid) << sh;
        pw |= (unsigned long long)(in && id < idq && pb_ok) << sh;
      }
    }
    am[u] = aw;
    mask_out[(size_t)(6 + u) * mask_stride] = pw;
  }
  if (n > S.K) {
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
          x[u] = at.slot((int)t);
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
    // the selected bucket's members within the K nearest, above q and below
    // ParticleBoundary, join the admitted words (each slot lies in one word of its row)
#pragma unroll 1
    for (int i = 0; i < m; ++i) {
      const int t = (int)L.slot[i][lane];
      const double d2 = dist2(xq, at.slot(t));
      const uint32_t id = tag_id(tag[t]);
      const bool ok = id > idq && (int)id < S.pb_gid && key_le(d2, id, td, ti);
#pragma unroll
      for (int u = 0; u < 6; ++u) {
        const int s = t - (rb[u >> 1] + 64 * (u & 1));
        const int len = rl[u >> 1] - 64 * (u & 1);
        if (ok && s >= 0 && s < 64 && s < len) am[u] |= 1ull << s;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < 6; ++u) mask_out[(size_t)u * mask_stride] = am[u];
  return true;
}

// The overflowing queries of this wavefront (more than K candidates in range,
// or a row range the compacted kernel cannot hold; ovf set on their lanes): each
// gets its K-th key, stamped with the step, and the step's flag makes the exact
// kernel run.  Every lane of the wavefront calls this.  With ROWS, each lane
// passes its own scene index, position and candidate rows, and the wavefront takes
// them from the query's lane instead of reading them again.
template <bool ROWS, class PosAt = GlobalPos>
__device__ void wave_overflow(const StepBufs& b, const DevScene* __restrict__ sc, int p, bool ovf, KthLds& L,
                              int sidx = 0, double2 xp = double2{}, const int* rbp = nullptr,
                              const int* rlp = nullptr, LaneSel* ls = nullptr,
                              const PosAt& at = GlobalPos{nullptr}) {
  unsigned long long m = __ballot(ovf);
  if (m == 0ull) return;
  const bool lead = (threadIdx.x & 63) == 0;
  if (lead) atomicOr(ovf_word(b), 1);
  // many overflowing lanes (the reference's smoothing length caps every query):
  // each finds its own key at once; what it leaves goes to the wavefront search
  if (ROWS && ls != nullptr && b.kmask != nullptr && __popcll(m) >= 8) {
    bool done = false;
    if (ovf) {
      const DevScene& S = sc[sidx];
      double kd2;
      uint32_t kid;
      if (S.nsec == 0 && lane_kth(S, p, xp, rbp, rlp, b.posA, b.tagA, *ls, kd2, kid, b.kmask + p, b.kmask_n,
                                        at)) {
        b.kth[p] = KthRec{kd2, kid, -b.it[1]};   // negated: kmask holds its pair partners
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
      b.kth[q] = KthRec{kd2, kid, b.it[1]};
      if (n < 0) atomicOr(&b.ctl[C_BROKEN], 1);
    }
  }
}

// K-th key of slot q in this step, or +inf when q did not overflow.
__device__ __forceinline__ void kth_of(const StepBufs& b, int q, double& kd2, uint32_t& kid) {
  const KthRec k = b.kth[q];
  if (k.step == b.it[1] || k.step == -b.it[1]) {
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

  __device__ __forceinline__ void init(const DevScene& S, uint32_t tg, float4 ap, float4 bp, float2, float4) {
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
                                      float2, float4) {
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
// is the pair's first index branch on the original ids, as a5 does).  TWO: stressType
// 2, dvdt's viscous term from the stresses of p and q (k_stress) instead of the
// single-step form.
template <bool TWO>
struct RefCtxT : PairBase {
  float vpx, vpy, lpx, lpy;   // predicted velocity, localVelocity
  float rhop, Pp, mp, rlp;    // predicted density, DynamicPressure, mass, localDensity
  float drp, drlp, r0;        // density and localDensity minus rho0; rho0
  float inv_mp, irp, irp2, Kp, gk, kap, mu;
  CSum sR, sRt;               // ddensitydt, ddensitydt_approx
  CSum sL, sC2, sLt;          // dpdt diffusion part; sum m_q (v_p - v_q).grad W; dpdt_approx diffusion part
  CSum sAx, sAy, sTx, sTy;    // dvdt, dvdt_approx
  float dr = 0.f, dP = 0.f, dvx = 0.f, dvy = 0.f;   // fluctuations
  float4 sp;                  // TWO: p's viscous stress {xx, xy, yy, 0}

  __device__ __forceinline__ void init(const DevScene& S, uint32_t tg, float4 ap, float4 bp, float2 cp,
                                       float4 spv) {
    init_base(S, tg);
    if constexpr (TWO) sp = spv;
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
                                      float2 cq, float4 sq) {
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
      if constexpr (TWO) {
        // ViscousForcesTwoStep (:1724-1768): m_q (S_p c_p + S_q c_q) . grad_p W with
        // c_p = rholoc_q / (rho_p^2 rho_q), c_q = rholoc_p / (rho_q^2 rho_p), the same
        // for either pair order
        const float cp = rlq * irp2 * irq, cq = rlp * irq * irq * irp;
        const float txx = sp.x * cp + sq.x * cq, txy = sp.y * cp + sq.y * cq, tyy = sp.z * cp + sq.z * cq;
        ax += mq * (txx * gx + txy * gy);
        ay += mq * (txy * gx + tyy * gy);
      } else {
        const float num = r0 * ((drlq + drq) - (drlp + drp)) + (drlq * drq - drlp * drp);
        const float dA = num * irp2 * irq * irq;
        const float sg = p_is_i ? dA : -dA;
        ax += mq * sg * sgx;
        ay += mq * sg * sgy;
      }
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

// The pair context of a scheme; UPH_MODE_REF2 is the reference scheme with
// stressType 2 (a template value only: such a handle's scenes say UPH_MODE_REF).
#define UPH_MODE_REF2 2
template <int MODE>
struct CtxSel {
  using type = PairCtx;
};
template <>
struct CtxSel<UPH_MODE_REF> {
  using type = RefCtxT<false>;
};
template <>
struct CtxSel<UPH_MODE_REF2> {
  using type = RefCtxT<true>;
};
template <int MODE>
using CtxOf = typename CtxSel<MODE>::type;

// The {localDensity, localPressure} mirror of slot q, read by the ref-mode scheme only.
template <int MODE>
__device__ __forceinline__ float2 mirror_c(const StepBufs& b, uint32_t q) {
  return MODE != UPH_MODE_CONFIG ? ld_off(b.mC, q) : make_float2(0.f, 0.f);
}

// The viscous stress of slot q, read with stressType 2 only.
template <int MODE>
__device__ __forceinline__ float4 mirror_s(const StepBufs& b, uint32_t q) {
  return MODE == UPH_MODE_REF2 ? ld_off(b.sst, q) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// A lazy handle's epilogue of slot p (after its state for the next step is written
// to the B set): the tag; the next step's predicted neighbour mirrors, by the same
// predict() the sort's reorder would run on this B-side state, so a step that is not
// re-sorted reads exactly what a re-sorted one would; and the stale test of the new
// position x1: the slot is stale when x1 lies outside the cell it was binned into
// (cellOf) grown by the margin m = 1 - 2h/cell (in cells; the grid's outer cells are
// open outward).  A query's 3x3 stencil is taken around its current cell, so while no
// particle is stale every neighbour within 2h lies in the stencil's binned cells: a
// neighbour binned two cells away would be more than (1 - m) cells = 2h away.  A stale
// slot marks its scene and the step (one atomic per scene and wavefront).
// Its loads come before its stores, and callers store the slot's state after it: on
// this GPU a wait for a load also waits for every store issued before it.
template <bool FLUID>
__device__ __forceinline__ void lazy_out(const DevScene& S, const StepBufs& b, int p, uint32_t tg, int itime,
                                         double2 x1, double2 v1, double rho1, float4 rt, double P1, float4 r2,
                                         float2 dv) {
  const int pk = b.cellOf[p];
  const Pred o = FLUID ? predict_fluid_next(S, v1, rho1, rt, P1, r2, dv)
                       : predict(S, b, tg, itime + 1, x1, v1, rho1, rt, P1, r2, dv);
  float4 ma, mb;
  float2 mc;
  mirror_of(S, o, rt.w, tg, ma, mb, mc);
  b.tagB[p] = tg;
  b.mA2[p] = ma;
  b.mB2[p] = mb;
  if (S.mode == 1) b.mC2[p] = mc;
  const int bx = pk & 0xffff, by = pk >> 16;
  const double u = (x1.x - S.xmin) * S.cell_inv, w = (x1.y - S.ymin) * S.cell_inv;
  const double m = S.skin_m;
  const bool stale = (bx > 0 && u < bx - m) || (bx < S.ncx - 1 && u >= bx + 1 + m) || (by > 0 && w < by - m) ||
                     (by < S.ncy - 1 && w >= by + 1 + m);
  const unsigned long long sm = __ballot(stale);
  if (sm == 0ull) return;
  const int par = itime & 1;
  const int lead = __ffsll((long long)sm) - 1;
  const int s0 = __shfl(S.sidx, lead, 64);
  if ((int)(threadIdx.x & 63) == lead) {
    atomicOr(&b.sstale[par * b.nscenes + s0], 1);
    atomicOr(&b.ctl[C_LSTALE + par], 1);
  } else if (stale && S.sidx != s0) {   // a wavefront of two scenes
    atomicOr(&b.sstale[par * b.nscenes + S.sidx], 1);
  }
}

// Boundary slot: wall state for the next step (deck + kinematics at t = (itime+1) dt).
// LAZY: a lazy re-sort handle's pair kernel (lazy_out, then the state).
template <bool LAZY = false>
__device__ __forceinline__ void wall_update(const DevScene& S, const StepBufs& b, int p, uint32_t tg, int itime,
                                            float mass) {
  const uint32_t id = tag_id(tg);
  double2 x = b.deckPos[id], v = b.deckVel[id];
  if (S.motion != 0) {
    double lvx;
    (void)wall_kin(S, b.wallKind[id], b.wallRef[id], S.t0 + (itime + 1) * S.dt, b.deckPos[id], x, v, &lvx);
  }
  const double rho = b.rhoA[p];
  const float4 rt = make_float4(0.f, 0.f, 0.f, mass);
  const double P = S.mode == 1 ? b.deckP[id] : 0.0;
  if (LAZY)
    lazy_out<false>(S, b, p, tg, itime, x, v, rho, rt, P, make_float4(0.f, 0.f, 0.f, 0.f), make_float2(0.f, 0.f));
  b.posB[p] = x;
  b.velB[p] = v;
  b.rhoB[p] = rho;
  b.rtB[p] = rt;
  if (S.mode == 1) b.presB[p] = P;
}

// Gravity and the FP64 leapfrog corrector (sph.cc:137-168) for fluid slot p (LAZY: a
// lazy re-sort handle's pair kernel, lazy_out, then the state).
template <bool LAZY = false>
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
  const double2 x1 = make_double2(xp.x + dt * v1x, xp.y + dt * v1y);
  const float4 rt = make_float4(Rt, (float)atx, (float)aty, mass);
  if (LAZY)
    lazy_out<true>(S, b, p, (1u << UPH_ID_BITS) | c.idp, itime, x1, make_double2(v1x, v1y), rho1, rt, 0.0,
             make_float4(0.f, 0.f, 0.f, 0.f), make_float2(0.f, 0.f));
  b.posB[p] = x1;
  b.velB[p] = make_double2(v1x, v1y);
  b.rhoB[p] = rho1;
  b.rtB[p] = rt;
  if (b.dbgRates) {
    // {R, R~, a~ (as the predictor reads them)} then the FP64 acceleration the
    // corrector used, in the same 32 bytes
    b.dbgRates[2 * p] = make_float4(R, Rt, (float)atx, (float)aty);
    *reinterpret_cast<double2*>(b.dbgRates + 2 * p + 1) = make_double2(ax, ay);
  }
}

// UPH_MODE_REF: gravity, the sponge layer and the FP64 leapfrog corrector
// (sph.cc:137-168) with DynamicPressure integrated like the density.
template <bool LAZY = false, bool TWO>
__device__ __forceinline__ void correct(const DevScene& S, const StepBufs& b, int p, int itime, double2 xp,
                                        double2 vn, float mass, const RefCtxT<TWO>& c) {
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
  const double2 x1 = make_double2(xp.x + dt * v1x, xp.y + dt * v1y);
  const float4 rt = make_float4(Rt, (float)atx, (float)aty, mass);
  const float4 r2 = make_float4(RPt, c.dr, c.dP, 0.f);
  const float2 dv = make_float2(c.dvx, c.dvy);
  if (LAZY)
    lazy_out<true>(S, b, p, (1u << UPH_ID_BITS) | c.idp, itime, x1, make_double2(v1x, v1y), rho1, rt, P1, r2, dv);
  b.posB[p] = x1;
  b.velB[p] = make_double2(v1x, v1y);
  b.rhoB[p] = rho1;
  b.presB[p] = P1;
  b.rtB[p] = rt;
  b.rt2B[p] = r2;
  b.dvB[p] = dv;
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
template <bool EXACT, int KT, int MODE, bool ONE = false, bool LAZY = false>
__device__ __forceinline__ bool pair_direct(int p, const StepBufs& b, const DevScene* __restrict__ sc) {
  const uint32_t tg = b.tagA[p];
  const DevScene& S = sc[ONE ? 0 : b.sceneOf[p]];
  const int itime = b.it[0];
  const double2 xp = b.posA[p];
  const double2 vn = b.velA[p];
  const float mass = b.mB[p].z;
  if (tag_type(tg) != 1) {
    wall_update<LAZY>(S, b, p, tg, itime, mass);
    return false;
  }
  {
    int rb[3], rl[3];
    candidate_rows(S, xp, b.cellStart, rb, rl);
    if (rl[0] + rl[1] + rl[2] > CAND_MAX) {   // a broken state: flag it, skip the query
      atomicOr(&b.ctl[C_BROKEN], 1);
      return false;
    }
  }
  CtxOf<MODE> c;
  c.init(S, tg, b.mA[p], b.mB[p], mirror_c<MODE>(b, (uint32_t)p), mirror_s<MODE>(b, (uint32_t)p));
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
    c.template add<KT>(S, ex, ey, tq, b.mA[q], bq, mirror_c<MODE>(b, (uint32_t)q), mirror_s<MODE>(b, (uint32_t)q));
  });
  if (!EXACT && S.nsec > 0 && cnt <= S.K) cnt += ghost_count(S, xp, b);
  if (!EXACT && cnt > S.K) return true;
  correct<LAZY>(S, b, p, itime, xp, vn, mass, c);
  return false;
}

// The pair sums of fluid particle p over the set bits of m[0..2] (slot bases
// rb[0..2]), in that order, into c.
template <int KT, int MODE, class Ctx = CtxOf<MODE>>
__device__ __forceinline__ void walk_pairs(const DevScene& S, const StepBufs& b, int p, double2 xp, Ctx& c,
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
  struct NbBase {
    double2 x;
    float4 a, b;
    float2 c;
  };
  struct NbS : NbBase {
    float4 s;
  };
  using Nb = typename std::conditional<MODE == UPH_MODE_REF2, NbS, NbBase>::type;
  auto load = [&](Nb& o, int q) {
    const uint32_t s = (uint32_t)(q >= 0 ? q : p);
    o.x = ld_off(b.posA, s);
    o.a = ld_off(b.mA, s);
    o.b = ld_off(b.mB, s);
    o.c = mirror_c<MODE>(b, s);
    if constexpr (MODE == UPH_MODE_REF2) o.s = mirror_s<MODE>(b, s);
  };
  auto pair = [&](const Nb& o) {
    const uint32_t tq = __float_as_uint(o.b.w);
    if (c.template admit<false>(S, tag_id(tq), 0.0, 0.0, 0u, b, 0)) {
      float4 sq = make_float4(0.f, 0.f, 0.f, 0.f);
      if constexpr (MODE == UPH_MODE_REF2) sq = o.s;
      c.template add<KT>(S, xp.x - o.x.x, xp.y - o.x.y, tq, o.a, o.b, o.c, sq);
    }
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
template <int KT, int MODE, bool LAZY = false>
__device__ __forceinline__ void compact_pairs(const DevScene& S, const StepBufs& b, int p, uint32_t tg, int itime,
                                              double2 xp, double2 vn, float mass, const int* rb,
                                              const unsigned long long* m) {
  CtxOf<MODE> c;
  c.init(S, tg, b.mA[p], b.mB[p], mirror_c<MODE>(b, (uint32_t)p), mirror_s<MODE>(b, (uint32_t)p));
  walk_pairs<KT, MODE>(S, b, p, xp, c, rb, m);
  correct<LAZY>(S, b, p, itime, xp, vn, mass, c);
}

// The capped path's pairs of fluid slot p by the same bit walk: pass 1 marks, in two
// 64-bit masks per candidate row (rows of up to 128 slots), the candidates in range
// that the K cap keeps; pass 2 walks them row by row, in slot order as the direct
// loop sums them.
// Returns false, leaving p to the direct loop, when a row is longer than 128 slots.
template <int KT, int MODE, bool ONE = false, bool LAZY = false>
__device__ __forceinline__ bool exact_pairs(int p, const StepBufs& b, const DevScene* __restrict__ sc) {
  const uint32_t tg = b.tagA[p];
  const DevScene& S = sc[ONE ? 0 : b.sceneOf[p]];
  const double2 xp = b.posA[p];
  int rb[3], rl[3];
  candidate_rows(S, xp, b.cellStart, rb, rl);
  if (rl[0] > 128 || rl[1] > 128 || rl[2] > 128) return false;
  const int itime = b.it[0];
  const double2 vn = b.velA[p];
  const float mass = b.mB[p].z;
  CtxOf<MODE> c;
  c.init(S, tg, b.mA[p], b.mB[p], mirror_c<MODE>(b, (uint32_t)p), mirror_s<MODE>(b, (uint32_t)p));
  double kpd2;
  uint32_t kpid;
  kth_of(b, p, kpd2, kpid);
  const double cut = S.cutoff;
  const uint32_t idp = tag_id(tg);
  const bool pb_ok = (int)idp < S.pb_gid;
  const int stamp = b.it[1];
  const bool have_mask = b.kmask != nullptr && b.kth[p].step == -stamp;
  // Row r's admitted candidates as bits of m[2r] (slots rb[r] + [0, 64)) and m[2r + 1]
  // (slots rb[r] + [64, 128)), by the rules of Search.cc:27-40 (admit<true>).  Pass 1a
  // tests distances and decides the candidates q with a larger original index, whose
  // admission needs only p's K-th key (q ranks within p's K nearest, q below
  // ParticleBoundary); it marks as pending those with a smaller one, whose admission
  // needs q's K-th key (p ranks within q's K nearest).  A query whose K-th key the
  // per-lane search found takes both sets from it instead.  Pass 1b reads the pending
  // candidates' keys, two in flight.
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
        // the per-lane K-th search left this word's admitted candidates above p and
        // its candidates below p, which wait for their own K-th keys
        am = b.kmask[(size_t)(2 * r + w) * b.kmask_n + p];
        pm = b.kmask[(size_t)(6 + 2 * r + w) * b.kmask_n + p];
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
  correct<LAZY>(S, b, p, itime, xp, vn, mass, c);
  return true;
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
// STG (every thread of the compacted kernel's block calling this): when the block's
// fluid slots share one scene and one cell row, pass 1 reads the candidates from a
// copy of the union of their three rows in LDS (spos, at most stg_max positions less a
// margin for the reads past a row's end; sred: 10 ints of LDS for the block's
// reduction).  A wavefront of two scenes returns only after the block's barriers.
template <int KT, int MODE, bool UNI, bool ONE = false, bool LAZY = false, bool STG = false>
__device__ __forceinline__ void compact_query(int n, int p, const StepBufs& b, const DevScene* __restrict__ sc,
                                              KthLds& lds, LaneSel* ls = nullptr, double2* spos = nullptr,
                                              int* sred = nullptr, int stg_max = 0) {
  const bool live = p < n;
  const int pp = live ? p : n - 1;
  const uint32_t tg = b.tagA[pp];
  const bool fluid = live && tag_type(tg) == 1;
  const int slane = ONE ? 0 : b.sceneOf[pp];
  const int sidx = ONE ? 0 : UNI ? __builtin_amdgcn_readfirstlane(slane) : slane;
  const bool mixed = !ONE && UNI && __ballot(slane != sidx) != 0ull;   // a wavefront of two scenes
  if (!STG && mixed) return;
  const DevScene& S = sc[sidx];
  const int itime = b.it[0];
  const double2 xp = b.posA[pp];
  const double cut = S.cutoff;
  int cx, cy;
  cell_of(S, xp.x, xp.y, cx, cy);   // in a wavefront of two scenes, a lane's rows below are not used
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
  }
  bool staged = false;
  int so[3] = {0, 0, 0};
  if constexpr (STG) {
    if (threadIdx.x == 0) {
      sred[0] = sred[2] = sred[3] = sred[4] = sred[8] = INT_MAX;
      sred[1] = sred[5] = sred[6] = sred[7] = sred[9] = INT_MIN;
    }
    __syncthreads();
    const int w0 = -wave_max(fluid ? -cy : -INT_MAX), w1 = wave_max(fluid ? cy : INT_MIN);
    const int w8 = -wave_max(fluid ? -slane : -INT_MAX), w9 = wave_max(fluid ? slane : INT_MIN);
    const int w2 = -wave_max(fluid ? -rb[0] : -INT_MAX), w3 = -wave_max(fluid ? -rb[1] : -INT_MAX);
    const int w4 = -wave_max(fluid ? -rb[2] : -INT_MAX);
    const int w5 = wave_max(fluid ? rb[0] + rl[0] : INT_MIN), w6 = wave_max(fluid ? rb[1] + rl[1] : INT_MIN);
    const int w7 = wave_max(fluid ? rb[2] + rl[2] : INT_MIN);
    if ((threadIdx.x & 63) == 0) {
      atomicMin(&sred[0], w0);
      atomicMax(&sred[1], w1);
      atomicMin(&sred[2], w2);
      atomicMin(&sred[3], w3);
      atomicMin(&sred[4], w4);
      atomicMax(&sred[5], w5);
      atomicMax(&sred[6], w6);
      atomicMax(&sred[7], w7);
      atomicMin(&sred[8], w8);
      atomicMax(&sred[9], w9);
    }
    __syncthreads();
    if (sred[0] == sred[1] && sred[8] == sred[9]) {
      // one cell row for every fluid slot: row r's ranges are contiguous cells of one
      // cell row, so their union is one range; a lane reads up to 66 slots past its
      // row's end (the wavefront's longest row, four per trip), kept inside spos
      const int lo0 = sred[2], lo1 = sred[3], lo2 = sred[4];
      const int len0 = sred[5] - lo0, len1 = sred[6] - lo1, len2 = sred[7] - lo2;
      if (len0 + len1 + len2 + 68 <= stg_max) {
        staged = true;
        for (int i = threadIdx.x; i < len0 + len1 + len2; i += blockDim.x)
          spos[i] = b.posA[i < len0 ? lo0 + i : i < len0 + len1 ? lo1 + (i - len0) : lo2 + (i - len0 - len1)];
        so[0] = -lo0;
        so[1] = len0 - lo1;
        so[2] = len0 + len1 - lo2;
        if (!fluid) {
          // its rows are empty, but pass 1 tests the wavefront's trip count of slots
          // from their starts (bits cleared after): start them inside the copy
          rb[0] = lo0;
          rb[1] = lo1;
          rb[2] = lo2;
        }
      }
    }
    __syncthreads();
    if (mixed) return;
  }
  auto pass1 = [&](auto at) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      // the trip count is the wave's maximum row length (uniform, in an SGPR); slots
      // past this lane's row end are tested too and their bits cleared below
      const int lmax = __builtin_amdgcn_readfirstlane(wave_max(rl[r]));
      int kr = rb[r];
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
          for (int u = 0; u < 4; ++u) xs[u] = at(r, kr + u);   // posA has POS_PAD zeroed slots past n
          kr += 4;
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
  };
  if (staged)
    pass1([&](int r, int s) { return spos[so[r] + s]; });
  else
    pass1([&](int, int s) { return b.posA[s]; });
  // every lane reaches wave_overflow: the work of a lane is a call of this
  bool ovf = false;
  auto run = [&]() {
    if (!live) return;
    const double2 vn = b.velA[p];
    const float mass = b.mB[p].z;
    if (!fluid) {
      wall_update<LAZY>(S, b, p, tg, itime, mass);
      return;
    }
    // drop the particle itself (always in range, in the row of the cell it was binned
    // into: its current cell's, or with a lazy re-sort possibly the row above or below)
    if (LAZY) {
      // written out per row: a loop over the row arrays here keeps them in memory,
      // and the compiler then moves them to LDS and reads the walk's row bases there
      const int s0 = p - rb[0], s1 = p - rb[1], s2 = p - rb[2];
      if (s0 >= 0 && s0 < rl[0]) m[0] &= ~(1ull << s0);
      if (s1 >= 0 && s1 < rl[1]) m[1] &= ~(1ull << s1);
      if (s2 >= 0 && s2 < rl[2]) m[2] &= ~(1ull << s2);
    } else {
      const int self = p - rb[1];
      if (self >= 0 && self < 64) m[1] &= ~(1ull << self);
    }
    int cnt = __popcll(m[0]) + __popcll(m[1]) + __popcll(m[2]);
    if (S.nsec > 0 && !bad && cnt <= S.K) cnt += ghost_count(S, xp, b);   // mirror rows rank too
    if (bad || cnt > S.K) {
      ovf = true;   // the exact kernel recomputes this step
      return;
    }
    compact_pairs<KT, MODE, LAZY>(S, b, p, tg, itime, xp, vn, mass, rb, m);
  };
  run();
  wave_overflow<true>(b, sc, p, ovf, lds, sidx, xp, rb, rlf, ls, GlobalPos{b.posA});
}

// ---------------------------------------------------------------- search

// Pairs of query p that SearchNeighbors emits (Search.cc:27-40): in range,
// j > i, j < ParticleBoundary, key <= the K-th key (kd2, kid).
__device__ __forceinline__ bool search_keeps(const DevScene& S, uint32_t idp, double d2, uint32_t idq, double kd2,
                                             uint32_t kid) {
  return idq > idp && (int)idq < S.pb_gid && key_le(d2, idq, kd2, kid);
}

// ================================================================== host

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
inline int dalloc(T** p, size_t count, Err& err) {
  *p = nullptr;
  if (count == 0) count = 1;
  HIPCHK(hipMalloc((void**)p, count * sizeof(T)), err);
  return 0;
}

inline int grid_of(int n) { return (n + BLK - 1) / BLK; }
inline int grid8_of(int n) { return ((grid_of(n) + 7) / 8) * 8; }

struct CellScan {
  int m = 0, nb = 0;
};


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
  // every scene's queries are expected to have more than K in range (4 pi (h/dx)^2
  // above K) and no scene has mirror rows: the capped pass recomputes every slot of
  // every step, so the step's first kernel only finds the K-th keys (k_kth)
  bool kth_all = false;
  // the scenes take stressType 2 (UPH_MODE_REF2 kernels, k_stress before the pair pass)
  bool stress2 = false;
  // lazy re-sort (b.lazy): the next step sorts the state from scratch (after uph_create
  // or a state set from outside), gathering the B set of view sort_view
  bool need_sort = true;
  int sort_view = 0;
};

// Host functions of uph_step.hip that the other translation units call.
void fill_scene(DevScene& S, const uph_scene& u, double cell_factor);
int alloc_bufs(StepBufs& b, int n, int ncells, int nblocks_scan, Err& err, bool search_only, bool capped = false);
void free_bufs(StepBufs& b);
int launch_scan(hipStream_t st, int m, int nb, int* count, int* partial, int* start);
int launch_sort(hipStream_t st, int n, int ncells, int nb, StepBufs& b, const DevScene* dsc, int search_only,
                bool step_ctl, Err& err, int view = 0);
int launch_step(uph_handle* h, hipEvent_t pa, hipEvent_t pb, hipEvent_t pc = nullptr);
int launch_call_step(uph_handle* h, int k, int n_steps, hipEvent_t pa, hipEvent_t pb, hipEvent_t pc);
// uph_pair.hip: the pair pass of a step (both views for a lazy handle), and a
// sort-every-step handle's capped pass.
int launch_pair_pass(uph_handle* h, hipStream_t st);
int launch_stress(uph_handle* h, hipStream_t st);
int launch_capped(uph_handle* h, hipStream_t st);
// The buffers as the state getters read them after the last call (waits for the
// stream): the A set holds the last step's input and B its output, in one slot order.
// A lazy handle's are those of the view its last pair kernel ran with.
int read_bufs(uph_t* h, StepBufs& out);
// The state of a lazy handle is about to be set from outside: the next step sorts it
// from the B set of the current read view.
int lazy_reset(uph_t* h);
int get_rates_f64(uph_t* h, int scene, float* R, float* Rt, double* a, float* at);
// uph_diag.hip: approxVorticity of scene `scene` of the last evaluation, copied to host out.
int vort_run(uph_t* h, int scene, double* out);
