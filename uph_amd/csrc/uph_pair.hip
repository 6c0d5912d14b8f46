// uph_pair.hip -- the step's pair kernels (compacted, direct and capped passes, the
// wavefronts of two scenes) and their launches; the device code is in uph_internal.h.
#include <climits>

#include "uph_internal.h"

// EXACT (the capped path) runs as a small grid-stride launch that returns at
// once unless the step's overflow flag is set.  ONE: a one-scene handle, whose scene
// constants stay in scalar registers (the capped reference-scheme pass spills 124
// bytes per lane instead of 160 and runs 4 % faster).
template <bool EXACT, int KT, int MODE, bool ONE = false, bool LAZY = false>
// Four wavefronts per SIMD (128 VGPRs): the configuration scheme needs 129, the
// reference scheme 179, which at this target spills and still runs the capped step
// faster than with three wavefronts (no spill) or five.
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4)))
void k_pair(int n, StepBufs b0, const DevScene* __restrict__ sc) {
  const StepBufs b = LAZY ? pair_view(b0) : b0;
  if (EXACT && (b.ctl[C_OVF] == 0 || b.ctl[C_BROKEN] != 0)) return;
  if (EXACT && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&b.ctl[C_NOVF], 1);
  __shared__ KthLds lds[EXACT ? 1 : BLK / 64];
  const int nlog = (n + BLK - 1) / BLK;
  // the trip count is the same for the whole block, so every wavefront reaches
  // wave_overflow with all of its lanes
  for (int lb = xcd_block(); lb < nlog; lb += gridDim.x) {
    const int p = lb * BLK + threadIdx.x;
    bool ovf = false;
    if (p < n) {
      // the capped step's fluid rows take the bit walk when their rows fit it
      const bool walked = EXACT && tag_type(b.tagA[p]) == 1 && exact_pairs<KT, MODE, ONE, LAZY>(p, b, sc);
      if (!walked) ovf = pair_direct<EXACT, KT, MODE, ONE, LAZY>(p, b, sc);
    }
    if (!EXACT) wave_overflow<false>(b, sc, p, ovf, lds[threadIdx.x >> 6]);
  }
}

// LANE: the handle's queries are expected to be capped, so the kernel carries the
// per-lane K-th key search and its LDS.
template <int PB, int KT, int MODE, bool ONE, bool LANE, bool LAZY>
// Occupancy targets, each measured faster than the compiler's own choice: the
// reference scheme's sums need about 133 VGPRs, capped at 128 (4 wavefronts per SIMD
// instead of 3, a few spilled registers); with the per-lane K-th search (LANE) the
// configuration scheme runs 6 wavefronts per SIMD (80 VGPRs) instead of 5.  The
// default instance (91 VGPRs, 5 wavefronts) measured slower at 6.
__global__ __launch_bounds__(PB)
__attribute__((amdgpu_waves_per_eu(MODE != UPH_MODE_CONFIG ? 4 : LANE ? 6 : 1)))
void k_pair_compact(int n, StepBufs b0, const DevScene* __restrict__ sc) {
  const StepBufs b = LAZY ? pair_view(b0) : b0;
  const int p = xcd_block() * PB + threadIdx.x;
  if constexpr (LANE) {
    __shared__ KthLaneLds lds[PB / 64];
    compact_query<KT, MODE, true, ONE, LAZY>(n, p, b, sc, lds[threadIdx.x >> 6].k, &lds[threadIdx.x >> 6].l);
  } else {
    // pass 1 reads its candidates from the block's staged rows where they fit (20 KB,
    // within the LDS of the kernel's five wavefronts per SIMD)
    constexpr int STG_N = PB == 128 ? 624 : 1248;
    __shared__ KthLds lds[PB / 64];
    __shared__ double2 spos[STG_N];
    __shared__ int sred[10];
    compact_query<KT, MODE, true, ONE, LAZY, true>(n, p, b, sc, lds[threadIdx.x >> 6], nullptr, spos, sred, STG_N);
  }
}

// The wavefronts whose slots belong to two scenes (listed by uph_create), with the
// scene taken per lane: one wavefront per block.
template <int KT, int MODE, bool LAZY>
__global__ __launch_bounds__(64) void k_pair_mixed(int n, StepBufs b0, const DevScene* __restrict__ sc,
                                                   const int* __restrict__ waves) {
  const StepBufs b = LAZY ? pair_view(b0) : b0;
  __shared__ KthLds lds[1];
  compact_query<KT, MODE, false, false, LAZY>(n, waves[blockIdx.x] * 64 + (int)threadIdx.x, b, sc, lds[0]);
}

// The step's first kernel for handles whose every query is expected to have more than
// K in range and that have no mirror rows (uph_handle::kth_all): the capped pass
// recomputes every slot of such a step, so this kernel only finds each fluid query's
// K-th key and pair partners (lane_kth; the wavefront search for the queries it leaves)
// and raises the step's overflow flag.  Each lane takes its own scene, so wavefronts
// of two scenes need no other pass.
//
// The search is bound by its candidate loads (each a 16-byte position per lane; about
// 720 per wavefront, served from L1 at about 21 cache accesses each), not by their
// misses (1 % of them reach L2) or its arithmetic.  So a one-scene block whose fluid
// slots share one cell row copies the union of their three candidate rows into LDS
// with coalesced loads (KTH_STAGE positions at most, else it reads posA as before),
// and the per-lane search reads its candidates there.
//
// Its registers would allow six wavefronts per SIMD; it runs four (KTH_BLOCK_LDS),
// since with more the candidate positions of the resident wavefronts outgrow the
// caches: on the capped 1 M reference-scheme step it took 684 us at six, 594 us at four
// and 605 us at three wavefronts per SIMD in one run, and 1,061 us at eight against
// 667 us at six in another.
constexpr int KTH_BLOCK_LDS = 40 * 1024;   // a quarter of the CU's 160 KB: four blocks of four wavefronts
constexpr int KTH_STAGE = 1272;            // staged positions: 20,352 bytes with the search's 20,480
static_assert(sizeof(KthLaneLds) * (BLK / 64) + sizeof(double2) * KTH_STAGE + 8 * sizeof(int) <= KTH_BLOCK_LDS,
              "k_kth LDS");
template <bool ONE>
__global__ __launch_bounds__(BLK) void k_kth(int n, StepBufs b, const DevScene* __restrict__ sc) {
  __shared__ KthLaneLds lds[BLK / 64];
  __shared__ double2 spos[KTH_STAGE];
  // the block's fluid cell rows (min, max) and, per candidate row r, the union of the
  // fluid slots' ranges [lo_r, hi_r)
  __shared__ int sred[8];
  const int p = xcd_block() * BLK + threadIdx.x;
  const bool live = p < n;
  const int pp = live ? p : n - 1;
  const bool fluid = live && tag_type(b.tagA[pp]) == 1;
  const int sidx = ONE ? 0 : b.sceneOf[pp];
  const DevScene& S = sc[sidx];
  const double2 xp = b.posA[pp];
  int rb[3], rl[3];
  candidate_rows(S, xp, b.cellStart, rb, rl);
  bool staged = false;
  int so[3] = {0, 0, 0}, hi[2] = {0, 0};
  if constexpr (ONE) {
    int cx, cy;
    cell_of(S, xp.x, xp.y, cx, cy);
    if (threadIdx.x == 0) {
      sred[0] = INT_MAX;
      sred[1] = INT_MIN;
      sred[2] = sred[3] = sred[4] = INT_MAX;
      sred[5] = sred[6] = sred[7] = INT_MIN;
    }
    __syncthreads();
    // wavefront minima and maxima first (a non-fluid lane contributes nothing), then
    // one LDS atomic per wavefront and value
    const int v[8] = {fluid ? cy : INT_MAX, fluid ? cy : INT_MIN, fluid ? rb[0] : INT_MAX,
                      fluid ? rb[1] : INT_MAX, fluid ? rb[2] : INT_MAX, fluid ? rb[0] + rl[0] : INT_MIN,
                      fluid ? rb[1] + rl[1] : INT_MIN, fluid ? rb[2] + rl[2] : INT_MIN};
    const int w0 = -wave_max(-v[0]), w1 = wave_max(v[1]);
    const int w2 = -wave_max(-v[2]), w3 = -wave_max(-v[3]), w4 = -wave_max(-v[4]);
    const int w5 = wave_max(v[5]), w6 = wave_max(v[6]), w7 = wave_max(v[7]);
    if ((threadIdx.x & 63) == 0) {
      atomicMin(&sred[0], w0);
      atomicMax(&sred[1], w1);
      atomicMin(&sred[2], w2);
      atomicMin(&sred[3], w3);
      atomicMin(&sred[4], w4);
      atomicMax(&sred[5], w5);
      atomicMax(&sred[6], w6);
      atomicMax(&sred[7], w7);
    }
    __syncthreads();
    // one cell row for every fluid slot of the block (and some fluid slot): row r's
    // ranges are contiguous cells of one cell row, so their union is one range
    int lo0 = 0, lo1 = 0, lo2 = 0, len0 = 0, len1 = 0, len2 = 0;
    if (sred[0] == sred[1]) {
      lo0 = sred[2];
      lo1 = sred[3];
      lo2 = sred[4];
      len0 = sred[5] - lo0;
      len1 = sred[6] - lo1;
      len2 = sred[7] - lo2;
      staged = len0 + len1 + len2 <= KTH_STAGE;
    }
    if (staged) {
      for (int i = threadIdx.x; i < len0 + len1 + len2; i += BLK)
        spos[i] = b.posA[i < len0 ? lo0 + i : i < len0 + len1 ? lo1 + (i - len0) : lo2 + (i - len0 - len1)];
      so[0] = -lo0;
      so[1] = len0 - lo1;
      so[2] = len0 + len1 - lo2;
      hi[0] = lo0 + len0;
      hi[1] = lo1 + len1;
    }
    __syncthreads();
  }
  // every lane of the wavefront reaches this, fluid or not
  if (staged)
    wave_overflow<true>(b, sc, p, fluid, lds[threadIdx.x >> 6].k, sidx, xp, rb, rl, &lds[threadIdx.x >> 6].l,
                        StagedPos{spos, {so[0], so[1], so[2]}, hi[0], hi[1]});
  else
    wave_overflow<true>(b, sc, p, fluid, lds[threadIdx.x >> 6].k, sidx, xp, rb, rl, &lds[threadIdx.x >> 6].l,
                        GlobalPos{b.posA});
}

// ClassicShearStress (derivatives.c:1686-1721) in gather form for stressType 2, before
// the pair pass: S_p = -mu sum_q m_q rholoc_q / (rho_q rho_p) (D g^T + g D^T) with
// D = v_p - vloc_q and g = grad W(x_p - x_q), the same for either order of the pair,
// over p's pairs: a fluid slot's admitted neighbours (Search.cc:33), a wall slot's
// fluid neighbours (a wall is only ever a pair's second particle), every neighbour in
// range while no query is capped.  FP32 terms from the step's mirrors, FP64 sums in
// slot order.  The sums of one slot, for walk_pairs (or the direct loop).
struct StressCtx : PairBase {
  bool fluid;
  int fend;   // first id past the scene's fluid particles
  float vpx, vpy, r0, irp;
  CSum sxx, sxy, syy;

  __device__ __forceinline__ void init(const DevScene& S, uint32_t tg, float4 ap, float4 bp, float2, float4) {
    init_base(S, tg);
    fluid = tag_type(tg) == 1;
    fend = S.poff + S.nf;
    vpx = ap.x;
    vpy = ap.y;
    r0 = S.frho0;
    irp = 1.0f / (r0 + bp.x);
  }

  template <bool EXACT>
  __device__ __forceinline__ bool admit(const DevScene&, uint32_t idq, double, double, uint32_t, const StepBufs&,
                                        int) const {
    return fluid ? (idq > idp ? (int)idq < pb : pb_ok) : (int)idq < fend;
  }

  template <int KT>
  __device__ __forceinline__ void add(const DevScene&, double ex, double ey, uint32_t, float4 aq, float4 bq,
                                      float2 cq, float4) {
#pragma clang fp contract(off)
    const float dx = (float)ex, dy = (float)ey;
    const float r2 = dx * dx + dy * dy;
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float f = grad_factor<KT>(r2, r2 * rinv, rinv);
    const float gx = f * dx, gy = f * dy;
    const float rhoq = r0 + bq.x, rlq = r0 + cq.x;
    const float vr = bq.z * rlq * __builtin_amdgcn_rcpf(rhoq) * irp;
    const float Dx = vpx - aq.z, Dy = vpy - aq.w;
    sxx.add(vr * (2.0f * Dx * gx));
    sxy.add(vr * (Dx * gy + Dy * gx));
    syy.add(vr * (2.0f * Dy * gy));
  }
};

// One thread per slot; the compacted kernel's pass 1 (bit masks of the in-range
// candidates of the three rows, a trip count uniform over the wavefront) and its bit
// walk; a slot whose row is longer than 64 slots takes the direct loop.  A fluid query
// with more than K in range sets C_SCAP (uph_sync reports it).  The result is
// {xx, xy (= yx), yy, 0}.
template <int KT, bool ONE>
__global__ __launch_bounds__(BLK) void k_stress(int n, StepBufs b, const DevScene* __restrict__ sc) {
  const int p = blockIdx.x * BLK + threadIdx.x;
  const bool live = p < n;
  const int pp = live ? p : n - 1;
  const DevScene& S = sc[ONE ? 0 : b.sceneOf[pp]];
  const uint32_t tg = b.tagA[pp];
  StressCtx c;
  c.init(S, tg, b.mA[pp], b.mB[pp], make_float2(0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f));
  const bool act = live && S.stress2 && (c.fluid || c.pb_ok);   // a wall at or past ParticleBoundary is in no pair
  const double2 xp = b.posA[pp];
  const double cut = S.cutoff;
  int cx, cy;
  cell_of(S, xp.x, xp.y, cx, cy);
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx < S.ncx - 1 ? cx + 1 : S.ncx - 1;
  int rb[3], rl[3];
  unsigned long long m[3];
  bool bad = false;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int row = cy - 1 + r;
    rb[r] = 0;
    rl[r] = 0;
    m[r] = 0ull;
    if (act && row >= 0 && row < S.ncy) {
      const int base = S.coff + row * S.ncx;
      rb[r] = b.cellStart[base + x0];
      rl[r] = b.cellStart[base + x1 + 1] - rb[r];
    }
    if (rl[r] > 64) {
      bad = true;
      rl[r] = 0;
    }
    const int lmax = __builtin_amdgcn_readfirstlane(wave_max(rl[r]));
    const double2* rp = b.posA + rb[r];
    unsigned long long bits = 0ull;
#pragma unroll 1
    for (int k = 0; k < lmax; k += 4) {
      double2 xs[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) xs[u] = rp[k + u];   // posA has POS_PAD zeroed slots past n
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double d2 = dist2(xp, xs[u]);
        bits |= (unsigned long long)(d2 < cut && k + u < rl[r]) << ((k + u) & 63);   // 0 past the row's end
      }
    }
    m[r] = bits;
  }
  // the slot itself, in whichever row holds it (rows written out, as in compact_query)
  const int s0 = p - rb[0], s1 = p - rb[1], s2 = p - rb[2];
  if (s0 >= 0 && s0 < rl[0]) m[0] &= ~(1ull << s0);
  if (s1 >= 0 && s1 < rl[1]) m[1] &= ~(1ull << s1);
  if (s2 >= 0 && s2 < rl[2]) m[2] &= ~(1ull << s2);
  if (act && !bad) {
    if (c.fluid && __popcll(m[0]) + __popcll(m[1]) + __popcll(m[2]) > S.K) atomicOr(&b.ctl[C_SCAP], 1);
    walk_pairs<KT, UPH_MODE_REF, StressCtx>(S, b, p, xp, c, rb, m);
  } else if (act) {
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
      if (!c.template admit<false>(S, tag_id(tq), d2, 0.0, 0u, b, q)) return;
      c.template add<KT>(S, ex, ey, tq, b.mA[q], bq, b.mC[q], make_float4(0.f, 0.f, 0.f, 0.f));
    });
    if (c.fluid && cnt > S.K) atomicOr(&b.ctl[C_SCAP], 1);
  }
  if (live) {
    const float mu = S.fmu;
    b.sst[p] = act ? make_float4(-mu * c.sxx.val(), -mu * c.sxy.val(), -mu * c.syy.val(), 0.f)
                   : make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

int launch_stress(uph_handle* h, hipStream_t st) {
  const bool one = h->nscenes == 1;
  if (h->kernel == UPH_KERNEL_POLY7) {
    if (one)
      hipLaunchKernelGGL((k_stress<UPH_KERNEL_POLY7, true>), dim3(grid_of(h->n)), dim3(BLK), 0, st, h->n, h->b, h->dsc);
    else
      hipLaunchKernelGGL((k_stress<UPH_KERNEL_POLY7, false>), dim3(grid_of(h->n)), dim3(BLK), 0, st, h->n, h->b,
                         h->dsc);
  } else {
    if (one)
      hipLaunchKernelGGL((k_stress<UPH_KERNEL_CUBIC, true>), dim3(grid_of(h->n)), dim3(BLK), 0, st, h->n, h->b, h->dsc);
    else
      hipLaunchKernelGGL((k_stress<UPH_KERNEL_CUBIC, false>), dim3(grid_of(h->n)), dim3(BLK), 0, st, h->n, h->b,
                         h->dsc);
  }
  HIPCHK(hipGetLastError(), h->err);
  return 0;
}

// The pair pass of one smoothing kernel KT and scheme MODE (compacted or direct, and
// the wavefronts of two scenes) over the buffers b; LAZY: a lazy re-sort handle's (no
// such handle is expected to be capped).
template <int KT, int MODE, bool LAZY>
static int launch_pair_pass_b(uph_handle* h, hipStream_t st, const StepBufs& b) {
  if (h->opts.pair_variant == 1) {
    hipLaunchKernelGGL((k_pair<false, KT, MODE, false, LAZY>), dim3(grid8_of(h->n)), dim3(BLK), 0, st, h->n, b,
                       h->dsc);
  } else if (!LAZY && MODE != UPH_MODE_REF2 && h->kth_all) {
    // the pair sums of such a step all come from the capped pass; the dynamic LDS only
    // brings the block to KTH_BLOCK_LDS (the one-scene instance also holds the staged rows)
    const int lane_lds = (int)(sizeof(KthLaneLds) * (BLK / 64));
    const int stage_lds = (int)(sizeof(double2) * KTH_STAGE + 8 * sizeof(int));
    if (h->nscenes == 1)
      hipLaunchKernelGGL((k_kth<true>), dim3(grid8_of(h->n)), dim3(BLK), KTH_BLOCK_LDS - lane_lds - stage_lds, st,
                         h->n, b, h->dsc);
    else
      hipLaunchKernelGGL((k_kth<false>), dim3(grid8_of(h->n)), dim3(BLK), KTH_BLOCK_LDS - lane_lds, st, h->n, b,
                         h->dsc);
  } else {
    const int pbk = h->opts.pair_block;
    const int g = (((h->n + pbk - 1) / pbk + 7) / 8) * 8;
    // a stressType 2 handle is never expected to be capped: no per-lane K-th instance
    const bool one = h->nscenes == 1, lane = !LAZY && MODE != UPH_MODE_REF2 && h->capped;
#define UPH_COMPACT(PBV, ONEV, LANEV)                                                                        \
  hipLaunchKernelGGL((k_pair_compact<PBV, KT, MODE, ONEV, LANEV, LAZY>), dim3(g), dim3(PBV), 0, st, h->n, b, h->dsc)
#define UPH_COMPACT_PB(ONEV, LANEV)              \
  do {                                           \
    if (pbk == 128) UPH_COMPACT(128, ONEV, LANEV); \
    else if (pbk == 512) UPH_COMPACT(512, ONEV, LANEV); \
    else UPH_COMPACT(256, ONEV, LANEV);          \
  } while (0)
    if (one && lane) UPH_COMPACT_PB(true, !LAZY && MODE != UPH_MODE_REF2);
    else if (one) UPH_COMPACT_PB(true, false);
    else if (lane) UPH_COMPACT_PB(false, !LAZY && MODE != UPH_MODE_REF2);
    else UPH_COMPACT_PB(false, false);
#undef UPH_COMPACT_PB
#undef UPH_COMPACT
    if (h->nmixed > 0)
      hipLaunchKernelGGL((k_pair_mixed<KT, MODE, LAZY>), dim3(h->nmixed), dim3(64), 0, st, h->n, b, h->dsc,
                         h->mixedWaves);
  }
  HIPCHK(hipGetLastError(), h->err);
  return 0;
}

template <int KT, int MODE>
static int launch_pair_pass_kt(uph_handle* h, hipStream_t st) {
  if constexpr (MODE == UPH_MODE_REF2) return launch_pair_pass_b<KT, MODE, false>(h, st, h->b);   // sorts every step
  else
    return h->b.lazy ? launch_pair_pass_b<KT, MODE, true>(h, st, h->b)
                     : launch_pair_pass_b<KT, MODE, false>(h, st, h->b);
}

// The exact capped fixup of a sort-every-step handle (returns at once unless the step
// overflowed).
template <int KT, int MODE>
static int launch_capped_kt(uph_handle* h, hipStream_t st) {
  const int gfix = std::min(grid8_of(h->n), FIX_GRID);
  if (h->nscenes == 1)
    hipLaunchKernelGGL((k_pair<true, KT, MODE, true>), dim3(gfix), dim3(BLK), 0, st, h->n, h->b, h->dsc);
  else
    hipLaunchKernelGGL((k_pair<true, KT, MODE>), dim3(gfix), dim3(BLK), 0, st, h->n, h->b, h->dsc);
  HIPCHK(hipGetLastError(), h->err);
  return 0;
}

int launch_pair_pass(uph_handle* h, hipStream_t st) {
  if (h->stress2)
    return h->kernel == UPH_KERNEL_POLY7 ? launch_pair_pass_kt<UPH_KERNEL_POLY7, UPH_MODE_REF2>(h, st)
                                         : launch_pair_pass_kt<UPH_KERNEL_CUBIC, UPH_MODE_REF2>(h, st);
  if (h->mode == UPH_MODE_REF)
    return h->kernel == UPH_KERNEL_POLY7 ? launch_pair_pass_kt<UPH_KERNEL_POLY7, UPH_MODE_REF>(h, st)
                                         : launch_pair_pass_kt<UPH_KERNEL_CUBIC, UPH_MODE_REF>(h, st);
  return h->kernel == UPH_KERNEL_POLY7 ? launch_pair_pass_kt<UPH_KERNEL_POLY7, UPH_MODE_CONFIG>(h, st)
                                       : launch_pair_pass_kt<UPH_KERNEL_CUBIC, UPH_MODE_CONFIG>(h, st);
}

int launch_capped(uph_handle* h, hipStream_t st) {
  if (h->stress2)
    return h->kernel == UPH_KERNEL_POLY7 ? launch_capped_kt<UPH_KERNEL_POLY7, UPH_MODE_REF2>(h, st)
                                         : launch_capped_kt<UPH_KERNEL_CUBIC, UPH_MODE_REF2>(h, st);
  if (h->mode == UPH_MODE_REF)
    return h->kernel == UPH_KERNEL_POLY7 ? launch_capped_kt<UPH_KERNEL_POLY7, UPH_MODE_REF>(h, st)
                                         : launch_capped_kt<UPH_KERNEL_CUBIC, UPH_MODE_REF>(h, st);
  return h->kernel == UPH_KERNEL_POLY7 ? launch_capped_kt<UPH_KERNEL_POLY7, UPH_MODE_CONFIG>(h, st)
                                       : launch_capped_kt<UPH_KERNEL_CUBIC, UPH_MODE_CONFIG>(h, st);
}
