// uph_kernels.h -- device data layout and kernels of the SPH hot path (gfx950).
//
// Layout in HBM (structure of arrays, one slot per particle, every scene of a
// batch concatenated; slots are re-sorted every step by (scene, cell, id)):
//   pos    double2  position (FP64)
//   vel    double2  velocity v_n at the start of the step (FP64)
//   rho    double   density rho_n (FP64)
//   rt     float4   {Rt, at.x, at.y, mass}: predicting rates of the previous step
//   tag    uint32   itype << 29 | global original index
//   mA     float4   {v.x, v.y, vloc.x, vloc.y} after the predictor (FP32 mirror)
//   mB     float4   {rho, P/rho^2, mass, tag bits} after the predictor (FP32 mirror)
// Buffers "A" are the sorted input of the pair kernel, buffers "B" its output
// in the same slot order; the reorder of the next step gathers B into A.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define UPH_ID_BITS 29
#define UPH_ID_MASK ((1u << UPH_ID_BITS) - 1u)

struct DevScene {
  // grid
  double xmin, ymin, cell_inv;
  int ncx, ncy, coff;       // cells, first global cell
  int poff, np, nf;         // first slot, particles, fluid particles
  int pb_gid;               // exclusive bound on a neighbour's global index
  int K, gravity, motion, gamma_int;
  double cutoff;            // squared search radius, (h*h)*4
  double h, dt, rho0, gamma, B, c0, g, r0;
  double m_amp, m_omega;
  double t0;                // time of step 0: wall kinematics run at t0 + itime * dt
  // FP32 copies for the pair sums
  float fh, finvh, fwalpha, fav, fc0, fk, fcd, fr0, frho0;
  float fk7;                // PolyOrderSeven gradient factor -2145 / (192512 pi h^4)
  float fw7;                // PolyOrderSeven W factor 2145 / (1347584 pi h^2)
  int kernel;               // UPH_KERNEL_*
  // UPH_MODE_REF
  int mode;
  double eps;               // weight of the velocity fluctuation in localVelocity
  float fK0, fgamma, fgk, fkap, fmu;  // rho0 c^2, gamma, gamma kappa, kappa, viscosity0
  // lid / gate speed, the gate's stop time, the sloshing tank's right-wall offset
  double lid_u, t_stop, flat_bottom;
  // mirror ("virtual") particles: the section codes of the scene's virtualType in
  // order (UPH_GS_*); its section tests as bits, one word per 32 fluid rows
  // (nwords) and section, section-major from word gw of the step's bit array; the id
  // of its first mirror row (after its np rows); a sloshing moving set whose loop
  // zeroes LeftBoundaryPosition; the boundary positions where the motion does not
  // set them; LocalDepth, dx
  int nsec, sec[5], gw, nwords, gbase, slosh_halt;
  double lbp0, rbp0, depth, dx;
  // lazy re-sort: the scene's index in its batch, and the margin (in cells) by which a
  // particle may leave the cell it was binned into before its scene is re-sorted
  int sidx;
  double skin_m;
  // stressType 2 (reference scheme): ClassicShearStress and ViscousForcesTwoStep
  int stress2;
};

// Sections of derivatives.c:3193-4192 (one loop over the fluid rows each).
enum {
  UPH_GS_LEFT0 = 1,  // x about 0: |x| < 2h, (-|x|, y)
  UPH_GS_LEFT,       // x about LeftBoundaryPosition: |x - L| < 2h, (L - |x - L|, y)
  UPH_GS_BOTTOM,     // y about 0: y < 2h and x <= FlatBottomLength, (x, -y)
  UPH_GS_CORNER_L,   // |x - L| < 2h and y < 2h: (L - |x - L|, -y)
  UPH_GS_CORNER_R,   // |x - R| < 2h and y < 2h: (R + |x - R|, -y)
  UPH_GS_RIGHT_FB,   // |x - FBL| < 2h: (FBL + |x - FBL|, y)
  UPH_GS_RIGHT,      // |x - R| < 2h: (R + |x - R|, y)
  UPH_GS_TOP_DX,     // |y - H - dx| < 2h: (x, H + 2 |y - H - dx|)
  UPH_GS_TOP,        // |y - H| < 2h: (x, H + 2 |y - H|)
  UPH_GS_SLOPE       // reflection in y = (x - FBL) / 15 within 3h, x > FBL
};

// Section list of a virtualType (derivatives.c:3195, 3471, 3626, 3783, 3907, 4032;
// virtualType 2's are commented out); returns its length.
__host__ __device__ inline int mirror_sections(int vtype, int* sec) {
  int m = 0;
  auto put = [&](int a, int b, int c, int d, int e) {
    const int v[5] = {a, b, c, d, e};
    for (int k = 0; k < 5; ++k)
      if (v[k]) sec[m++] = v[k];
  };
  switch (vtype) {
    case 1: put(UPH_GS_LEFT, UPH_GS_BOTTOM, UPH_GS_CORNER_L, UPH_GS_SLOPE, 0); break;
    case 3: put(UPH_GS_LEFT0, UPH_GS_BOTTOM, UPH_GS_TOP_DX, UPH_GS_CORNER_L, UPH_GS_RIGHT_FB); break;
    case 4:
    case 8: put(UPH_GS_LEFT0, UPH_GS_BOTTOM, UPH_GS_CORNER_L, UPH_GS_RIGHT_FB, 0); break;
    case 5: put(UPH_GS_LEFT, UPH_GS_CORNER_L, UPH_GS_CORNER_R, UPH_GS_RIGHT, 0); break;
    case 7: put(UPH_GS_LEFT0, UPH_GS_BOTTOM, UPH_GS_TOP, UPH_GS_CORNER_L, UPH_GS_RIGHT_FB); break;
    default: break;
  }
  return m;
}

// Section `sec`'s test of a fluid row at x and, when it passes, its mirror
// position g and mirrored velocity v and localVelocity lv, in FP64 with the
// reference's expressions (no contraction, so the positions are bit-exact).
__host__ __device__ inline bool mirror_row(int sec, double lbp, double rbp, double fbl, double H, double dx,
                                           double h, double2 x, double2& g, double2& v, double2& lv) {
#pragma clang fp contract(off)
  const double h2 = 2.0 * h;
  double sx = 1.0, sy = 1.0;
  bool on = false;
  switch (sec) {
    case UPH_GS_LEFT0: {
      const double d = fabs(x.x - 0.0 * fbl);
      on = d < h2; g = make_double2(0.0 - d, x.y); sx = -1.0;
      break;
    }
    case UPH_GS_LEFT: {
      const double d = fabs(x.x - lbp);
      on = d < h2; g = make_double2(lbp - d, x.y); sx = -1.0;
      break;
    }
    case UPH_GS_BOTTOM:
      on = (x.y < h2) && (x.x <= fbl); g = make_double2(x.x, -x.y); sy = -1.0;
      break;
    case UPH_GS_CORNER_L: {
      const double d = fabs(x.x - lbp);
      on = (d < h2) && (x.y < h2); g = make_double2(lbp - d, -x.y); sx = sy = -1.0;
      break;
    }
    case UPH_GS_CORNER_R: {
      const double d = fabs(x.x - rbp);
      on = (d < h2) && (x.y < h2); g = make_double2(rbp + d, -x.y); sx = sy = -1.0;
      break;
    }
    case UPH_GS_RIGHT_FB: {
      const double d = fabs(x.x - fbl);
      on = d < h2; g = make_double2(fbl + d, x.y); sx = -1.0;
      break;
    }
    case UPH_GS_RIGHT: {
      const double d = fabs(x.x - rbp);
      on = d < h2; g = make_double2(rbp + d, x.y); sx = -1.0;
      break;
    }
    case UPH_GS_TOP_DX: {
      const double d = fabs(x.y - H - dx);
      on = d < h2; g = make_double2(x.x, H + 2 * d); sy = -1.0;
      break;
    }
    case UPH_GS_TOP: {
      const double d = fabs(x.y - H);
      on = d < h2; g = make_double2(x.x, H + 2 * d); sy = -1.0;
      break;
    }
    case UPH_GS_SLOPE: {
      const double slope = 1.0 / 15.0;   // System.c:309
      const double denoms = 1 + slope * slope;
      const double nums = 1 - slope * slope;
      const double yIntercept = slope * fbl;
      const double d = (1.0 / sqrt(denoms)) * fabs(slope * x.x - x.y - yIntercept);
      on = (d < 3.0 * h) && (x.x > fbl);
      g = make_double2((nums * x.x + 2.0 * slope * x.y + 2.0 * slope * yIntercept) / denoms,
                       (2.0 * slope * x.x - nums * x.y - 2.0 * yIntercept) / denoms);
      if (on) {
        v = make_double2((nums * v.x + 2.0 * slope * v.y) / denoms, (2.0 * slope * v.x - nums * v.y) / denoms);
        lv = make_double2((nums * lv.x + 2.0 * slope * lv.y) / denoms, (2.0 * slope * lv.x - nums * lv.y) / denoms);
      }
      return on;
    }
    default:
      return false;
  }
  if (on) {
    v = make_double2(sx * v.x, sy * v.y);
    lv = make_double2(sx * lv.x, sy * lv.y);
  }
  return on;
}

// LeftBoundaryPosition and RightBoundaryPosition of a derivatives(t) call, as
// inputbp leaves them (inputbp.c:192-283): the piston's position, the sloshing
// tank's L(t) (0 past a moving row of another itype) and FlatBottomLength + L(t),
// 0 in the no-motion branch, otherwise the values the scene was given.
__host__ __device__ inline void boundary_lines(const DevScene& S, double t, double& lbp, double& rbp) {
  lbp = S.lbp0;
  rbp = S.rbp0;
  if (S.motion == 1 || S.motion == 5) {
    const double L = 0.5 * S.m_amp * (1.0 - cos(S.m_omega * t));
    if (S.motion == 1) {
      lbp = L;
    } else {
      lbp = S.slosh_halt ? 0.0 : L;
      rbp = S.flat_bottom + L;
    }
  } else if (S.motion == 0 || S.motion == 2) {
    lbp = 0.0;
  }
}

__device__ __forceinline__ uint32_t tag_id(uint32_t t) { return t & UPH_ID_MASK; }
__device__ __forceinline__ uint32_t tag_type(uint32_t t) { return t >> UPH_ID_BITS; }

// Cell coordinates, clamped into the grid (NaN-safe).
__device__ __forceinline__ void cell_of(const DevScene& S, double x, double y, int& cx, int& cy) {
  double fx = floor((x - S.xmin) * S.cell_inv);
  double fy = floor((y - S.ymin) * S.cell_inv);
  fx = fmin(fmax(fx, 0.0), (double)(S.ncx - 1));
  fy = fmin(fmax(fy, 0.0), (double)(S.ncy - 1));
  cx = (int)fx;
  cy = (int)fy;
}

// FP64 squared distance, dx*dx + dy*dy evaluated without contraction.
__device__ __forceinline__ double dist2(double2 a, double2 b) {
  const double ex = a.x - b.x;
  const double ey = a.y - b.y;
  return __dadd_rn(__dmul_rn(ex, ex), __dmul_rn(ey, ey));
}

// Lexicographic (d2, id) <= (kd2, kid).
__device__ __forceinline__ bool key_le(double d2, uint32_t id, double kd2, uint32_t kid) {
  return d2 < kd2 || (d2 == kd2 && id <= kid);
}

// Visit the candidate slots of a particle at `p`: the 3x3 cells around its
// cell, which in row-major cell order are three contiguous slot ranges.
template <class F>
__device__ __forceinline__ void for_candidates(const DevScene& S, double2 p, const int* __restrict__ cellStart, F&& f) {
  int cx, cy;
  cell_of(S, p.x, p.y, cx, cy);
  const int x0 = cx > 0 ? cx - 1 : 0;
  const int x1 = cx < S.ncx - 1 ? cx + 1 : S.ncx - 1;
#pragma unroll 1
  for (int row = cy - 1; row <= cy + 1; ++row) {
    if (row < 0 || row >= S.ncy) continue;
    const int base = S.coff + row * S.ncx;
    const int b = cellStart[base + x0];
    const int e = cellStart[base + x1 + 1];
#pragma unroll 1
    for (int q = b; q < e; ++q) f(q);
  }
}

// Sum of FP32 pair terms accumulated in FP64: each term enters exactly, so the only
// rounding is FP64's, and an add costs a conversion and one FP64 add (FP64 issues at
// the rate of unpacked FP32 here) where compensated FP32 summation costs four FP32
// operations, in the same two registers.
struct CSum {
  double s;
  __device__ __forceinline__ CSum() : s(0.0) {}
  __device__ __forceinline__ void add(float x) { s += (double)x; }
  __device__ __forceinline__ float val() const { return (float)s; }
};

// Load of base[idx] through a 32-bit byte offset from a uniform base (lets the
// compiler use the scalar-base + vector-offset addressing form).
template <class T>
__device__ __forceinline__ T ld_off(const T* __restrict__ base, uint32_t idx) {
  return *(const T*)((const char*)base + (size_t)(idx * (uint32_t)sizeof(T)));
}

// Wavefront maximum of an int.
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Tait equation of state in FP64 (derivatives.c:1674-1682); integer gamma by
// repeated squaring, otherwise pow().
__device__ __forceinline__ double tait(const DevScene& S, double rho) {
  const double x = rho / S.rho0;
  double xp;
  if (S.gamma_int > 0) {
    double r = 1.0, b = x;
    int e = S.gamma_int;
    while (e) {
      if (e & 1) r *= b;
      b *= b;
      e >>= 1;
    }
    xp = r;
  } else {
    xp = pow(x, S.gamma);
  }
  return S.B * (xp - 1.0);
}

// Kinematics of a moving-set boundary particle at time t (inputbp.c:192-283):
// kind 0 is static, 1 follows the scene's motion, 2 / 3 / 4 are the sloshing
// tank's left wall, right wall and bottom (deck itype 2, 3, 1), and 5 / 6 / 7 the
// same rows after a moving row of another itype, past which the reference's loop
// has zeroed LeftBoundaryPosition and LeftBoundaryVelocity (inputbp.c:270-275;
// RightBoundaryPosition was taken before the loop and keeps moving).  ref is the
// particle's gate height (tempPositionGate) or its sloshing-bottom offset.  Sets
// x and v from the deck position xd, and *lvx where the reference assigns the
// localVelocity x component (returns true then).
__host__ __device__ inline bool wall_kin(const DevScene& S, int kind, double ref, double t, double2 xd, double2& x,
                                         double2& v, double* lvx) {
  if (kind == 0) return false;
  switch (S.motion) {
    case 1:   // piston, absolute position (inputbp.c:194-207)
      x.x = 0.5 * S.m_amp * (1.0 - cos(S.m_omega * t));
      v.x = 0.5 * S.m_amp * S.m_omega * sin(S.m_omega * t);
      *lvx = v.x;
      return true;
    case 2:   // sinusoidal offset of the deck position
      x.x = xd.x + S.m_amp * sin(S.m_omega * t);
      v.x = S.m_amp * S.m_omega * cos(S.m_omega * t);
      *lvx = v.x;
      return true;
    case 3:   // lid (inputbp.c:209-223)
      x.x = S.lid_u * t + xd.x;
      v.x = S.lid_u;
      *lvx = S.lid_u;
      return true;
    case 4: {   // gate (inputbp.c:225-243)
      const bool on = t > 0.0 && t < S.t_stop;
      x.y = S.lid_u * (on ? t : S.t_stop) + ref;
      v.y = on ? S.lid_u : 0.0;
      return false;
    }
    case 5: {   // sloshing tank (inputbp.c:244-276)
      const bool halted = kind >= 5;
      const int k = halted ? kind - 3 : kind;
      const double Lt = 0.5 * S.m_amp * (1.0 - cos(S.m_omega * t));
      const double L = halted ? 0.0 : Lt;
      x.x = k == 2 ? L : (k == 3 ? S.flat_bottom + Lt : ref + L);
      v.x = halted ? 0.0 : 0.5 * S.m_amp * S.m_omega * sin(S.m_omega * t);
      return false;
    }
    default:
      return false;
  }
}

// Motion kind and reference value of boundary row i of a scene (n_fluid fluid
// rows, then n_fixed fixed, then n_moving moving rows), from its deck itype.
// Called for the rows in order: *halted (0 before the first row) records that a
// sloshing-tank moving row of an itype other than 1, 2, 3 has been passed.
__host__ __device__ inline int wall_class(int motion, int i, int nf, int nfix, int ity, double dx, double gate_y,
                                          double* ref, int* halted) {
  *ref = 0.0;
  if (motion == 0 || i < nf + nfix) return 0;
  if (motion == 4) *ref = gate_y;
  if (motion != 5) return 1;
  const int h = *halted ? 3 : 0;
  if (ity == 2) return 2 + h;
  if (ity == 3) return 3 + h;
  if (ity == 1) {
    *ref = (double)(i - nf - nfix) * dx * 0.5;
    return 4 + h;
  }
  *halted = 1;
  return 0;
}
