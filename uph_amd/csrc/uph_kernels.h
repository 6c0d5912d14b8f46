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
};

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

// Kahan-compensated FP32 sum (c carries the rounding error of the last add).
struct CSum {
  float s, c;
  __device__ __forceinline__ CSum() : s(0.f), c(0.f) {}
  __device__ __forceinline__ void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ float val() const { return s - c; }
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

// Wall kinematics at time t for a moving (itype 3) particle (inputbp.c:194-207).
__device__ __forceinline__ void wall_motion(const DevScene& S, double t, double xdeck, double& x, double& vx) {
  if (S.motion == 1) {
    x = 0.5 * S.m_amp * (1.0 - cos(S.m_omega * t));
    vx = 0.5 * S.m_amp * S.m_omega * sin(S.m_omega * t);
  } else if (S.motion == 2) {
    x = xdeck + S.m_amp * sin(S.m_omega * t);
    vx = S.m_amp * S.m_omega * cos(S.m_omega * t);
  }
}
