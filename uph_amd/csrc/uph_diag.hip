// uph_diag.hip -- diagnostics of a handle's state: energies and probes, the second
// probe, approxVorticity.
#include "uph_internal.h"

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
  StepBufs rb;
  if (read_bufs(h, rb)) return -1;
  hipLaunchKernelGGL(k_diag, dim3(nb), dim3(DIAG_BLK), 0, h->stream, S.poff, S.np, rb.tagA, rb.posB,
                     rb.velB, rb.rhoB, h->b.massD, h->mode == UPH_MODE_REF ? rb.presB : nullptr, h->dsc,
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
int vort_run(uph_t* h, int scene, double* out) {
  const DevScene& S = h->hs[scene];
  if (S.nf > h->vort_cap) {
    if (h->vort) HIPCHK(hipFree(h->vort), h->err);
    h->vort = nullptr;
    HIPCHK(hipMalloc((void**)&h->vort, (size_t)S.nf * sizeof(double)), h->err);
    h->vort_cap = S.nf;
  }
  StepBufs rb;
  if (read_bufs(h, rb)) return -1;
  if (h->kernel == UPH_KERNEL_POLY7)
    hipLaunchKernelGGL(k_vort<UPH_KERNEL_POLY7>, dim3(grid_of(S.np)), dim3(BLK), 0, h->stream, S.poff, S.np, rb,
                       h->dsc, scene, h->vort);
  else
    hipLaunchKernelGGL(k_vort<UPH_KERNEL_CUBIC>, dim3(grid_of(S.np)), dim3(BLK), 0, h->stream, S.poff, S.np, rb,
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
