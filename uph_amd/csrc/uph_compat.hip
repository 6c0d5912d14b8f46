// uph_compat.hip -- the derivatives(System*, double) drop-in and the transfer of a
// scene's state between host arrays and a handle.
#include "uph_internal.h"

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
  StepBufs rb;
  if (read_bufs(h, rb)) return -1;
  hipLaunchKernelGGL(k_state_out, dim3(grid_of(S.np)), dim3(BLK), 0, h->stream, S.poff, S.np, rb.tagA, rb.posB,
                     rb.velB, rb.rhoB, (double2*)pos, (double2*)vel, rho);
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
  // into the B set of the state getters' view; a lazy handle sorts it from scratch next
  StepBufs rb;
  if (read_bufs(h, rb) || lazy_reset(h)) return -1;
  hipLaunchKernelGGL(k_state_in, dim3(grid_of(S.np)), dim3(BLK), 0, h->stream, S.poff, S.np, S.nf, rb.tagA,
                     rb.posB, rb.velB, rb.rhoB, (const double2*)pos, (const double2*)vel, rho,
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
          compat_motion(s) != UPH_MOTION_SLOSH ? s->RightBoundaryPosition : 0.0, (double)s->stressType};
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
  u.stress_type = s->stressType;
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
  o.sort_every_step = 1;   // each call restarts the step from the caller's state
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
