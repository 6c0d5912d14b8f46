/*
 * uph_run.c -- C host driver: the reference's `sph <nsteps>` (src/bin/sph.cc)
 * with the time loop on the GPU through libuph_gpu.so.
 *
 * Run it in a directory holding the reference's input decks:
 *   parameters.dat (34 values, System.c:60-94), FluidParticle.dat,
 *   FixedBoundaryParticle.dat, MovingBoundaryParticle.dat, SolidParticle.dat
 * (12-column rows, input.c:57-72 / inputbp.c:88-150).
 *
 *   uph_run <nsteps> [--fluid FILE] [--h-factor F] [--alpha A] [--kernel cubic|poly7]
 *           [--scheme config|ref] [--graph] [--compat] [--out FILE] [--no-output] [--probe2 X Y]
 *
 * --compat runs the time loop on the host in the shape of sph.cc:88-168 over a
 * System object (include/uph_system.h) and calls the library's derivatives()
 * once per step, as the reference's own main does; without it the whole loop
 * runs on the device through uph_step.
 * Constants follow System.c:292-313: c = sqrt(g H)/Mach (sqrt(2 g H)/Mach when
 * movingBoundary == 2), B = rho0 c^2 / gamma, eta = sqrt(Nn / 4 pi), hsml = eta dx
 * (input.c:107) unless --h-factor sets h = F dx.  The scheme is the library's:
 * cubic spline (--kernel poly7: the PolyOrderSeven pair of CalculateKernel,
 * derivatives.c:767-823, device loop only), Tait pressure in the momentum equation, artificial viscosity
 * `alpha` (deck field 15 unless --alpha), walls reset from the decks, the
 * piston of inputbp.c:194-207 when movingBoundary == 1.  --scheme ref runs the
 * reference's own run-phase scheme instead (DynamicPressure integrated from the
 * deck column, viscous forces with viscosity0, fluctuations weighted by epsilon,
 * diffusion thermalDiffusivity0; with --kernel poly7 this is the reference's
 * configuration as its decks set it), device loop only.  After the run the
 * fluid state is written in the deck format (default FluidParticleFinal.dat).
 * --probe2 places the second pressure probe (PointPressureTwo.dat and the fifth
 * column of PointPressure.dat; default: the first probe's location).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/uph.h"
#include "../../include/uph_system.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
  int n;
  double *pos, *vel, *rho, *pres, *mass, *lvel, *gate;   /* gate: column 9, tempPositionGate */
  double *tpos;                                         /* column 8, tempPosition */
  int *itype;
} deck_t;

static int read_deck(const char *path, deck_t *d) {
  memset(d, 0, sizeof *d);
  FILE *f = fopen(path, "r");
  if (!f) {
    fprintf(stderr, "uph_run: cannot open %s\n", path);
    return -1;
  }
  if (fscanf(f, "%d", &d->n) != 1 || d->n < 0) {
    fprintf(stderr, "uph_run: bad count in %s\n", path);
    fclose(f);
    return -1;
  }
  size_t n = d->n > 0 ? (size_t)d->n : 1;
  d->pos = calloc(2 * n, sizeof(double));
  d->vel = calloc(2 * n, sizeof(double));
  d->lvel = calloc(2 * n, sizeof(double));
  d->rho = calloc(n, sizeof(double));
  d->pres = calloc(n, sizeof(double));
  d->mass = calloc(n, sizeof(double));
  d->gate = calloc(n, sizeof(double));
  d->tpos = calloc(n, sizeof(double));
  d->itype = calloc(n, sizeof(int));
  for (int i = 0; i < d->n; i++) {
    int got = fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %d %lf %lf", &d->pos[2 * i],
                     &d->pos[2 * i + 1], &d->vel[2 * i], &d->vel[2 * i + 1], &d->rho[i], &d->pres[i],
                     &d->mass[i], &d->tpos[i], &d->gate[i], &d->itype[i], &d->lvel[2 * i], &d->lvel[2 * i + 1]);
    if (got != 12) {
      fprintf(stderr, "uph_run: %s row %d: expected 12 columns\n", path, i + 1);
      fclose(f);
      return -1;
    }
  }
  fclose(f);
  return 0;
}

static void free_deck(deck_t *d) {
  free(d->pos); free(d->vel); free(d->lvel); free(d->rho); free(d->pres); free(d->mass); free(d->gate);
  free(d->tpos);
  free(d->itype);
}

/* One printStep output (see write_output). */
typedef struct snap {
  int n, nf;
  const double *pos, *vel, *rho, *mass, *pres, *rholoc;
  const int *itype;
  double h, t;
  double ke, pe, probe_p, probe_norm;
  int probe_n;
  const int *gate_flag;             /* damGateFlag of every row (input.c:148-159) */
  const double *vort;               /* approxVorticity of the fluid rows (nf), or NULL */
  double probe2_p;                  /* second probe (derivatives.c:2340-2393) */
  int probe2_n;
  const double *tpos, *tgate;       /* deck columns 8 and 9 of the fluid rows (FluidParticleInit.dat) */
} snap_t;
static int write_output(const snap_t *o, int itime);
static int write_init(const snap_t *o, const double *lvel);

/* n rows of 2 doubles over one zeroed block (the reference's CreateMatrix layout) */
static double **rows2(int n) {
  size_t m = n > 0 ? (size_t)n : 1;
  double **r = malloc(m * sizeof(double *));
  double *b = calloc(2 * m, sizeof(double));
  for (size_t i = 0; i < m; i++) r[i] = b + 2 * i;
  return r;
}

static void free_rows2(double **r) {
  if (r) free(r[0]);
  free(r);
}

static double *vec(int n) { return calloc(n > 0 ? (size_t)n : 1, sizeof(double)); }

/* The host time loop of sph.cc:88-168 with the fluctuation fields zero (the
 * config-mode scheme) and derivatives() from the library.  On return s holds
 * the state after nsteps. */
static const int *gate_flag_of;   /* damGateFlag of the run's rows */

static int run_compat(uph_system *s, int nsteps, int print_step, int init_steps) {
  const int nf = s->NumberOfFluidParticles, np = s->MaxNumberOfParticles;
  const double dt = s->timeStep;
  for (int itime = 1; itime <= nsteps; itime++) {
    const double t = itime * dt;
    if (itime != 1) {
      /* predictor (sph.cc:91-126) */
      memcpy(s->density_min, s->density, sizeof(double) * nf);
      memcpy(s->DynamicPressure_min, s->DynamicPressure, sizeof(double) * nf);
      memcpy(s->localDensity, s->density, sizeof(double) * nf);
      memcpy(s->localPressure, s->DynamicPressure, sizeof(double) * nf);
      memcpy(s->localVelocity[0], s->Velocity[0], sizeof(double) * 2 * np);
      memcpy(s->Velocity_min[0], s->Velocity[0], sizeof(double) * 2 * np);
      for (int i = 0; i < nf; i++) {
        s->density[i] += 0.5 * dt * s->ddensitydt_approx[i];
        s->localDensity[i] = s->localDensity[i] + 0.5 * dt * s->ddensitydt_approx[i] + s->densityFluctuation[i];
        s->DynamicPressure[i] += 0.5 * dt * s->dpdt_approx[i];
        s->localPressure[i] = s->localPressure[i] + 0.5 * dt * s->dpdt_approx[i] + s->pressureFluctuation[i];
      }
      for (int i = 0; i < nf; i++)
        for (int d = 0; d < 2; d++) {
          s->Velocity_min[i][d] = s->Velocity[i][d];
          s->Velocity_xsph[i][d] += 0.5 * dt * s->dvdt[i][d];
          s->Velocity[i][d] += 0.5 * dt * s->dvdt_approx[i][d];
          s->localVelocity[i][d] = s->Velocity[i][d] + 0.5 * dt * s->dvdt_approx[i][d] +
                                   s->epsilon * s->velocityFluctuation[i][d];
        }
    } else {
      memcpy(s->Velocity_xsph[0], s->Velocity[0], sizeof(double) * 2 * np);
    }
    if (uph_derivatives(s, t) != 0) {
      fprintf(stderr, "uph_run: %s\n", uph_derivatives_last_error());
      return -1;
    }
    /* corrector (sph.cc:137-168) */
    if (itime == 1) {
      for (int i = 0; i < nf; i++) {
        s->localDensity[i] = s->density[i] + 0.5 * dt * s->ddensitydt[i] + s->densityFluctuation[i];
        s->density[i] += 0.5 * dt * s->ddensitydt[i];
        s->localPressure[i] = s->DynamicPressure[i] + 0.5 * dt * s->dpdt[i] + s->pressureFluctuation[i];
        s->DynamicPressure[i] += 0.5 * dt * s->dpdt[i];
      }
      for (int i = 0; i < nf; i++)
        for (int d = 0; d < 2; d++) {
          s->Velocity_xsph[i][d] = s->Velocity[i][d] + 0.5 * dt * s->dvdt[i][d] +
                                   s->epsilon * s->velocityFluctuation[i][d];
          s->localVelocity[i][d] = s->Velocity[i][d] + 0.5 * dt * s->dvdt[i][d] +
                                   s->epsilon * s->velocityFluctuation[i][d];
          s->Velocity[i][d] += 0.5 * dt * s->dvdt[i][d];
          s->Position[i][d] += dt * s->Velocity[i][d];
        }
    } else {
      for (int i = 0; i < nf; i++) {
        s->localDensity[i] = s->density_min[i] + dt * s->ddensitydt[i] + s->densityFluctuation[i];
        s->density[i] = s->density_min[i] + dt * s->ddensitydt[i];
        s->localPressure[i] = s->DynamicPressure_min[i] + dt * s->dpdt[i] + s->pressureFluctuation[i];
        s->DynamicPressure[i] = s->DynamicPressure_min[i] + dt * s->dpdt[i];
      }
      for (int i = 0; i < nf; i++)
        for (int d = 0; d < 2; d++) {
          s->Velocity_xsph[i][d] = s->Velocity_min[i][d] + dt * s->dvdt[i][d] +
                                   s->epsilon * s->velocityFluctuation[i][d];
          s->localVelocity[i][d] = s->Velocity_min[i][d] + dt * s->dvdt[i][d] +
                                   s->epsilon * s->velocityFluctuation[i][d];
          s->Velocity[i][d] = s->Velocity_min[i][d] + dt * s->dvdt[i][d];
          s->Position[i][d] += dt * s->Velocity[i][d];
        }
    }
    if (print_step > 0 && itime % print_step == 0) {
      /* sph.cc:168-180: the rows are contiguous (rows2), as CreateMatrix makes them;
       * fluid, boundary and mirror rows (output.c:28) */
      const int rows = nf + s->NumberOfBoundaryParticles + s->NumberOfVirtualParticles;
      snap_t o = {rows, nf, s->Position[0], s->Velocity[0], s->density, s->mass, s->DynamicPressure,
                  s->localDensity, s->itype, s->hsml[0], itime * dt, s->totalKineticEnergy,
                  s->potentialEnergy, s->pointPressure, s->pressureNorm, s->pointNeighbors,
                  gate_flag_of, s->approxVorticity, s->pointPressureTwo, s->pointNeighborsTwo,
                  s->tempPosition, s->tempPositionGate};
      if (write_output(&o, itime) != 0) return -1;
      if (itime == init_steps && write_init(&o, s->localVelocity[0]) != 0) return -1;
    }
  }
  return 0;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Output of sph.cc's printStep (output.c:19-130): the snapshot out<itime>.dat of
 * every particle, and one line appended to TotalKineticEnergy.dat, PointPressure.dat
 * and PointPressureTwo.dat.  specificDissipationRate (column 13) is written as 0:
 * the reference accumulates it over every step (derivatives.c:2573-2628 adds into it
 * and nothing resets it after input.c:128), from viscous pair sums that feed nothing
 * else, so it would cost a pair pass per step; for the same reason the dissipation
 * and control-volume series (PowerCalc.dat, ControlVolPower.dat,
 * ControlVolKinetic.dat, output.c:93-112) are not written. */

static int write_output(const snap_t *o, int itime) {
  char name[64];
  snprintf(name, sizeof name, "out%d.dat", itime);
  FILE *f = fopen(name, "w");
  FILE *fke = fopen("TotalKineticEnergy.dat", "a"), *fpp = fopen("PointPressure.dat", "a");
  FILE *fp2 = fopen("PointPressureTwo.dat", "a");
  if (!f || !fke || !fpp || !fp2) {
    fprintf(stderr, "uph_run: cannot write the output of step %d\n", itime);
    if (f) fclose(f);
    if (fke) fclose(fke);
    if (fpp) fclose(fpp);
    if (fp2) fclose(fp2);
    return -1;
  }
  for (int i = 0; i < o->n; i++)
    fprintf(f, "%d \t%.10lf\t%.10lf\t%.10lf\t%.10lf\t%.10lf\t%.10lf\t%.10lf\t%.10lf\t%d\t%.10lf\t%d\t%.10lf\t%.10lf\n",
            i, o->pos[2 * i], o->pos[2 * i + 1], o->vel[2 * i], o->vel[2 * i + 1], o->rho[i], o->mass[i], o->h,
            o->pres[i], o->gate_flag ? o->gate_flag[i] : 0, (o->vort && i < o->nf) ? o->vort[i] : 0.0,
            o->itype[i], 0.0, o->rholoc[i]);
  fprintf(fke, "%.10lf\t%.10lf %.10lf\n", o->t, o->ke, o->pe);
  fprintf(fpp, "%.10lf\t%.10lf\t%d\t%.10lf\t%.10lf\n", o->t, o->probe_p, o->probe_n, o->probe_norm, o->probe2_p);
  fprintf(fp2, "%.10lf\t%d\t%.10lf\n", o->t, o->probe2_n, o->probe2_p);
  fclose(f);
  fclose(fke);
  fclose(fpp);
  fclose(fp2);
  return 0;
}

/* FluidParticleInit.dat at step NumberOfInitializationSteps (output.c:116-128): the
 * fluid rows in the deck format, a restart point. */
static int write_init(const snap_t *o, const double *lvel) {
  FILE *f = fopen("FluidParticleInit.dat", "a");
  if (!f) return -1;
  fprintf(f, "%d\n", o->nf);
  for (int i = 0; i < o->nf; i++)
    fprintf(f, "%.10lf\t%.10lf\t%.10lf\t%.10lf\t%.10lf\t%.10lf\t%.10lf\t%.10lf\t%.10lf\t%d\t%.10lf\t%.10lf\n",
            o->pos[2 * i], o->pos[2 * i + 1], o->vel[2 * i], o->vel[2 * i + 1], o->rho[i], o->pres[i], o->mass[i],
            o->tpos ? o->tpos[i] : 0.0, o->tgate ? o->tgate[i] : 0.0, o->itype[i], lvel[2 * i], lvel[2 * i + 1]);
  fclose(f);
  return 0;
}

/* sph.cc:54-80 empties the series files at start */
static void truncate_series(void) {
  const char *names[] = {"TotalKineticEnergy.dat", "PointPressure.dat", "FluidParticleInit.dat",
                         "PointPressureTwo.dat"};
  for (int k = 0; k < 4; k++) {
    FILE *f = fopen(names[k], "w");
    if (f) fclose(f);
  }
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "Give a maximum number of time step\n");
    return 1;
  }
  int nsteps = atoi(argv[1]);
  const char *fluid_file = "FluidParticle.dat", *out_file = "FluidParticleFinal.dat";
  double h_factor = 0.0, alpha = -1.0;
  int use_graph = 0, compat = 0, write = 1, kernel = UPH_KERNEL_CUBIC, mode = UPH_MODE_CONFIG;
  int have_probe2 = 0;
  double probe2[2] = {0.0, 0.0};
  for (int a = 2; a < argc; a++) {
    if (!strcmp(argv[a], "--fluid") && a + 1 < argc) fluid_file = argv[++a];
    else if (!strcmp(argv[a], "--h-factor") && a + 1 < argc) h_factor = atof(argv[++a]);
    else if (!strcmp(argv[a], "--alpha") && a + 1 < argc) alpha = atof(argv[++a]);
    else if (!strcmp(argv[a], "--out") && a + 1 < argc) out_file = argv[++a];
    else if (!strcmp(argv[a], "--kernel") && a + 1 < argc) {
      const char *k = argv[++a];
      if (!strcmp(k, "cubic")) kernel = UPH_KERNEL_CUBIC;
      else if (!strcmp(k, "poly7")) kernel = UPH_KERNEL_POLY7;
      else {
        fprintf(stderr, "uph_run: unknown kernel %s (cubic, poly7)\n", k);
        return 1;
      }
    }
    else if (!strcmp(argv[a], "--scheme") && a + 1 < argc) {
      const char *k = argv[++a];
      if (!strcmp(k, "config")) mode = UPH_MODE_CONFIG;
      else if (!strcmp(k, "ref")) mode = UPH_MODE_REF;
      else {
        fprintf(stderr, "uph_run: unknown scheme %s (config, ref)\n", k);
        return 1;
      }
    }
    else if (!strcmp(argv[a], "--probe2") && a + 2 < argc) {
      probe2[0] = atof(argv[++a]);
      probe2[1] = atof(argv[++a]);
      have_probe2 = 1;
    }
    else if (!strcmp(argv[a], "--graph")) use_graph = 1;
    else if (!strcmp(argv[a], "--compat")) compat = 1;
    else if (!strcmp(argv[a], "--no-output")) write = 0;
    else {
      fprintf(stderr, "uph_run: unknown option %s\n", argv[a]);
      return 1;
    }
  }
  double v[34];
  FILE *fp = fopen("parameters.dat", "r");
  if (!fp) {
    fprintf(stderr, "cannot find parameter file parameters.dat\n");
    return 1;
  }
  for (int k = 0; k < 34; k++)
    if (fscanf(fp, "%lf", &v[k]) != 1) {
      fprintf(stderr, "uph_run: parameters.dat has fewer than 34 values\n");
      fclose(fp);
      return 1;
    }
  fclose(fp);
  if ((int)v[25] != 1) {
    /* flagInitSimul 0 selects the initialization phase (frozen particles, inputbp.c:75,
       derivatives.c:4197) */
    fprintf(stderr, "uph_run: flagInitSimul %g: only the run phase (flagInitSimul 1) is supported\n", v[25]);
    return 1;
  }
  /* field map of System.c:60-94 (0-based) */
  const double stroke = v[0], period = v[1], depth = v[2], rho0 = v[4], gamma = v[5], g = v[6];
  const double dx = v[7], mach = v[11], alpha_deck = v[14], dt = v[18], nn = v[20];
  const double k0 = v[23], c0d = v[24];
  const int gravity = (int)v[26], moving = (int)v[28];
  const int print_step = write ? (int)v[16] : 0, init_steps = (int)v[29];

  deck_t fl, fx, mv, so;
  if (read_deck(fluid_file, &fl) || read_deck("FixedBoundaryParticle.dat", &fx) ||
      read_deck("MovingBoundaryParticle.dat", &mv) || read_deck("SolidParticle.dat", &so))
    return 1;
  if (so.n != 0) {
    fprintf(stderr, "uph_run: solid particles are not supported (SolidParticle.dat has %d)\n", so.n);
    return 1;
  }
  const int n = fl.n + fx.n + mv.n;
  deck_t all;
  memset(&all, 0, sizeof all);
  all.pos = malloc(sizeof(double) * 2 * n); all.vel = malloc(sizeof(double) * 2 * n);
  all.lvel = malloc(sizeof(double) * 2 * n); all.rho = malloc(sizeof(double) * n);
  all.mass = malloc(sizeof(double) * n); all.itype = malloc(sizeof(int) * n);
  all.pres = malloc(sizeof(double) * n); all.gate = malloc(sizeof(double) * n);
  all.tpos = malloc(sizeof(double) * n);
  int *gflag = calloc(n > 0 ? (size_t)n : 1, sizeof(int));
  deck_t *parts[3] = {&fl, &fx, &mv};
  int o = 0;
  double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
  for (int p = 0; p < 3; p++)
    for (int i = 0; i < parts[p]->n; i++, o++) {
      for (int d = 0; d < 2; d++) {
        all.pos[2 * o + d] = parts[p]->pos[2 * i + d];
        all.vel[2 * o + d] = parts[p]->vel[2 * i + d];
        all.lvel[2 * o + d] = parts[p]->lvel[2 * i + d];
        if (all.pos[2 * o + d] < lo[d]) lo[d] = all.pos[2 * o + d];
        if (all.pos[2 * o + d] > hi[d]) hi[d] = all.pos[2 * o + d];
      }
      all.rho[o] = parts[p]->rho[i];
      all.pres[o] = parts[p]->pres[i];
      all.gate[o] = parts[p]->gate[i];
      all.tpos[o] = parts[p]->tpos[i];
      /* input.c:148-159: fluid rows of the dam column (x <= 0.38 at input) 1, the rest -1;
       * boundary rows 0 (inputbp.c:311) */
      if (p == 0) gflag[o] = parts[p]->pos[2 * i] <= 0.38 ? 1 : -1;
      all.mass[o] = parts[p]->mass[i];
      all.itype[o] = parts[p]->itype[i];
    }
  const double c = (moving == 2 ? sqrt(2.0 * g * depth) : sqrt(g * depth)) / mach;
  const double eta = sqrt(nn / (4.0 * M_PI));
  uph_scene s;
  memset(&s, 0, sizeof s);
  s.n_fluid = fl.n; s.n_fixed = fx.n; s.n_moving = mv.n;
  s.pos = all.pos; s.vel = all.vel; s.rho = all.rho; s.mass = all.mass; s.itype = all.itype;
  s.lvel = all.lvel;
  s.dx = dx;
  s.h = h_factor > 0 ? h_factor * dx : eta * dx;
  s.rho0 = rho0; s.gamma = gamma; s.c0 = c; s.B = (1. / gamma) * rho0 * c * c; s.g = g;
  s.gravity = gravity; s.alpha_av = alpha >= 0 ? alpha : alpha_deck;
  s.k_spring = k0; s.c_damp = c0d; s.dt = dt; s.K = 40;
  /* inputbp.c:192-283: movingBoundary 1, 3, 2, else virtualType 5 */
  s.motion = moving == 1 ? UPH_MOTION_PISTON : moving == 3 ? UPH_MOTION_LID : moving == 2 ? UPH_MOTION_GATE
           : (int)v[15] == 5 ? UPH_MOTION_SLOSH : UPH_MOTION_NONE;
  s.lid_u = v[19]; s.depth = v[2]; s.flat_bottom = v[3]; s.gate_y = all.gate;
  s.vtype = (int)v[15];   /* virtualType: mirror particles (derivatives.c:3193-4192) */
  s.m_amp = stroke; s.m_omega = 2.0 * M_PI / period;
  s.kernel = kernel;
  s.mode = mode;
  s.pres = all.pres;
  s.kappa = v[27]; s.mu = v[10]; s.eps = v[13];   /* thermalDiffusivity0, viscosity0, epsilon */
  s.stress_type = (int)v[32];                       /* stressType */
  /* the second probe: pressureProbeXX/YY, which the reference never assigns
   * (System.c:130-131 commented out): --probe2, else the first probe's location */
  if (!have_probe2) {
    probe2[0] = v[30];
    probe2[1] = v[31];
  }
  gate_flag_of = gflag;
  const double margin = 10 * s.h + (moving == 1 ? stroke : 0.0);
  s.box[0] = lo[0] - margin; s.box[1] = lo[1] - margin;
  s.box[2] = hi[0] + margin; s.box[3] = hi[1] + 2 * (hi[1] - lo[1]);
  if (compat && uph_derivatives_scheme(mode, kernel) != 0) {
    fprintf(stderr, "uph_run: %s\n", uph_derivatives_last_error());
    return 1;
  }
  if (compat) {
    /* a System as the reference's input() leaves it (System.c, input.c, inputbp.c) */
    uph_system sy;
    memset(&sy, 0, sizeof sy);
    sy.dim = 2;
    sy.NumberOfFluidParticles = fl.n;
    sy.NumberOfFixedBoundaryParticles = fx.n;
    sy.NumberOfMovingBoundaryParticles = mv.n;
    sy.NumberOfBoundaryParticles = fx.n + mv.n;
    /* rows for the mirror particles derivatives() appends (at most 5 per fluid row;
     * the reference reserves 100, System.c:195) */
    const int nm = n + 5 * fl.n;
    sy.MaxNumberOfParticles = nm;
    sy.Position = rows2(nm); sy.Velocity = rows2(nm); sy.localVelocity = rows2(nm);
    sy.Velocity_min = rows2(nm); sy.Velocity_xsph = rows2(nm);
    sy.dvdt = rows2(nm); sy.dvdt_approx = rows2(nm); sy.velocityFluctuation = rows2(nm);
    sy.mass = vec(nm); sy.density = vec(nm); sy.density_min = vec(nm); sy.localDensity = vec(nm);
    sy.Pressure = vec(nm); sy.DynamicPressure = vec(nm); sy.DynamicPressure_min = vec(nm);
    sy.localPressure = vec(nm); sy.dpdt = vec(nm); sy.dpdt_approx = vec(nm); sy.hsml = vec(nm);
    sy.ddensitydt = vec(nm); sy.ddensitydt_approx = vec(nm);
    sy.densityFluctuation = vec(nm); sy.pressureFluctuation = vec(nm);
    sy.itype = calloc(nm, sizeof(int));
    o = 0;
    for (int p = 0; p < 3; p++)
      for (int i = 0; i < parts[p]->n; i++, o++) {
        for (int d = 0; d < 2; d++) {
          sy.Position[o][d] = parts[p]->pos[2 * i + d];
          sy.Velocity[o][d] = parts[p]->vel[2 * i + d];
          sy.localVelocity[o][d] = parts[p]->lvel[2 * i + d];
        }
        sy.density[o] = parts[p]->rho[i];
        /* input.c:109-110 sets localDensity = rho0 and localPressure = 0; config mode
         * starts localDensity at the density */
        sy.localDensity[o] = mode == UPH_MODE_REF ? rho0 : parts[p]->rho[i];
        sy.localPressure[o] = 0.0;
        sy.DynamicPressure[o] = parts[p]->pres[i];
        sy.mass[o] = parts[p]->mass[i];
        sy.itype[o] = parts[p]->itype[i];
        sy.hsml[o] = s.h;
      }
    sy.dx = dx;
    sy.density0 = rho0; sy.AdiabaticConstant = gamma; sy.Gravity = g; sy.gravityFlag = gravity;
    sy.SpeedOfSound = c; sy.CompressionFactor = s.B; sy.alpha = s.alpha_av;
    sy.springConstant0 = k0; sy.dampingConstant0 = c0d; sy.timeStep = dt; sy.HowManyNeighbors = 40;
    sy.flagInitSimul = 1;
    sy.movingBoundary = moving; sy.stroke = stroke; sy.wavePeriod = period; sy.omega = 2.0 * M_PI / period;
    sy.virtualType = (int)v[15]; sy.lidVelocity = v[19]; sy.LocalDepth = v[2]; sy.FlatBottomLength = v[3];
    sy.tempPositionGate = all.gate;
    sy.pressureProbeX = v[30]; sy.pressureProbeY = v[31];
    sy.pressureProbeXX = probe2[0]; sy.pressureProbeYY = probe2[1];
    sy.approxVorticity = vec(nm);
    sy.tempPosition = all.tpos;
    sy.viscosity0 = v[10]; sy.epsilon = v[13]; sy.thermalDiffusivity0 = v[27];
    sy.stressType = (int)v[32];
    if (print_step > 0) truncate_series();
    double tc = now_s();
    if (run_compat(&sy, nsteps, print_step, init_steps) != 0) return 1;
    double elc = now_s() - tc;
    printf("%d fluid + %d wall particles, %d steps in %.3f s (host loop, derivatives() on the GPU): "
           "%.4g particle-updates/s\n", fl.n, fx.n + mv.n, nsteps, elc, elc > 0 ? (double)fl.n * nsteps / elc : 0.0);
    /* the diagnostics of the last derivatives() call (the state before the last corrector) */
    printf("diagnostics: kinetic %.12e potential %.12e probe %.12e (%d neighbours)\n", sy.totalKineticEnergy,
           sy.potentialEnergy, sy.pointPressure, sy.pointNeighbors);
    FILE *fc = fopen(out_file, "w");
    if (!fc) {
      fprintf(stderr, "uph_run: cannot write %s\n", out_file);
      return 1;
    }
    fprintf(fc, "%d\n", fl.n);
    for (int i = 0; i < fl.n; i++)
      fprintf(fc, "%.10lf %.10lf %.10lf %.10lf %.10lf %.10lf %.10lf %.10lf %.10lf %d %.10lf %.10lf\n",
              sy.Position[i][0], sy.Position[i][1], sy.Velocity[i][0], sy.Velocity[i][1], sy.density[i],
              sy.DynamicPressure[i], sy.mass[i], 0.0, 0.0, 1, sy.localVelocity[i][0], sy.localVelocity[i][1]);
    fclose(fc);
    free_rows2(sy.Position); free_rows2(sy.Velocity); free_rows2(sy.localVelocity);
    free_rows2(sy.Velocity_min); free_rows2(sy.Velocity_xsph); free_rows2(sy.dvdt);
    free_rows2(sy.dvdt_approx); free_rows2(sy.velocityFluctuation);
    free(sy.mass); free(sy.density); free(sy.density_min); free(sy.localDensity); free(sy.Pressure);
    free(sy.DynamicPressure); free(sy.DynamicPressure_min); free(sy.localPressure); free(sy.dpdt);
    free(sy.dpdt_approx); free(sy.hsml); free(sy.ddensitydt); free(sy.ddensitydt_approx);
    free(sy.densityFluctuation); free(sy.pressureFluctuation); free(sy.itype); free(sy.approxVorticity);
    free(all.pos); free(all.vel); free(all.lvel); free(all.rho); free(all.mass); free(all.itype); free(all.pres); free(all.gate); free(all.tpos); free(gflag);
    free_deck(&fl); free_deck(&fx); free_deck(&mv); free_deck(&so);
    return 0;
  }
  uph_opts opts = {use_graph, 0, 0, 256, 1.0, 0, 0.0};
  uph_t *h = NULL;
  if (uph_create(&s, 1, &opts, &h) != 0) {
    fprintf(stderr, "uph_run: %s\n", uph_last_error(NULL));
    return 1;
  }
  /* room for the mirror rows the snapshots list after the particle rows (output.c:28) */
  const int gcap = uph_ghost_capacity(h, 0) > 0 ? uph_ghost_capacity(h, 0) : 0;
  const int nx = n + gcap;
  double *pos = malloc(sizeof(double) * 2 * nx), *vel = malloc(sizeof(double) * 2 * nx);
  double *rho = malloc(sizeof(double) * nx);
  double t0 = now_s();
  if (print_step <= 0 || print_step > nsteps) {
    if (uph_step(h, nsteps) != 0 || uph_sync(h) < 0) {
      fprintf(stderr, "uph_run: %s\n", uph_last_error(h));
      uph_destroy(h);
      return 1;
    }
  } else {
    /* printStep output: the next interval is queued before the host writes the
     * files of this one, so the device keeps stepping during the I/O */
    truncate_series();
    double *pres = malloc(sizeof(double) * nx), *rhol = malloc(sizeof(double) * nx);
    double *massx = malloc(sizeof(double) * nx);
    double *vort = malloc(sizeof(double) * (fl.n > 0 ? (size_t)fl.n : 1));
    int *itypex = malloc(sizeof(int) * nx), *gflagx = calloc(nx, sizeof(int));
    memcpy(gflagx, gflag, sizeof(int) * n);
    memcpy(massx, all.mass, sizeof(double) * n);
    memcpy(itypex, all.itype, sizeof(int) * n);
    for (int i = n; i < nx; i++) itypex[i] = 4;   /* mirror rows (derivatives.c:3220) */
    int done = 0, first = print_step < nsteps ? print_step : nsteps;
    if (uph_step(h, first) != 0) {
      fprintf(stderr, "uph_run: %s\n", uph_last_error(h));
      return 1;
    }
    while (done < nsteps) {
      const int chunk = print_step < nsteps - done ? print_step : nsteps - done;
      uph_diag dg;
      double p2 = 0.0;
      int n2 = 0;
      if (uph_get_state(h, 0, pos, vel, rho) != 0 || uph_diagnostics(h, 0, v[30], v[31], &dg) != 0 ||
          uph_probe_two(h, 0, probe2[0], probe2[1], &p2, &n2) != 0 || uph_vorticity(h, 0, vort) != 0) {
        fprintf(stderr, "uph_run: %s\n", uph_last_error(h));
        return 1;
      }
      done += chunk;
      /* everything the snapshot lists is read before the next interval is queued */
      int ng = 0;
      if (done % print_step == 0) {
        for (int i = 0; i < n; i++) {
          pres[i] = i < fl.n ? s.B * (pow(rho[i] / rho0, gamma) - 1.0) : (i < fl.n + fx.n ? fx.pres[i - fl.n] : mv.pres[i - fl.n - fx.n]);
          rhol[i] = i < fl.n ? rho[i] : rho0;
        }
        if (mode == UPH_MODE_REF && uph_get_pressure(h, 0, pres) != 0) {   /* DynamicPressure */
          fprintf(stderr, "uph_run: %s\n", uph_last_error(h));
          return 1;
        }
        ng = uph_get_ghosts(h, 0, gcap, pos + 2 * (size_t)n, vel + 2 * (size_t)n, NULL, rho + n, rhol + n,
                            pres + n, NULL, massx + n);
        if (ng < 0) {
          fprintf(stderr, "uph_run: %s\n", uph_last_error(h));
          return 1;
        }
      }
      if (done < nsteps) {
        const int next = print_step < nsteps - done ? print_step : nsteps - done;
        if (uph_step(h, next) != 0) {
          fprintf(stderr, "uph_run: %s\n", uph_last_error(h));
          return 1;
        }
      }
      if (done % print_step == 0) {
        snap_t o = {n + ng, fl.n, pos, vel, rho, massx, pres, rhol, itypex, s.h, done * dt, dg.kinetic_energy,
                    dg.potential_energy, dg.probe_pressure, dg.probe_norm, dg.probe_neighbors, gflagx, vort, p2,
                    n2, all.tpos, all.gate};
        if (write_output(&o, done) != 0) return 1;
        if (done == init_steps && write_init(&o, vel) != 0) return 1;
      }
    }
    free(massx);
    free(itypex);
    free(gflagx);
    free(vort);
    if (uph_sync(h) < 0) {
      fprintf(stderr, "uph_run: %s\n", uph_last_error(h));
      return 1;
    }
    free(pres);
    free(rhol);
  }
  double el = now_s() - t0;
  printf("%d fluid + %d wall particles, %d steps in %.3f s: %.4g particle-updates/s\n", fl.n,
         fx.n + mv.n, nsteps, el, el > 0 ? (double)fl.n * nsteps / el : 0.0);
  uph_diag dg;
  if (uph_diagnostics(h, 0, v[30], v[31], &dg) == 0)
    printf("diagnostics: kinetic %.12e potential %.12e probe %.12e (%d neighbours)\n", dg.kinetic_energy,
           dg.potential_energy, dg.probe_pressure, dg.probe_neighbors);
  if (uph_get_state(h, 0, pos, vel, rho) != 0) {
    fprintf(stderr, "uph_run: %s\n", uph_last_error(h));
    uph_destroy(h);
    return 1;
  }
  FILE *fo = fopen(out_file, "w");
  if (!fo) {
    fprintf(stderr, "uph_run: cannot write %s\n", out_file);
    uph_destroy(h);
    return 1;
  }
  double *pfin = malloc(sizeof(double) * n);
  if (mode == UPH_MODE_REF && uph_get_pressure(h, 0, pfin) != 0) {
    fprintf(stderr, "uph_run: %s\n", uph_last_error(h));
    uph_destroy(h);
    return 1;
  }
  fprintf(fo, "%d\n", fl.n);
  for (int i = 0; i < fl.n; i++) {
    const double P = mode == UPH_MODE_REF ? pfin[i] : s.B * (pow(rho[i] / rho0, gamma) - 1.0);
    fprintf(fo, "%.10lf %.10lf %.10lf %.10lf %.10lf %.10lf %.10lf %.10lf %.10lf %d %.10lf %.10lf\n",
            pos[2 * i], pos[2 * i + 1], vel[2 * i], vel[2 * i + 1], rho[i], P, all.mass[i], 0.0, 0.0, 1,
            vel[2 * i], vel[2 * i + 1]);
  }
  fclose(fo);
  uph_destroy(h);
  free(pos); free(vel); free(rho); free(pfin);
  free(all.pos); free(all.vel); free(all.lvel); free(all.rho); free(all.mass); free(all.itype); free(all.pres); free(all.gate); free(all.tpos); free(gflag);
  free_deck(&fl); free_deck(&fx); free_deck(&mv); free_deck(&so);
  return 0;
}
