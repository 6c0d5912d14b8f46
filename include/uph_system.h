/* uph_system.h -- binary-compatible view of the reference solver's state object.
 *
 * The reference passes one `System` (src/lib/System.h:13-153, created by
 * CreateSystem, System.c:38-345) to every solver routine, among them
 * `void derivatives(System *sys, double t)` (derivatives.h:12, called from
 * sph.cc:135).  `uph_system` has the same members, types and order, so a
 * `System*` built by the reference's own code can be handed to this library's
 * `derivatives()` unchanged.  Member names follow the reference so that code
 * written against it reads the same; the grouping comments are this library's.
 * tests/test_abi.py compares every member offset against the reference header
 * where that header is available.
 *
 * Arrays are host memory owned by the caller; `double **` members are row
 * pointers (one row of `dim` doubles per particle).  Particles are ordered
 * fluid, solid, fixed boundary, moving boundary, frozen (inputbp.c:117-160). */
#ifndef UPH_SYSTEM_H
#define UPH_SYSTEM_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct uph_system {
  /* geometry and particle counts */
  int dim;
  int *itype;
  int nx;
  int ny;
  double Lx, Ly;
  double dx, dy;
  int NumberOfInteractingParticles;
  int MaxNumberOfParticles;
  int NumberOfFluidParticles;
  int NumberOfBoundaryParticles;
  int NumberOfFixedBoundaryParticles;
  int NumberOfMovingBoundaryParticles;
  int NumberOfFrozenParticles;
  int NumberOfSolidParticles;
  int NumberOfVirtualParticles;
  int nearNeighbor;

  /* wave-maker wall state */
  double LeftBoundaryPosition, LeftBoundaryVelocity, RightBoundaryPosition, LeftBoundaryAcceleration;

  /* per-particle state and rates */
  double **tmp;
  double **Position;
  double **Velocity;
  double **Velocity_xsph;
  double *mass;
  double *Pressure;
  double *Energy;
  double *hsml;
  double *density;
  double **Velocity_min;
  double *ddensitydt;
  double **dvdt;
  double **av;
  double *density_min;
  double density0;
  double *viscosity;
  double *ViscousStressxx, *ViscousStressxy, *ViscousStressyx, *ViscousStressyy;
  double coreCoefficient;
  double machNumber;
  double *DynamicPressure, *DynamicPressure_min, *dpdt;
  double *localDensity_min;
  double *meanCompressibility, *meanSpecDensity;
  double pressureKernelNorm;

  /* fluctuation fields of the reference scheme */
  double *densityFluctuation, *pressureFluctuation;
  double **velocityFluctuation;
  double *normalizationFactor;
  double damping_time, tstep;
  int dampingIterations;

  /* local (smoothed) fields, kernel values, run control */
  double *localDensity, *localPressure, **localVelocity;
  double *wc, **dwcdx;
  double timeStep, pecletNumber, viscosity0;
  int virtualType;
  double *tempPosition, *tempPositionGate, gateVelocity;
  int printStep, screenStep;
  double lennardSelect, lidVelocity, wallViscosity;
  double totalKineticEnergy;
  int pointNeighbors;
  double pointPressure, *tempPointPressure;
  double initialEnergy, potentialEnergy;
  double *wp;
  int selectBoundaryForce;
  double springConstant0, dampingConstant0;
  double *springConstant, *dampingConstant;
  int *pj_pair;
  double *tempPointVolume;
  double **gradNormFactor;
  int flagInitSimul;
  double *localCompressibility;
  int gravityFlag;
  double *thermalDiffusivity, thermalDiffusivity0;
  int movingBoundary;
  int NumberOfInitializationSteps;
  double consistencyDensityOne, consistencyDensityTwo, consistencyDensityThree;
  double densityUncertainty;
  double *specificVolume;
  double **Position_min;
  double pressureNorm;
  double *Volume;
  double *vorticity, *approxVorticity;
  double pressureProbeX, pressureProbeY;
  double pressureProbeXX, pressureProbeYY;
  double *tempPointPressureTwo;
  int pointNeighborsTwo;
  double pointPressureTwo;
  double *xSpongeVelocity, *ySpongeVelocity;
  double *spongeFactor, spongeFactor0, spongePosition0;
  int waveAbsorption;

  /* diagnostics */
  double enstrophyPower;
  int stressType;
  double *specificDissipationPower, dissipationPower;
  double *specificBoundaryDissipationRate, boundaryDissipationRate;
  double *specificDissipationRate, dissipationRate, *specificDissipationRateTemp;
  int disspationControlVolume;
  double *specificTurbulentDissipationRate, *turbulentKineticEnergy;
  double *controlKineticEnergy, *controlDissipationRate, *controlMaxVelocity, *controlVolumeX;
  int controlVolumeNumber;
  int *damGateFlag;

  /* predicting rates */
  double *ddensitydt_approx;
  double *dpdt_approx;
  double **dvdt_approx;

  /* kernel normalisations and physical constants */
  double alpha_cspline, alpha_kernelFour, alpha_deconvolutionFour, alpha_kernelFive, alpha_deconvolutionFive;
  double alpha_wendlandFour, alpha_deconvolutionSix;
  double alpha;
  double SpeedOfSound;
  double slope;
  double stroke, omega, wavePeriod;
  double eta;
  double Gravity;
  double LennardCoefficient;
  double LocalDepth, FlatBottomLength;
  double epsilon;

  /* pair list */
  int *i_pair, *j_pair;
  int *NeighborsSize;
  int *i_inflow;
  double *rij2, *w;
  double **dwdx;
  int inflowParticles;
  double CompressionFactor;
  double AdiabaticConstant;
  int niac;
  int HowManyNeighbors;
} uph_system;

/* The reference's rate evaluation, config-mode scheme, on the GPU
 * (replaces derivatives(), derivatives.h:12 / derivatives.c:3137).
 *
 * Reads the fluid rows' Position, Velocity, localVelocity and density (the
 * scheme keeps localDensity equal to density for fluid), mass, itype, hsml[0]
 * and the constants density0, AdiabaticConstant, SpeedOfSound,
 * CompressionFactor, Gravity, gravityFlag, alpha, springConstant0,
 * dampingConstant0, dx, timeStep, HowManyNeighbors, flagInitSimul (must be 1,
 * the run phase), movingBoundary, virtualType, stroke, omega, lidVelocity,
 * LocalDepth, FlatBottomLength and tempPositionGate.  Writes, for fluid rows: ddensitydt, ddensitydt_approx, dvdt,
 * dvdt_approx; Pressure and DynamicPressure = Tait(density); dpdt and
 * dpdt_approx = (dP/drho) times the two density rates; the fluctuation fields
 * (where allocated) = 0; the diagnostics of derivatives.c:4317-4324
 * (totalKineticEnergy, potentialEnergy, pointNeighbors, pointPressure,
 * pressureNorm, consistencyDensity*, densityUncertainty) of the state it was
 * called with, probe at pressureProbeX/Y; pointPressureTwo and pointNeighborsTwo
 * at pressureProbeXX/YY (derivatives.c:2340-2393); approxVorticity of the fluid rows
 * (derivatives.c:2840-2885) where the System has that array.  Boundary rows are reset to the state they had at the
 * first call (the deck, as inputbp re-reads it every step) with the moving rows'
 * kinematics of inputbp.c:192-283 at time t: movingBoundary 1 (piston), 3 (lid
 * at lidVelocity), 2 (gate rising at lidVelocity from tempPositionGate until
 * 1.5 (LocalDepth - 15 dx) / lidVelocity), else virtualType 5 (sloshing tank:
 * walls of itype 2, 3 and bottom of itype 1, FlatBottomLength apart).  Solid,
 * frozen and virtual particles are not supported (error).
 *
 * The first call builds a device context for the System's sizes and
 * constants; later calls with the same System, sizes and constants reuse it
 * (a change rebuilds it, and the boundary deck is taken again).  One context
 * per process: calls are serialised.
 *
 * uph_derivatives returns 0, or -1 with a message in uph_derivatives_last_error;
 * derivatives() has the reference's void signature: on failure it prints the
 * message to stderr and sets the fluid rates to NaN. */
int uph_derivatives(uph_system *sys, double t);
void derivatives(uph_system *sys, double t);
const char *uph_derivatives_last_error(void);

/* The scheme derivatives() evaluates, for the whole process (no counterpart in the
 * reference, whose derivatives() is its run-phase scheme): mode UPH_MODE_CONFIG
 * (default: the description above) or UPH_MODE_REF, kernel UPH_KERNEL_CUBIC or
 * UPH_KERNEL_POLY7 (include/uph.h).  Without a call, the environment variable
 * UPH_DERIVATIVES_SCHEME selects it at the first derivatives() call: "ref" (the
 * reference's scheme with its PolyOrderSeven kernel), "ref-cubic" or "config".
 *
 * UPH_MODE_REF evaluates derivatives.c:4196-4315 as the reference's decks run it
 * (flagInitSimul 1, spring walls; stressType 1, or 2 as the System says) at the
 * System's state: it reads
 * the fluid rows' DynamicPressure, localDensity, localPressure and localVelocity
 * as the caller's host loop (sph.cc:90-168) left them, plus thermalDiffusivity0,
 * viscosity0 and epsilon, and writes ddensitydt, ddensitydt_approx, dvdt,
 * dvdt_approx, dpdt, dpdt_approx, densityFluctuation, pressureFluctuation,
 * velocityFluctuation and Pressure (Tait, output only); DynamicPressure stays the
 * caller's.  Returns 0, or -1 for an unknown mode or kernel. */
int uph_derivatives_scheme(int mode, int kernel);

#ifdef __cplusplus
}
#endif

#endif
