/*
 * uph.h -- C ABI of libuph_gpu.so, the MI355X (gfx950) implementation of the
 * per-step SPH hot path of Uph.
 *
 * Two kinds of entry points:
 *
 *  1. Drop-in replacements for the reference's own solver seam:
 *       SearchNeighbors   -- same signature and contract as src/lib/Search.cc:5
 *                            (declared extern at src/lib/derivatives.c:34-42,
 *                            called from derivatives.c:1076-1086).
 *
 *  2. A device-resident stepping API: the leapfrog of src/bin/sph.cc:90-168
 *     with the rates of derivatives() (src/lib/derivatives.c:3137) computed on
 *     the GPU, so a host driver runs many steps -- of one scene or of a batch of
 *     independent scenes -- with no per-step host round trip.  The reference has
 *     no library step function (its integrator lives in main), so this mirrors
 *     the reference's call order rather than a reference symbol:
 *       uph_create   ~ CreateSystem + input   (System.c:38, input.c:22)
 *       uph_step     ~ N iterations of the loop body sph.cc:90-168
 *       uph_get_state~ reading sys->Position / Velocity / density
 *       uph_destroy  ~ DestroySystem          (System.c:347)
 *
 * Conventions: every function returns 0 (or a count) on success and a negative
 * value on failure, never exits the process; uph_last_error() describes the
 * last failure on a handle (or of uph_create when the handle is NULL).  One
 * host thread per handle; each handle owns one HIP stream and its device
 * memory; the caller owns every host buffer.  Particle arrays are in the
 * reference's order (fluid, solid, fixed boundary, moving boundary) and
 * positions/velocities are interleaved (x0, y0, x1, y1, ...).
 */
#ifndef UPH_H_
#define UPH_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- drop-in */

/*
 * Replaces SearchNeighbors (src/lib/Search.cc:5-44).  For every fluid query
 * s < NumberOfFluidParticles, takes the NumberOfNeighbors nearest particles of
 * all NumberOfParticles other than s itself (ranked by squared distance, ties
 * by index) and emits (s, j, d2) for those with j > s, d2 < dist and
 * j < ParticleBoundary, query-major, each query's pairs in rank order.
 * Squared distances are FP64 dx*dx + dy*dy with dx = x_s - x_j, uncontracted.
 * Returns the pair count niac.  Like the reference it trusts the caller's
 * buffers to hold every pair; uph_search_neighbors_cap adds a capacity check.
 * Returns -1 on a device error.
 */
int SearchNeighbors(const double **Position, int *i_pair, int *j_pair, double *rij2,
                    const int NumberOfFluidParticles, const int NumberOfParticles,
                    int ParticleBoundary, const int NumberOfNeighbors, const double dist);

/* SearchNeighbors on a contiguous N x 2 position array with a pair capacity:
 * returns -2 (writing nothing) when more than `capacity` pairs qualify. */
int uph_search_neighbors_cap(const double *pos_xy, int *i_pair, int *j_pair, double *rij2,
                             int n_fluid, int n_particles, int particle_boundary,
                             int n_neighbors, double dist, long capacity);

/* uph_search_neighbors_cap without any allocation: no hipMalloc, hipFree, pinned
 * or host heap allocation in the call.  No counterpart in the reference; it is for
 * callers that cannot allocate (serving loops, code whose device memory is fixed in
 * advance).  Every device buffer of the search lives in one file-scope device array
 * of 32.5 MB that each GPU the library runs on holds until the process ends; its
 * device address is looked up once per device at the first call there.
 * Limits: n_particles <= 65536 and at most 1,572,864 pairs; beyond them it returns
 * -3 with a message and writes nothing.  -2: more pairs than `capacity` (nothing
 * written); -1: invalid input or a device error.  Inputs and outputs are host
 * arrays, as for uph_search_neighbors_cap; the work runs on the null stream and the
 * call returns when the pairs are in the caller's buffers.  Every call on one GPU
 * shares that GPU's array, which keeps the last call's contents (each call clears
 * what it needs zeroed): calls from several host threads of this process are
 * serialized by a lock, and two calls that overlapped on one GPU would race on it. */
int uph_search_neighbors_noalloc(const double *pos_xy, int *i_pair, int *j_pair, double *rij2,
                                 int n_fluid, int n_particles, int particle_boundary,
                                 int n_neighbors, double dist, long capacity);

/* Message of the last failure of SearchNeighbors / uph_search_neighbors_cap /
 * uph_search_neighbors_noalloc. */
const char *uph_search_last_error(void);

/* ---------------------------------------------------------- stepping API */

/* Kinematics of the moving boundary particles (the n_moving rows; inputbp.c:192-283). */
enum {
  UPH_MOTION_NONE = 0,
  UPH_MOTION_PISTON = 1,      /* movingBoundary 1: x = S/2 (1 - cos wt), vx = S w/2 sin wt, localVelocity x = vx */
  UPH_MOTION_SINE_OFFSET = 2, /* x = x_deck + A sin wt, vx = A w cos wt (no reference counterpart) */
  UPH_MOTION_LID = 3,         /* movingBoundary 3: x = x_deck + U t, vx = U, localVelocity x = U */
  UPH_MOTION_GATE = 4,        /* movingBoundary 2: y = U t + gate_y for 0 < t < t_stop, then U t_stop + gate_y,
                                 vy = U then 0; t_stop = 1.5 (depth - 15 dx) / U */
  UPH_MOTION_SLOSH = 5        /* virtualType 5: with L = S/2 (1 - cos wt): itype 2 rows at x = L, itype 3 rows at
                                 x = flat_bottom + L, itype 1 rows (the k-th moving row) at x = k dx/2 + L;
                                 vx = S w/2 sin wt */
};

/* Smoothing kernel of the pair sums. */
enum {
  UPH_KERNEL_CUBIC = 0,       /* cubic spline, Kernel (derivatives.c:44-66) */
  UPH_KERNEL_POLY7 = 1        /* PolyOrderSeven W and its gradient as CalculateKernel evaluates them
                                 (derivatives.c:767-823, 1099-1101): grad W = -a0 (1 - q^2/4)^7
                                 (2048 - 3584 q^2 + ... + 305 q^12) dx / h^2, a0 = 2145/(192512 pi h^2) */
};

/* Scheme of the pair sums and the leapfrog. */
enum {
  UPH_MODE_CONFIG = 0,        /* Tait pressure in the momentum terms, artificial viscosity alpha_av */
  UPH_MODE_REF = 1            /* the reference's run phase (derivatives.c:4196-4315, flagInitSimul = 1,
                                 stressType = 1): DynamicPressure integrated by the pressure transport
                                 equation (:1569-1672, :1254-1331), viscous forces (:1770-1839,
                                 :1448-1515), density / pressure / velocity fluctuations
                                 (:2060-2171) in the local fields of the leapfrog (sph.cc:90-168) */
};

/* One independent scene.  Arrays hold n = n_fluid + n_fixed + n_moving rows. */
typedef struct uph_scene {
  int n_fluid, n_fixed, n_moving;
  const double *pos;    /* 2n */
  const double *vel;    /* 2n */
  const double *rho;    /* n  : density (walls keep their deck density)   */
  const double *mass;   /* n  */
  const int *itype;     /* n  : 1 fluid, 2 fixed, 3 moving, 5 spring wall   */
  const double *lvel;   /* 2n : localVelocity deck columns                  */
  double dx;            /* lattice spacing; spring range r0 = 3 dx          */
  double h;             /* smoothing length (all particles)                 */
  double rho0, gamma, c0, B;  /* Tait: P = B((rho/rho0)^gamma - 1)          */
  double g;
  int gravity;          /* 1: a_y -= g on fluid                            */
  double alpha_av;      /* Monaghan artificial viscosity alpha (beta = 0)   */
  double k_spring, c_damp;    /* springConstant0, dampingConstant0          */
  double dt;
  int K;                /* neighbour cap of the search (40 in the reference) */
  int motion;           /* UPH_MOTION_*                                     */
  double m_amp, m_omega;
  double box[4];        /* xmin, ymin, xmax, ymax covered by the cell grid  */
  int kernel;           /* UPH_KERNEL_*; the same for every scene of a batch */
  const double *pres;   /* n  : DynamicPressure deck column (UPH_MODE_REF; NULL: zero) */
  int mode;             /* UPH_MODE_*; the same for every scene of a batch */
  double kappa, mu, eps;      /* thermalDiffusivity0, viscosity0, epsilon (UPH_MODE_REF) */
  double lid_u;         /* lidVelocity U (UPH_MOTION_LID, UPH_MOTION_GATE)        */
  double depth;         /* LocalDepth (UPH_MOTION_GATE's stop time)               */
  double flat_bottom;   /* FlatBottomLength (UPH_MOTION_SLOSH's right wall)       */
  const double *gate_y; /* n: deck column 9, tempPositionGate (UPH_MOTION_GATE; NULL: zero) */
  int vtype;            /* virtualType: mirror ("virtual") particles of derivatives.c:3193-4192,
                           which the search ranks among the K nearest and never pairs;
                           1, 3, 4, 5, 7, 8 have them, other values none */
  double lbp, rbp;      /* LeftBoundaryPosition, RightBoundaryPosition where the motion does
                           not set them (UPH_MOTION_LID, UPH_MOTION_GATE, and rbp without
                           UPH_MOTION_SLOSH); the mirror sections read them */
  int stress_type;      /* stressType (UPH_MODE_REF, derivatives.c:4231-4244): 2 = the viscous
                           stress of ClassicShearStress (:1686-1721) and its divergence,
                           ViscousForcesTwoStep (:1724-1768), in dvdt; other values the
                           single-step viscous forces (:1770-1839).  The same for every scene
                           of a batch; 2 needs a handle without mirror rows whose queries are
                           never capped (the stress sums take every neighbour in range) */
} uph_scene;

typedef struct uph_opts {
  int use_graph;        /* 1: replay each step as one captured hipGraph     */
  int keep_rates;       /* 1: keep the last step's rates for uph_get_rates */
  int pair_variant;     /* 0: compacted two-pass pair kernel (default), 1: direct loop */
  int pair_block;       /* work-group size of the compacted pair kernel: 128, 256 (default), 512 */
  double cell_factor;   /* cell edge = cell_factor x search radius, 1 <= f <= 4 (default 1) */
  int sort_every_step;  /* 1: re-sort the particles into their cells every step; -1: lazy
                           re-sort, where the implementation allows it (no mirror rows, no
                           scene expected to be capped, no hipGraph): cells grow by the skin
                           and the particles are re-sorted only when one of them has left its
                           cell by more than the growth; pair sets are unchanged.  0 (default):
                           lazy for handles of at most 16 scenes, every step for larger
                           batches (a batch re-sorts when any of its scenes needs it) */
  double skin;          /* lazy re-sort: cell edge x (1 + skin); <= 0: 0.02 for one scene,
                           0.05 for a batch (at most 1) */
} uph_opts;

typedef struct uph_handle uph_t;

int uph_create(const uph_scene *scenes, int n_scenes, const uph_opts *opts, uph_t **out);
void uph_destroy(uph_t *h);

/* Advance every scene by n_steps leapfrog steps (asynchronous on the handle's stream). */
int uph_step(uph_t *h, int n_steps);

/* uph_step in one persistent launch: all n_steps whole steps (binning, scan, reorder,
 * pair sums, capped pass) run in one cooperative kernel whose blocks meet at grid-wide
 * barriers between the phases (7 per step), instead of 7 launches per step.  For
 * handles without mirror rows (virtualType sections), whose queries are not expected to
 * be capped (4 pi (h/dx)^2 <= 3/4 K) and which use the compacted pair kernel; -1
 * otherwise.  Results equal uph_step's bit for bit.  No counterpart in the reference.
 * A barrier that waits too long gives up rather than hang; uph_sync then fails. */
int uph_step_persistent(uph_t *h, int n_steps);

/* Wait for the handle's stream; returns the number of steps in which some fluid
 * particle had more than K neighbours in range (exact capped path taken). */
int uph_sync(uph_t *h);

/* Current state of one scene in the reference's order (any pointer may be NULL):
 * fluid rows after the last step; moving boundary rows where the last step's
 * derivatives(t) placed them (inputbp.c:192-283), at the deck before the first step. */
int uph_get_state(uph_t *h, int scene, double *pos, double *vel, double *rho);

/* Mirror rows of one scene from the last step's derivatives() (derivatives.c:3193-4192):
 * the rows NumberOfVirtualParticles counts, in the reference's order (section by
 * section, sources in index order), with the fields derivatives.c copies into them
 * from their source at that call: Position, Velocity, localVelocity (2 per row),
 * density, localDensity, DynamicPressure (ref mode; config mode: the Tait pressure of
 * the density), localPressure, mass.  Returns the number of rows, or -1 when it
 * exceeds cap (any pointer may be NULL).  uph_ghost_capacity: the most rows a step
 * can make (one per fluid row and section). */
int uph_get_ghosts(uph_t *h, int scene, int cap, double *pos, double *vel, double *lvel, double *rho,
                   double *rholoc, double *P, double *Ploc, double *mass);
int uph_ghost_capacity(uph_t *h, int scene);

/* Rates of the last step (needs opts.keep_rates): R = ddensitydt, Rt = ddensitydt_approx,
 * a = dvdt, at = dvdt_approx; fluid rows only (n_fluid, n_fluid, 2 n_fluid, 2 n_fluid). */
int uph_get_rates(uph_t *h, int scene, float *R, float *Rt, float *a, float *at);

/* UPH_MODE_REF: DynamicPressure of one scene in the reference's order (n values;
 * walls hold their deck value).  UPH_MODE_CONFIG: the Tait pressure of the
 * density. */
int uph_get_pressure(uph_t *h, int scene, double *P);

/* UPH_MODE_REF rates of the last step (needs opts.keep_rates), fluid rows:
 * dpdt, dpdt_approx, densityFluctuation, pressureFluctuation (n_fluid each),
 * velocityFluctuation (2 n_fluid).  Any pointer may be NULL. */
int uph_get_ref_rates(uph_t *h, int scene, float *RP, float *RPt, float *drho, float *dP, float *dv);

/* Completed steps. */
int uph_itime(const uph_t *h);

/* The cells' growth of a lazy re-sort handle (uph_opts.skin after its defaults), 0 for a
 * handle that re-sorts every step.  No counterpart in the reference. */
double uph_skin(const uph_t *h);

/* Lazy re-sort (uph_opts.sort_every_step <= 0): the re-sorts run so far, counting the
 * first step's sort and one after each state set from outside (waits for the
 * handle's stream); -1 for a handle that re-sorts every step.  No counterpart in the
 * reference, whose kd-tree is rebuilt from scratch every step (Search.cc:5-44). */
int uph_resorts(uph_t *h);

/* Run n_steps with HIP events on the handle's stream: *ms_total spans the
 * whole loop, *ms_pair is the mean duration of the fused pair kernel. */
int uph_time_steps(uph_t *h, int n_steps, double *ms_total, double *ms_pair);

/* As uph_time_steps, with the step split into its phases (means per step, any
 * pointer may be NULL): *ms_sort the binning, scan and reorder kernels (and the
 * mirror rows' kernels), *ms_pair the compacted pair pass, *ms_exact the capped
 * (K-nearest) pass that follows it, which returns at once unless some query of
 * the step had more than K neighbours in range.  With all three NULL no event is
 * recorded between the kernels, and *ms_total times the bare loop. */
int uph_time_phases(uph_t *h, int n_steps, double *ms_total, double *ms_sort, double *ms_pair,
                    double *ms_exact);

const char *uph_last_error(const uph_t *h);

/* Diagnostics that sph.cc's output reports, over a scene's current fluid state:
 * TotalKineticEnergy (derivatives.c:2254-2263), NeighborsToPressurePoint
 * (:2290-2338; the pressure is the Tait pressure of each particle) and
 * DensityConsistencyFunctions (:3090-3117, whose first value the reference
 * always reports as 0).  probe_pressure is 0 and probe_norm 1 when no particle
 * is within 3 h of the probe. */
typedef struct uph_diag {
  double kinetic_energy, potential_energy;
  double probe_pressure, probe_norm;
  int probe_neighbors;
  double consistency_one, consistency_two, consistency_three, density_uncertainty;
} uph_diag;
int uph_diagnostics(uph_t *h, int scene, double probe_x, double probe_y, uph_diag *out);

/* The second pressure probe (NeighborsToPressurePointTwo + EnsembleAveragedPointPressureTwo,
 * derivatives.c:2340-2393, written by output.c:77-92): the mean of the positive
 * DynamicPressure of the fluid rows strictly within 2h of (x, y), and their count (0 and
 * 0 when none).  The reference reads the location from pressureProbeXX/YY, which nothing
 * assigns (System.c:130-131 is commented out); here the caller names it. */
int uph_probe_two(uph_t *h, int scene, double x, double y, double *pressure, int *neighbors);

/* approxVorticity of the fluid rows (ApproximateVortcity, derivatives.c:2840-2885) as the
 * last step's derivatives() evaluation computed it (positions at the start of the step,
 * predicted velocities and densities, the step's pair set), n_fluid values in the
 * reference's order. */
int uph_vorticity(uph_t *h, int scene, double *out);

/* Device-side state exchange with the surrogate (no counterpart in the reference).
 * Arrays are device pointers in the reference's order (pos/vel: 2n doubles);
 * uph_state_to_device writes what uph_get_state returns.
 * uph_state_from_device loads the fluid rows (walls keep their state); with
 * restart != 0 the next step takes the first-step form of the leapfrog
 * (sph.cc:154-167) with localVelocity = velocity, and uph_itime restarts at 0
 * (the wall kinematics keep their clock).
 * Both calls, and uph_graph_build, synchronise the device before touching caller
 * buffers (which may be in use on any stream) and return with their results
 * complete. */
int uph_state_to_device(uph_t *h, int scene, double *pos, double *vel, double *rho);
int uph_state_from_device(uph_t *h, int scene, const double *pos, const double *vel, const double *rho,
                          int restart);

/* Radius graph of n device positions (2n doubles): receivers are the first n_recv
 * particles; an edge (i <- j) exists for every j != i with squared distance < dist.
 * off (n+1): CSR by receiver (0 edges for non-receivers); send/recv (E): edge ends in
 * candidate order; soff (n+1), sperm (E): for node j, the ids of its edges as a sender
 * (both NULL: no sender view is built).  Returns -2 (and *n_edges) when E exceeds
 * capacity. */
int uph_graph_build(const double *pos, int n, int n_recv, double dist, int *off, int *send, int *recv,
                    int *soff, int *sperm, long capacity, long *n_edges);
const char *uph_graph_last_error(void);

/* ------------------------------------------ surrogate (graph network) kernels
 *
 * No counterpart in the reference: the learned acceleration surrogate that runs
 * beside the plain solver.  Device buffers are raw pointers from uph_dev_alloc;
 * `stream` is a hipStream_t (NULL: the default stream).  Matrices are row-major;
 * BF16 arrays hold __bf16, and N, K (unless trans_a) and every leading
 * dimension must be multiples of 8.
 */
int uph_dev_alloc(void **p, size_t bytes);
int uph_dev_free(void *p);
int uph_dev_upload(void *dst, const void *src, size_t bytes);
int uph_dev_download(void *dst, const void *src, size_t bytes);
int uph_dev_copy(void *dst, const void *src, size_t bytes);
int uph_dev_zero(void *p, size_t bytes);
int uph_dev_sync(void);
const char *uph_gnn_last_error(void);

/* C = act(op(A) B + bias) with BF16 MFMA and FP32 accumulation.  trans bit 0:
 * op(A) is the transpose of the K x M matrix A (else A is M x K); bit 1: B is
 * given as its N x K transpose (row stride ldb; K a multiple of 8), else K x N.
 * splits > 1 divides K among work-groups (FP32 partials in `work`,
 * M*N*splits floats); accumulate adds into an FP32 C (needs `work`, M*N
 * floats).  act: 0 none, 1 ReLU. */
int uph_gemm_bf16(int M, int N, int K, const void *A, int lda, int trans, const void *B, int ldb,
                  const float *bias, void *C, int ldc, int out_bf16, int act, int splits, float *work,
                  int accumulate, void *stream);
int uph_cast_f32_bf16(long n, const float *x, void *y, void *stream);
int uph_cast_bf16_f32(long n, const void *x, float *y, void *stream);
int uph_relu_bwd(long n, const float *dy, const void *y, void *dz, void *stream);
int uph_add_f32(long n, const float *x, float *y, void *stream);
int uph_colsum_f32(int M, int N, const float *x, int ldx, float *out, int accumulate, void *stream);
/* out[i] = sum_{e in [off[i], off[i+1])} x[idx ? idx[e] : e] (rows of F values) */
int uph_segsum(int nseg, int F, const int *off, const int *idx, const void *x, int ldx, int x_bf16,
               float *out, int ldo, int accumulate, void *stream);
int uph_gather_rows(int n, int F, const int *rows, const void *x, int ldx, void *y, int ldy, int esize,
                    void *stream);
int uph_adam(long n, float *w, const float *g, float *m, float *v, void *w_bf16, float lr, float b1,
             float b2, float eps, int t, void *stream);
int uph_mse(int M, int C, const float *y, int ldy, const float *t, int ldt, const int *mask, int count,
            float *dy, float *loss_part, void *stream);
int uph_transpose_bf16(int K, int N, const float *w, void *wt, void *stream);
int uph_gather_add2(long E, int F, const int *ra, const float *a, const int *rb, const float *b, int ldab,
                    float *z, int ldz, void *stream);
int uph_relu_to_bf16(long M, int F, const float *z, int ldz, void *y, int ldy, void *stream);
int uph_rows_to_bf16(long M, int F, const float *x, int ldx, void *y, int ldy, void *stream);
int uph_fill_col_bf16(long M, int ld, int col, float v, void *y, void *stream);
int uph_relu_bwd_rows(long M, int F, const float *dy, int lddy, const void *y, int ldy, void *dz, int ldz,
                      void *stream);
int uph_add_rows_f32(long M, int F, const float *x, int ldx, float *y, int ldy, void *stream);
/* Surrogate inputs and update: node features from 5 macro-step velocities (hist: 5 x n x 2,
 * newest first) and a wall flag; edge features (x_recv - x_send)/R, |r|/R; semi-implicit
 * Euler with the predicted acceleration; SPH summation density over the graph. */
int uph_gns_node_feats(int n, int n_recv, const double *hist, float inv_vscale, void *out, void *stream);
int uph_gns_edge_feats(long E, const int *send, const int *recv, const double *pos, double invR, void *out,
                       void *stream);
int uph_gns_integrate(int n_recv, double *pos, double *vel, const float *acc, int lda, double dt,
                      double ascale, void *stream);
int uph_sph_density(int n_recv, const int *off, const int *send, const double *pos, const double *mass,
                    double h, double *rho, void *stream);
/* One message-passing step on the edges at width 128, fused: z = [e,1] A + zb[recv] +
 * zc[send], u = relu(z), m = [u,1] D, e += m, eb_out = bf16(e), m_out = m.  eb_in/eb_out
 * (may alias), u1_out: E x 136 BF16 rows (latent, the constant 1, zeros); e, m_out: E x 128
 * FP32; zb, zc: n x 128 FP32; At, Dt: the weights transposed, 128 x 136 BF16.  u1_out
 * (optional) receives u for the backward pass. */
int uph_gns_edge_step(long E, const void *eb_in, void *eb_out, float *e, const float *zb, const float *zc,
                      const int *recv, const int *send, const void *At, const void *Dt, float *m_out,
                      void *u1_out, void *stream);
/* FP64 elementwise helpers of the rollout: out = a x + b y (macro-step velocities from
 * positions); out = x * num / den (density carried across a surrogate block). */
int uph_lincomb_f64(long n, double a, const double *x, double b, const double *y, double *out, void *stream);
int uph_scale_ratio_f64(long n, const double *x, const double *num, const double *den, double *out,
                        void *stream);

/* The rollout network's edge kernels (no edge latent; width W <= 128).  Receivers are the
 * rows i < n_recv of the receiver-ordered CSR (off, send); r_e = ((x_i - x_j)/R, |.|/R, 1)
 * from the FP64 positions (pos: 2n doubles); A: FP32 rows 0..3 (row stride ldA);
 * zb, zc: FP32 node rows.
 * uph_sn_edge_agg: s_i = sum_e relu(r_e A + zb_i + zc_j) in columns 0..W-1 of s (row
 *   stride lds >= W + 1), the edge count in column W, zeros up to lds.
 * uph_sn_edge_bwd: with ds (the gradient of s's first W columns): dzb_i and dzc_j, the
 *   sums of g_e = ds_i [pre_e > 0] over i's and j's edges (W columns, row stride lddz),
 *   and dA (rows 0..3) += sum_e r_e^T g_e; part: 1024 W floats of scratch.  Every sum
 *   runs in a fixed order (dzc through the sender-ordered permutation soff, sperm). */
int uph_sn_edge_agg(int n, int n_recv, int W, const int *off, const int *send, const double *pos, double invR,
                    const float *A, int ldA, const float *zb, int ldzb, const float *zc, int ldzc, float *s,
                    int lds, void *stream);
int uph_sn_edge_bwd(int n, int n_recv, int W, const int *off, const int *send, const int *soff, const int *sperm,
                    const int *recv, const double *pos, double invR, const float *A, int ldA, const float *zb,
                    int ldzb, const float *zc, int ldzc, const float *ds, int ldds, float *dzb, float *dzc,
                    int lddz, float *dA, float *part, void *stream);
/* The rollout network's inference at width 128, fused: mode 0 the node encoder (xn: n x 16
 * BF16; L1t, L2t the transposed encoder weights) writing h and step 0's projections; mode 1
 * one message-passing step (the edge sums from A, zb, zc; the node MLP D, E, F, G; h
 * updated in place) writing the next step's projections zb_out, zc_out (Bt, Ct); mode 2
 * the last step with the decoder (P1t, P2t) writing out (n x 8 FP32, columns 0..1 used).
 * Transposed weights are BF16 with row stride 136 (the bias at column 128; Ft: 128 x 128,
 * P2t: 8 rows); zb, zb_out, h: n x 128 FP32; zc, zc_out: n x 128 BF16.  Buffers a mode
 * does not read or write may be NULL. */
int uph_sn_fused(int mode, int n, int n_recv, const int *off, const int *send, const double *pos, double invR,
                 const float *A, const float *zb, const void *zc, const void *Dt, const void *Et, const void *Ft,
                 const void *Gt, float *h, const void *Bt, const void *Ct, float *zb_out, void *zc_out,
                 const void *P1t, const void *P2t, float *out, const void *xn, const void *L1t, const void *L2t,
                 void *stream);
const char *uph_sn_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* UPH_H_ */
