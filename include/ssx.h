/*
 * ssx.h -- C ABI of the MI355X (gfx950) elliptic-solve engine for SSENSS.
 *
 * All compute runs in hand-written HIP kernels on one GPU; the host side is
 * C.  Vectors are element-local "E-vectors" exactly as the reference stores
 * a double**** field: one contiguous block of nel*n^3 doubles, node (e,k,j,i)
 * at ((e*n+k)*n+j)*n+i, i fastest (reference utils.c:64-81).  Unless a
 * function says "host", vector arguments are DEVICE pointers.
 *
 * Every call is ordered on the context's HIP stream (ssx_set_stream).
 * Functions that return a number to the host (ssx_dot, the solvers)
 * synchronise that stream before returning; the others only enqueue work.
 *
 * Status: every function returns SSX_OK (0) or a positive SSX_E* code;
 * ssx_last_error() gives a message for the last failure on this thread.
 *
 * Reference interfaces replaced (file:line in the SSENSS src/ tree):
 *   ssx_create          connect_elements + initialise_preconditioners,
 *                       init.c:38-50, precon.c:200-209 (the upload hook)
 *   ssx_ax              discrete_laplace_operator poisson.c:34-56,
 *                       discrete_helmholtz_operator helmholtz.c:33-49
 *   ssx_opeval          global_opeval solver.c:191-198
 *   ssx_dss             direct_stiffness_summation global.c:241-257
 *   ssx_dss_average     direct_stiffness_averaging global.c:259-268
 *   ssx_dot             global_dot_product global.c:270-282
 *   ssx_mask            mask_dirichlet_boundaries poisson.c:89-99
 *   ssx_nullspace       subtract_nullspace solver.c:200-213
 *   ssx_restrict        multigrid_restriction precon.c:283-291
 *   ssx_prolong         multigrid_prolongation precon.c:293-300
 *   ssx_jacobi          jacobi_preconditioner precon.c:311-324
 *   ssx_mg_precon       multigrid_preconditioner precon.c:227-236
 *   ssx_cg              conjugate_gradient solver.c:33-81
 *   ssx_fcg             flexible_conjugate_gradient solver.c:83-157
 *   ssx_solve_poisson_cg    solve_poisson_cg solver.c:159-173 (host arrays)
 *   ssx_solve_helmholtz_cg  solve_helmholtz_cg solver.c:175-189 (host arrays)
 *   ssx_particles_set_velocity  compute_reference_domain_velocity
 *                       geometry.c:188-198
 *   ssx_particles_step  integrate_particle_equations particles.c:182-218
 *   ssx_particles_*_host    the same on host arrays (the shim's entry)
 */
#ifndef SSX_H
#define SSX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSX_OK        0
#define SSX_EINVAL    1   /* bad argument                                 */
#define SSX_ENOMEM    2   /* host or device allocation failed             */
#define SSX_EHIP      3   /* a HIP runtime call failed                    */
#define SSX_ENOTCONV  4   /* solver stopped at max_iter above eps         */
#define SSX_EMESH     5   /* connectivity the engine cannot assemble      */
#define SSX_ENODEV    6   /* no GPU visible                               */
#define SSX_ENUMERIC  7   /* NaN or Inf residual                          */

/* variable codes, reference poisson.h:24-27 (enum PRESSURE ...) */
#define SSX_PRESSURE    0   /* no Dirichlet mask, null space removed      */
#define SSX_X_VELOCITY  1   /* Dirichlet mask, no null space              */
#define SSX_Y_VELOCITY  2
#define SSX_Z_VELOCITY  3

#define SSX_OP_POISSON    0
#define SSX_OP_HELMHOLTZ  1

#define SSX_PC_JACOBI  0
#define SSX_PC_MG      1

/* ssx_level_data selectors */
#define SSX_DATA_GFAC      0   /* 6*nel*n^3: G*w*detJ, [e][xx,xy,xz,yy,yz,zz][node] */
#define SSX_DATA_MASS      1   /* w*detJ                                   */
#define SSX_DATA_DIAG_P    2   /* inverse assembled Poisson diagonal       */
#define SSX_DATA_DIAG_H    3   /* inverse assembled Helmholtz diagonal     */
#define SSX_DATA_SCALE     4   /* 1 / multiplicity                         */
#define SSX_DATA_MASK      5   /* Dirichlet mask                           */
#define SSX_DATA_INVMASS   6   /* 1 / DSS(w*detJ), level 0                 */
#define SSX_DATA_X         7   /* coordinates (x, y, z)                    */
#define SSX_DATA_Y         8
#define SSX_DATA_Z         9

typedef struct ssx_ctx ssx_ctx;

/* Mesh as SSENSS holds it after read_grid (dataio.c:197-227) and
 * set_boundary_conditions (geometry.c:230-237): GLL coordinates per element
 * and the grid.h5 connectivity fields (dataio.c:45-62).  Neighbour ids are
 * element numbers >= 0 or boundary codes < 0 (-1 no-slip, -2 free-slip). */
typedef struct ssx_mesh {
	int n;                   /* GLL points per direction (polynomial order + 1) */
	int nel;
	const double *x, *y, *z;        /* nel*n^3 each, host */
	const int *neigh;               /* nel*6: W,E,S,N,B,T */
	const int *edge_neigh;          /* nel*12 */
	const int *edge_size;           /* nel*12 */
	const int *corn_neigh;          /* nel*8 */
	const int *corn_size;           /* nel*8 */
	double dt, visc;                /* Helmholtz lambda = (11/6)/(dt*visc) */
	int svvn;                       /* SVV cut-off mode, reference main.c:54 */
	double svva;                    /* SVV amplitude, reference main.c:55 */
} ssx_mesh;

typedef struct ssx_stats {
	int iterations;          /* loop bodies of the outer solver            */
	int converged;           /* residual <= eps                            */
	double residual;         /* final sqrt(<r,r>_w)                        */
	long long fine_ax;       /* fine-level operator applications          */
	long long coarse_iters;  /* coarsest-level CG iterations (all cycles)  */
	long long smooth_iters;  /* smoother CG iterations (all levels)        */
	long long vcycles;
	double seconds;          /* wall time of the solve                     */
} ssx_stats;

const char *ssx_last_error(void);
const char *ssx_version(void);

/* ---- structured mesh generator (host only; no device needed) ----
 * A box of mx*my*mz hexahedra on [lo,hi] with GLL nodes, connectivity in the
 * conventions of the reference's misc/rectangular_grid.py:80-135 (element
 * id = i + j*mx + k*mx*my; faces/edges/corners wrap periodically, then
 * boundary faces with bc[s] < 0 get that code with edge/corner size 0).
 * bc[6] per face W,E,S,N,B,T: 0 periodic, -1 no-slip, -2 free-slip.
 * deform: 0 none; 1 x += a sin(pi u)sin(pi v)sin(pi w) applied to each
 * coordinate, with (u,v,w) the unit-cube coordinates; 2 tanh wall
 * stretching of y with factor a: y = tanh(a eta)/tanh(a), eta in [-1,1].
 * Arrays are caller-allocated with the sizes of ssx_mesh. */
int ssx_mesh_box(int n, int mx, int my, int mz, const double lo[3],
		const double hi[3], const int bc[6], int deform, double amp,
		double *x, double *y, double *z, int *neigh, int *edge_neigh,
		int *edge_size, int *corn_neigh, int *corn_size);

/* ---- host-side assembly helpers (no device needed) ----
 * ssx_dss_host: direct stiffness summation of a HOST E-vector whose elements
 * have n points per direction (any level), with the shared-node groups and
 * add order the device kernel uses (reference global.c:241-257).
 * ssx_geometry_host: level-0 premultiplied factors G*w*detJ (6*nel*n^3,
 * layout of SSX_DATA_GFAC) and w*detJ (nel*n^3) of a mesh (reference
 * geometry.c:102-135, poisson.c:58-66). */
int ssx_dss_host(const ssx_mesh *mesh, int n, double *u);
/* ssx_dss_gather_host: the same sum in the gather form the fused CG update
 * evaluates on the device (each copy adds its partners' values, read from
 * per-copy partner lists, in its own add order); equal to ssx_dss_host bit
 * for bit. */
int ssx_dss_gather_host(const ssx_mesh *mesh, int n, double *u);
int ssx_geometry_host(const ssx_mesh *mesh, double *gfac, double *mass);

/* ---- context ---- */
int ssx_create(ssx_ctx **ctx, const ssx_mesh *mesh);
void ssx_destroy(ssx_ctx *ctx);
/* Run all later work of ctx on this hipStream_t.  NULL is the HIP null
 * (default) stream, e.g. torch's default stream; a context starts on a
 * non-blocking stream of its own, which ssx_own_stream() returns. */
int ssx_set_stream(ssx_ctx *ctx, void *hip_stream);
void *ssx_own_stream(ssx_ctx *ctx);
void *ssx_get_stream(ssx_ctx *ctx);
int ssx_synchronize(ssx_ctx *ctx);
/* levels, points per direction of a level, element count, local nodes */
int ssx_nlevels(const ssx_ctx *ctx);
int ssx_level_n(const ssx_ctx *ctx, int level);
int ssx_nel(const ssx_ctx *ctx);
size_t ssx_level_ntot(const ssx_ctx *ctx, int level);
double ssx_lambda(const ssx_ctx *ctx);
/* copy level data (SSX_DATA_*) to a host array */
int ssx_level_data(ssx_ctx *ctx, int level, int which, double *host);
/* shared-node groups of a level by multiplicity 2, 4, 8 */
int ssx_dss_groups(const ssx_ctx *ctx, int level, long long counts[3]);
/* Jacobi diagonal choice: 0 Poisson, 1 Helmholtz, -1 by var (default) */
int ssx_set_diag_choice(ssx_ctx *ctx, int choice);
/* options: SSX_OPT_FUSED_CG 1 (default) runs each CG iteration as four or
 * five fused kernels with the stopping test on the device; 0 runs the plain
 * one-kernel-per-vector-operation loop with a host check per iteration */
#define SSX_OPT_FUSED_CG 1
/* SSX_OPT_SMOOTHER_PRECISION: precision of the p-multigrid smoother and
 * order transfer on all but the coarsest level: SSX_PREC_FP64 (default, the
 * reference algorithm), SSX_PREC_FP32 (FP32 fields and factors) or
 * SSX_PREC_BF16 (FP32 fields, bfloat16 geometric factors).  The outer FCG,
 * its inner products, the residual update and the coarsest-level CG stay
 * FP64; inner products always accumulate in FP64. */
#define SSX_OPT_SMOOTHER_PRECISION 2
/* SSX_OPT_PERSISTENT_CG: 1 runs the fused CG of every multigrid
 * level with n <= 4 (the coarse levels) as one cooperative launch with
 * grid-wide barriers instead of six launches per iteration; 0 keeps the
 * per-iteration launches (default); a value 2..1024 also fixes the number of
 * workgroups of that launch (default: as many as the device keeps
 * co-resident).  Same algorithm and stopping test; inner products are
 * summed in a different fixed order. */
#define SSX_OPT_PERSISTENT_CG 3
/* SSX_OPT_LP_COARSE_RTOL: k in [1,15]; with a reduced-precision smoother
 * the FP64 coarsest solve stops at max(eps, 10^-k |f|_w) (default k = 3):
 * its correction feeds FP32 levels, so digits beyond theirs are lost. */
#define SSX_OPT_LP_COARSE_RTOL 4
/* SSX_OPT_FUSED_REDUCE: 1 lets the last workgroup of the DSS and of the
 * update kernel form alpha and beta (three launches per fused CG
 * iteration); 0 (default) uses separate single-workgroup reductions
 * (five).  Same summation order: bit-identical results.  Off by default:
 * every workgroup's device-scope release (an L2 write-back per XCD) costs
 * more than the two launches it saves. */
#define SSX_OPT_FUSED_REDUCE 5
/* SSX_OPT_COARSE_ASSEMBLED: 1 (default) runs the p-multigrid coarsest-level
 * Jacobi-PCG over unique nodes with the level's operator assembled once
 * (CSR; at n = 2 a 27-point stencil) instead of on E-vectors with the
 * element operator and DSS, two launches per iteration; 2 the same with
 * four launches per iteration (separate reductions; bit-identical to 1);
 * 0 keeps the E-vector CG.  Same algorithm and stopping test; the
 * E-vector form sums in another order (equal to rounding). */
#define SSX_OPT_COARSE_ASSEMBLED 6
/* SSX_OPT_GATHER_DSS: 1 (default) assembles A p inside the fused CG update:
 * every copy adds its partners' values in its own add order, read from
 * per-copy partner lists, so the DSS launch and its write-back go away;
 * 0 runs the group DSS as a launch of its own.  Bit-identical results. */
#define SSX_OPT_GATHER_DSS 7
/* SSX_OPT_PERSIST_ABORT_TEST: test hook of the persistent CG's failure
 * path: 1 makes the next persistent launch start with its abort word set,
 * so every block leaves at its first grid barrier, as when a barrier wait
 * gives up; the solve then returns SSX_EHIP.  Cleared by that launch. */
#define SSX_OPT_PERSIST_ABORT_TEST 8
/* SSX_OPT_PARTICLE_WALK: how ssx_particles_step / ssx_tracers_step (and
 * ssx_particles_step_host) take the particles, after a counting sort by
 * owning element: 1 (default) one relocation pass per launch, the
 * particles still moving compacted between passes; 0 the whole walk of
 * each particle in one lane of one launch.  Bit-identical results. */
#define SSX_OPT_PARTICLE_WALK 9
/* SSX_OPT_COARSE_DIRECT: m > 0 solves the p-multigrid coarsest level
 * directly, with its assembled operator inverted once (dense, rocSOLVER
 * Cholesky), when it has at most m unique nodes and the smoother runs in
 * reduced precision (SSX_OPT_SMOOTHER_PRECISION != FP64, whose coarse solve
 * stops at a relative tolerance anyway).  The inverse takes 8 m^2 bytes of
 * device memory (m = 16384: 2 GiB).  0 (default): Jacobi-PCG as the
 * reference does.  A coarse operator that is not positive definite (other
 * than the pressure's constant null space) falls back to the CG. */
#define SSX_OPT_COARSE_DIRECT 10
/* SSX_OPT_COARSE_SPARSE: m > 0 solves the p-multigrid coarsest level of a
 * reduced-precision V-cycle (as SSX_OPT_COARSE_DIRECT) by a sparse
 * Cholesky factor when it has at most m unique nodes and more than the
 * dense solve takes: nested-dissection order, multifrontal factorisation
 * at the first solve of each operator (lambda, null space) on the host,
 * fronts of 1024 rows and more on the device through rocSOLVER / rocBLAS
 * (about 0.5 s for 30k nodes, 3 s for 100k), and the two triangular solves
 * on the device, a few launches per level of the separator tree.  Default
 * SSX_COARSE_SPARSE_DEFAULT; 0: Jacobi-PCG to the cycle's relative
 * tolerance.  A coarse operator that is not positive definite falls back
 * to the CG. */
#define SSX_OPT_COARSE_SPARSE 11
#define SSX_COARSE_SPARSE_DEFAULT 131072
#define SSX_PREC_FP64 0
#define SSX_PREC_FP32 1
#define SSX_PREC_BF16 2
int ssx_set_option(ssx_ctx *ctx, int option, int value);

/* ---- operators on device E-vectors ---- */
int ssx_ax(ssx_ctx *ctx, int level, int op, const double *u, double *au);
int ssx_opeval(ssx_ctx *ctx, int level, int op, int var, const double *p,
		double *ap);
int ssx_dss(ssx_ctx *ctx, int level, double *u);
int ssx_dss_average(ssx_ctx *ctx, int level, double *u);
int ssx_dot(ssx_ctx *ctx, int level, const double *x, const double *y,
		double *out_host);
int ssx_mask(ssx_ctx *ctx, int level, int var, double *u);
int ssx_nullspace(ssx_ctx *ctx, int level, int var, double *u);
int ssx_restrict(ssx_ctx *ctx, int level, const double *fine, double *coarse);
int ssx_prolong(ssx_ctx *ctx, int level, const double *coarse, double *fine);
int ssx_jacobi(ssx_ctx *ctx, int level, int var, const double *x, double *px);

/* ---- solvers on device E-vectors (f is masked in place, u is the initial
 * guess and the result, as in the reference) ---- */
int ssx_cg(ssx_ctx *ctx, int level, int op, int var, double *f, double *u,
		double eps, int max_iter, ssx_stats *stats);
int ssx_fcg(ssx_ctx *ctx, int op, int var, int precon, double *f, double *u,
		double eps, int max_iter, ssx_stats *stats);
int ssx_mg_precon(ssx_ctx *ctx, int op, int var, double eps, const double *x,
		double *px);

/* ---- inertial point particles (particles.c), device arrays ----
 * State as struct particle (particles.h:25-41): owning element, previous
 * element, reference coordinates (x,y,z) in [-1,1]^3 of that element,
 * velocity (u,v,w) in reference coordinates, time of the last update t,
 * response time tau.  reinit (may be NULL) is set to 1 for a particle the
 * reference would re-seed at random (more than 10 relocation passes, or no
 * way back into an element): it is restarted at the element centre. */
typedef struct ssx_particles {
	long n;
	int *el, *old_el;
	double *x, *y, *z, *u, *v, *w, *t, *tau;
	int *reinit;
} ssx_particles;
/* compute_reference_domain_velocity (geometry.c:188-198): shift the
 * reference-space velocity history one level and set ur = J^-1 (u,v,w) from
 * physical device E-vectors (level 0) */
int ssx_particles_set_velocity(ssx_ctx *ctx, const double *u,
		const double *v, const double *w);
/* integrate_particle_equations (particles.c:182-218): advance every
 * particle from its t to t_now (params->t) with history step dt
 * (params->dt); *max_passes (may be NULL) receives the largest relocation
 * pass count.  Synchronous: returns after the step has finished.  Particles
 * whose el is not an element index are left untouched and the call returns
 * SSX_EINVAL. */
int ssx_particles_step(ssx_ctx *ctx, const ssx_particles *p, double t_now,
		double dt, int *max_passes);
/* the same two calls on HOST arrays (velocity E-vectors; a particle list
 * whose pointers are host memory), staged through device memory the
 * context keeps; synchronous */
int ssx_particles_set_velocity_host(ssx_ctx *ctx, const double *u,
		const double *v, const double *w);
int ssx_particles_step_host(ssx_ctx *ctx, const ssx_particles *p,
		double t_now, double dt, int *max_passes);
/* RK2 tracers (dx/dt = u_f(x, t), no inertia), the tracer mode of the same
 * walk (no counterpart in the reference; its particles are inertial, RK4):
 * every tracer goes from its t to t_now by Heun's method in the reference
 * coordinates of its element, with the velocity history of
 * ssx_particles_set_velocity.  A stage that leaves [-1,1]^3 is replaced by
 * a straight move to the boundary along the slope that left (k1 if the
 * predictor left, else (k1 + k2)/2), a hop across the closest face with
 * the crossed coordinate set to the opposite face, and a new step over the
 * remaining time; at a wall the tracer stays on the wall for the rest of
 * the step; after more than 10 passes it restarts at the element centre
 * with reinit set.  u, v, w receive the fluid velocity (reference space) at
 * the end point; tau is not read (may be NULL).  Synchronous; a tracer
 * whose el is not an element index is left untouched and the call returns
 * SSX_EINVAL. */
int ssx_tracers_step(ssx_ctx *ctx, const ssx_particles *p, double t_now,
		double dt, int *max_passes);
int ssx_tracers_step_host(ssx_ctx *ctx, const ssx_particles *p,
		double t_now, double dt, int *max_passes);
/* How many particles entered each relocation pass of the last particle or
 * tracer step that ran pass by pass (SSX_OPT_PARTICLE_WALK 1):
 * counts[k], k < SSX_PASS_COUNTS (pass 0: every particle of the step).
 * No counterpart in the reference (a diagnostic of the walk's load).  All
 * zero before such a step. */
#define SSX_PASS_COUNTS 12
int ssx_particles_pass_counts(ssx_ctx *ctx, int counts[SSX_PASS_COUNTS]);
/* ssx_particles_step_host without any device allocation.  No counterpart
 * in the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * host lists are staged through a fixed device array the library holds on
 * each GPU it runs on (until the process ends), for at most
 * SSX_NOALLOC_MAX_PARTICLES particles (more: SSX_EINVAL).  The particle
 * tables must exist already: call ssx_particles_set_velocity(_host) first
 * (else SSX_EINVAL).  All calls on one GPU share that array: calls that
 * overlap on one GPU race.  Results are those of ssx_particles_step_host. */
#define SSX_NOALLOC_MAX_PARTICLES 131072
int ssx_particles_step_host_noalloc(ssx_ctx *ctx, const ssx_particles *p,
		double t_now, double dt, int *max_passes);

/* ---- Navier-Stokes time step (navier_stokes.c:224-235) ----
 * The reference's splitting scheme on the context's mesh, every field in
 * device memory: OIFS convection (three RK4 sub-integrations with dt/2
 * steps, navier_stokes.c:71-92, time.c:84-152) with the dealiased
 * skew-symmetric convection operator at nconv = 3n/2 points per direction
 * (convection.c:31-140, main.c:53); the pressure right-hand side
 * (navier_stokes.c:115-146, stokes.c:107-130), EXT3 pressure estimate and
 * FCG + p-MG pressure solve (ssx_fcg, :148-159); the velocity right-hand
 * side, EXT3 estimate and the three Helmholtz Jacobi-PCG solves (:161-222;
 * with batch_velocity as one ssx_batch_cg of three problems); BDF3/EXT3
 * weights (time.c:51-73); solves to 1e-10 in the weighted norm
 * (residual_tolerance, navier_stokes.c:94-99).  The velocity solves need
 * the new pressure (navier_stokes.c:184), so a step's pressure solve runs
 * before them.  Dirichlet values are zero (the reference's default
 * boundary functions, structs.c:67-68); the body force g is constant
 * (main.c:104-106).  The context's lambda must be (11/6)/(dt visc);
 * n <= 10 (nconv <= 15). */
typedef struct ssx_ns ssx_ns;
typedef struct ssx_ns_params {
	double dt, visc, dens;     /* params->dt, visc, dens                 */
	double t;                  /* time of the newest velocity level      */
	double g[3];               /* body force gx, gy, gz                  */
	int batch_velocity;        /* 1: the velocity solves as one batch    */
} ssx_ns_params;
#define SSX_NS_U 0
#define SSX_NS_V 1
#define SSX_NS_W 2
#define SSX_NS_P 3
int ssx_ns_create(ssx_ns **ns, ssx_ctx *ctx, const ssx_ns_params *prm);
void ssx_ns_destroy(ssx_ns *ns);
/* initialise_from_function (poisson.c:245-282) from arrays: vel[l][c] the
 * velocity component c (device E-vector) at t - (2-l) dt, l = 0, 1, 2;
 * the convection and double-curl histories of the two older levels are
 * formed as there, the pressure history is zero */
int ssx_ns_init(ssx_ns *ns, const double *const vel[3][3]);
/* one time step; stats (nullable): [0] pressure FCG, [1..3] velocity CG */
int ssx_ns_step(ssx_ns *ns, ssx_stats stats[4]);
/* device E-vector of a field (SSX_NS_U..P) at history level 0 (newest),
 * 1 or 2; valid until the next ssx_ns_step */
double *ssx_ns_field(ssx_ns *ns, int which, int level);
/* copy that field into the device E-vector dst (ordered on the context's
 * stream) */
int ssx_ns_read(ssx_ns *ns, int which, int level, double *dst);
double ssx_ns_time(const ssx_ns *ns);

/* ---- batches of independent problems ----
 * Jacobi-PCG (conjugate_gradient, solver.c:33-81) on level 0 of `count`
 * contexts at once: problem i applies op[i] for variable var[i] of ctx[i]
 * (any element counts and orders; a context may appear more than once, e.g.
 * for three velocity components).  Each CG stage is one launch for the
 * whole batch (the operator: one per distinct order); with use_graph the
 * iterations replay a captured HIP graph.  The batch owns its work vectors;
 * the contexts must outlive it.  It runs on a stream of its own unless
 * ssx_batch_set_stream names another (graph replay needs a non-null
 * stream).  ssx_batch_cg: f[i], u[i] device E-vectors of problem i (f
 * masked in place, u initial guess in / solution out), stopping test
 * sqrt(<r,r>_w) <= eps[i] or max_iter; stats (nullable) gets `count`
 * entries.  Returns SSX_ENOTCONV if any problem stopped above its eps.
 * The batch stream first waits (by an event) for the work already enqueued
 * on every context's stream, so f and u may still be in flight there; the
 * call returns after the batch stream has finished.  The contexts' Jacobi
 * diagonal choice, gather-form option and operator launch shape are read
 * again at every call. */
typedef struct ssx_batch ssx_batch;
int ssx_batch_create(ssx_batch **batch, ssx_ctx *const *ctx, const int *op,
		const int *var, int count);
void ssx_batch_destroy(ssx_batch *batch);
int ssx_batch_set_stream(ssx_batch *batch, void *hip_stream);
int ssx_batch_cg(ssx_batch *batch, double *const *f, double *const *u,
		const double *eps, int max_iter, int use_graph, ssx_stats *stats);
/* Every problem of the batch with the reference's solver for it, all
 * advancing together: precon[i] = SSX_PC_MG solves problem i with
 * flexible_conjugate_gradient (solver.c:83-157) preconditioned by the
 * p-multigrid V-cycle (precon.c:227-271) -- the pressure solve of
 * navier_stokes.c:157, or an ensemble member; SSX_PC_JACOBI with
 * conjugate_gradient (solver.c:33-81) -- the velocity solves of
 * navier_stokes.c:219.  The multigrid problems take one FCG iteration at a
 * time in lock-step; in each V-cycle the smoother CG of every level (15
 * iterations, early exit at eps) and the coarsest-level CG (500, to eps)
 * of all problems run as batched passes, one launch per CG stage for the
 * whole batch (the operator: one per order, operator and precision), and
 * the Jacobi problems iterate along in every pass until they converge.
 * A multigrid problem uses its context's level vectors and FCG history and
 * its smoother precision (SSX_OPT_SMOOTHER_PRECISION, with the coarsest
 * level in FP64 stopping at max(eps, 10^-k |f|_w) as in ssx_fcg), so a
 * context may carry at most one multigrid problem of a batch (any number
 * of Jacobi ones) and must not be used by other calls meanwhile.  f, u,
 * eps, max_iter and the return codes are those of ssx_batch_cg (max_iter
 * bounds the FCG iterations of a multigrid problem and the CG iterations
 * of a Jacobi one); stats get the FCG figures of ssx_fcg.  With no
 * multigrid problem this is ssx_batch_cg.  The multigrid passes are
 * launched eagerly (use_graph applies to an all-Jacobi batch). */
int ssx_batch_solve(ssx_batch *batch, const int *precon, double *const *f,
		double *const *u, const double *eps, int max_iter, int use_graph,
		ssx_stats *stats);

/* ---- operator launch autotuning ----
 * The element operator has SSX_AX_NVARIANTS launch shapes (elements per
 * workgroup; non-temporal streaming of factors and result; derivative rows
 * in registers or LDS).  ssx_create
 * times every shape on every level in FP64 and keeps the fastest; the
 * reduced-precision copies are tuned when first built.  prec_mask: bit p
 * for precision SSX_PREC_p.  Results are bit-identical across shapes.
 * ssx_ax_variant returns the shape in use (us, nullable: the measured
 * microseconds per launch of each shape); ssx_set_ax_variant forces one. */
#define SSX_AX_NVARIANTS 8
int ssx_autotune(ssx_ctx *ctx, int prec_mask);
int ssx_ax_variant(const ssx_ctx *ctx, int level, int prec,
		double us[SSX_AX_NVARIANTS]);
int ssx_set_ax_variant(ssx_ctx *ctx, int level, int prec, int variant);

/* ---- reference entry points on HOST E-vectors: f (masked in place) and u
 * (initial guess in, solution out), FCG + p-multigrid for Poisson and
 * Jacobi PCG for Helmholtz, max_iter 5000 ---- */
int ssx_solve_poisson_cg(ssx_ctx *ctx, double *f, double *u, int var,
		double eps, ssx_stats *stats);
int ssx_solve_helmholtz_cg(ssx_ctx *ctx, double *f, double *u, int var,
		double eps, ssx_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
