/*
 * ssx_oracle.h -- CPU restatement of the SSENSS elliptic-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU path is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py, and by nothing else.  The product library
 * (ssenss_amd/csrc) never links or calls it.
 *
 * It restates, in plain C99 over flat arrays, the algorithms of the reference
 * files below (citations are file:line in the SSENSS source tree, src/):
 *
 *   GLL basis            basis.c:31-52, 75-188; SVV matrix deriv.c:81-112
 *   geometry             geometry.c:102-135, deriv.c:68-79, tripalg.c:165-234
 *   Poisson Ax           poisson.c:34-66, deriv.c:29-66, tripalg.c:224-234
 *   Helmholtz Ax         helmholtz.c:33-49
 *   DSS                  global.c:38-177, 241-257; wiring init.c:68-130, 298-320
 *   weighted dot, avg    global.c:259-282
 *   mask / null space    poisson.c:89-143, solver.c:200-213
 *   PCG / FCG            solver.c:33-157, 191-198
 *   p-multigrid          precon.c:58-83, 121-183, 227-324
 *   Jacobi diagonal      helmholtz.c:51-77, precon.c:336-359
 *   scaling/mask/mass    precon.c:361-413
 *   order transfer       interpolation.c:31-39, 106-176
 *
 * Parity status of each function is recorded beside its definition in
 * ssx_oracle.c ("pinned by": golden vectors from the compiled reference,
 * closed forms, or invariants).
 */
#ifndef SSX_ORACLE_H
#define SSX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NMAX 16     /* largest points-per-direction supported */
#define ORC_LMAX 10     /* multigrid levels, structs.c:58 (new_preconditioner(n, 10)) */

/* var codes, poisson.h:24-27 */
enum { ORC_PRESSURE = 0, ORC_X_VELOCITY = 1, ORC_Y_VELOCITY = 2, ORC_Z_VELOCITY = 3 };
/* operator codes */
enum { ORC_OP_POISSON = 0, ORC_OP_HELMHOLTZ = 1 };
/* preconditioner codes */
enum { ORC_PC_JACOBI = 0, ORC_PC_MG = 1 };

typedef struct orc_ctx orc_ctx;

/* Mesh as the reference holds it after read_grid (dataio.c:197-227) with
 * file n equal to run n: per-element GLL coordinates (i fastest, then j, k,
 * element) and the connectivity fields of grid.h5 (neighbour ids >= 0, or a
 * negative boundary code: -1 no-slip/Dirichlet, -2 free-slip). */
typedef struct orc_mesh {
	int n;                  /* GLL points per direction */
	int nel;
	const double *x, *y, *z;        /* nel*n^3 each */
	const int *neigh;               /* nel*6  (W,E,S,N,B,T) */
	const int *edge_neigh;          /* nel*12 */
	const int *edge_size;           /* nel*12 */
	const int *corn_neigh;          /* nel*8  */
	const int *corn_size;           /* nel*8  */
	double dt, visc;                /* lambda = (11/6)/(dt*visc), helmholtz.c:38 */
	int svvn;                       /* main.c:54: 2n/3 */
	double svva;                    /* main.c:55: 0 */
} orc_mesh;

/* --- basis (stand-alone, for checking against golden vectors) --- */
/* points xi[n], weights rho[n], d[n*n] with d[j*n+i] = l_j'(xi_i),
 * dsvv[n*n] same layout; any output pointer may be NULL. */
void orc_gll_basis(int n, int svvn, double svva, double *xi, double *rho,
		double *d, double *dsvv);
/* J[p*nf+i] = l_p^{coarse}(xi_i^{fine}), interpolation.c:115-123 */
void orc_interp_matrix(int nfrom, int nto, double *j);

/* --- context --- */
orc_ctx *orc_create(const orc_mesh *mesh);
void orc_destroy(orc_ctx *c);
int orc_nlevels(const orc_ctx *c);
int orc_level_n(const orc_ctx *c, int level);
int orc_nel(const orc_ctx *c);
double orc_lambda(const orc_ctx *c);

/* Level data, each nel*n_l^3 unless stated: which = 0..5 G (xx,xy,xz,yy,yz,zz),
 * 6 detJ, 7 poisson diag^-1, 8 helmholtz diag^-1, 9 stiffsum scale (1/mult),
 * 10 dirichlet mask, 11 inverse mass (level 0 only), 12..14 x,y,z coords,
 * 15 reference derivative matrix d (n*n), 16 svv matrix (n*n),
 * 17 prolongation matrix into this level from level+1 (n_l*n_{l+1}, layout
 * [p*n_l+i]).  Returns NULL if absent. */
const double *orc_level_data(const orc_ctx *c, int level, int which);

/* element operators on an E-vector at `level`, no mask, no DSS */
void orc_ax(orc_ctx *c, int level, int op, const double *u, double *au);
/* global_opeval: element operator, mask (var != PRESSURE), DSS */
void orc_opeval(orc_ctx *c, int level, int op, int var, const double *p,
		double *ap);
void orc_dss(orc_ctx *c, int level, double *u);
void orc_dss_average(orc_ctx *c, int level, double *u);
double orc_dot(orc_ctx *c, int level, const double *x, const double *y);
void orc_mask(orc_ctx *c, int level, int var, double *u);
void orc_subtract_nullspace(orc_ctx *c, int level, int var, double *u);
/* restriction level-1 -> level (J^T then DSS-average), prolongation
 * level+1 -> level (J), interpolation.c:136-176, precon.c:283-300 */
void orc_restrict(orc_ctx *c, int level, const double *fine, double *coarse);
void orc_prolong(orc_ctx *c, int level, const double *coarse, double *fine);

/* Jacobi-preconditioned CG at `level` (solver.c:33-81). f is masked in
 * place, u is the initial guess and the result. Returns the final weighted
 * residual norm; *iters receives the number of loop bodies executed. */
double orc_cg(orc_ctx *c, int level, int op, int var, double *f, double *u,
		double eps, int max_iter, int *iters);
/* flexible CG at level 0 (solver.c:83-157); precon ORC_PC_MG selects the
 * multigrid V-cycle with mmax = 10, ORC_PC_JACOBI mmax = 1. */
double orc_fcg(orc_ctx *c, int op, int var, int precon, double *f, double *u,
		double eps, int max_iter, int *iters);
/* p-multigrid preconditioner application (precon.c:227-236) */
void orc_mg_precon(orc_ctx *c, int op, int var, double eps, const double *x,
		double *px);
/* Diagonal used by the Jacobi preconditioner for `var` (precon.c:211-225):
 * 0 = poisson diag, 1 = helmholtz diag.  Default: var != PRESSURE -> 1.
 * Setting -1 restores the default. */
void orc_set_diag_choice(orc_ctx *c, int choice);
/* Cumulative counts since creation: [0] fine-level opeval calls,
 * [1] coarse CG iterations, [2] smoother CG iterations, [3] V-cycles. */
void orc_counters(const orc_ctx *c, long long *out4);
/* --- particles (particles.c) --- */
/* particle state as struct particle (particles.h:25-41), single rank */
typedef struct orc_particle {
	long id;
	int el, old_el, status;
	double x, y, z;         /* reference coordinates in element el */
	double u, v, w;         /* velocity in reference coordinates   */
	double t, tau;
	int reinit;             /* the reference would re-seed it randomly */
} orc_particle;
/* Point interpolation of an element field f (n^3, [k][j][i]) at reference
 * coordinates (interpolation.c:41-104): full tensor Lagrange for n <= 5,
 * else the 5-point stencil around the point. */
double orc_interp_point(int n, const double *xi, const double *f, double x,
		double y, double z);
/* compute_reference_domain_velocity (geometry.c:188-198): shifts the
 * reference-space velocity history (time.c:30-41) and sets
 * ur = J^-1 (u,v,w) from physical E-vectors at level 0. */
void orc_particles_set_velocity(orc_ctx *c, const double *u, const double *v,
		const double *w);
/* integrate_particle_equations (particles.c:182-218) to time t_now with
 * time step dt (params->t, params->dt); returns the number of passes. */
int orc_particles_step(orc_ctx *c, orc_particle *p, long n, double t_now,
		double dt);
/* RK2 tracers (dx/dt = u_f(x, t), no inertia): the tracer mode of the
 * particle walk, an extension with no reference counterpart (pinned by
 * closed forms: uniform flow, linear fields).  Each tracer goes from its t
 * to t_now by Heun's method in the reference coordinates of its element.
 * A stage that leaves [-1,1]^3 is replaced by a straight move to the
 * boundary along the slope that left (k1 if the predictor left, else
 * (k1 + k2)/2), a hop to the neighbour across the closest face with the
 * crossed coordinate set to the opposite face, and a new step over the
 * remaining time.  At a wall the tracer stays on the wall for the rest of
 * the step.  More than 10 passes: restart at the element centre, reinit
 * flag set.  u, v, w receive the fluid velocity (reference space) at the
 * end point; tau is not used.  Returns the largest pass count. */
int orc_tracers_step(orc_ctx *c, orc_particle *p, long n, double t_now,
		double dt);
/* physical coordinates of a particle (particles.c:539-544) */
void orc_particle_position(orc_ctx *c, const orc_particle *p, double *xyz);

/* OpenMP threads used by the element loops (1 = serial) */
void orc_set_threads(int nthreads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
