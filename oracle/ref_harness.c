/*
 * ref_harness.c -- drives the compiled reference (oracle/_ref/libssenss_ref.so,
 * built from the SSENSS sources where they lie) to write golden vectors.
 * TEST INFRASTRUCTURE ONLY; runs in the build container, never on the GPU
 * box.  Called by tests/golden/make_golden.py.
 *
 *   ref_harness basis <out.bin>
 *       for n = 2..16: xi, rho, deriv_matrix, svv_deriv_matrix
 *       (new_gll_basis basis.c:31-52, compute_svv_derivative_matrix
 *       deriv.c:81-104 with svvn = 2n/3, svva = 0) and, for every
 *       multigrid pair, the interpolation matrix (interpolation.c:125-134)
 *   ref_harness mesh <dir>
 *       reads <dir>/mesh.bin and <dir>/in_*.bin, sets the element structures
 *       up as main.c:60-79 does (read_grid replaced by filling the same
 *       fields: dataio.c:407-452 with file n = run n), and writes
 *       <dir>/ref_*.bin and <dir>/ref_scalars.txt.
 *   ref_harness time <dir> <seconds>
 *       the reference's own code timed on the element structures of
 *       <dir>/mesh.bin: the Poisson element operator over all elements
 *       (discrete_laplace_operator, poisson.c:34-56) repeated for about
 *       <seconds>, then one Dirichlet Poisson FCG + p-MG solve (Poisson
 *       diagonal, X_VELOCITY mask) of <dir>/in_f.bin to 1e-10 ||b||_w from
 *       a zero guess; prints "ax_seconds <per application> ax_reps <k>
 *       solve_seconds <s> residual <r> eps <eps>".
 *   ref_harness ns <dir>
 *       Navier-Stokes time steps: the element structures as main.c:45-79
 *       with the convective geometry (the grid interpolated to nconv =
 *       3n/2 points, as read_grid does for conv_geom), the velocity history
 *       from <dir>/ns_vel.bin (u,v,w at t0-2dt, t0-dt, t0) set up as
 *       initialise_from_function (poisson.c:245-282) sets it from
 *       functions, then <dir>/ns.bin's step count of main.c:102-107
 *       (source term g, integrate_navier_stokes navier_stokes.c:224-235);
 *       writes <dir>/ref_ns_%d.bin (u, v, w, p) after each step.
 *       ns.bin: double dt, visc, dens, t0, gx, gy, gz; int nsteps.
 *
 * Mesh file (little endian): int n, nel; double dt, visc; int svvn;
 * double svva; int neigh[nel*6], edge_neigh[nel*12], edge_size[nel*12],
 * corn_neigh[nel*8], corn_size[nel*8]; double x,y,z[nel*n^3].
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mpi.h>
#include "utils.h"
#include "structs.h"
#include "basis.h"
#include "tripalg.h"
#include "geometry.h"
#include "deriv.h"
#include "interpolation.h"
#include "init.h"
#include "precon.h"
#include "poisson.h"
#include "helmholtz.h"
#include "global.h"
#include "solver.h"
#include "particles.h"
#include "time.h"
#include "convection.h"
#include "stokes.h"
#include "navier_stokes.h"
#include "test.h"
#include <unistd.h>

static void die(const char *m)
{
	fprintf(stderr, "ref_harness: %s\n", m);
	exit(2);
}

static void *rd(FILE *f, size_t sz, size_t n)
{
	void *p = malloc(sz * (n ? n : 1));
	if (!p || fread(p, sz, n, f) != n)
		die("short read");
	return p;
}

static void wr(const char *dir, const char *name, const double *a, size_t n)
{
	char path[1024];
	snprintf(path, sizeof(path), "%s/%s", dir, name);
	FILE *f = fopen(path, "wb");
	if (!f || fwrite(a, sizeof(double), n, f) != n)
		die("write failed");
	fclose(f);
}

static double *rdvec(const char *dir, const char *name, size_t n)
{
	char path[1024];
	snprintf(path, sizeof(path), "%s/%s", dir, name);
	FILE *f = fopen(path, "rb");
	if (!f)
		die(path);
	double *a = rd(f, sizeof(double), n);
	fclose(f);
	return a;
}

static void do_basis(const char *out)
{
	FILE *f = fopen(out, "wb");
	if (!f)
		die("open basis out");
	for (int n = 2; n <= 16; ++n) {
		struct basis *b = new_gll_basis(n);
		compute_svv_derivative_matrix(b, 2 * n / 3, 0.0);
		fwrite(b->points, sizeof(double), n, f);
		fwrite(b->weights, sizeof(double), n, f);
		fwrite(*b->deriv_matrix, sizeof(double), n * n, f);
		fwrite(*b->svv_deriv_matrix, sizeof(double), n * n, f);
		free_basis(b);
	}
	/* multigrid pairs nf -> nc = (nf-1)/2 + 1 (>= 2) */
	for (int nf = 3; nf <= 16; ++nf) {
		int nc = (nf - 1) / 2 + 1;
		if (nc < 2)
			nc = 2;
		struct basis *bf = new_gll_basis(nf), *bc = new_gll_basis(nc);
		create_interpolation_matrices(bf, bc);
		fwrite(*bf->interp_matrix, sizeof(double), nf * nc, f); /* coarse->fine */
		fwrite(*bc->interp_matrix, sizeof(double), nf * nc, f); /* fine->coarse */
		free_basis(bf);
		free_basis(bc);
	}
	fclose(f);
}

/* binary search for the iteration count of a deterministic solve: the
 * smallest k for which the solver stops at or below eps */
typedef double (*solver_fn)(struct element **, double ****, double ****, int,
		void (*)(), void (*)(), double, int);

static double ****vec4(int n, int nel)
{
	return new_4d_array(n, n, n, nel);
}

static void set4(double ****a, const double *src, size_t nt)
{
	memcpy(***a, src, nt * sizeof(double));
}

static int count_iters(struct element **sel, solver_fn fn, const double *f,
		const double *u0, int n, int nel, int var, void (*op)(),
		void (*pc)(), double eps, double *err_final, double *u_out)
{
	const size_t nt = (size_t) nel * n * n * n;
	double ****F = vec4(n, nel), ****U = vec4(n, nel);
	set4(F, f, nt);
	set4(U, u0, nt);
	double err = fn(sel, F, U, var, op, pc, eps, 5000);
	*err_final = err;
	memcpy(u_out, ***U, nt * sizeof(double));
	if (!(err <= eps)) {
		free_4d_array(F);
		free_4d_array(U);
		return -1;
	}
	int lo = 1, hi = 5000;
	while (lo < hi) {
		int mid = (lo + hi) / 2;
		set4(F, f, nt);
		set4(U, u0, nt);
		double e = fn(sel, F, U, var, op, pc, eps, mid);
		if (e <= eps)
			hi = mid;
		else
			lo = mid + 1;
	}
	free_4d_array(F);
	free_4d_array(U);
	return lo;
}

static void do_mesh(const char *dir)
{
	char path[1024];
	snprintf(path, sizeof(path), "%s/mesh.bin", dir);
	FILE *f = fopen(path, "rb");
	if (!f)
		die("mesh.bin");
	int hdr[2];
	if (fread(hdr, sizeof(int), 2, f) != 2)
		die("mesh header");
	const int n = hdr[0], nel = hdr[1];
	double dtv[2], svva;
	int svvn;
	if (fread(dtv, sizeof(double), 2, f) != 2 || fread(&svvn, sizeof(int), 1, f)
			!= 1 || fread(&svva, sizeof(double), 1, f) != 1)
		die("mesh params");
	int *neigh = rd(f, sizeof(int), (size_t) nel * 6);
	int *eneigh = rd(f, sizeof(int), (size_t) nel * 12);
	int *esize = rd(f, sizeof(int), (size_t) nel * 12);
	int *cneigh = rd(f, sizeof(int), (size_t) nel * 8);
	int *csize = rd(f, sizeof(int), (size_t) nel * 8);
	const int n3 = n * n * n;
	const size_t nt = (size_t) nel * n3;
	double *X = rd(f, sizeof(double), nt), *Y = rd(f, sizeof(double), nt),
		*Z = rd(f, sizeof(double), nt);
	fclose(f);

	/* main.c:45-79 with read_grid's field filling (dataio.c:407-452) */
	struct communicator *comm = new_communicator(0, 1, MPI_COMM_WORLD,
			MPI_INFO_NULL);
	struct parameters *params = new_parameters(0, 1);
	params->n = n;
	params->nel = nel;
	params->neltot = nel;
	params->nconv = (3 * n) / 2;
	params->svvn = svvn;
	params->svva = svva;
	params->dt = dtv[0];
	params->visc = dtv[1];
	struct basis *gll = new_gll_basis(n);
	struct basis *conv = new_gll_basis(params->nconv);
	create_interpolation_matrices(gll, conv);
	compute_svv_derivative_matrix(gll, params->svvn, params->svva);
	struct geometry **geom = new_geometries(gll, nel, nel);
	for (int e = 0; e < nel; ++e) {
		geom[e]->id = e;
		for (int s = 0; s < 6; ++s)
			geom[e]->neigh_rot[s] = 0;
		memcpy(geom[e]->edge_size, esize + e * 12, 12 * sizeof(int));
		memcpy(geom[e]->corner_size, csize + e * 8, 8 * sizeof(int));
		get_neighbour_elnums_procs(geom[e]->neigh_el, geom[e]->neigh_proc,
				neigh + e * 6, nel, 1, 6);
		get_neighbour_elnums_procs(geom[e]->neigh_edge_el,
				geom[e]->neigh_edge_proc, eneigh + e * 12, nel, 1, 12);
		get_neighbour_elnums_procs(geom[e]->neigh_corn_el,
				geom[e]->neigh_corn_proc, cneigh + e * 8, nel, 1, 8);
		set_boundary_conditions(geom[e]);
		memcpy(**geom[e]->grid->x, X + (size_t) e * n3, n3 * sizeof(double));
		memcpy(**geom[e]->grid->y, Y + (size_t) e * n3, n3 * sizeof(double));
		memcpy(**geom[e]->grid->z, Z + (size_t) e * n3, n3 * sizeof(double));
		compute_geometry_transformations(geom[e], gll);
	}
	/* the convective geometry is not on the elliptic path; the element
	 * constructor only stores the pointer */
	struct element **sel = new_elements(geom, geom, gll, conv, params, comm);
	connect_elements(sel);
	initialise_preconditioners(sel);
	const int nl = (*sel)->precon->nlevels;

	/* geometry: 6 G arrays + detJ, level 0 */
	{
		double *g = malloc(sizeof(double) * 7 * nt);
		struct trimat *gm;
		for (int e = 0; e < nel; ++e) {
			gm = geom[e]->geom_mat;
			double ***arr[7] = { gm->x.x, gm->x.y, gm->x.z, gm->y.y, gm->y.z,
				gm->z.z, geom[e]->jac_det };
			for (int a = 0; a < 7; ++a)
				memcpy(g + a * nt + (size_t) e * n3, **arr[a],
						n3 * sizeof(double));
		}
		wr(dir, "ref_geom.bin", g, 7 * nt);
		free(g);
	}
	/* level data */
	for (int l = 0; l < nl; ++l) {
		const int nn = (*sel)->precon->nmg[l];
		const size_t ntl = (size_t) nel * nn * nn * nn;
		double *d = malloc(sizeof(double) * 4 * ntl);
		memcpy(d, **(*sel)->precon->poisson_diag[l], ntl * sizeof(double));
		memcpy(d + ntl, **(*sel)->precon->helmholtz_diag[l],
				ntl * sizeof(double));
		memcpy(d + 2 * ntl, **(*sel)->precon->stiffsum_scale[l],
				ntl * sizeof(double));
		memcpy(d + 3 * ntl, **(*sel)->precon->dirichlet_mask[l],
				ntl * sizeof(double));
		snprintf(path, sizeof(path), "ref_level%d.bin", l);
		wr(dir, path, d, 4 * ntl);
		free(d);
	}
	wr(dir, "ref_invmass.bin", **(*sel)->precon->inverse_mass, nt);

	double *u = rdvec(dir, "in_u.bin", nt);
	double *fr = rdvec(dir, "in_f.bin", nt);
	double ****U = vec4(n, nel), ****AU = vec4(n, nel);
	set4(U, u, nt);
	/* element operators (no mask, no DSS) */
	for (int e = 0; e < nel; ++e)
		discrete_laplace_operator(sel[e], U[e], AU[e]);
	wr(dir, "ref_ax_poisson.bin", ***AU, nt);
	for (int e = 0; e < nel; ++e)
		discrete_helmholtz_operator(sel[e], U[e], AU[e]);
	wr(dir, "ref_ax_helmholtz.bin", ***AU, nt);
	/* assembled operators */
	global_opeval(sel, U, AU, PRESSURE, discrete_laplace_operator);
	wr(dir, "ref_opeval_pressure.bin", ***AU, nt);
	global_opeval(sel, U, AU, X_VELOCITY, discrete_helmholtz_operator);
	wr(dir, "ref_opeval_helm_vel.bin", ***AU, nt);
	/* DSS of a random field and of an integer-valued field */
	set4(AU, u, nt);
	direct_stiffness_summation(sel, AU);
	wr(dir, "ref_dss_u.bin", ***AU, nt);
	for (size_t q = 0; q < nt; ++q)
		(***AU)[q] = (double) ((q * 2654435761u) % 1000003u);
	direct_stiffness_summation(sel, AU);
	wr(dir, "ref_dss_int.bin", ***AU, nt);
	set4(AU, u, nt);
	direct_stiffness_averaging(sel, AU);
	wr(dir, "ref_dss_avg.bin", ***AU, nt);
	/* inner products */
	double ****ONE = vec4(n, nel);
	ones(***ONE, nt);
	double dots[3];
	dots[0] = global_dot_product(sel, U, U);
	dots[1] = global_dot_product(sel, U, ONE);
	dots[2] = global_dot_product(sel, ONE, ONE);
	wr(dir, "ref_dots.bin", dots, 3);
	/* null-space projection and mask */
	set4(AU, u, nt);
	subtract_nullspace(sel, AU, PRESSURE);
	wr(dir, "ref_nullspace.bin", ***AU, nt);
	/* order transfer level 0 <-> 1 per element (interpolation.c:136-176) */
	if (nl > 1) {
		const int nc = (*sel)->precon->nmg[1];
		const size_t ntc = (size_t) nel * nc * nc * nc;
		double ****C1 = vec4(nc, nel);
		for (int e = 0; e < nel; ++e)
			integrate_to_grid(sel[e]->precon->basis[2], C1[e],
					sel[e]->precon->basis[1], U[e]);
		wr(dir, "ref_restrict_raw.bin", ***C1, ntc);
		double *v1 = rdvec(dir, "in_v1.bin", ntc);
		set4(C1, v1, ntc);
		for (int e = 0; e < nel; ++e)
			interpolate_to_grid(sel[e]->precon->basis[1], AU[e],
					sel[e]->precon->basis[2], C1[e]);
		wr(dir, "ref_prolong.bin", ***AU, nt);
		free(v1);
		free_4d_array(C1);
	}
	/* solver runs.  eps = 1e-10 * ||b||_w, b = the given right-hand side */
	double ****FB = vec4(n, nel);
	set4(FB, fr, nt);
	const double bnorm = sqrt(global_dot_product(sel, FB, FB));
	const double eps = 1e-10 * bnorm;
	double *zero = calloc(nt, sizeof(double)), *sol = malloc(sizeof(double) * nt);
	double res[16];
	int nres = 0;
	res[nres++] = eps;
	/* (1) Helmholtz, Jacobi PCG, var X_VELOCITY (solve_helmholtz_cg) */
	{
		double err;
		int it = count_iters(sel, conjugate_gradient, fr, zero, n, nel,
				X_VELOCITY, discrete_helmholtz_operator,
				jacobi_preconditioner, eps, &err, sol);
		res[nres++] = it;
		res[nres++] = err;
		wr(dir, "ref_cg_helm.bin", sol, nt);
	}
	/* (2) pressure Poisson, FCG + p-multigrid (solve_poisson_cg) */
	{
		double err;
		int it = count_iters(sel, flexible_conjugate_gradient, fr, zero, n,
				nel, PRESSURE, discrete_laplace_operator,
				multigrid_preconditioner, eps, &err, sol);
		res[nres++] = it;
		res[nres++] = err;
		wr(dir, "ref_fcg_pressure.bin", sol, nt);
	}
	/* (3) one multigrid preconditioner application, PRESSURE */
	{
		double ****R = vec4(n, nel), ****PR = vec4(n, nel);
		set4(R, fr, nt);
		multigrid_preconditioner(sel, R, PR, PRESSURE, eps,
				discrete_laplace_operator);
		wr(dir, "ref_mg_pressure.bin", ***PR, nt);
		free_4d_array(R);
		free_4d_array(PR);
	}
	/* (4) Dirichlet Poisson with the Poisson diagonal: var X_VELOCITY
	 * masks, helmholtz_diag aliased to poisson_diag at every level */
	{
		for (int e = 0; e < nel; ++e)
			for (int l = 0; l < nl; ++l)
				sel[e]->precon->helmholtz_diag[l] =
					sel[e]->precon->poisson_diag[l];
		double err;
		int it = count_iters(sel, flexible_conjugate_gradient, fr, zero, n,
				nel, X_VELOCITY, discrete_laplace_operator,
				multigrid_preconditioner, eps, &err, sol);
		res[nres++] = it;
		res[nres++] = err;
		wr(dir, "ref_fcg_dirichlet.bin", sol, nt);
		double ****R = vec4(n, nel), ****PR = vec4(n, nel);
		set4(R, fr, nt);
		mask_dirichlet_boundaries(sel, R, X_VELOCITY);
		multigrid_preconditioner(sel, R, PR, X_VELOCITY, eps,
				discrete_laplace_operator);
		wr(dir, "ref_mg_dirichlet.bin", ***PR, nt);
		free_4d_array(R);
		free_4d_array(PR);
	}
	res[nres++] = bnorm;
	wr(dir, "ref_scalars.bin", res, nres);

	/* particles (particles.c:182-218) with injected positions: the
	 * reference seeds them from the wall clock (particles.c:152) */
	snprintf(path, sizeof(path), "%s/in_particles.bin", dir);
	if (access(path, R_OK) == 0) {
		/* in_particles.bin: int np; double t_now, dt, s[3];
		 * then per particle el, x, y, z, u, v, w, t, tau (doubles) */
		FILE *pf = fopen(path, "rb");
		int np;
		double hdr2[5];
		if (fread(&np, sizeof(int), 1, pf) != 1
				|| fread(hdr2, sizeof(double), 5, pf) != 5)
			die("particles header");
		double *rec = rd(pf, sizeof(double), (size_t) np * 9);
		fclose(pf);
		double *vel = rdvec(dir, "in_vel.bin", 3 * nt);
		/* three velocity levels, oldest first, each shifted into the
		 * history by compute_reference_domain_velocity (main.c:84) */
		for (int lev = 0; lev < 3; ++lev) {
			for (int e = 0; e < nel; ++e)
				for (int a = 0; a < 3; ++a) {
					double *dst = ***sel[e]->u->ind[a];
					const double *src = vel + a * nt + (size_t) e * n3;
					for (int q = 0; q < n3; ++q)
						dst[q] = hdr2[2 + lev] * src[q];
				}
			compute_reference_domain_velocity(sel);
		}
		params->t = hdr2[0];
		params->dt = hdr2[1];
		struct plist pl;
		pl.n = np;
		pl.ntot = np;
		pl.nbuff = np;
		pl.ntypes = 1;
		pl.tau = NULL;
		pl.nproc = 1;
		pl.mpi_comm = MPI_COMM_WORLD;
		pl.list = calloc(np, sizeof(struct particle));
		for (int i = 0; i < np; ++i) {
			const double *r = rec + (size_t) i * 9;
			struct particle *q = pl.list + i;
			q->id = i;
			q->type = 0;
			q->el = (int) r[0];
			q->old_el = q->el;
			q->x = r[1]; q->y = r[2]; q->z = r[3];
			q->u = r[4]; q->v = r[5]; q->w = r[6];
			q->t = r[7];
			q->tau = r[8];
			q->proc = 0;
			q->old_proc = 0;
			q->status = 0;
		}
		integrate_particle_equations(&pl, sel);
		/* relocate_particles reorders the list: write back by id */
		double *out = calloc((size_t) np * 8, sizeof(double));
		for (int i = 0; i < np; ++i) {
			const struct particle *q = pl.list + i;
			double *o = out + (size_t) q->id * 8;
			o[0] = q->el;
			o[1] = q->x; o[2] = q->y; o[3] = q->z;
			o[4] = q->u; o[5] = q->v; o[6] = q->w;
			o[7] = q->t;
		}
		wr(dir, "ref_particles.bin", out, (size_t) np * 8);
		free(out);
		free(pl.list);
		free(rec);
		free(vel);
	}
	free(zero);
	free(sol);
	free(u);
	free(fr);
}

static double wall(void)
{
	return MPI_Wtime();
}

static struct element **ns_setup(const char *dir, const double *prm,
		int *n_out, int *nel_out);

static void do_time(const char *dir, double seconds)
{
	const double prm[7] = { 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.0 };
	int n, nel;
	struct element **sel = ns_setup(dir, prm, &n, &nel);
	const size_t nt = (size_t) nel * n * n * n;
	double *fr = rdvec(dir, "in_f.bin", nt);
	double ****F = vec4(n, nel), ****U = vec4(n, nel), ****AU = vec4(n, nel);
	set4(U, fr, nt);
	for (int e = 0; e < nel; ++e)           /* warm-up */
		discrete_laplace_operator(sel[e], U[e], AU[e]);
	int reps = 0;
	double t0 = wall(), t1;
	do {
		for (int e = 0; e < nel; ++e)
			discrete_laplace_operator(sel[e], U[e], AU[e]);
		++reps;
		t1 = wall();
	} while (t1 - t0 < seconds);
	const double ax = (t1 - t0) / reps;
	set4(F, fr, nt);
	memset(***U, 0, nt * sizeof(double));
	const double eps = 1e-10 * sqrt(global_dot_product(sel, F, F));
	const int nl = (*sel)->precon->nlevels;
	for (int e = 0; e < nel; ++e)
		for (int l = 0; l < nl; ++l)
			sel[e]->precon->helmholtz_diag[l] =
				sel[e]->precon->poisson_diag[l];
	t0 = wall();
	const double err = flexible_conjugate_gradient(sel, F, U, X_VELOCITY,
			discrete_laplace_operator, multigrid_preconditioner, eps, 5000);
	t1 = wall();
	printf("ax_seconds %.6e ax_reps %d solve_seconds %.6f residual %.6e "
			"eps %.6e\n", ax, reps, t1 - t0, err, eps);
	free(fr);
}

/* the mesh file and both geometries (main.c:45-79) */
static struct element **ns_setup(const char *dir, const double *prm,
		int *n_out, int *nel_out)
{
	char path[1024];
	snprintf(path, sizeof(path), "%s/mesh.bin", dir);
	FILE *f = fopen(path, "rb");
	if (!f)
		die("mesh.bin");
	int hdr[2];
	if (fread(hdr, sizeof(int), 2, f) != 2)
		die("mesh header");
	const int n = hdr[0], nel = hdr[1];
	double dtv[2], svva;
	int svvn;
	if (fread(dtv, sizeof(double), 2, f) != 2 || fread(&svvn, sizeof(int), 1, f)
			!= 1 || fread(&svva, sizeof(double), 1, f) != 1)
		die("mesh params");
	int *neigh = rd(f, sizeof(int), (size_t) nel * 6);
	int *eneigh = rd(f, sizeof(int), (size_t) nel * 12);
	int *esize = rd(f, sizeof(int), (size_t) nel * 12);
	int *cneigh = rd(f, sizeof(int), (size_t) nel * 8);
	int *csize = rd(f, sizeof(int), (size_t) nel * 8);
	const int n3 = n * n * n;
	const size_t nt = (size_t) nel * n3;
	double *X = rd(f, sizeof(double), nt), *Y = rd(f, sizeof(double), nt),
		*Z = rd(f, sizeof(double), nt);
	fclose(f);
	struct communicator *comm = new_communicator(0, 1, MPI_COMM_WORLD,
			MPI_INFO_NULL);
	struct parameters *params = new_parameters(0, 1);
	params->n = n;
	params->nel = nel;
	params->neltot = nel;
	params->nconv = (3 * n) / 2;                 /* main.c:53 */
	params->svvn = svvn;
	params->svva = svva;
	params->dt = prm[0];
	params->visc = prm[1];
	params->dens = prm[2];
	params->t = prm[3];
	params->gx = prm[4];
	params->gy = prm[5];
	params->gz = prm[6];
	struct basis *gll = new_gll_basis(n);
	struct basis *conv = new_gll_basis(params->nconv);
	create_interpolation_matrices(gll, conv);
	compute_svv_derivative_matrix(gll, params->svvn, params->svva);
	struct geometry **geom = new_geometries(gll, nel, nel);
	struct geometry **cgeom = new_geometries(conv, nel, nel);
	double ***tmp = new_3d_array(n, n, n);
	for (int e = 0; e < nel; ++e) {
		struct geometry *gs[2] = { geom[e], cgeom[e] };
		for (int k = 0; k < 2; ++k) {
			struct geometry *g = gs[k];
			g->id = e;
			for (int s = 0; s < 6; ++s)
				g->neigh_rot[s] = 0;
			memcpy(g->edge_size, esize + e * 12, 12 * sizeof(int));
			memcpy(g->corner_size, csize + e * 8, 8 * sizeof(int));
			get_neighbour_elnums_procs(g->neigh_el, g->neigh_proc,
					neigh + e * 6, nel, 1, 6);
			get_neighbour_elnums_procs(g->neigh_edge_el,
					g->neigh_edge_proc, eneigh + e * 12, nel, 1, 12);
			get_neighbour_elnums_procs(g->neigh_corn_el,
					g->neigh_corn_proc, cneigh + e * 8, nel, 1, 8);
			set_boundary_conditions(g);
		}
		memcpy(**geom[e]->grid->x, X + (size_t) e * n3, n3 * sizeof(double));
		memcpy(**geom[e]->grid->y, Y + (size_t) e * n3, n3 * sizeof(double));
		memcpy(**geom[e]->grid->z, Z + (size_t) e * n3, n3 * sizeof(double));
		compute_geometry_transformations(geom[e], gll);
		/* dataio.c:433-441: the file's grid interpolated to the basis */
		const double *src[3] = { X, Y, Z };
		double ***dst[3] = { cgeom[e]->grid->x, cgeom[e]->grid->y,
			cgeom[e]->grid->z };
		for (int a = 0; a < 3; ++a) {
			memcpy(**tmp, src[a] + (size_t) e * n3, n3 * sizeof(double));
			interpolate_to_grid(conv, dst[a], gll, tmp);
		}
		compute_geometry_transformations(cgeom[e], conv);
	}
	free_3d_array(tmp);
	struct element **sel = new_elements(geom, cgeom, gll, conv, params, comm);
	connect_elements(sel);
	initialise_preconditioners(sel);
	free(neigh); free(eneigh); free(esize); free(cneigh); free(csize);
	free(X); free(Y); free(Z);
	*n_out = n;
	*nel_out = nel;
	return sel;
}

static void do_ns(const char *dir)
{
	char path[1024];
	snprintf(path, sizeof(path), "%s/ns.bin", dir);
	FILE *f = fopen(path, "rb");
	if (!f)
		die("ns.bin");
	double prm[7];
	int nsteps;
	if (fread(prm, sizeof(double), 7, f) != 7
			|| fread(&nsteps, sizeof(int), 1, f) != 1)
		die("ns params");
	fclose(f);
	int n, nel;
	struct element **sel = ns_setup(dir, prm, &n, &nel);
	const int n3 = n * n * n;
	const size_t nt = (size_t) nel * n3;
	double *vel = rdvec(dir, "ns_vel.bin", 9 * nt);
	/* initialise_from_function (poisson.c:245-282) with the levels
	 * given as arrays instead of evaluated functions */
	for (int j = 0; j < 3; ++j) {
		for (int e = 0; e < nel; ++e) {
			time_shift(sel[e]->u, sel[e]->old->u, sel[e]->old->old->u,
					sel[e]->ntot);
			for (int a = 0; a < 3; ++a)
				memcpy(***sel[e]->u->ind[a], vel + (3 * j + a) * nt
						+ (size_t) e * n3, n3 * sizeof(double));
			if (j >= 2)
				continue;
			time_shift(sel[e]->conv, sel[e]->old->conv,
					sel[e]->old->old->conv, sel[e]->ntot);
			time_shift(sel[e]->rot, sel[e]->old->rot,
					sel[e]->old->old->rot, sel[e]->ntot);
			convection_operator(sel[e], sel[e]->u, sel[e]->u, sel[e]->conv);
			double_curl_operator(sel[e], sel[e]->u, sel[e]->rot);
		}
	}
	struct parameters *params = (*sel)->params;
	double *out = malloc(sizeof(double) * 4 * nt);
	for (int step = 0; step < nsteps; ++step) {
		/* main.c:104-107 */
		set_source_term(sel, X_SOURCE, one_func, params->gx);
		set_source_term(sel, Y_SOURCE, one_func, params->gy);
		set_source_term(sel, Z_SOURCE, one_func, params->gz);
		integrate_navier_stokes(sel);
		for (int e = 0; e < nel; ++e) {
			for (int a = 0; a < 3; ++a)
				memcpy(out + a * nt + (size_t) e * n3,
						***sel[e]->u->ind[a], n3 * sizeof(double));
			memcpy(out + 3 * nt + (size_t) e * n3, **sel[e]->p,
					n3 * sizeof(double));
		}
		snprintf(path, sizeof(path), "ref_ns_%d.bin", step + 1);
		wr(dir, path, out, 4 * nt);
	}
	free(out);
	free(vel);
}

int main(int argc, char **argv)
{
	MPI_Init(&argc, &argv);
	if (argc < 3)
		die("usage: ref_harness basis <out> | mesh <dir> | ns <dir> | "
				"time <dir> <seconds>");
	if (!strcmp(argv[1], "basis"))
		do_basis(argv[2]);
	else if (!strcmp(argv[1], "mesh"))
		do_mesh(argv[2]);
	else if (!strcmp(argv[1], "ns"))
		do_ns(argv[2]);
	else if (!strcmp(argv[1], "time") && argc > 3)
		do_time(argv[2], atof(argv[3]));
	else
		die("unknown task");
	MPI_Finalize();
	return 0;
}
