/*
 * ssx_shim.c -- SSENSS-side binding of the ssx engine (single MPI rank).
 *
 * Compiled into the reference build (src/), against its own headers:
 *   solve_poisson_cg    replaces solver.c:159-173 (FCG + p-MG on the GPU)
 *   solve_helmholtz_cg  replaces solver.c:175-189 (Jacobi-PCG on the GPU)
 *   ssx_shim_init       called after initialise_preconditioners (main.c:79):
 *                       uploads the mesh (coordinates, connectivity) once
 * Build: gcc -std=c99 -c -I<ssenss>/src -I<repo>/include ssx_shim.c, link
 * with -L<repo>/ssenss_amd -lssx.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "structs.h"
#include "tripalg.h"
#include "geometry.h"
#include "basis.h"
#include "poisson.h"
#include "solver.h"
#include "ssx.h"

ssx_ctx *g_ctx;          /* shared with ssx_particles_shim.c */
static double *g_u;

/* after initialise_preconditioners(sel), main.c:79 */
void ssx_shim_init(struct element **sel)
{
	const int nel = (*sel)->nel, n = (*sel)->basis->n;
	const size_t n3 = (size_t) n * n * n, nt = n3 * nel;
	double *x = malloc(nt * sizeof(double)), *y = malloc(nt * sizeof(double)),
		*z = malloc(nt * sizeof(double));
	int *nb = malloc(sizeof(int) * 6 * nel), *en = malloc(sizeof(int) * 12 * nel),
		*es = malloc(sizeof(int) * 12 * nel), *cn = malloc(sizeof(int) * 8 * nel),
		*cs = malloc(sizeof(int) * 8 * nel);
	for (int e = 0; e < nel; ++e) {
		struct geometry *g = sel[e]->geom;
		memcpy(x + e * n3, **g->grid->x, n3 * sizeof(double));
		memcpy(y + e * n3, **g->grid->y, n3 * sizeof(double));
		memcpy(z + e * n3, **g->grid->z, n3 * sizeof(double));
		memcpy(nb + 6 * e, g->neigh_el, 6 * sizeof(int));
		memcpy(en + 12 * e, g->neigh_edge_el, 12 * sizeof(int));
		memcpy(es + 12 * e, g->edge_size, 12 * sizeof(int));
		memcpy(cn + 8 * e, g->neigh_corn_el, 8 * sizeof(int));
		memcpy(cs + 8 * e, g->corner_size, 8 * sizeof(int));
	}
	const struct parameters *p = (*sel)->params;
	ssx_mesh m = { n, nel, x, y, z, nb, en, es, cn, cs, p->dt, p->visc,
		p->svvn, p->svva };
	if (ssx_create(&g_ctx, &m) != SSX_OK) {
		fprintf(stderr, "ssx_create: %s\n", ssx_last_error());
		exit(1);
	}
	g_u = malloc(nt * sizeof(double));
	free(x); free(y); free(z); free(nb); free(en); free(es); free(cn); free(cs);
}

static void gather(struct element **sel, int var, double *u)
{
	const size_t n3 = (*sel)->basis->ntot;
	for (int e = 0; e < (*sel)->nel; ++e)
		memcpy(u + e * n3, **variable_array(sel[e], var), n3 * sizeof(double));
}

static void scatter(struct element **sel, int var, const double *u)
{
	const size_t n3 = (*sel)->basis->ntot;
	for (int e = 0; e < (*sel)->nel; ++e)
		memcpy(**variable_array(sel[e], var), u + e * n3, n3 * sizeof(double));
}

/* solver.c:159-173 */
void solve_poisson_cg(struct element **sel, double ****f, int var, double eps)
{
	ssx_stats st;
	gather(sel, var, g_u);
	int rc = ssx_solve_poisson_cg(g_ctx, ***f, g_u, var, eps, &st);
	if (rc != SSX_OK && rc != SSX_ENOTCONV)
		fprintf(stderr, "ssx_solve_poisson_cg: %s\n", ssx_last_error());
	scatter(sel, var, g_u);
}

/* solver.c:175-189 */
void solve_helmholtz_cg(struct element **sel, double ****f, int var, double eps)
{
	ssx_stats st;
	gather(sel, var, g_u);
	int rc = ssx_solve_helmholtz_cg(g_ctx, ***f, g_u, var, eps, &st);
	if (rc != SSX_OK && rc != SSX_ENOTCONV)
		fprintf(stderr, "ssx_solve_helmholtz_cg: %s\n", ssx_last_error());
	scatter(sel, var, g_u);
}
