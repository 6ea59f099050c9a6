/*
 * ssx_particles_shim.c -- SSENSS-side binding of the ssx particle path
 * (single MPI rank; uses g_ctx from ssx_shim.c).
 *
 *   compute_reference_domain_velocity  wraps geometry.c:188-198 (renamed
 *                       there to ref_compute_reference_domain_velocity) and
 *                       uploads the reference-space velocity
 *   integrate_particle_equations  replaces particles.c:182-218 (the walk on
 *                       the GPU; re-seeding by the reference's own
 *                       initialise_particle_position)
 * Build as ssx_shim.c, with -I/opt/conda/include (or wherever mpi.h is).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "structs.h"
#include "tripalg.h"
#include "geometry.h"
#include "basis.h"
#include "particles.h"
#include "ssx.h"

extern ssx_ctx *g_ctx;
void ref_compute_reference_domain_velocity(struct element **sel);

static void *grow(void *p, size_t *cap, size_t need)
{
	if (need <= *cap)
		return p;
	free(p);
	*cap = need;
	p = malloc(need);
	if (!p) {
		fprintf(stderr, "ssx particles shim: out of memory\n");
		exit(1);
	}
	return p;
}

/* geometry.c:188-198: the original body (renamed) keeps the host copy of
 * ur that initialise_particle_position reads; the device copy follows */
void compute_reference_domain_velocity(struct element **sel)
{
	static double *buf;
	static size_t cap;
	const int nel = (*sel)->nel;
	const size_t n3 = (*sel)->basis->ntot, nt = n3 * nel;
	ref_compute_reference_domain_velocity(sel);
	buf = grow(buf, &cap, 3 * nt * sizeof(double));
	for (int e = 0; e < nel; ++e) {
		memcpy(buf + e * n3, **sel[e]->u->x, n3 * sizeof(double));
		memcpy(buf + nt + e * n3, **sel[e]->u->y, n3 * sizeof(double));
		memcpy(buf + 2 * nt + e * n3, **sel[e]->u->z, n3 * sizeof(double));
	}
	if (ssx_particles_set_velocity_host(g_ctx, buf, buf + nt, buf + 2 * nt)
			!= SSX_OK) {
		fprintf(stderr, "ssx_particles_set_velocity_host: %s\n",
				ssx_last_error());
		exit(1);
	}
}

/* particles.c:182-218 (one rank: no relocation between ranks) */
void integrate_particle_equations(struct plist *p, struct element **sel)
{
	static double *d;
	static int *iv;
	static size_t dcap, icap;
	const size_t n = p->n;
	const struct parameters *prm = (*sel)->params;
	d = grow(d, &dcap, (n ? n : 1) * 8 * sizeof(double));
	iv = grow(iv, &icap, (n ? n : 1) * 3 * sizeof(int));
	double *x = d, *y = d + n, *z = d + 2 * n, *u = d + 3 * n,
		*v = d + 4 * n, *w = d + 5 * n, *t = d + 6 * n, *tau = d + 7 * n;
	int *el = iv, *old_el = iv + n, *re = iv + 2 * n;
	for (size_t i = 0; i < n; ++i) {
		const struct particle *q = &p->list[i];
		el[i] = q->el; old_el[i] = q->old_el;
		x[i] = q->x; y[i] = q->y; z[i] = q->z;
		u[i] = q->u; v[i] = q->v; w[i] = q->w;
		t[i] = q->t; tau[i] = q->tau;
	}
	ssx_particles ps = { (long) n, el, old_el, x, y, z, u, v, w, t, tau, re };
	if (ssx_particles_step_host(g_ctx, &ps, prm->t, prm->dt, NULL)
			!= SSX_OK) {
		fprintf(stderr, "ssx_particles_step_host: %s\n", ssx_last_error());
		exit(1);
	}
	for (size_t i = 0; i < n; ++i) {
		struct particle *q = &p->list[i];
		q->el = el[i]; q->old_el = old_el[i];
		q->x = x[i]; q->y = y[i]; q->z = z[i];
		q->u = u[i]; q->v = v[i]; q->w = w[i];
		q->t = t[i];
		if (re[i]) {            /* particles.c:198-201: random re-seed */
			initialise_particle_position(q, sel[q->el]);
			q->t = prm->t;
		}
	}
}
