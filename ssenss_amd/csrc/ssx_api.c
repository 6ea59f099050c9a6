/*
 * ssx_api.c -- C host side of the ssx engine: context set-up, upload, and
 * the solver control flow (PCG, flexible CG, p-multigrid V-cycle).  All
 * vector work is enqueued as gfx950 kernels on the context stream; scalars
 * (alpha, beta, inner products) stay in device memory, and the host reads
 * one residual per iteration to decide termination.
 *
 * The algorithms follow the reference exactly (citations: SSENSS src/):
 *   conjugate_gradient           solver.c:33-81
 *   flexible_conjugate_gradient  solver.c:83-157
 *   global_opeval                solver.c:191-198
 *   subtract_nullspace           solver.c:200-213
 *   multigrid V-cycle            precon.c:227-309
 *   set-up of levels/diagonals   precon.c:58-83, 121-209, 336-413,
 *                                helmholtz.c:51-77, geometry.c:102-135
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <hip/hip_runtime_api.h>
#include "ssx_internal.h"

#define HIPC(call)                                                      \
	do {                                                            \
		hipError_t e_ = (call);                                 \
		if (e_ != hipSuccess) {                                 \
			ssx_set_error("%s: %s", #call, hipGetErrorString(e_)); \
			return SSX_EHIP;                                \
		}                                                       \
	} while (0)
#define TRY(call)                                                       \
	do {                                                            \
		int rc_ = (call);                                       \
		if (rc_ != SSX_OK)                                      \
			return rc_;                                     \
	} while (0)

const char *ssx_version(void)
{
	return "ssx 0.1 (gfx950)";
}

static double now_s(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

static int dalloc(double **p, size_t n)
{
	if (hipMalloc((void **) p, (n ? n : 1) * sizeof(double)) != hipSuccess) {
		*p = NULL;
		ssx_set_error("hipMalloc of %zu doubles failed", n);
		return SSX_ENOMEM;
	}
	return SSX_OK;
}

static int up(double **d, const double *h, size_t n)
{
	TRY(dalloc(d, n));
	HIPC(hipMemcpy(*d, h, n * sizeof(double), hipMemcpyHostToDevice));
	return SSX_OK;
}

/* ------------------------------------------------------------------ */
/* host geometry of one level (geometry.c:102-135, deriv.c:68-79,       */
/* tripalg.c:165-234, helmholtz.c:51-77)                               */
/* ------------------------------------------------------------------ */
static void level_geometry(const ssx_basis *b, int nel, const double *x,
		const double *y, const double *z, double lambda, double *gf,
		double *mass, double *dp, double *dh)
{
	const int n = b->n, n2 = n * n, n3 = n2 * n;
	double *w3 = malloc(sizeof(double) * n3);
	double *G = malloc(sizeof(double) * 6 * n3);
	for (int k = 0; k < n; ++k)
		for (int j = 0; j < n; ++j)
			for (int i = 0; i < n; ++i)
				w3[(k * n + j) * n + i] = b->rho[i] * b->rho[j] * b->rho[k];
	const double *D = b->d;
	for (int e = 0; e < nel; ++e) {
		const size_t off = (size_t) e * n3;
		const double *c[3] = { x + off, y + off, z + off };
		for (int k = 0; k < n; ++k)
		for (int j = 0; j < n; ++j)
		for (int i = 0; i < n; ++i) {
			const int q = (k * n + j) * n + i;
			double J[3][3];
			for (int a = 0; a < 3; ++a) {
				double s0 = 0.0, s1 = 0.0, s2 = 0.0;
				for (int l = 0; l < n; ++l) {
					s0 += D[i * n + l] * c[a][(k * n + j) * n + l];
					s1 += D[j * n + l] * c[a][(k * n + l) * n + i];
					s2 += D[k * n + l] * c[a][(l * n + j) * n + i];
				}
				J[a][0] = s0;
				J[a][1] = s1;
				J[a][2] = s2;
			}
			const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
			const double c01 = J[1][0] * J[2][2] - J[1][2] * J[2][0];
			const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
			const double det = J[0][0] * c00 - J[0][1] * c01 + J[0][2] * c02;
			double Ji[3][3];
			Ji[0][0] = c00 / det;
			Ji[0][1] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) / det;
			Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
			Ji[1][0] = -c01 / det;
			Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
			Ji[1][2] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]) / det;
			Ji[2][0] = c02 / det;
			Ji[2][1] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]) / det;
			Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
			/* G = Jinv Jinv^T (rows of the inverse Jacobian) */
			double g6[6];
			const int pa[6] = { 0, 0, 0, 1, 1, 2 }, pb[6] = { 0, 1, 2, 1, 2, 2 };
			for (int s = 0; s < 6; ++s)
				g6[s] = Ji[pa[s]][0] * Ji[pb[s]][0]
					+ Ji[pa[s]][1] * Ji[pb[s]][1]
					+ Ji[pa[s]][2] * Ji[pb[s]][2];
			const double wj = w3[q] * det;
			for (int s = 0; s < 6; ++s) {
				G[s * n3 + q] = g6[s];
				gf[(size_t) e * 6 * n3 + (size_t) s * n3 + q] = g6[s] * wj;
			}
			mass[off + q] = wj;
		}
		/* approximate Jacobi diagonal: no detJ, no cross terms
		 * (helmholtz.c:51-68), lambda term for the Helmholtz one */
		for (int k = 0; k < n; ++k)
		for (int j = 0; j < n; ++j)
		for (int i = 0; i < n; ++i) {
			const int q = (k * n + j) * n + i;
			double t = 0.0;
			for (int p = 0; p < n; ++p) {
				const int q1 = (k * n + j) * n + p, q2 = (k * n + p) * n + i,
					q3 = (p * n + j) * n + i;
				t += w3[q1] * (D[p * n + i] * D[p * n + i]) * G[0 * n3 + q1];
				t += w3[q2] * (D[p * n + j] * D[p * n + j]) * G[3 * n3 + q2];
				t += w3[q3] * (D[p * n + k] * D[p * n + k]) * G[5 * n3 + q3];
			}
			dp[off + q] = t;
			dh[off + q] = t + lambda * mass[off + q];
		}
	}
	free(G);
	free(w3);
}

/* coordinates of a coarser level: the finer level's coordinate polynomial
 * evaluated at the coarse GLL points (precon.c:144-152) */
static void coarsen_coords(const ssx_basis *bf, const ssx_basis *bc, int nel,
		const double *xf, double *xc)
{
	const int nf = bf->n, nc = bc->n;
	double A[SSX_NMAX * SSX_NMAX];
	ssx_interp(bf->xi, nf, bc->xi, nc, A);  /* A[p*nc+i] = l_p^f(xi_i^c) */
	double *t1 = malloc(sizeof(double) * nf * nf * nc);
	double *t2 = malloc(sizeof(double) * nf * nc * nc);
	for (int e = 0; e < nel; ++e) {
		const double *u = xf + (size_t) e * nf * nf * nf;
		double *v = xc + (size_t) e * nc * nc * nc;
		for (int kj = 0; kj < nf * nf; ++kj)
			for (int i = 0; i < nc; ++i) {
				double s = 0.0;
				for (int p = 0; p < nf; ++p)
					s += A[p * nc + i] * u[kj * nf + p];
				t1[kj * nc + i] = s;
			}
		for (int k = 0; k < nf; ++k)
			for (int j = 0; j < nc; ++j)
				for (int i = 0; i < nc; ++i) {
					double s = 0.0;
					for (int l = 0; l < nf; ++l)
						s += t1[(k * nf + l) * nc + i] * A[l * nc + j];
					t2[(k * nc + j) * nc + i] = s;
				}
		for (int k = 0; k < nc; ++k)
			for (int ji = 0; ji < nc * nc; ++ji) {
				double s = 0.0;
				for (int l = 0; l < nf; ++l)
					s += t2[l * nc * nc + ji] * A[l * nc + k];
				v[k * nc * nc + ji] = s;
			}
	}
	free(t1);
	free(t2);
}

static int setup_level(ssx_ctx *c, const ssx_mesh *m, int l, double *hx,
		double *hy, double *hz)
{
	ssx_level *L = &c->lev[l];
	const int n = L->n, n3 = L->n3;
	const size_t nt = L->ntot;
	int rc;
	double *mult = NULL;
	double *gf = malloc(sizeof(double) * 6 * nt);
	double *mass = malloc(sizeof(double) * nt);
	double *dp = malloc(sizeof(double) * nt);
	double *dh = malloc(sizeof(double) * nt);
	double *one = malloc(sizeof(double) * nt);
	if (!gf || !mass || !dp || !dh || !one) {
		free(gf); free(mass); free(dp); free(dh); free(one);
		ssx_set_error("out of host memory (level %d)", l);
		return SSX_ENOMEM;
	}
	level_geometry(&L->b, c->nel, hx, hy, hz, c->lambda, gf, mass, dp, dh);
	rc = ssx_build_groups(m, n, L->grp, L->ngrp);
	if (rc != SSX_OK)
		goto out;
	/* assembled diagonals and multiplicity (precon.c:336-381) */
	ssx_host_dss(L->grp, L->ngrp, dp);
	ssx_host_dss(L->grp, L->ngrp, dh);
	for (size_t q = 0; q < nt; ++q) {
		dp[q] = 1.0 / dp[q];
		dh[q] = 1.0 / dh[q];
		one[q] = 1.0;
	}
	ssx_host_dss(L->grp, L->ngrp, one);
	double nun = 0.0;
	for (size_t q = 0; q < nt; ++q) {
		one[q] = 1.0 / one[q];
		nun += one[q];
	}
	L->nunique = nun;
	mult = malloc(sizeof(double) * nt);
	if (!mult) {
		rc = SSX_ENOMEM;
		goto out;
	}
	memcpy(mult, one, sizeof(double) * nt);
	rc = up(&L->gf, gf, 6 * nt);
	if (!rc) rc = up(&L->mass, mass, nt);
	if (!rc) rc = up(&L->diag_p, dp, nt);
	if (!rc) rc = up(&L->diag_h, dh, nt);
	if (!rc) rc = up(&L->scale, one, nt);
	if (rc)
		goto out;
	if (l == 0) {            /* inverse mass, precon.c:397-413 */
		memcpy(one, mass, sizeof(double) * nt);
		ssx_host_dss(L->grp, L->ngrp, one);
		for (size_t q = 0; q < nt; ++q)
			one[q] = 1.0 / one[q];
		rc = up(&c->inv_mass, one, nt);
		if (rc)
			goto out;
	}
	/* Dirichlet mask (precon.c:383-395, poisson.c:101-143) */
	for (int e = 0; e < c->nel; ++e) {
		int bits = 0;
		for (int s = 0; s < 6; ++s)
			if (m->neigh[e * 6 + s] == -1)
				bits |= 1 << s;
		for (int k = 0; k < n; ++k)
			for (int j = 0; j < n; ++j)
				for (int i = 0; i < n; ++i) {
					int z0 = ((bits & 1) && i == 0) || ((bits & 2) && i == n - 1)
						|| ((bits & 4) && j == 0) || ((bits & 8) && j == n - 1)
						|| ((bits & 16) && k == 0) || ((bits & 32) && k == n - 1);
					one[(size_t) e * n3 + (k * n + j) * n + i] = z0 ? 0.0 : 1.0;
				}
	}
	rc = up(&L->mask, one, nt);
	if (rc)
		goto out;
	if (l > 0 && l == c->nlev - 1) {   /* coarsest level: assembled form */
		rc = ssx_coarse_build(c, l, gf, mass, dp, dh, one);
		if (rc)
			goto out;
	}
	{
		double dm[2 * SSX_NMAX * SSX_NMAX];
		memcpy(dm, L->b.d, sizeof(double) * n * n);
		memcpy(dm + n * n, L->b.dsvv, sizeof(double) * n * n);
		rc = up(&L->dmat, dm, 2 * (size_t) n * n);
		if (rc)
			goto out;
	}
	for (int mi = 0; mi < 3; ++mi) {
		const int mult = 2 << mi;
		int *d;
		size_t bytes = sizeof(int) * (size_t) (L->ngrp[mi] ? L->ngrp[mi] : 1)
			* mult;
		if (hipMalloc((void **) &d, bytes) != hipSuccess) {
			rc = SSX_ENOMEM;
			ssx_set_error("hipMalloc groups");
			goto out;
		}
		if (L->ngrp[mi] && hipMemcpy(d, L->grp[mi], sizeof(int) * (size_t)
				L->ngrp[mi] * mult, hipMemcpyHostToDevice) != hipSuccess) {
			rc = SSX_EHIP;
			ssx_set_error("upload groups");
			goto out;
		}
		free(L->grp[mi]);
		L->grp[mi] = d;
	}
	{
		/* bits 0-5: Dirichlet faces; bits 8-13: faces with a neighbour */
		int *bcb = malloc(sizeof(int) * c->nel);
		for (int e = 0; e < c->nel; ++e) {
			int bits = 0;
			for (int s = 0; s < 6; ++s) {
				if (m->neigh[e * 6 + s] == -1)
					bits |= 1 << s;
				if (m->neigh[e * 6 + s] >= 0)
					bits |= 256 << s;
			}
			bcb[e] = bits;
		}
		/* the fused inner products take multiplicity 1 for nodes on no
		 * face with a neighbour; check that against the assembled
		 * multiplicity (scale still holds 1/mult in `mult`) */
		L->fused_ok = 1;
		for (int e = 0; e < c->nel && L->fused_ok; ++e) {
			const int eb = bcb[e];
			for (int k = 0; k < n && L->fused_ok; ++k)
				for (int j = 0; j < n; ++j)
					for (int i = 0; i < n; ++i) {
						int sh = ((eb & 256) && i == 0) || ((eb & 512) && i == n - 1)
							|| ((eb & 1024) && j == 0) || ((eb & 2048) && j == n - 1)
							|| ((eb & 4096) && k == 0) || ((eb & 8192) && k == n - 1);
						double w = mult[(size_t) e * n3 + (k * n + j) * n + i];
						if ((w < 1.0) != sh) {
							L->fused_ok = 0;
							break;
						}
					}
		}
		if (hipMalloc((void **) &L->bc, sizeof(int) * c->nel) != hipSuccess
				|| hipMemcpy(L->bc, bcb, sizeof(int) * c->nel,
					hipMemcpyHostToDevice) != hipSuccess) {
			free(bcb);
			rc = SSX_EHIP;
			ssx_set_error("upload bc");
			goto out;
		}
		free(bcb);
	}
	if (L->fused_ok) {       /* partner lists for the gather-form DSS */
		int *tab = NULL, *soff = NULL, P = 0;
		rc = ssx_build_gather(m, n, &tab, &soff, &P);
		if (rc)
			goto out;
		const size_t tb = sizeof(int) * (size_t) c->nel * (P ? P : 1);
		if (hipMalloc((void **) &L->gtab, tb) != hipSuccess
				|| hipMalloc((void **) &L->gsoff, sizeof(int) * n3)
					!= hipSuccess
				|| (P && hipMemcpy(L->gtab, tab, sizeof(int) * (size_t)
					c->nel * P, hipMemcpyHostToDevice) != hipSuccess)
				|| hipMemcpy(L->gsoff, soff, sizeof(int) * n3,
					hipMemcpyHostToDevice) != hipSuccess) {
			free(tab);
			free(soff);
			rc = SSX_EHIP;
			ssx_set_error("upload gather tables");
			goto out;
		}
		free(tab);
		free(soff);
		L->gP = P;
	}
	rc = dalloc(&L->r, nt);
	if (!rc) rc = dalloc(&L->z, nt);
	if (!rc) rc = dalloc(&L->p, nt);
	if (!rc) rc = dalloc(&L->ap, nt);
	if (!rc) rc = dalloc(&L->mg_r, nt);
	if (!rc) rc = dalloc(&L->mg_fn, nt);
	if (!rc) rc = dalloc(&L->mg_e, nt);
out:
	free(gf); free(mass); free(dp); free(dh); free(one); free(mult);
	return rc;
}

int ssx_create(ssx_ctx **out, const ssx_mesh *m)
{
	*out = NULL;
	if (!m || m->n < SSX_NMIN || m->n > SSX_NMAX || m->nel < 1) {
		ssx_set_error("ssx_create: n must be in [%d,%d] and nel >= 1",
				SSX_NMIN, SSX_NMAX);
		return SSX_EINVAL;
	}
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
		ssx_set_error("ssx_create: no GPU visible");
		return SSX_ENODEV;
	}
	ssx_ctx *c = calloc(1, sizeof(*c));
	if (!c)
		return SSX_ENOMEM;
	if (hipGetDevice(&c->device) != hipSuccess) {
		free(c);
		ssx_set_error("ssx_create: hipGetDevice failed");
		return SSX_EHIP;
	}
	c->nel = m->nel;
	c->n = m->n;
	c->diag_choice = -1;
	c->fused = 1;
	c->persist = 0;
	c->p_elem_walk = 1;
	c->coarse_sparse = SSX_COARSE_SPARSE_DEFAULT;
	c->fused_red = 0;
	c->coarse_csr = 1;
	c->gather_upd = 1;
	c->lp_coarse_rtol = pow(10.0, -SSX_LP_COARSE_RTOL_EXP);
	c->lambda = (11.0 / 6.0) / (m->dt * m->visc);
	int nl = SSX_LMAX;
	for (int i = 0; i < SSX_LMAX; ++i) {     /* precon.c:62-70 */
		int v = (m->n - 1) / (1 << i) + 1;
		c->nmg[i] = v > 2 ? v : 2;
		if (c->nmg[i] == 2) {
			nl = i + 1;
			break;
		}
	}
	c->nlev = nl;
	hipStream_t st;
	if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
		free(c);
		ssx_set_error("hipStreamCreate failed");
		return SSX_EHIP;
	}
	c->own_stream = st;
	c->stream = st;
	c->h_neigh = malloc(sizeof(int) * 6 * (size_t) m->nel);
	if (c->h_neigh)
		memcpy(c->h_neigh, m->neigh, sizeof(int) * 6 * (size_t) m->nel);
	int rc = SSX_OK;
	/* coordinates of the previous (finer) level; level 0 keeps its own */
	double *px = NULL, *py = NULL, *pz = NULL;
	for (int l = 0; l < nl && rc == SSX_OK; ++l) {
		ssx_level *L = &c->lev[l];
		L->n = c->nmg[l];
		L->n3 = L->n * L->n * L->n;
		L->ntot = (size_t) c->nel * L->n3;
		ssx_basis_init(&L->b, L->n, m->svvn, m->svva);
		double *nx = malloc(sizeof(double) * L->ntot);
		double *ny = malloc(sizeof(double) * L->ntot);
		double *nz = malloc(sizeof(double) * L->ntot);
		if (!nx || !ny || !nz) {
			free(nx); free(ny); free(nz);
			ssx_set_error("out of host memory (coordinates)");
			rc = SSX_ENOMEM;
			break;
		}
		if (l == 0) {
			memcpy(nx, m->x, sizeof(double) * L->ntot);
			memcpy(ny, m->y, sizeof(double) * L->ntot);
			memcpy(nz, m->z, sizeof(double) * L->ntot);
		} else {
			const ssx_basis *bf = &c->lev[l - 1].b;
			coarsen_coords(bf, &L->b, c->nel, px, nx);
			coarsen_coords(bf, &L->b, c->nel, py, ny);
			coarsen_coords(bf, &L->b, c->nel, pz, nz);
		}
		if (l >= 2) {
			free(px); free(py); free(pz);
		}
		px = nx; py = ny; pz = nz;
		if (l == 0) {
			L->h_x = nx; L->h_y = ny; L->h_z = nz;
		}
		rc = setup_level(c, m, l, nx, ny, nz);
	}
	/* the coarsest level keeps its coordinates too (the nested-dissection
	 * order of its sparse direct solve is taken from them) */
	if (px && px != c->lev[0].h_x) {
		if (rc == SSX_OK && nl > 1 && !c->lev[nl - 1].h_x) {
			ssx_level *Lc = &c->lev[nl - 1];
			Lc->h_x = px; Lc->h_y = py; Lc->h_z = pz;
		} else {
			free(px); free(py); free(pz);
		}
	}
	/* prolongation matrices */
	for (int l = 0; l + 1 < nl && rc == SSX_OK; ++l) {
		double P[SSX_NMAX * SSX_NMAX];
		ssx_interp(c->lev[l + 1].b.xi, c->lev[l + 1].n, c->lev[l].b.xi,
				c->lev[l].n, P);
		rc = up(&c->lev[l].prol, P, (size_t) c->lev[l + 1].n * c->lev[l].n);
	}
	if (rc == SSX_OK) {
		const size_t nt = c->lev[0].ntot;
		size_t np = (size_t) 12 * SSX_NPART;   /* FCG: 10 dots at once */
		for (int l = 0; l < nl; ++l) {
			/* one operator partial per element, then DSS partials */
			size_t q = (size_t) c->nel
				+ (size_t) ssxk_dss_blocks(c->lev[l].ngrp);
			if (q > np)
				np = q;
		}
		c->npart = np;
		rc = dalloc(&c->scal, SSX_NSCAL);
		if (!rc) rc = dalloc(&c->part, np);
		if (!rc && hipMalloc((void **) &c->pbar, 4 * sizeof(unsigned))
				!= hipSuccess)
			rc = SSX_ENOMEM;
		if (!rc && (hipMalloc((void **) &c->redctr, 4 * sizeof(unsigned))
				!= hipSuccess || hipMemset(c->redctr, 0,
				4 * sizeof(unsigned)) != hipSuccess))
			rc = SSX_ENOMEM;
		if (!rc && hipMalloc((void **) &c->cg_flags, sizeof(int) * 2 * SSX_LMAX)
				!= hipSuccess) {
			rc = SSX_ENOMEM;
			ssx_set_error("hipMalloc flags");
		}
		if (!rc && hipHostMalloc((void **) &c->h_flags,
				sizeof(int) * 2 * SSX_LMAX, 0) != hipSuccess) {
			rc = SSX_ENOMEM;
			ssx_set_error("hipHostMalloc flags");
		}
		if (!rc) rc = dalloc(&c->fcg_r, nt);
		if (!rc) rc = dalloc(&c->fcg_z, nt);
		if (!rc && hipHostMalloc((void **) &c->h_scal,
				SSX_NSCAL * sizeof(double), 0) != hipSuccess) {
			rc = SSX_ENOMEM;
			ssx_set_error("hipHostMalloc");
		}
		if (!rc)
			HIPC(hipMemset(c->scal, 0, SSX_NSCAL * sizeof(double)));
	}
	if (rc == SSX_OK)
		rc = ssx_autotune(c, 1 << SSX_PREC_FP64);
	if (rc != SSX_OK) {
		ssx_destroy(c);
		return rc;
	}
	*out = c;
	return SSX_OK;
}

void ssx_destroy(ssx_ctx *c)
{
	if (!c)
		return;
	hipStreamSynchronize((hipStream_t) c->stream);
	for (int l = 0; l < c->nlev; ++l) {
		ssx_level *L = &c->lev[l];
		hipFree(L->gf); hipFree(L->mass); hipFree(L->diag_p);
		hipFree(L->diag_h); hipFree(L->scale); hipFree(L->mask);
		hipFree(L->dmat); hipFree(L->prol); hipFree(L->bc);
		for (int mi = 0; mi < 3; ++mi)
			hipFree(L->grp[mi]);
		hipFree(L->gtab); hipFree(L->gsoff);
		hipFree(L->r); hipFree(L->z); hipFree(L->p); hipFree(L->ap);
		hipFree(L->mg_r); hipFree(L->mg_fn); hipFree(L->mg_e);
		ssx_coarse_free(&L->cs);
		free(L->h_x); free(L->h_y); free(L->h_z);
		hipFree(L->gf_lp); hipFree(L->mass_lp); hipFree(L->diagp_lp);
		hipFree(L->diagh_lp); hipFree(L->scale_lp); hipFree(L->mask_lp);
		hipFree(L->dmat_lp); hipFree(L->prol_lp);
		hipFree(L->r_lp); hipFree(L->z_lp); hipFree(L->p_lp);
		hipFree(L->ap_lp); hipFree(L->mgr_lp); hipFree(L->fn_lp);
		hipFree(L->e_lp);
	}
	hipFree(c->inv_mass);
	hipFree(c->cg_flags);
	hipFree(c->pbar);
	hipFree(c->redctr);
	if (c->h_flags)
		hipHostFree(c->h_flags);
	hipFree(c->scal);
	hipFree(c->part);
	hipFree(c->fcg_r);
	hipFree(c->fcg_z);
	for (int i = 0; i < 11; ++i)
		hipFree(c->fcg_p[i]);
	for (int i = 0; i < 10; ++i)
		hipFree(c->fcg_ap[i]);
	hipFree(c->d_f);
	hipFree(c->d_u);
	hipFree(c->p_stage);
	hipFree(c->p_perm);
	hipFree(c->p_bucket);
	hipFree(c->p_work);
	hipFree(c->p_pcnt);
	ssx_coarse_direct_free(c);
	free(c->h_neigh);
	hipFree(c->d_neigh);
	for (int a = 0; a < 9; ++a) {
		hipFree(c->p_jac[a]);
		hipFree(c->p_jinv[a]);
		hipFree(c->p_ur[a / 3][a % 3]);
	}
	hipFree(c->p_xi);
	hipFree(c->p_maxpass);
	if (c->h_scal)
		hipHostFree(c->h_scal);
	if (c->own_stream)
		hipStreamDestroy((hipStream_t) c->own_stream);
	free(c);
}

int ssx_set_stream(ssx_ctx *c, void *s)
{
	if (!c)
		return SSX_EINVAL;
	HIPC(hipStreamSynchronize((hipStream_t) c->stream));
	c->stream = s;
	return SSX_OK;
}

void *ssx_get_stream(ssx_ctx *c) { return c ? c->stream : NULL; }
void *ssx_own_stream(ssx_ctx *c) { return c ? c->own_stream : NULL; }

int ssx_synchronize(ssx_ctx *c)
{
	HIPC(hipStreamSynchronize((hipStream_t) c->stream));
	return SSX_OK;
}

int ssx_nlevels(const ssx_ctx *c) { return c->nlev; }
int ssx_nel(const ssx_ctx *c) { return c->nel; }
double ssx_lambda(const ssx_ctx *c) { return c->lambda; }

int ssx_level_n(const ssx_ctx *c, int l)
{
	return (l >= 0 && l < c->nlev) ? c->lev[l].n : -1;
}

size_t ssx_level_ntot(const ssx_ctx *c, int l)
{
	return (l >= 0 && l < c->nlev) ? c->lev[l].ntot : 0;
}

int ssx_set_option(ssx_ctx *c, int option, int value)
{
	if (!c)
		return SSX_EINVAL;
	switch (option) {
	case SSX_OPT_FUSED_CG:
		c->fused = value != 0;
		return SSX_OK;
	case SSX_OPT_COARSE_ASSEMBLED:
		if (value < 0 || value > 2) {
			ssx_set_error("coarse option %d out of [0,2]", value);
			return SSX_EINVAL;
		}
		c->coarse_csr = value;
		return SSX_OK;
	case SSX_OPT_GATHER_DSS:
		c->gather_upd = value != 0;
		return SSX_OK;
	case SSX_OPT_FUSED_REDUCE:
		c->fused_red = value != 0;
		return SSX_OK;
	case SSX_OPT_LP_COARSE_RTOL:
		if (value < 1 || value > 15) {
			ssx_set_error("coarse tolerance exponent %d out of [1,15]",
					value);
			return SSX_EINVAL;
		}
		c->lp_coarse_rtol = pow(10.0, -value);
		return SSX_OK;
	case SSX_OPT_PERSIST_ABORT_TEST:
		c->persist_abort_test = value != 0;
		return SSX_OK;
	case SSX_OPT_COARSE_DIRECT:
		if (value < 0 || value > 46340) {
			ssx_set_error("direct coarse solve: %d unknowns out of range",
					value);
			return SSX_EINVAL;
		}
		c->coarse_direct = value;
		return SSX_OK;
	case SSX_OPT_COARSE_SPARSE:
		if (value < 0) {
			ssx_set_error("sparse coarse solve: %d unknowns out of range",
					value);
			return SSX_EINVAL;
		}
		c->coarse_sparse = value;
		return SSX_OK;
	case SSX_OPT_PARTICLE_WALK:
		if (value != 0 && value != 1) {
			ssx_set_error("particle walk option %d not 0 or 1", value);
			return SSX_EINVAL;
		}
		c->p_elem_walk = value;
		return SSX_OK;
	case SSX_OPT_PERSISTENT_CG:
		if (value < 0 || value > 1024) {
			ssx_set_error("persistent CG option %d out of range", value);
			return SSX_EINVAL;
		}
		c->persist = value != 0;
		c->persist_grid = value > 1 ? value : 0;
		return SSX_OK;
	case SSX_OPT_SMOOTHER_PRECISION:
		if (value != SSX_PREC_FP64 && value != SSX_PREC_FP32
				&& value != SSX_PREC_BF16) {
			ssx_set_error("precision %d unsupported", value);
			return SSX_EINVAL;
		}
		c->smooth_prec = value;
		return SSX_OK;
	default:
		ssx_set_error("unknown option %d", option);
		return SSX_EINVAL;
	}
}

int ssx_set_diag_choice(ssx_ctx *c, int choice)
{
	if (choice < -1 || choice > 1)
		return SSX_EINVAL;
	c->diag_choice = choice;
	return SSX_OK;
}

int ssx_dss_groups(const ssx_ctx *c, int l, long long counts[3])
{
	if (l < 0 || l >= c->nlev)
		return SSX_EINVAL;
	for (int i = 0; i < 3; ++i)
		counts[i] = c->lev[l].ngrp[i];
	return SSX_OK;
}

int ssx_level_data(ssx_ctx *c, int l, int which, double *host)
{
	if (l < 0 || l >= c->nlev || !host)
		return SSX_EINVAL;
	ssx_level *L = &c->lev[l];
	const double *src = NULL;
	size_t n = L->ntot;
	switch (which) {
	case SSX_DATA_GFAC: src = L->gf; n = 6 * L->ntot; break;
	case SSX_DATA_MASS: src = L->mass; break;
	case SSX_DATA_DIAG_P: src = L->diag_p; break;
	case SSX_DATA_DIAG_H: src = L->diag_h; break;
	case SSX_DATA_SCALE: src = L->scale; break;
	case SSX_DATA_MASK: src = L->mask; break;
	case SSX_DATA_INVMASS: src = l == 0 ? c->inv_mass : NULL; break;
	case SSX_DATA_X: case SSX_DATA_Y: case SSX_DATA_Z:
		if (l != 0)
			return SSX_EINVAL;
		memcpy(host, which == SSX_DATA_X ? L->h_x : which == SSX_DATA_Y
				? L->h_y : L->h_z, sizeof(double) * n);
		return SSX_OK;
	default: return SSX_EINVAL;
	}
	if (!src)
		return SSX_EINVAL;
	HIPC(hipStreamSynchronize((hipStream_t) c->stream));
	HIPC(hipMemcpy(host, src, n * sizeof(double), hipMemcpyDeviceToHost));
	return SSX_OK;
}

/* ------------------------------------------------------------------ */
/* operators                                                            */
/* ------------------------------------------------------------------ */

static int chk(const ssx_ctx *c, int l)
{
	if (!c || l < 0 || l >= c->nlev) {
		ssx_set_error("bad context or level %d", l);
		return SSX_EINVAL;
	}
	return SSX_OK;
}

int ssx_ax(ssx_ctx *c, int l, int op, const double *u, double *au)
{
	TRY(chk(c, l));
	ssx_level *L = &c->lev[l];
	return ssxk_ax(c->stream, L->n, L->axv[0], c->nel, op, u, au, L->gf,
			L->mass,
			L->dmat, c->lambda, L->bc, 0, NULL, NULL);
}

int ssx_dss(ssx_ctx *c, int l, double *u)
{
	TRY(chk(c, l));
	return ssxk_dss_all(c->stream, c->lev[l].grp, c->lev[l].ngrp, u, NULL,
			NULL, 0, NULL);
}

int ssx_opeval(ssx_ctx *c, int l, int op, int var, const double *p,
		double *ap)
{
	TRY(chk(c, l));
	ssx_level *L = &c->lev[l];
	if (l == 0)
		c->cnt[0]++;
	/* element operator, mask (var != PRESSURE), DSS: solver.c:191-198 */
	TRY(ssxk_ax(c->stream, L->n, L->axv[0], c->nel, op, p, ap, L->gf,
			L->mass,
			L->dmat, c->lambda, L->bc, var != SSX_PRESSURE, NULL, NULL));
	return ssxk_dss_all(c->stream, L->grp, L->ngrp, ap, NULL, NULL, 0, NULL);
}

int ssx_dss_average(ssx_ctx *c, int l, double *u)
{
	TRY(chk(c, l));
	/* DSS then x 1/mult; 1/mult is exactly 1/M on a group of M copies and
	 * 1 elsewhere, so the group kernel scales in place (global.c:259-268) */
	return ssxk_dss_all(c->stream, c->lev[l].grp, c->lev[l].ngrp, u, NULL,
			NULL, 1, NULL);
}

static int dot_to(ssx_ctx *c, int l, const double *x, const double *y,
		int dst)
{
	ssx_level *L = &c->lev[l];
	TRY(ssxk_dot(c->stream, L->ntot, x, L->scale, y, c->part, 0));
	return ssxk_reduce(c->stream, c->part, 0, c->scal, dst);
}

/* global_opeval's element loop and mask (solver.c:191-198) without the
 * DSS, into the level's Ap buffer: the gather-form kernels assemble it as
 * they read it */
static int opeval_raw(ssx_ctx *c, int l, int op, int var, const double *p)
{
	ssx_level *L = &c->lev[l];
	if (l == 0)
		c->cnt[0]++;
	return ssxk_ax(c->stream, L->n, L->axv[0], c->nel, op, p, L->ap, L->gf,
			L->mass, L->dmat, c->lambda, L->bc, var != SSX_PRESSURE, NULL,
			NULL);
}

static int read_scal(ssx_ctx *c, int slot, double *v)
{
	HIPC(hipMemcpyAsync(c->h_scal + slot, c->scal + slot, sizeof(double),
			hipMemcpyDeviceToHost, (hipStream_t) c->stream));
	HIPC(hipStreamSynchronize((hipStream_t) c->stream));
	*v = c->h_scal[slot];
	return SSX_OK;
}

int ssx_dot(ssx_ctx *c, int l, const double *x, const double *y, double *out)
{
	TRY(chk(c, l));
	TRY(dot_to(c, l, x, y, SC_TMP));
	return read_scal(c, SC_TMP, out);
}

int ssx_mask(ssx_ctx *c, int l, int var, double *u)
{
	TRY(chk(c, l));
	if (var == SSX_PRESSURE)          /* poisson.c:91-92 */
		return SSX_OK;
	return ssxk_mul(c->stream, c->lev[l].ntot, u, c->lev[l].mask, u);
}

int ssx_nullspace(ssx_ctx *c, int l, int var, double *u)
{
	TRY(chk(c, l));
	if (var != SSX_PRESSURE)          /* solver.c:202-203 */
		return SSX_OK;
	ssx_level *L = &c->lev[l];
	TRY(ssxk_dot(c->stream, L->ntot, u, L->scale, NULL, c->part, 0));
	TRY(ssxk_reduce(c->stream, c->part, 0, c->scal, SC_NULLC));
	return ssxk_shift_s(c->stream, L->ntot, c->scal, SC_NULLC, L->nunique, u);
}

int ssx_restrict(ssx_ctx *c, int l, const double *fine, double *coarse)
{
	if (l < 1 || l >= c->nlev)
		return SSX_EINVAL;
	TRY(ssxk_restrict(c->stream, c->lev[l - 1].n, c->lev[l].n, c->nel,
			c->lev[l - 1].prol, fine, coarse));
	return ssx_dss_average(c, l, coarse);   /* precon.c:283-291 */
}

int ssx_prolong(ssx_ctx *c, int l, const double *coarse, double *fine)
{
	if (l < 0 || l + 1 >= c->nlev)
		return SSX_EINVAL;
	return ssxk_prolong(c->stream, c->lev[l].n, c->lev[l + 1].n, c->nel,
			c->lev[l].prol, coarse, fine, 0);
}

static const double *jdiag(ssx_ctx *c, int l, int var)
{
	int ch = c->diag_choice;
	if (ch < 0)
		ch = var == SSX_PRESSURE ? 0 : 1;      /* precon.c:211-225 */
	return ch == 0 ? c->lev[l].diag_p : c->lev[l].diag_h;
}

int ssx_jacobi(ssx_ctx *c, int l, int var, const double *x, double *px)
{
	TRY(chk(c, l));
	return ssxk_mul(c->stream, c->lev[l].ntot, x, jdiag(c, l, var), px);
}

/* ------------------------------------------------------------------ */
/* solvers                                                              */
/* ------------------------------------------------------------------ */

/* conjugate_gradient, solver.c:33-81, Jacobi preconditioner at level l:
 * one launch per vector operation and one host read of the residual per
 * iteration (kept as the plain path and for meshes where the fused inner
 * products do not apply) */
static int cg_level_plain(ssx_ctx *c, int l, int op, int var, double *f,
		double *u, double eps, int max_iter, int *iters, double *err_out)
{
	ssx_level *L = &c->lev[l];
	const size_t nt = L->ntot;
	double *r = L->r, *z = L->z, *p = L->p, *ap = L->ap;
	void *st = c->stream;
	const double *mask = var == SSX_PRESSURE ? NULL : L->mask;
	TRY(ssx_mask(c, l, var, f));
	TRY(ssx_opeval(c, l, op, var, u, r));
	TRY(ssxk_axpby(st, nt, 1.0, f, NULL, -1.0, r));
	TRY(ssx_nullspace(c, l, var, r));
	TRY(ssx_jacobi(c, l, var, r, z));
	TRY(ssx_nullspace(c, l, var, z));
	TRY(ssxk_copy(st, p, z, nt));
	TRY(dot_to(c, l, r, z, SC_RDZ));
	int j = 0, it = 0;
	double err = 0.0;
	do {
		TRY(ssx_opeval(c, l, op, var, p, ap));
		TRY(dot_to(c, l, p, ap, SC_PAP));
		TRY(ssxk_scalar(st, c->scal, 0, SC_ALPHA, SC_RDZ, SC_PAP, 0.0));
		TRY(ssxk_axpy_s(st, nt, c->scal, SC_ALPHA, 1.0, p, u));
		TRY(ssxk_axpy_s(st, nt, c->scal, SC_ALPHA, -1.0, ap, r));
		TRY(ssx_jacobi(c, l, var, r, z));
		TRY(ssx_nullspace(c, l, var, z));
		TRY(ssxk_scalar(st, c->scal, 2, SC_RDZ_OLD, SC_RDZ, 0, 0.0));
		TRY(dot_to(c, l, r, z, SC_RDZ));
		TRY(ssxk_scalar(st, c->scal, 0, SC_BETA, SC_RDZ, SC_RDZ_OLD, 0.0));
		TRY(ssxk_xpby_s(st, nt, z, c->scal, SC_BETA, p, mask));
		TRY(dot_to(c, l, r, r, SC_RR));
		double rr;
		TRY(read_scal(c, SC_RR, &rr));
		err = sqrt(rr);
		++it;
		if (isnan(err) || isinf(err)) {
			ssx_set_error("CG level %d: residual is %g", l, err);
			if (iters)
				*iters = it;
			*err_out = err;
			return SSX_ENUMERIC;
		}
	} while (err > eps && ++j < max_iter);
	if (iters)
		*iters = it;
	*err_out = err;
	return SSX_OK;
}

/* The start of conjugate_gradient (solver.c:44-56) at level l in FP64
 * storage, on the level's r, z, p, ap: f masked, r = f - A u (r = f when
 * x_zero: the guess is known to be zero), the null space removed, z =
 * D^-1 r, p = z, and <r,z>_w into scal[o + CG_RDZ] through the partials
 * `part` (SSX_NPART doubles).  rtol > 0: also <r,r>_w into scal[o + CG_RR]
 * and the stopping tolerance max(eps, rtol |r|_w) into scal[o + CG_TOL]. */
int ssx_cg_start(ssx_ctx *c, int l, int op, int var, double *f,
		const double *u, int x_zero, double *part, double *scal, int o,
		double eps, double rtol)
{
	ssx_level *L = &c->lev[l];
	const size_t nt = L->ntot;
	double *r = L->r, *z = L->z, *p = L->p, *ap = L->ap;
	void *st = c->stream;
	const int pres = var == SSX_PRESSURE;
	const double *dinv = jdiag(c, l, var);
	TRY(ssx_mask(c, l, var, f));
	if (!x_zero && !pres && c->gather_upd && L->gP > 0) {
		/* r = f - A u in one pass over the unassembled A u (bit-identical
		 * to the DSS launch and axpby below).  z, p and <r,z> stay
		 * separate here: the fully fused start (used by the FP32
		 * smoother, where it is bit-identical) moved the FP64 32^3
		 * solution by ~1e-16 after a few outer iterations */
		TRY(opeval_raw(c, l, op, var, u));
		TRY(ssxk_resid_g(st, 0, nt, L->n3, L->gP, L->gsoff, L->gtab, ap, f,
				NULL, r, NULL, NULL, NULL, NULL));
		TRY(ssxk_mul(st, nt, r, dinv, z));
		TRY(ssxk_copy(st, p, z, nt));
	} else {
		if (x_zero) {
			TRY(ssxk_copy(st, r, f, nt));
		} else {
			TRY(ssx_opeval(c, l, op, var, u, r));
			TRY(ssxk_axpby(st, nt, 1.0, f, NULL, -1.0, r));
		}
		TRY(ssx_nullspace(c, l, var, r));
		TRY(ssxk_mul(st, nt, r, dinv, z));
		TRY(ssx_nullspace(c, l, var, z));
		TRY(ssxk_copy(st, p, z, nt));
	}
	TRY(ssxk_dot(st, nt, r, L->scale, z, part, 0));
	TRY(ssxk_reduce(st, part, 0, scal, o + CG_RDZ));
	if (rtol > 0.0) {       /* stop at max(eps, rtol * |r0|_w) */
		TRY(ssxk_dot(st, nt, r, L->scale, r, part, 0));
		TRY(ssxk_reduce(st, part, 0, scal, o + CG_RR));
		TRY(ssxk_cg_tol(st, scal, o + CG_TOL, o + CG_RR, eps, rtol));
	}
	return SSX_OK;
}

/*
 * The same iteration fused: Ax (from the second iteration on with the
 * direction update p = (z - c) + beta p of the previous one) and the sum of
 * p (A_e p) over all local copies, which for p equal on the copies of each
 * shared node is <p, DSS(Ap)>_w; the DSS; a reduce that forms alpha; the
 * fused update (u, r, z and the <r,z>, <r,r> partials); a reduce that
 * forms beta and the stopping decision.  The stopping test of solver.c:73-74
 * runs on the device: once it fires every later kernel of the solve returns
 * at once, so the host enqueues iterations in chunks and synchronises once
 * per chunk.  x_zero: the initial guess is known to be zero (A*0 = 0 is
 * exact, so skipping the product changes nothing).
 */
static int cg_level_fused(ssx_ctx *c, int l, int op, int var, double *f,
		double *u, double eps, double rtol, int max_iter, int x_zero,
		int *iters, double *err_out)
{
	ssx_level *L = &c->lev[l];
	const size_t nt = L->ntot;
	double *r = L->r, *z = L->z, *p = L->p, *ap = L->ap;
	void *st = c->stream;
	const int pres = var == SSX_PRESSURE;
	const double *mask = pres ? NULL : L->mask;
	const double *dinv = jdiag(c, l, var);
	const int o = SC_CG(l);
	int *flags = c->cg_flags + 2 * l;
	TRY(ssx_cg_start(c, l, op, var, f, u, x_zero, c->part, c->scal, o, eps,
			rtol));
	TRY(ssxk_set_ints(st, flags, 0, 2));
	const double eps_dev = rtol > 0.0 ? -1.0 : eps;
	if (c->persist && L->n <= 4) {
		int grid = c->persist_grid;
		if (grid <= 0) {
			TRY(ssxk_cg_persist_max_grid(L->n, &grid));
			/* partials: 6 doubles per block in part */
			if ((size_t) grid * 6 > c->npart)
				grid = (int) (c->npart / 6);
			if (grid > 1024)
				grid = 1024;
		}
		TRY(ssxk_set_ints(st, (int *) c->pbar, 0, 4));
		if (c->persist_abort_test) {
			/* test hook: the launch starts aborted, so every block leaves
			 * at its first barrier, as when a barrier wait gives up */
			TRY(ssxk_set_ints(st, (int *) c->pbar + 2, 1, 1));
			c->persist_abort_test = 0;
		}
		TRY(ssxk_cg_persist(st, L->n, grid, c->nel, op, !pres, pres,
				max_iter, eps_dev, c->lambda, (double) L->nunique, L->gf,
				L->mass, L->dmat, dinv, L->scale, mask, L->bc, L->grp,
				L->ngrp, u, r, z, p, ap, c->part, c->scal, o, flags,
				c->pbar));
		hipStream_t hs = (hipStream_t) st;
		HIPC(hipMemcpyAsync(c->h_flags, flags, 2 * sizeof(int),
				hipMemcpyDeviceToHost, hs));
		HIPC(hipMemcpyAsync(c->h_scal + o + CG_RR, c->scal + o + CG_RR,
				sizeof(double), hipMemcpyDeviceToHost, hs));
		HIPC(hipStreamSynchronize(hs));
		if (c->h_flags[0] < 0) {
			ssx_set_error("CG level %d: a grid barrier of the persistent "
					"solve timed out", l);
			return SSX_EHIP;
		}
	} else {
	const long long nba = c->nel;   /* operator partials: one per element */
	const long long nbd = ssxk_dss_blocks(L->ngrp);
	const ssx_lastred lr0 = { c->redctr, c->part, nba, c->scal, o,
		max_iter, pres, flags, eps_dev, (double) L->nunique };
	const ssx_lastred lr1 = { c->redctr + 1, c->part, SSX_NPART, c->scal, o,
		max_iter, pres, flags, eps_dev, (double) L->nunique };
	int launched = 0;
	int chunk = c->cg_hint[l] > 0 ? c->cg_hint[l] + 1 : 8;
	/* the solution step u += alpha p of each iteration is taken by the
	 * next operator launch, from its load of p; the last one by a flush */
	const int defer = !c->fused_red;
	for (;;) {
		if (chunk > max_iter - launched)
			chunk = max_iter - launched;
		for (int k = 0; k < chunk; ++k) {
			/* from the second iteration on, the operator kernel first forms
			 * p = (z - c) + beta p (solver.c:70-72) */
			if (launched + k == 0)
				TRY(ssxk_ax(st, L->n, L->axv[0], c->nel, op, p, ap, L->gf,
						L->mass,
						L->dmat, c->lambda, L->bc, !pres, flags, c->part));
			else
				TRY(ssxk_ax_pupd(st, L->n, L->axv[0], c->nel, op, z, p,
						c->scal,
						o + CG_BETA, pres ? o + CG_NULLC : -1, ap, L->gf,
						L->mass, L->dmat, c->lambda, L->bc, !pres, flags,
						c->part, defer ? u : NULL, o + CG_ALPHA));
			if (c->fused_red && nbd > 0) {
				TRY(ssxk_dss_all_red(st, L->grp, L->ngrp, ap, NULL, NULL,
						flags, &lr0));
				TRY(ssxk_cg_update_red(st, nt, c->scal, o + CG_ALPHA, p, ap,
						u, r, z, dinv, L->scale, pres, c->part, flags, &lr1));
				continue;
			}
			if (c->gather_upd && L->gP > 0) {
				/* alpha needs only the operator's partials; the
				 * update then assembles Ap as it reads it */
				TRY(ssxk_cg_reduce(st, 0, c->part, nba, c->scal, o, flags,
						eps_dev, max_iter, L->nunique, pres));
				TRY(ssxk_cg_update_g(st, nt, L->n3, L->gP, L->gsoff,
						L->gtab, c->scal, o + CG_ALPHA, p, ap,
						defer ? NULL : u, r, z, dinv, pres, c->part, flags));
			} else {
				TRY(ssxk_dss_all(st, L->grp, L->ngrp, ap, NULL, NULL, 0,
						flags));
				TRY(ssxk_cg_reduce(st, 0, c->part, nba, c->scal, o, flags,
						eps_dev, max_iter, L->nunique, pres));
				TRY(ssxk_cg_update(st, nt, c->scal, o + CG_ALPHA, p, ap,
						defer ? NULL : u, r, z, dinv, L->scale, pres, c->part,
						flags));
			}
			TRY(ssxk_cg_reduce(st, 1, c->part, 0, c->scal, o, flags,
					eps_dev, max_iter, L->nunique, pres));
		}
		launched += chunk;
		hipStream_t hs = (hipStream_t) st;
		HIPC(hipMemcpyAsync(c->h_flags, flags, 2 * sizeof(int),
				hipMemcpyDeviceToHost, hs));
		HIPC(hipMemcpyAsync(c->h_scal + o + CG_RR, c->scal + o + CG_RR,
				sizeof(double), hipMemcpyDeviceToHost, hs));
		HIPC(hipStreamSynchronize(hs));
		if (c->h_flags[0] || launched >= max_iter)
			break;
		chunk = 4;
	}
	if (defer && launched > 0)
		TRY(ssxk_cg_uflush(st, nt, c->scal, o + CG_ALPHA, p, u));
	}
	const int it = c->h_flags[1];
	const double err = sqrt(c->h_scal[o + CG_RR]);
	c->cg_hint[l] = it;
	if (iters)
		*iters = it;
	*err_out = err;
	if (isnan(err) || isinf(err)) {
		ssx_set_error("CG level %d: residual is %g", l, err);
		return SSX_ENUMERIC;
	}
	return SSX_OK;
}

static int cg_level(ssx_ctx *c, int l, int op, int var, double *f, double *u,
		double eps, int max_iter, int x_zero, int *iters, double *err_out)
{
	if (c->fused && c->lev[l].fused_ok)
		return cg_level_fused(c, l, op, var, f, u, eps, 0.0, max_iter,
				x_zero, iters, err_out);
	return cg_level_plain(c, l, op, var, f, u, eps, max_iter, iters,
			err_out);
}

static int mg_recurse(ssx_ctx *c, int l, int op, int var, double *f,
		double *x, double eps)
{
	double err;
	int it;
	if (l == c->nlev - 1) {          /* precon.c:242-244 */
		if (c->coarse_csr && c->lev[l].cs.nu > 0)
			TRY(ssx_coarse_cg(c, l, op, var, f, x, eps, 0.0, 500, l > 0, &it,
					&err));
		else
			TRY(cg_level(c, l, op, var, f, x, eps, 500, l > 0, &it, &err));
		c->cnt[1] += it;
		return SSX_OK;
	}
	ssx_level *L = &c->lev[l], *C = &c->lev[l + 1];
	void *st = c->stream;
	double *r = L->mg_r, *fn = C->mg_fn, *e = C->mg_e;
	if (l == 0 && c->gather_upd && L->gP > 0) {   /* precon.c:302-309 */
		TRY(opeval_raw(c, l, op, var, x));
		TRY(ssxk_resid_g(st, 0, L->ntot, L->n3, L->gP, L->gsoff, L->gtab,
				L->ap, f, NULL, r, NULL, NULL, NULL, NULL));
	} else if (l == 0) {
		TRY(ssx_opeval(c, l, op, var, x, r));
		TRY(ssxk_axpby(st, L->ntot, 1.0, f, NULL, -1.0, r));
	} else {
		TRY(ssxk_copy(st, r, f, L->ntot));
	}
	TRY(ssx_restrict(c, l + 1, r, fn));
	TRY(ssxk_fill(st, e, 0.0, C->ntot));
	TRY(mg_recurse(c, l + 1, op, var, fn, e, eps));
	/* x += J e (prolongation fused with the add, precon.c:262-263) */
	TRY(ssxk_prolong(st, L->n, C->n, c->nel, L->prol, e, x, 1));
	TRY(cg_level(c, l, op, var, f, x, eps, 15, 0, &it, &err));
	c->cnt[2] += it;
	return SSX_OK;
}

/* ------------------------------------------------------------------ */
/* mixed precision: FP32 (or FP32 + BF16 factors) smoother and transfer */
/* on every level but the coarsest, which stays FP64                    */
/* ------------------------------------------------------------------ */

static int falloc(float **p, size_t n)
{
	if (hipMalloc((void **) p, (n ? n : 1) * sizeof(float)) != hipSuccess) {
		*p = NULL;
		ssx_set_error("hipMalloc of %zu floats failed", n);
		return SSX_ENOMEM;
	}
	return SSX_OK;
}

static int cvt_new(ssx_ctx *c, float **dst, const double *src, size_t n)
{
	if (!*dst)
		TRY(falloc(dst, n));
	return ssxk_cvt(c->stream, SSX_PREC_FP64, SSX_PREC_FP32, src, *dst, n);
}

static int build_lp(ssx_ctx *c, int prec)
{
	int tuned = 0;
	for (int l = 0; l < c->nlev; ++l)
		if (!c->lev[l].fused_ok) {
			ssx_set_error("level %d: reduced-precision smoother needs the "
					"fused CG path", l);
			return SSX_EMESH;
		}
	for (int l = 0; l < c->nlev; ++l) {
		ssx_level *L = &c->lev[l];
		const size_t nt = L->ntot;
		if (!L->fn_lp)
			TRY(falloc(&L->fn_lp, nt));
		if (!L->e_lp)
			TRY(falloc(&L->e_lp, nt));
		if (l == c->nlev - 1)
			continue;                       /* coarsest: FP64 */
		if (L->lp_prec == prec)
			continue;                       /* already converted */
		hipFree(L->gf_lp);
		L->gf_lp = NULL;
		const size_t bytes = 6 * nt * (prec == SSX_PREC_BF16 ? 2 : 4);
		if (hipMalloc(&L->gf_lp, bytes) != hipSuccess) {
			ssx_set_error("hipMalloc low-precision factors");
			return SSX_ENOMEM;
		}
		TRY(ssxk_cvt(c->stream, SSX_PREC_FP64, prec, L->gf, L->gf_lp, 6 * nt));
		TRY(cvt_new(c, &L->mass_lp, L->mass, nt));
		TRY(cvt_new(c, &L->diagp_lp, L->diag_p, nt));
		TRY(cvt_new(c, &L->diagh_lp, L->diag_h, nt));
		TRY(cvt_new(c, &L->scale_lp, L->scale, nt));
		TRY(cvt_new(c, &L->mask_lp, L->mask, nt));
		TRY(cvt_new(c, &L->dmat_lp, L->dmat, 2 * (size_t) L->n * L->n));
		TRY(cvt_new(c, &L->prol_lp, L->prol,
				(size_t) c->lev[l + 1].n * L->n));
		if (!L->r_lp) {
			TRY(falloc(&L->r_lp, nt));
			TRY(falloc(&L->z_lp, nt));
			TRY(falloc(&L->p_lp, nt));
			TRY(falloc(&L->ap_lp, nt));
			TRY(falloc(&L->mgr_lp, nt));
		}
		L->lp_prec = prec;
		tuned = 1;
	}
	if (tuned)
		TRY(ssx_autotune(c, 1 << prec));
	return SSX_OK;
}

static const float *jdiag_lp(ssx_ctx *c, int l, int var)
{
	int ch = c->diag_choice;
	if (ch < 0)
		ch = var == SSX_PRESSURE ? 0 : 1;
	return ch == 0 ? c->lev[l].diagp_lp : c->lev[l].diagh_lp;
}

static int opeval_lp(ssx_ctx *c, int l, int op, int var, const float *p,
		float *ap)
{
	ssx_level *L = &c->lev[l];
	if (l == 0)
		c->cnt[0]++;
	TRY(ssxk_ax_lp(c->stream, L->lp_prec, L->n, L->axv[L->lp_prec], c->nel,
			op, p, ap, L->gf_lp,
			L->mass_lp, L->dmat_lp, c->lambda, L->bc, var != SSX_PRESSURE,
			NULL, NULL));
	return ssxk_dss_all_lp(c->stream, L->grp, L->ngrp, ap, NULL, NULL, 0,
			NULL);
}

/* opeval_lp without the DSS, into the level's FP32 Ap buffer */
static int opeval_raw_lp(ssx_ctx *c, int l, int op, int var, const float *p)
{
	ssx_level *L = &c->lev[l];
	if (l == 0)
		c->cnt[0]++;
	return ssxk_ax_lp(c->stream, L->lp_prec, L->n, L->axv[L->lp_prec],
			c->nel, op, p, L->ap_lp, L->gf_lp, L->mass_lp, L->dmat_lp,
			c->lambda, L->bc, var != SSX_PRESSURE, NULL, NULL);
}

static int nullspace_lp(ssx_ctx *c, int l, int var, float *u)
{
	if (var != SSX_PRESSURE)
		return SSX_OK;
	ssx_level *L = &c->lev[l];
	TRY(ssxk_dot_lp(c->stream, L->ntot, u, L->scale_lp, NULL, c->part, 0));
	TRY(ssxk_reduce(c->stream, c->part, 0, c->scal, SC_NULLC));
	return ssxk_shift_s_lp(c->stream, L->ntot, c->scal, SC_NULLC,
			L->nunique, u);
}

/* The start of conjugate_gradient (solver.c:44-56) at level l in FP32
 * storage (the reduced-precision smoother), on the level's FP32 r, z, p,
 * ap: f masked, r = f - A u, the null space removed, z = D^-1 r, p = z,
 * and <r,z>_w (accumulated in FP64) into scal[o + CG_RDZ] through the
 * partials `part` (SSX_NPART doubles). */
int ssx_cg_start_lp(ssx_ctx *c, int l, int op, int var, float *f,
		const float *u, double *part, double *scal, int o)
{
	ssx_level *L = &c->lev[l];
	const size_t nt = L->ntot;
	float *r = L->r_lp, *z = L->z_lp, *p = L->p_lp, *ap = L->ap_lp;
	void *st = c->stream;
	const int pres = var == SSX_PRESSURE;
	const float *mask = pres ? NULL : L->mask_lp;
	const float *dinv = jdiag_lp(c, l, var);
	if (!pres && c->gather_upd && L->gP > 0) {
		/* f = f mask, r = f - A u, z = D^-1 r, p = z and <r,z>_w in one
		 * pass over the unassembled A u (bit-identical to the sequence
		 * below) */
		TRY(opeval_raw_lp(c, l, op, var, u));
		TRY(ssxk_resid_g_lp(st, 1, nt, L->n3, L->gP, L->gsoff, L->gtab, ap,
				f, mask, r, z, p, dinv, part));
	} else {
		if (mask)
			TRY(ssxk_mul_lp(st, nt, f, mask, f));
		TRY(opeval_lp(c, l, op, var, u, r));
		TRY(ssxk_axpby_lp(st, nt, 1.0, f, NULL, -1.0, r));
		TRY(nullspace_lp(c, l, var, r));
		TRY(ssxk_mul_lp(st, nt, r, dinv, z));
		TRY(nullspace_lp(c, l, var, z));
		TRY(ssxk_copy_lp(st, p, z, nt));
		TRY(ssxk_dot_lp(st, nt, r, L->scale_lp, z, part, 0));
	}
	return ssxk_reduce(st, part, 0, scal, o + CG_RDZ);
}

/* conjugate_gradient (solver.c:33-81) in FP32 storage: the fused six-kernel
 * iteration of cg_level_fused with FP64 scalars and inner products */
static int cg_level_lp(ssx_ctx *c, int l, int op, int var, float *f,
		float *u, double eps, int max_iter, int *iters)
{
	ssx_level *L = &c->lev[l];
	const size_t nt = L->ntot;
	float *r = L->r_lp, *z = L->z_lp, *p = L->p_lp, *ap = L->ap_lp;
	void *st = c->stream;
	const int pres = var == SSX_PRESSURE;
	const float *dinv = jdiag_lp(c, l, var);
	const int o = SC_CG(l);
	int *flags = c->cg_flags + 2 * l;
	TRY(ssx_cg_start_lp(c, l, op, var, f, u, c->part, c->scal, o));
	TRY(ssxk_set_ints(st, flags, 0, 2));
	const long long nba = c->nel;   /* operator partials: one per element */
	/* The solution step stays in the update here: taking it in the next
	 * operator launch (as cg_level_fused does) made the FP32 and BF16
	 * smoothers slower (32^3 n = 8 solve 0.256 -> 0.266 s; measured again
	 * with the operator's loads pipelined: FP32 0.193 -> 0.197 s, BF16
	 * factors 0.183 -> 0.185 s, bit-identical results) */
	const int defer = 0;
	for (int k = 0; k < max_iter; ++k) {
		if (k == 0)
			TRY(ssxk_ax_lp(st, L->lp_prec, L->n, L->axv[L->lp_prec], c->nel,
					op, p, ap, L->gf_lp,
					L->mass_lp, L->dmat_lp, c->lambda, L->bc, !pres, flags,
					c->part));
		else
			TRY(ssxk_ax_pupd_lp(st, L->lp_prec, L->n, L->axv[L->lp_prec],
					c->nel, op, z, p,
					c->scal, o + CG_BETA, pres ? o + CG_NULLC : -1, ap,
					L->gf_lp, L->mass_lp, L->dmat_lp, c->lambda, L->bc, !pres,
					flags, c->part, defer ? u : NULL, o + CG_ALPHA));
		if (c->fused_red && ssxk_dss_blocks(L->ngrp) > 0) {
			const ssx_lastred lr0 = { c->redctr, c->part, nba, c->scal, o,
				max_iter, pres, flags, eps, (double) L->nunique };
			const ssx_lastred lr1 = { c->redctr + 1, c->part, SSX_NPART,
				c->scal, o, max_iter, pres, flags, eps, (double) L->nunique };
			TRY(ssxk_dss_all_red_lp(st, L->grp, L->ngrp, ap, NULL, NULL,
					flags, &lr0));
			TRY(ssxk_cg_update_red_lp(st, nt, c->scal, o + CG_ALPHA, p, ap, u,
					r, z, dinv, L->scale_lp, pres, c->part, flags, &lr1));
			continue;
		}
		if (c->gather_upd && L->gP > 0) {
			TRY(ssxk_cg_reduce(st, 0, c->part, nba, c->scal, o, flags, eps,
					max_iter, L->nunique, pres));
			TRY(ssxk_cg_update_g_lp(st, nt, L->n3, L->gP, L->gsoff, L->gtab,
					c->scal, o + CG_ALPHA, p, ap, defer ? NULL : u, r, z,
					dinv, pres, c->part, flags));
		} else {
			TRY(ssxk_dss_all_lp(st, L->grp, L->ngrp, ap, NULL, NULL, 0,
					flags));
			TRY(ssxk_cg_reduce(st, 0, c->part, nba, c->scal, o, flags, eps,
					max_iter, L->nunique, pres));
			TRY(ssxk_cg_update_lp(st, nt, c->scal, o + CG_ALPHA, p, ap,
					defer ? NULL : u, r, z, dinv, L->scale_lp, pres, c->part,
					flags));
		}
		TRY(ssxk_cg_reduce(st, 1, c->part, 0, c->scal, o, flags, eps,
				max_iter, L->nunique, pres));
	}
	if (defer && max_iter > 0)
		TRY(ssxk_cg_uflush_lp(st, nt, c->scal, o + CG_ALPHA, p, u));
	if (iters) {
		hipStream_t hs = (hipStream_t) st;
		HIPC(hipMemcpyAsync(c->h_flags, flags, 2 * sizeof(int),
				hipMemcpyDeviceToHost, hs));
		HIPC(hipStreamSynchronize(hs));
		*iters = c->h_flags[1];
	}
	return SSX_OK;
}

/* V-cycle of precon.c:238-271 with the level-l problem held in
 * lev[l].fn_lp (right-hand side) and lev[l].e_lp (solution) */
static int mg_recurse_lp(ssx_ctx *c, int l, int op, int var, double eps)
{
	ssx_level *L = &c->lev[l];
	void *st = c->stream;
	if (l == c->nlev - 1) {          /* coarsest level in FP64 */
		double err;
		int it;
		/* Right-hand side: FP64 and DSS-averaged by the caller.  The
		 * correction is handed to FP32 levels, so the solve stops at
		 * max(eps, lp_coarse_rtol * |f|_w): below that its digits are
		 * lost in the conversion, and FP32 rounding of the fine residual
		 * leaves a component in the null space of singular problems
		 * (periodic velocity) that CG cannot remove. */
		TRY(ssxk_fill(st, L->mg_e, 0.0, L->ntot));
		if (c->coarse_csr && L->cs.nu > 0)
			TRY(ssx_coarse_cg(c, l, op, var, L->mg_fn, L->mg_e, eps,
					c->lp_coarse_rtol, 500, 1, &it, &err));
		else
			TRY(cg_level_fused(c, l, op, var, L->mg_fn, L->mg_e, eps,
					c->lp_coarse_rtol, 500, 1, &it, &err));
		c->cnt[1] += it;
		return ssxk_cvt(st, SSX_PREC_FP64, SSX_PREC_FP32, L->mg_e, L->e_lp,
				L->ntot);
	}
	ssx_level *C = &c->lev[l + 1];
	float *r = L->mgr_lp;
	if (l == 0 && c->gather_upd && L->gP > 0) {
		TRY(opeval_raw_lp(c, l, op, var, L->e_lp));
		TRY(ssxk_resid_g_lp(st, 0, L->ntot, L->n3, L->gP, L->gsoff, L->gtab,
				L->ap_lp, L->fn_lp, NULL, r, NULL, NULL, NULL, NULL));
	} else if (l == 0) {
		TRY(opeval_lp(c, l, op, var, L->e_lp, r));
		TRY(ssxk_axpby_lp(st, L->ntot, 1.0, L->fn_lp, NULL, -1.0, r));
	} else {
		TRY(ssxk_copy_lp(st, r, L->fn_lp, L->ntot));
	}
	TRY(ssxk_restrict_lp(st, L->n, C->n, c->nel, L->prol_lp, r, C->fn_lp));
	if (l + 1 == c->nlev - 1) {
		/* The coarsest solve runs in FP64 to eps.  Its right-hand side is
		 * averaged over the copies of each shared node in FP64: an FP32
		 * average leaves the copies unequal at FP32 rounding, a part of
		 * the residual that no assembled correction removes, and on the
		 * singular pressure problem CG then stalls above eps and breaks
		 * down. */
		TRY(ssxk_cvt(st, SSX_PREC_FP32, SSX_PREC_FP64, C->fn_lp, C->mg_fn,
				C->ntot));
		TRY(ssxk_dss_all(st, C->grp, C->ngrp, C->mg_fn, NULL, NULL, 1,
				NULL));
	} else {
		TRY(ssxk_dss_all_lp(st, C->grp, C->ngrp, C->fn_lp, NULL, NULL, 1,
				NULL));
	}
	TRY(ssxk_fill_lp(st, C->e_lp, 0.0, C->ntot));
	TRY(mg_recurse_lp(c, l + 1, op, var, eps));
	TRY(ssxk_prolong_lp(st, L->n, C->n, c->nel, L->prol_lp, C->e_lp,
			L->e_lp, 1));
	int it = 15;
	TRY(cg_level_lp(c, l, op, var, L->fn_lp, L->e_lp, eps, 15, NULL));
	c->cnt[2] += it;
	return SSX_OK;
}

static int mg_precon_lp(ssx_ctx *c, int op, int var, double eps,
		const double *x, double *px)
{
	TRY(build_lp(c, c->smooth_prec));
	ssx_level *L = &c->lev[0];
	void *st = c->stream;
	c->cnt[3]++;
	/* FP32 right-hand side and the Jacobi guess in one pass */
	TRY(ssxk_d2f_jacobi(st, L->ntot, x, L->fn_lp, jdiag_lp(c, 0, var),
			L->e_lp));
	TRY(mg_recurse_lp(c, 0, op, var, eps));
	/* the FP32 smoother leaves the copies of a shared node unequal at FP32
	 * rounding; averaging them in FP64 returns an assembled vector, as the
	 * FP64 cycle does, so the FCG iterate stays continuous.  With the
	 * partner lists the conversion and the average are one gather pass. */
	if (c->gather_upd && L->gP > 0)
		return ssxk_f2d_average_g(st, L->ntot, L->n3, L->gP, L->gsoff,
				L->gtab, L->e_lp, px);
	TRY(ssxk_cvt(st, SSX_PREC_FP32, SSX_PREC_FP64, L->e_lp, px, L->ntot));
	return ssxk_dss_all(st, L->grp, L->ngrp, px, NULL, NULL, 1, NULL);
}

/* ------------------------------------------------------------------ */
/* the V-cycle in stages, for batches that run the smoother and coarse  */
/* CG of many problems in shared launches: the same kernels as          */
/* mg_recurse / mg_recurse_lp, cut where those call the level CG        */
/* ------------------------------------------------------------------ */

int ssx_lp_ready(ssx_ctx *c, int prec)
{
	return build_lp(c, prec);
}

const double *ssx_jdiag(ssx_ctx *c, int l, int var)
{
	return jdiag(c, l, var);
}

const float *ssx_jdiag_lp(ssx_ctx *c, int l, int var)
{
	return jdiag_lp(c, l, var);
}

/* initial guess of the cycle (precon.c:229-235): px = D^-1 x; reduced
 * precision: FP32 right-hand side and guess on level 0 */
int ssx_vc_begin(ssx_ctx *c, int prec, int var, const double *x, double *px)
{
	ssx_level *L = &c->lev[0];
	c->cnt[3]++;
	if (prec == SSX_PREC_FP64)
		return ssx_jacobi(c, 0, var, x, px);
	return ssxk_d2f_jacobi(c->stream, L->ntot, x, L->fn_lp,
			jdiag_lp(c, 0, var), L->e_lp);
}

/* descent from level l (not the coarsest) to l + 1: the residual (level 0)
 * or the level's right-hand side, restricted and DSS-averaged, and a zero
 * initial error on l + 1 (precon.c:246-260) */
int ssx_vc_down(ssx_ctx *c, int prec, int l, int op, int var,
		const double *x, const double *px)
{
	ssx_level *L = &c->lev[l], *C = &c->lev[l + 1];
	void *st = c->stream;
	if (prec == SSX_PREC_FP64) {
		double *r = L->mg_r;
		if (l == 0 && c->gather_upd && L->gP > 0) {   /* precon.c:302-309 */
			TRY(opeval_raw(c, l, op, var, px));
			TRY(ssxk_resid_g(st, 0, L->ntot, L->n3, L->gP, L->gsoff,
					L->gtab, L->ap, (double *) x, NULL, r, NULL, NULL,
					NULL, NULL));
		} else if (l == 0) {
			TRY(ssx_opeval(c, l, op, var, px, r));
			TRY(ssxk_axpby(st, L->ntot, 1.0, x, NULL, -1.0, r));
		} else {
			TRY(ssxk_copy(st, r, L->mg_fn, L->ntot));
		}
		TRY(ssx_restrict(c, l + 1, r, C->mg_fn));
		return ssxk_fill(st, C->mg_e, 0.0, C->ntot);
	}
	float *r = L->mgr_lp;
	if (l == 0 && c->gather_upd && L->gP > 0) {
		TRY(opeval_raw_lp(c, l, op, var, L->e_lp));
		TRY(ssxk_resid_g_lp(st, 0, L->ntot, L->n3, L->gP, L->gsoff, L->gtab,
				L->ap_lp, L->fn_lp, NULL, r, NULL, NULL, NULL, NULL));
	} else if (l == 0) {
		TRY(opeval_lp(c, l, op, var, L->e_lp, r));
		TRY(ssxk_axpby_lp(st, L->ntot, 1.0, L->fn_lp, NULL, -1.0, r));
	} else {
		TRY(ssxk_copy_lp(st, r, L->fn_lp, L->ntot));
	}
	TRY(ssxk_restrict_lp(st, L->n, C->n, c->nel, L->prol_lp, r, C->fn_lp));
	if (l + 1 == c->nlev - 1) {
		/* the coarsest right-hand side in FP64, averaged in FP64 (as
		 * mg_recurse_lp) */
		TRY(ssxk_cvt(st, SSX_PREC_FP32, SSX_PREC_FP64, C->fn_lp, C->mg_fn,
				C->ntot));
		TRY(ssxk_dss_all(st, C->grp, C->ngrp, C->mg_fn, NULL, NULL, 1,
				NULL));
		return ssxk_fill(st, C->mg_e, 0.0, C->ntot);
	}
	TRY(ssxk_dss_all_lp(st, C->grp, C->ngrp, C->fn_lp, NULL, NULL, 1, NULL));
	return ssxk_fill_lp(st, C->e_lp, 0.0, C->ntot);
}

/* after the coarsest solve (reduced precision): its FP64 correction in
 * FP32 for the level above */
int ssx_vc_coarse_done(ssx_ctx *c, int prec)
{
	ssx_level *L = &c->lev[c->nlev - 1];
	if (prec == SSX_PREC_FP64 || c->nlev < 2)
		return SSX_OK;
	return ssxk_cvt(c->stream, SSX_PREC_FP64, SSX_PREC_FP32, L->mg_e,
			L->e_lp, L->ntot);
}

/* ascent to level l: the prolonged correction of l + 1 added to the
 * level's solution (precon.c:262-263) */
int ssx_vc_up(ssx_ctx *c, int prec, int l, double *px)
{
	ssx_level *L = &c->lev[l], *C = &c->lev[l + 1];
	if (prec == SSX_PREC_FP64)
		return ssxk_prolong(c->stream, L->n, C->n, c->nel, L->prol, C->mg_e,
				l == 0 ? px : L->mg_e, 1);
	return ssxk_prolong_lp(c->stream, L->n, C->n, c->nel, L->prol_lp,
			C->e_lp, L->e_lp, 1);
}

/* end of the cycle (reduced precision): the FP32 result averaged in FP64
 * into px, an assembled vector (as mg_precon_lp) */
int ssx_vc_end(ssx_ctx *c, int prec, double *px)
{
	ssx_level *L = &c->lev[0];
	void *st = c->stream;
	if (prec == SSX_PREC_FP64)
		return SSX_OK;
	if (c->gather_upd && L->gP > 0)
		return ssxk_f2d_average_g(st, L->ntot, L->n3, L->gP, L->gsoff,
				L->gtab, L->e_lp, px);
	TRY(ssxk_cvt(st, SSX_PREC_FP32, SSX_PREC_FP64, L->e_lp, px, L->ntot));
	return ssxk_dss_all(st, L->grp, L->ngrp, px, NULL, NULL, 1, NULL);
}

int ssx_mg_precon(ssx_ctx *c, int op, int var, double eps, const double *x,
		double *px)
{
	TRY(chk(c, 0));
	/* a single level is the coarsest, which is FP64 in every mode */
	if (c->smooth_prec != SSX_PREC_FP64 && c->nlev > 1)
		return mg_precon_lp(c, op, var, eps, x, px);
	c->cnt[3]++;
	TRY(ssx_jacobi(c, 0, var, x, px));      /* precon.c:230 */
	return mg_recurse(c, 0, op, var, (double *) x, px, eps);
}

int ssx_cg(ssx_ctx *c, int l, int op, int var, double *f, double *u,
		double eps, int max_iter, ssx_stats *stats)
{
	TRY(chk(c, l));
	long long c0[4];
	memcpy(c0, c->cnt, sizeof(c0));
	double t0 = now_s(), err;
	int it = 0;
	int rc;
	if (c->coarse_csr && l == c->nlev - 1 && c->lev[l].cs.nu > 0) {
		rc = ssx_mask(c, l, var, f);               /* solver.c:46 */
		if (rc == SSX_OK)
			rc = ssx_coarse_cg(c, l, op, var, f, u, eps, 0.0, max_iter, 0,
					&it, &err);
	} else {
		rc = cg_level(c, l, op, var, f, u, eps, max_iter, 0, &it, &err);
	}
	if (rc == SSX_OK)
		rc = ssx_synchronize(c);
	if (stats) {
		memset(stats, 0, sizeof(*stats));
		stats->iterations = it;
		stats->residual = err;
		stats->converged = err <= eps;
		stats->fine_ax = c->cnt[0] - c0[0];
		stats->seconds = now_s() - t0;
	}
	if (rc == SSX_OK && !(err <= eps))
		rc = SSX_ENOTCONV;
	return rc;
}

/* flexible_conjugate_gradient (solver.c:83-157) in three phases, so a
 * batch can interleave many problems: begin (history buffers, f masked,
 * r = f - A u, null space), then per iteration the preconditioner into
 * fcg_z followed by ssx_fcg_iter (orthogonalisation, direction, A p, step
 * and <r,r>_w into scal[SC_RR]; enqueued only), and end (the rotated
 * history handed back to the context). */
int ssx_fcg_begin(ssx_ctx *c, int op, int var, int precon, double *f,
		const double *u, ssx_fcg_state *s)
{
	TRY(chk(c, 0));
	const int mmax = precon == SSX_PC_MG ? 10 : 1;   /* solver.c:90-92 */
	const size_t nt = c->lev[0].ntot;
	s->mmax = 0;             /* set once the history is complete */
	for (int i = 0; i < mmax + 1; ++i)
		if (!c->fcg_p[i])
			TRY(dalloc(&c->fcg_p[i], nt));
	for (int i = 0; i < mmax; ++i)
		if (!c->fcg_ap[i])
			TRY(dalloc(&c->fcg_ap[i], nt));
	for (int i = 0; i < 11; ++i)
		s->p[i] = c->fcg_p[i];
	for (int i = 0; i < 10; ++i)
		s->ap[i] = c->fcg_ap[i];
	s->mmax = mmax;
	s->j = 0;
	TRY(ssx_mask(c, 0, var, f));
	TRY(ssx_opeval(c, 0, op, var, u, c->fcg_r));
	TRY(ssxk_axpby(c->stream, nt, 1.0, f, NULL, -1.0, c->fcg_r));
	return ssx_nullspace(c, 0, var, c->fcg_r);
}

int ssx_fcg_iter(ssx_ctx *c, int op, int var, double *u, ssx_fcg_state *s)
{
	ssx_level *L = &c->lev[0];
	const size_t nt = L->ntot;
	void *st = c->stream;
	double **p = s->p, **ap = s->ap;
	double *r = c->fcg_r, *z = c->fcg_z;
	const double *mask = var == SSX_PRESSURE ? NULL : L->mask;
	const long j = s->j;
	const int mmax = s->mmax;
	int m = (int) j, mm = (int) (j % (mmax + 1));
	if (mm < 1)
		mm = 1;
	if (mm < m)
		m = mm;
	TRY(ssx_nullspace(c, 0, var, z));
	/* beta_k = <z, Ap_k>_w / <p_k, Ap_k>_w, k < m, in one pass */
	TRY(ssxk_multidot(st, nt, z, L->scale, (const double *const *) ap,
			m, c->part, c->scal, SC_ZAP0, SC_PAP0, SC_BETA0));
	double *tp = p[m];                 /* p[k] -> p[k+1], k < m */
	for (int k = m; k > 0; --k)
		p[k] = p[k - 1];
	p[0] = tp;
	/* p_0 = mask (z - sum_k beta_k p_{k+1}), one pass */
	TRY(ssxk_multi_axpy(st, nt, z, (const double *const *) (p + 1), m,
			c->scal, SC_BETA0, mask, p[0]));
	int m2 = (int) (j + 1), mm2 = (int) ((j + 1) % (mmax + 1));
	if (mm2 < 1)
		mm2 = 1;
	if (mm2 < m2)
		m2 = mm2;
	double *ta = ap[m2 - 1];           /* ap[k] -> ap[k+1], k < m2-1 */
	for (int k = m2 - 1; k > 0; --k)
		ap[k] = ap[k - 1];
	ap[0] = ta;
	TRY(ssx_opeval(c, 0, op, var, p[0], ap[0]));
	for (int k = m2 - 1; k > 0; --k)
		TRY(ssxk_scalar(st, c->scal, 2, SC_PAP0 + k, SC_PAP0 + k - 1, 0,
				0.0));
	/* <p_0, Ap_0>_w and <p_0, r>_w in one pass */
	const double *pr2[2] = { ap[0], r };
	TRY(ssxk_multidot(st, nt, p[0], L->scale, pr2, 2, c->part, c->scal,
			SC_PAPN, -1, 0));
	TRY(ssxk_scalar(st, c->scal, 2, SC_PAP0, SC_PAPN, 0, 0.0));
	TRY(ssxk_scalar(st, c->scal, 0, SC_ALPHA, SC_PAPN + 1, SC_PAPN, 0.0));
	/* u += alpha p_0; r -= alpha Ap_0; <r,r>_w */
	TRY(ssxk_fcg_step(st, nt, c->scal, SC_ALPHA, p[0], ap[0], u, r,
			L->scale, c->part));
	TRY(ssxk_reduce(st, c->part, 0, c->scal, SC_RR));
	s->j = j + 1;
	return SSX_OK;
}

void ssx_fcg_end(ssx_ctx *c, const ssx_fcg_state *s)
{
	if (s->mmax == 0)        /* ssx_fcg_begin did not complete */
		return;
	/* keep the rotated history buffers owned by the context */
	for (int i = 0; i < 11; ++i)
		c->fcg_p[i] = s->p[i];
	for (int i = 0; i < 10; ++i)
		c->fcg_ap[i] = s->ap[i];
}

int ssx_fcg(ssx_ctx *c, int op, int var, int precon, double *f, double *u,
		double eps, int max_iter, ssx_stats *stats)
{
	TRY(chk(c, 0));
	ssx_fcg_state s;
	s.mmax = 0;
	long long c0[4];
	memcpy(c0, c->cnt, sizeof(c0));
	double t0 = now_s();
	int rc = ssx_fcg_begin(c, op, var, precon, f, u, &s);
	if (rc != SSX_OK) {
		ssx_fcg_end(c, &s);
		return rc;
	}
	int it = 0;
	double err = 0.0;
	do {
		if (precon == SSX_PC_MG)
			rc = ssx_mg_precon(c, op, var, eps, c->fcg_r, c->fcg_z);
		else
			rc = ssx_jacobi(c, 0, var, c->fcg_r, c->fcg_z);
		if (rc == SSX_OK)
			rc = ssx_fcg_iter(c, op, var, u, &s);
		double rr = 0.0;
		if (rc == SSX_OK)
			rc = read_scal(c, SC_RR, &rr);
		if (rc != SSX_OK)
			break;
		err = sqrt(rr);
		++it;
		if (isnan(err) || isinf(err)) {
			rc = SSX_ENUMERIC;
			ssx_set_error("FCG: residual is %g", err);
			break;
		}
	} while (err > eps && s.j < max_iter);
	ssx_fcg_end(c, &s);
	if (rc == SSX_OK)
		rc = ssx_synchronize(c);
	if (stats) {
		memset(stats, 0, sizeof(*stats));
		stats->iterations = it;
		stats->residual = err;
		stats->converged = err <= eps;
		stats->fine_ax = c->cnt[0] - c0[0];
		stats->coarse_iters = c->cnt[1] - c0[1];
		stats->smooth_iters = c->cnt[2] - c0[2];
		stats->vcycles = c->cnt[3] - c0[3];
		stats->seconds = now_s() - t0;
	}
	if (rc == SSX_OK && !(err <= eps))
		rc = SSX_ENOTCONV;
	return rc;
}

int ssx_dss_host(const ssx_mesh *m, int n, double *u)
{
	if (!m || !u || n < SSX_NMIN || n > SSX_NMAX || m->nel < 1)
		return SSX_EINVAL;
	int *grp[3];
	long long ng[3];
	TRY(ssx_build_groups(m, n, grp, ng));
	ssx_host_dss(grp, ng, u);
	for (int i = 0; i < 3; ++i)
		free(grp[i]);
	return SSX_OK;
}

int ssx_dss_gather_host(const ssx_mesh *m, int n, double *u)
{
	if (!m || !u || n < SSX_NMIN || n > SSX_NMAX || m->nel < 1)
		return SSX_EINVAL;
	const int n3 = n * n * n;
	const size_t nt = (size_t) m->nel * n3;
	int *tab, *soff, P;
	TRY(ssx_build_gather(m, n, &tab, &soff, &P));
	double *v = malloc(sizeof(double) * nt);
	if (!v) {
		free(tab);
		free(soff);
		return SSX_ENOMEM;
	}
	memcpy(v, u, sizeof(double) * nt);
	for (size_t q = 0; q < nt; ++q) {
		const int code = soff[q % n3];
		const int *pt = tab + (q / n3) * (size_t) P + (code >> 3);
		double s = v[q];
		for (int k = 0; k < (code & 7) && pt[k] >= 0; ++k)
			s = s + v[pt[k]];
		u[q] = s;
	}
	free(v);
	free(tab);
	free(soff);
	return SSX_OK;
}

int ssx_geometry_host(const ssx_mesh *m, double *gfac, double *mass)
{
	if (!m || !gfac || !mass || m->n < SSX_NMIN || m->n > SSX_NMAX
			|| m->nel < 1)
		return SSX_EINVAL;
	const size_t nt = (size_t) m->nel * m->n * m->n * m->n;
	double *dp = malloc(sizeof(double) * nt), *dh = malloc(sizeof(double) * nt);
	if (!dp || !dh) {
		free(dp);
		free(dh);
		return SSX_ENOMEM;
	}
	ssx_basis b;
	ssx_basis_init(&b, m->n, m->svvn, m->svva);
	level_geometry(&b, m->nel, m->x, m->y, m->z, 0.0, gfac, mass, dp, dh);
	free(dp);
	free(dh);
	return SSX_OK;
}

/* host-array entry points, solver.c:159-189 */
static int host_solve(ssx_ctx *c, double *f, double *u, int var, double eps,
		int helm, ssx_stats *stats)
{
	const size_t nt = c->lev[0].ntot;
	if (!c->d_f)
		TRY(dalloc(&c->d_f, nt));
	if (!c->d_u)
		TRY(dalloc(&c->d_u, nt));
	hipStream_t st = (hipStream_t) c->stream;
	HIPC(hipMemcpyAsync(c->d_f, f, nt * sizeof(double),
			hipMemcpyHostToDevice, st));
	HIPC(hipMemcpyAsync(c->d_u, u, nt * sizeof(double),
			hipMemcpyHostToDevice, st));
	int rc = helm ? ssx_cg(c, 0, SSX_OP_HELMHOLTZ, var, c->d_f, c->d_u, eps,
			5000, stats)
		: ssx_fcg(c, SSX_OP_POISSON, var, SSX_PC_MG, c->d_f, c->d_u, eps,
			5000, stats);
	if (rc != SSX_OK && rc != SSX_ENOTCONV)
		return rc;
	HIPC(hipMemcpyAsync(f, c->d_f, nt * sizeof(double),
			hipMemcpyDeviceToHost, st));
	HIPC(hipMemcpyAsync(u, c->d_u, nt * sizeof(double),
			hipMemcpyDeviceToHost, st));
	HIPC(hipStreamSynchronize(st));
	return rc;
}

int ssx_solve_poisson_cg(ssx_ctx *c, double *f, double *u, int var,
		double eps, ssx_stats *stats)
{
	TRY(chk(c, 0));
	return host_solve(c, f, u, var, eps, 0, stats);
}

int ssx_solve_helmholtz_cg(ssx_ctx *c, double *f, double *u, int var,
		double eps, ssx_stats *stats)
{
	TRY(chk(c, 0));
	return host_solve(c, f, u, var, eps, 1, stats);
}

/* ------------------------------------------------------------------ */
/* particles                                                            */
/* ------------------------------------------------------------------ */

/* J and J^-1 of level 0 in the reference's operation order
 * (deriv.c:68-79 derivatives summed in index order, tripalg.c:165-222
 * determinant and cofactors), so that particle positions and owning
 * elements follow the reference bit for bit */
/* J (jac[3r+s] = D_s x_r), J^-1 and (det nullable) det J at every node of
 * a basis, with the reference's cofactor formulas and operation order
 * (compute_jacobian deriv.c:68-79, compute_trimat_determinant
 * tripalg.c:165-182, invert_trimat :201-222) */
void ssx_jacobians(const ssx_basis *b, int nel, const double *x,
		const double *y, const double *z, double *const jac[9],
		double *const jinv[9], double *det_out)
{
	const int n = b->n, n3 = n * n * n;
	const double *D = b->d;
	static const int k0[3] = { 1, 0, 0 }, k1[3] = { 2, 2, 1 },
		k2[3] = { 2, 0, 1 }, k3[3] = { 1, 2, 0 };
	static const double sg[3] = { 1.0, -1.0, 1.0 };
	for (int e = 0; e < nel; ++e) {
		const size_t off = (size_t) e * n3;
		const double *c[3] = { x + off, y + off, z + off };
		for (int k = 0; k < n; ++k)
		for (int j = 0; j < n; ++j)
		for (int i = 0; i < n; ++i) {
			const int q = (k * n + j) * n + i;
			double a[3][3], ai[3][3];
			for (int r = 0; r < 3; ++r) {
				double s0 = 0.0, s1 = 0.0, s2 = 0.0;
				for (int l = 0; l < n; ++l) {
					s0 += D[i * n + l] * c[r][(k * n + j) * n + l];
					s1 += c[r][(k * n + l) * n + i] * D[j * n + l];
					s2 += c[r][(l * n + j) * n + i] * D[k * n + l];
				}
				a[r][0] = s0;
				a[r][1] = s1;
				a[r][2] = s2;
			}
			double t, det;
			t = a[1][1] * a[2][2];
			t = -a[1][2] * a[2][1] + t;
			det = a[0][0] * t;
			t = a[1][0] * a[2][2];
			t = -a[1][2] * a[2][0] + t;
			det = -a[0][1] * t + det;
			t = a[1][0] * a[2][1];
			t = -a[1][1] * a[2][0] + t;
			det = a[0][2] * t + det;
			for (int ii = 0; ii < 3; ++ii)
				for (int jj = 0; jj < 3; ++jj) {
					double u = a[k0[jj]][k3[ii]] * a[k1[jj]][k2[ii]];
					u = -a[k0[jj]][k2[ii]] * a[k1[jj]][k3[ii]] + u;
					ai[ii][jj] = sg[jj] * u / det;
				}
			for (int r = 0; r < 3; ++r)
				for (int s = 0; s < 3; ++s) {
					jac[r * 3 + s][off + q] = a[r][s];
					jinv[r * 3 + s][off + q] = ai[r][s];
				}
			if (det_out)
				det_out[off + q] = det;
		}
	}
}

static int particles_init(ssx_ctx *c)
{
	if (c->p_jac[0])
		return SSX_OK;
	ssx_level *L = &c->lev[0];
	const size_t nt = L->ntot;
	double *hj[9], *hi[9];
	int rc = SSX_OK;
	for (int a = 0; a < 9; ++a) {
		hj[a] = malloc(sizeof(double) * nt);
		hi[a] = malloc(sizeof(double) * nt);
		if (!hj[a] || !hi[a])
			rc = SSX_ENOMEM;
	}
	if (rc == SSX_OK) {
		ssx_jacobians(&L->b, c->nel, L->h_x, L->h_y, L->h_z, hj, hi, NULL);
		for (int a = 0; a < 9 && rc == SSX_OK; ++a) {
			rc = up(&c->p_jac[a], hj[a], nt);
			if (!rc)
				rc = up(&c->p_jinv[a], hi[a], nt);
			if (!rc)
				rc = dalloc(&c->p_ur[a / 3][a % 3], nt);
			if (!rc)
				HIPC(hipMemset(c->p_ur[a / 3][a % 3], 0, nt * sizeof(double)));
		}
	}
	for (int a = 0; a < 9; ++a) {
		free(hj[a]);
		free(hi[a]);
	}
	if (rc == SSX_OK)
		rc = up(&c->p_xi, L->b.xi, (size_t) L->n);
	if (rc == SSX_OK) {
		if (hipMalloc((void **) &c->d_neigh, sizeof(int) * 6 * c->nel)
				!= hipSuccess || hipMalloc((void **) &c->p_maxpass,
				2 * sizeof(int)) != hipSuccess) {
			ssx_set_error("hipMalloc particle tables");
			return SSX_ENOMEM;
		}
		HIPC(hipMemcpy(c->d_neigh, c->h_neigh, sizeof(int) * 6 * c->nel,
				hipMemcpyHostToDevice));
		/* the tracer walk enters a neighbour on the face opposite the one
		 * it left: right when the neighbour across face s names this
		 * element across face s^1 (faces W,E,S,N,B,T) */
		c->p_oriented = 1;
		for (int e = 0; e < c->nel && c->p_oriented; ++e)
			for (int f = 0; f < 6; ++f) {
				const int nb = c->h_neigh[6 * e + f];
				if (nb >= 0 && c->h_neigh[6 * nb + (f ^ 1)] != e) {
					c->p_oriented = 0;
					break;
				}
			}
	}
	if (rc != SSX_OK)
		ssx_set_error("particle set-up failed");
	return rc;
}

int ssx_particles_set_velocity(ssx_ctx *c, const double *u, const double *v,
		const double *w)
{
	TRY(chk(c, 0));
	TRY(particles_init(c));
	/* time_shift (time.c:37-41): now -> old -> older, by rotation */
	for (int a = 0; a < 3; ++a) {
		double *older = c->p_ur[2][a];
		c->p_ur[2][a] = c->p_ur[1][a];
		c->p_ur[1][a] = c->p_ur[0][a];
		c->p_ur[0][a] = older;
	}
	return ssxk_particles_ur(c->stream, c->lev[0].ntot,
			(const double *const *) c->p_jinv, u, v, w, c->p_ur[0][0],
			c->p_ur[0][1], c->p_ur[0][2]);
}

/* one step of the particle walk on device arrays: inertial particles
 * (particles.c) or, with tracers, RK2 tracers (tau not used) */
/* the processing-order buffers for n particles (grown on demand) */
static int particle_order_bufs(ssx_ctx *c, long n)
{
	if (!c->p_bucket && hipMalloc((void **) &c->p_bucket, sizeof(int)
			* (size_t) ssxk_particle_order_ints(c->nel, c->lev[0].n))
			!= hipSuccess) {
		c->p_bucket = NULL;
		ssx_set_error("hipMalloc of the particle buckets");
		return SSX_ENOMEM;
	}
	if (!c->p_pcnt && hipMalloc((void **) &c->p_pcnt,
			2 * sizeof(int) * (SSX_WALK_PASSES + 1)) != hipSuccess) {
		c->p_pcnt = NULL;
		ssx_set_error("hipMalloc of the particle pass counts");
		return SSX_ENOMEM;
	}
	if (c->p_perm_cap >= n)
		return SSX_OK;
	hipFree(c->p_perm);
	hipFree(c->p_work);
	c->p_perm = c->p_work = NULL;
	c->p_perm_cap = 0;
	if (hipMalloc((void **) &c->p_perm, sizeof(int) * (size_t) n)
			!= hipSuccess || hipMalloc((void **) &c->p_work,
			5 * sizeof(int) * (size_t) n) != hipSuccess) {
		hipFree(c->p_perm);
		c->p_perm = NULL;
		ssx_set_error("hipMalloc of the particle order (%ld)", n);
		return SSX_ENOMEM;
	}
	c->p_perm_cap = n;
	return SSX_OK;
}

int ssx_walk_dev(ssx_ctx *c, const ssx_particles *p, double t_now,
		double dt, int *max_passes, int tracers, int *counters)
{
	TRY(chk(c, 0));
	if (!p || p->n < 0 || (p->n > 0 && (!p->el || !p->old_el || !p->x
			|| !p->y || !p->z || !p->u || !p->v || !p->w || !p->t
			|| (!tracers && !p->tau)))) {
		ssx_set_error("bad particle list");
		return SSX_EINVAL;
	}
	if (!(dt > 0.0)) {
		ssx_set_error("particle step: dt must be positive");
		return SSX_EINVAL;
	}
	const int own = counters == NULL;
	if (counters) {
		/* callers with their own counters allocate nothing: the tables
		 * must exist already (ssx_particles_set_velocity sets them up) */
		if (!c->p_maxpass) {
			ssx_set_error("particle step: set the velocity "
					"(ssx_particles_set_velocity) first");
			return SSX_EINVAL;
		}
	} else {
		TRY(particles_init(c));
		counters = c->p_maxpass;
	}
	if (tracers && !c->p_oriented) {
		ssx_set_error("tracer step: a neighbour across face s does not name "
				"the element across the opposite face; the tracer walk "
				"needs consistently oriented elements");
		return SSX_EMESH;
	}
	const double *ur[9];
	for (int a = 0; a < 9; ++a)
		ur[a] = c->p_ur[a / 3][a % 3];
	/* The allocating entries take the particles element by element (a
	 * counting sort into context buffers), one relocation pass per launch
	 * with the unfinished particles compacted between passes (option 1,
	 * the default; 0: the whole walk in one launch).  Callers with their
	 * own counters allocate nothing and take the particles as stored, the
	 * whole walk in one launch.  The results are the same bits in every
	 * form. */
	const int *perm = NULL;
	TRY(ssxk_set_ints(c->stream, counters, 0, 2));
	if (own && p->n > 0) {
		TRY(particle_order_bufs(c, p->n));
		TRY(ssxk_particle_order(c->stream, p->n, p->el, p->z, c->p_xi,
				c->nel, c->lev[0].n, c->p_bucket, c->p_perm));
		perm = c->p_perm;
		if (c->p_elem_walk > 0) {
			c->p_pass_np = p->n;
			if (tracers)
				TRY(ssxk_tracers_passes(c->stream, c->lev[0].n, c->nel,
						c->p_xi, ur, c->d_neigh, t_now, dt, p->n, perm,
						c->p_work, c->p_pcnt, p->el, p->old_el, p->x, p->y,
						p->z, p->u, p->v, p->w, p->t, p->reinit, counters));
			else
				TRY(ssxk_particles_passes(c->stream, c->lev[0].n, c->nel,
						c->p_xi, ur, (const double *const *) c->p_jac,
						(const double *const *) c->p_jinv, c->d_neigh,
						t_now, dt, p->n, perm, c->p_bucket, c->p_work,
						c->p_pcnt, p->el,
						p->old_el, p->x, p->y, p->z, p->u, p->v, p->w, p->t,
						p->tau, p->reinit, counters));
			goto launched;
		}
	}
	if (tracers)
		TRY(ssxk_tracers_step(c->stream, c->lev[0].n, c->nel, c->p_xi, ur,
				c->d_neigh, t_now, dt, p->n, perm, p->el, p->old_el, p->x,
				p->y, p->z, p->u, p->v, p->w, p->t, p->reinit, counters));
	else
		TRY(ssxk_particles_step(c->stream, c->lev[0].n, c->nel, c->p_xi, ur,
				(const double *const *) c->p_jac,
				(const double *const *) c->p_jinv, c->d_neigh, t_now, dt,
				p->n, perm, p->el, p->old_el, p->x, p->y, p->z, p->u, p->v,
				p->w, p->t, p->tau, p->reinit, counters));
launched:;
	int fl[2] = { 0, 0 };
	HIPC(hipMemcpyAsync(fl, counters, sizeof(fl), hipMemcpyDeviceToHost,
			(hipStream_t) c->stream));
	HIPC(hipStreamSynchronize((hipStream_t) c->stream));
	if (max_passes)
		*max_passes = fl[0];
	if (fl[1]) {
		ssx_set_error("%d particles name no element of the mesh (skipped)",
				fl[1]);
		return SSX_EINVAL;
	}
	return SSX_OK;
}

int ssx_particles_pass_counts(ssx_ctx *c, int counts[SSX_PASS_COUNTS])
{
	TRY(chk(c, 0));
	if (!counts) {
		ssx_set_error("ssx_particles_pass_counts: null array");
		return SSX_EINVAL;
	}
	memset(counts, 0, SSX_PASS_COUNTS * sizeof(int));
	if (!c->p_pcnt || !c->p_pass_np)
		return SSX_OK;
	int d[SSX_WALK_PASSES + 1];
	hipStream_t st = (hipStream_t) c->stream;
	HIPC(hipMemcpyAsync(d, c->p_pcnt, sizeof(d), hipMemcpyDeviceToHost, st));
	HIPC(hipStreamSynchronize(st));
	/* d[k] counts the particles that entered pass k, k >= 1 */
	counts[0] = (int) c->p_pass_np;
	for (int k = 1; k < SSX_PASS_COUNTS && k < SSX_WALK_PASSES; ++k)
		counts[k] = d[k];
	return SSX_OK;
}

int ssx_particles_step(ssx_ctx *c, const ssx_particles *p, double t_now,
		double dt, int *max_passes)
{
	return ssx_walk_dev(c, p, t_now, dt, max_passes, 0, NULL);
}

int ssx_tracers_step(ssx_ctx *c, const ssx_particles *p, double t_now,
		double dt, int *max_passes)
{
	return ssx_walk_dev(c, p, t_now, dt, max_passes, 1, NULL);
}

/* host-array variants: stage through device memory held by the context */
static int stage_bytes(ssx_ctx *c, size_t bytes)
{
	if (c->p_stage_bytes >= bytes)
		return SSX_OK;
	hipFree(c->p_stage);
	c->p_stage = NULL;
	c->p_stage_bytes = 0;
	if (hipMalloc(&c->p_stage, bytes ? bytes : 1) != hipSuccess) {
		c->p_stage = NULL;
		ssx_set_error("hipMalloc of %zu bytes of particle staging", bytes);
		return SSX_ENOMEM;
	}
	c->p_stage_bytes = bytes;
	return SSX_OK;
}

int ssx_particles_set_velocity_host(ssx_ctx *c, const double *u,
		const double *v, const double *w)
{
	TRY(chk(c, 0));
	if (!u || !v || !w) {
		ssx_set_error("null velocity array");
		return SSX_EINVAL;
	}
	const size_t nt = c->lev[0].ntot;
	TRY(stage_bytes(c, 3 * nt * sizeof(double)));
	double *d = (double *) c->p_stage;
	hipStream_t st = (hipStream_t) c->stream;
	HIPC(hipMemcpyAsync(d, u, nt * sizeof(double), hipMemcpyHostToDevice,
			st));
	HIPC(hipMemcpyAsync(d + nt, v, nt * sizeof(double),
			hipMemcpyHostToDevice, st));
	HIPC(hipMemcpyAsync(d + 2 * nt, w, nt * sizeof(double),
			hipMemcpyHostToDevice, st));
	TRY(ssx_particles_set_velocity(c, d, d + nt, d + 2 * nt));
	HIPC(hipStreamSynchronize(st));
	return SSX_OK;
}

size_t ssx_walk_stage_bytes(size_t n)
{
	return n * (8 * sizeof(double) + 3 * sizeof(int));
}

int ssx_walk_host(ssx_ctx *c, const ssx_particles *h, double t_now,
		double dt, int *max_passes, int tracers, void *stage,
		size_t stage_cap, int *counters)
{
	TRY(chk(c, 0));
	if (!h || h->n < 0 || (h->n > 0 && (!h->el || !h->old_el || !h->x
			|| !h->y || !h->z || !h->u || !h->v || !h->w || !h->t
			|| (!tracers && !h->tau)))) {
		ssx_set_error("bad particle list");
		return SSX_EINVAL;
	}
	const size_t n = (size_t) h->n;
	if (n == 0) {
		if (max_passes)
			*max_passes = 0;
		return SSX_OK;
	}
	/* layout: 8 double columns x..tau, then el, old_el, reinit */
	if (stage) {
		if (ssx_walk_stage_bytes(n) > stage_cap) {
			ssx_set_error("particle step: %zu particles need %zu bytes of "
					"staging, more than the %zu available", n,
					ssx_walk_stage_bytes(n), stage_cap);
			return SSX_EINVAL;
		}
	} else {
		TRY(stage_bytes(c, ssx_walk_stage_bytes(n)));
		stage = c->p_stage;
	}
	double *dd = (double *) stage;
	int *di = (int *) (dd + 8 * n);
	double *hcol[8] = { h->x, h->y, h->z, h->u, h->v, h->w, h->t, h->tau };
	hipStream_t st = (hipStream_t) c->stream;
	for (int k = 0; k < 8; ++k)
		if (hcol[k])
			HIPC(hipMemcpyAsync(dd + k * n, hcol[k], n * sizeof(double),
					hipMemcpyHostToDevice, st));
	HIPC(hipMemcpyAsync(di, h->el, n * sizeof(int), hipMemcpyHostToDevice,
			st));
	HIPC(hipMemcpyAsync(di + n, h->old_el, n * sizeof(int),
			hipMemcpyHostToDevice, st));
	ssx_particles d = { h->n, di, di + n, dd, dd + n, dd + 2 * n, dd + 3 * n,
		dd + 4 * n, dd + 5 * n, dd + 6 * n, dd + 7 * n, di + 2 * n };
	const int rc = ssx_walk_dev(c, &d, t_now, dt, max_passes, tracers,
			counters);
	if (rc != SSX_OK && rc != SSX_EINVAL)
		return rc;
	/* results come back also when some particles named no element */
	for (int k = 0; k < 7; ++k)
		HIPC(hipMemcpyAsync(hcol[k], dd + k * n, n * sizeof(double),
				hipMemcpyDeviceToHost, st));
	HIPC(hipMemcpyAsync(h->el, di, n * sizeof(int), hipMemcpyDeviceToHost,
			st));
	HIPC(hipMemcpyAsync(h->old_el, di + n, n * sizeof(int),
			hipMemcpyDeviceToHost, st));
	if (h->reinit)
		HIPC(hipMemcpyAsync(h->reinit, di + 2 * n, n * sizeof(int),
				hipMemcpyDeviceToHost, st));
	HIPC(hipStreamSynchronize(st));
	return rc;
}

int ssx_particles_step_host(ssx_ctx *c, const ssx_particles *h, double t_now,
		double dt, int *max_passes)
{
	return ssx_walk_host(c, h, t_now, dt, max_passes, 0, NULL, 0, NULL);
}

int ssx_tracers_step_host(ssx_ctx *c, const ssx_particles *h, double t_now,
		double dt, int *max_passes)
{
	return ssx_walk_host(c, h, t_now, dt, max_passes, 1, NULL, 0, NULL);
}

/* ------------------------------------------------------------------ */
/* operator launch autotuner                                            */
/* ------------------------------------------------------------------ */

/* average time in microseconds of `reps` Poisson operator launches of
 * variant v on level l in precision prec, on the context's stream */
static int time_ax(ssx_ctx *c, int l, int prec, int v, int reps,
		double *us)
{
	ssx_level *L = &c->lev[l];
	hipStream_t st = (hipStream_t) c->stream;
	hipEvent_t e0, e1;
	HIPC(hipEventCreate(&e0));
	HIPC(hipEventCreate(&e1));
	int rc = SSX_OK;
	for (int r = -2; r < reps && rc == SSX_OK; ++r) {
		if (r == 0 && hipEventRecord(e0, st) != hipSuccess)
			rc = SSX_EHIP;
		if (rc)
			break;
		if (prec == SSX_PREC_FP64)
			rc = ssxk_ax(c->stream, L->n, v, c->nel, SSX_OP_POISSON, L->p,
					L->ap, L->gf, L->mass, L->dmat, c->lambda, L->bc, 1, NULL,
					c->part);
		else
			rc = ssxk_ax_lp(c->stream, prec, L->n, v, c->nel, SSX_OP_POISSON,
					L->p_lp, L->ap_lp, L->gf_lp, L->mass_lp, L->dmat_lp,
					c->lambda, L->bc, 1, NULL, c->part);
	}
	float ms = 0.0f;
	if (rc == SSX_OK && (hipEventRecord(e1, st) != hipSuccess
			|| hipEventSynchronize(e1) != hipSuccess
			|| hipEventElapsedTime(&ms, e0, e1) != hipSuccess)) {
		ssx_set_error("autotune: event timing failed");
		rc = SSX_EHIP;
	}
	hipEventDestroy(e0);
	hipEventDestroy(e1);
	*us = 1e3 * ms / reps;
	return rc;
}

int ssx_autotune(ssx_ctx *c, int prec_mask)
{
	if (!c)
		return SSX_EINVAL;
	for (int prec = 0; prec < 3; ++prec) {
		if (!(prec_mask & (1 << prec)))
			continue;
		for (int l = 0; l < c->nlev; ++l) {
			ssx_level *L = &c->lev[l];
			if (prec != SSX_PREC_FP64 && (l == c->nlev - 1
					|| L->lp_prec != prec || !L->p_lp))
				continue;       /* no copies in that precision here */
			/* operands: any finite values */
			if (prec == SSX_PREC_FP64)
				TRY(ssxk_fill(c->stream, L->p, 1.0, L->ntot));
			else
				TRY(ssxk_fill_lp(c->stream, L->p_lp, 1.0, L->ntot));
			/* enough launches for about a millisecond of work */
			const int reps = L->ntot > (1 << 22) ? 5
				: L->ntot > (1 << 18) ? 20 : 50;
			double best = 0.0;
			int bv = 0;
			for (int v = 0; v < SSX_AX_NVARIANTS; ++v) {
				double us;
				TRY(time_ax(c, l, prec, v, reps, &us));
				c->ax_us[l][prec][v] = us;
				if (v == 0 || us < best) {
					best = us;
					bv = v;
				}
			}
			L->axv[prec] = bv;
		}
	}
	return SSX_OK;
}

int ssx_ax_variant(const ssx_ctx *c, int l, int prec,
		double us[SSX_AX_NVARIANTS])
{
	if (!c || l < 0 || l >= c->nlev || prec < 0 || prec > 2)
		return -1;
	if (us)
		for (int v = 0; v < SSX_AX_NVARIANTS; ++v)
			us[v] = c->ax_us[l][prec][v];
	return c->lev[l].axv[prec];
}

int ssx_set_ax_variant(ssx_ctx *c, int l, int prec, int v)
{
	if (!c || l < 0 || l >= c->nlev || prec < 0 || prec > 2 || v < 0
			|| v >= SSX_AX_NVARIANTS) {
		ssx_set_error("bad level, precision or variant");
		return SSX_EINVAL;
	}
	c->lev[l].axv[prec] = v;
	return SSX_OK;
}
