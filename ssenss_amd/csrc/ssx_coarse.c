/*
 * ssx_coarse.c -- the coarsest p-multigrid level solved over unique nodes.
 *
 * At the coarsest level (n = 2: trilinear elements, every node a vertex
 * shared by up to 8 elements) an E-vector holds each unknown up to 8
 * times, and every CG iteration pays the element operator, the DSS and
 * vector updates on that redundant storage.  Here the level's Poisson and
 * Helmholtz operators are assembled once at set-up into one CSR structure
 * over the unique nodes (27 entries per row at n = 2), and the coarse
 * Jacobi-PCG (conjugate_gradient, solver.c:33-81, as called from
 * multigrid_recurse, precon.c:242-244) runs on unique-node vectors:
 *
 *   set-up: f_u = f at one copy of each node (f is averaged over copies),
 *           mask, r = f - A x, null space, z = D^-1 r, p = z, <r,z>
 *   loop:   ap = mask (A p) with p = mask((z - c) + beta p) formed on the
 *           fly (ping-pong p buffers), reduce -> alpha, update u, r, z and
 *           the partials, reduce -> beta and the stopping test
 *
 * On continuous vectors the weighted inner product <x,y>_w of the E-vector
 * form is the plain sum over unique nodes, so this is the same algorithm
 * with the same stopping test; only the order of the sums differs.  The
 * solution is written back to every copy.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "ssx_internal.h"

#define HIPQ(x)                                                              \
	do {                                                                 \
		hipError_t e_ = (x);                                         \
		if (e_ != hipSuccess) {                                      \
			ssx_set_error("%s: %s", #x, hipGetErrorString(e_));   \
			return SSX_EHIP;                                     \
		}                                                            \
	} while (0)
#define TRYQ(x)                                                              \
	do {                                                                 \
		int rc_ = (x);                                               \
		if (rc_ != SSX_OK)                                           \
			return rc_;                                          \
	} while (0)

/* the element operator of the kernels (poisson.c:34-56, helmholtz.c:33-49)
 * on one host element: D_il at d[i*n+l]; ds = SVV matrix for Helmholtz */
static void host_ax_elem(int n, int op, const double *d, const double *ds,
		const double *g, const double *mass, double lambda,
		const double *u, double *au, double *t)
{
	const int n3 = n * n * n;
	const double *Dx = op == SSX_OP_HELMHOLTZ ? ds : d;
	double *t0 = t, *t1 = t + n3, *t2 = t + 2 * n3;
	for (int k = 0; k < n; ++k)
		for (int j = 0; j < n; ++j)
			for (int i = 0; i < n; ++i) {
				double gx = 0.0, gy = 0.0, gz = 0.0;
				for (int l = 0; l < n; ++l) {
					gx += Dx[i * n + l] * u[(k * n + j) * n + l];
					gy += d[j * n + l] * u[(k * n + l) * n + i];
					gz += d[k * n + l] * u[(l * n + j) * n + i];
				}
				const int q = (k * n + j) * n + i;
				const double gxx = g[q], gxy = g[n3 + q], gxz = g[2 * n3 + q];
				const double gyy = g[3 * n3 + q], gyz = g[4 * n3 + q];
				const double gzz = g[5 * n3 + q];
				t0[q] = gxx * gx + gxy * gy + gxz * gz;
				t1[q] = gxy * gx + gyy * gy + gyz * gz;
				t2[q] = gxz * gx + gyz * gy + gzz * gz;
			}
	const double *Dt = op == SSX_OP_HELMHOLTZ ? ds : d;
	for (int k = 0; k < n; ++k)
		for (int j = 0; j < n; ++j)
			for (int i = 0; i < n; ++i) {
				double v = 0.0;
				for (int l = 0; l < n; ++l)
					v += d[l * n + i] * t0[(k * n + j) * n + l]
						+ Dt[l * n + j] * t1[(k * n + l) * n + i];
				for (int l = 0; l < n; ++l)
					v += Dt[l * n + k] * t2[(l * n + j) * n + i];
				const int q = (k * n + j) * n + i;
				if (op == SSX_OP_HELMHOLTZ)
					v += lambda * (mass[q] * u[q]);
				au[q] = v;
			}
}

void ssx_coarse_free(ssx_coarse *cs)
{
	hipFree(cs->rowptr); hipFree(cs->col);
	hipFree(cs->valp); hipFree(cs->valh);
	hipFree(cs->rep); hipFree(cs->cmap);
	hipFree(cs->dinvp); hipFree(cs->dinvh); hipFree(cs->mask);
	hipFree(cs->ones);
	hipFree(cs->f); hipFree(cs->x); hipFree(cs->r); hipFree(cs->z);
	hipFree(cs->p0); hipFree(cs->p1); hipFree(cs->ap);
	memset(cs, 0, sizeof(*cs));
}

static int dup(void **d, const void *h, size_t bytes)
{
	if (hipMalloc(d, bytes ? bytes : 1) != hipSuccess) {
		*d = NULL;
		ssx_set_error("hipMalloc of %zu bytes (coarse level)", bytes);
		return SSX_ENOMEM;
	}
	if (h && bytes)
		HIPQ(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
	return SSX_OK;
}

/*
 * Build at set-up, while the level's groups, factors, mass, inverse
 * diagonals and mask are still on the host.
 */
int ssx_coarse_build(ssx_ctx *c, int l, const double *gf, const double *mass,
		const double *dp, const double *dh, const double *mask)
{
	ssx_level *L = &c->lev[l];
	ssx_coarse *cs = &L->cs;
	const int n = L->n, n3 = L->n3, nel = c->nel;
	const size_t nt = L->ntot;
	int *cmap = malloc(sizeof(int) * nt);
	long long *gid = malloc(sizeof(long long) * nt);
	if (!cmap || !gid) {
		free(cmap); free(gid);
		ssx_set_error("out of host memory (coarse level)");
		return SSX_ENOMEM;
	}
	/* copy -> group, then unique ids in order of the smallest copy */
	for (size_t q = 0; q < nt; ++q)
		gid[q] = -1;
	for (int mi = 0; mi < 3; ++mi) {
		const int M = 2 << mi;
		const long long G = L->ngrp[mi];
		for (long long g = 0; g < G; ++g)
			for (int t = 0; t < M; ++t)
				gid[L->grp[mi][t * G + g]] = ((long long) mi << 40) | g;
	}
	for (size_t q = 0; q < nt; ++q)
		cmap[q] = -1;
	int nu = 0;
	int *rep = malloc(sizeof(int) * nt);
	if (!rep) {
		free(cmap); free(gid);
		return SSX_ENOMEM;
	}
	for (size_t q = 0; q < nt; ++q) {
		if (cmap[q] >= 0)
			continue;
		rep[nu] = (int) q;
		if (gid[q] < 0) {
			cmap[q] = nu;
		} else {
			const int mi = (int) (gid[q] >> 40);
			const long long g = gid[q] & ((1LL << 40) - 1);
			const int M = 2 << mi;
			const long long G = L->ngrp[mi];
			for (int t = 0; t < M; ++t)
				cmap[L->grp[mi][t * G + g]] = nu;
		}
		++nu;
	}
	free(gid);
	/* element matrices, both operators, rows gathered per unique node */
	const int per = 8 * n3;                /* row entries: <= 8 elements */
	int *cnt = calloc(nu, sizeof(int));
	int *rc_col = malloc(sizeof(int) * (size_t) nu * per);
	double *rvp = malloc(sizeof(double) * (size_t) nu * per);
	double *rvh = malloc(sizeof(double) * (size_t) nu * per);
	double *ae = malloc(sizeof(double) * 2 * (size_t) n3 * n3);
	double *ub = malloc(sizeof(double) * n3 * 5);
	int rc = SSX_OK;
	if (!cnt || !rc_col || !rvp || !rvh || !ae || !ub) {
		ssx_set_error("out of host memory (coarse level)");
		rc = SSX_ENOMEM;
	}
	for (int e = 0; e < nel && rc == SSX_OK; ++e) {
		const double *g = gf + (size_t) e * 6 * n3;
		const double *ms = mass + (size_t) e * n3;
		for (int o = 0; o < 2; ++o)
			for (int b = 0; b < n3; ++b) {
				memset(ub, 0, sizeof(double) * n3);
				ub[b] = 1.0;
				host_ax_elem(n, o, L->b.d, L->b.dsvv, g, ms, c->lambda, ub,
						ub + n3, ub + 2 * n3);
				for (int a = 0; a < n3; ++a)
					ae[((size_t) o * n3 + a) * n3 + b] = ub[n3 + a];
			}
		for (int a = 0; a < n3; ++a) {
			const int row = cmap[(size_t) e * n3 + a];
			for (int b = 0; b < n3; ++b) {
				const int k = cnt[row]++;
				if (k >= per) {
					ssx_set_error("coarse level: node in more than 8 "
							"elements");
					rc = SSX_EMESH;
					break;
				}
				rc_col[(size_t) row * per + k] = cmap[(size_t) e * n3 + b];
				rvp[(size_t) row * per + k] = ae[(size_t) a * n3 + b];
				rvh[(size_t) row * per + k] = ae[((size_t) n3 + a) * n3 + b];
			}
			if (rc)
				break;
		}
	}
	/* sort each row by column (stable: contributions keep element order)
	 * and sum duplicates */
	int *rowptr = malloc(sizeof(int) * (nu + 1));
	long long nnz = 0;
	if (rc == SSX_OK && !rowptr)
		rc = SSX_ENOMEM;
	for (int u = 0; u < nu && rc == SSX_OK; ++u) {
		int *cc = rc_col + (size_t) u * per;
		double *vp = rvp + (size_t) u * per, *vh = rvh + (size_t) u * per;
		const int m = cnt[u];
		for (int a = 1; a < m; ++a) {         /* insertion sort */
			const int kc = cc[a];
			const double kp = vp[a], kh = vh[a];
			int b = a - 1;
			while (b >= 0 && cc[b] > kc) {
				cc[b + 1] = cc[b];
				vp[b + 1] = vp[b];
				vh[b + 1] = vh[b];
				--b;
			}
			cc[b + 1] = kc;
			vp[b + 1] = kp;
			vh[b + 1] = kh;
		}
		int w = 0;
		for (int a = 0; a < m; ++a) {
			if (w > 0 && cc[w - 1] == cc[a]) {
				vp[w - 1] += vp[a];
				vh[w - 1] += vh[a];
			} else {
				cc[w] = cc[a];
				vp[w] = vp[a];
				vh[w] = vh[a];
				++w;
			}
		}
		cnt[u] = w;
		rowptr[u] = (int) nnz;
		nnz += w;
	}
	int *col = NULL;
	double *valp = NULL, *valh = NULL;
	if (rc == SSX_OK) {
		rowptr[nu] = (int) nnz;
		col = malloc(sizeof(int) * nnz);
		valp = malloc(sizeof(double) * nnz);
		valh = malloc(sizeof(double) * nnz);
		if (!col || !valp || !valh)
			rc = SSX_ENOMEM;
	}
	for (int u = 0; u < nu && rc == SSX_OK; ++u) {
		memcpy(col + rowptr[u], rc_col + (size_t) u * per,
				sizeof(int) * cnt[u]);
		memcpy(valp + rowptr[u], rvp + (size_t) u * per,
				sizeof(double) * cnt[u]);
		memcpy(valh + rowptr[u], rvh + (size_t) u * per,
				sizeof(double) * cnt[u]);
	}
	free(cnt); free(rc_col); free(rvp); free(rvh); free(ae); free(ub);
	/* per-unique-node diagonals, mask, ones */
	double *tmp = rc == SSX_OK ? malloc(sizeof(double) * 4 * (size_t) nu)
		: NULL;
	if (rc == SSX_OK && !tmp)
		rc = SSX_ENOMEM;
	if (rc == SSX_OK) {
		for (int u = 0; u < nu; ++u) {
			tmp[u] = dp[rep[u]];
			tmp[nu + u] = dh[rep[u]];
			tmp[2 * nu + u] = mask[rep[u]];
			tmp[3 * nu + u] = 1.0;
		}
		cs->nu = nu;
		cs->nnz = nnz;
		const size_t vb = sizeof(double) * (size_t) nu;
		rc = dup((void **) &cs->rowptr, rowptr, sizeof(int) * (nu + 1));
		if (!rc) rc = dup((void **) &cs->col, col, sizeof(int) * nnz);
		if (!rc) rc = dup((void **) &cs->valp, valp, sizeof(double) * nnz);
		if (!rc) rc = dup((void **) &cs->valh, valh, sizeof(double) * nnz);
		if (!rc) rc = dup((void **) &cs->rep, rep, sizeof(int) * nu);
		if (!rc) rc = dup((void **) &cs->cmap, cmap, sizeof(int) * nt);
		if (!rc) rc = dup((void **) &cs->dinvp, tmp, vb);
		if (!rc) rc = dup((void **) &cs->dinvh, tmp + nu, vb);
		if (!rc) rc = dup((void **) &cs->mask, tmp + 2 * nu, vb);
		if (!rc) rc = dup((void **) &cs->ones, tmp + 3 * nu, vb);
		double **vec[7] = { &cs->f, &cs->x, &cs->r, &cs->z, &cs->p0, &cs->p1,
			&cs->ap };
		for (int k = 0; k < 7 && !rc; ++k)
			rc = dup((void **) vec[k], NULL, vb);
	}
	free(tmp); free(rowptr); free(col); free(valp); free(valh);
	free(cmap); free(rep);
	if (rc != SSX_OK)
		ssx_coarse_free(cs);
	return rc;
}

/* dot of unique-node vectors into scal[dst] (weight 1) */
static int udot(ssx_ctx *c, const double *x, const double *y, int dst)
{
	ssx_coarse *cs = &c->lev[c->nlev - 1].cs;
	TRYQ(ssxk_dot(c->stream, (size_t) cs->nu, x, cs->ones, y, c->part, 0));
	return ssxk_reduce(c->stream, c->part, 0, c->scal, dst);
}

/* u -= <u,1>/nu (subtract_nullspace, solver.c:200-213, unique nodes) */
static int unull(ssx_ctx *c, double *x)
{
	ssx_coarse *cs = &c->lev[c->nlev - 1].cs;
	TRYQ(ssxk_dot(c->stream, (size_t) cs->nu, x, cs->ones, NULL, c->part, 0));
	TRYQ(ssxk_reduce(c->stream, c->part, 0, c->scal, SC_NULLC));
	return ssxk_shift_s(c->stream, (size_t) cs->nu, c->scal, SC_NULLC,
			(double) cs->nu, x);
}

/* ---- direct coarsest solve (SSX_OPT_COARSE_DIRECT) ----
 * The assembled operator, masked rows and columns replaced by the identity
 * (and for a singular one, the pressure or an all-periodic Poisson level,
 * the rank-one term <1,.>/nu added, so that the inverse maps zero-mean
 * data to the zero-mean solution), is factored by
 * rocSOLVER's Cholesky and inverted once per (operator, lambda, null
 * space); a solve is then one dense product over the unique nodes. */
void ssx_coarse_direct_free(ssx_ctx *c)
{
	ssx_sd_free(c->sd);
	c->sd = NULL;
	hipFree(c->sd_gmap);
	hipFree(c->sd_smap);
	c->sd_gmap = c->sd_smap = NULL;
	c->sd_nu = 0;
	hipFree(c->cd_inv);
	c->cd_inv = NULL;
	c->cd_n = 0;
	if (c->cd_blas)
		rocblas_destroy_handle((rocblas_handle) c->cd_blas);
	c->cd_blas = NULL;
}

/* the context's rocBLAS handle on stream st, created at first use with
 * atomics disallowed: rocBLAS may otherwise sum partial results with
 * atomics in an order that depends on timing, and the factors must come
 * out the same bits every run */
static int coarse_blas(ssx_ctx *c, hipStream_t st)
{
	if (!c->cd_blas) {
		rocblas_handle h;
		if (rocblas_create_handle(&h) != rocblas_status_success) {
			ssx_set_error("rocblas_create_handle failed");
			return SSX_EHIP;
		}
		c->cd_blas = h;
		if (rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed)
				!= rocblas_status_success) {
			ssx_set_error("rocblas_set_atomics_mode failed");
			return SSX_EHIP;
		}
	}
	if (rocblas_set_stream((rocblas_handle) c->cd_blas, st)
			!= rocblas_status_success) {
		ssx_set_error("rocblas_set_stream failed");
		return SSX_EHIP;
	}
	return SSX_OK;
}

static int coarse_direct_build(ssx_ctx *c, int l, int op, int pres)
{
	ssx_coarse *cs = &c->lev[l].cs;
	if (c->cd_inv && c->cd_n == cs->nu && c->cd_op == op
			&& c->cd_pres == pres && c->cd_lambda == c->lambda)
		return c->cd_bad ? SSX_ENUMERIC : SSX_OK;
	const int nu = cs->nu;
	hipStream_t st = (hipStream_t) c->stream;
	if (c->cd_n != nu) {
		hipFree(c->cd_inv);
		c->cd_inv = NULL;
		c->cd_n = 0;
		if (hipMalloc((void **) &c->cd_inv, sizeof(double) * (size_t) nu
				* nu) != hipSuccess) {
			c->cd_inv = NULL;
			ssx_set_error("hipMalloc of the %d x %d coarse inverse", nu, nu);
			return SSX_ENOMEM;
		}
		c->cd_n = nu;
	}
	TRYQ(coarse_blas(c, st));
	rocblas_handle h = (rocblas_handle) c->cd_blas;
	c->cd_op = op;
	c->cd_pres = pres;
	c->cd_lambda = c->lambda;
	c->cd_bad = 1;                  /* until the factorisation succeeds */
	/* singular: the pressure, and the Poisson operator of a level with no
	 * Dirichlet node (all periodic), both with the constants as null
	 * space */
	int singular = pres;
	if (!pres && op == SSX_OP_POISSON) {
		TRYQ(udot(c, cs->mask, cs->ones, SC_TMP));
		double msum = 0.0;
		HIPQ(hipMemcpyAsync(&msum, c->scal + SC_TMP, sizeof(double),
				hipMemcpyDeviceToHost, st));
		HIPQ(hipStreamSynchronize(st));
		singular = msum >= (double) nu - 0.5;
	}
	c->cd_sing = singular;
	const double *val = op == SSX_OP_HELMHOLTZ ? cs->valh : cs->valp;
	TRYQ(ssxk_csr_dense(st, nu, cs->rowptr, cs->col, val,
			pres ? NULL : cs->mask, singular ? 1.0 / nu : 0.0, c->cd_inv));
	/* the level's CG flags (its CG is what this replaces) hold the infos */
	rocblas_int *info = (rocblas_int *) (c->cg_flags + 2 * l);
	rocblas_int hinfo[2] = { 0, 0 };
	if (rocsolver_dpotrf(h, rocblas_fill_lower, nu, c->cd_inv, nu, info)
			!= rocblas_status_success
			|| rocsolver_dpotri(h, rocblas_fill_lower, nu, c->cd_inv, nu,
					info + 1) != rocblas_status_success) {
		ssx_set_error("rocSOLVER Cholesky of the coarse operator failed");
		return SSX_EHIP;
	}
	HIPQ(hipMemcpyAsync(hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost,
			st));
	HIPQ(hipStreamSynchronize(st));
	if (hinfo[0] != 0 || hinfo[1] != 0)
		return SSX_ENUMERIC;        /* not definite: keep the CG */
	TRYQ(ssxk_dense_sym(st, nu, c->cd_inv));
	c->cd_bad = 0;
	return SSX_OK;
}

/* x = A^-1 f over the unique nodes (the right-hand side masked, and for
 * a singular operator made zero-mean first) */
static int coarse_direct_solve(ssx_ctx *c, ssx_level *L, int pres,
		const double *f, double *x)
{
	ssx_coarse *cs = &L->cs;
	void *st = c->stream;
	const int nu = cs->nu;
	TRYQ(ssxk_gather_u(st, nu, cs->rep, f, cs->f));
	if (!pres)
		TRYQ(ssxk_mul(st, (size_t) nu, cs->f, cs->mask, cs->f));
	if (c->cd_sing)
		TRYQ(unull(c, cs->f));
	TRYQ(ssxk_dense_mv(st, nu, c->cd_inv, cs->f, cs->x));
	if (c->cd_sing)
		TRYQ(unull(c, cs->x));
	return ssxk_scatter_u(st, L->ntot, cs->cmap, cs->x, x);
}

/* ---- sparse direct coarsest solve (SSX_OPT_COARSE_SPARSE) ----
 * The same system as the dense solve above, without its dense inverse:
 * the masked nodes are left out (their solution is zero), and for a
 * singular operator one node is pinned to zero instead of the rank-one
 * term (the zero-mean right-hand side makes the pinned system's solution a
 * solution of the singular one; the null space is then removed as in the
 * CG).  The factor (ssx_sdirect.c) is built on the host once per
 * (operator, lambda, null space) and solved on the device. */
static int coarse_sparse_build(ssx_ctx *c, int l, int op, int pres)
{
	ssx_level *L = &c->lev[l];
	ssx_coarse *cs = &L->cs;
	const int nu = cs->nu;
	if (c->sd_nu == nu && c->sd_op == op && c->sd_pres == pres
			&& c->sd_lambda == c->lambda && (c->sd || c->sd_bad))
		return c->sd_bad ? SSX_ENUMERIC : SSX_OK;
	ssx_sd_free(c->sd);
	c->sd = NULL;
	hipFree(c->sd_gmap);
	hipFree(c->sd_smap);
	c->sd_gmap = c->sd_smap = NULL;
	c->sd_nu = nu;
	c->sd_op = op;
	c->sd_pres = pres;
	c->sd_lambda = c->lambda;
	c->sd_bad = 1;                  /* until the factorisation succeeds */
	hipStream_t st = (hipStream_t) c->stream;
	HIPQ(hipStreamSynchronize(st));
	const long long nnz = cs->nnz;
	int *rp = malloc(sizeof(int) * ((size_t) nu + 1));
	int *ci = malloc(sizeof(int) * (size_t) nnz);
	double *av = malloc(sizeof(double) * (size_t) nnz);
	double *mk = malloc(sizeof(double) * (size_t) nu);
	int *rep = malloc(sizeof(int) * (size_t) nu);
	int *rid = malloc(sizeof(int) * (size_t) nu);
	int *map = malloc(sizeof(int) * (size_t) nu);
	int *rr = malloc(sizeof(int) * ((size_t) nu + 1));
	int *rc_ = malloc(sizeof(int) * (size_t) nnz);
	double *ra = malloc(sizeof(double) * (size_t) nnz);
	double *xyz = malloc(sizeof(double) * 3 * (size_t) nu);
	int rc = (rp && ci && av && mk && rep && rid && map && rr && rc_ && ra
			&& xyz) ? SSX_OK : SSX_ENOMEM;
	if (rc != SSX_OK)
		ssx_set_error("out of host memory (sparse coarse factor)");
	if (rc == SSX_OK && !L->h_x) {
		ssx_set_error("coarsest level without host coordinates");
		rc = SSX_EINVAL;
	}
	const double *val = op == SSX_OP_HELMHOLTZ ? cs->valh : cs->valp;
	if (rc == SSX_OK && (hipMemcpy(rp, cs->rowptr, sizeof(int) * (nu + 1),
				hipMemcpyDeviceToHost) != hipSuccess
			|| hipMemcpy(ci, cs->col, sizeof(int) * nnz,
				hipMemcpyDeviceToHost) != hipSuccess
			|| hipMemcpy(av, val, sizeof(double) * nnz,
				hipMemcpyDeviceToHost) != hipSuccess
			|| hipMemcpy(mk, cs->mask, sizeof(double) * nu,
				hipMemcpyDeviceToHost) != hipSuccess
			|| hipMemcpy(rep, cs->rep, sizeof(int) * nu,
				hipMemcpyDeviceToHost) != hipSuccess)) {
		ssx_set_error("hipMemcpy of the coarse operator failed");
		rc = SSX_EHIP;
	}
	int n = 0, singular = pres;
	if (rc == SSX_OK) {
		if (!pres && op == SSX_OP_POISSON) {
			double msum = 0.0;
			for (int u = 0; u < nu; ++u)
				msum += mk[u];
			singular = msum >= (double) nu - 0.5;
		}
		/* included nodes; a singular system pins its last node */
		for (int u = 0; u < nu; ++u)
			rid[u] = (pres || mk[u] != 0.0) ? 0 : -1;
		if (singular)
			for (int u = nu - 1; u >= 0; --u)
				if (rid[u] == 0) {
					rid[u] = -1;
					break;
				}
		for (int u = 0; u < nu; ++u)
			if (rid[u] == 0) {
				map[n] = u;
				rid[u] = n++;
			}
		long long m = 0;
		rr[0] = 0;
		for (int k = 0; k < n; ++k) {
			const int u = map[k];
			for (int q = rp[u]; q < rp[u + 1]; ++q)
				if (rid[ci[q]] >= 0) {
					rc_[m] = rid[ci[q]];
					ra[m] = av[q];
					++m;
				}
			rr[k + 1] = (int) m;
			const size_t e = (size_t) rep[u];
			xyz[3 * (size_t) k] = L->h_x[e];
			xyz[3 * (size_t) k + 1] = L->h_y[e];
			xyz[3 * (size_t) k + 2] = L->h_z[e];
		}
	}
	c->sd_sing = singular;
	ssx_sdfactor *f = NULL;
	/* the large fronts go to the device through rocSOLVER / rocBLAS */
	if (rc == SSX_OK)
		rc = coarse_blas(c, st);
	if (rc == SSX_OK && n > 0)
		rc = ssx_sd_factor(n, rr, rc_, ra, xyz, c->cd_blas, st, &f);
	else if (rc == SSX_OK)
		rc = SSX_ENUMERIC;          /* nothing to factor: keep the CG */
	if (rc == SSX_OK) {
		/* the factor's system index -> unique node: gather and scatter
		 * go straight between unique-node vectors and the factor */
		for (int i = 0; i < n; ++i)
			f->perm[i] = map[f->perm[i]];
		rc = ssx_sd_upload(f);
	}
	/* and straight between the level's E-vectors and the factor (a
	 * non-singular level: the masked nodes' solution is zero) */
	int *cmap = NULL, *gm = NULL, *sm = NULL;
	const size_t nt = L->ntot;
	if (rc == SSX_OK) {
		cmap = malloc(sizeof(int) * nt);
		gm = malloc(sizeof(int) * (size_t) n);
		sm = malloc(sizeof(int) * nt);
		if (!cmap || !gm || !sm) {
			ssx_set_error("out of host memory (sparse coarse maps)");
			rc = SSX_ENOMEM;
		} else if (hipMemcpy(cmap, cs->cmap, sizeof(int) * nt,
				hipMemcpyDeviceToHost) != hipSuccess) {
			ssx_set_error("hipMemcpy of the coarse copy map failed");
			rc = SSX_EHIP;
		}
	}
	if (rc == SSX_OK) {
		for (int u = 0; u < nu; ++u)
			rid[u] = -1;            /* reused: new index of node u */
		for (int i = 0; i < n; ++i) {
			gm[i] = rep[f->perm[i]];
			rid[f->perm[i]] = i;
		}
		for (size_t q = 0; q < nt; ++q)
			sm[q] = rid[cmap[q]];
		if (hipMalloc((void **) &c->sd_gmap, sizeof(int) * (size_t) n)
				!= hipSuccess
				|| hipMalloc((void **) &c->sd_smap, sizeof(int) * nt)
				!= hipSuccess
				|| hipMemcpy(c->sd_gmap, gm, sizeof(int) * (size_t) n,
					hipMemcpyHostToDevice) != hipSuccess
				|| hipMemcpy(c->sd_smap, sm, sizeof(int) * nt,
					hipMemcpyHostToDevice) != hipSuccess) {
			ssx_set_error("device maps of the sparse coarse solve");
			rc = SSX_ENOMEM;
		}
	}
	free(cmap); free(gm); free(sm);
	free(rp); free(ci); free(av); free(mk); free(rep); free(rid);
	free(map); free(rr); free(rc_); free(ra); free(xyz);
	if (rc != SSX_OK) {
		ssx_sd_free(f);
		return rc;
	}
	c->sd = f;
	c->sd_bad = 0;
	return SSX_OK;
}

static int coarse_sparse_solve(ssx_ctx *c, ssx_level *L, int pres,
		const double *f, double *x)
{
	ssx_coarse *cs = &L->cs;
	void *st = c->stream;
	const int nu = cs->nu;
	if (!c->sd_sing)                /* E-vector -> factor -> E-vector */
		return ssx_sd_solve_dev_map(c->sd, st, f, c->sd_gmap, x,
				c->sd_smap, (int) L->ntot);
	TRYQ(ssxk_gather_u(st, nu, cs->rep, f, cs->f));
	if (!pres)
		TRYQ(ssxk_mul(st, (size_t) nu, cs->f, cs->mask, cs->f));
	TRYQ(unull(c, cs->f));
	TRYQ(ssxk_fill(st, cs->x, 0.0, (size_t) nu));
	TRYQ(ssx_sd_solve_dev(c->sd, st, cs->f, cs->x));
	TRYQ(unull(c, cs->x));
	return ssxk_scatter_u(st, L->ntot, cs->cmap, cs->x, x);
}

int ssx_coarse_cg(ssx_ctx *c, int l, int op, int var, const double *f,
		double *x, double eps, double rtol, int max_iter, int x_zero,
		int *iters, double *err_out)
{
	ssx_level *L = &c->lev[l];
	ssx_coarse *cs = &L->cs;
	if (cs->nu <= 0) {
		ssx_set_error("level %d has no assembled operator", l);
		return SSX_EINVAL;
	}
	/* the direct solve replaces the CG of the reduced-precision cycle (a
	 * relative tolerance, a zero start) on small levels */
	if (c->coarse_direct > 0 && cs->nu <= c->coarse_direct && rtol > 0.0
			&& x_zero) {
		const int pres = var == SSX_PRESSURE;
		const int rc = coarse_direct_build(c, l, op, pres);
		if (rc == SSX_OK) {
			TRYQ(coarse_direct_solve(c, L, pres, f, x));
			if (iters)
				*iters = 1;
			*err_out = 0.0;
			return SSX_OK;
		}
		if (rc != SSX_ENUMERIC)
			return rc;
	}
	if (c->coarse_sparse > 0 && cs->nu <= c->coarse_sparse && rtol > 0.0
			&& x_zero) {
		const int pres = var == SSX_PRESSURE;
		const int rc = coarse_sparse_build(c, l, op, pres);
		if (rc == SSX_OK) {
			TRYQ(coarse_sparse_solve(c, &c->lev[l], pres, f, x));
			if (iters)
				*iters = 1;
			*err_out = 0.0;
			return SSX_OK;
		}
		if (rc != SSX_ENUMERIC)
			return rc;
	}
	void *st = c->stream;
	const int nu = cs->nu;
	const int pres = var == SSX_PRESSURE;
	const double *val = op == SSX_OP_HELMHOLTZ ? cs->valh : cs->valp;
	int ch = c->diag_choice;
	if (ch < 0)
		ch = pres ? 0 : 1;                /* precon.c:211-225 */
	const double *dinv = ch == 0 ? cs->dinvp : cs->dinvh;
	const double *mask = pres ? NULL : cs->mask;
	const int o = SC_CG(l);
	int *flags = c->cg_flags + 2 * l;
	TRYQ(ssxk_gather_u(st, nu, cs->rep, f, cs->f));
	if (mask)
		TRYQ(ssxk_mul(st, (size_t) nu, cs->f, mask, cs->f));
	if (x_zero) {
		TRYQ(ssxk_fill(st, cs->x, 0.0, (size_t) nu));
		TRYQ(ssxk_copy(st, cs->r, cs->f, (size_t) nu));
	} else {
		TRYQ(ssxk_gather_u(st, nu, cs->rep, x, cs->x));
		TRYQ(ssxk_csr_pupd(st, nu, cs->rowptr, cs->col, val, cs->x, cs->p1,
				NULL, mask, c->scal, 0, -1, 1, cs->ap, NULL, NULL));
		TRYQ(ssxk_copy(st, cs->r, cs->f, (size_t) nu));
		TRYQ(ssxk_axpby(st, (size_t) nu, -1.0, cs->ap, NULL, 1.0, cs->r));
	}
	if (pres)
		TRYQ(unull(c, cs->r));
	TRYQ(ssxk_mul(st, (size_t) nu, cs->r, dinv, cs->z));
	if (pres)
		TRYQ(unull(c, cs->z));
	TRYQ(ssxk_copy(st, cs->p0, cs->z, (size_t) nu));
	TRYQ(udot(c, cs->r, cs->z, o + CG_RDZ));
	TRYQ(ssxk_set_ints(st, flags, 0, 2));
	double eps_dev = eps;
	if (rtol > 0.0) {
		TRYQ(udot(c, cs->r, cs->r, o + CG_RR));
		TRYQ(ssxk_cg_tol(st, c->scal, o + CG_TOL, o + CG_RR, eps, rtol));
		eps_dev = -1.0;
	}
	const long long nbs = ssxk_csr_blocks(nu);
	/* one node per lane: no idle workgroups on the small level */
	const int nbu = (nu + 255) / 256 < SSX_NPART ? (nu + 255) / 256
		: SSX_NPART;
	double *pa = cs->p0, *pb = cs->p1;      /* p before / after update */
	int launched = 0;
	int chunk = c->cg_hint[l] > 0 ? c->cg_hint[l] + 1 : 8;
	/* two launches per iteration (each kernel replays the reduction it
	 * needs); SSX_OPT_COARSE_ASSEMBLED = 2 keeps the four-launch form */
	const int two = c->coarse_csr == 1
		&& 4 * (size_t) SSX_NPART + (size_t) nbs <= c->npart;
	double *ps = c->part + 4 * (size_t) SSX_NPART;   /* SpMV partials */
	double *pu = c->part;                            /* update partials */
	for (;;) {
		if (chunk > max_iter - launched)
			chunk = max_iter - launched;
		for (int k = 0; k < chunk && two; ++k) {
			const int j = launched + k;
			TRYQ(ssxk_csr_cg_spmv(st, nu, cs->rowptr, cs->col, val, pa, pb,
					cs->z, mask, c->scal, o, j, eps_dev, max_iter,
					(double) nu, pres, cs->ap, ps, pu, nbu, flags, 0));
			TRYQ(ssxk_csr_cg_update(st, nu, c->scal, o, j, pb, cs->ap,
					cs->x, cs->r, cs->z, dinv, cs->ones, pres, ps, (int) nbs,
					pu, nbu, flags));
			double *t = pa;
			pa = pb;
			pb = t;
		}
		if (two)      /* the stopping test of the last iteration */
			TRYQ(ssxk_csr_cg_spmv(st, nu, cs->rowptr, cs->col, val, pa, pb,
					cs->z, mask, c->scal, o, launched + chunk, eps_dev,
					max_iter, (double) nu, pres, cs->ap, ps, pu, nbu, flags,
					1));
		for (int k = 0; k < chunk && !two; ++k) {
			const int first = launched + k == 0;
			TRYQ(ssxk_csr_pupd(st, nu, cs->rowptr, cs->col, val, pa, pb,
					cs->z, mask, c->scal, o + CG_BETA,
					pres ? o + CG_NULLC : -1, first, cs->ap, c->part, flags));
			TRYQ(ssxk_cg_reduce(st, 0, c->part, nbs, c->scal, o, flags,
					eps_dev, max_iter, (double) nu, pres));
			TRYQ(ssxk_cg_update_nb(st, (size_t) nu, c->scal, o + CG_ALPHA,
					pb, cs->ap, cs->x, cs->r, cs->z, dinv, cs->ones, pres,
					c->part, flags, nbu));
			TRYQ(ssxk_cg_reduce(st, 1, c->part, nbu, c->scal, o, flags,
					eps_dev, max_iter, (double) nu, pres));
			double *t = pa;
			pa = pb;
			pb = t;
		}
		launched += chunk;
		hipStream_t hs = (hipStream_t) st;
		HIPQ(hipMemcpyAsync(c->h_flags, flags, 2 * sizeof(int),
				hipMemcpyDeviceToHost, hs));
		HIPQ(hipMemcpyAsync(c->h_scal + o + CG_RR, c->scal + o + CG_RR,
				sizeof(double), hipMemcpyDeviceToHost, hs));
		HIPQ(hipStreamSynchronize(hs));
		if (c->h_flags[0] || launched >= max_iter)
			break;
		chunk = 4;
	}
	TRYQ(ssxk_scatter_u(st, L->ntot, cs->cmap, cs->x, x));
	const int it = c->h_flags[1];
	const double err = sqrt(c->h_scal[o + CG_RR]);
	c->cg_hint[l] = it;
	if (iters)
		*iters = it;
	*err_out = err;
	if (isnan(err) || isinf(err)) {
		ssx_set_error("coarse CG level %d: residual is %g", l, err);
		return SSX_ENUMERIC;
	}
	return SSX_OK;
}
