/*
 * ssx_sdirect.c -- a sparse direct solver for the p-multigrid coarsest
 * level: nested-dissection ordering from node coordinates, multifrontal
 * Cholesky on the host (once per operator), and the two triangular solves
 * on the device, one launch per level of the separator tree.
 *
 * The coarse CG of the reduced-precision V-cycle (precon.c:242-244 with a
 * relative tolerance) is bound by launch latency: tens of dependent
 * iterations of a few microseconds of work each.  A factorisation turns the
 * solve into a fixed, short sequence of launches whose work is dense:
 *
 *   ordering   recursive bisection of the node set at the median of its
 *              longest coordinate extent; the separator is the set of nodes
 *              of one half adjacent (in the matrix graph) to the other, so
 *              periodic connections are separated too.  Each separator is
 *              one supernode whose children are the two halves' trees;
 *              subsets of at most SD_LEAF nodes are leaf supernodes.  The
 *              new order numbers each subtree contiguously, separator last.
 *   symbolic   U_s, the rows of supernode s below its own columns: the
 *              nodes after the subtree adjacent to the subtree (from its own
 *              columns and its children's U), ascending.
 *   numeric    the front of s (|C_s| + |U_s| square, lower half) gets the
 *              matrix entries of its columns and the children's update
 *              matrices (extend-add), and is factored partially: L11, L21,
 *              the update matrix F22 - L21 L21^T for the parent.  Stored:
 *              M = L11^-1 (so a solve step is a product, not a substitution
 *              chain) and K = L21 M, row- and column-major.
 *   solve      forward (leaves to root): b~ = b_C - (children's updates on
 *              C), y_C = M b~, t_s = K b~ + (children's updates on U_s);
 *              backward (root to leaves): x_C = M^T y_C - K^T x_U (which is
 *              M^T (y_C - L21^T x_U)).  Both products of a step read the
 *              same vector, so each level is one launch in each direction.
 *              Every update reaches its destination through one parent
 *              gather with the children in a fixed order: no atomics, the
 *              result is the same on every run.
 */
#pragma GCC optimize("O3")
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "ssx_internal.h"

#define SD_LEAF 64                   /* largest leaf supernode            */

typedef struct {
	int n;
	const int *rp, *ci;
	const double *xyz;
	int *perm;                       /* new index -> node                 */
	int next;                        /* next new index                    */
	int *mark;                       /* scratch per node                  */
	ssx_sdnode *node;
	int nsn, cap;
	int *first;                      /* per supernode: first new index of
	                                    its subtree                       */
	int err;
} nd_t;

static int cmp_key(const void *a, const void *b)
{
	const double *x = a, *y = b;
	return x[0] < y[0] ? -1 : x[0] > y[0] ? 1 : (x[1] < y[1] ? -1 : x[1] > y[1]);
}

static int new_node(nd_t *d)
{
	if (d->nsn == d->cap) {
		const int cap = d->cap ? 2 * d->cap : 256;
		ssx_sdnode *nn = realloc(d->node, sizeof(ssx_sdnode) * cap);
		int *nf = realloc(d->first, sizeof(int) * cap);
		if (!nn || !nf) {
			if (nn) d->node = nn;
			if (nf) d->first = nf;
			d->err = 1;
			return -1;
		}
		d->node = nn;
		d->first = nf;
		d->cap = cap;
	}
	const int s = d->nsn++;
	memset(&d->node[s], 0, sizeof(ssx_sdnode));
	d->node[s].ch0 = d->node[s].ch1 = -1;
	d->node[s].g0 = d->node[s].g1 = -1;
	return s;
}

/* order the nodes set[0..m) (m >= 1); returns the root supernode of their
 * tree */
static int nd_rec(nd_t *d, int *set, int m)
{
	if (d->err)
		return -1;
	const int start = d->next;
	if (m <= SD_LEAF) {
		const int s = new_node(d);
		if (s < 0)
			return -1;
		for (int k = 0; k < m; ++k)
			d->perm[d->next++] = set[k];
		d->node[s].c0 = start;
		d->node[s].nc = m;
		d->first[s] = start;
		return s;
	}
	/* longest extent */
	double lo[3], hi[3];
	for (int a = 0; a < 3; ++a) {
		lo[a] = INFINITY;
		hi[a] = -INFINITY;
	}
	for (int k = 0; k < m; ++k)
		for (int a = 0; a < 3; ++a) {
			const double v = d->xyz[3 * (size_t) set[k] + a];
			lo[a] = v < lo[a] ? v : lo[a];
			hi[a] = v > hi[a] ? v : hi[a];
		}
	int ax = 0;
	for (int a = 1; a < 3; ++a)
		if (hi[a] - lo[a] > hi[ax] - lo[ax])
			ax = a;
	/* sort by (coordinate, node id): the split is the median rank */
	double *key = malloc(sizeof(double) * 2 * (size_t) m);
	if (!key) {
		d->err = 1;
		return -1;
	}
	for (int k = 0; k < m; ++k) {
		key[2 * k] = d->xyz[3 * (size_t) set[k] + ax];
		key[2 * k + 1] = (double) set[k];
	}
	qsort(key, m, 2 * sizeof(double), cmp_key);
	for (int k = 0; k < m; ++k)
		set[k] = (int) key[2 * k + 1];
	free(key);
	const int h = m / 2;
	/* mark: 1 first half (A), 2 second half (B) */
	for (int k = 0; k < m; ++k)
		d->mark[set[k]] = k < h ? 1 : 2;
	/* separator: nodes of A with a neighbour in B */
	int *a2 = malloc(sizeof(int) * (size_t) m);
	if (!a2) {
		d->err = 1;
		return -1;
	}
	int na = 0, ns = 0;
	int *sv = malloc(sizeof(int) * (size_t) (h > 0 ? h : 1));
	if (!sv) {
		free(a2);
		d->err = 1;
		return -1;
	}
	for (int k = 0; k < h; ++k) {
		const int v = set[k];
		int cut = 0;
		for (int q = d->rp[v]; q < d->rp[v + 1] && !cut; ++q)
			cut = d->mark[d->ci[q]] == 2;
		if (cut)
			sv[ns++] = v;
		else
			a2[na++] = v;
	}
	for (int k = 0; k < m; ++k)
		d->mark[set[k]] = 0;
	/* B is set[h..m) */
	int *bset = set + h;
	const int nb = m - h;
	int ca = -1, cb = -1;
	if (na > 0)
		ca = nd_rec(d, a2, na);
	free(a2);
	if (nb > 0 && !d->err)
		cb = nd_rec(d, bset, nb);
	if (d->err) {
		free(sv);
		return -1;
	}
	const int s = new_node(d);
	if (s < 0) {
		free(sv);
		return -1;
	}
	d->node[s].c0 = d->next;
	d->node[s].nc = ns;
	for (int k = 0; k < ns; ++k)
		d->perm[d->next++] = sv[k];
	free(sv);
	d->node[s].ch0 = ca >= 0 ? ca : cb;
	d->node[s].ch1 = ca >= 0 ? cb : -1;
	d->first[s] = start;
	return s;
}

/* ---- dense kernels of the numeric factorisation (column-major) ---- */

/* F[i][j] -= sum_{k in [kb, ke)} F[i][k] F[j][k] for the lower part of
 * columns j >= ke: 8 x 4 tiles with the sums in registers (each entry of
 * the block's columns is loaded once per tile, not once per product);
 * built for AVX2/FMA where the host has them */
__attribute__((target_clones("arch=haswell", "default")))
void ssx_sd_trailing_update(double *F, int nr, int kb, int ke)
{
	for (int j0 = ke; j0 < nr; j0 += 4) {
		const int nj = nr - j0 < 4 ? nr - j0 : 4;
		for (int i0 = j0; i0 < nr; i0 += 8) {
			const int ni = nr - i0 < 8 ? nr - i0 : 8;
			double acc[4][8];
			for (int jj = 0; jj < 4; ++jj)
				for (int ii = 0; ii < 8; ++ii)
					acc[jj][ii] = 0.0;
			if (ni == 8 && nj == 4) {
				for (int k = kb; k < ke; ++k) {
					const double *fk = F + (size_t) k * nr;
					const double *a = fk + i0;
					const double b0 = fk[j0], b1 = fk[j0 + 1];
					const double b2 = fk[j0 + 2], b3 = fk[j0 + 3];
					for (int ii = 0; ii < 8; ++ii) {
						acc[0][ii] += a[ii] * b0;
						acc[1][ii] += a[ii] * b1;
						acc[2][ii] += a[ii] * b2;
						acc[3][ii] += a[ii] * b3;
					}
				}
			} else {
				for (int k = kb; k < ke; ++k) {
					const double *fk = F + (size_t) k * nr;
					for (int jj = 0; jj < nj; ++jj)
						for (int ii = 0; ii < ni; ++ii)
							acc[jj][ii] += fk[i0 + ii] * fk[j0 + jj];
				}
			}
			for (int jj = 0; jj < nj; ++jj) {
				double *fj = F + (size_t) (j0 + jj) * nr;
				for (int ii = 0; ii < ni; ++ii)
					if (i0 + ii >= j0 + jj)
						fj[i0 + ii] -= acc[jj][ii];
			}
		}
	}
}

/* partial Cholesky of the first nc columns of the nr x nr lower front F;
 * F22 becomes the update matrix.  Returns 0, or 1 on a non-positive pivot */
static int front_factor(double *F, int nr, int nc, double *flops)
{
	const int B = 32;
	for (int kb = 0; kb < nc; kb += B) {
		const int ke = kb + B < nc ? kb + B : nc;
		/* the block's columns, left-looking inside the block */
		for (int j = kb; j < ke; ++j) {
			double *fj = F + (size_t) j * nr;
			for (int k = kb; k < j; ++k) {
				const double *fk = F + (size_t) k * nr;
				const double l = fk[j];
				for (int i = j; i < nr; ++i)
					fj[i] -= fk[i] * l;
			}
			const double dgn = fj[j];
			if (!(dgn > 0.0))
				return 1;
			const double r = sqrt(dgn);
			const double ir = 1.0 / r;
			fj[j] = r;
			for (int i = j + 1; i < nr; ++i)
				fj[i] *= ir;
		}
		ssx_sd_trailing_update(F, nr, kb, ke);
	}
	double fl = 0.0;
	for (int j = 0; j < nc; ++j)
		fl += (double) (nr - j) * (nr - j);
	*flops += fl;
	return 0;
}

/* M = L^-1 for the n x n lower triangular L (leading dimension ld); M
 * column-major n x n, upper part zero */
static void tri_inverse(const double *L, int ld, int n, double *M)
{
	memset(M, 0, sizeof(double) * (size_t) n * n);
	for (int j = 0; j < n; ++j) {
		double *mj = M + (size_t) j * n;
		mj[j] = 1.0 / L[j + (size_t) j * ld];
		/* column j of the inverse: forward substitution of e_j, by
		 * columns of L (saxpy form) */
		for (int k = j; k < n; ++k) {
			const double v = k == j ? mj[j] : mj[k] / L[k + (size_t) k * ld];
			mj[k] = v;
			const double *lk = L + (size_t) k * ld;
			for (int i = k + 1; i < n; ++i)
				mj[i] -= lk[i] * v;
		}
	}
}

/* ---- large fronts on the device (rocSOLVER / rocBLAS) ----
 * The dense work of a front grows with the cube of its size and the top
 * separators' fronts hold most of it: those of at least SD_GPU_FRONT rows
 * are factored on the device: potrf of F11, trsm for L21, syrk for the
 * update matrix, trtri for M and trmm for K = L21 M; M, K and the update
 * matrix come back to the host. */
#define SD_GPU_FRONT 1024

typedef struct {
	rocblas_handle h;
	hipStream_t st;
	double *dF, *dM, *dK;
	size_t capF, capM, capK;
	rocblas_int *dinfo;
} sd_gpu;

static int sd_grow(double **p, size_t *cap, size_t n)
{
	if (n <= *cap)
		return SSX_OK;
	hipFree(*p);
	*p = NULL;
	*cap = 0;
	if (hipMalloc((void **) p, sizeof(double) * n) != hipSuccess) {
		*p = NULL;
		ssx_set_error("hipMalloc of a %zu-entry front", n);
		return SSX_ENOMEM;
	}
	*cap = n;
	return SSX_OK;
}

static void sd_gpu_free(sd_gpu *g)
{
	hipFree(g->dF); hipFree(g->dM); hipFree(g->dK); hipFree(g->dinfo);
}

/* F (host, nr x nr column-major, lower part): M = L11^-1 (nc x nc), K
 * (nu x nc) and the update matrix Up (nu x nu) on return; SSX_ENUMERIC if
 * F11 is not positive definite */
static int front_gpu(sd_gpu *g, const double *F, int nr, int nc, double *M,
		double *K, double *Up)
{
	const int nu = nr - nc;
	const double one = 1.0, mone = -1.0;
	int rc = sd_grow(&g->dF, &g->capF, (size_t) nr * nr);
	if (!rc) rc = sd_grow(&g->dM, &g->capM, (size_t) nc * nc);
	if (!rc && nu > 0) rc = sd_grow(&g->dK, &g->capK, (size_t) nu * nc);
	if (!rc && !g->dinfo && hipMalloc((void **) &g->dinfo,
			2 * sizeof(rocblas_int)) != hipSuccess) {
		g->dinfo = NULL;
		rc = SSX_ENOMEM;
	}
	if (rc)
		return rc;
	rocblas_int info[2] = { 0, 0 };
	/* blocking copies and a device-wide wait on either side of the
	 * library calls: the front and its results live in pageable host
	 * memory, and nothing here may depend on how the runtime orders an
	 * asynchronous pageable copy against library kernels */
	int bad = hipStreamSynchronize(g->st) != hipSuccess;
	bad |= hipMemcpy(g->dF, F, sizeof(double) * (size_t) nr * nr,
			hipMemcpyHostToDevice) != hipSuccess;
	bad |= rocsolver_dpotrf(g->h, rocblas_fill_lower, nc, g->dF, nr,
			g->dinfo) != rocblas_status_success;
	if (nu > 0) {
		bad |= rocblas_dtrsm(g->h, rocblas_side_right, rocblas_fill_lower,
				rocblas_operation_transpose, rocblas_diagonal_non_unit, nu,
				nc, &one, g->dF, nr, g->dF + nc, nr)
				!= rocblas_status_success;
		bad |= rocblas_dsyrk(g->h, rocblas_fill_lower,
				rocblas_operation_none, nu, nc, &mone, g->dF + nc, nr, &one,
				g->dF + nc + (size_t) nc * nr, nr) != rocblas_status_success;
	}
	bad |= hipDeviceSynchronize() != hipSuccess;
	bad |= hipMemcpy2D(g->dM, sizeof(double) * nc, g->dF,
			sizeof(double) * nr, sizeof(double) * nc, nc,
			hipMemcpyDeviceToDevice) != hipSuccess;
	bad |= rocsolver_dtrtri(g->h, rocblas_fill_lower,
			rocblas_diagonal_non_unit, nc, g->dM, nc, g->dinfo + 1)
			!= rocblas_status_success;
	if (nu > 0)
		bad |= rocblas_dtrmm(g->h, rocblas_side_right, rocblas_fill_lower,
				rocblas_operation_none, rocblas_diagonal_non_unit, nu, nc,
				&one, g->dM, nc, g->dF + nc, nr, g->dK, nu)
				!= rocblas_status_success;
	bad |= hipStreamSynchronize(g->st) != hipSuccess;
	bad |= hipDeviceSynchronize() != hipSuccess;
	bad |= hipMemcpy(M, g->dM, sizeof(double) * (size_t) nc * nc,
			hipMemcpyDeviceToHost) != hipSuccess;
	if (nu > 0) {
		bad |= hipMemcpy(K, g->dK, sizeof(double) * (size_t) nu * nc,
				hipMemcpyDeviceToHost) != hipSuccess;
		bad |= hipMemcpy2D(Up, sizeof(double) * nu,
				g->dF + nc + (size_t) nc * nr, sizeof(double) * nr,
				sizeof(double) * nu, nu, hipMemcpyDeviceToHost)
				!= hipSuccess;
	}
	bad |= hipMemcpy(info, g->dinfo, sizeof(info), hipMemcpyDeviceToHost)
			!= hipSuccess;
	if (bad) {
		ssx_set_error("sparse direct: device factorisation of a front of "
				"%d rows failed", nr);
		return SSX_EHIP;
	}
	if (info[0] != 0 || info[1] != 0)
		return SSX_ENUMERIC;
	/* the strict upper part of M: zero, as the host path leaves it */
	for (int j = 1; j < nc; ++j)
		memset(M + (size_t) j * nc, 0, sizeof(double) * (size_t) j);
	return SSX_OK;
}

void ssx_sd_free(ssx_sdfactor *f)
{
	if (!f)
		return;
	free(f->node); free(f->perm); free(f->lev_off); free(f->lev_sn);
	free(f->val); free(f->idx); free(f->wk_off); free(f->wk_vec);
	free(f->wk);
	hipFree(f->d_node); hipFree(f->d_val);
	hipFree(f->d_idx); hipFree(f->d_perm); hipFree(f->d_wk);
	hipFree(f->d_b); hipFree(f->d_y); hipFree(f->d_x); hipFree(f->d_t);
	free(f);
}

static int sd_factor_run(int n, const int *rp, const int *ci,
		const double *a, const double *xyz, void *blas, void *stream,
		ssx_sdfactor **out)
{
	sd_gpu gpu = { (rocblas_handle) blas, (hipStream_t) stream, NULL, NULL,
		NULL, 0, 0, 0, NULL };
	*out = NULL;
	if (n <= 0) {
		ssx_set_error("sparse direct: empty system");
		return SSX_EINVAL;
	}
	ssx_sdfactor *f = calloc(1, sizeof(*f));
	nd_t d = { 0 };
	int *set = malloc(sizeof(int) * (size_t) n);
	int *inv = NULL;
	d.n = n;
	d.rp = rp;
	d.ci = ci;
	d.xyz = xyz;
	d.perm = malloc(sizeof(int) * (size_t) n);
	d.mark = calloc((size_t) n, sizeof(int));
	if (!f || !set || !d.perm || !d.mark) {
		free(f); free(set); free(d.perm); free(d.mark);
		ssx_set_error("out of host memory (sparse direct ordering)");
		return SSX_ENOMEM;
	}
	for (int i = 0; i < n; ++i)
		set[i] = i;
	nd_rec(&d, set, n);
	free(set);
	int rc = d.err || d.next != n ? SSX_ENOMEM : SSX_OK;
	if (rc != SSX_OK)
		ssx_set_error("sparse direct: nested dissection failed");
	f->n = n;
	f->perm = d.perm;
	f->node = d.node;
	f->nsn = d.nsn;
	/* inverse permutation */
	inv = rc == SSX_OK ? malloc(sizeof(int) * (size_t) n) : NULL;
	if (rc == SSX_OK && !inv)
		rc = SSX_ENOMEM;
	for (int i = 0; rc == SSX_OK && i < n; ++i)
		inv[d.perm[i]] = i;
	/* ---- symbolic: U of every supernode (they are in post-order) ---- */
	int **U = rc == SSX_OK ? calloc((size_t) f->nsn, sizeof(int *)) : NULL;
	if (rc == SSX_OK && !U)
		rc = SSX_ENOMEM;
	int *mk = d.mark;                /* reused: stamp per new index */
	for (int i = 0; i < n; ++i)
		mk[i] = -1;
	int *lst = rc == SSX_OK ? malloc(sizeof(int) * (size_t) n) : NULL;
	if (rc == SSX_OK && !lst)
		rc = SSX_ENOMEM;
	for (int s = 0; rc == SSX_OK && s < f->nsn; ++s) {
		ssx_sdnode *p = &f->node[s];
		const int end = p->c0 + p->nc;
		int m = 0;
		for (int j = 0; j < p->nc; ++j) {
			const int v = d.perm[p->c0 + j];
			for (int q = rp[v]; q < rp[v + 1]; ++q) {
				const int w = inv[ci[q]];
				if (w >= end && mk[w] != s) {
					mk[w] = s;
					lst[m++] = w;
				}
			}
		}
		const int ch[2] = { p->ch0, p->ch1 };
		for (int c = 0; c < 2; ++c) {
			if (ch[c] < 0)
				continue;
			const ssx_sdnode *q = &f->node[ch[c]];
			for (int k = 0; k < q->nu; ++k) {
				const int w = U[ch[c]][k];
				if (w >= end && mk[w] != s) {
					mk[w] = s;
					lst[m++] = w;
				}
			}
		}
		/* ascending (counting over the stamp is not ordered) */
		int *u = malloc(sizeof(int) * (size_t) (m > 0 ? m : 1));
		if (!u) {
			rc = SSX_ENOMEM;
			break;
		}
		memcpy(u, lst, sizeof(int) * (size_t) m);
		/* insertion into order by a simple sort */
		for (int k = 1; k < m; ++k) {
			const int v = u[k];
			int q = k - 1;
			while (q >= 0 && u[q] > v) {
				u[q + 1] = u[q];
				--q;
			}
			u[q + 1] = v;
		}
		U[s] = u;
		p->nu = m;
	}
	free(lst);
	/* ---- sizes and offsets ---- */
	long long nval = 0, nidx = 0, nt = 0;
	for (int s = 0; rc == SSX_OK && s < f->nsn; ++s) {
		ssx_sdnode *p = &f->node[s];
		p->moff = nval;
		nval += (long long) p->nc * p->nc;
		p->koff = nval;
		nval += (long long) p->nu * p->nc;
		p->mroff = nval;
		nval += (long long) p->nc * p->nc;
		p->kroff = nval;
		nval += (long long) p->nu * p->nc;
		p->toff = nt;
		nt += p->nu;
		p->uoff = (int) nidx;
		nidx += p->nu;
		const int nr = p->nc + p->nu;
		if (p->ch0 >= 0) {
			p->g0 = (int) nidx;
			nidx += nr;
		}
		if (p->ch1 >= 0) {
			p->g1 = (int) nidx;
			nidx += nr;
		}
	}
	if (rc == SSX_OK && (nidx > 0x7fffffffLL)) {
		ssx_set_error("sparse direct: index table too large");
		rc = SSX_EINVAL;
	}
	f->nval = nval;
	f->nidx = nidx;
	f->nt = nt;
	f->val = rc == SSX_OK ? malloc(sizeof(double) * (size_t) (nval > 0 ? nval : 1))
		: NULL;
	f->idx = rc == SSX_OK ? malloc(sizeof(int) * (size_t) (nidx > 0 ? nidx : 1))
		: NULL;
	if (rc == SSX_OK && (!f->val || !f->idx))
		rc = SSX_ENOMEM;
	/* U lists and the children's gather maps: position in R_s of each
	 * entry of the child's U (-1 where none maps) -- stored per
	 * destination: g[r] = index k in the child's U landing at r */
	int *pos = d.mark;               /* position in R_s of a new index */
	for (int i = 0; i < n; ++i)
		pos[i] = -1;
	for (int s = 0; rc == SSX_OK && s < f->nsn; ++s) {
		ssx_sdnode *p = &f->node[s];
		const int nr = p->nc + p->nu;
		memcpy(f->idx + p->uoff, U[s], sizeof(int) * (size_t) p->nu);
		for (int j = 0; j < p->nc; ++j)
			pos[p->c0 + j] = j;
		for (int k = 0; k < p->nu; ++k)
			pos[U[s][k]] = p->nc + k;
		const int ch[2] = { p->ch0, p->ch1 };
		const int go[2] = { p->g0, p->g1 };
		for (int c = 0; c < 2; ++c) {
			if (ch[c] < 0)
				continue;
			int *g = f->idx + go[c];
			for (int r = 0; r < nr; ++r)
				g[r] = -1;
			const ssx_sdnode *q = &f->node[ch[c]];
			for (int k = 0; k < q->nu; ++k) {
				const int r = pos[U[ch[c]][k]];
				if (r < 0) {
					ssx_set_error("sparse direct: update outside the "
							"parent front");
					rc = SSX_EINVAL;
					break;
				}
				g[r] = k;
			}
		}
		for (int j = 0; j < p->nc; ++j)
			pos[p->c0 + j] = -1;
		for (int k = 0; k < p->nu; ++k)
			pos[U[s][k]] = -1;
	}
	/* ---- numeric: fronts in post-order, update matrices on a stack ---- */
	double **upd = rc == SSX_OK ? calloc((size_t) f->nsn, sizeof(double *))
		: NULL;
	if (rc == SSX_OK && !upd)
		rc = SSX_ENOMEM;
	f->flops = 0.0;
	for (int s = 0; rc == SSX_OK && s < f->nsn; ++s) {
		ssx_sdnode *p = &f->node[s];
		const int nc = p->nc, nu = p->nu, nr = nc + nu;
		double *F = calloc((size_t) nr * nr > 0 ? (size_t) nr * nr : 1,
				sizeof(double));
		if (!F) {
			rc = SSX_ENOMEM;
			break;
		}
		for (int j = 0; j < nc; ++j)
			pos[p->c0 + j] = j;
		for (int k = 0; k < nu; ++k)
			pos[U[s][k]] = nc + k;
		/* the matrix entries of the columns (lower part) */
		for (int j = 0; j < nc; ++j) {
			const int v = d.perm[p->c0 + j];
			for (int q = rp[v]; q < rp[v + 1]; ++q) {
				const int w = inv[ci[q]];
				if (w < p->c0 + j)
					continue;
				const int r = pos[w];
				if (r < 0) {
					ssx_set_error("sparse direct: entry outside the front");
					rc = SSX_EINVAL;
					break;
				}
				F[r + (size_t) j * nr] += a[q];
			}
		}
		/* extend-add of the children's update matrices */
		const int ch[2] = { p->ch0, p->ch1 };
		for (int c = 0; c < 2 && rc == SSX_OK; ++c) {
			if (ch[c] < 0)
				continue;
			const ssx_sdnode *q = &f->node[ch[c]];
			const int m = q->nu;
			const double *Uc = upd[ch[c]];
			const int *uc = U[ch[c]];
			for (int jb = 0; jb < m; ++jb) {
				const int cj = pos[uc[jb]];
				for (int ib = jb; ib < m; ++ib)
					F[pos[uc[ib]] + (size_t) cj * nr] += Uc[ib + (size_t) jb * m];
			}
			free(upd[ch[c]]);
			upd[ch[c]] = NULL;
		}
		for (int j = 0; j < nc; ++j)
			pos[p->c0 + j] = -1;
		for (int k = 0; k < nu; ++k)
			pos[U[s][k]] = -1;
		const int dev = blas && nr >= SD_GPU_FRONT && nc > 0;
		if (rc == SSX_OK && dev) {
			double *Up = nu > 0 ? malloc(sizeof(double) * (size_t) nu * nu)
				: NULL;
			if (nu > 0 && !Up)
				rc = SSX_ENOMEM;
			if (rc == SSX_OK)
				rc = front_gpu(&gpu, F, nr, nc, f->val + p->moff,
						f->val + p->koff, Up);
			if (rc == SSX_ENUMERIC)
				ssx_set_error("sparse direct: the coarse operator is not "
						"positive definite (supernode %d)", s);
			if (rc == SSX_OK) {
				upd[s] = Up;
				f->ndev++;
				double fl = 0.0;
				for (int j = 0; j < nc; ++j)
					fl += (double) (nr - j) * (nr - j);
				f->flops += fl + (double) nu * nc * nc;
			} else {
				free(Up);
			}
		}
		if (rc == SSX_OK && !dev && front_factor(F, nr, nc, &f->flops)) {
			ssx_set_error("sparse direct: the coarse operator is not "
					"positive definite (supernode %d)", s);
			rc = SSX_ENUMERIC;
		}
		if (rc == SSX_OK && !dev) {
			double *M = f->val + p->moff;
			tri_inverse(F, nr, nc, M);
			/* K = L21 M (nu x nc): the rows below apply to b~ directly,
			 * so y = M b~ and t = K b~ are independent products */
			double *K = f->val + p->koff;
			memset(K, 0, sizeof(double) * (size_t) nu * nc);
			for (int j = 0; j < nc; ++j) {
				double *kj = K + (size_t) j * nu;
				for (int i = j; i < nc; ++i) {
					const double m = M[i + (size_t) j * nc];
					const double *li = F + nc + (size_t) i * nr;
					for (int r = 0; r < nu; ++r)
						kj[r] += li[r] * m;
				}
			}
			f->flops += (double) nu * nc * nc;
			if (nu > 0) {
				double *Up = malloc(sizeof(double) * (size_t) nu * nu);
				if (!Up) {
					rc = SSX_ENOMEM;
				} else {
					for (int j = 0; j < nu; ++j)
						memcpy(Up + (size_t) j * nu,
								F + nc + (size_t) (nc + j) * nr,
								sizeof(double) * (size_t) nu);
					upd[s] = Up;
				}
			}
		}
		if (rc == SSX_OK) {
			const double *M = f->val + p->moff, *K = f->val + p->koff;
			/* row-major copies for the forward products */
			double *Mr = f->val + p->mroff, *Kr = f->val + p->kroff;
			for (int j = 0; j < nc; ++j)
				for (int i = 0; i < nc; ++i)
					Mr[(size_t) i * nc + j] = M[i + (size_t) j * nc];
			for (int j = 0; j < nc; ++j)
				for (int r = 0; r < nu; ++r)
					Kr[(size_t) r * nc + j] = K[r + (size_t) j * nu];
		}
		free(F);
	}
	if (upd) {
		for (int s = 0; s < f->nsn; ++s)
			free(upd[s]);
		free(upd);
	}
	if (U) {
		for (int s = 0; s < f->nsn; ++s)
			free(U[s]);
		free(U);
	}
	sd_gpu_free(&gpu);
	free(inv);
	free(d.mark);
	free(d.first);
	/* ---- levels: leaves 0, a parent one above its highest child ---- */
	int *lev = rc == SSX_OK ? malloc(sizeof(int) * (size_t) f->nsn) : NULL;
	if (rc == SSX_OK && !lev)
		rc = SSX_ENOMEM;
	int nlev = 0;
	for (int s = 0; rc == SSX_OK && s < f->nsn; ++s) {
		const ssx_sdnode *p = &f->node[s];
		int l = 0;
		if (p->ch0 >= 0 && lev[p->ch0] + 1 > l)
			l = lev[p->ch0] + 1;
		if (p->ch1 >= 0 && lev[p->ch1] + 1 > l)
			l = lev[p->ch1] + 1;
		lev[s] = l;
		if (l + 1 > nlev)
			nlev = l + 1;
	}
	if (rc == SSX_OK) {
		f->nlev = nlev;
		f->lev_off = calloc((size_t) nlev + 1, sizeof(int));
		f->lev_sn = malloc(sizeof(int) * (size_t) f->nsn);
		if (!f->lev_off || !f->lev_sn)
			rc = SSX_ENOMEM;
	}
	if (rc == SSX_OK) {
		for (int s = 0; s < f->nsn; ++s)
			f->lev_off[lev[s] + 1]++;
		for (int l = 0; l < nlev; ++l)
			f->lev_off[l + 1] += f->lev_off[l];
		int *fill = calloc((size_t) nlev, sizeof(int));
		if (!fill) {
			rc = SSX_ENOMEM;
		} else {
			for (int s = 0; s < f->nsn; ++s)
				f->lev_sn[f->lev_off[lev[s]] + fill[lev[s]]++] = s;
			free(fill);
		}
		/* the device solve stages a supernode's vectors in LDS */
		for (int s = 0; s < f->nsn && rc == SSX_OK; ++s) {
			const ssx_sdnode *p = &f->node[s];
			if (p->nc + p->nu > SD_MAX_VEC) {
				ssx_set_error("sparse direct: supernode of %d + %d rows "
						"beyond the solve's limit %d", p->nc, p->nu,
						SD_MAX_VEC);
				rc = SSX_ENUMERIC;
			}
		}
	}
	free(lev);
	if (rc != SSX_OK) {
		ssx_sd_free(f);
		return rc;
	}
	*out = f;
	return SSX_OK;
}

/* |x - x*|_max / |x*|_max for the host solve of A x = A x*, x* a fixed
 * vector in [1, 2): about 1e-13 for a sound factor of these operators */
static int sd_check(const ssx_sdfactor *f, const int *rp, const int *ci,
		const double *a, double *err)
{
	const int n = f->n;
	double *xs = malloc(sizeof(double) * (size_t) n);
	double *b = calloc((size_t) n, sizeof(double));
	double *x = malloc(sizeof(double) * (size_t) n);
	int rc = xs && b && x ? SSX_OK : SSX_ENOMEM;
	if (rc == SSX_OK) {
		for (int i = 0; i < n; ++i)
			xs[i] = 1.0 + (double) (((unsigned) i * 2654435761u) >> 22)
				/ 1024.0;
		for (int i = 0; i < n; ++i) {
			double s = 0.0;
			for (int q = rp[i]; q < rp[i + 1]; ++q)
				s += a[q] * xs[ci[q]];
			b[i] = s;
		}
		rc = ssx_sd_solve_host(f, b, x);
	}
	if (rc == SSX_OK) {
		double e = 0.0, m = 0.0;
		for (int i = 0; i < n; ++i) {
			const double d = fabs(x[i] - xs[i]);
			e = d > e || d != d ? d : e;
			m = fabs(xs[i]) > m ? fabs(xs[i]) : m;
		}
		*err = e / m;
	}
	free(xs); free(b); free(x);
	return rc;
}

/* A factor with fronts from rocSOLVER / rocBLAS is checked before use:
 * the host solve of A x = A x* must return x* to 1e-9, or the factor is
 * computed again on the host alone.  A device front reported as not
 * positive definite is also settled by the host factorisation: under a
 * kernel profiler the device fronts have failed that way on operators the
 * host factors soundly. */
int ssx_sd_factor(int n, const int *rp, const int *ci, const double *a,
		const double *xyz, void *blas, void *stream, ssx_sdfactor **out)
{
	int rc = sd_factor_run(n, rp, ci, a, xyz, blas, stream, out);
	if (rc == SSX_ENUMERIC && blas)      /* the host decides definiteness */
		return sd_factor_run(n, rp, ci, a, xyz, NULL, NULL, out);
	if (rc != SSX_OK || (*out)->ndev == 0)
		return rc;
	double err = 0.0;
	rc = sd_check(*out, rp, ci, a, &err);
	if (rc == SSX_OK && err <= 1e-9)
		return SSX_OK;
	ssx_sd_free(*out);
	*out = NULL;
	if (rc != SSX_OK)
		return rc;
	return sd_factor_run(n, rp, ci, a, xyz, NULL, NULL, out);
}

/* The solve on the host, step for step as on the device (a check of the
 * factorisation and of the device kernels).  b and x in the system's own
 * numbering.
 *   forward (levels up):  b~ = b_C - (children's t on C),
 *                         y_C = M b~,  t = K b~ + (children's t on U)
 *   backward (levels down): x_C = M^T y_C - K^T x_U
 * (K = L21 M, so M^T (y - L21^T x_U) = M^T y - K^T x_U). */
int ssx_sd_solve_host(const ssx_sdfactor *f, const double *b, double *x)
{
	const int n = f->n;
	double *bn = malloc(sizeof(double) * (size_t) n);
	double *y = malloc(sizeof(double) * (size_t) n);
	double *xn = malloc(sizeof(double) * (size_t) n);
	double *t = malloc(sizeof(double) * (size_t) (f->nt > 0 ? f->nt : 1));
	double *w = malloc(sizeof(double) * (size_t) n);
	if (!bn || !y || !xn || !t || !w) {
		free(bn); free(y); free(xn); free(t); free(w);
		ssx_set_error("out of host memory (sparse direct solve)");
		return SSX_ENOMEM;
	}
	for (int i = 0; i < n; ++i)
		bn[i] = b[f->perm[i]];
	for (int l = 0; l < f->nlev; ++l)
		for (int k = f->lev_off[l]; k < f->lev_off[l + 1]; ++k) {
			const ssx_sdnode *p = &f->node[f->lev_sn[k]];
			const int nc = p->nc, nu = p->nu;
			const double *M = f->val + p->moff, *K = f->val + p->koff;
			const int *g0 = p->g0 >= 0 ? f->idx + p->g0 : NULL;
			const int *g1 = p->g1 >= 0 ? f->idx + p->g1 : NULL;
			const double *t0 = p->ch0 >= 0 ? t + f->node[p->ch0].toff : NULL;
			const double *t1 = p->ch1 >= 0 ? t + f->node[p->ch1].toff : NULL;
			for (int i = 0; i < nc; ++i) {
				double v = bn[p->c0 + i];
				if (g0 && g0[i] >= 0)
					v -= t0[g0[i]];
				if (g1 && g1[i] >= 0)
					v -= t1[g1[i]];
				w[i] = v;
			}
			for (int i = 0; i < nc; ++i) {
				double s = 0.0;
				for (int j = 0; j <= i; ++j)
					s += M[i + (size_t) j * nc] * w[j];
				y[p->c0 + i] = s;
			}
			for (int r = 0; r < nu; ++r) {
				double s = 0.0;
				for (int j = 0; j < nc; ++j)
					s += K[r + (size_t) j * nu] * w[j];
				double ps = 0.0;
				if (g0 && g0[nc + r] >= 0)
					ps += t0[g0[nc + r]];
				if (g1 && g1[nc + r] >= 0)
					ps += t1[g1[nc + r]];
				t[p->toff + r] = s + ps;
			}
		}
	for (int l = f->nlev - 1; l >= 0; --l)
		for (int k = f->lev_off[l]; k < f->lev_off[l + 1]; ++k) {
			const ssx_sdnode *p = &f->node[f->lev_sn[k]];
			const int nc = p->nc, nu = p->nu;
			const double *M = f->val + p->moff, *K = f->val + p->koff;
			const int *u = f->idx + p->uoff;
			for (int j = 0; j < nc; ++j) {
				double s = 0.0, q = 0.0;
				for (int i = j; i < nc; ++i)
					s += M[i + (size_t) j * nc] * y[p->c0 + i];
				for (int r = 0; r < nu; ++r)
					q += K[r + (size_t) j * nu] * xn[u[r]];
				w[j] = s - q;
			}
			for (int j = 0; j < nc; ++j)
				xn[p->c0 + j] = w[j];
		}
	for (int i = 0; i < n; ++i)
		x[f->perm[i]] = xn[i];
	free(bn); free(y); free(xn); free(t); free(w);
	return SSX_OK;
}

/* test hooks: factor a host system and solve it on the host, reporting
 * the supernode count, levels, factor entries, flops, the largest
 * supernode (columns, rows below) and the factor check's error */
int ssxh_sd_test(int n, const int *rp, const int *ci, const double *a,
		const double *xyz, const double *b, double *x, double *info)
{
	ssx_sdfactor *f = NULL;
	const int rc = ssx_sd_factor(n, rp, ci, a, xyz, NULL, NULL, &f);
	if (rc != SSX_OK)
		return rc;
	const int rs = ssx_sd_solve_host(f, b, x);
	if (info) {
		int maxnc = 0, maxnu = 0;
		for (int s = 0; s < f->nsn; ++s) {
			maxnc = f->node[s].nc > maxnc ? f->node[s].nc : maxnc;
			maxnu = f->node[s].nu > maxnu ? f->node[s].nu : maxnu;
		}
		info[0] = f->nsn;
		info[1] = f->nlev;
		info[2] = (double) f->nval;
		info[3] = f->flops;
		info[4] = maxnc;
		info[5] = maxnu;
		double err = -1.0;
		if (sd_check(f, rp, ci, a, &err) != SSX_OK)
			err = -1.0;
		info[6] = err;
	}
	ssx_sd_free(f);
	return rs;
}

/* ---- device side ---- */

static int sd_dup(void **dst, const void *src, size_t bytes)
{
	*dst = NULL;
	if (hipMalloc(dst, bytes > 0 ? bytes : 8) != hipSuccess) {
		*dst = NULL;
		ssx_set_error("hipMalloc of %zu bytes (sparse direct)", bytes);
		return SSX_ENOMEM;
	}
	if (src && bytes > 0 && hipMemcpy(*dst, src, bytes,
			hipMemcpyHostToDevice) != hipSuccess) {
		ssx_set_error("hipMemcpy (sparse direct upload)");
		return SSX_EHIP;
	}
	return SSX_OK;
}

int ssx_sd_upload(ssx_sdfactor *f)
{
	/* work tables: per level, the forward outputs (y rows, then t rows) of
	 * every supernode and the backward ones (x rows), in blocks sized by
	 * their row length */
	long long tot = 0;
	for (int s = 0; s < f->nsn; ++s) {
		const ssx_sdnode *p = &f->node[s];
		tot += ssxk_sd_blocks(p->nc + p->nu, p->nc)
			+ ssxk_sd_blocks(p->nc, p->nc + p->nu);
	}
	f->wk_off = malloc(sizeof(int) * (2 * (size_t) f->nlev + 1));
	f->wk_vec = calloc(2 * (size_t) f->nlev, sizeof(int));
	f->wk = malloc(sizeof(ssx_sdwork) * (size_t) (tot > 0 ? tot : 1));
	if (!f->wk_off || !f->wk_vec || !f->wk) {
		ssx_set_error("out of host memory (sparse direct tables)");
		return SSX_ENOMEM;
	}
	int m = 0;
	for (int l = 0; l < f->nlev; ++l)
		for (int d = 0; d < 2; ++d) {
			f->wk_off[2 * l + d] = m;
			for (int k = f->lev_off[l]; k < f->lev_off[l + 1]; ++k) {
				const int s = f->lev_sn[k];
				const ssx_sdnode *p = &f->node[s];
				const int nb = d == 0 ? ssxk_sd_blocks(p->nc + p->nu, p->nc)
					: ssxk_sd_blocks(p->nc, p->nc + p->nu);
				const int vec = d == 0 ? p->nc : p->nc + p->nu;
				if (vec > f->wk_vec[2 * l + d])
					f->wk_vec[2 * l + d] = vec;
				for (int b = 0; b < nb; ++b) {
					f->wk[m].sn = s;
					f->wk[m].blk = b;
					++m;
				}
			}
		}
	f->wk_off[2 * f->nlev] = m;
	const size_t vn = sizeof(double) * (size_t) f->n;
	int rc = sd_dup((void **) &f->d_node, f->node,
			sizeof(ssx_sdnode) * (size_t) f->nsn);
	if (!rc) rc = sd_dup((void **) &f->d_idx, f->idx,
			sizeof(int) * (size_t) f->nidx);
	if (!rc) rc = sd_dup((void **) &f->d_perm, f->perm,
			sizeof(int) * (size_t) f->n);
	if (!rc) rc = sd_dup((void **) &f->d_val, f->val,
			sizeof(double) * (size_t) f->nval);
	if (!rc) rc = sd_dup((void **) &f->d_wk, f->wk,
			sizeof(ssx_sdwork) * (size_t) m);
	if (!rc) rc = sd_dup((void **) &f->d_b, NULL, vn);
	if (!rc) rc = sd_dup((void **) &f->d_y, NULL, vn);
	if (!rc) rc = sd_dup((void **) &f->d_x, NULL, vn);
	if (!rc) rc = sd_dup((void **) &f->d_t, NULL,
			sizeof(double) * (size_t) f->nt);
	/* the host copies of the values are no longer needed */
	if (!rc) {
		free(f->val);
		f->val = NULL;
	}
	return rc;
}

static int sd_levels(ssx_sdfactor *f, void *st);

int ssx_sd_solve_dev_map(ssx_sdfactor *f, void *st, const double *src,
		const int *gmap, double *dst, const int *smap, int m)
{
	int rc = ssxk_sd_gather(st, f->n, gmap, src, f->d_b);
	if (rc == SSX_OK)
		rc = sd_levels(f, st);
	if (rc == SSX_OK)
		rc = ssxk_sd_scatter_map(st, m, smap, f->d_x, dst);
	return rc;
}

int ssx_sd_solve_dev(ssx_sdfactor *f, void *st, const double *b, double *x)
{
	int rc = ssxk_sd_gather(st, f->n, f->d_perm, b, f->d_b);
	if (rc == SSX_OK)
		rc = sd_levels(f, st);
	if (rc == SSX_OK)
		rc = ssxk_sd_scatter(st, f->n, f->d_perm, f->d_x, x);
	return rc;
}

/* one forward and one backward launch per level of the tree: d_b -> d_x */
static int sd_levels(ssx_sdfactor *f, void *st)
{
	int rc = SSX_OK;
	for (int l = 0; rc == SSX_OK && l < f->nlev; ++l) {
		const int *o = f->wk_off + 2 * l;
		rc = ssxk_sd_fwd(st, o[1] - o[0], f->d_wk + o[0], f->d_node,
				f->d_val, f->d_idx, f->d_b, f->d_y, f->d_t,
				f->wk_vec[2 * l]);
	}
	for (int l = f->nlev - 1; rc == SSX_OK && l >= 0; --l) {
		const int *o = f->wk_off + 2 * l;
		rc = ssxk_sd_bwd(st, o[2] - o[1], f->d_wk + o[1], f->d_node,
				f->d_val, f->d_idx, f->d_y, f->d_x, f->wk_vec[2 * l + 1]);
	}
	return rc;
}
