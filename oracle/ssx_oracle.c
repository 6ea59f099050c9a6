/*
 * ssx_oracle.c -- CPU restatement of the SSENSS elliptic-solve path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU path and the CPU
 * baseline of bench.py.  Never linked into the product library.
 *
 * Every function below follows the reference algorithm it cites (file:line
 * in the SSENSS src/ tree) in the same operation order where the reference
 * fixes one (pointwise products, the accumulation order of the 3x3 algebra,
 * the DSS add order).  Small dense products that the reference hands to BLAS
 * dgemm are summed here in index order; BLAS fixes no order, so those agree
 * with the reference to rounding only.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   gll basis, svv matrix     golden vectors from the compiled reference
 *                             (tests/golden), closed forms for n <= 5,
 *                             sum(rho) = 2, D*1 = 0, D*xi = 1, SBP identity
 *   interpolation matrices    golden vectors; polynomial reproduction
 *   geometry                  golden vectors; affine closed form
 *   ax (poisson, helmholtz)   golden vectors; constants -> 0, symmetry,
 *                             brute-force quadrature assembly at n = 3
 *   dss, scale, mask, mass    golden vectors (bit-exact on integer fields)
 *   dot, null space           golden vectors; <1,1>_w = unique node count
 *   restrict / prolong        golden vectors; polynomial reproduction
 *   cg, fcg, mg               golden iteration counts and solutions
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ssx_oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

#define NMAX ORC_NMAX
#define LMAX ORC_LMAX

typedef struct {
	int n, n3;
	double xi[NMAX], rho[NMAX];
	double d[NMAX * NMAX];          /* d[j*n+i] = l_j'(xi_i) = D_ij    */
	double dsvv[NMAX * NMAX];       /* same layout                      */
	double *w3;                     /* rho_i rho_j rho_k, [k][j][i]     */
} basis_t;

typedef struct {
	basis_t b;
	int n, n3;
	size_t ntot;                    /* nel * n3 */
	double *x, *y, *z;
	double *g[6];                   /* xx, xy, xz, yy, yz, zz */
	double *detj;
	double *diag_p, *diag_h;        /* inverses of the assembled diagonals */
	double *scale;                  /* 1 / multiplicity */
	double *mask;
	double nunique;                 /* <1,1>_w */
	double *prol;                   /* [p*n+i] = l_p^{level+1}(xi_i^{level}) */
	double *bufs;                   /* DSS pre-sum copies, nel*(6n^2+12n+8) */
	int **slot_nodes;               /* per slot 0..25: node offsets */
} level_t;

struct orc_ctx {
	int nel, nlev;
	int nmg[LMAX];
	level_t lev[LMAX];
	double *inv_mass;
	double *jac[9], *jinv[9];       /* level 0: J_ab, (J^-1)_ab         */
	double *ur[3][3];               /* [time level: 0 now, 1 old, 2 older][comp] */
	const int *neigh;               /* face neighbours (mesh)           */
	int *neigh_copy;
	double pdt, pt;                 /* params->dt, params->t for particles */
	int *bc;                        /* per element Dirichlet bit mask */
	int *part;                      /* nel*26 partner (elid*26+slot) or -1 */
	double lambda;
	int diag_choice;
	long long cnt[4];
	/* solver scratch, allocated on first use */
	double *mg_r[LMAX], *mg_fn[LMAX], *mg_e[LMAX];
	double *cg_r[LMAX], *cg_z[LMAX], *cg_p[LMAX], *cg_ap[LMAX];
};

static void *xcalloc(size_t n, size_t s)
{
	void *p = calloc(n ? n : 1, s);
	if (!p) {
		fprintf(stderr, "ssx_oracle: out of memory\n");
		abort();
	}
	return p;
}

/* ------------------------------------------------------------------ */
/* GLL basis: basis.c:75-188, deriv.c:81-112                           */
/* ------------------------------------------------------------------ */

/* Legendre P_n(x) by the three-term recurrence, basis.c:75-89 */
static double leg(double x, int n)
{
	double l0 = 1.0, l1 = x, l2 = 0.0;
	if (n <= 0)
		return l0;
	if (n == 1)
		return l1;
	for (int i = 1; i < n; ++i) {
		l2 = ((2.0 * (double) i + 1.0) / ((double) i + 1.0)) * x * l1
			- (((double) i) / ((double) i + 1.0)) * l0;
		l0 = l1;
		l1 = l2;
	}
	return l2;
}

/* P_n'(x), basis.c:91-101 */
static double legd(double x, int n)
{
	if (pow(x, 2) == 1.0)
		return pow(x, n - 1) * 0.5 * ((double) n) * ((double) (n + 1));
	return (x * leg(x, n) - leg(x, n - 1)) * ((double) n) / (pow(x, 2) - 1.0);
}

/* Roots of P_m in increasing order.  The reference takes them as the
 * eigenvalues of the Legendre Jacobi matrix (basis.c:103-121, LAPACK dgeev);
 * they only seed the GLL Newton iteration, so Newton on P_m itself, run to
 * full precision, is used here. */
static void gl_roots(int m, double *r)
{
	for (int i = 0; i < m; ++i) {
		double x = -cos(M_PI * ((double) i + 0.75) / ((double) m + 0.5));
		for (int it = 0; it < 100; ++it) {
			double p = leg(x, m);
			double dp = ((double) m) * (x * p - leg(x, m - 1))
				/ (x * x - 1.0);
			double dx = p / dp;
			x -= dx;
			if (fabs(dx) < 1e-17)
				break;
		}
		r[i] = x;
	}
	for (int i = 1; i < m; ++i)     /* insertion sort, ascending */
		for (int j = i; j > 0 && r[j - 1] > r[j]; --j) {
			double t = r[j];
			r[j] = r[j - 1];
			r[j - 1] = t;
		}
}

/* GLL points: Newton on (1-x^2) P_N'(x) from GL midpoints, basis.c:123-148 */
static void gll_points(int n, double *xi)
{
	const double eps = 1.e-12;
	xi[0] = -1.0;
	xi[n - 1] = 1.0;
	if (n < 3)
		return;
	double gl[NMAX];
	gl_roots(n - 1, gl);
	for (int i = 2; i < n; ++i) {
		xi[i - 1] = 0.5 * (gl[i - 2] + gl[i - 1]);
		double tmp = 0.0;
		while (fabs(xi[i - 1] - tmp) > eps) {
			tmp = xi[i - 1];
			xi[i - 1] = tmp + ((1.0 - pow(tmp, 2)) * legd(tmp, n - 1))
				/ (((double) (n - 1)) * ((double) n)
				* leg(tmp, n - 1));
		}
	}
}

/* basis.c:150-158 */
static void gll_weights(const double *xi, int n, double *rho)
{
	for (int i = 0; i < n; ++i)
		rho[i] = 2.0 / (((double) (n - 1)) * ((double) n)
				* pow(leg(xi[i], n - 1), 2));
}

/* basis.c:160-177: d[j][i] = l_j'(xi_i) */
static void deriv_matrix(const double *xi, int n, double *d)
{
	for (int j = 0; j < n; ++j)
		for (int i = 0; i < n; ++i)
			d[j * n + i] = (i == j) ? 0.0 : leg(xi[i], n - 1)
				/ (leg(xi[j], n - 1) * (xi[i] - xi[j]));
	d[0] = -((double) n) * ((double) (n - 1)) / 4.0;
	d[(n - 1) * n + (n - 1)] = -d[0];
}

/* deriv.c:106-112 */
static double svv_func(int n, int m, int i)
{
	if (i <= m)
		return 0.0;
	return exp(-pow(((double) (n - i)) / ((double) (m - i)), 2));
}

/* deriv.c:81-104 with basis.c:190-209.  Matrices in the reference layout
 * a[j][i] = A(i,j) (column-major A); products as in linalg.c:35-43. */
static void svv_matrix(const double *xi, const double *rho, const double *d,
		int n, int m, double c, double *out)
{
	double leg_t[NMAX * NMAX], inv_t[NMAX * NMAX], tmp[NMAX * NMAX];
	for (int j = 0; j < n; ++j)
		for (int i = 0; i < n; ++i) {
			leg_t[j * n + i] = rho[j] * leg(xi[j], i) * ((double) i + 0.5);
			inv_t[j * n + i] = leg(xi[i], j);
		}
	/* TMP(i,j) = sum_l LEG(i,l) D(l,j), LEG(i,l) = leg_t[l][i] */
	for (int j = 0; j < n; ++j)
		for (int i = 0; i < n; ++i) {
			double s = 0.0;
			for (int l = 0; l < n; ++l)
				s += leg_t[l * n + i] * d[j * n + l];
			tmp[j * n + i] = s;
		}
	for (int j = 0; j < n; ++j)
		for (int i = 0; i < n; ++i)
			tmp[j * n + i] *= sqrt(1.0 + c * svv_func(n - 1, m, i));
	for (int j = 0; j < n; ++j)
		for (int i = 0; i < n; ++i) {
			double s = 0.0;
			for (int l = 0; l < n; ++l)
				s += inv_t[l * n + i] * tmp[j * n + l];
			out[j * n + i] = s;
		}
}

static void basis_init(basis_t *b, int n, int svvn, double svva)
{
	b->n = n;
	b->n3 = n * n * n;
	gll_points(n, b->xi);
	gll_weights(b->xi, n, b->rho);
	deriv_matrix(b->xi, n, b->d);
	svv_matrix(b->xi, b->rho, b->d, n, svvn, svva, b->dsvv);
	b->w3 = xcalloc((size_t) b->n3, sizeof(double));
	for (int k = 0; k < n; ++k)             /* basis.c:179-188 */
		for (int j = 0; j < n; ++j)
			for (int i = 0; i < n; ++i)
				b->w3[(k * n + j) * n + i] =
					b->rho[i] * b->rho[j] * b->rho[k];
}

void orc_gll_basis(int n, int svvn, double svva, double *xi, double *rho,
		double *d, double *dsvv)
{
	basis_t b;
	basis_init(&b, n, svvn, svva);
	if (xi)
		memcpy(xi, b.xi, sizeof(double) * n);
	if (rho)
		memcpy(rho, b.rho, sizeof(double) * n);
	if (d)
		memcpy(d, b.d, sizeof(double) * n * n);
	if (dsvv)
		memcpy(dsvv, b.dsvv, sizeof(double) * n * n);
	free(b.w3);
}

/* Lagrange polynomial l_j of nodes x[0..n-1] at xi, interpolation.c:31-39 */
static double lagrange(double xi, const double *x, int j, int n)
{
	double li = 1.0;
	for (int i = 0; i < j; ++i)
		li *= (xi - x[i]) / (x[j] - x[i]);
	for (int i = j + 1; i < n; ++i)
		li *= (xi - x[i]) / (x[j] - x[i]);
	return li;
}

/* out[p*nto+i] = l_p^{from}(xi_i^{to}), interpolation.c:115-123 */
static void interp_matrix(const double *xfrom, int nfrom, const double *xto,
		int nto, double *out)
{
	for (int p = 0; p < nfrom; ++p)
		for (int i = 0; i < nto; ++i)
			out[p * nto + i] = lagrange(xto[i], xfrom, p, nfrom);
}

void orc_interp_matrix(int nfrom, int nto, double *j)
{
	double xf[NMAX], xt[NMAX];
	gll_points(nfrom, xf);
	gll_points(nto, xt);
	interp_matrix(xf, nfrom, xt, nto, j);
}

/* ------------------------------------------------------------------ */
/* Tensor contractions (deriv.c:29-66, interpolation.c:136-176)         */
/* m is in the reference layout m[l*n+i] = M_il                         */
/* ------------------------------------------------------------------ */

/* out = (M (x) I (x) I) u along xi:  out_kji = sum_l M_il u_kjl */
static void fwd_x(const double *m, int n, const double *u, double *out)
{
	for (int kj = 0; kj < n * n; ++kj)
		for (int i = 0; i < n; ++i) {
			double s = 0.0;
			for (int l = 0; l < n; ++l)
				s += m[l * n + i] * u[kj * n + l];
			out[kj * n + i] = s;
		}
}

/* out_kji = sum_l M_jl u_kli */
static void fwd_y(const double *m, int n, const double *u, double *out)
{
	for (int k = 0; k < n; ++k)
		for (int j = 0; j < n; ++j)
			for (int i = 0; i < n; ++i) {
				double s = 0.0;
				for (int l = 0; l < n; ++l)
					s += u[(k * n + l) * n + i] * m[l * n + j];
				out[(k * n + j) * n + i] = s;
			}
}

/* out_kji = sum_l M_kl u_lji */
static void fwd_z(const double *m, int n, const double *u, double *out)
{
	const int n2 = n * n;
	for (int k = 0; k < n; ++k)
		for (int ji = 0; ji < n2; ++ji) {
			double s = 0.0;
			for (int l = 0; l < n; ++l)
				s += u[l * n2 + ji] * m[l * n + k];
			out[k * n2 + ji] = s;
		}
}

/* transposes: out_kji (+)= sum_l M_li t_kjl, etc. (deriv.c:54-66) */
static void tr_x(const double *m, int n, const double *t, double *out)
{
	for (int kj = 0; kj < n * n; ++kj)
		for (int i = 0; i < n; ++i) {
			double s = 0.0;
			for (int l = 0; l < n; ++l)
				s += m[i * n + l] * t[kj * n + l];
			out[kj * n + i] = s;
		}
}

static void tr_y_add(const double *m, int n, const double *t, double *out)
{
	for (int k = 0; k < n; ++k)
		for (int j = 0; j < n; ++j)
			for (int i = 0; i < n; ++i) {
				double s = 0.0;
				for (int l = 0; l < n; ++l)
					s += t[(k * n + l) * n + i] * m[j * n + l];
				out[(k * n + j) * n + i] = s + out[(k * n + j) * n + i];
			}
}

static void tr_z_add(const double *m, int n, const double *t, double *out)
{
	const int n2 = n * n;
	for (int k = 0; k < n; ++k)
		for (int ji = 0; ji < n2; ++ji) {
			double s = 0.0;
			for (int l = 0; l < n; ++l)
				s += t[l * n2 + ji] * m[k * n + l];
			out[k * n2 + ji] = s + out[k * n2 + ji];
		}
}

/* Tensor transfer v = (A (x) A (x) A) u with A(i,p) = a[p*m+i], u of n^3,
 * v of m^3 (interpolate_to_grid, interpolation.c:136-155). */
static void transfer(const double *a, int m, int n, const double *u,
		double *v, double *t1, double *t2)
{
	for (int kj = 0; kj < n * n; ++kj)
		for (int i = 0; i < m; ++i) {
			double s = 0.0;
			for (int p = 0; p < n; ++p)
				s += a[p * m + i] * u[kj * n + p];
			t1[kj * m + i] = s;
		}
	for (int k = 0; k < n; ++k)
		for (int j = 0; j < m; ++j)
			for (int i = 0; i < m; ++i) {
				double s = 0.0;
				for (int l = 0; l < n; ++l)
					s += t1[(k * n + l) * m + i] * a[l * m + j];
				t2[(k * m + j) * m + i] = s;
			}
	for (int k = 0; k < m; ++k)
		for (int ji = 0; ji < m * m; ++ji) {
			double s = 0.0;
			for (int l = 0; l < n; ++l)
				s += t2[l * m * m + ji] * a[l * m + k];
			v[k * m * m + ji] = s;
		}
}

/* Transposed transfer v = (A^T (x) A^T (x) A^T) u with A(i,p) = a[p*n+i]
 * (i < n fine, p < m coarse), u of n^3, v of m^3 (integrate_to_grid,
 * interpolation.c:157-176). */
static void transfer_t(const double *a, int m, int n, const double *u,
		double *v, double *t1, double *t2)
{
	for (int kj = 0; kj < n * n; ++kj)
		for (int p = 0; p < m; ++p) {
			double s = 0.0;
			for (int c = 0; c < n; ++c)
				s += a[p * n + c] * u[kj * n + c];
			t1[kj * m + p] = s;
		}
	for (int k = 0; k < n; ++k)
		for (int j = 0; j < m; ++j)
			for (int i = 0; i < m; ++i) {
				double s = 0.0;
				for (int l = 0; l < n; ++l)
					s += t1[(k * n + l) * m + i] * a[j * n + l];
				t2[(k * m + j) * m + i] = s;
			}
	for (int k = 0; k < m; ++k)
		for (int ji = 0; ji < m * m; ++ji) {
			double s = 0.0;
			for (int l = 0; l < n; ++l)
				s += t2[l * m * m + ji] * a[k * n + l];
			v[k * m * m + ji] = s;
		}
}

/* ------------------------------------------------------------------ */
/* Geometry: geometry.c:102-135, deriv.c:68-79, tripalg.c:165-234       */
/* ------------------------------------------------------------------ */

static void geometry_element(const basis_t *b, const double *x,
		const double *y, const double *z, double *g[6], double *detj,
		size_t off, double **jout, double **jiout)
{
	const int n = b->n, n3 = b->n3;
	double *jac = xcalloc((size_t) 9 * n3, sizeof(double));
	const double *c[3] = { x, y, z };
	for (int a = 0; a < 3; ++a) {           /* J_ab = D_b x_a */
		fwd_x(b->d, n, c[a], jac + (a * 3 + 0) * n3);
		fwd_y(b->d, n, c[a], jac + (a * 3 + 1) * n3);
		fwd_z(b->d, n, c[a], jac + (a * 3 + 2) * n3);
	}
	for (int q = 0; q < n3; ++q) {
		double a[3][3], ai[3][3];
		for (int r = 0; r < 3; ++r)
			for (int s = 0; s < 3; ++s)
				a[r][s] = jac[(r * 3 + s) * n3 + q];
		/* determinant, tripalg.c:165-182 */
		double det = 0.0;
		{
			double t;
			t = a[1][1] * a[2][2];
			t = -a[1][2] * a[2][1] + t;
			det = a[0][0] * t;
			t = a[1][0] * a[2][2];
			t = -a[1][2] * a[2][0] + t;
			det = -a[0][1] * t + det;
			t = a[1][0] * a[2][1];
			t = -a[1][1] * a[2][0] + t;
			det = a[0][2] * t + det;
		}
		/* inverse by cofactors, tripalg.c:201-222 */
		static const int k0[3] = { 1, 0, 0 }, k1[3] = { 2, 2, 1 },
			k2[3] = { 2, 0, 1 }, k3[3] = { 1, 2, 0 };
		static const double sg[3] = { 1.0, -1.0, 1.0 };
		for (int i = 0; i < 3; ++i)
			for (int j = 0; j < 3; ++j) {
				double t = a[k0[j]][k3[i]] * a[k1[j]][k2[i]];
				t = -a[k0[j]][k2[i]] * a[k1[j]][k3[i]] + t;
				ai[i][j] = sg[j] * t / det;
			}
		/* G_ia = sum_j Jinv_aj Jinv_ij, geometry.c:130-135 with
		 * tripalg.c:224-234 (first term replaces, then adds) */
		double gm[3][3];
		for (int i = 0; i < 3; ++i)
			for (int aa = 0; aa < 3; ++aa) {
				double s = ai[aa][0] * ai[i][0];
				s = ai[aa][1] * ai[i][1] + s;
				s = ai[aa][2] * ai[i][2] + s;
				gm[i][aa] = s;
			}
		/* symmetric storage: row 1 and 2 writes overwrite the aliased
		 * upper entries with the same products (tripalg.c:91-132) */
		if (jout)
			for (int r = 0; r < 3; ++r)
				for (int s2 = 0; s2 < 3; ++s2) {
					jout[r * 3 + s2][off + q] = a[r][s2];
					jiout[r * 3 + s2][off + q] = ai[r][s2];
				}
		g[0][off + q] = gm[0][0];
		g[1][off + q] = gm[1][0];
		g[2][off + q] = gm[2][0];
		g[3][off + q] = gm[1][1];
		g[4][off + q] = gm[2][1];
		g[5][off + q] = gm[2][2];
		detj[off + q] = det;
	}
	free(jac);
}

/* ------------------------------------------------------------------ */
/* DSS wiring: init.c:68-130, 298-320; buffers global.c:38-177          */
/* slots 0..5 faces W,E,S,N,B,T, 6..17 edges, 18..25 corners           */
/* ------------------------------------------------------------------ */

static int slot_len(int s, int n)
{
	return s < 6 ? n * n : (s < 18 ? n : 1);
}

/* node offset (k*n+j)*n+i of position t in slot s, global.c:38-166 */
static int slot_node(int s, int t, int n)
{
	const int m = n - 1;
	int a = t / n, b = t % n, k, j, i;
	switch (s) {
	case 0: k = a; j = b; i = 0; break;      /* w[k][j] = u[k][j][0]   */
	case 1: k = a; j = b; i = m; break;
	case 2: k = a; j = 0; i = b; break;      /* s[k][i] = u[k][0][i]   */
	case 3: k = a; j = m; i = b; break;
	case 4: k = 0; j = a; i = b; break;      /* b[j][i] = u[0][j][i]   */
	case 5: k = m; j = a; i = b; break;
	case 6: k = t; j = 0; i = 0; break;      /* sw */
	case 7: k = t; j = 0; i = m; break;      /* se */
	case 8: k = t; j = m; i = 0; break;      /* nw */
	case 9: k = t; j = m; i = m; break;      /* ne */
	case 10: k = 0; j = t; i = 0; break;     /* bw */
	case 11: k = 0; j = t; i = m; break;     /* be */
	case 12: k = m; j = t; i = 0; break;     /* tw */
	case 13: k = m; j = t; i = m; break;     /* te */
	case 14: k = 0; j = 0; i = t; break;     /* bs */
	case 15: k = 0; j = m; i = t; break;     /* bn */
	case 16: k = m; j = 0; i = t; break;     /* ts */
	case 17: k = m; j = m; i = t; break;     /* tn */
	default: {
		int cc = s - 18;                    /* global.c:157-166 */
		i = (cc % 2) * m;
		j = ((cc / 2) % 2) * m;
		k = (cc / 4) * m;
	}
	}
	return (k * n + j) * n + i;
}

static int slot_offset(int s, int n)
{
	if (s < 6)
		return s * n * n;
	if (s < 18)
		return 6 * n * n + (s - 6) * n;
	return 6 * n * n + 12 * n + (s - 18);
}

static int wire(orc_ctx *c, const orc_mesh *m)
{
	const int nel = m->nel;
	c->part = xcalloc((size_t) nel * 26, sizeof(int));
	for (int e = 0; e < nel; ++e)
		for (int s = 0; s < 26; ++s) {
			const int *nb;
			int ns, base, loc, sz = 1;
			if (s < 6) {
				nb = m->neigh; ns = 6; base = 0; loc = s;
			} else if (s < 18) {
				nb = m->edge_neigh; ns = 12; base = 6; loc = s - 6;
				sz = m->edge_size[e * 12 + loc];
			} else {
				nb = m->corn_neigh; ns = 8; base = 18; loc = s - 18;
				sz = m->corn_size[e * 8 + loc];
			}
			int el = nb[e * ns + loc];
			c->part[e * 26 + s] = -1;
			if (el < 0 || sz == 0)
				continue;
			if (sz != 1 || el >= nel)
				return -1;      /* init.c:183: size must be 1 */
			int si = -1;            /* first slot of el pointing back */
			for (int q = 0; q < ns; ++q)
				if (nb[el * ns + q] == e) {
					si = q;
					break;
				}
			if (si < 0)
				return -1;
			c->part[e * 26 + s] = el * 26 + base + si;
		}
	return 0;
}

static void dss(orc_ctx *c, int level, double *u)
{
	level_t *L = &c->lev[level];
	const int n = L->n, n3 = L->n3;
	const int bl = 6 * n * n + 12 * n + 8;
	/* phase 1: every element copies its wired faces/edges/corners
	 * (global.c:243-247) */
#pragma omp parallel for schedule(static)
	for (int e = 0; e < c->nel; ++e)
		for (int s = 0; s < 26; ++s) {
			if (c->part[e * 26 + s] < 0)
				continue;
			double *buf = L->bufs + (size_t) e * bl + slot_offset(s, n);
			const int *nd = L->slot_nodes[s];
			for (int t = 0; t < slot_len(s, n); ++t)
				buf[t] = u[(size_t) e * n3 + nd[t]];
		}
	/* phase 2: faces, then edges, then corners of the partners are added
	 * in slot order (global.c:252-256) */
#pragma omp parallel for schedule(static)
	for (int e = 0; e < c->nel; ++e)
		for (int s = 0; s < 26; ++s) {
			int pp = c->part[e * 26 + s];
			if (pp < 0)
				continue;
			int pe = pp / 26, ps = pp % 26;
			const double *buf = L->bufs + (size_t) pe * bl
				+ slot_offset(ps, n);
			const int *nd = L->slot_nodes[s];
			for (int t = 0; t < slot_len(s, n); ++t)
				u[(size_t) e * n3 + nd[t]] += buf[t];
		}
}

/* ------------------------------------------------------------------ */
/* Element operators: poisson.c:34-66, helmholtz.c:33-49                */
/* ------------------------------------------------------------------ */

static void ax_element(const level_t *L, int op, double lambda, size_t off,
		const double *u, double *au, double *scr)
{
	const int n = L->n, n3 = L->n3;
	const double *dd = L->b.d;
	const double *ds = (op == ORC_OP_HELMHOLTZ) ? L->b.dsvv : L->b.d;
	double *g0 = scr, *g1 = scr + n3, *g2 = scr + 2 * n3;
	double *t0 = scr + 3 * n3, *t1 = scr + 4 * n3, *t2 = scr + 5 * n3;
	/* helmholtz.c:41-44: the swapped matrix reaches the xi-forward and
	 * the eta/zeta-transpose products only */
	fwd_x(ds, n, u, g0);
	fwd_y(dd, n, u, g1);
	fwd_z(dd, n, u, g2);
	const double *gxx = L->g[0] + off, *gxy = L->g[1] + off,
		*gxz = L->g[2] + off, *gyy = L->g[3] + off,
		*gyz = L->g[4] + off, *gzz = L->g[5] + off;
	const double *w = L->b.w3, *jd = L->detj + off;
	for (int q = 0; q < n3; ++q) {
		double a = gxx[q] * g0[q];
		a = gxy[q] * g1[q] + a;
		a = gxz[q] * g2[q] + a;
		double b = gxy[q] * g0[q];
		b = gyy[q] * g1[q] + b;
		b = gyz[q] * g2[q] + b;
		double cc = gxz[q] * g0[q];
		cc = gyz[q] * g1[q] + cc;
		cc = gzz[q] * g2[q] + cc;
		t0[q] = (a * w[q]) * jd[q];     /* poisson.c:58-66 */
		t1[q] = (b * w[q]) * jd[q];
		t2[q] = (cc * w[q]) * jd[q];
	}
	tr_x(dd, n, t0, au);
	tr_y_add(ds, n, t1, au);
	tr_z_add(ds, n, t2, au);
	if (op == ORC_OP_HELMHOLTZ)
		for (int q = 0; q < n3; ++q)
			au[q] = lambda * ((u[q] * w[q]) * jd[q]) + au[q];
}

void orc_ax(orc_ctx *c, int level, int op, const double *u, double *au)
{
	level_t *L = &c->lev[level];
#pragma omp parallel
	{
		double *scr = xcalloc((size_t) 6 * L->n3, sizeof(double));
#pragma omp for schedule(static)
		for (int e = 0; e < c->nel; ++e) {
			size_t off = (size_t) e * L->n3;
			ax_element(L, op, c->lambda, off, u + off, au + off, scr);
		}
		free(scr);
	}
}

/* ------------------------------------------------------------------ */
/* Vector primitives (linalg.c:45-164, global.c:259-282)                */
/* ------------------------------------------------------------------ */

double orc_dot(orc_ctx *c, int level, const double *x, const double *y)
{
	level_t *L = &c->lev[level];
	const int n3 = L->n3;
	double *part = xcalloc((size_t) c->nel, sizeof(double));
#pragma omp parallel for schedule(static)
	for (int e = 0; e < c->nel; ++e) {
		double s = 0.0;
		size_t off = (size_t) e * n3;
		for (int q = 0; q < n3; ++q)
			s += (x[off + q] * L->scale[off + q]) * y[off + q];
		part[e] = s;
	}
	double s = 0.0;
	for (int e = 0; e < c->nel; ++e)
		s += part[e];
	free(part);
	return s;
}

void orc_dss(orc_ctx *c, int level, double *u)
{
	dss(c, level, u);
}

void orc_dss_average(orc_ctx *c, int level, double *u)
{
	level_t *L = &c->lev[level];
	dss(c, level, u);
	for (size_t q = 0; q < L->ntot; ++q)
		u[q] = u[q] * L->scale[q];
}

void orc_mask(orc_ctx *c, int level, int var, double *u)
{
	if (var == ORC_PRESSURE)        /* poisson.c:91-92 */
		return;
	level_t *L = &c->lev[level];
	for (size_t q = 0; q < L->ntot; ++q)
		u[q] = u[q] * L->mask[q];
}

void orc_subtract_nullspace(orc_ctx *c, int level, int var, double *u)
{
	if (var != ORC_PRESSURE)        /* solver.c:200-213 */
		return;
	level_t *L = &c->lev[level];
	double s = 0.0;
	for (size_t q = 0; q < L->ntot; ++q)
		s += (u[q] * L->scale[q]) * 1.0;
	const double a = -s / L->nunique;
	for (size_t q = 0; q < L->ntot; ++q)
		u[q] = a * 1.0 + u[q];
}

void orc_opeval(orc_ctx *c, int level, int op, int var, const double *p,
		double *ap)
{
	if (level == 0)
		c->cnt[0]++;
	orc_ax(c, level, op, p, ap);    /* solver.c:191-198 */
	orc_mask(c, level, var, ap);
	dss(c, level, ap);
}

static void level_scratch(orc_ctx *c, int l)
{
	size_t nt = c->lev[l].ntot;
	if (!c->cg_r[l]) {
		c->cg_r[l] = xcalloc(nt, sizeof(double));
		c->cg_z[l] = xcalloc(nt, sizeof(double));
		c->cg_p[l] = xcalloc(nt, sizeof(double));
		c->cg_ap[l] = xcalloc(nt, sizeof(double));
		c->mg_r[l] = xcalloc(nt, sizeof(double));
		c->mg_fn[l] = xcalloc(nt, sizeof(double));
		c->mg_e[l] = xcalloc(nt, sizeof(double));
	}
}

void orc_restrict(orc_ctx *c, int level, const double *fine, double *coarse)
{
	level_t *F = &c->lev[level - 1], *C = &c->lev[level];
	const int nf = F->n, nc = C->n;
#pragma omp parallel
	{
		double *t1 = xcalloc((size_t) nf * nf * nf, sizeof(double));
		double *t2 = xcalloc((size_t) nf * nf * nf, sizeof(double));
#pragma omp for schedule(static)
		for (int e = 0; e < c->nel; ++e)
			transfer_t(F->prol, nc, nf, fine + (size_t) e * F->n3,
					coarse + (size_t) e * C->n3, t1, t2);
		free(t1);
		free(t2);
	}
	orc_dss_average(c, level, coarse);      /* precon.c:283-291 */
}

void orc_prolong(orc_ctx *c, int level, const double *coarse, double *fine)
{
	level_t *F = &c->lev[level], *C = &c->lev[level + 1];
	const int nf = F->n, nc = C->n;
#pragma omp parallel
	{
		double *t1 = xcalloc((size_t) nf * nf * nf, sizeof(double));
		double *t2 = xcalloc((size_t) nf * nf * nf, sizeof(double));
#pragma omp for schedule(static)
		for (int e = 0; e < c->nel; ++e)
			transfer(F->prol, nf, nc, coarse + (size_t) e * C->n3,
					fine + (size_t) e * F->n3, t1, t2);
		free(t1);
		free(t2);
	}
}

static const double *jacobi_diag(orc_ctx *c, int level, int var)
{
	int ch = c->diag_choice;
	if (ch < 0)
		ch = (var == ORC_PRESSURE) ? 0 : 1;     /* precon.c:211-225 */
	return ch == 0 ? c->lev[level].diag_p : c->lev[level].diag_h;
}

static void jacobi(orc_ctx *c, int level, int var, const double *x,
		double *px)
{
	const double *dg = jacobi_diag(c, level, var);
	for (size_t q = 0; q < c->lev[level].ntot; ++q)
		px[q] = x[q] * dg[q];           /* precon.c:317-324 */
}

/* ------------------------------------------------------------------ */
/* Solvers: solver.c:33-157, precon.c:227-309                           */
/* ------------------------------------------------------------------ */

static double cg_level(orc_ctx *c, int l, int op, int var, double *f,
		double *u, double eps, int max_iter, int *iters, int kind)
{
	level_scratch(c, l);
	const size_t nt = c->lev[l].ntot;
	double *r = c->cg_r[l], *z = c->cg_z[l], *p = c->cg_p[l],
		*ap = c->cg_ap[l];
	orc_mask(c, l, var, f);
	orc_opeval(c, l, op, var, u, r);
	for (size_t q = 0; q < nt; ++q)
		r[q] = 1.0 * f[q] + (-1.0 * r[q]);
	orc_subtract_nullspace(c, l, var, r);
	jacobi(c, l, var, r, z);
	orc_subtract_nullspace(c, l, var, z);
	memcpy(p, z, nt * sizeof(double));
	double rdz = orc_dot(c, l, r, z);
	int j = 0, it = 0;
	double err;
	do {
		orc_opeval(c, l, op, var, p, ap);
		const double alpha = rdz / orc_dot(c, l, p, ap);
		for (size_t q = 0; q < nt; ++q) {
			u[q] = alpha * p[q] + u[q];
			r[q] = -alpha * ap[q] + r[q];
		}
		jacobi(c, l, var, r, z);
		orc_subtract_nullspace(c, l, var, z);
		const double rdz_old = rdz;
		rdz = orc_dot(c, l, r, z);
		const double beta = rdz / rdz_old;
		for (size_t q = 0; q < nt; ++q)
			p[q] = 1.0 * z[q] + beta * p[q];
		orc_mask(c, l, var, p);
		++it;
	} while ((err = sqrt(orc_dot(c, l, r, r))) > eps && ++j < max_iter);
	if (iters)
		*iters = it;
	if (kind == 1)
		c->cnt[1] += it;
	else if (kind == 2)
		c->cnt[2] += it;
	return err;
}

double orc_cg(orc_ctx *c, int level, int op, int var, double *f, double *u,
		double eps, int max_iter, int *iters)
{
	return cg_level(c, level, op, var, f, u, eps, max_iter, iters, 0);
}

static double mg_recurse(orc_ctx *c, int l, int op, int var, double *f,
		double *x, double eps)
{
	if (l == c->nlev - 1)           /* precon.c:242-244 */
		return cg_level(c, l, op, var, f, x, eps, 500, NULL, 1);
	level_scratch(c, l);
	level_scratch(c, l + 1);
	const size_t nt = c->lev[l].ntot, nc = c->lev[l + 1].ntot;
	double *r = c->mg_r[l], *fn = c->mg_fn[l + 1], *e = c->mg_e[l + 1];
	if (l == 0) {                   /* compute_residual, precon.c:302-309 */
		orc_opeval(c, l, op, var, x, r);
		for (size_t q = 0; q < nt; ++q)
			r[q] = 1.0 * f[q] + (-1.0 * r[q]);
	} else {
		memcpy(r, f, nt * sizeof(double));
	}
	orc_restrict(c, l + 1, r, fn);
	memset(e, 0, nc * sizeof(double));
	mg_recurse(c, l + 1, op, var, fn, e, eps);
	orc_prolong(c, l, e, r);
	for (size_t q = 0; q < nt; ++q)
		x[q] = 1.0 * r[q] + x[q];
	return cg_level(c, l, op, var, f, x, eps, 15, NULL, 2);
}

void orc_mg_precon(orc_ctx *c, int op, int var, double eps, const double *x,
		double *px)
{
	/* precon.c:227-236; f is masked in place by the smoother, so work on
	 * the caller's vector exactly as the reference does */
	c->cnt[3]++;
	jacobi(c, 0, var, x, px);
	mg_recurse(c, 0, op, var, (double *) x, px, eps);
}

double orc_fcg(orc_ctx *c, int op, int var, int precon, double *f, double *u,
		double eps, int max_iter, int *iters)
{
	const int mmax = (precon == ORC_PC_MG) ? 10 : 1;
	const size_t nt = c->lev[0].ntot;
	double *r = xcalloc(nt, sizeof(double)), *z = xcalloc(nt, sizeof(double));
	double *p[11], *ap[10], pap[10] = { 0 }, beta[10] = { 0 };
	for (int i = 0; i < mmax + 1; ++i)
		p[i] = xcalloc(nt, sizeof(double));
	for (int i = 0; i < mmax; ++i)
		ap[i] = xcalloc(nt, sizeof(double));
	orc_mask(c, 0, var, f);
	orc_opeval(c, 0, op, var, u, r);
	for (size_t q = 0; q < nt; ++q)
		r[q] = 1.0 * f[q] + (-1.0 * r[q]);
	orc_subtract_nullspace(c, 0, var, r);
	long j = 0;
	int it = 0;
	double err;
	do {
		int m = (int) j, mm = (int) (j % (mmax + 1));
		if (mm < 1)
			mm = 1;
		if (mm < m)
			m = mm;
		if (precon == ORC_PC_MG)
			orc_mg_precon(c, op, var, eps, r, z);
		else
			jacobi(c, 0, var, r, z);
		orc_subtract_nullspace(c, 0, var, z);
		for (int k = 0; k < m; ++k)
			beta[k] = orc_dot(c, 0, z, ap[k]) / pap[k];
		/* p[k] -> p[k+1] for k < m (solver.c:121-122), by rotation */
		double *tp = p[m];
		for (int k = m; k > 0; --k)
			p[k] = p[k - 1];
		p[0] = tp;
		memcpy(p[0], z, nt * sizeof(double));
		for (int k = 1; k < m + 1; ++k)
			for (size_t q = 0; q < nt; ++q)
				p[0][q] = -beta[k - 1] * p[k][q] + p[0][q];
		orc_mask(c, 0, var, p[0]);
		int m2 = (int) (j + 1), mm2 = (int) ((j + 1) % (mmax + 1));
		if (mm2 < 1)
			mm2 = 1;
		if (mm2 < m2)
			m2 = mm2;
		double *ta = ap[m2 - 1];        /* ap[k] -> ap[k+1], k < m2-1 */
		for (int k = m2 - 1; k > 0; --k)
			ap[k] = ap[k - 1];
		ap[0] = ta;
		orc_opeval(c, 0, op, var, p[0], ap[0]);
		for (int k = m2 - 1; k > 0; --k)
			pap[k] = pap[k - 1];
		pap[0] = orc_dot(c, 0, p[0], ap[0]);
		const double alpha = orc_dot(c, 0, p[0], r) / pap[0];
		for (size_t q = 0; q < nt; ++q) {
			u[q] = alpha * p[0][q] + u[q];
			r[q] = -alpha * ap[0][q] + r[q];
		}
		++it;
	} while ((err = sqrt(orc_dot(c, 0, r, r))) > eps && ++j < max_iter);
	if (iters)
		*iters = it;
	for (int i = 0; i < mmax + 1; ++i)
		free(p[i]);
	for (int i = 0; i < mmax; ++i)
		free(ap[i]);
	free(r);
	free(z);
	return err;
}

/* ------------------------------------------------------------------ */
/* Setup: precon.c:58-83, 121-209, 336-413                              */
/* ------------------------------------------------------------------ */

static void level_alloc(orc_ctx *c, level_t *L)
{
	L->ntot = (size_t) c->nel * L->n3;
	L->x = xcalloc(L->ntot, sizeof(double));
	L->y = xcalloc(L->ntot, sizeof(double));
	L->z = xcalloc(L->ntot, sizeof(double));
	for (int a = 0; a < 6; ++a)
		L->g[a] = xcalloc(L->ntot, sizeof(double));
	L->detj = xcalloc(L->ntot, sizeof(double));
	L->diag_p = xcalloc(L->ntot, sizeof(double));
	L->diag_h = xcalloc(L->ntot, sizeof(double));
	L->scale = xcalloc(L->ntot, sizeof(double));
	L->mask = xcalloc(L->ntot, sizeof(double));
	const int n = L->n;
	L->bufs = xcalloc((size_t) c->nel * (6 * n * n + 12 * n + 8),
			sizeof(double));
	L->slot_nodes = xcalloc(26, sizeof(int *));
	for (int s = 0; s < 26; ++s) {
		L->slot_nodes[s] = xcalloc((size_t) slot_len(s, n), sizeof(int));
		for (int t = 0; t < slot_len(s, n); ++t)
			L->slot_nodes[s][t] = slot_node(s, t, n);
	}
}

/* approximate diagonal, helmholtz.c:51-68 */
static double diag_entry(const level_t *L, size_t off, int i, int j, int k,
		double lambda)
{
	const int n = L->n;
	const double *d = L->b.d, *w = L->b.w3;
	const double *g11 = L->g[0] + off, *g22 = L->g[3] + off,
		*g33 = L->g[5] + off;
	double tmp = 0.0;
#define IX(kk, jj, ii) (((kk) * n + (jj)) * n + (ii))
	for (int p = 0; p < n; ++p) {
		tmp += w[IX(k, j, p)] * pow(d[i * n + p], 2.0) * g11[IX(k, j, p)];
		tmp += w[IX(k, p, i)] * pow(d[j * n + p], 2.0) * g22[IX(k, p, i)];
		tmp += w[IX(p, j, i)] * pow(d[k * n + p], 2.0) * g33[IX(p, j, i)];
	}
	tmp += lambda * w[IX(k, j, i)] * L->detj[off + IX(k, j, i)];
#undef IX
	return tmp;
}

orc_ctx *orc_create(const orc_mesh *m)
{
	if (m->n < 2 || m->n > NMAX || m->nel < 1)
		return NULL;
	orc_ctx *c = xcalloc(1, sizeof(*c));
	c->nel = m->nel;
	c->diag_choice = -1;
	c->lambda = (11.0 / 6.0) / (m->dt * m->visc);
	/* level orders, precon.c:62-70 */
	int nl = LMAX;
	for (int i = 0; i < LMAX; ++i) {
		int v = (m->n - 1) / (1 << i) + 1;
		c->nmg[i] = v > 2 ? v : 2;
		if (c->nmg[i] == 2) {
			nl = i + 1;
			break;
		}
	}
	c->nlev = nl;
	if (wire(c, m) != 0) {
		fprintf(stderr, "ssx_oracle: unsupported connectivity\n");
		free(c->part);
		free(c);
		return NULL;
	}
	c->bc = xcalloc((size_t) m->nel, sizeof(int));
	for (int e = 0; e < m->nel; ++e)         /* geometry.c:230-237 */
		for (int s = 0; s < 6; ++s)
			if (m->neigh[e * 6 + s] == -1)
				c->bc[e] |= 1 << s;
	for (int l = 0; l < nl; ++l) {
		level_t *L = &c->lev[l];
		L->n = c->nmg[l];
		L->n3 = L->n * L->n * L->n;
		basis_init(&L->b, L->n, m->svvn, m->svva);
		level_alloc(c, L);
	}
	/* coordinates: level 0 as given, coarser ones interpolated from the
	 * next finer level (precon.c:137-153) */
	memcpy(c->lev[0].x, m->x, c->lev[0].ntot * sizeof(double));
	memcpy(c->lev[0].y, m->y, c->lev[0].ntot * sizeof(double));
	memcpy(c->lev[0].z, m->z, c->lev[0].ntot * sizeof(double));
	for (int l = 0; l + 1 < nl; ++l) {
		level_t *F = &c->lev[l], *C = &c->lev[l + 1];
		F->prol = xcalloc((size_t) F->n * C->n, sizeof(double));
		interp_matrix(C->b.xi, C->n, F->b.xi, F->n, F->prol);
		double *q = xcalloc((size_t) F->n * C->n, sizeof(double));
		interp_matrix(F->b.xi, F->n, C->b.xi, C->n, q);
		double *t1 = xcalloc((size_t) F->n3, sizeof(double));
		double *t2 = xcalloc((size_t) F->n3, sizeof(double));
		for (int e = 0; e < c->nel; ++e) {
			size_t fo = (size_t) e * F->n3, co = (size_t) e * C->n3;
			transfer(q, C->n, F->n, F->x + fo, C->x + co, t1, t2);
			transfer(q, C->n, F->n, F->y + fo, C->y + co, t1, t2);
			transfer(q, C->n, F->n, F->z + fo, C->z + co, t1, t2);
		}
		free(t1);
		free(t2);
		free(q);
	}
	for (int a = 0; a < 9; ++a) {
		c->jac[a] = xcalloc(c->lev[0].ntot, sizeof(double));
		c->jinv[a] = xcalloc(c->lev[0].ntot, sizeof(double));
	}
	for (int tl = 0; tl < 3; ++tl)
		for (int a = 0; a < 3; ++a)
			c->ur[tl][a] = xcalloc(c->lev[0].ntot, sizeof(double));
	c->neigh_copy = xcalloc((size_t) m->nel * 6, sizeof(int));
	memcpy(c->neigh_copy, m->neigh, sizeof(int) * 6 * (size_t) m->nel);
	for (int l = 0; l < nl; ++l) {
		level_t *L = &c->lev[l];
		for (int e = 0; e < c->nel; ++e) {
			size_t off = (size_t) e * L->n3;
			geometry_element(&L->b, L->x + off, L->y + off, L->z + off,
					L->g, L->detj, off, l == 0 ? c->jac : NULL,
					l == 0 ? c->jinv : NULL);
		}
	}
	/* Jacobi diagonals, precon.c:336-359 */
	for (int l = 0; l < nl; ++l) {
		level_t *L = &c->lev[l];
		const int n = L->n;
		double *a = xcalloc(L->ntot, sizeof(double));
		for (int pass = 0; pass < 2; ++pass) {
			const double lam = (11.0 / 6.0) * (double) (pass != 0)
				/ (m->dt * m->visc);
			for (int e = 0; e < c->nel; ++e) {
				size_t off = (size_t) e * L->n3;
				for (int k = 0; k < n; ++k)
					for (int j = 0; j < n; ++j)
						for (int i = 0; i < n; ++i)
							a[off + (k * n + j) * n + i] =
								diag_entry(L, off, i, j, k, lam);
			}
			dss(c, l, a);
			double *dst = pass ? L->diag_h : L->diag_p;
			for (size_t q = 0; q < L->ntot; ++q)
				dst[q] = 1.0 * 1.0 / a[q];
		}
		/* multiplicity scaling, precon.c:361-381 */
		for (size_t q = 0; q < L->ntot; ++q)
			a[q] = 1.0;
		dss(c, l, a);
		for (size_t q = 0; q < L->ntot; ++q)
			L->scale[q] = 1.0 * 1.0 / a[q];
		/* Dirichlet mask, precon.c:383-395 with poisson.c:101-143 */
		for (size_t q = 0; q < L->ntot; ++q)
			L->mask[q] = 1.0;
		for (int e = 0; e < c->nel; ++e) {
			double *mk = L->mask + (size_t) e * L->n3;
			for (int s = 0; s < 6; ++s) {
				if (!(c->bc[e] & (1 << s)))
					continue;
				for (int t = 0; t < n * n; ++t)
					mk[slot_node(s, t, n)] = 0.0;
			}
		}
		double s = 0.0;
		for (size_t q = 0; q < L->ntot; ++q)
			s += (1.0 * L->scale[q]) * 1.0;
		L->nunique = s;
		free(a);
	}
	/* inverse mass, precon.c:397-413 */
	{
		level_t *L = &c->lev[0];
		c->inv_mass = xcalloc(L->ntot, sizeof(double));
		double *y = xcalloc(L->ntot, sizeof(double));
		for (int e = 0; e < c->nel; ++e)
			for (int q = 0; q < L->n3; ++q) {
				size_t o = (size_t) e * L->n3 + q;
				y[o] = (1.0 * L->b.w3[q]) * L->detj[o];
			}
		dss(c, 0, y);
		for (size_t q = 0; q < L->ntot; ++q)
			c->inv_mass[q] = 1.0 * 1.0 / y[q];
		free(y);
	}
	return c;
}

void orc_destroy(orc_ctx *c)
{
	if (!c)
		return;
	for (int l = 0; l < c->nlev; ++l) {
		level_t *L = &c->lev[l];
		free(L->b.w3);
		free(L->x);
		free(L->y);
		free(L->z);
		for (int a = 0; a < 6; ++a)
			free(L->g[a]);
		free(L->detj);
		free(L->diag_p);
		free(L->diag_h);
		free(L->scale);
		free(L->mask);
		free(L->prol);
		free(L->bufs);
		for (int s = 0; s < 26; ++s)
			free(L->slot_nodes[s]);
		free(L->slot_nodes);
		free(c->cg_r[l]);
		free(c->cg_z[l]);
		free(c->cg_p[l]);
		free(c->cg_ap[l]);
		free(c->mg_r[l]);
		free(c->mg_fn[l]);
		free(c->mg_e[l]);
	}
	free(c->inv_mass);
	for (int a = 0; a < 9; ++a) {
		free(c->jac[a]);
		free(c->jinv[a]);
	}
	for (int tl = 0; tl < 3; ++tl)
		for (int a = 0; a < 3; ++a)
			free(c->ur[tl][a]);
	free(c->neigh_copy);
	free(c->bc);
	free(c->part);
	free(c);
}

int orc_nlevels(const orc_ctx *c) { return c->nlev; }
int orc_level_n(const orc_ctx *c, int level) { return c->lev[level].n; }
int orc_nel(const orc_ctx *c) { return c->nel; }
double orc_lambda(const orc_ctx *c) { return c->lambda; }
void orc_set_diag_choice(orc_ctx *c, int choice) { c->diag_choice = choice; }

void orc_counters(const orc_ctx *c, long long *out4)
{
	for (int i = 0; i < 4; ++i)
		out4[i] = c->cnt[i];
}

const double *orc_level_data(const orc_ctx *c, int level, int which)
{
	if (level < 0 || level >= c->nlev)
		return NULL;
	const level_t *L = &c->lev[level];
	switch (which) {
	case 0: case 1: case 2: case 3: case 4: case 5:
		return L->g[which];
	case 6: return L->detj;
	case 7: return L->diag_p;
	case 8: return L->diag_h;
	case 9: return L->scale;
	case 10: return L->mask;
	case 11: return level == 0 ? c->inv_mass : NULL;
	case 12: return L->x;
	case 13: return L->y;
	case 14: return L->z;
	case 15: return L->b.d;
	case 16: return L->b.dsvv;
	case 17: return L->prol;
	default: return NULL;
	}
}

void orc_set_threads(int nthreads)
{
#ifdef _OPENMP
	omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#else
	(void) nthreads;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* Particles: interpolation.c:41-104, particles.c:182-402, 493-539      */
/* ------------------------------------------------------------------ */

/* interpolation.c:98-104 */
static int first_point(const double *x, double xp, int n, int m)
{
	for (int i = 0; i < n - m + 1; ++i)
		if (x[i] > xp)
			return (i - m / 2) > 0 ? i - m / 2 : 0;
	return n - m;
}

double orc_interp_point(int n, const double *xi, const double *f, double x,
		double y, double z)
{
	if (n <= 5) {           /* lagrangian_interpolation, :41-65 */
		double lx[NMAX], ly[NMAX], lz[NMAX], tmp[NMAX * NMAX];
		for (int i = 0; i < n; ++i) {
			lx[i] = lagrange(x, xi, i, n);
			ly[i] = lagrange(y, xi, i, n);
			lz[i] = lagrange(z, xi, i, n);
		}
		/* dgemv: tmp[r] = sum_k f[k][r] lz[k]; then sum_j tmp[j][i] ly[j];
		 * then the dot with lx */
		for (int r = 0; r < n * n; ++r) {
			double s = 0.0;
			for (int k = 0; k < n; ++k)
				s += f[k * n * n + r] * lz[k];
			tmp[r] = s;
		}
		double out = 0.0;
		for (int i = 0; i < n; ++i) {
			double s = 0.0;
			for (int j = 0; j < n; ++j)
				s += tmp[j * n + i] * ly[j];
			out += lx[i] * s;
		}
		return out;
	}
	const int m = 5;        /* :67-96 */
	const int i1 = first_point(xi, x, n, m), j1 = first_point(xi, y, n, m),
		k1 = first_point(xi, z, n, m);
	double lx[5], ly[5], lz[5];
	for (int i = 0; i < m; ++i) {
		lx[i] = lagrange(x, xi + i1, i, m);
		ly[i] = lagrange(y, xi + j1, i, m);
		lz[i] = lagrange(z, xi + k1, i, m);
	}
	double tmp = 0.0;
	for (int k = 0; k < m; ++k)
		for (int j = 0; j < m; ++j)
			for (int i = 0; i < m; ++i)
				tmp += f[((k1 + k) * n + j1 + j) * n + i1 + i]
					* lx[i] * ly[j] * lz[k];
	return tmp;
}

void orc_particles_set_velocity(orc_ctx *c, const double *u, const double *v,
		const double *w)
{
	const size_t nt = c->lev[0].ntot;
	const double *uu[3] = { u, v, w };
	/* time_shift: old -> older, now -> old (time.c:37-41) */
	for (int a = 0; a < 3; ++a) {
		memcpy(c->ur[2][a], c->ur[1][a], nt * sizeof(double));
		memcpy(c->ur[1][a], c->ur[0][a], nt * sizeof(double));
	}
	/* ur_i = sum_j (J^-1)_ij u_j, first term replaces (tripalg.c:224-234) */
	for (size_t q = 0; q < nt; ++q)
		for (int i = 0; i < 3; ++i) {
			double s = c->jinv[i * 3 + 0][q] * uu[0][q];
			s = c->jinv[i * 3 + 1][q] * uu[1][q] + s;
			s = c->jinv[i * 3 + 2][q] * uu[2][q] + s;
			c->ur[0][i][q] = s;
		}
}

enum { P_NORMAL, P_COLLISION, P_OUTSIDE };

static double field_at(orc_ctx *c, const double *f, int el, double x,
		double y, double z)
{
	const level_t *L = &c->lev[0];
	return orc_interp_point(L->n, L->b.xi, f + (size_t) el * L->n3, x, y, z);
}

/* particles.c:493-515 */
static double fluid_velocity(orc_ctx *c, int el, const double *y, double t,
		int comp)
{
	const double eps = 1.0e-6, T = c->pt, dt = c->pdt;
	double u[3] = { 0.0, 0.0, 0.0 };
	const double d[3] = { t - T + 2.0 * dt, t - T + dt, t - T };
	const double a = 0.5 * d[1] * d[2] / pow(dt, 2);
	const double b = -d[0] * d[2] / pow(dt, 2);
	const double cc = 0.5 * d[0] * d[1] / pow(dt, 2);
	if (fabs(a) > eps)
		u[0] = field_at(c, c->ur[2][comp], el, y[0], y[1], y[2]);
	if (fabs(b) > eps)
		u[1] = field_at(c, c->ur[1][comp], el, y[0], y[1], y[2]);
	if (fabs(cc) > eps)
		u[2] = field_at(c, c->ur[0][comp], el, y[0], y[1], y[2]);
	return a * u[0] + b * u[1] + cc * u[2];
}

static void prhs(orc_ctx *c, const orc_particle *p, const double *y,
		double t, double *f)
{
	const double a = 1.0 / p->tau;          /* particles.c:228-236 */
	for (int i = 0; i < 3; ++i) {
		f[i] = y[i + 3];
		f[i + 3] = a * (fluid_velocity(c, p->el, y, t, i) - y[i + 3]);
	}
}

static int crossed(const double *y)
{
	for (int i = 0; i < 3; ++i)
		if (fabs(y[i]) > 1.0)
			return 1;
	return 0;
}

static void reinit(orc_ctx *c, orc_particle *p)
{
	/* the reference re-seeds with rand() (particles.c:170-180); the
	 * restatement places it at the element centre and flags it */
	double y[6] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
	p->x = p->y = p->z = 0.0;
	p->u = fluid_velocity(c, p->el, y, c->pt, 0);
	p->v = fluid_velocity(c, p->el, y, c->pt, 1);
	p->w = fluid_velocity(c, p->el, y, c->pt, 2);
	p->reinit = 1;
}

/* particles.c:238-269; force: ignore intermediate crossings */
static void rk4(orc_ctx *c, orc_particle *p, double dt, int force)
{
	static const double a[4] = { 0.0, 0.5, 0.5, 1.0 };
	double y[6], f[4][6];
	const double y0[6] = { p->x, p->y, p->z, p->u, p->v, p->w };
	for (int j = 0; j < 4; ++j) {
		const double t = p->t + a[j] * dt;
		for (int i = 0; i < 6; ++i)
			y[i] = j == 0 ? y0[i] : y0[i] + a[j] * dt * f[j - 1][i];
		if (j > 0 && crossed(y) && !force) {
			p->status = P_OUTSIDE;
			return;
		}
		prhs(c, p, y, t, f[j]);
	}
	for (int i = 0; i < 6; ++i)
		y[i] = y0[i] + dt * (f[0][i] / 6.0 + f[1][i] / 3.0 + f[2][i] / 3.0
				+ f[3][i] / 6.0);
	if (crossed(y)) {
		p->status = P_OUTSIDE;
		return;
	}
	p->x = y[0]; p->y = y[1]; p->z = y[2];
	p->u = y[3]; p->v = y[4]; p->w = y[5];
	p->t = c->pt;
}

static double dmin2(double a, double b)
{
	return a < b ? a : b;                   /* utils.c:220-226 */
}

static double step_to_boundary(const double *y, const double *f)
{
	double d[3];                            /* particles.c:310-316 */
	for (int i = 0; i < 3; ++i)
		d[i] = (copysign(1.0, f[i]) - y[i]) / f[i];
	return dmin2(dmin2(d[0], d[1]), d[2]);
}

/* particles.c:271-308 */
static void to_boundary(orc_ctx *c, orc_particle *p, double dt)
{
	static const double a[4] = { 0.0, 0.5, 0.5, 1.0 };
	double y[6], f[5][6];
	const double y0[6] = { p->x, p->y, p->z, p->u, p->v, p->w };
	double *yn[6] = { &p->x, &p->y, &p->z, &p->u, &p->v, &p->w };
	double dtau = 0.0;
	for (int j = 0; j < 4; ++j) {
		const double t = p->t + a[j] * dtau;
		for (int i = 0; i < 6; ++i)
			y[i] = j == 0 ? y0[i] : y0[i] + a[j] * dtau * f[j - 1][i];
		prhs(c, p, y, t, f[j]);
		dtau = step_to_boundary(y0, f[j]);
	}
	for (int i = 0; i < 6; ++i)
		f[4][i] = (f[0][i] / 6.0 + f[1][i] / 3.0 + f[2][i] / 3.0
				+ f[3][i] / 6.0);
	dtau = step_to_boundary(y0, f[4]);
	if (dtau < dt) {
		p->t += dtau;
		for (int i = 0; i < 6; ++i)
			*yn[i] += dtau * f[4][i];
	} else {
		p->status = P_NORMAL;
		rk4(c, p, dt, 1);
		if (p->status != P_NORMAL) {
			p->status = P_NORMAL;
			reinit(c, p);
			p->t = c->pt;
		}
	}
}

/* particles.c:380-393 */
static int closest(const orc_particle *p)
{
	const double x[3] = { p->x, p->y, p->z };
	double hmin = 2.0;
	int k = 0;
	for (int i = 0; i < 6; ++i) {
		const double h = fabs(x[i / 2] - ((i % 2) ? 1.0 : -1.0));
		if (h < hmin) {
			hmin = h;
			k = i;
		}
	}
	return k;
}

/* physical_particle_velocity / parametric_particle_velocity, :517-539;
 * the callers overwrite u, v, w one component at a time, so later
 * components see the already converted earlier ones (kept as is) */
static double phys_vel(orc_ctx *c, const orc_particle *p, int comp)
{
	double dx[3];
	for (int i = 0; i < 3; ++i)
		dx[i] = field_at(c, c->jac[comp * 3 + i], p->el, p->x, p->y, p->z);
	return p->u * dx[0] + p->v * dx[1] + p->w * dx[2];
}

static double param_vel(orc_ctx *c, const orc_particle *p, int comp)
{
	double dx[3];                   /* tinv_jac[i][comp] = (J^-1)[comp][i] */
	for (int i = 0; i < 3; ++i)
		dx[i] = field_at(c, c->jinv[comp * 3 + i], p->el, p->x, p->y, p->z);
	return p->u * dx[0] + p->v * dx[1] + p->w * dx[2];
}

/* particles.c:333-349 */
static void new_element(orc_ctx *c, orc_particle *p)
{
	double *u[3] = { &p->u, &p->v, &p->w };
	const int j = closest(p);
	const int nb = c->neigh_copy[p->el * 6 + j];
	if (nb >= 0) {
		for (int i = 0; i < 3; ++i)
			*u[i] = phys_vel(c, p, i);
		p->el = nb;
	} else {
		p->status = P_COLLISION;
		*u[j / 2] *= -1.0;
	}
}

/* particles.c:351-378 (neigh_rot is always 0) */
static void new_location(orc_ctx *c, orc_particle *p)
{
	const double y0[6] = { p->x, p->y, p->z, p->u, p->v, p->w };
	double *yn[6] = { &p->x, &p->y, &p->z, &p->u, &p->v, &p->w };
	double x2[2];
	int k = closest(p);
	int j = 0;
	for (int i = 0; i < 3; ++i)
		if (i != k / 2)
			x2[j++] = y0[i];
	for (int i = 0; i < 6; ++i)
		if (c->neigh_copy[p->el * 6 + i] == p->old_el)
			k = i;
	*yn[k / 2] = (k % 2) ? 1.0 : -1.0;
	j = 0;
	for (int i = 0; i < 3; ++i) {
		if (i != k / 2)
			*yn[i] = x2[j++];
		*yn[i + 3] = param_vel(c, p, i);
	}
	p->old_el = p->el;
}

int orc_particles_step(orc_ctx *c, orc_particle *ps, long n, double t_now,
		double dt)
{
	c->pt = t_now;
	c->pdt = dt;
	int maxpass = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max:maxpass)
	for (long q = 0; q < n; ++q) {
		orc_particle *p = ps + q;
		p->status = P_OUTSIDE;
		int j = 0;
		do {                            /* particles.c:187-217 */
			if (p->status != P_NORMAL) {
				p->status = P_NORMAL;
				if (j > 10) {
					reinit(c, p);
					p->t = c->pt;
				} else {
					rk4(c, p, c->pt - p->t, 0);
					if (p->status == P_OUTSIDE)
						to_boundary(c, p, 2.0 * dt);
					if (p->status == P_OUTSIDE)
						new_element(c, p);
				}
			}
			if (p->status == P_OUTSIDE)
				new_location(c, p);
			++j;
		} while (p->status != P_NORMAL);
		if (j > maxpass)
			maxpass = j;
	}
	return maxpass;
}

/* RK2 tracers (header).  Every product and sum is written out in the
 * order the device kernel evaluates it. */
static double tstep_to_boundary(const double *y, const double *f)
{
	double d[3];                    /* an axis without motion: never */
	for (int i = 0; i < 3; ++i)
		d[i] = f[i] != 0.0 ? (copysign(1.0, f[i]) - y[i]) / f[i] : INFINITY;
	return dmin2(dmin2(d[0], d[1]), d[2]);
}

static void tracer_velocity(orc_ctx *c, int el, const double *y, double t,
		double *f)
{
	for (int i = 0; i < 3; ++i)
		f[i] = fluid_velocity(c, el, y, t, i);
}

int orc_tracers_step(orc_ctx *c, orc_particle *ps, long n, double t_now,
		double dt)
{
	c->pt = t_now;
	c->pdt = dt;
	int maxpass = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max:maxpass)
	for (long q = 0; q < n; ++q) {
		orc_particle *p = ps + q;
		double x[3] = { p->x, p->y, p->z };
		int j = 0;
		for (;;) {
			if (j > 10) {                   /* as particles.c:195-198 */
				x[0] = x[1] = x[2] = 0.0;
				p->reinit = 1;
				p->t = t_now;
				break;
			}
			++j;
			const double h = t_now - p->t;
			if (!(h > 0.0))
				break;
			double k1[3], k2[3], xp[3], xn[3], s[3];
			tracer_velocity(c, p->el, x, p->t, k1);
			for (int i = 0; i < 3; ++i)
				xp[i] = x[i] + h * k1[i];
			int out = crossed(xp);
			if (!out) {
				tracer_velocity(c, p->el, xp, p->t + h, k2);
				for (int i = 0; i < 3; ++i) {
					s[i] = 0.5 * (k1[i] + k2[i]);
					xn[i] = x[i] + h * s[i];
				}
				if (!crossed(xn)) {
					x[0] = xn[0]; x[1] = xn[1]; x[2] = xn[2];
					p->t = t_now;
					break;
				}
			} else {
				s[0] = k1[0]; s[1] = k1[1]; s[2] = k1[2];
			}
			double dtau = tstep_to_boundary(x, s);
			if (!(dtau < h))
				dtau = h;
			if (dtau < 0.0)
				dtau = 0.0;
			for (int i = 0; i < 3; ++i)
				x[i] = x[i] + dtau * s[i];
			p->t = p->t + dtau;
			orc_particle tmp = *p;
			tmp.x = x[0]; tmp.y = x[1]; tmp.z = x[2];
			const int k = closest(&tmp);
			const int nb = c->neigh_copy[p->el * 6 + k];
			x[k / 2] = (k % 2) ? 1.0 : -1.0;
			if (nb >= 0) {
				/* the entry face: the face of nb that leads back to
				 * the element left, the opposite one first; the other
				 * two coordinates keep their order */
				int kin = k ^ 1;
				if (c->neigh_copy[nb * 6 + kin] != p->el)
					for (int i = 0; i < 6; ++i)
						if (c->neigh_copy[nb * 6 + i] == p->el)
							kin = i;
				const int d0 = k / 2, d = kin / 2;
				const double xa = d0 == 0 ? x[1] : x[0];
				const double xb = d0 == 2 ? x[1] : x[2];
				for (int i = 0, j = 0; i < 3; ++i)
					if (i != d)
						x[i] = j++ == 0 ? xa : xb;
				x[d] = (kin % 2) ? 1.0 : -1.0;
				p->old_el = p->el;
				p->el = nb;
			} else {                        /* wall: stay on it */
				p->t = t_now;
				break;
			}
		}
		p->x = x[0]; p->y = x[1]; p->z = x[2];
		double uf[3];
		tracer_velocity(c, p->el, x, t_now, uf);
		p->u = uf[0]; p->v = uf[1]; p->w = uf[2];
		if (j > maxpass)
			maxpass = j;
	}
	return maxpass;
}

void orc_particle_position(orc_ctx *c, const orc_particle *p, double *xyz)
{
	const level_t *L = &c->lev[0];
	const double *xs[3] = { L->x, L->y, L->z };
	for (int a = 0; a < 3; ++a)
		xyz[a] = field_at(c, xs[a], p->el, p->x, p->y, p->z);
}
