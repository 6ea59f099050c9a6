/*
 * ssx_setup.c -- host-side set-up of the ssx engine (C).
 *
 * GLL basis, spectral-vanishing-viscosity derivative matrix, order-transfer
 * matrices, and the shared-node groups that drive the atomic-free direct
 * stiffness summation on the device.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ssx_internal.h"

static __thread char g_err[512];

void ssx_set_error(const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof(g_err), fmt, ap);
	va_end(ap);
}

const char *ssx_last_error(void)
{
	return g_err;
}

/* P_N(x) and P_{N-1}(x) by the Bonnet recurrence */
static void legendre2(int N, double x, double *pn, double *pnm1)
{
	double p0 = 1.0, p1 = x;
	if (N == 0) {
		*pn = 1.0;
		*pnm1 = 0.0;
		return;
	}
	for (int k = 1; k < N; ++k) {
		double p2 = ((2.0 * k + 1.0) * x * p1 - (double) k * p0)
			/ (double) (k + 1);
		p0 = p1;
		p1 = p2;
	}
	*pn = p1;
	*pnm1 = p0;
}

static double legendre(int N, double x)
{
	double a, b;
	legendre2(N, x, &a, &b);
	return a;
}

/*
 * GLL points of order N = n-1: the end points and the roots of P_N'.
 * Newton on (1-x^2) P_N' written through the identity
 * (1-x^2) P_N' = N (P_{N-1} - x P_N), whose derivative is -N(N+1) P_N,
 * started from the Chebyshev-Gauss-Lobatto points.  This is the same point
 * set as basis.c:123-148 (there: dgeev-seeded Newton to 1e-12).
 */
static void gll_points(int n, double *xi)
{
	const int N = n - 1;
	xi[0] = -1.0;
	xi[N] = 1.0;
	for (int i = 1; i < N; ++i) {
		double x = -cos(M_PI * (double) i / (double) N);
		for (int it = 0; it < 100; ++it) {
			double pn, pm;
			legendre2(N, x, &pn, &pm);
			double f = (double) N * (pm - x * pn);
			double df = -(double) N * (double) (N + 1) * pn;
			double dx = f / df;
			x -= dx;
			if (fabs(dx) <= 1e-16)
				break;
		}
		xi[i] = x;
	}
	if (n % 2 == 1)
		xi[N / 2] = 0.0;
}

void ssx_basis_init(ssx_basis *b, int n, int svvn, double svva)
{
	const int N = n - 1;
	memset(b, 0, sizeof(*b));
	b->n = n;
	gll_points(n, b->xi);
	double pN[SSX_NMAX];
	for (int i = 0; i < n; ++i) {
		pN[i] = legendre(N, b->xi[i]);
		b->rho[i] = 2.0 / ((double) N * (double) (N + 1) * pN[i] * pN[i]);
	}
	/* D_il = l_l'(xi_i) */
	for (int i = 0; i < n; ++i)
		for (int l = 0; l < n; ++l)
			b->d[i * n + l] = (i == l) ? 0.0
				: pN[i] / (pN[l] * (b->xi[i] - b->xi[l]));
	b->d[0] = -(double) N * (double) (N + 1) / 4.0;
	b->d[n * n - 1] = (double) N * (double) (N + 1) / 4.0;
	/* SVV matrix: Dsvv = V F L D with L_rl = (r+1/2) rho_l P_r(xi_l) the
	 * discrete Legendre transform, F = diag(sqrt(1 + c s_r)),
	 * s_r = exp(-((N-r)/(m-r))^2) for r > m else 0, V_ir = P_r(xi_i)
	 * (deriv.c:81-112, basis.c:190-209).  With c = 0, Dsvv = D up to
	 * rounding. */
	double L[SSX_NMAX * SSX_NMAX], V[SSX_NMAX * SSX_NMAX];
	double T[SSX_NMAX * SSX_NMAX];
	for (int r = 0; r < n; ++r)
		for (int l = 0; l < n; ++l) {
			L[r * n + l] = ((double) r + 0.5) * b->rho[l]
				* legendre(r, b->xi[l]);
			V[l * n + r] = legendre(r, b->xi[l]);
		}
	for (int r = 0; r < n; ++r) {
		double s = 0.0;
		if (r > svvn) {
			double q = (double) (N - r) / (double) (svvn - r);
			s = exp(-q * q);
		}
		double f = sqrt(1.0 + svva * s);
		for (int l = 0; l < n; ++l) {
			double acc = 0.0;
			for (int p = 0; p < n; ++p)
				acc += L[r * n + p] * b->d[p * n + l];
			T[r * n + l] = acc * f;
		}
	}
	for (int i = 0; i < n; ++i)
		for (int l = 0; l < n; ++l) {
			double acc = 0.0;
			for (int r = 0; r < n; ++r)
				acc += V[i * n + r] * T[r * n + l];
			b->dsvv[i * n + l] = acc;
		}
}

/* out[p*nto+i] = l_p(xto_i) for the Lagrange basis on xfrom */
void ssx_interp(const double *xfrom, int nfrom, const double *xto, int nto,
		double *out)
{
	for (int p = 0; p < nfrom; ++p)
		for (int i = 0; i < nto; ++i) {
			double v = 1.0;
			for (int q = 0; q < nfrom; ++q)
				if (q != p)
					v *= (xto[i] - xfrom[q]) / (xfrom[p] - xfrom[q]);
			out[p * nto + i] = v;
		}
}

/* ------------------------------------------------------------------ */
/* Shared-node groups                                                  */
/*                                                                     */
/* The reference assembles with per-element pointer wiring           */
/* (init.c:68-130): element e's face/edge/corner slot s is tied to the */
/* FIRST slot of the neighbour that points back at e, pre-sum copies   */
/* of all slots are taken, and each element then adds its partners'    */
/* copies in slot order faces W,E,S,N,B,T, edges 0..11, corners 0..7   */
/* (global.c:241-257).  Every copy of a shared node therefore becomes  */
/* own + partners in that order.  Here the same lists are built once,  */
/* grouped by node, and checked to follow the slot-bit pattern the     */
/* device kernel evaluates: with the m copies of a node numbered by    */
/* t = 0..m-1, copy t sums t, t^1, t^2, t^4, t^3, t^5, t^6, t^7 (m=8), */
/* t, t^1, t^2, t^3 (m=4) or t, t^1 (m=2).                             */
/* ------------------------------------------------------------------ */

/* slot s (0..25), position t -> local node (k*n+j)*n+i */
static int slot_node(int s, int t, int n)
{
	const int m = n - 1;
	int a = t / n, b = t % n, k, j, i;
	switch (s) {
	case 0: k = a; j = b; i = 0; break;
	case 1: k = a; j = b; i = m; break;
	case 2: k = a; j = 0; i = b; break;
	case 3: k = a; j = m; i = b; break;
	case 4: k = 0; j = a; i = b; break;
	case 5: k = m; j = a; i = b; break;
	case 6: k = t; j = 0; i = 0; break;
	case 7: k = t; j = 0; i = m; break;
	case 8: k = t; j = m; i = 0; break;
	case 9: k = t; j = m; i = m; break;
	case 10: k = 0; j = t; i = 0; break;
	case 11: k = 0; j = t; i = m; break;
	case 12: k = m; j = t; i = 0; break;
	case 13: k = m; j = t; i = m; break;
	case 14: k = 0; j = 0; i = t; break;
	case 15: k = 0; j = m; i = t; break;
	case 16: k = m; j = 0; i = t; break;
	case 17: k = m; j = m; i = t; break;
	default: {
		int c = s - 18;
		i = (c % 2) * m;
		j = ((c / 2) % 2) * m;
		k = (c / 4) * m;
	}
	}
	return (k * n + j) * n + i;
}

/* position of local node (k,j,i) in slot s, or -1 */
static int slot_pos(int s, int k, int j, int i, int n)
{
	const int m = n - 1;
	switch (s) {
	case 0: return i == 0 ? k * n + j : -1;
	case 1: return i == m ? k * n + j : -1;
	case 2: return j == 0 ? k * n + i : -1;
	case 3: return j == m ? k * n + i : -1;
	case 4: return k == 0 ? j * n + i : -1;
	case 5: return k == m ? j * n + i : -1;
	case 6: return (j == 0 && i == 0) ? k : -1;
	case 7: return (j == 0 && i == m) ? k : -1;
	case 8: return (j == m && i == 0) ? k : -1;
	case 9: return (j == m && i == m) ? k : -1;
	case 10: return (k == 0 && i == 0) ? j : -1;
	case 11: return (k == 0 && i == m) ? j : -1;
	case 12: return (k == m && i == 0) ? j : -1;
	case 13: return (k == m && i == m) ? j : -1;
	case 14: return (k == 0 && j == 0) ? i : -1;
	case 15: return (k == 0 && j == m) ? i : -1;
	case 16: return (k == m && j == 0) ? i : -1;
	case 17: return (k == m && j == m) ? i : -1;
	default: {
		int c = s - 18;
		return (i == (c % 2) * m && j == ((c / 2) % 2) * m
				&& k == (c / 4) * m) ? 0 : -1;
	}
	}
}

/* partner table part[e*26+s] = el*26+slot or -1 (init.c:68-130) */
static int *wire(const ssx_mesh *m)
{
	int *part = malloc(sizeof(int) * (size_t) m->nel * 26);
	if (!part)
		return NULL;
	for (int e = 0; e < m->nel; ++e)
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
			const int el = nb[e * ns + loc];
			part[e * 26 + s] = -1;
			if (el < 0 || sz == 0)
				continue;
			if (sz != 1 || el >= m->nel) {
				ssx_set_error("element %d slot %d: size %d / neighbour %d"
						" not supported", e, s, sz, el);
				free(part);
				return NULL;
			}
			int si = -1;
			for (int q = 0; q < ns; ++q)
				if (nb[el * ns + q] == e) {
					si = q;
					break;
				}
			if (si < 0) {
				ssx_set_error("element %d slot %d: neighbour %d does not"
						" point back", e, s, el);
				free(part);
				return NULL;
			}
			part[e * 26 + s] = el * 26 + base + si;
		}
	return part;
}

/* ordered contribution list of copy (e,q): own, then partners by slot */
static int copy_list(const int *part, int n, int e, int q, long long *out)
{
	const int n3 = n * n * n;
	const int i = q % n, j = (q / n) % n, k = q / (n * n);
	int c = 0;
	out[c++] = (long long) e * n3 + q;
	for (int s = 0; s < 26; ++s) {
		int pp = part[e * 26 + s];
		if (pp < 0)
			continue;
		int t = slot_pos(s, k, j, i, n);
		if (t < 0)
			continue;
		if (c >= 8)
			return -1;
		out[c++] = (long long) (pp / 26) * n3 + slot_node(pp % 26, t, n);
	}
	return c;
}

static long long uf_find(long long *par, long long a)
{
	while (par[a] != a) {
		par[a] = par[par[a]];
		a = par[a];
	}
	return a;
}

static const int xor_order8[8] = { 0, 1, 2, 4, 3, 5, 6, 7 };

int ssx_build_groups(const ssx_mesh *m, int n, int *grp[3], long long ng[3])
{
	const int n3 = n * n * n;
	const long long ntot = (long long) m->nel * n3;
	grp[0] = grp[1] = grp[2] = NULL;
	ng[0] = ng[1] = ng[2] = 0;
	if (ntot >= (1LL << 31)) {
		ssx_set_error("mesh too large for 32-bit node indices");
		return SSX_EINVAL;
	}
	int *part = wire(m);
	if (!part)
		return SSX_EMESH;
	long long *par = malloc(sizeof(long long) * (size_t) ntot);
	int *cnt = calloc((size_t) ntot, sizeof(int));
	if (!par || !cnt) {
		free(part); free(par); free(cnt);
		ssx_set_error("out of host memory building groups");
		return SSX_ENOMEM;
	}
	for (long long a = 0; a < ntot; ++a)
		par[a] = a;
	long long lst[8];
	int rc = SSX_OK;
	for (int e = 0; e < m->nel && rc == SSX_OK; ++e)
		for (int q = 0; q < n3; ++q) {
			int len = copy_list(part, n, e, q, lst);
			if (len < 0) {
				rc = SSX_EMESH;
				ssx_set_error("more than 8 copies of a node");
				break;
			}
			for (int t = 1; t < len; ++t) {
				long long ra = uf_find(par, lst[0]);
				long long rb = uf_find(par, lst[t]);
				if (ra != rb) {
					if (ra < rb)
						par[rb] = ra;
					else
						par[ra] = rb;
				}
			}
		}
	if (rc != SSX_OK)
		goto out;
	/* the root of each group is its smallest copy index */
	for (long long a = 0; a < ntot; ++a)
		cnt[uf_find(par, a)]++;
	long long tot[9] = { 0 };
	for (long long a = 0; a < ntot; ++a)
		if (par[a] == a) {
			int c = cnt[a];
			if (c != 1 && c != 2 && c != 4 && c != 8) {
				ssx_set_error("node shared by %d copies (not 1,2,4,8)", c);
				rc = SSX_EMESH;
				goto out;
			}
			tot[c]++;
		}
	const int mults[3] = { 2, 4, 8 };
	for (int mi = 0; mi < 3; ++mi) {
		ng[mi] = tot[mults[mi]];
		grp[mi] = malloc(sizeof(int) * (size_t) (ng[mi] ? ng[mi] : 1)
				* mults[mi]);
		if (!grp[mi]) {
			rc = SSX_ENOMEM;
			goto out;
		}
	}
	long long fill[3] = { 0, 0, 0 };
	for (long long a = 0; a < ntot && rc == SSX_OK; ++a) {
		if (par[a] != a || cnt[a] == 1)
			continue;
		const int c = cnt[a];
		const int mi = c == 2 ? 0 : (c == 4 ? 1 : 2);
		const long long G = ng[mi], g = fill[mi]++;
		/* slots from the list of the root copy: the k-th partner in its
		 * list sits at slot xor_order[k] (XOR with 0 is the slot) */
		int len = copy_list(part, n, (int) (a / n3), (int) (a % n3), lst);
		if (len != c) {
			rc = SSX_EMESH;
			ssx_set_error("copy %lld lists %d of %d copies", a, len, c);
			break;
		}
		long long slot_copy[8];
		for (int t = 0; t < c; ++t)
			slot_copy[c == 8 ? xor_order8[t] : t] = lst[t];
		/* every copy must list the group in the slot-bit order */
		for (int t = 0; t < c && rc == SSX_OK; ++t) {
			long long l2[8];
			int len2 = copy_list(part, n, (int) (slot_copy[t] / n3),
					(int) (slot_copy[t] % n3), l2);
			if (len2 != c) {
				rc = SSX_EMESH;
				break;
			}
			for (int k = 0; k < c; ++k) {
				int want = t ^ (c == 8 ? xor_order8[k] : k);
				if (l2[k] != slot_copy[want]) {
					rc = SSX_EMESH;
					ssx_set_error("copy %lld: add order differs from the"
							" slot-bit pattern", slot_copy[t]);
					break;
				}
			}
		}
		for (int t = 0; t < c; ++t)
			grp[mi][(long long) t * G + g] = (int) slot_copy[t];
	}
out:
	free(part);
	free(par);
	free(cnt);
	if (rc != SSX_OK)
		for (int mi = 0; mi < 3; ++mi) {
			free(grp[mi]);
			grp[mi] = NULL;
			ng[mi] = 0;
		}
	return rc;
}

/* host DSS with the same summation order as the device kernel */
void ssx_host_dss(int *grp[3], const long long ng[3], double *u)
{
	const int mults[3] = { 2, 4, 8 };
	for (int mi = 0; mi < 3; ++mi) {
		const int c = mults[mi];
		const long long G = ng[mi];
		for (long long g = 0; g < G; ++g) {
			double v[8], s[8];
			for (int t = 0; t < c; ++t)
				v[t] = u[grp[mi][(long long) t * G + g]];
			for (int t = 0; t < c; ++t) {
				double a = v[t];
				for (int k = 1; k < c; ++k)
					a = a + v[t ^ (c == 8 ? xor_order8[k] : k)];
				s[t] = a;
			}
			for (int t = 0; t < c; ++t)
				u[grp[mi][(long long) t * G + g]] = s[t];
		}
	}
}

/* Per-copy partner lists for the DSS fused into the CG update.
 *
 * Each shared copy (e,q) keeps its partners in its own add order, i.e. the
 * list copy_list builds (own first, then partners by slot; global.c:
 * 241-257), so own + partners left to right is the DSS value of that copy.
 * A node on one face has at most 1 partner, on an edge at most 3, at a
 * corner at most 7: element e's block of *P entries holds, for every
 * surface node in local order, that many entries, unused ones -1.
 * soff[q] = (offset of node q in the block) << 3 | capacity. */
int ssx_build_gather(const ssx_mesh *m, int n, int **tab, int **soff,
		int *P)
{
	const int n3 = n * n * n;
	*tab = NULL;
	*soff = malloc(sizeof(int) * (size_t) n3);
	if (!*soff)
		return SSX_ENOMEM;
	int off = 0;
	for (int q = 0; q < n3; ++q) {
		const int i = q % n, j = (q / n) % n, k = q / (n * n);
		const int on = (i == 0 || i == n - 1) + (j == 0 || j == n - 1)
			+ (k == 0 || k == n - 1);
		const int cap = on == 0 ? 0 : on == 1 ? 1 : on == 2 ? 3 : 7;
		(*soff)[q] = off << 3 | cap;
		off += cap;
	}
	*P = off;
	int *part = wire(m);
	if (!part) {
		free(*soff);
		*soff = NULL;
		return SSX_EMESH;
	}
	int *t = malloc(sizeof(int) * (size_t) m->nel * (off ? off : 1));
	if (!t) {
		free(part);
		free(*soff);
		*soff = NULL;
		return SSX_ENOMEM;
	}
	long long lst[8];
	int rc = SSX_OK;
	for (int e = 0; e < m->nel && rc == SSX_OK; ++e)
		for (int q = 0; q < n3; ++q) {
			const int cap = (*soff)[q] & 7;
			if (!cap)
				continue;
			int *dst = t + (size_t) e * off + ((*soff)[q] >> 3);
			const int len = copy_list(part, n, e, q, lst);
			if (len < 1 || len - 1 > cap) {
				ssx_set_error("copy (%d,%d): %d partners for %d slots", e,
						q, len - 1, cap);
				rc = SSX_EMESH;
				break;
			}
			for (int k = 0; k < cap; ++k)
				dst[k] = k + 1 < len ? (int) lst[k + 1] : -1;
		}
	free(part);
	if (rc != SSX_OK) {
		free(t);
		free(*soff);
		*soff = NULL;
		return rc;
	}
	*tab = t;
	return SSX_OK;
}
