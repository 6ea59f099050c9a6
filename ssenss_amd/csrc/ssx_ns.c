/*
 * ssx_ns.c -- the incompressible Navier-Stokes time step of SSENSS around
 * the elliptic solves, on the device: integrate_navier_stokes
 * (navier_stokes.c:224-235) with
 *   - OIFS convection: three RK4 sub-integrations with dt/2 steps of the
 *     pure-advection problem started from the interpolated velocity
 *     history (lagrangian_convection_integration navier_stokes.c:71-92,
 *     integrate_convection_rk4 time.c:84-130, add_interpolated_velocity
 *     time.c:132-152);
 *   - the dealiased skew-symmetric convection operator at nconv = 3n/2
 *     points per direction (convection.c:31-140, main.c:53);
 *   - the pressure right-hand side (navier_stokes.c:115-146,
 *     estimate_pressure_gradient stokes.c:107-130), the EXT3 pressure
 *     estimate (:101-113) and the FCG + p-MG pressure solve (:148-159);
 *   - the velocity right-hand side (:177-201), the EXT3 velocity estimate
 *     (:161-175) and the three Helmholtz Jacobi-PCG solves (:203-222), as
 *     one batch of three problems;
 *   - BDF3/EXT3 weights (time.c:51-73).
 * Every field stays in device memory; the operations keep the reference's
 * order (each per-element loop of the reference becomes one pass over all
 * elements).  Dirichlet boundary values are zero, the reference's default
 * boundary functions (structs.c:67-68); the body force is constant
 * (main.c:104-106).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "ssx_internal.h"

#define HIPN(x)                                                              \
	do {                                                                 \
		hipError_t e_ = (x);                                         \
		if (e_ != hipSuccess) {                                      \
			ssx_set_error("%s: %s", #x, hipGetErrorString(e_));   \
			return SSX_EHIP;                                     \
		}                                                            \
	} while (0)
#define TRYN(x)                                                              \
	do {                                                                 \
		int rc_ = (x);                                               \
		if (rc_ != SSX_OK)                                           \
			return rc_;                                          \
	} while (0)

#define NS_TOL 1.0e-10          /* residual_tolerance, navier_stokes.c:33-99 */

/* kernels (ssx_nskernels.hip) */
int ssxk_ns_jac(void *st, size_t nt, int n, const double *dm,
		const double *const a[3], double *const out[9]);
int ssxk_ns_curl(void *st, size_t nt, const double *const J[9],
		const double *const ti[9], const double *mass, double *const out[3]);
int ssxk_ns_div(void *st, size_t nt, const double *const J[9],
		const double *const ti[9], const double *mass, double scale,
		double *out);
int ssxk_ns_grad(void *st, size_t nt, int n, const double *dm,
		const double *p, const double *mass, const double *const ti[9],
		double *const out[3]);
int ssxk_ns_conv(void *st, size_t nt, int n, const double *dm,
		const double *const u[3], const double *const v[3],
		const double *const Jv[9], const double *const ti[9],
		const double *mass, double *const out[3]);
size_t ssxk_ns_conv_full_bytes(int n, int nc);
int ssxk_ns_conv_full(void *st, int nel, int n, int nc, const double *prol,
		const double *dm, const double *const u[3], const double *const v[3],
		const double *const ti[9], const double *mass, double *const out[3]);
int ssxk_ns_surf(void *st, int nel, int n, const int *bc, int dirichlet,
		const double *wn, const double *g, double *f);
int ssxk_ns_zero_dir(void *st, size_t nt, int n, const int *bc,
		double *const u[3]);
int ssxk_ns_comb(void *st, size_t nt, const double *const x[3][3],
		const double a[3], const int use[3], double *const y[3]);
int ssxk_ns_mul3(void *st, size_t nt, double alpha, const double *const x[3],
		const double *const w[3], double beta, double *const z[3]);

struct ssx_ns {
	ssx_ctx *c;
	int n, nc, nel;
	size_t nt, ntc;
	double dt, visc, dens, t, g[3];
	int batched;
	ssx_batch *vb;             /* the three velocity solves */
	/* geometry: level 0 (tinv, face weights x normals per direction) and
	 * the convective level */
	double *ti[9], *wn[3];
	double *dm_c, *prol_c, *ti_c[9], *mass_c;
	/* fields: [0] newest level */
	double *u[3][3], *p[3], *conv[3][3], *rot[3][3];
	/* scratch at n */
	double *ut[3], *rhs[3], *a[3], *b[3], *J[9];
	double *rkf[4][3], *rky[4][3];
	/* scratch at nconv */
	double *cu[3], *cv[3], *cJ[9], *co[3];
	int conv_elem;               /* convection one workgroup per element */
	int inited;
	long long cnt_solve[4];
};

static int nalloc(double **p, size_t n)
{
	if (hipMalloc((void **) p, (n ? n : 1) * sizeof(double)) != hipSuccess) {
		*p = NULL;
		ssx_set_error("hipMalloc of %zu doubles failed (Navier-Stokes)", n);
		return SSX_ENOMEM;
	}
	return SSX_OK;
}

static int nup(double **d, const double *h, size_t n)
{
	TRYN(nalloc(d, n));
	HIPN(hipMemcpy(*d, h, n * sizeof(double), hipMemcpyHostToDevice));
	return SSX_OK;
}

void ssx_ns_destroy(ssx_ns *ns)
{
	if (!ns)
		return;
	if (ns->vb)
		ssx_batch_destroy(ns->vb);
	double **all[] = { ns->ti, ns->ti + 3, ns->ti + 6, ns->wn, ns->ti_c,
		ns->ti_c + 3, ns->ti_c + 6, ns->ut, ns->rhs, ns->a, ns->b, ns->J,
		ns->J + 3, ns->J + 6, ns->cu, ns->cv, ns->cJ, ns->cJ + 3, ns->cJ + 6,
		ns->co, ns->p };
	for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); ++k)
		for (int c = 0; c < 3; ++c)
			hipFree(all[k][c]);
	for (int l = 0; l < 3; ++l)
		for (int c = 0; c < 3; ++c) {
			hipFree(ns->u[l][c]);
			hipFree(ns->conv[l][c]);
			hipFree(ns->rot[l][c]);
		}
	for (int s = 0; s < 4; ++s)
		for (int c = 0; c < 3; ++c) {
			hipFree(ns->rkf[s][c]);
			hipFree(ns->rky[s][c]);
		}
	hipFree(ns->dm_c);
	hipFree(ns->prol_c);
	hipFree(ns->mass_c);
	free(ns);
}

/* geometry of the step: tinv_jac and the face normals at level 0
 * (compute_geometry_transformations geometry.c:102-113,
 * compute_surface_normal :137-186), and at the convective level the
 * grid interpolated to nconv points (as read_grid builds conv_geom,
 * dataio.c:433-441), its tinv_jac and w detJ */
static int ns_geometry(ssx_ns *ns)
{
	ssx_ctx *c = ns->c;
	const ssx_level *L = &c->lev[0];
	const int n = ns->n, nc = ns->nc, nel = ns->nel;
	const int n2 = n * n, n3 = n2 * n, nc3 = nc * nc * nc;
	const size_t nt = ns->nt, ntc = ns->ntc;
	int rc = SSX_OK;
	double *jac[9], *jinv[9], *wn[3], *tmp = NULL;
	double *xc[3] = { NULL, NULL, NULL }, *det = NULL;
	for (int k = 0; k < 9; ++k)
		jac[k] = jinv[k] = NULL;
	for (int k = 0; k < 9; ++k) {
		jac[k] = malloc(sizeof(double) * (ntc > nt ? ntc : nt));
		jinv[k] = malloc(sizeof(double) * (ntc > nt ? ntc : nt));
		if (!jac[k] || !jinv[k])
			rc = SSX_ENOMEM;
	}
	for (int d = 0; d < 3; ++d) {
		wn[d] = malloc(sizeof(double) * (size_t) nel * 6 * n2);
		xc[d] = malloc(sizeof(double) * ntc);
		if (!wn[d] || !xc[d])
			rc = SSX_ENOMEM;
	}
	det = malloc(sizeof(double) * ntc);
	tmp = malloc(sizeof(double) * ntc);
	if (!det || !tmp)
		rc = SSX_ENOMEM;
	if (rc != SSX_OK) {
		ssx_set_error("out of host memory (Navier-Stokes geometry)");
		goto out;
	}
	ssx_jacobians(&L->b, nel, L->h_x, L->h_y, L->h_z, jac, jinv, NULL);
	/* tinv_jac[a][b] = (J^-1)[b][a] */
	for (int a = 0; a < 3 && rc == SSX_OK; ++a)
		for (int b = 0; b < 3 && rc == SSX_OK; ++b)
			rc = nup(&ns->ti[3 * a + b], jinv[3 * b + a], nt);
	/* face normals (compute_surface_normal), times w_b w_a */
	const int m = n - 1;
	for (int e = 0; e < nel && rc == SSX_OK; ++e)
		for (int s = 0; s < 6; ++s)
			for (int a = 0; a < n; ++a)
				for (int b = 0; b < n; ++b) {
					int k, j, i;
					switch (s) {
					case 0: k = a; j = b; i = 0; break;
					case 1: k = a; j = b; i = m; break;
					case 2: k = a; j = 0; i = b; break;
					case 3: k = a; j = m; i = b; break;
					case 4: k = 0; j = a; i = b; break;
					default: k = m; j = a; i = b; break;
					}
					const size_t q = (size_t) e * n3 + (k * n + j) * n + i;
					double J[3][3];
					for (int r = 0; r < 3; ++r)
						for (int t = 0; t < 3; ++t)
							J[r][t] = jac[3 * r + t][q];
					double nv[3];
					if (s == 4 || s == 5) {       /* b, t: xi x eta */
						const double sg = s == 4 ? -1.0 : 1.0;
						nv[0] = sg * (J[1][0] * J[2][1] - J[2][0] * J[1][1]);
						nv[1] = -sg * (J[0][0] * J[2][1] - J[2][0] * J[0][1]);
						nv[2] = sg * (J[0][0] * J[1][1] - J[1][0] * J[0][1]);
					} else if (s == 2 || s == 3) {  /* s, n: xi x zeta */
						const double sg = s == 2 ? 1.0 : -1.0;
						nv[0] = sg * (J[1][0] * J[2][2] - J[2][0] * J[1][2]);
						nv[1] = -sg * (J[0][0] * J[2][2] - J[2][0] * J[0][2]);
						nv[2] = sg * (J[0][0] * J[1][2] - J[1][0] * J[0][2]);
					} else {                         /* w, e: eta x zeta */
						const double sg = s == 0 ? -1.0 : 1.0;
						nv[0] = sg * (J[1][1] * J[2][2] - J[2][1] * J[1][2]);
						nv[1] = -sg * (J[0][1] * J[2][2] - J[2][1] * J[0][2]);
						nv[2] = sg * (J[0][1] * J[1][2] - J[1][1] * J[0][2]);
					}
					const double ww = L->b.rho[b] * L->b.rho[a];
					for (int d = 0; d < 3; ++d)
						wn[d][((size_t) e * 6 + s) * n2 + a * n + b] =
							ww * nv[d];
				}
	for (int d = 0; d < 3 && rc == SSX_OK; ++d)
		rc = nup(&ns->wn[d], wn[d], (size_t) nel * 6 * n2);
	/* convective level */
	ssx_basis bc;
	ssx_basis_init(&bc, nc, (2 * nc) / 3, 0.0);
	double prol[SSX_NMAX * SSX_NMAX];
	ssx_interp(L->b.xi, n, bc.xi, nc, prol);      /* prol[p*nc+i] */
	const double *src[3] = { L->h_x, L->h_y, L->h_z };
	for (int d = 0; d < 3; ++d)
		for (int e = 0; e < nel; ++e) {
			/* interpolate_to_grid: x-, then y-, then z-direction */
			const double *u = src[d] + (size_t) e * n3;
			double *o = xc[d] + (size_t) e * nc3;
			double *t1 = tmp, *t2 = tmp + (size_t) nc * n2;
			for (int k = 0; k < n; ++k)
				for (int j = 0; j < n; ++j)
					for (int i = 0; i < nc; ++i) {
						double s = 0.0;
						for (int p = 0; p < n; ++p)
							s += prol[p * nc + i] * u[(k * n + j) * n + p];
						t1[(k * n + j) * nc + i] = s;
					}
			for (int k = 0; k < n; ++k)
				for (int j = 0; j < nc; ++j)
					for (int i = 0; i < nc; ++i) {
						double s = 0.0;
						for (int p = 0; p < n; ++p)
							s += t1[(k * n + p) * nc + i] * prol[p * nc + j];
						t2[(k * nc + j) * nc + i] = s;
					}
			for (int k = 0; k < nc; ++k)
				for (int ji = 0; ji < nc * nc; ++ji) {
					double s = 0.0;
					for (int p = 0; p < n; ++p)
						s += t2[p * nc * nc + ji] * prol[p * nc + k];
					o[k * nc * nc + ji] = s;
				}
		}
	ssx_jacobians(&bc, nel, xc[0], xc[1], xc[2], jac, jinv, det);
	for (int a = 0; a < 3 && rc == SSX_OK; ++a)
		for (int b = 0; b < 3 && rc == SSX_OK; ++b)
			rc = nup(&ns->ti_c[3 * a + b], jinv[3 * b + a], ntc);
	for (int e = 0; e < nel; ++e)
		for (int k = 0; k < nc; ++k)
			for (int j = 0; j < nc; ++j)
				for (int i = 0; i < nc; ++i) {
					const size_t q = (size_t) e * nc3 + (k * nc + j) * nc + i;
					det[q] = (bc.rho[i] * bc.rho[j] * bc.rho[k]) * det[q];
				}
	if (rc == SSX_OK)
		rc = nup(&ns->mass_c, det, ntc);
	if (rc == SSX_OK)
		rc = nup(&ns->dm_c, bc.d, (size_t) nc * nc);
	if (rc == SSX_OK)
		rc = nup(&ns->prol_c, prol, (size_t) n * nc);
out:
	for (int k = 0; k < 9; ++k) {
		free(jac[k]);
		free(jinv[k]);
	}
	for (int d = 0; d < 3; ++d) {
		free(wn[d]);
		free(xc[d]);
	}
	free(det);
	free(tmp);
	return rc;
}

int ssx_ns_create(ssx_ns **out, ssx_ctx *c, const ssx_ns_params *prm)
{
	if (!out || !c || !prm) {
		ssx_set_error("ssx_ns_create: null argument");
		return SSX_EINVAL;
	}
	*out = NULL;
	const int n = c->lev[0].n, nc = (3 * n) / 2;      /* main.c:53 */
	if (nc > 15) {
		ssx_set_error("ssx_ns_create: n = %d gives %d convective points "
				"per direction, more than 15", n, nc);
		return SSX_EINVAL;
	}
	if (!(prm->dt > 0.0) || !(prm->visc > 0.0) || !(prm->dens > 0.0)) {
		ssx_set_error("ssx_ns_create: dt, visc and dens must be positive");
		return SSX_EINVAL;
	}
	const double lam = (11.0 / 6.0) / (prm->dt * prm->visc);
	if (fabs(lam - c->lambda) > 1e-12 * fabs(lam)) {
		ssx_set_error("ssx_ns_create: the context's Helmholtz lambda %g is "
				"not (11/6)/(dt visc) = %g", c->lambda, lam);
		return SSX_EINVAL;
	}
	ssx_ns *ns = calloc(1, sizeof(*ns));
	if (!ns)
		return SSX_ENOMEM;
	ns->c = c;
	ns->n = n;
	ns->nc = nc;
	ns->nel = c->nel;
	ns->nt = c->lev[0].ntot;
	ns->ntc = (size_t) c->nel * nc * nc * nc;
	ns->dt = prm->dt;
	ns->visc = prm->visc;
	ns->dens = prm->dens;
	ns->t = prm->t;
	for (int d = 0; d < 3; ++d)
		ns->g[d] = prm->g[d];
	ns->batched = prm->batch_velocity;
	/* the whole convection operator one element per workgroup where its
	 * LDS fits the device, else the transfer kernels and the node-per-lane
	 * operator with the Jacobian of v stored (cJ) */
	{
		int dev = 0, lim = 0;
		const size_t need = ssxk_ns_conv_full_bytes(n, nc);
		ns->conv_elem = hipGetDevice(&dev) == hipSuccess
			&& hipDeviceGetAttribute(&lim,
					hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess
			&& need <= (size_t) lim;
	}
	int rc = ns_geometry(ns);
	const size_t nt = ns->nt, ntc = ns->ntc;
	for (int l = 0; l < 3 && rc == SSX_OK; ++l)
		for (int d = 0; d < 3 && rc == SSX_OK; ++d) {
			rc = nalloc(&ns->u[l][d], nt);
			if (rc == SSX_OK)
				rc = nalloc(&ns->conv[l][d], nt);
			if (rc == SSX_OK)
				rc = nalloc(&ns->rot[l][d], nt);
		}
	for (int d = 0; d < 3 && rc == SSX_OK; ++d) {
		double **v[] = { &ns->p[d], &ns->ut[d], &ns->rhs[d], &ns->a[d],
			&ns->b[d], &ns->cu[d], &ns->cv[d], &ns->co[d] };
		const size_t sz[] = { nt, nt, nt, nt, nt, ntc, ntc, ntc };
		/* the nc-point fields only for the unfused convection */
		for (int k = 0; k < (ns->conv_elem ? 5 : 8) && rc == SSX_OK; ++k)
			rc = nalloc(v[k], sz[k]);
		for (int s = 0; s < 4 && rc == SSX_OK; ++s) {
			rc = nalloc(&ns->rkf[s][d], nt);
			if (rc == SSX_OK)
				rc = nalloc(&ns->rky[s][d], nt);
		}
	}
	for (int k = 0; k < 9 && rc == SSX_OK; ++k) {
		rc = nalloc(&ns->J[k], nt);
		if (rc == SSX_OK && !ns->conv_elem)
			rc = nalloc(&ns->cJ[k], ntc);
	}
	/* zero histories (the element constructor's zeroed arrays) */
	for (int l = 0; l < 3 && rc == SSX_OK; ++l) {
		rc = ssxk_fill(c->stream, ns->p[l], 0.0, nt);
		for (int d = 0; d < 3 && rc == SSX_OK; ++d) {
			rc = ssxk_fill(c->stream, ns->u[l][d], 0.0, nt);
			if (rc == SSX_OK)
				rc = ssxk_fill(c->stream, ns->conv[l][d], 0.0, nt);
			if (rc == SSX_OK)
				rc = ssxk_fill(c->stream, ns->rot[l][d], 0.0, nt);
		}
	}
	if (rc == SSX_OK && ns->batched) {
		ssx_ctx *cs[3] = { c, c, c };
		const int ops[3] = { SSX_OP_HELMHOLTZ, SSX_OP_HELMHOLTZ,
			SSX_OP_HELMHOLTZ };
		const int vars[3] = { SSX_X_VELOCITY, SSX_Y_VELOCITY,
			SSX_Z_VELOCITY };
		rc = ssx_batch_create(&ns->vb, cs, ops, vars, 3);
	}
	if (rc != SSX_OK) {
		ssx_ns_destroy(ns);
		return rc;
	}
	*out = ns;
	return SSX_OK;
}

static void rotate3(double *h[3][3])
{
	for (int d = 0; d < 3; ++d) {
		double *t = h[2][d];
		h[2][d] = h[1][d];
		h[1][d] = h[0][d];
		h[0][d] = t;
	}
}

/* convection_operator(u, v) -> out at n (convection.c:31-69): u and v
 * interpolated to the convective points, the skew-symmetric operator
 * there, integrated back (J^T) */
static int ns_convection(ssx_ns *ns, double *const u[3], double *const v[3],
		double *const out[3])
{
	void *st = ns->c->stream;
	const int n = ns->n, nc = ns->nc, nel = ns->nel;
	if (ns->conv_elem)          /* the whole operator per element */
		return ssxk_ns_conv_full(st, nel, n, nc, ns->prol_c, ns->dm_c,
				(const double *const *) u, (const double *const *) v,
				(const double *const *) ns->ti_c, ns->mass_c, out);
	for (int d = 0; d < 3; ++d) {
		TRYN(ssxk_prolong(st, nc, n, nel, ns->prol_c, u[d], ns->cu[d], 0));
		TRYN(ssxk_prolong(st, nc, n, nel, ns->prol_c, v[d], ns->cv[d], 0));
	}
	TRYN(ssxk_ns_jac(st, ns->ntc, nc, ns->dm_c, (const double *const *) ns->cv,
			ns->cJ));
	TRYN(ssxk_ns_conv(st, ns->ntc, nc, ns->dm_c,
			(const double *const *) ns->cu, (const double *const *) ns->cv,
			(const double *const *) ns->cJ, (const double *const *) ns->ti_c,
			ns->mass_c, ns->co));
	for (int d = 0; d < 3; ++d)
		TRYN(ssxk_restrict(st, nc, n, nel, ns->prol_c, ns->co[d], out[d]));
	return SSX_OK;
}

/* double_curl_operator (stokes.c:94-105) -> out */
static int ns_double_curl(ssx_ns *ns, double *const u[3], double *const out[3])
{
	void *st = ns->c->stream;
	const ssx_level *L = &ns->c->lev[0];
	const double *const *ti = (const double *const *) ns->ti;
	TRYN(ssxk_ns_jac(st, ns->nt, ns->n, L->dmat, (const double *const *) u,
			ns->J));
	TRYN(ssxk_ns_curl(st, ns->nt, (const double *const *) ns->J, ti, NULL,
			ns->a));
	TRYN(ssxk_ns_jac(st, ns->nt, ns->n, L->dmat,
			(const double *const *) ns->a, ns->J));
	return ssxk_ns_curl(st, ns->nt, (const double *const *) ns->J, ti, L->mass,
			out);
}

/* y += alpha U(t), the velocity history interpolated to time t with
 * Lagrange weights, terms of weight |w| <= 1e-6 left out
 * (add_interpolated_velocity, time.c:132-152) */
static int ns_add_interp(ssx_ns *ns, double *const y[3], double t,
		double alpha)
{
	const double t0 = ns->t, dt0 = ns->dt;
	const double d[3] = { t - t0 + 2.0 * dt0, t - t0 + dt0, t - t0 };
	const double a = 0.5 * d[1] * d[2] / pow(dt0, 2);
	const double b = -d[0] * d[2] / pow(dt0, 2);
	const double c = 0.5 * d[0] * d[1] / pow(dt0, 2);
	const double eps = 1.0e-6;
	const double *x[3][3];
	for (int k = 0; k < 3; ++k) {
		x[0][k] = ns->u[2][k];
		x[1][k] = ns->u[1][k];
		x[2][k] = ns->u[0][k];
	}
	const double co[3] = { a * alpha, b * alpha, c * alpha };
	const int use[3] = { fabs(a) > eps, fabs(b) > eps, fabs(c) > eps };
	return ssxk_ns_comb(ns->c->stream, ns->nt, x, co, use, y);
}

/* f = M^-1 DSS(f) per component (trivec_inverse_mass_matrix,
 * time.c:154-167) */
static int ns_inv_mass(ssx_ns *ns, double *const f[3])
{
	ssx_ctx *c = ns->c;
	const ssx_level *L = &c->lev[0];
	const double *im[3] = { c->inv_mass, c->inv_mass, c->inv_mass };
	for (int d = 0; d < 3; ++d)
		TRYN(ssxk_dss_all(c->stream, L->grp, L->ngrp, f[d], NULL, NULL, 0,
				NULL));
	return ssxk_ns_mul3(c->stream, ns->nt, 1.0, (const double *const *) f,
			im, 0.0, f);
}

/* integrate_convection_rk4 (time.c:84-130) on ut */
static int ns_rk4(ssx_ns *ns, double t0, double t, double h, double beta)
{
	void *st = ns->c->stream;
	static const double ra[4] = { 0.0, 0.5, 0.5, 1.0 };
	static const double rb[4] = { 1.0, 2.0, 2.0, 1.0 };
	const size_t nt = ns->nt;
	TRYN(ns_add_interp(ns, ns->ut, t0, beta));
	int k = 0;
	while (t0 + (double) k++ * h < t - 0.5 * h) {
		for (int i = 0; i < 4; ++i) {
			for (int d = 0; d < 3; ++d)
				TRYN(ssxk_copy(st, ns->rky[i][d], ns->ut[d], nt));
			if (i > 0) {
				const double *x[3][3] = { { ns->rkf[i - 1][0],
					ns->rkf[i - 1][1], ns->rkf[i - 1][2] } };
				const double co[3] = { -h * ra[i], 0.0, 0.0 };
				const int use[3] = { 1, 0, 0 };
				TRYN(ssxk_ns_comb(st, nt, x, co, use, ns->rky[i]));
			}
			for (int d = 0; d < 3; ++d)
				TRYN(ssxk_fill(st, ns->a[d], 0.0, nt));
			TRYN(ns_add_interp(ns, ns->a, t0 + ((double) (k - 1) + ra[i])
					* h, 1.0));
			TRYN(ns_convection(ns, ns->a, ns->rky[i], ns->rkf[i]));
			TRYN(ns_inv_mass(ns, ns->rkf[i]));
		}
		for (int i = 0; i < 4; ++i) {
			const double *x[3][3] = { { ns->rkf[i][0], ns->rkf[i][1],
				ns->rkf[i][2] } };
			const double co[3] = { -h * rb[i] / 6.0, 0.0, 0.0 };
			const int use[3] = { 1, 0, 0 };
			TRYN(ssxk_ns_comb(st, nt, x, co, use, ns->ut));
		}
	}
	return SSX_OK;
}

int ssx_ns_init(ssx_ns *ns, const double *const vel[3][3])
{
	if (!ns || !vel) {
		ssx_set_error("ssx_ns_init: null argument");
		return SSX_EINVAL;
	}
	void *st = ns->c->stream;
	/* initialise_from_function (poisson.c:245-282) with the levels as
	 * arrays: each level shifted into the history, convection and double
	 * curl of the two older ones */
	for (int j = 0; j < 3; ++j) {
		rotate3(ns->u);
		for (int d = 0; d < 3; ++d) {
			if (!vel[j][d]) {
				ssx_set_error("ssx_ns_init: null velocity level %d", j);
				return SSX_EINVAL;
			}
			TRYN(ssxk_copy(st, ns->u[0][d], vel[j][d], ns->nt));
		}
		if (j >= 2)
			continue;
		rotate3(ns->conv);
		rotate3(ns->rot);
		TRYN(ns_convection(ns, ns->u[0], ns->u[0], ns->conv[0]));
		TRYN(ns_double_curl(ns, ns->u[0], ns->rot[0]));
	}
	for (int l = 0; l < 3; ++l)
		TRYN(ssxk_fill(st, ns->p[l], 0.0, ns->nt));
	ns->inited = 1;
	return SSX_OK;
}

int ssx_ns_step(ssx_ns *ns, ssx_stats stats[4])
{
	if (!ns || !ns->inited) {
		ssx_set_error("ssx_ns_step: call ssx_ns_init first");
		return SSX_EINVAL;
	}
	ssx_ctx *c = ns->c;
	const ssx_level *L = &c->lev[0];
	void *st = c->stream;
	const size_t nt = ns->nt;
	const double dt = ns->dt;
	ssx_stats sloc[4];
	memset(sloc, 0, sizeof(sloc));
	/* -- lagrangian_convection_integration (navier_stokes.c:71-92) -- */
	rotate3(ns->conv);
	TRYN(ns_convection(ns, ns->u[0], ns->u[0], ns->conv[0]));
	for (int d = 0; d < 3; ++d)
		TRYN(ssxk_fill(st, ns->ut[d], 0.0, nt));
	const double t = ns->t, h = dt / 2.0;
	TRYN(ns_rk4(ns, t - 2.0 * dt, t - dt, h, 1.0 / 3.0));
	TRYN(ns_rk4(ns, t - dt, t, h, -3.0 / 2.0));
	TRYN(ns_rk4(ns, t, t + dt, h, 3.0));
	{   /* u~ += dt f, f = g */
		for (int d = 0; d < 3; ++d)
			TRYN(ssxk_fill(st, ns->a[d], ns->g[d], nt));
		const double *x[3][3] = { { ns->a[0], ns->a[1], ns->a[2] } };
		const double co[3] = { dt, 0.0, 0.0 };
		const int use[3] = { 1, 0, 0 };
		TRYN(ssxk_ns_comb(st, nt, x, co, use, ns->ut));
	}
	/* -- compute_pressure_rhs (navier_stokes.c:115-146) -- */
	{
		/* estimate_pressure_gradient (stokes.c:107-130) into b: f w detJ,
		 * minus EXT3 of the convection, minus visc EXT3 of the double
		 * curl, times dens */
		for (int d = 0; d < 3; ++d)
			TRYN(ssxk_fill(st, ns->a[d], ns->g[d], nt));
		const double *ms[3] = { L->mass, L->mass, L->mass };
		TRYN(ssxk_ns_mul3(st, nt, 1.0, (const double *const *) ns->a, ms, 0.0,
				ns->b));
		rotate3(ns->rot);
		TRYN(ns_double_curl(ns, ns->u[0], ns->rot[0]));
		const double *xc[3][3], *xr[3][3];
		for (int l = 0; l < 3; ++l)
			for (int d = 0; d < 3; ++d) {
				xc[l][d] = ns->conv[l][d];
				xr[l][d] = ns->rot[l][d];
			}
		const int use[3] = { 1, 1, 1 };
		/* stiffly_stable EXPLICIT3 with c = -1, then c = -visc */
		const double cc[3] = { 3.0 * -1.0, -3.0 * -1.0, 1.0 * -1.0 };
		const double cr[3] = { 3.0 * -ns->visc, -3.0 * -ns->visc,
			1.0 * -ns->visc };
		TRYN(ssxk_ns_comb(st, nt, xc, cc, use, ns->b));
		TRYN(ssxk_ns_comb(st, nt, xr, cr, use, ns->b));
		TRYN(ssxk_ns_mul3(st, nt, ns->dens, (const double *const *) ns->b,
				NULL, 0.0, ns->b));
		/* DSS and inverse mass of each component */
		TRYN(ns_inv_mass(ns, ns->b));
		/* rhs = -dens/dt div u~ (weak), surface terms on Dirichlet faces */
		TRYN(ssxk_ns_jac(st, nt, ns->n, L->dmat,
				(const double *const *) ns->ut, ns->J));
		TRYN(ssxk_ns_div(st, nt, (const double *const *) ns->J,
				(const double *const *) ns->ti, L->mass, -ns->dens / dt,
				ns->rhs[0]));
		for (int d = 0; d < 3; ++d)
			TRYN(ssxk_ns_surf(st, ns->nel, ns->n, L->bc, 1, ns->wn[d],
					ns->b[d], ns->rhs[0]));
		TRYN(ssxk_dss_all(st, L->grp, L->ngrp, ns->rhs[0], NULL, NULL, 0,
				NULL));
	}
	/* estimate_pressure (navier_stokes.c:101-113) and the solve */
	{
		/* the three-component combination on (p, p, p): the three
		 * outputs are equal, a[0] is kept */
		const double *x[3][3] = { { ns->p[0], ns->p[0], ns->p[0] },
			{ ns->p[1], ns->p[1], ns->p[1] },
			{ ns->p[2], ns->p[2], ns->p[2] } };
		for (int d = 0; d < 3; ++d)
			TRYN(ssxk_fill(st, ns->a[d], 0.0, nt));
		const double co[3] = { 3.0, -3.0, 1.0 };
		const int use[3] = { 1, 1, 1 };
		TRYN(ssxk_ns_comb(st, nt, x, co, use, ns->a));
		double *tp = ns->p[2];
		ns->p[2] = ns->p[1];
		ns->p[1] = ns->p[0];
		ns->p[0] = tp;
		TRYN(ssxk_copy(st, ns->p[0], ns->a[0], nt));
	}
	int rc = ssx_fcg(c, SSX_OP_POISSON, SSX_PRESSURE, SSX_PC_MG, ns->rhs[0],
			ns->p[0], NS_TOL, 5000, &sloc[0]);
	if (rc != SSX_OK && rc != SSX_ENOTCONV)
		return rc;
	/* -- compute_velocity_rhs (navier_stokes.c:177-201) -- */
	TRYN(ssxk_ns_mul3(st, nt, 1.0 / (dt * ns->visc),
			(const double *const *) ns->ut, NULL, 0.0, ns->ut));
	TRYN(ssxk_ns_grad(st, nt, ns->n, L->dmat, ns->p[0], L->mass,
			(const double *const *) ns->ti, ns->b));
	{
		const double *ms[3] = { L->mass, L->mass, L->mass };
		TRYN(ssxk_ns_mul3(st, nt, 1.0, (const double *const *) ns->ut, ms,
				0.0, ns->rhs));
	}
	for (int d = 0; d < 3; ++d)
		TRYN(ssxk_ns_surf(st, ns->nel, ns->n, L->bc, 0, ns->wn[d], ns->p[0],
				ns->b[d]));
	{
		const double *x[3][3] = { { ns->b[0], ns->b[1], ns->b[2] } };
		const double co[3] = { -1.0 / (ns->visc * ns->dens), 0.0, 0.0 };
		const int use[3] = { 1, 0, 0 };
		TRYN(ssxk_ns_comb(st, nt, x, co, use, ns->rhs));
	}
	for (int d = 0; d < 3; ++d)
		TRYN(ssxk_dss_all(st, L->grp, L->ngrp, ns->rhs[d], NULL, NULL, 0,
				NULL));
	/* estimate_velocity (navier_stokes.c:161-175) */
	{
		const double *x[3][3];
		for (int l = 0; l < 3; ++l)
			for (int d = 0; d < 3; ++d)
				x[l][d] = ns->u[l][d];
		for (int d = 0; d < 3; ++d)
			TRYN(ssxk_fill(st, ns->a[d], 0.0, nt));
		const double co[3] = { 3.0, -3.0, 1.0 };
		const int use[3] = { 1, 1, 1 };
		TRYN(ssxk_ns_comb(st, nt, x, co, use, ns->a));
		rotate3(ns->u);
		for (int d = 0; d < 3; ++d)
			TRYN(ssxk_copy(st, ns->u[0][d], ns->a[d], nt));
		TRYN(ssxk_ns_zero_dir(st, nt, ns->n, L->bc, ns->u[0]));
	}
	/* the three Helmholtz solves (navier_stokes.c:218-221) */
	if (ns->batched) {
		double *fs[3] = { ns->rhs[0], ns->rhs[1], ns->rhs[2] };
		double *us[3] = { ns->u[0][0], ns->u[0][1], ns->u[0][2] };
		const double eps[3] = { NS_TOL, NS_TOL, NS_TOL };
		rc = ssx_batch_cg(ns->vb, fs, us, eps, 5000, 1, sloc + 1);
		if (rc != SSX_OK && rc != SSX_ENOTCONV)
			return rc;
	} else {
		for (int d = 0; d < 3; ++d) {
			rc = ssx_cg(c, 0, SSX_OP_HELMHOLTZ, SSX_X_VELOCITY + d,
					ns->rhs[d], ns->u[0][d], NS_TOL, 5000, &sloc[1 + d]);
			if (rc != SSX_OK && rc != SSX_ENOTCONV)
				return rc;
		}
	}
	ns->t += dt;                     /* navier_stokes.c:234 */
	if (stats)
		memcpy(stats, sloc, sizeof(sloc));
	return SSX_OK;
}

double *ssx_ns_field(ssx_ns *ns, int which, int level)
{
	if (!ns || level < 0 || level > 2)
		return NULL;
	if (which >= SSX_NS_U && which <= SSX_NS_W)
		return ns->u[level][which - SSX_NS_U];
	if (which == SSX_NS_P)
		return ns->p[level];
	return NULL;
}

int ssx_ns_read(ssx_ns *ns, int which, int level, double *dst)
{
	double *src = ssx_ns_field(ns, which, level);
	if (!src || !dst) {
		ssx_set_error("ssx_ns_read: bad field, level or destination");
		return SSX_EINVAL;
	}
	return ssxk_copy(ns->c->stream, dst, src, ns->nt);
}

double ssx_ns_time(const ssx_ns *ns)
{
	return ns ? ns->t : 0.0;
}
