// ssx_particles.hip -- inertial point particles on gfx950: spectral point
// interpolation and the element walk of SSENSS's particles.c, one thread
// per particle.
//
// Each thread runs the reference's per-particle relocation loop
// (particles.c:187-217): RK4 of dx/dt = v, dv/dt = (u_f - v)/tau in the
// element's reference coordinates; on leaving [-1,1]^3 a linearised step to
// the boundary, a hop to the closest face's neighbour (or a reflection at a
// wall), re-entry on the neighbour's matching face, and another pass.  The
// passes of different particles are independent, so the reference's global
// "while any particle is unfinished" loop is the per-particle loop here.
//
// Floating-point contraction is off in this file: every product and sum is
// rounded where the C reference rounds it, so positions and owning elements
// come out bit-identical to the reference's for n > 5 (where it interpolates
// with plain loops) and to rounding for n <= 5 (where it calls BLAS dgemv).
#include <hip/hip_runtime.h>
#include "ssx_internal.h"

#pragma clang fp contract(off)

#define PMAXN SSX_NMAX

struct PField {
	int n;
	int nel;
	const double *xi;       // n GLL points
	const double *ur[3][3]; // [history level: 0 now, 1 old, 2 older][comp]
	const double *jac[9];   // J_ab
	const double *jinv[9];  // (J^-1)_ab
	const int *neigh;       // nel*6
	double T, dt;           // params->t, params->dt
};

// the GLL points of the field, staged in LDS by each kernel (the field
// argument itself stays unmodified, so its arrays keep global loads)
static __shared__ double s_xi[PMAXN];

// Lagrange polynomial, interpolation.c:31-39
__device__ __forceinline__ double lagr(double xi, const double *x, int j, int n)
{
	double li = 1.0;
	for (int i = 0; i < j; ++i)
		li *= (xi - x[i]) / (x[j] - x[i]);
	for (int i = j + 1; i < n; ++i)
		li *= (xi - x[i]) / (x[j] - x[i]);
	return li;
}

__device__ __forceinline__ int first_pt(const double *x, double xp, int n, int m)
{
	for (int i = 0; i < n - m + 1; ++i)
		if (x[i] > xp)
			return (i - m / 2) > 0 ? i - m / 2 : 0;
	return n - m;
}

// interpolation weights of one point (shared by all fields at that point)
struct IW {
	int full;               // n <= 5: full tensor product
	int i1, j1, k1;
	double lx[5], ly[5], lz[5];        // n <= 5 or the 5-point stencil
};

__device__ __forceinline__ void weights(const PField &F, double x, double y, double z, IW &w)
{
	const int n = F.n;
	w.full = n <= 5;
	if (w.full) {
		w.i1 = w.j1 = w.k1 = 0;
		for (int i = 0; i < n; ++i) {
			w.lx[i] = lagr(x, s_xi, i, n);
			w.ly[i] = lagr(y, s_xi, i, n);
			w.lz[i] = lagr(z, s_xi, i, n);
		}
		return;
	}
	w.i1 = first_pt(s_xi, x, n, 5);
	w.j1 = first_pt(s_xi, y, n, 5);
	w.k1 = first_pt(s_xi, z, n, 5);
	for (int i = 0; i < 5; ++i) {
		w.lx[i] = lagr(x, s_xi + w.i1, i, 5);
		w.ly[i] = lagr(y, s_xi + w.j1, i, 5);
		w.lz[i] = lagr(z, s_xi + w.k1, i, 5);
	}
}

// polynomial_interpolation, interpolation.c:41-96
__device__ __forceinline__ double interp(const PField &F, const IW &w, const double *f)
{
	const int n = F.n;
	if (w.full) {
		double out = 0.0;
		for (int i = 0; i < n; ++i) {
			double s = 0.0;
			for (int j = 0; j < n; ++j) {
				double t = 0.0;
				for (int k = 0; k < n; ++k)
					t += f[(k * n + j) * n + i] * w.lz[k];
				s += t * w.ly[j];
			}
			out += w.lx[i] * s;
		}
		return out;
	}
	double tmp = 0.0;
	for (int k = 0; k < 5; ++k)
		for (int j = 0; j < 5; ++j)
			for (int i = 0; i < 5; ++i)
				tmp += f[((w.k1 + k) * n + w.j1 + j) * n + w.i1 + i]
					* w.lx[i] * w.ly[j] * w.lz[k];
	return tmp;
}

struct PState {
	int el, old_el, status, reinit;
	double s[6];            // x, y, z, u, v, w
	double t, tau;
};

enum { P_NORMAL, P_COLLISION, P_OUTSIDE };

// fluid_point_velocity, particles.c:493-515 (all three components)
__device__ __forceinline__ void fluid(const PField &F, int el, const double *y, double t,
		double *uf)
{
	const double eps = 1.0e-6;
	const double d0 = t - F.T + 2.0 * F.dt, d1 = t - F.T + F.dt, d2 = t - F.T;
	const double dt2 = F.dt * F.dt;
	const double a = 0.5 * d1 * d2 / dt2;
	const double b = -d0 * d2 / dt2;
	const double c = 0.5 * d0 * d1 / dt2;
	IW w;
	weights(F, y[0], y[1], y[2], w);
	const size_t off = (size_t) el * F.n * F.n * F.n;
	for (int comp = 0; comp < 3; ++comp) {
		double u0 = 0.0, u1 = 0.0, u2 = 0.0;
		if (fabs(a) > eps)
			u0 = interp(F, w, F.ur[2][comp] + off);
		if (fabs(b) > eps)
			u1 = interp(F, w, F.ur[1][comp] + off);
		if (fabs(c) > eps)
			u2 = interp(F, w, F.ur[0][comp] + off);
		uf[comp] = a * u0 + b * u1 + c * u2;
	}
}

__device__ __forceinline__ void rhs(const PField &F, const PState &p, const double *y,
		double t, double *f)
{
	const double a = 1.0 / p.tau;           // particles.c:228-236
	double uf[3];
	fluid(F, p.el, y, t, uf);
	for (int i = 0; i < 3; ++i) {
		f[i] = y[i + 3];
		f[i + 3] = a * (uf[i] - y[i + 3]);
	}
}

__device__ __forceinline__ int crossed(const double *y)
{
	for (int i = 0; i < 3; ++i)
		if (fabs(y[i]) > 1.0)
			return 1;
	return 0;
}

// the reference re-seeds with rand() (particles.c:170-180); here the
// particle restarts at the element centre with the fluid velocity and is
// flagged
__device__ __forceinline__ void reinit(const PField &F, PState &p)
{
	double y[6] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
	double uf[3];
	fluid(F, p.el, y, F.T, uf);
	p.s[0] = p.s[1] = p.s[2] = 0.0;
	p.s[3] = uf[0];
	p.s[4] = uf[1];
	p.s[5] = uf[2];
	p.reinit = 1;
}

// particles.c:238-269
__device__ __forceinline__ void rk4(const PField &F, PState &p, double dt, int force)
{
	const double a[4] = { 0.0, 0.5, 0.5, 1.0 };
	double y[6], f[4][6];
	double y0[6];
	for (int i = 0; i < 6; ++i)
		y0[i] = p.s[i];
	for (int j = 0; j < 4; ++j) {
		const double t = p.t + a[j] * dt;
		for (int i = 0; i < 6; ++i)
			y[i] = j == 0 ? y0[i] : y0[i] + a[j] * dt * f[j - 1][i];
		if (j > 0 && crossed(y) && !force) {
			p.status = P_OUTSIDE;
			return;
		}
		rhs(F, p, y, t, f[j]);
	}
	for (int i = 0; i < 6; ++i)
		y[i] = y0[i] + dt * (f[0][i] / 6.0 + f[1][i] / 3.0 + f[2][i] / 3.0
				+ f[3][i] / 6.0);
	if (crossed(y)) {
		p.status = P_OUTSIDE;
		return;
	}
	for (int i = 0; i < 6; ++i)
		p.s[i] = y[i];
	p.t = F.T;
}

__device__ __forceinline__ double dmin2(double a, double b)
{
	return a < b ? a : b;                   // utils.c:220-226
}

__device__ __forceinline__ double step_to_boundary(const double *y, const double *f)
{
	double d[3];                            // particles.c:310-316
	for (int i = 0; i < 3; ++i)
		d[i] = (copysign(1.0, f[i]) - y[i]) / f[i];
	return dmin2(dmin2(d[0], d[1]), d[2]);
}

// particles.c:271-308
__device__ __forceinline__ void to_boundary(const PField &F, PState &p, double dt)
{
	const double a[4] = { 0.0, 0.5, 0.5, 1.0 };
	double y[6], f[5][6], y0[6];
	for (int i = 0; i < 6; ++i)
		y0[i] = p.s[i];
	double dtau = 0.0;
	for (int j = 0; j < 4; ++j) {
		const double t = p.t + a[j] * dtau;
		for (int i = 0; i < 6; ++i)
			y[i] = j == 0 ? y0[i] : y0[i] + a[j] * dtau * f[j - 1][i];
		rhs(F, p, y, t, f[j]);
		dtau = step_to_boundary(y0, f[j]);
	}
	for (int i = 0; i < 6; ++i)
		f[4][i] = (f[0][i] / 6.0 + f[1][i] / 3.0 + f[2][i] / 3.0
				+ f[3][i] / 6.0);
	dtau = step_to_boundary(y0, f[4]);
	if (dtau < dt) {
		p.t += dtau;
		for (int i = 0; i < 6; ++i)
			p.s[i] += dtau * f[4][i];
	} else {
		p.status = P_NORMAL;
		rk4(F, p, dt, 1);
		if (p.status != P_NORMAL) {
			p.status = P_NORMAL;
			reinit(F, p);
			p.t = F.T;
		}
	}
}

// particles.c:380-393
__device__ __forceinline__ int closest3(const double *s)
{
	double hmin = 2.0;
	int k = 0;
	for (int i = 0; i < 6; ++i) {
		const double h = fabs(s[i / 2] - ((i % 2) ? 1.0 : -1.0));
		if (h < hmin) {
			hmin = h;
			k = i;
		}
	}
	return k;
}

__device__ __forceinline__ int closest(const PState &p)
{
	return closest3(p.s);
}

// physical / parametric particle velocity, particles.c:517-539; the
// callers overwrite u, v, w one component at a time, so later components
// see the already converted earlier ones (kept as in the reference)
__device__ __forceinline__ double conv_vel(const PField &F, const PState &p,
		const double *const *m, int comp)
{
	IW w;
	weights(F, p.s[0], p.s[1], p.s[2], w);
	const size_t off = (size_t) p.el * F.n * F.n * F.n;
	double dx[3];
	for (int i = 0; i < 3; ++i)
		dx[i] = interp(F, w, m[comp * 3 + i] + off);
	return p.s[3] * dx[0] + p.s[4] * dx[1] + p.s[5] * dx[2];
}

// particles.c:333-349 (single rank)
__device__ __forceinline__ void new_element(const PField &F, PState &p)
{
	const int j = closest(p);
	const int nb = F.neigh[p.el * 6 + j];
	if (nb >= 0) {
		for (int i = 0; i < 3; ++i)
			p.s[3 + i] = conv_vel(F, p, F.jac, i);
		p.el = nb;
	} else {
		p.status = P_COLLISION;
		p.s[3 + j / 2] *= -1.0;
	}
}

// particles.c:351-378 (neigh_rot is always 0)
__device__ __forceinline__ void new_location(const PField &F, PState &p)
{
	double y0[3] = { p.s[0], p.s[1], p.s[2] };
	double x2[2];
	int k = closest(p);
	int j = 0;
	for (int i = 0; i < 3; ++i)
		if (i != k / 2)
			x2[j++] = y0[i];
	for (int i = 0; i < 6; ++i)
		if (F.neigh[p.el * 6 + i] == p.old_el)
			k = i;
	p.s[k / 2] = (k % 2) ? 1.0 : -1.0;
	j = 0;
	for (int i = 0; i < 3; ++i) {
		if (i != k / 2)
			p.s[i] = x2[j++];
		p.s[i + 3] = conv_vel(F, p, F.jinv, i);
	}
	p.old_el = p.el;
}

__global__ void __launch_bounds__(64)
k_particles_step(const PField F, long n, int *__restrict__ el,
		int *__restrict__ old_el, double *__restrict__ x,
		double *__restrict__ y, double *__restrict__ z,
		double *__restrict__ u, double *__restrict__ v,
		double *__restrict__ w, double *__restrict__ t,
		const double *__restrict__ tau, int *__restrict__ reinit_out,
		int *__restrict__ flags)
{
	if (threadIdx.x < F.n)
		s_xi[threadIdx.x] = F.xi[threadIdx.x];
	__syncthreads();
	const long q = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= n)
		return;
	PState p;
	p.el = el[q];
	if (p.el < 0 || p.el >= F.nel) {        // not an element: report, skip
		atomicAdd(flags + 1, 1);
		return;
	}
	p.old_el = old_el[q];
	p.s[0] = x[q]; p.s[1] = y[q]; p.s[2] = z[q];
	p.s[3] = u[q]; p.s[4] = v[q]; p.s[5] = w[q];
	p.t = t[q];
	p.tau = tau[q];
	p.reinit = 0;
	p.status = P_OUTSIDE;
	int j = 0;
	do {                                    // particles.c:187-217
		if (p.status != P_NORMAL) {
			p.status = P_NORMAL;
			if (j > 10) {
				reinit(F, p);
				p.t = F.T;
			} else {
				rk4(F, p, F.T - p.t, 0);
				if (p.status == P_OUTSIDE)
					to_boundary(F, p, 2.0 * F.dt);
				if (p.status == P_OUTSIDE)
					new_element(F, p);
			}
		}
		if (p.status == P_OUTSIDE)
			new_location(F, p);
		++j;
	} while (p.status != P_NORMAL);
	el[q] = p.el;
	old_el[q] = p.old_el;
	x[q] = p.s[0]; y[q] = p.s[1]; z[q] = p.s[2];
	u[q] = p.s[3]; v[q] = p.s[4]; w[q] = p.s[5];
	t[q] = p.t;
	if (reinit_out)
		reinit_out[q] = p.reinit;
	atomicMax(flags, j);                    // a maximum: order-independent
}

// ---------------------------------------------------------------------
// RK2 tracers (dx/dt = u_f(x, t), no inertia): the tracer mode of the
// walk, an extension with no reference counterpart.  Heun's method in the
// reference coordinates of the owning element; a stage that leaves
// [-1,1]^3 is replaced by a straight move to the boundary along the slope
// that left (k1 if the predictor left, else (k1 + k2)/2), a hop across the
// closest face with the crossed coordinate set to the opposite face, and
// a new step over the remaining time.  At a wall the tracer stays on the
// wall for the rest of the step; after more than 10 passes it restarts at
// the element centre (flagged), as the inertial walk does.
// ---------------------------------------------------------------------

__device__ __forceinline__ double tstep_to_boundary(const double *y, const double *f)
{
	double d[3];                    // an axis without motion: never
	for (int i = 0; i < 3; ++i)
		d[i] = f[i] != 0.0 ? (copysign(1.0, f[i]) - y[i]) / f[i] : INFINITY;
	return dmin2(dmin2(d[0], d[1]), d[2]);
}

__global__ void __launch_bounds__(64)
k_tracers_step(const PField F, long n, int *__restrict__ el,
		int *__restrict__ old_el, double *__restrict__ x,
		double *__restrict__ y, double *__restrict__ z,
		double *__restrict__ u, double *__restrict__ v,
		double *__restrict__ w, double *__restrict__ t,
		int *__restrict__ reinit_out, int *__restrict__ flags)
{
	if (threadIdx.x < F.n)
		s_xi[threadIdx.x] = F.xi[threadIdx.x];
	__syncthreads();
	const long q = (long) blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= n)
		return;
	int e = el[q];
	if (e < 0 || e >= F.nel) {              // not an element: report, skip
		atomicAdd(flags + 1, 1);
		return;
	}
	int oe = old_el[q], re = 0;
	double xs[3] = { x[q], y[q], z[q] };
	double tt = t[q];
	int j = 0;
	for (;;) {
		if (j > 10) {
			xs[0] = xs[1] = xs[2] = 0.0;
			re = 1;
			tt = F.T;
			break;
		}
		++j;
		const double h = F.T - tt;
		if (!(h > 0.0))
			break;
		double k1[3], k2[3], xp[3], xn[3], sl[3];
		fluid(F, e, xs, tt, k1);
		for (int i = 0; i < 3; ++i)
			xp[i] = xs[i] + h * k1[i];
		if (!crossed(xp)) {
			fluid(F, e, xp, tt + h, k2);
			for (int i = 0; i < 3; ++i) {
				sl[i] = 0.5 * (k1[i] + k2[i]);
				xn[i] = xs[i] + h * sl[i];
			}
			if (!crossed(xn)) {
				xs[0] = xn[0]; xs[1] = xn[1]; xs[2] = xn[2];
				tt = F.T;
				break;
			}
		} else {
			sl[0] = k1[0]; sl[1] = k1[1]; sl[2] = k1[2];
		}
		double dtau = tstep_to_boundary(xs, sl);
		if (!(dtau < h))
			dtau = h;
		if (dtau < 0.0)
			dtau = 0.0;
		for (int i = 0; i < 3; ++i)
			xs[i] = xs[i] + dtau * sl[i];
		tt = tt + dtau;
		const int k = closest3(xs);
		const int nb = F.neigh[e * 6 + k];
		xs[k / 2] = (k % 2) ? 1.0 : -1.0;
		if (nb >= 0) {
			oe = e;
			e = nb;
			xs[k / 2] = (k % 2) ? -1.0 : 1.0;
		} else {                        // wall: stay on it
			tt = F.T;
			break;
		}
	}
	double uf[3];
	fluid(F, e, xs, F.T, uf);
	el[q] = e;
	old_el[q] = oe;
	x[q] = xs[0]; y[q] = xs[1]; z[q] = xs[2];
	u[q] = uf[0]; v[q] = uf[1]; w[q] = uf[2];
	t[q] = tt;
	if (reinit_out)
		reinit_out[q] = re;
	atomicMax(flags, j);
}

struct P9 {
	const double *p[9];
};

// ur_i = sum_j (J^-1)_ij u_j (geometry.c:188-198, tripalg.c:224-234)
__global__ void k_particles_ur(size_t nt, P9 jinv,
		const double *__restrict__ u, const double *__restrict__ v,
		const double *__restrict__ w, double *__restrict__ u0,
		double *__restrict__ u1, double *__restrict__ u2)
{
	const double *uu[3] = { u, v, w };
	double *out[3] = { u0, u1, u2 };
	for (size_t q = (size_t) blockIdx.x * blockDim.x + threadIdx.x; q < nt;
			q += (size_t) gridDim.x * blockDim.x) {
		const double a = uu[0][q], b = uu[1][q], c = uu[2][q];
		for (int i = 0; i < 3; ++i) {
			double s = jinv.p[i * 3 + 0][q] * a;
			s = jinv.p[i * 3 + 1][q] * b + s;
			s = jinv.p[i * 3 + 2][q] * c + s;
			out[i][q] = s;
		}
	}
}

extern "C" int ssxk_particles_ur(void *stream, size_t nt,
		const double *const jinv[9], const double *u, const double *v,
		const double *w, double *u0, double *u1, double *u2)
{
	P9 d_jinv;
	for (int a = 0; a < 9; ++a)
		d_jinv.p[a] = jinv[a];
	size_t b = (nt + 255) / 256;
	if (b > 16384)
		b = 16384;
	hipLaunchKernelGGL(k_particles_ur, dim3((unsigned) b), dim3(256), 0,
			(hipStream_t) stream, nt, d_jinv, u, v, w, u0, u1, u2);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) {
		ssx_set_error("k_particles_ur: %s", hipGetErrorString(e));
		return SSX_EHIP;
	}
	return SSX_OK;
}

extern "C" int ssxk_tracers_step(void *stream, int n_pts, int nel,
		const double *xi, const double *const ur[9], const int *neigh,
		double T, double dt, long np, int *el, int *old_el, double *x,
		double *y, double *z, double *u, double *v, double *w, double *t,
		int *reinit, int *flags)
{
	PField F;
	F.n = n_pts;
	F.nel = nel;
	F.xi = xi;
	for (int l = 0; l < 3; ++l)
		for (int c = 0; c < 3; ++c)
			F.ur[l][c] = ur[l * 3 + c];
	for (int a = 0; a < 9; ++a) {
		F.jac[a] = nullptr;
		F.jinv[a] = nullptr;
	}
	F.neigh = neigh;
	F.T = T;
	F.dt = dt;
	if (np <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_tracers_step, dim3((unsigned) ((np + 63) / 64)),
			dim3(64), 0, (hipStream_t) stream, F, np, el, old_el, x, y, z, u,
			v, w, t, reinit, flags);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) {
		ssx_set_error("k_tracers_step: %s", hipGetErrorString(e));
		return SSX_EHIP;
	}
	return SSX_OK;
}

extern "C" int ssxk_particles_step(void *stream, int n_pts, int nel,
		const double *xi,
		const double *const ur[9], const double *const jac[9],
		const double *const jinv[9], const int *neigh, double T, double dt,
		long np, int *el, int *old_el, double *x, double *y, double *z,
		double *u, double *v, double *w, double *t, const double *tau,
		int *reinit, int *flags)
{
	PField F;
	F.n = n_pts;
	F.nel = nel;
	F.xi = xi;
	for (int l = 0; l < 3; ++l)
		for (int c = 0; c < 3; ++c)
			F.ur[l][c] = ur[l * 3 + c];
	for (int a = 0; a < 9; ++a) {
		F.jac[a] = jac[a];
		F.jinv[a] = jinv[a];
	}
	F.neigh = neigh;
	F.T = T;
	F.dt = dt;
	if (np <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_particles_step, dim3((unsigned) ((np + 63) / 64)),
			dim3(64), 0, (hipStream_t) stream, F, np, el, old_el, x, y, z, u,
			v, w, t, tau, reinit, flags);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) {
		ssx_set_error("k_particles_step: %s", hipGetErrorString(e));
		return SSX_EHIP;
	}
	return SSX_OK;
}
