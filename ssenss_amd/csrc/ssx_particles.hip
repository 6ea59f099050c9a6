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
// Processing order: the walk may take the particles in an order `perm`
// (a counting sort by owning element, k_porder_*), so that the 64 lanes of
// a wavefront interpolate from one or two elements' field data and their
// gathers hit the same cache lines.  A particle's result does not depend on
// the order, so the outputs are the same bits in any order.
//
// Registers: the interpolation width is a template argument (NF = n for
// n <= 5, the full tensor product; NF = 0 for the 5-point stencil of
// n > 5), the Runge-Kutta stages keep one slope and a running sum rather
// than all slopes, and the coordinate that a hop sets is chosen by selects,
// so no array is indexed by a run-time value and nothing spills to scratch.
//
// Floating-point contraction is off in this file: every product and sum is
// rounded where the C reference rounds it, so positions and owning elements
// come out bit-identical to the reference's for n > 5 (where it interpolates
// with plain loops) and to rounding for n <= 5 (where it calls BLAS dgemv).
#include <hip/hip_runtime.h>
#include "ssx_internal.h"
#include "ssx_device.h"

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

// v[i] for a run-time i in [0, 3), by selects
__device__ __forceinline__ double pick3(const double *v, int i)
{
	return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
}

// v[i] = a for a run-time i in [0, 3), by selects
__device__ __forceinline__ void put3(double *v, int i, double a)
{
#pragma unroll
	for (int k = 0; k < 3; ++k)
		v[k] = k == i ? a : v[k];
}

// Lagrange polynomial j of the M points x, interpolation.c:31-39
template <int M>
__device__ __forceinline__ double lagr(double xi, const double *x, int j)
{
	double li = 1.0;
#pragma unroll
	for (int i = 0; i < M; ++i)
		if (i != j)
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

// interpolation weights of one point (shared by all fields at that point):
// NF > 0 the full tensor product over n = NF points, NF = 0 the 5-point
// stencil starting at (i1, j1, k1)
template <int NF>
struct IW {
	static constexpr int M = NF > 0 ? NF : 5;
	int i1, j1, k1;
	double lx[M], ly[M], lz[M];
};

template <int NF>
__device__ __forceinline__ void weights(const PField &F, double x, double y,
		double z, IW<NF> &w)
{
	constexpr int M = IW<NF>::M;
	if (NF > 0) {
		w.i1 = w.j1 = w.k1 = 0;
	} else {
		w.i1 = first_pt(s_xi, x, F.n, 5);
		w.j1 = first_pt(s_xi, y, F.n, 5);
		w.k1 = first_pt(s_xi, z, F.n, 5);
	}
#pragma unroll
	for (int i = 0; i < M; ++i) {
		w.lx[i] = lagr<M>(x, s_xi + w.i1, i);
		w.ly[i] = lagr<M>(y, s_xi + w.j1, i);
		w.lz[i] = lagr<M>(z, s_xi + w.k1, i);
	}
}

// five consecutive doubles (a stencil row) in three loads: two of 16 bytes
// at 8-byte alignment and one of 8
typedef double pdbl2 __attribute__((ext_vector_type(2), aligned(8)));

__device__ __forceinline__ void row5(const double *p, double *r)
{
	const pdbl2 a = *(const pdbl2 *) p;
	const pdbl2 b = *(const pdbl2 *) (p + 2);
	r[0] = a.x;
	r[1] = a.y;
	r[2] = b.x;
	r[3] = b.y;
	r[4] = p[4];
}

// polynomial_interpolation, interpolation.c:41-96, of three fields at one
// point in one pass: each field's sum is formed in the reference's order,
// and the three independent sums overlap their latencies
template <int NF>
__device__ __forceinline__ void interp3(const PField &F, const IW<NF> &w,
		const double *f0, const double *f1, const double *f2, double *out)
{
	if (NF > 0) {
		constexpr int n = NF > 0 ? NF : 1;
		double o0 = 0.0, o1 = 0.0, o2 = 0.0;
#pragma unroll
		for (int i = 0; i < n; ++i) {
			double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
			for (int j = 0; j < n; ++j) {
				double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
				for (int k = 0; k < n; ++k) {
					const int q = (k * n + j) * n + i;
					t0 += f0[q] * w.lz[k];
					t1 += f1[q] * w.lz[k];
					t2 += f2[q] * w.lz[k];
				}
				s0 += t0 * w.ly[j];
				s1 += t1 * w.ly[j];
				s2 += t2 * w.ly[j];
			}
			o0 += w.lx[i] * s0;
			o1 += w.lx[i] * s1;
			o2 += w.lx[i] * s2;
		}
		out[0] = o0;
		out[1] = o1;
		out[2] = o2;
		return;
	}
	const int n = F.n;
	double a0 = 0.0, a1 = 0.0, a2 = 0.0;
	const int b = (w.k1 * n + w.j1) * n + w.i1;
	// NF = 0: one stencil plane per loop trip (fewer registers, for the
	// first pass over every particle); NF = -1: the whole stencil unrolled
	// (more loads in flight, for the few particles of the later passes)
	constexpr int KU = NF < 0 ? 5 : 1;
#pragma unroll KU
	for (int k = 0; k < 5; ++k) {
		const double lz = k == 0 ? w.lz[0] : k == 1 ? w.lz[1]
			: k == 2 ? w.lz[2] : k == 3 ? w.lz[3] : w.lz[4];
#pragma unroll
		for (int j = 0; j < 5; ++j) {
			const int q = b + (k * n + j) * n;
			double r0[5], r1[5], r2[5];
			row5(f0 + q, r0);
			row5(f1 + q, r1);
			row5(f2 + q, r2);
#pragma unroll
			for (int i = 0; i < 5; ++i) {
				a0 += r0[i] * w.lx[i] * w.ly[j] * lz;
				a1 += r1[i] * w.lx[i] * w.ly[j] * lz;
				a2 += r2[i] * w.lx[i] * w.ly[j] * lz;
			}
		}
	}
	out[0] = a0;
	out[1] = a1;
	out[2] = a2;
}

struct PState {
	int el, old_el, status, reinit;
	double s[6];            // x, y, z, u, v, w
	double t, tau;
};

enum { P_NORMAL, P_COLLISION, P_OUTSIDE };

// ---------------------------------------------------------------------
// Cooperative evaluation (NF = COOP, the 5-point stencil): one particle per
// 64-lane workgroup, all lanes holding the same state.  The lanes share out
// the stencil's products f * lx * ly * lz (rounded as in interp3), each
// field's sum is formed by one lane in interp3's order, and the sums are
// broadcast: every stencil load of an evaluation is in flight at once,
// which is what the last relocation passes, with few particles each, need.
// ---------------------------------------------------------------------
#define COOP (-2)

static __shared__ double s_coop[9 * 125];

__device__ __forceinline__ double sel5(const double *v, int i)
{
	return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : i == 3 ? v[3]
		: v[4];
}

// Sums over the stencil of fields f(q) for q < nf (lane q's sum), each
// field's products in stencil order; use(q) false leaves the sum 0.
template <typename Fld, typename Use>
__device__ __forceinline__ double coop_sums(int n, const IW<COOP> &w, int nf,
		Fld field, Use use)
{
	const int lane = (int) threadIdx.x;
	const int base = (w.k1 * n + w.j1) * n + w.i1;
	for (int t = lane; t < nf * 125; t += 64) {
		const int q = t / 125, pt = t - 125 * q;
		if (use(q)) {
			const int k = pt / 25, j = pt / 5 - 5 * k, i = pt - 5 * (pt / 5);
			s_coop[t] = field(q)[base + (k * n + j) * n + i] * sel5(w.lx, i)
				* sel5(w.ly, j) * sel5(w.lz, k);
		}
	}
	__syncthreads();
	double acc = 0.0;
	if (lane < nf && use(lane)) {
		const double *sp = s_coop + lane * 125;
#pragma unroll 25
		for (int pt = 0; pt < 125; ++pt)
			acc += sp[pt];
	}
	__syncthreads();
	return acc;
}

// fluid_point_velocity, particles.c:493-515 (all three components)
template <int NF>
__device__ __forceinline__ void fluid(const PField &F, int el,
		const double *y, double t, double *uf)
{
	const double eps = 1.0e-6;
	const double d0 = t - F.T + 2.0 * F.dt, d1 = t - F.T + F.dt, d2 = t - F.T;
	const double dt2 = F.dt * F.dt;
	const double a = 0.5 * d1 * d2 / dt2;
	const double b = -d0 * d2 / dt2;
	const double c = 0.5 * d0 * d1 / dt2;
	IW<NF> w;
	weights<NF>(F, y[0], y[1], y[2], w);
	const size_t off = (size_t) el * F.n * F.n * F.n;
	// u[l][comp] for the history levels whose weight exceeds eps (level
	// 0 = now, weight c; 1 = old, b; 2 = older, a), else 0
	if constexpr (NF == COOP) {
		const bool use0 = fabs(c) > eps, use1 = fabs(b) > eps,
			  use2 = fabs(a) > eps;
		// the field pointers copied out first: a lambda that captured the
		// argument structure would need its address, a private copy
		const double *u00 = F.ur[0][0] + off, *u01 = F.ur[0][1] + off,
			  *u02 = F.ur[0][2] + off, *u10 = F.ur[1][0] + off,
			  *u11 = F.ur[1][1] + off, *u12 = F.ur[1][2] + off,
			  *u20 = F.ur[2][0] + off, *u21 = F.ur[2][1] + off,
			  *u22 = F.ur[2][2] + off;
		const double s = coop_sums(F.n, w, 9, [=](int q) {
			return q == 0 ? u00 : q == 1 ? u01 : q == 2 ? u02 : q == 3 ? u10
				: q == 4 ? u11 : q == 5 ? u12 : q == 6 ? u20 : q == 7 ? u21
				: u22;
		}, [=](int q) { return q < 3 ? use0 : q < 6 ? use1 : use2; });
		double u[9];
#pragma unroll
		for (int q = 0; q < 9; ++q)
			u[q] = __shfl(s, q);
#pragma unroll
		for (int comp = 0; comp < 3; ++comp)
			uf[comp] = a * u[6 + comp] + b * u[3 + comp] + c * u[comp];
	} else {
		const double wl[3] = { c, b, a };
		double u[3][3];
#pragma unroll
		for (int l = 0; l < 3; ++l) {
			u[l][0] = u[l][1] = u[l][2] = 0.0;
			if (fabs(wl[l]) > eps)
				interp3<NF>(F, w, F.ur[l][0] + off, F.ur[l][1] + off,
						F.ur[l][2] + off, u[l]);
		}
#pragma unroll
		for (int comp = 0; comp < 3; ++comp)
			uf[comp] = a * u[2][comp] + b * u[1][comp] + c * u[0][comp];
	}
}

template <int NF>
__device__ __forceinline__ void rhs(const PField &F,
		const PState &p, const double *y, double t, double *f)
{
	const double a = 1.0 / p.tau;           // particles.c:228-236
	double uf[3];
	fluid<NF>(F, p.el, y, t, uf);
#pragma unroll
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
template <int NF>
__device__ __forceinline__ void reinit(const PField &F,
		PState &p)
{
	double y[6] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
	double uf[3];
	fluid<NF>(F, p.el, y, F.T, uf);
	p.s[0] = p.s[1] = p.s[2] = 0.0;
	p.s[3] = uf[0];
	p.s[4] = uf[1];
	p.s[5] = uf[2];
	p.reinit = 1;
}

// the RK4 stage nodes 0, 1/2, 1/2, 1 and the weights' divisors 6, 3, 3, 6
__device__ __forceinline__ double rk_node(int j)
{
	return j == 0 ? 0.0 : (j == 3 ? 1.0 : 0.5);
}

__device__ __forceinline__ double rk_div(int j)
{
	return (j == 0 || j == 3) ? 6.0 : 3.0;
}

// particles.c:238-269.  The slope sum f0/6 + f1/3 + f2/3 + f3/6 is built
// stage by stage in the order the reference adds it.
template <int NF>
__device__ __forceinline__ void rk4(const PField &F,
		PState &p, double dt, int force)
{
	double y0[6], y[6], f[6], acc[6];
#pragma unroll
	for (int i = 0; i < 6; ++i) {
		y0[i] = p.s[i];
		f[i] = acc[i] = 0.0;
	}
	for (int j = 0; j < 4; ++j) {
		const double aj = rk_node(j);
		const double t = p.t + aj * dt;
#pragma unroll
		for (int i = 0; i < 6; ++i)
			y[i] = j == 0 ? y0[i] : y0[i] + aj * dt * f[i];
		if (j > 0 && crossed(y) && !force) {
			p.status = P_OUTSIDE;
			return;
		}
		rhs<NF>(F, p, y, t, f);
		const double dv = rk_div(j);
#pragma unroll
		for (int i = 0; i < 6; ++i)
			acc[i] = j == 0 ? f[i] / dv : acc[i] + f[i] / dv;
	}
#pragma unroll
	for (int i = 0; i < 6; ++i)
		y[i] = y0[i] + dt * acc[i];
	if (crossed(y)) {
		p.status = P_OUTSIDE;
		return;
	}
#pragma unroll
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
#pragma unroll
	for (int i = 0; i < 3; ++i)
		d[i] = (copysign(1.0, f[i]) - y[i]) / f[i];
	return dmin2(dmin2(d[0], d[1]), d[2]);
}

// particles.c:271-308 (the slope sum built stage by stage, as in rk4)
template <int NF>
__device__ __forceinline__ void to_boundary(const PField &F,
		PState &p, double dt)
{
	double y[6], f[6], acc[6], y0[6];
#pragma unroll
	for (int i = 0; i < 6; ++i) {
		y0[i] = p.s[i];
		f[i] = acc[i] = 0.0;
	}
	double dtau = 0.0;
	for (int j = 0; j < 4; ++j) {
		const double aj = rk_node(j);
		const double t = p.t + aj * dtau;
#pragma unroll
		for (int i = 0; i < 6; ++i)
			y[i] = j == 0 ? y0[i] : y0[i] + aj * dtau * f[i];
		rhs<NF>(F, p, y, t, f);
		dtau = step_to_boundary(y0, f);
		const double dv = rk_div(j);
#pragma unroll
		for (int i = 0; i < 6; ++i)
			acc[i] = j == 0 ? f[i] / dv : acc[i] + f[i] / dv;
	}
	dtau = step_to_boundary(y0, acc);
	if (dtau < dt) {
		p.t += dtau;
#pragma unroll
		for (int i = 0; i < 6; ++i)
			p.s[i] += dtau * acc[i];
	} else {
		p.status = P_NORMAL;
		rk4<NF>(F, p, dt, 1);
		if (p.status != P_NORMAL) {
			p.status = P_NORMAL;
			reinit<NF>(F, p);
			p.t = F.T;
		}
	}
}

// particles.c:380-393
__device__ __forceinline__ int closest3(const double *s)
{
	double hmin = 2.0;
	int k = 0;
#pragma unroll
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
template <int NF>
__device__ __forceinline__ double conv_vel(const PField &F, const PState &p,
		const double *const *m, int comp)
{
	IW<NF> w;
	weights<NF>(F, p.s[0], p.s[1], p.s[2], w);
	const size_t off = (size_t) p.el * F.n * F.n * F.n;
	// row comp of the 3x3 factor, picked by selects (an index into the
	// argument's pointer table would copy the table to scratch)
	const double *r0 = comp == 0 ? m[0] : comp == 1 ? m[3] : m[6];
	const double *r1 = comp == 0 ? m[1] : comp == 1 ? m[4] : m[7];
	const double *r2 = comp == 0 ? m[2] : comp == 1 ? m[5] : m[8];
	double dx[3];
	if constexpr (NF == COOP) {
		const double *f0 = r0 + off, *f1 = r1 + off, *f2 = r2 + off;
		const double s = coop_sums(F.n, w, 3, [=](int q) {
			return q == 0 ? f0 : q == 1 ? f1 : f2;
		}, [](int) { return true; });
		dx[0] = __shfl(s, 0);
		dx[1] = __shfl(s, 1);
		dx[2] = __shfl(s, 2);
	} else {
		interp3<NF>(F, w, r0 + off, r1 + off, r2 + off, dx);
	}
	return p.s[3] * dx[0] + p.s[4] * dx[1] + p.s[5] * dx[2];
}

// particles.c:333-349 (single rank)
template <int NF>
__device__ __forceinline__ void new_element(const PField &F, PState &p)
{
	const int j = closest(p);
	const int nb = F.neigh[p.el * 6 + j];
	if (nb >= 0) {
		for (int i = 0; i < 3; ++i)
			p.s[3 + i] = conv_vel<NF>(F, p, F.jac, i);
		p.el = nb;
	} else {
		p.status = P_COLLISION;
		put3(p.s + 3, j / 2, pick3(p.s + 3, j / 2) * -1.0);
	}
}

// particles.c:351-378 (neigh_rot is always 0): the entry face's
// coordinate is set to +-1, the other two keep their values in order
template <int NF>
__device__ __forceinline__ void new_location(const PField &F, PState &p)
{
	const double y0[3] = { p.s[0], p.s[1], p.s[2] };
	int k = closest(p);
	const int d0 = k / 2;
	const double x2a = d0 == 0 ? y0[1] : y0[0];     // x2[0], x2[1]: the
	const double x2b = d0 == 2 ? y0[1] : y0[2];     // other two, in order
	for (int i = 0; i < 6; ++i)
		if (F.neigh[p.el * 6 + i] == p.old_el)
			k = i;
	const int d = k / 2;
	put3(p.s, d, (k % 2) ? 1.0 : -1.0);
	int j = 0;
	for (int i = 0; i < 3; ++i) {
		if (i != d) {
			put3(p.s, i, j == 0 ? x2a : x2b);
			++j;
		}
		put3(p.s + 3, i, conv_vel<NF>(F, p, F.jinv, i));
	}
	p.old_el = p.el;
}

// one iteration of particles.c:187-217's relocation loop; j counts them
template <int NF>
__device__ __forceinline__ void particle_iter(const PField &F,
		PState &p, int &j)
{
	if (p.status != P_NORMAL) {
		p.status = P_NORMAL;
		if (j > 10) {
			reinit<NF>(F, p);
			p.t = F.T;
		} else {
			rk4<NF>(F, p, F.T - p.t, 0);
			if (p.status == P_OUTSIDE)
				to_boundary<NF>(F, p, 2.0 * F.dt);
			if (p.status == P_OUTSIDE)
				new_element<NF>(F, p);
		}
	}
	if (p.status == P_OUTSIDE)
		new_location<NF>(F, p);
	++j;
}

__device__ __forceinline__ void load_particle(PState &p, long q,
		const int *__restrict__ old_el, const double *__restrict__ x,
		const double *__restrict__ y, const double *__restrict__ z,
		const double *__restrict__ u, const double *__restrict__ v,
		const double *__restrict__ w, const double *__restrict__ t,
		const double *__restrict__ tau)
{
	p.old_el = old_el[q];
	p.s[0] = x[q]; p.s[1] = y[q]; p.s[2] = z[q];
	p.s[3] = u[q]; p.s[4] = v[q]; p.s[5] = w[q];
	p.t = t[q];
	p.tau = tau[q];
}

__device__ __forceinline__ void store_particle(const PState &p, long q,
		int *__restrict__ el, int *__restrict__ old_el,
		double *__restrict__ x, double *__restrict__ y,
		double *__restrict__ z, double *__restrict__ u,
		double *__restrict__ v, double *__restrict__ w,
		double *__restrict__ t)
{
	el[q] = p.el;
	old_el[q] = p.old_el;
	x[q] = p.s[0]; y[q] = p.s[1]; z[q] = p.s[2];
	u[q] = p.s[3]; v[q] = p.s[4]; w[q] = p.s[5];
	t[q] = p.t;
}

// particles.c:187-217 for particle q, all passes
template <int NF>
__device__ __forceinline__ void particle_walk(const PField &F,
		long q, int *__restrict__ el, int *__restrict__ old_el,
		double *__restrict__ x, double *__restrict__ y,
		double *__restrict__ z, double *__restrict__ u,
		double *__restrict__ v, double *__restrict__ w,
		double *__restrict__ t, const double *__restrict__ tau,
		int *__restrict__ reinit_out, int *__restrict__ flags)
{
	PState p;
	p.el = el[q];
	if (p.el < 0 || p.el >= F.nel) {        // not an element: report, skip
		atomicAdd(flags + 1, 1);
		return;
	}
	load_particle(p, q, old_el, x, y, z, u, v, w, t, tau);
	p.reinit = 0;
	p.status = P_OUTSIDE;
	int j = 0;
	do
		particle_iter<NF>(F, p, j);
	while (p.status != P_NORMAL);
	store_particle(p, q, el, old_el, x, y, z, u, v, w, t);
	if (reinit_out)
		reinit_out[q] = p.reinit;
	atomicMax(flags, j);                    // a maximum: order-independent
}

// Appends q to list (count *cnt) in the lanes where want holds: one atomic
// per wavefront, each lane's slot its rank among those lanes.  Called by
// every active lane at the same point.
__device__ __forceinline__ void append(int *__restrict__ list,
		int *__restrict__ cnt, long q, bool want)
{
	const unsigned long long m = __ballot(want);
	if (!m)
		return;
	const int lane = (int) __lane_id();
	const int leader = __ffsll((unsigned long long) m) - 1;
	int base = 0;
	if (lane == leader)
		base = atomicAdd(cnt, __popcll(m));
	base = __shfl(base, leader);
	if (want)
		list[base + __popcll(m & ((1ull << lane) - 1ull))] = (int) q;
}

// pass state of an unfinished particle: status, iterations so far, re-seed
__device__ __forceinline__ int pst_pack(const PState &p, int j)
{
	return p.status | (j << 2) | (p.reinit << 7);
}

// A relocation pass in two phases, so that the lanes of a wavefront do
// the same work: phase A runs the head of the loop body (the re-seed after
// 11 passes, else the RK4 attempt) over the pass's particles; the ones
// whose attempt left the element (status OUTSIDE, state unchanged) go to
// phase B, which runs the rest of the body (step to the boundary, new
// element, entry point).  Together they are particle_iter.
//
// Counts: ca[k] particles enter pass k (ca[0] = np), cm[k] enter its phase
// B.  Pass 0 reads the first np entries of in (null: all particles as
// stored); a pass k > 0 reads in[0 .. ca[k]) and the pass state in pst.

// the state of particle q at the start of pass `pass`
__device__ __forceinline__ void pass_state(int pass, long q,
		const int *__restrict__ pst, PState &p, int &j)
{
	if (pass == 0) {
		p.status = P_OUTSIDE;
		p.reinit = 0;
		j = 0;
	} else {
		const int st = pst[q];
		p.status = st & 3;
		j = (st >> 2) & 31;
		p.reinit = (st >> 7) & 1;
	}
}

__device__ __forceinline__ void particle_done(const PState &p, long q, int j,
		int *__restrict__ reinit_out, int *__restrict__ flags)
{
	if (reinit_out)
		reinit_out[q] = p.reinit;
	atomicMax(flags, j);                    // a maximum: order-independent
}

template <int NF>
__global__ void __launch_bounds__(64)
k_particles_a(const PField F, long np, int pass, long coop_max,
		const int *__restrict__ in, int *__restrict__ mid,
		int *__restrict__ ca, int *__restrict__ cm, int *__restrict__ pst,
		int *__restrict__ el, int *__restrict__ old_el,
		double *__restrict__ x, double *__restrict__ y,
		double *__restrict__ z, double *__restrict__ u,
		double *__restrict__ v, double *__restrict__ w,
		double *__restrict__ t, const double *__restrict__ tau,
		int *__restrict__ reinit_out, int *__restrict__ flags)
{
	// pass 0 (a grid the size of the list) in XCD-contiguous block order:
	// each XCD walks one range of the sorted list, so an element's fields
	// are fetched into one L2.  The later passes' grids outnumber their
	// lists, and round-robin keeps every XCD busy.
	const long lb = pass == 0 ? (long) xcd_block() : (long) blockIdx.x;
	const long m = pass == 0 ? np : (long) ca[pass];
	if (lb * 64 >= m || (pass > 0 && m <= coop_max))
		return;
	if (threadIdx.x < F.n)
		s_xi[threadIdx.x] = F.xi[threadIdx.x];
	__syncthreads();
	for (long i = lb * 64 + threadIdx.x; i < m; i += (long) gridDim.x * 64) {
		const long q = in ? (long) in[i] : i;
		bool more = false;
		PState p;
		p.el = el[q];
		if (pass == 0 && (p.el < 0 || p.el >= F.nel)) {
			atomicAdd(flags + 1, 1);        // not an element: report, skip
		} else {
			int j;
			pass_state(pass, q, pst, p, j);
			load_particle(p, q, old_el, x, y, z, u, v, w, t, tau);
			p.status = P_NORMAL;            // particles.c:189-196
			if (j > 10) {
				reinit<NF>(F, p);
				p.t = F.T;
			} else {
				rk4<NF>(F, p, F.T - p.t, 0);
			}
			more = p.status == P_OUTSIDE;
			if (more) {
				pst[q] = pst_pack(p, j);
			} else {
				++j;
				store_particle(p, q, el, old_el, x, y, z, u, v, w, t);
				particle_done(p, q, j, reinit_out, flags);
			}
		}
		append(mid, cm + pass, q, more);
	}
}

template <int NF>
__global__ void __launch_bounds__(64)
k_particles_b(const PField F, int pass, long coop_max,
		const int *__restrict__ mid, int *__restrict__ out,
		int *__restrict__ ca, int *__restrict__ cm, int *__restrict__ pst,
		int *__restrict__ el, int *__restrict__ old_el,
		double *__restrict__ x, double *__restrict__ y,
		double *__restrict__ z, double *__restrict__ u,
		double *__restrict__ v, double *__restrict__ w,
		double *__restrict__ t, const double *__restrict__ tau,
		int *__restrict__ reinit_out, int *__restrict__ flags)
{
	const long lb = blockIdx.x;
	const long m = cm[pass];
	if (lb * 64 >= m || (pass > 0 && ca[pass] <= coop_max))
		return;
	if (threadIdx.x < F.n)
		s_xi[threadIdx.x] = F.xi[threadIdx.x];
	__syncthreads();
	for (long i = lb * 64 + threadIdx.x; i < m; i += (long) gridDim.x * 64) {
		const long q = mid[i];
		PState p;
		p.el = el[q];
		int j;
		pass_state(1, q, pst, p, j);
		load_particle(p, q, old_el, x, y, z, u, v, w, t, tau);
		to_boundary<NF>(F, p, 2.0 * F.dt);  // particles.c:197-205
		if (p.status == P_OUTSIDE)
			new_element<NF>(F, p);
		if (p.status == P_OUTSIDE)
			new_location<NF>(F, p);
		++j;
		store_particle(p, q, el, old_el, x, y, z, u, v, w, t);
		const bool more = p.status != P_NORMAL;
		if (more)
			pst[q] = pst_pack(p, j);
		else
			particle_done(p, q, j, reinit_out, flags);
		append(out, ca + pass + 1, q, more);
	}
}

// one particle per lane, in the order perm (null: as stored), every field
// read from the global arrays
template <int NF>
__global__ void __launch_bounds__(64)
k_particles_step(const PField F, long n, const int *__restrict__ perm,
		int *__restrict__ el, int *__restrict__ old_el,
		double *__restrict__ x, double *__restrict__ y,
		double *__restrict__ z, double *__restrict__ u,
		double *__restrict__ v, double *__restrict__ w,
		double *__restrict__ t, const double *__restrict__ tau,
		int *__restrict__ reinit_out, int *__restrict__ flags)
{
	if (threadIdx.x < F.n)
		s_xi[threadIdx.x] = F.xi[threadIdx.x];
	__syncthreads();
	const long i = (long) xcd_block() * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	particle_walk<NF>(F, perm ? (long) perm[i] : i, el, old_el, x, y, z,
			u, v, w, t, tau, reinit_out, flags);
}

// A later pass with at most coop_max particles: one particle per 64-lane
// workgroup, the field evaluations cooperative (COOP).  Every lane runs the
// particle's control flow on the same values; lane 0 writes.
__global__ void __launch_bounds__(64)
k_particles_coop(const PField F, int pass, long coop_max,
		const int *__restrict__ list, int *__restrict__ next,
		int *__restrict__ cnt, int *__restrict__ pst,
		int *__restrict__ el, int *__restrict__ old_el,
		double *__restrict__ x, double *__restrict__ y,
		double *__restrict__ z, double *__restrict__ u,
		double *__restrict__ v, double *__restrict__ w,
		double *__restrict__ t, const double *__restrict__ tau,
		int *__restrict__ reinit_out, int *__restrict__ flags)
{
	const long m = cnt[pass];
	if (m > coop_max || (long) blockIdx.x >= m)
		return;
	if (threadIdx.x < F.n)
		s_xi[threadIdx.x] = F.xi[threadIdx.x];
	__syncthreads();
	for (long i = blockIdx.x; i < m; i += gridDim.x) {
		const long q = list[i];
		PState p;
		p.el = el[q];
		const int st = pst[q];
		p.status = st & 3;
		int j = (st >> 2) & 31;
		p.reinit = (st >> 7) & 1;
		load_particle(p, q, old_el, x, y, z, u, v, w, t, tau);
		particle_iter<COOP>(F, p, j);
		if (threadIdx.x == 0) {
			store_particle(p, q, el, old_el, x, y, z, u, v, w, t);
			if (p.status != P_NORMAL) {
				pst[q] = pst_pack(p, j);
				next[atomicAdd(cnt + pass + 1, 1)] = (int) q;
			} else {
				particle_done(p, q, j, reinit_out, flags);
			}
		}
	}
}

// ---------------------------------------------------------------------
// RK2 tracers (dx/dt = u_f(x, t), no inertia): the tracer mode of the
// walk, an extension with no reference counterpart.  Heun's method in the
// reference coordinates of the owning element; a stage that leaves
// [-1,1]^3 is replaced by a straight move to the boundary along the slope
// that left (k1 if the predictor left, else (k1 + k2)/2), a hop across the
// closest face onto the face of the neighbour that leads back (the
// opposite face for aligned elements), and a new step over the remaining
// time.  At a wall the tracer stays on the
// wall for the rest of the step; after more than 10 passes it restarts at
// the element centre (flagged), as the inertial walk does.
// ---------------------------------------------------------------------

__device__ __forceinline__ double tstep_to_boundary(const double *y, const double *f)
{
	double d[3];                    // an axis without motion: never
#pragma unroll
	for (int i = 0; i < 3; ++i)
		d[i] = f[i] != 0.0 ? (copysign(1.0, f[i]) - y[i]) / f[i] : INFINITY;
	return dmin2(dmin2(d[0], d[1]), d[2]);
}

// one iteration of the tracer loop; true when the tracer has arrived
template <int NF>
__device__ __forceinline__ bool tracer_iter(const PField &F,
		int &e, int &oe, double *xs, double &tt, int &j, int &re)
{
	if (j > 10) {
		xs[0] = xs[1] = xs[2] = 0.0;
		re = 1;
		tt = F.T;
		return true;
	}
	++j;
	const double h = F.T - tt;
	if (!(h > 0.0))
		return true;
	double k1[3], k2[3], xp[3], xn[3], sl[3];
	fluid<NF>(F, e, xs, tt, k1);
#pragma unroll
	for (int i = 0; i < 3; ++i)
		xp[i] = xs[i] + h * k1[i];
	if (!crossed(xp)) {
		fluid<NF>(F, e, xp, tt + h, k2);
#pragma unroll
		for (int i = 0; i < 3; ++i) {
			sl[i] = 0.5 * (k1[i] + k2[i]);
			xn[i] = xs[i] + h * sl[i];
		}
		if (!crossed(xn)) {
			xs[0] = xn[0]; xs[1] = xn[1]; xs[2] = xn[2];
			tt = F.T;
			return true;
		}
	} else {
		sl[0] = k1[0]; sl[1] = k1[1]; sl[2] = k1[2];
	}
	double dtau = tstep_to_boundary(xs, sl);
	if (!(dtau < h))
		dtau = h;
	if (dtau < 0.0)
		dtau = 0.0;
#pragma unroll
	for (int i = 0; i < 3; ++i)
		xs[i] = xs[i] + dtau * sl[i];
	tt = tt + dtau;
	const int k = closest3(xs);
	const int nb = F.neigh[e * 6 + k];
	put3(xs, k / 2, (k % 2) ? 1.0 : -1.0);
	if (nb >= 0) {
		// the entry face: the face of nb that leads back to e, the
		// opposite one first (aligned elements; two elements periodic in
		// one direction lead back across both faces), else found as
		// new_location finds it; the other two coordinates keep their order
		int kin = k ^ 1;
		if (F.neigh[nb * 6 + kin] != e)
			for (int i = 0; i < 6; ++i)
				if (F.neigh[nb * 6 + i] == e)
					kin = i;
		const int d0 = k / 2, d = kin / 2;
		const double xa = d0 == 0 ? xs[1] : xs[0];
		const double xb = d0 == 2 ? xs[1] : xs[2];
		int jj = 0;
#pragma unroll
		for (int i = 0; i < 3; ++i)
			if (i != d)
				put3(xs, i, jj++ == 0 ? xa : xb);
		put3(xs, d, (kin % 2) ? 1.0 : -1.0);
		oe = e;
		e = nb;
		return false;
	}
	tt = F.T;                               // wall: stay on it
	return true;
}

// the fluid velocity at the arrival point, and the tracer's outputs
template <int NF>
__device__ __forceinline__ void tracer_finish(const PField &F, bool writer,
		long q, int e, int oe, const double *xs, double tt, int j, int re,
		int *__restrict__ el, int *__restrict__ old_el,
		double *__restrict__ x, double *__restrict__ y,
		double *__restrict__ z, double *__restrict__ u,
		double *__restrict__ v, double *__restrict__ w,
		double *__restrict__ t, int *__restrict__ reinit_out,
		int *__restrict__ flags)
{
	double uf[3];
	fluid<NF>(F, e, xs, F.T, uf);
	if (!writer)
		return;
	el[q] = e;
	old_el[q] = oe;
	x[q] = xs[0]; y[q] = xs[1]; z[q] = xs[2];
	u[q] = uf[0]; v[q] = uf[1]; w[q] = uf[2];
	t[q] = tt;
	if (reinit_out)
		reinit_out[q] = re;
	atomicMax(flags, j);
}

template <int NF>
__device__ __forceinline__ void tracer_walk(const PField &F,
		long q, int *__restrict__ el, int *__restrict__ old_el,
		double *__restrict__ x, double *__restrict__ y,
		double *__restrict__ z, double *__restrict__ u,
		double *__restrict__ v, double *__restrict__ w,
		double *__restrict__ t, int *__restrict__ reinit_out,
		int *__restrict__ flags)
{
	int e = el[q];
	if (e < 0 || e >= F.nel) {              // not an element: report, skip
		atomicAdd(flags + 1, 1);
		return;
	}
	int oe = old_el[q], re = 0;
	double xs[3] = { x[q], y[q], z[q] };
	double tt = t[q];
	int j = 0;
	while (!tracer_iter<NF>(F, e, oe, xs, tt, j, re))
		;
	tracer_finish<NF>(F, true, q, e, oe, xs, tt, j, re, el, old_el, x, y, z,
			u,
			v, w, t, reinit_out, flags);
}

// one pass of the tracer loop (as particle_pass_one; the pass state is the
// iteration count)
template <int NF>
__device__ __forceinline__ void tracer_pass_one(const PField &F,
		int pass, long q, int *__restrict__ next,
		int *__restrict__ cnt, int *__restrict__ pst,
		int *__restrict__ el, int *__restrict__ old_el,
		double *__restrict__ x, double *__restrict__ y,
		double *__restrict__ z, double *__restrict__ u,
		double *__restrict__ v, double *__restrict__ w,
		double *__restrict__ t, int *__restrict__ reinit_out,
		int *__restrict__ flags)
{
	bool more = false;
	int e = el[q];
	if (pass == 0 && (e < 0 || e >= F.nel)) {
		atomicAdd(flags + 1, 1);            // not an element: report, skip
	} else {
		int oe = old_el[q], re = 0, j = pass == 0 ? 0 : pst[q];
		double xs[3] = { x[q], y[q], z[q] };
		double tt = t[q];
		if (tracer_iter<NF>(F, e, oe, xs, tt, j, re)) {
			tracer_finish<NF>(F, true, q, e, oe, xs, tt, j, re, el, old_el, x,
					y, z, u, v, w, t, reinit_out, flags);
		} else {
			more = true;
			pst[q] = j;
			el[q] = e;
			old_el[q] = oe;
			x[q] = xs[0]; y[q] = xs[1]; z[q] = xs[2];
			t[q] = tt;
		}
	}
	append(next, cnt + pass + 1, q, more);
}

template <int NF>
__global__ void __launch_bounds__(64)
k_tracers_step(const PField F, long n, const int *__restrict__ perm,
		int *__restrict__ el, int *__restrict__ old_el,
		double *__restrict__ x, double *__restrict__ y,
		double *__restrict__ z, double *__restrict__ u,
		double *__restrict__ v, double *__restrict__ w,
		double *__restrict__ t, int *__restrict__ reinit_out,
		int *__restrict__ flags)
{
	if (threadIdx.x < F.n)
		s_xi[threadIdx.x] = F.xi[threadIdx.x];
	__syncthreads();
	const long i0 = (long) xcd_block() * blockDim.x + threadIdx.x;
	if (i0 >= n)
		return;
	tracer_walk<NF>(F, perm ? (long) perm[i0] : i0, el, old_el, x, y, z,
			u, v, w, t, reinit_out, flags);
}

template <int NF>
__global__ void __launch_bounds__(64)
k_tracers_pass(const PField F, long np, int pass, long coop_max,
		const int *__restrict__ list, int *__restrict__ next,
		int *__restrict__ cnt, int *__restrict__ pst,
		int *__restrict__ el, int *__restrict__ old_el,
		double *__restrict__ x, double *__restrict__ y,
		double *__restrict__ z, double *__restrict__ u,
		double *__restrict__ v, double *__restrict__ w,
		double *__restrict__ t, int *__restrict__ reinit_out,
		int *__restrict__ flags)
{
	const long lb = pass == 0 ? (long) xcd_block() : (long) blockIdx.x;
	const long m = pass == 0 ? np : (long) cnt[pass];
	if (lb * 64 >= m || (pass > 0 && m <= coop_max))
		return;
	if (threadIdx.x < F.n)
		s_xi[threadIdx.x] = F.xi[threadIdx.x];
	__syncthreads();
	for (long i = lb * 64 + threadIdx.x; i < m; i += (long) gridDim.x * 64)
		tracer_pass_one<NF>(F, pass, list ? (long) list[i] : i, next, cnt,
				pst, el, old_el, x, y, z, u, v, w, t, reinit_out, flags);
}

// the tracers of a later pass with at most coop_max of them, one per
// 64-lane workgroup (as k_particles_coop)
__global__ void __launch_bounds__(64)
k_tracers_coop(const PField F, int pass, long coop_max,
		const int *__restrict__ list, int *__restrict__ next,
		int *__restrict__ cnt, int *__restrict__ pst,
		int *__restrict__ el, int *__restrict__ old_el,
		double *__restrict__ x, double *__restrict__ y,
		double *__restrict__ z, double *__restrict__ u,
		double *__restrict__ v, double *__restrict__ w,
		double *__restrict__ t, int *__restrict__ reinit_out,
		int *__restrict__ flags)
{
	const long m = cnt[pass];
	if (m > coop_max || (long) blockIdx.x >= m)
		return;
	if (threadIdx.x < F.n)
		s_xi[threadIdx.x] = F.xi[threadIdx.x];
	__syncthreads();
	const bool lead = threadIdx.x == 0;
	for (long i = blockIdx.x; i < m; i += gridDim.x) {
		const long q = list[i];
		int e = el[q], oe = old_el[q], re = 0, j = pst[q];
		double xs[3] = { x[q], y[q], z[q] };
		double tt = t[q];
		if (tracer_iter<COOP>(F, e, oe, xs, tt, j, re)) {
			tracer_finish<COOP>(F, lead, q, e, oe, xs, tt, j, re, el, old_el,
					x, y, z, u, v, w, t, reinit_out, flags);
		} else if (lead) {
			pst[q] = j;
			el[q] = e;
			old_el[q] = oe;
			x[q] = xs[0]; y[q] = xs[1]; z[q] = xs[2];
			t[q] = tt;
			next[atomicAdd(cnt + pass + 1, 1)] = (int) q;
		}
	}
}

// ---------------------------------------------------------------------
// Processing order: a counting sort of the particles by owning element
// and, for the 5-point stencil (n > 5), by the stencil's first plane in z,
// so that the lanes of a wavefront read few distinct rows of field data.
// cnt has nkey + 1 slots (the last for particles naming no element) and
// room for the scan's chunk totals; after the scan it holds each bucket's
// start, and k_porder_place hands out the slots.  The order inside a bucket follows the atomics and may differ
// run to run; no result depends on it.
// ---------------------------------------------------------------------

struct PKey {
	const int *el;
	const double *z;         // reference coordinate z of each particle
	const double *xi;        // GLL points
	int nel, n;
	int sub;                 // stencil starts per element (1: key = el)
};

__device__ __forceinline__ int pkey(const PKey &K, long i)
{
	const int e = K.el[i];
	if (e < 0 || e >= K.nel)
		return K.nel * K.sub;
	if (K.sub == 1)
		return e;
	// first_pt of the z coordinate, from global memory
	const double zp = K.z[i];
	int k1 = K.n - 5;
	for (int a = 0; a < K.n - 4; ++a)
		if (K.xi[a] > zp) {
			k1 = a - 2 > 0 ? a - 2 : 0;
			break;
		}
	return e * K.sub + k1;
}

// a fill kernel rather than a memset node, so the sequence can be captured
__global__ void __launch_bounds__(256)
k_porder_zero(int m, int *__restrict__ cnt)
{
	for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256)
		cnt[i] = 0;
}

// the particles to order: list[0 .. *len) (list null: 0 .. n)
__global__ void __launch_bounds__(256)
k_porder_count(long n, const int *__restrict__ list,
		const int *__restrict__ len, PKey K, int *__restrict__ cnt)
{
	const long m = len ? (long) *len : n;
	for (long i = (long) blockIdx.x * 256 + threadIdx.x; i < m;
			i += (long) gridDim.x * 256)
		atomicAdd(cnt + pkey(K, list ? (long) list[i] : i), 1);
}

// Exclusive prefix sum of cnt[0..m) in place, in chunks of SCAN_CHUNK
// buckets (8 per lane of a 1024-lane workgroup): k_porder_sums writes
// each chunk's total, k_porder_scan_sums scans the totals in one
// workgroup, k_porder_scan_chunks scans each chunk from its offset.
#define SCAN_PER 8
#define SCAN_CHUNK (1024 * SCAN_PER)

// the 1024 lanes' exclusive prefix sums of v, and the total
__device__ __forceinline__ int block_excl_scan(int v, int *sh, int *total)
{
	sh[threadIdx.x] = v;
	__syncthreads();
	for (int d = 1; d < 1024; d <<= 1) {
		const int a = threadIdx.x >= (unsigned) d ? sh[threadIdx.x - d] : 0;
		__syncthreads();
		sh[threadIdx.x] += a;
		__syncthreads();
	}
	const int incl = sh[threadIdx.x];
	*total = sh[1023];
	__syncthreads();
	return incl - v;
}

__global__ void __launch_bounds__(1024)
k_porder_sums(int m, const int *__restrict__ cnt, int *__restrict__ sums)
{
	__shared__ int sh[1024];
	const long b = (long) blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER;
	int v = 0;
#pragma unroll
	for (int k = 0; k < SCAN_PER; ++k)
		if (b + k < m)
			v += cnt[b + k];
	int total;
	block_excl_scan(v, sh, &total);
	if (threadIdx.x == 0)
		sums[blockIdx.x] = total;
}

// exclusive scan of the chunk totals (at most 1024 x 8 of them)
__global__ void __launch_bounds__(1024)
k_porder_scan_sums(int nb, int *__restrict__ sums)
{
	__shared__ int sh[1024];
	const int b = threadIdx.x * SCAN_PER;
	int loc[SCAN_PER], v = 0;
#pragma unroll
	for (int k = 0; k < SCAN_PER; ++k) {
		loc[k] = b + k < nb ? sums[b + k] : 0;
		v += loc[k];
	}
	int total;
	int run = block_excl_scan(v, sh, &total);
#pragma unroll
	for (int k = 0; k < SCAN_PER; ++k)
		if (b + k < nb) {
			sums[b + k] = run;
			run += loc[k];
		}
}

__global__ void __launch_bounds__(1024)
k_porder_scan_chunks(int m, int *__restrict__ cnt,
		const int *__restrict__ sums)
{
	__shared__ int sh[1024];
	const long b = (long) blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER;
	int loc[SCAN_PER], v = 0;
#pragma unroll
	for (int k = 0; k < SCAN_PER; ++k) {
		loc[k] = b + k < m ? cnt[b + k] : 0;
		v += loc[k];
	}
	int total;
	int run = block_excl_scan(v, sh, &total) + sums[blockIdx.x];
#pragma unroll
	for (int k = 0; k < SCAN_PER; ++k)
		if (b + k < m) {
			cnt[b + k] = run;
			run += loc[k];
		}
}

__global__ void __launch_bounds__(256)
k_porder_place(long n, const int *__restrict__ list,
		const int *__restrict__ len, PKey K, int *__restrict__ cnt,
		int *__restrict__ perm)
{
	const long m = len ? (long) *len : n;
	for (long i = (long) blockIdx.x * 256 + threadIdx.x; i < m;
			i += (long) gridDim.x * 256) {
		const long q = list ? (long) list[i] : i;
		perm[atomicAdd(cnt + pkey(K, q), 1)] = (int) q;
	}
}

static PKey make_key(const int *el, const double *z, const double *xi,
		int nel, int n_pts)
{
	PKey K;
	K.el = el;
	K.z = z;
	K.xi = xi;
	K.nel = nel;
	K.n = n_pts;
	K.sub = n_pts > 5 ? n_pts - 4 : 1;
	return K;
}

// perm[0 .. count) = the particles list[0 .. *len) (list null: 0 .. n)
// in key order; cnt: m buckets and the scan's chunk totals
static void order_launch(hipStream_t st, long n, const int *list,
		const int *len, const PKey &K, long m, int *cnt, int *perm)
{
	const long nb = (m + SCAN_CHUNK - 1) / SCAN_CHUNK;
	long b = (n + 255) / 256;
	if (b > 4096)
		b = 4096;
	const long bz = (m + 255) / 256 < 1024 ? (m + 255) / 256 : 1024;
	hipLaunchKernelGGL(k_porder_zero, dim3((unsigned) bz), dim3(256), 0, st,
			(int) m, cnt);
	hipLaunchKernelGGL(k_porder_count, dim3((unsigned) b), dim3(256), 0, st,
			n, list, len, K, cnt);
	int *sums = cnt + m;
	hipLaunchKernelGGL(k_porder_sums, dim3((unsigned) nb), dim3(1024), 0, st,
			(int) m, cnt, sums);
	hipLaunchKernelGGL(k_porder_scan_sums, dim3(1), dim3(1024), 0, st,
			(int) nb, sums);
	hipLaunchKernelGGL(k_porder_scan_chunks, dim3((unsigned) nb), dim3(1024),
			0, st, (int) m, cnt, sums);
	hipLaunchKernelGGL(k_porder_place, dim3((unsigned) b), dim3(256), 0, st,
			n, list, len, K, cnt, perm);
}

// buckets of the processing order for n points per direction (+1: the
// particles naming no element)
extern "C" long ssxk_particle_order_buckets(int nel, int n_pts)
{
	return (long) nel * (n_pts > 5 ? n_pts - 4 : 1) + 1;
}

// ints of scratch the order needs: the buckets, then the chunk totals
extern "C" long ssxk_particle_order_ints(int nel, int n_pts)
{
	const long m = ssxk_particle_order_buckets(nel, n_pts);
	return m + (m + SCAN_CHUNK - 1) / SCAN_CHUNK;
}

extern "C" int ssxk_particle_order(void *stream, long np, const int *el,
		const double *z, const double *xi, int nel, int n_pts, int *cnt,
		int *perm)
{
	if (np <= 0)
		return SSX_OK;
	const long m = ssxk_particle_order_buckets(nel, n_pts);
	const long nb = (m + SCAN_CHUNK - 1) / SCAN_CHUNK;
	if (np > 0x7fffffffL || nb > 1024L * SCAN_PER || n_pts < 2
			|| n_pts > PMAXN) {
		ssx_set_error("particle order: %ld particles, %d elements, n = %d",
				np, nel, n_pts);
		return SSX_EINVAL;
	}
	order_launch((hipStream_t) stream, np, nullptr, nullptr,
			make_key(el, z, xi, nel, n_pts), m, cnt, perm);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) {
		ssx_set_error("particle order kernels: %s", hipGetErrorString(e));
		return SSX_EHIP;
	}
	return SSX_OK;
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

static PField make_field(int n_pts, int nel, const double *xi,
		const double *const ur[9], const double *const jac[9],
		const double *const jinv[9], const int *neigh, double T, double dt)
{
	PField F;
	F.n = n_pts;
	F.nel = nel;
	F.xi = xi;
	for (int l = 0; l < 3; ++l)
		for (int c = 0; c < 3; ++c)
			F.ur[l][c] = ur[l * 3 + c];
	for (int a = 0; a < 9; ++a) {
		F.jac[a] = jac ? jac[a] : nullptr;
		F.jinv[a] = jinv ? jinv[a] : nullptr;
	}
	F.neigh = neigh;
	F.T = T;
	F.dt = dt;
	return F;
}

// the kernel instance for n points: the full product for n <= 5, the
// 5-point stencil above
#define SSX_PDISPATCH(KERNEL, NPTS, ...)                                     \
	do {                                                                     \
		switch (NPTS) {                                                      \
		case 2: hipLaunchKernelGGL((KERNEL<2>), __VA_ARGS__); break;         \
		case 3: hipLaunchKernelGGL((KERNEL<3>), __VA_ARGS__); break;         \
		case 4: hipLaunchKernelGGL((KERNEL<4>), __VA_ARGS__); break;         \
		case 5: hipLaunchKernelGGL((KERNEL<5>), __VA_ARGS__); break;         \
		default: hipLaunchKernelGGL((KERNEL<0>), __VA_ARGS__); break;        \
		}                                                                    \
	} while (0)

// The pass-by-pass walks.  work holds np ints for each of: X, Y (the
// particles entering alternate passes), Z (a pass's phase B), the pass
// state, and W (pass 1's list sorted); cnt holds ca[0 .. SSX_WALK_PASSES] then cm[...].
// Pass k reads its list (pass 0: perm) and writes the next pass's (X for
// even k, Y for odd).  A later pass of at most SSX_COOP_MAX particles runs
// one particle per workgroup (the *_coop kernels, for the 5-point
// stencil), a larger one a particle per lane; each kernel reads the pass's
// count and returns at once when the other form takes it.
#define SSX_COOP_MAX 8192

static long lane_grid(long np)
{
	const long g = (np + 63) / 64;
	return g < 2048 ? g : 2048;
}

#define SSX_PCASE(NN, NAME, ...)                                             \
	case NN: NAME<NN>(__VA_ARGS__); break;
#define SSX_PSWITCH(NPTS, NAME, ...)                                         \
	switch (NPTS) {                                                          \
	SSX_PCASE(2, NAME, __VA_ARGS__) SSX_PCASE(3, NAME, __VA_ARGS__)          \
	SSX_PCASE(4, NAME, __VA_ARGS__) SSX_PCASE(5, NAME, __VA_ARGS__)          \
	default: NAME<0>(__VA_ARGS__); break;                                    \
	}

// Pass 1 takes the particles that changed element in pass 0 sorted again
// (by their new element: bucket, the order's scratch, reused), W the sorted
// list.
template <int NF>
static void particle_passes(hipStream_t st, const PField &F, long np,
		const int *perm, int *bucket, int *work, int *cnt, int *el,
		int *old_el, double *x, double *y, double *z, double *u, double *v,
		double *w, double *t, const double *tau, int *reinit, int *flags)
{
	constexpr int NL = NF == 0 ? -1 : NF;  // later passes: unrolled stencil
	const long cmax = NF == 0 ? SSX_COOP_MAX : 0;
	int *ca = cnt, *cm = cnt + SSX_WALK_PASSES + 1;
	int *X = work, *Y = work + np, *Z = work + 2 * np, *pst = work + 3 * np,
		*W = work + 4 * np;
	hipLaunchKernelGGL(k_porder_zero, dim3(1), dim3(256), 0, st,
			2 * (SSX_WALK_PASSES + 1), cnt);
	for (int pass = 0; pass < SSX_WALK_PASSES; ++pass) {
		const int *in = pass == 0 ? perm : ((pass - 1) & 1) ? Y : X;
		int *out = (pass & 1) ? Y : X;
		if (pass == 1) {
			order_launch(st, np, X, ca + 1, make_key(el, z, F.xi, F.nel, F.n),
					ssxk_particle_order_buckets(F.nel, F.n), bucket, W);
			in = W;
		}
		if (pass == 0)
			hipLaunchKernelGGL(k_particles_a<NF>, dim3((unsigned) ((np + 63)
					/ 64)), dim3(64), 0, st, F, np, 0, 0L, in, Z, ca, cm, pst,
					el, old_el, x, y, z, u, v, w, t, tau, reinit, flags);
		else
			hipLaunchKernelGGL(k_particles_a<NL>, dim3((unsigned) lane_grid(np)),
					dim3(64), 0, st, F, np, pass, cmax, in, Z, ca, cm, pst, el,
					old_el, x, y, z, u, v, w, t, tau, reinit, flags);
		hipLaunchKernelGGL(k_particles_b<NL>, dim3((unsigned) lane_grid(np)),
				dim3(64), 0, st, F, pass, cmax, Z, out, ca, cm, pst, el, old_el,
				x, y, z, u, v, w, t, tau, reinit, flags);
		if (pass > 0 && cmax > 0)
			hipLaunchKernelGGL(k_particles_coop, dim3((unsigned) (np < 2048
					? np : 2048)), dim3(64), 0, st, F, pass, cmax, in, out, ca,
					pst, el, old_el, x, y, z, u, v, w, t, tau, reinit, flags);
	}
}

template <int NF>
static void tracer_passes(hipStream_t st, const PField &F, long np,
		const int *perm, int *work, int *cnt, int *el, int *old_el, double *x,
		double *y, double *z, double *u, double *v, double *w, double *t,
		int *reinit, int *flags)
{
	constexpr int NL = NF == 0 ? -1 : NF;
	const long cmax = NF == 0 ? SSX_COOP_MAX : 0;
	int *X = work, *Y = work + np, *pst = work + 3 * np;
	hipLaunchKernelGGL(k_porder_zero, dim3(1), dim3(256), 0, st,
			2 * (SSX_WALK_PASSES + 1), cnt);
	hipLaunchKernelGGL(k_tracers_pass<NF>, dim3((unsigned) ((np + 63) / 64)),
			dim3(64), 0, st, F, np, 0, 0L, perm, X, cnt, pst, el, old_el, x, y,
			z, u, v, w, t, reinit, flags);
	for (int pass = 1; pass < SSX_WALK_PASSES; ++pass) {
		const int *in = ((pass - 1) & 1) ? Y : X;
		int *out = (pass & 1) ? Y : X;
		hipLaunchKernelGGL(k_tracers_pass<NL>, dim3((unsigned) lane_grid(np)),
				dim3(64), 0, st, F, np, pass, cmax, in, out, cnt, pst, el,
				old_el, x, y, z, u, v, w, t, reinit, flags);
		if (cmax > 0)
			hipLaunchKernelGGL(k_tracers_coop, dim3((unsigned) (np < 2048 ? np
					: 2048)), dim3(64), 0, st, F, pass, cmax, in, out, cnt, pst,
					el, old_el, x, y, z, u, v, w, t, reinit, flags);
	}
}

extern "C" int ssxk_particles_passes(void *stream, int n_pts, int nel,
		const double *xi, const double *const ur[9],
		const double *const jac[9], const double *const jinv[9],
		const int *neigh, double T, double dt, long np, const int *perm,
		int *bucket, int *work, int *cnt, int *el, int *old_el, double *x,
		double *y, double *z, double *u, double *v, double *w, double *t,
		const double *tau, int *reinit, int *flags)
{
	if (np <= 0)
		return SSX_OK;
	if (n_pts < 2 || n_pts > PMAXN || nel <= 0 || np > 0x7fffffffL / 3) {
		ssx_set_error("particle step: %d points per direction, %d elements, "
				"%ld particles", n_pts, nel, np);
		return SSX_EINVAL;
	}
	const PField F = make_field(n_pts, nel, xi, ur, jac, jinv, neigh, T, dt);
	SSX_PSWITCH(n_pts, particle_passes, (hipStream_t) stream, F, np, perm,
			bucket, work, cnt, el, old_el, x, y, z, u, v, w, t, tau, reinit,
			flags)
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) {
		ssx_set_error("particle passes: %s", hipGetErrorString(e));
		return SSX_EHIP;
	}
	return SSX_OK;
}

extern "C" int ssxk_tracers_passes(void *stream, int n_pts, int nel,
		const double *xi, const double *const ur[9], const int *neigh,
		double T, double dt, long np, const int *perm, int *work, int *cnt,
		int *el, int *old_el, double *x, double *y, double *z, double *u,
		double *v, double *w, double *t, int *reinit, int *flags)
{
	if (np <= 0)
		return SSX_OK;
	if (n_pts < 2 || n_pts > PMAXN || nel <= 0 || np > 0x7fffffffL / 3) {
		ssx_set_error("tracer step: %d points per direction, %d elements, "
				"%ld tracers", n_pts, nel, np);
		return SSX_EINVAL;
	}
	const PField F = make_field(n_pts, nel, xi, ur, nullptr, nullptr, neigh,
			T, dt);
	SSX_PSWITCH(n_pts, tracer_passes, (hipStream_t) stream, F, np, perm, work,
			cnt, el, old_el, x, y, z, u, v, w, t, reinit, flags)
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) {
		ssx_set_error("tracer passes: %s", hipGetErrorString(e));
		return SSX_EHIP;
	}
	return SSX_OK;
}

extern "C" int ssxk_tracers_step(void *stream, int n_pts, int nel,
		const double *xi, const double *const ur[9], const int *neigh,
		double T, double dt, long np, const int *perm, int *el, int *old_el,
		double *x, double *y, double *z, double *u, double *v, double *w,
		double *t, int *reinit, int *flags)
{
	if (np <= 0)
		return SSX_OK;
	if (n_pts < 2 || n_pts > PMAXN) {
		ssx_set_error("tracer step: %d points per direction", n_pts);
		return SSX_EINVAL;
	}
	const PField F = make_field(n_pts, nel, xi, ur, nullptr, nullptr, neigh,
			T, dt);
	SSX_PDISPATCH(k_tracers_step, n_pts, dim3((unsigned) ((np + 63) / 64)),
			dim3(64), 0, (hipStream_t) stream, F, np, perm, el, old_el, x, y,
			z, u, v, w, t, reinit, flags);
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
		long np, const int *perm, int *el, int *old_el, double *x, double *y,
		double *z, double *u, double *v, double *w, double *t,
		const double *tau, int *reinit, int *flags)
{
	if (np <= 0)
		return SSX_OK;
	if (n_pts < 2 || n_pts > PMAXN) {
		ssx_set_error("particle step: %d points per direction", n_pts);
		return SSX_EINVAL;
	}
	const PField F = make_field(n_pts, nel, xi, ur, jac, jinv, neigh, T, dt);
	SSX_PDISPATCH(k_particles_step, n_pts,
			dim3((unsigned) ((np + 63) / 64)), dim3(64), 0,
			(hipStream_t) stream, F, np, perm, el, old_el, x, y, z, u, v, w, t,
			tau, reinit, flags);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) {
		ssx_set_error("k_particles_step: %s", hipGetErrorString(e));
		return SSX_EHIP;
	}
	return SSX_OK;
}
