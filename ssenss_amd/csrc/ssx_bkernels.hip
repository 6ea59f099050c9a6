// ssx_bkernels.hip -- one launch per CG stage for a batch of independent
// problems of any element count and order (conjugate_gradient,
// solver.c:33-81, Jacobi preconditioner, lock-step iterations).
//
// Work tables map each workgroup to (problem, local workgroup); every
// problem has its own done flag, scalars and partial sums, so a converged
// problem's workgroups return at once while the others go on.  The
// operator stage is one launch per distinct order (the kernel shape
// depends on n, and one per operator); DSS, update and the two reductions
// are one launch each.  The host drops converged problems from the tables.
// Per-problem partials are summed in a fixed order: results do not depend
// on the other problems of the batch.
#include <hip/hip_runtime.h>
#include "ssx_internal.h"
#include "ssx_device.h"
#include "ssx_ax.h"
#include "ssx_batch.h"

#define LAUNCH_OK(name)                                                    \
	do {                                                               \
		hipError_t e_ = hipGetLastError();                         \
		if (e_ != hipSuccess) {                                    \
			ssx_set_error("%s: %s", name, hipGetErrorString(e_)); \
			return SSX_EHIP;                                   \
		}                                                          \
	} while (0)

// operator (PUPD: with the direction update of the previous iteration,
// and before it the previous iteration's solution step u += alpha p_old,
// which the update kernel leaves to it); one table per (order, operator)
// group
template <int N, int OP, bool PUPD, int V>
__global__ void __launch_bounds__((AxShape<N, V>::BLOCK))
k_bax(const ssx_bprob *__restrict__ pr, const ssx_bwork *__restrict__ tab)
{
	const ssx_bwork w = tab[blockIdx.x];
	const ssx_bprob *d = pr + w.prob;
	if (d->flags[0])
		return;
	const int ic = d->pressure ? CG_NULLC : -1;
	// PUPD reads z and p and writes p; otherwise p is the operand
	const double *u = PUPD ? nullptr : gptr(d->p);
	double *pio = PUPD ? gptr(d->p) : nullptr;
	const double pd = ax_team<N, OP, double, double, PUPD ? 2 : 0, V>(
			w.blk, d->nel,
			u, gptr(d->ap), gptr(d->gf), gptr(d->mass), gptr(d->dmat),
			d->lambda, gptr(d->bc), d->apply_mask, gptr(d->z), pio,
			d->scal, CG_BETA, ic, PUPD ? gptr(d->u) : nullptr, CG_ALPHA);
	ax_elem_partial<N, V>(pd, d->part, w.blk, d->nel);
}

// DSS of Ap; local workgroups number the pairs, then the quadruples, then
// the octuples of the problem
__global__ void __launch_bounds__(256)
k_bdss(const ssx_bprob *__restrict__ pr, const ssx_bwork *__restrict__ tab)
{
	const ssx_bwork w = tab[blockIdx.x];
	const ssx_bprob *d = pr + w.prob;
	if (d->flags[0])
		return;
	const long long b2 = (d->G2 + 255) / 256, b4 = (d->G4 + 255) / 256;
	const long long b = w.blk;
	if (b < b2) {
		const long long g = b * 256 + threadIdx.x;
		if (g < d->G2)
			dss_group<2, double>(d->G2, g, d->g2, d->ap, nullptr, 0);
	} else if (b < b2 + b4) {
		const long long g = (b - b2) * 256 + threadIdx.x;
		if (g < d->G4)
			dss_group<4, double>(d->G4, g, d->g4, d->ap, nullptr, 0);
	} else {
		const long long g = (b - b2 - b4) * 256 + threadIdx.x;
		if (g < d->G8)
			dss_group<8, double>(d->G8, g, d->g8, d->ap, nullptr, 0);
	}
}

// r -= alpha Ap; z = D^-1 r; partials of <r,z>_w, <r,r>_w, <z,1>_w,
// <r,1>_w over this workgroup's chunk.  The solution step u += alpha p is
// left to the next operator launch (or the batch's final flush).  A problem with the gather
// form (gP > 0) had no DSS launch: each copy adds its partners' values to
// its own in its own add order (the group kernel's value bit for bit) and
// its weight is 1/M from the partner count; the lane takes two of its
// nodes at a time and issues their list and partner loads before using
// any.  Otherwise Ap is assembled and the weight is scale.
__global__ void __launch_bounds__(256)
k_bupdate(const ssx_bprob *__restrict__ pr, const ssx_bwork *__restrict__ tab)
{
	const ssx_bwork w = tab[xcd_block()];
	const ssx_bprob *d = pr + w.prob;
	if (d->flags[0])
		return;
	__shared__ double sh[256];
	const double a = d->scal[CG_ALPHA];
	const int gP = d->gP;
	const long long n3 = (long long) d->n * d->n * d->n;
	// the problem's fields in registers: the loop's stores cannot then be
	// taken to alias them
	const long long ntot = d->ntot;
	const int pressure = d->pressure;
	const SSX_GLOBAL int *__restrict__ gsoff =
		(const SSX_GLOBAL int *) d->gsoff;
	const SSX_GLOBAL int *__restrict__ gtab = (const SSX_GLOBAL int *) d->gtab;
	const SSX_GLOBAL double *__restrict__ ap =
		(const SSX_GLOBAL double *) d->ap;
	const SSX_GLOBAL double *__restrict__ dinv =
		(const SSX_GLOBAL double *) d->dinv;
	const SSX_GLOBAL double *__restrict__ scale =
		(const SSX_GLOBAL double *) d->scale;
	SSX_GLOBAL double *__restrict__ r = (SSX_GLOBAL double *) d->r;
	SSX_GLOBAL double *__restrict__ z = (SSX_GLOBAL double *) d->z;
	double acc[4] = { 0.0, 0.0, 0.0, 0.0 };
	const long long q0 = (long long) w.blk * SSX_BATCH_CHUNK;
	for (int t = 0; t < SSX_BATCH_CHUNK / 256; t += 2) {
		long long qq[2];
		double s[2], wt[2], pv[2][7];
		int ix[2][7];
#pragma unroll
		for (int v = 0; v < 2; ++v) {
			const long long q = q0 + (t + v) * 256 + threadIdx.x;
			qq[v] = q;
			s[v] = 0.0;
			wt[v] = 1.0;
#pragma unroll
			for (int k = 0; k < 7; ++k)
				ix[v][k] = -1;
			if (q >= ntot)
				continue;
			s[v] = ap[q];
			if (gP > 0) {
				const long long e = q / n3;
				const int code = gsoff[q - e * n3];
				const SSX_GLOBAL int *pt = gtab + e * gP + (code >> 3);
#pragma unroll
				for (int k = 0; k < 7; ++k)
					ix[v][k] = k < (code & 7) ? pt[k] : -1;
			} else {
				wt[v] = scale[q];
			}
		}
#pragma unroll
		for (int v = 0; v < 2; ++v)
#pragma unroll
			for (int k = 0; k < 7; ++k)
				pv[v][k] = ix[v][k] >= 0 ? ap[ix[v][k]] : 0.0;
#pragma unroll
		for (int v = 0; v < 2; ++v) {
			const long long q = qq[v];
			if (q >= ntot)
				break;
			double sv = s[v];
			if (gP > 0) {
				int m = 1;
				// own + partners left to right; unused entries are -1 and
				// come last
#pragma unroll
				for (int k = 0; k < 7; ++k)
					if (ix[v][k] >= 0) {
						sv = sv + pv[v][k];
						++m;
					}
				wt[v] = 1.0 / m;
			}
			const double rv = -a * sv + r[q];
			r[q] = rv;
			const double zv = rv * dinv[q];
			z[q] = zv;
			const double ws = wt[v];
			acc[0] += (rv * ws) * zv;
			acc[1] += (rv * ws) * rv;
			if (pressure) {
				acc[2] += zv * ws;
				acc[3] += rv * ws;
			}
		}
	}
	double *base = gptr(d->part) + d->nel + d->nbd;
	for (int c = 0; c < 4; ++c) {
		sh[threadIdx.x] = acc[c];
		__syncthreads();
		for (int st = 128; st > 0; st >>= 1) {
			if ((int) threadIdx.x < st)
				sh[threadIdx.x] += sh[threadIdx.x + st];
			__syncthreads();
		}
		if (threadIdx.x == 0)
			base[(long long) c * d->nbu + w.blk] = sh[0];
		__syncthreads();
	}
}

// one workgroup per problem.  mode 0: <p,Ap>_w (the operator partials)
// -> alpha.  mode 1: the
// null-space constant, rdz, beta, the iteration count and the stopping
// test sqrt(<r,r>_w) <= eps or count >= maxit (or NaN), as k_cg_reduce
__global__ void __launch_bounds__(1024)
k_breduce(int mode, const ssx_bprob *__restrict__ pr,
		const int *__restrict__ list)
{
	const ssx_bprob *d = pr + (list ? list[blockIdx.x] : (int) blockIdx.x);
	if (d->flags[0])
		return;
	__shared__ double sh[4][1024];
	const int nq = mode == 0 ? 1 : (d->pressure ? 4 : 2);
	const double *base = gptr(d->part) + (mode == 0 ? 0 : d->nel + d->nbd);
	const long long m = mode == 0 ? d->nel : d->nbu;
	// the lane's partials are added in index order but loaded 8 at a time
	// (the first of every channel together), so their latencies overlap
	const long long t = threadIdx.x;
	double first[4];
#pragma unroll
	for (int c = 0; c < 4; ++c) {
		// branch-free: an unused slot loads base[0] and discards it
		const bool use = c < nq && t < m;
		const double v = base[use ? (long long) c * m + t : 0];
		first[c] = use ? v : 0.0;
	}
	for (int c = 0; c < 4; ++c) {
		double acc = 0.0;
		if (c < nq && t < m) {
			const double *pp = base + (long long) c * m;
			acc += first[c];
			// 8 loads in flight; past m a slot loads pp[0] unused
			for (long long q = t + 1024; q < m; q += 8 * 1024) {
				double v[8];
#pragma unroll
				for (int i = 0; i < 8; ++i) {
					const long long qi = q + i * 1024;
					v[i] = pp[qi < m ? qi : 0];
				}
#pragma unroll
				for (int i = 0; i < 8; ++i)
					if (q + i * 1024 < m)
						acc += v[i];
			}
		}
		sh[c][threadIdx.x] = acc;
	}
	__syncthreads();
	for (int st = 512; st > 0; st >>= 1) {
		if ((int) threadIdx.x < st)
			for (int c = 0; c < 4; ++c)
				sh[c][threadIdx.x] += sh[c][threadIdx.x + st];
		__syncthreads();
	}
	if (threadIdx.x != 0)
		return;
	double *s = d->scal;
	if (mode == 0) {
		s[CG_PAP] = sh[0][0];
		s[CG_ALPHA] = s[CG_RDZ] / sh[0][0];
		return;
	}
	double rz = sh[0][0], rr = sh[1][0], c = 0.0;
	if (d->pressure) {
		c = sh[2][0] / d->nunique;
		rz = rz - c * sh[3][0];
	}
	s[CG_NULLC] = c;
	s[CG_BETA] = rz / s[CG_RDZ];
	s[CG_RDZ] = rz;
	s[CG_RR] = rr;
	const int it = d->flags[1] + 1;
	d->flags[1] = it;
	const double tol = d->eps >= 0.0 ? d->eps : s[CG_TOL];
	if (sqrt(rr) <= tol || it >= d->maxit || !(rr == rr))
		d->flags[0] = 1;
}

// start of a pass: flags cleared and the scalars that the first fused
// direction update and solution step read set to zero
__global__ void k_btask_reset(ssx_bprob *__restrict__ pr,
		const int *__restrict__ list, int nt)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= nt)
		return;
	const ssx_bprob *d = pr + list[t];
	d->flags[0] = 0;
	d->flags[1] = 0;
	d->scal[CG_ALPHA] = 0.0;
	d->scal[CG_BETA] = 0.0;
	d->scal[CG_NULLC] = 0.0;
}

// u += alpha p over one update chunk of an FP64 task (the step the update
// leaves to the next operator launch); the arithmetic of that launch
__global__ void __launch_bounds__(256)
k_bflush(const ssx_bprob *__restrict__ pr, const ssx_bwork *__restrict__ tab)
{
	const ssx_bwork w = tab[blockIdx.x];
	const ssx_bprob *d = pr + w.prob;
	const double a = d->scal[CG_ALPHA];
	const long long ntot = d->ntot;
	const SSX_GLOBAL double *__restrict__ p =
		(const SSX_GLOBAL double *) d->p;
	SSX_GLOBAL double *__restrict__ u = (SSX_GLOBAL double *) d->u;
	const long long q0 = (long long) w.blk * SSX_BATCH_CHUNK;
	for (int t = 0; t < SSX_BATCH_CHUNK / 256; ++t) {
		const long long q = q0 + t * 256 + threadIdx.x;
		if (q < ntot)
			u[q] = a * p[q] + u[q];
	}
}

template <int N, int OP, bool PU, int V>
static void launch_bax(hipStream_t st, const ssx_bprob *pr,
		const ssx_bwork *tab, long long ntab)
{
	hipLaunchKernelGGL((k_bax<N, OP, PU, V>), dim3((unsigned) ntab),
			dim3(AxShape<N, V>::BLOCK), 0, st, pr, tab);
}

template <int N, int V>
static void launch_bax_nv(hipStream_t st, int op, int pupd,
		const ssx_bprob *pr, const ssx_bwork *tab, long long ntab)
{
	if (op == SSX_OP_HELMHOLTZ) {
		if (pupd)
			launch_bax<N, 1, true, V>(st, pr, tab, ntab);
		else
			launch_bax<N, 1, false, V>(st, pr, tab, ntab);
	} else {
		if (pupd)
			launch_bax<N, 0, true, V>(st, pr, tab, ntab);
		else
			launch_bax<N, 0, false, V>(st, pr, tab, ntab);
	}
}

extern "C" int ssxk_bax(void *stream, int n, int op, int v, int pupd,
		const ssx_bprob *pr, const ssx_bwork *tab, long long ntab)
{
	if (ntab <= 0)
		return SSX_OK;
	hipStream_t st = (hipStream_t) stream;
	switch (n * SSX_AX_NVARIANTS + v) {
#define CASEV(NN, VV) case NN * SSX_AX_NVARIANTS + VV: \
		launch_bax_nv<NN, VV>(st, op, pupd, pr, tab, ntab); break;
#define CASE(NN) CASEV(NN, 0) CASEV(NN, 1) CASEV(NN, 2) CASEV(NN, 3) \
		CASEV(NN, 4) CASEV(NN, 5) CASEV(NN, 6) CASEV(NN, 7)
	CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9)
	CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
#undef CASEV
	default:
		ssx_set_error("batched operator: n=%d variant %d unsupported", n, v);
		return SSX_EINVAL;
	}
	LAUNCH_OK("k_bax");
	return SSX_OK;
}

extern "C" int ssxk_bax_teams(int n, int v)
{
	return ssx_ax_teams(n, v);
}

extern "C" int ssxk_bdss(void *stream, const ssx_bprob *pr,
		const ssx_bwork *tab, long long ntab)
{
	if (ntab <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_bdss, dim3((unsigned) ntab), dim3(256), 0,
			(hipStream_t) stream, pr, tab);
	LAUNCH_OK("k_bdss");
	return SSX_OK;
}

extern "C" int ssxk_bupdate(void *stream, const ssx_bprob *pr,
		const ssx_bwork *tab, long long ntab)
{
	if (ntab <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_bupdate, dim3((unsigned) ntab), dim3(256), 0,
			(hipStream_t) stream, pr, tab);
	LAUNCH_OK("k_bupdate");
	return SSX_OK;
}

extern "C" int ssxk_breduce(void *stream, int mode, const ssx_bprob *pr,
		int np)
{
	hipLaunchKernelGGL(k_breduce, dim3(np), dim3(1024), 0,
			(hipStream_t) stream, mode, pr, (const int *) nullptr);
	LAUNCH_OK("k_breduce");
	return SSX_OK;
}

extern "C" int ssxk_breduce_list(void *stream, int mode, const ssx_bprob *pr,
		const int *list, int nt)
{
	if (nt <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_breduce, dim3(nt), dim3(1024), 0,
			(hipStream_t) stream, mode, pr, list);
	LAUNCH_OK("k_breduce");
	return SSX_OK;
}

extern "C" int ssxk_btask_reset(void *stream, const ssx_bprob *pr,
		const int *list, int nt)
{
	if (nt <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_btask_reset, dim3((nt + 255) / 256), dim3(256), 0,
			(hipStream_t) stream, (ssx_bprob *) pr, list, nt);
	LAUNCH_OK("k_btask_reset");
	return SSX_OK;
}

extern "C" int ssxk_bflush(void *stream, const ssx_bprob *pr,
		const ssx_bwork *tab, long long ntab)
{
	if (ntab <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_bflush, dim3((unsigned) ntab), dim3(256), 0,
			(hipStream_t) stream, pr, tab);
	LAUNCH_OK("k_bflush");
	return SSX_OK;
}
