// ssx_bkernels_lp.hip -- the batched CG stages of ssx_bkernels.hip for
// reduced-precision tasks: the p-multigrid smoother of a batch of problems
// in FP32 fields (geometric factors FP32 or BF16), inner products
// accumulated in FP64 (the single-problem FP32 smoother's arithmetic).
//
// The operator always applies the fused direction update p = (z - c) +
// beta p (a pass starts with beta = c = 0 and p = z, so the first
// iteration leaves p = z); the solution step u += alpha p is taken in the
// update, as the single-problem FP32 smoother takes it.
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

template <int N, int OP, typename G, int V>
__global__ void __launch_bounds__((AxShape<N, V>::BLOCK))
k_bax_lp(const ssx_bprob *__restrict__ pr, const ssx_bwork *__restrict__ tab)
{
	const ssx_bwork w = tab[blockIdx.x];
	const ssx_bprob *d = pr + w.prob;
	if (d->flags[0])
		return;
	const int ic = d->pressure ? CG_NULLC : -1;
	const double pd = ax_team<N, OP, float, G, 1, V>(w.blk, d->nel,
			nullptr, gptr(d->ap_f), gptr((const G *) d->gf_lp),
			gptr(d->mass_f), gptr(d->dmat_f), (float) d->lambda,
			gptr(d->bc), d->apply_mask, gptr(d->z_f), gptr(d->p_f),
			d->scal, CG_BETA, ic, nullptr, CG_ALPHA);
	ax_elem_partial<N, V>(pd, d->part, w.blk, d->nel);
}

// DSS of Ap (tasks without partner lists); local workgroups number the
// pairs, then the quadruples, then the octuples
__global__ void __launch_bounds__(256)
k_bdss_lp(const ssx_bprob *__restrict__ pr, const ssx_bwork *__restrict__ tab)
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
			dss_group<2, float>(d->G2, g, d->g2, d->ap_f, nullptr, 0);
	} else if (b < b2 + b4) {
		const long long g = (b - b2) * 256 + threadIdx.x;
		if (g < d->G4)
			dss_group<4, float>(d->G4, g, d->g4, d->ap_f, nullptr, 0);
	} else {
		const long long g = (b - b2 - b4) * 256 + threadIdx.x;
		if (g < d->G8)
			dss_group<8, float>(d->G8, g, d->g8, d->ap_f, nullptr, 0);
	}
}

// u += alpha p; r -= alpha Ap; z = D^-1 r; FP64 partials of <r,z>_w,
// <r,r>_w, <z,1>_w, <r,1>_w over this workgroup's chunk.  With partner
// lists Ap is assembled here (own + partners in the copy's add order,
// weight 1/M from the count); otherwise Ap was assembled by k_bdss_lp and
// the weight is scale.
__global__ void __launch_bounds__(256)
k_bupdate_lp(const ssx_bprob *__restrict__ pr,
		const ssx_bwork *__restrict__ tab)
{
	const ssx_bwork w = tab[xcd_block()];
	const ssx_bprob *d = pr + w.prob;
	if (d->flags[0])
		return;
	__shared__ double sh[256];
	const float a = (float) d->scal[CG_ALPHA];
	const int gP = d->gP;
	const long long n3 = (long long) d->n * d->n * d->n;
	const long long ntot = d->ntot;
	const int pressure = d->pressure;
	const SSX_GLOBAL int *__restrict__ gsoff =
		(const SSX_GLOBAL int *) d->gsoff;
	const SSX_GLOBAL int *__restrict__ gtab = (const SSX_GLOBAL int *) d->gtab;
	const SSX_GLOBAL float *__restrict__ ap =
		(const SSX_GLOBAL float *) d->ap_f;
	const SSX_GLOBAL float *__restrict__ dinv =
		(const SSX_GLOBAL float *) d->dinv_f;
	const SSX_GLOBAL float *__restrict__ scale =
		(const SSX_GLOBAL float *) d->scale_f;
	const SSX_GLOBAL float *__restrict__ p = (const SSX_GLOBAL float *) d->p_f;
	SSX_GLOBAL float *__restrict__ u = (SSX_GLOBAL float *) d->u_f;
	SSX_GLOBAL float *__restrict__ r = (SSX_GLOBAL float *) d->r_f;
	SSX_GLOBAL float *__restrict__ z = (SSX_GLOBAL float *) d->z_f;
	double acc[4] = { 0.0, 0.0, 0.0, 0.0 };
	const long long q0 = (long long) w.blk * SSX_BATCH_CHUNK;
	for (int t = 0; t < SSX_BATCH_CHUNK / 256; t += 2) {
		long long qq[2];
		float s[2], wt[2], pv[2][7], rq[2], dq[2], pq[2], uq[2];
		int ix[2][7];
#pragma unroll
		for (int v = 0; v < 2; ++v) {
			const long long q = q0 + (t + v) * 256 + threadIdx.x;
			qq[v] = q;
			s[v] = 0.0f;
			wt[v] = 1.0f;
			rq[v] = dq[v] = pq[v] = uq[v] = 0.0f;
#pragma unroll
			for (int k = 0; k < 7; ++k)
				ix[v][k] = -1;
			if (q >= ntot)
				continue;
			s[v] = ap[q];
			rq[v] = r[q];
			dq[v] = dinv[q];
			pq[v] = p[q];
			uq[v] = u[q];
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
				pv[v][k] = ix[v][k] >= 0 ? ap[ix[v][k]] : 0.0f;
#pragma unroll
		for (int v = 0; v < 2; ++v) {
			const long long q = qq[v];
			if (q >= ntot)
				break;
			float sv = s[v];
			if (gP > 0) {
				int m = 1;
#pragma unroll
				for (int k = 0; k < 7; ++k)
					if (ix[v][k] >= 0) {
						sv = sv + pv[v][k];
						++m;
					}
				wt[v] = 1.0f / (float) m;
			}
			u[q] = a * pq[v] + uq[v];
			const float rv = -a * sv + rq[v];
			r[q] = rv;
			const float zv = rv * dq[v];
			z[q] = zv;
			const float ws = wt[v];
			acc[0] += (double) (rv * ws) * (double) zv;
			acc[1] += (double) (rv * ws) * (double) rv;
			if (pressure) {
				acc[2] += (double) (zv * ws);
				acc[3] += (double) (rv * ws);
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

template <int N, typename G, int V>
static void launch_bax_lp(hipStream_t st, int op, const ssx_bprob *pr,
		const ssx_bwork *tab, long long ntab)
{
	if (op == SSX_OP_HELMHOLTZ)
		hipLaunchKernelGGL((k_bax_lp<N, 1, G, V>), dim3((unsigned) ntab),
				dim3(AxShape<N, V>::BLOCK), 0, st, pr, tab);
	else
		hipLaunchKernelGGL((k_bax_lp<N, 0, G, V>), dim3((unsigned) ntab),
				dim3(AxShape<N, V>::BLOCK), 0, st, pr, tab);
}

template <typename G>
static int bax_lp_dispatch(hipStream_t st, int n, int op, int v,
		const ssx_bprob *pr, const ssx_bwork *tab, long long ntab)
{
	switch (n * SSX_AX_NVARIANTS + v) {
#define CASEV(NN, VV) case NN * SSX_AX_NVARIANTS + VV: \
		launch_bax_lp<NN, G, VV>(st, op, pr, tab, ntab); break;
#define CASE(NN) CASEV(NN, 0) CASEV(NN, 1) CASEV(NN, 2) CASEV(NN, 3) \
		CASEV(NN, 4) CASEV(NN, 5) CASEV(NN, 6) CASEV(NN, 7)
	CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9)
	CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
#undef CASEV
	default:
		ssx_set_error("batched FP32 operator: n=%d variant %d unsupported",
				n, v);
		return SSX_EINVAL;
	}
	LAUNCH_OK("k_bax_lp");
	return SSX_OK;
}

extern "C" int ssxk_bax_lp(void *stream, int prec, int n, int op, int v,
		const ssx_bprob *pr, const ssx_bwork *tab, long long ntab)
{
	if (ntab <= 0)
		return SSX_OK;
	hipStream_t st = (hipStream_t) stream;
	if (prec == SSX_PREC_BF16)
		return bax_lp_dispatch<bf16_t>(st, n, op, v, pr, tab, ntab);
	if (prec == SSX_PREC_FP32)
		return bax_lp_dispatch<float>(st, n, op, v, pr, tab, ntab);
	ssx_set_error("batched operator: precision %d is not reduced", prec);
	return SSX_EINVAL;
}

extern "C" int ssxk_bdss_lp(void *stream, const ssx_bprob *pr,
		const ssx_bwork *tab, long long ntab)
{
	if (ntab <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_bdss_lp, dim3((unsigned) ntab), dim3(256), 0,
			(hipStream_t) stream, pr, tab);
	LAUNCH_OK("k_bdss_lp");
	return SSX_OK;
}

extern "C" int ssxk_bupdate_lp(void *stream, const ssx_bprob *pr,
		const ssx_bwork *tab, long long ntab)
{
	if (ntab <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_bupdate_lp, dim3((unsigned) ntab), dim3(256), 0,
			(hipStream_t) stream, pr, tab);
	LAUNCH_OK("k_bupdate_lp");
	return SSX_OK;
}
