// ssx_kernels.hip -- gfx950 kernels of the ssx elliptic-solve engine.
//
// Layout: E-vectors u[e][k][j][i] (i fastest).  Geometric factors are
// premultiplied G~_ab = G_ab * w * detJ, stored per element as six
// contiguous n^3 blocks [e][xx,xy,xz,yy,yz,zz][node] so that every factor
// streams from HBM in fully coalesced runs.
//
// Precisions (template parameters T = field type, G = factor type):
//   double/double  the FP64 path (outer solver, coarsest level, parity)
//   float/float    FP32 multigrid smoother and order transfer
//   float/bf16     FP32 fields with bfloat16 geometric factors
// Inner products always accumulate in FP64.
//
// Kernels:
//   k_ax            the element operator: template in ssx_ax.h, launchers
//                   here (one instantiation file per precision)
//   k_dss_all<T>    direct stiffness summation over shared-node groups of
//                   M = 2, 4, 8 copies; each copy becomes own + partners in
//                   the reference's add order (bit-identical result).
//   k_cg_*          fused CG iteration with device-resident scalars and a
//                   device-side stopping test.
//   k_restrict / k_prolong   tensor-product order transfer per element.
//   k_dot / k_reduce         deterministic two-pass weighted inner product.
#include <hip/hip_runtime.h>
#include "ssx_internal.h"
#include "ssx_device.h"

#define LAUNCH_CHECK(name)                                                   \
	do {                                                                 \
		hipError_t e_ = hipGetLastError();                           \
		if (e_ != hipSuccess) {                                      \
			ssx_set_error("%s: %s", name, hipGetErrorString(e_)); \
			return SSX_EHIP;                                     \
		}                                                            \
	} while (0)

static inline dim3 vgrid(size_t n)
{
	size_t b = (n + 255) / 256;
	if (b > 16384)
		b = 16384;
	if (b == 0)
		b = 1;
	return dim3((unsigned) b);
}

#ifndef SSX_GATHER_U
#define SSX_GATHER_U 2
#endif

#define GRID_STRIDE(q, n) \
	for (size_t q = (size_t) blockIdx.x * blockDim.x + threadIdx.x; q < (n); \
			q += (size_t) gridDim.x * blockDim.x)

// ---------------------------------------------------------------------
// Operator launchers (kernels in ssx_ax.h, one file per precision)
// ---------------------------------------------------------------------

#include "ssx_ax.h"

extern "C" int ssxk_ax_dispatch_f64(void *stream, int n, int v, int nel,
		int op, const void *u, void *au, const void *gf, const void *mass,
		const void *dmat, double lambda, const int *bc, int apply_mask,
		const int *done, double *part, const void *pu);
extern "C" int ssxk_ax_dispatch_f32(void *stream, int n, int v, int nel,
		int op, const void *u, void *au, const void *gf, const void *mass,
		const void *dmat, double lambda, const int *bc, int apply_mask,
		const int *done, double *part, const void *pu);
extern "C" int ssxk_ax_dispatch_bf16(void *stream, int n, int v, int nel,
		int op, const void *u, void *au, const void *gf, const void *mass,
		const void *dmat, double lambda, const int *bc, int apply_mask,
		const int *done, double *part, const void *pu);

extern "C" int ssxk_ax(void *stream, int n, int v, int nel, int op,
		const double *u, double *au, const double *gf, const double *mass,
		const double *dmat, double lambda, const int *bc, int apply_mask,
		const int *done, double *part)
{
	return ssxk_ax_dispatch_f64(stream, n, v, nel, op, u, au, gf, mass, dmat,
			lambda, bc, apply_mask, done, part, nullptr);
}

extern "C" int ssxk_ax_pupd(void *stream, int n, int v, int nel, int op,
		const double *z, double *p, const double *scal, int ib, int ic,
		double *ap, const double *gf, const double *mass, const double *dmat,
		double lambda, const int *bc, int apply_mask, const int *done,
		double *part, double *u, int ia)
{
	const PUpd pu = { z, p, scal, ib, ic, u, ia };
	return ssxk_ax_dispatch_f64(stream, n, v, nel, op, nullptr, ap, gf, mass,
			dmat, lambda, bc, apply_mask, done, part, &pu);
}

extern "C" int ssxk_ax_lp(void *stream, int prec, int n, int v, int nel,
		int op, const void *u, void *au, const void *gf, const void *mass,
		const void *dmat, double lambda, const int *bc, int apply_mask,
		const int *done, double *part)
{
	if (prec == SSX_PREC_BF16)
		return ssxk_ax_dispatch_bf16(stream, n, v, nel, op, u, au, gf, mass,
				dmat, lambda, bc, apply_mask, done, part, nullptr);
	return ssxk_ax_dispatch_f32(stream, n, v, nel, op, u, au, gf, mass, dmat,
			lambda, bc, apply_mask, done, part, nullptr);
}

extern "C" int ssxk_ax_pupd_lp(void *stream, int prec, int n, int v, int nel,
		int op, const void *z, void *p, const double *scal, int ib, int ic,
		void *ap, const void *gf, const void *mass, const void *dmat,
		double lambda, const int *bc, int apply_mask, const int *done,
		double *part, void *u, int ia)
{
	if (u) {
		ssx_set_error("ssxk_ax_pupd_lp: the deferred solution step is FP64 "
				"only");
		return SSX_EINVAL;
	}
	const PUpd pu = { z, p, scal, ib, ic, u, ia };
	if (prec == SSX_PREC_BF16)
		return ssxk_ax_dispatch_bf16(stream, n, v, nel, op, nullptr, ap, gf,
				mass, dmat, lambda, bc, apply_mask, done, part, &pu);
	return ssxk_ax_dispatch_f32(stream, n, v, nel, op, nullptr, ap, gf, mass,
			dmat, lambda, bc, apply_mask, done, part, &pu);
}

// ---------------------------------------------------------------------
// CG scalar logic shared by k_cg_reduce and the last-workgroup reductions
// ---------------------------------------------------------------------

// mode 0: s[o+PAP] = pap; s[o+ALPHA] = s[o+RDZ] / pap
__device__ __forceinline__ void cg_scalars0(double pap, double *scal, int o)
{
	scal[o + CG_PAP] = pap;
	scal[o + CG_ALPHA] = scal[o + CG_RDZ] / pap;
}

// mode 1: null-space constant c = zs/nunique; rdz = rz - c rs;
// beta = rdz / rdz_old; iteration count ++, done when sqrt(rr) <= eps
// (eps < 0: s[o+TOL]) or count >= maxit or NaN
__device__ __forceinline__ void cg_scalars1(double rz, double rr, double zs,
		double rs, double *scal, int o, int *flags, double eps, int maxit,
		double nunique, int pressure)
{
	double c = 0.0;
	if (pressure) {
		c = zs / nunique;
		rz = rz - c * rs;
	}
	scal[o + CG_NULLC] = c;
	scal[o + CG_BETA] = rz / scal[o + CG_RDZ];
	scal[o + CG_RDZ] = rz;
	scal[o + CG_RR] = rr;
	const int it = flags[1] + 1;
	flags[1] = it;
	const double tol = eps >= 0.0 ? eps : scal[o + CG_TOL];
	if (sqrt(rr) <= tol || it >= maxit || !(rr == rr))
		flags[0] = 1;
}

// In a 256-lane workgroup, the sum of pp[0..m) in exactly the order of a
// 1024-lane k_cg_reduce: lane t holds the accumulators of virtual lanes
// t, t+256, t+512, t+768 and folds them as the first two tree levels do.
__device__ double sum_as_1024(const double *pp, long long m, double *sh)
{
	double a[4];
#pragma unroll
	for (int v = 0; v < 4; ++v) {
		double acc = 0.0;
		for (long long q = threadIdx.x + 256 * v; q < m; q += 1024)
			acc += __hip_atomic_load(pp + q, __ATOMIC_RELAXED,
					__HIP_MEMORY_SCOPE_AGENT);
		a[v] = acc;
	}
	a[0] += a[2];
	a[1] += a[3];
	a[0] += a[1];
	sh[threadIdx.x] = a[0];
	__syncthreads();
	for (int st = 128; st > 0; st >>= 1) {
		if ((int) threadIdx.x < st)
			sh[threadIdx.x] += sh[threadIdx.x + st];
		__syncthreads();
	}
	const double t = sh[0];
	__syncthreads();
	return t;
}

// After thread 0 stored this workgroup's partials: true in every lane of
// the workgroup that arrives last (its partials and everyone else's are
// then visible to it).  The last one resets the counter.
__device__ bool last_workgroup(unsigned *counter)
{
	__shared__ int s_last;
	__syncthreads();
	if (threadIdx.x == 0) {
		__threadfence();
		const unsigned prev = atomicAdd(counter, 1u);
		s_last = prev == gridDim.x - 1;
		if (s_last) {
			atomicExch(counter, 0u);
			__threadfence();
		}
	}
	__syncthreads();
	return s_last != 0;
}

// ---------------------------------------------------------------------
// Direct stiffness summation over shared-node groups
// ---------------------------------------------------------------------

// One launch for all three multiplicities: blocks [0,b2) take the pairs,
// [b2,b2+b4) the quadruples, the rest the octuples.
template <typename T, bool RED>
__global__ void __launch_bounds__(256)
k_dss_all(long long G2, long long G4, long long G8, const int *__restrict__ i2,
		const int *__restrict__ i4, const int *__restrict__ i8,
		T *__restrict__ u, const T *__restrict__ p,
		double *__restrict__ part, int average, const int *__restrict__ done,
		ssx_lastred lr)
{
	if (done && *done)
		return;
	const long long b2 = (G2 + 255) / 256, b4 = (G4 + 255) / 256;
	long long b = blockIdx.x;
	double pd = 0.0;
	if (b < b2) {
		long long g = b * 256 + threadIdx.x;
		if (g < G2)
			pd = dss_group<2, T>(G2, g, i2, u, p, average);
	} else if (b < b2 + b4) {
		long long g = (b - b2) * 256 + threadIdx.x;
		if (g < G4)
			pd = dss_group<4, T>(G4, g, i4, u, p, average);
	} else {
		long long g = (b - b2 - b4) * 256 + threadIdx.x;
		if (g < G8)
			pd = dss_group<8, T>(G8, g, i8, u, p, average);
	}
	if (part) {
		__shared__ double sh[256];
		sh[threadIdx.x] = pd;
		__syncthreads();
#pragma unroll
		for (int st = 128; st > 0; st >>= 1) {
			if ((int) threadIdx.x < st)
				sh[threadIdx.x] += sh[threadIdx.x + st];
			__syncthreads();
		}
		if (threadIdx.x == 0)
			part[blockIdx.x] = sh[0];
	}
	if (RED && last_workgroup(lr.counter)) {
		__shared__ double sr[256];
		const double pap = sum_as_1024(lr.part, lr.cnt, sr);
		if (threadIdx.x == 0)
			cg_scalars0(pap, lr.scal, lr.o);
	}
}

extern "C" long long ssxk_dss_blocks(const long long ng[3])
{
	return (ng[0] + 255) / 256 + (ng[1] + 255) / 256 + (ng[2] + 255) / 256;
}

template <typename T>
static int launch_dss(hipStream_t st, int *const grp[3], const long long ng[3],
		void *u, const void *p, double *part, int average, const int *done,
		const ssx_lastred *lr = nullptr)
{
	long long nb = ssxk_dss_blocks(ng);
	if (nb == 0)
		return SSX_OK;
	if (lr)
		hipLaunchKernelGGL((k_dss_all<T, true>), dim3((unsigned) nb),
				dim3(256), 0, st, ng[0], ng[1], ng[2], grp[0], grp[1], grp[2],
				(T *) u, (const T *) p, part, average, done, *lr);
	else
		hipLaunchKernelGGL((k_dss_all<T, false>), dim3((unsigned) nb),
				dim3(256), 0, st, ng[0], ng[1], ng[2], grp[0], grp[1], grp[2],
				(T *) u, (const T *) p, part, average, done, ssx_lastred{});
	LAUNCH_CHECK("k_dss_all");
	return SSX_OK;
}

extern "C" int ssxk_dss_all_red(void *stream, int *const grp[3],
		const long long ng[3], double *u, const double *p, double *part,
		const int *done, const ssx_lastred *lr)
{
	return launch_dss<double>((hipStream_t) stream, grp, ng, u, p, part, 0,
			done, lr);
}

extern "C" int ssxk_dss_all_red_lp(void *stream, int *const grp[3],
		const long long ng[3], void *u, const void *p, double *part,
		const int *done, const ssx_lastred *lr)
{
	return launch_dss<float>((hipStream_t) stream, grp, ng, u, p, part, 0,
			done, lr);
}

extern "C" int ssxk_dss_all(void *stream, int *const grp[3],
		const long long ng[3], double *u, const double *p, double *part,
		int average, const int *done)
{
	return launch_dss<double>((hipStream_t) stream, grp, ng, u, p, part,
			average, done);
}

extern "C" int ssxk_dss_all_lp(void *stream, int *const grp[3],
		const long long ng[3], void *u, const void *p, double *part,
		int average, const int *done)
{
	return launch_dss<float>((hipStream_t) stream, grp, ng, u, p, part,
			average, done);
}

// ---------------------------------------------------------------------
// Fused CG iteration kernels (solver.c:58-74)
// ---------------------------------------------------------------------

__device__ __forceinline__ void block_sum4(double *v, double (*sh)[256])
{
#pragma unroll
	for (int c = 0; c < 4; ++c)
		sh[c][threadIdx.x] = v[c];
	__syncthreads();
#pragma unroll
	for (int st = 128; st >= 64; st >>= 1) {
		if ((int) threadIdx.x < st) {
#pragma unroll
			for (int c = 0; c < 4; ++c)
				sh[c][threadIdx.x] += sh[c][threadIdx.x + st];
		}
		__syncthreads();
	}
	// steps below 64 in wavefront 0 by shuffles (the same additions in the
	// same order); the totals go back to sh[c][0]
	if (threadIdx.x < 64) {
		double w[4];
#pragma unroll
		for (int c = 0; c < 4; ++c)
			w[c] = sh[c][threadIdx.x];
#pragma unroll
		for (int st = 32; st > 0; st >>= 1)
#pragma unroll
			for (int c = 0; c < 4; ++c)
				w[c] += __shfl_down(w[c], st, 64);
		if (threadIdx.x == 0)
#pragma unroll
			for (int c = 0; c < 4; ++c)
				sh[c][0] = w[c];
	}
	__syncthreads();
}

// u += a p; r -= a ap; z = dinv r; partial <r,z>_w, <r,r>_w and, for the
// pressure null space, <z,1>_w and <r,1>_w (FP64 accumulation).  UPD
// false (the launchers pick it for u null): the solution step is taken
// elsewhere (by the next operator launch and a final k_cg_uflush) and
// neither u nor p is touched.
template <typename T, bool RED, bool UPD>
__global__ void __launch_bounds__(256)
k_cg_update(size_t n, const double *__restrict__ scal, int ia,
		const T *__restrict__ p, const T *__restrict__ ap,
		T *__restrict__ u, T *__restrict__ r, T *__restrict__ z,
		const T *__restrict__ dinv, const T *__restrict__ scale, int pressure,
		double *__restrict__ part, const int *__restrict__ done,
		ssx_lastred lr)
{
	if (done && *done)
		return;
	__shared__ double sh[4][256];
	const T a = (T) scal[ia];
	double acc[4] = { 0.0, 0.0, 0.0, 0.0 };
	GRID_STRIDE(q, n) {
		if (UPD)
			u[q] = a * p[q] + u[q];
		const T rv = -a * ap[q] + r[q];
		r[q] = rv;
		const T zv = rv * dinv[q];
		z[q] = zv;
		const T w = scale[q];
		acc[0] += (double) (rv * w) * (double) zv;
		acc[1] += (double) (rv * w) * (double) rv;
		if (pressure) {
			acc[2] += (double) (zv * w);
			acc[3] += (double) (rv * w);
		}
	}
	block_sum4(acc, sh);
	if (threadIdx.x < 4)
		part[(size_t) threadIdx.x * SSX_NPART + blockIdx.x] = sh[threadIdx.x][0];
	if (RED && last_workgroup(lr.counter)) {
		// the last workgroup: every lane's stores of all four partials
		// happened before the counter (threads 1-3 included: the fence in
		// last_workgroup follows the block barrier)
		double v[4] = { 0.0, 0.0, 0.0, 0.0 };
		const int nq = lr.pressure ? 4 : 2;
		for (int c = 0; c < nq; ++c)
			v[c] = sum_as_1024(lr.part + (size_t) c * SSX_NPART, SSX_NPART,
					sh[0]);
		if (threadIdx.x == 0)
			cg_scalars1(v[0], v[1], v[2], v[3], lr.scal, lr.o, lr.flags,
					lr.eps, lr.maxit, lr.nunique, lr.pressure);
	}
}

// k_cg_update with the DSS of Ap done as a gather: every copy adds its
// partners' unassembled values to its own in its own add order (tab, the
// order the group kernel evaluates, so the assembled value is bit-identical)
// and takes its weight 1/M from the length of that list (M = 1, 2, 4, 8:
// exact).  Ap stays unassembled, so no copy is written before it is read.
// Same grid and per-thread node order as k_cg_update, so the partials are
// identical.  Each lane takes U of its grid-stride nodes at a time and
// issues the table and partner loads of all U before using any: the
// partner loads depend on the table loads, which depend on soff, and one
// node at a time leaves those three round trips unoverlapped.
template <typename T, int U, bool UPD>
__global__ void __launch_bounds__(256)
k_cg_update_g(unsigned n, unsigned n3, int P, const int *__restrict__ soff,
		const int *__restrict__ tab, const double *__restrict__ scal, int ia,
		const T *__restrict__ p, const T *__restrict__ ap,
		T *__restrict__ u, T *__restrict__ r, T *__restrict__ z,
		const T *__restrict__ dinv, int pressure,
		double *__restrict__ part, const int *__restrict__ done)
{
	if (done && *done)
		return;
	__shared__ double sh[4][256];
	constexpr bool HOIST = sizeof(T) == 4;
	const T a = (T) scal[ia];
	const unsigned stride = gridDim.x * 256u;
	const unsigned lb = xcd_block();
	double acc[4] = { 0.0, 0.0, 0.0, 0.0 };
	for (unsigned q0 = lb * 256u + threadIdx.x; q0 < n;
			q0 += U * stride) {
		int ix[U][7];
		T own[U], pv[U][7], rq[U], dq[U], pq[U], uq[U];
#pragma unroll
		for (int v = 0; v < U; ++v) {
			const unsigned q = q0 + v * stride;
			own[v] = T(0);
			rq[v] = dq[v] = pq[v] = uq[v] = T(0);
			if (q < n) {
				const unsigned e = q / n3;
				const int c = soff[q - e * n3];
				own[v] = ap[q];
				// FP32: the node's own operands go out with the lists
				// (measured faster; in FP64 the extra registers cost more)
				if (HOIST) {
					rq[v] = r[q];
					dq[v] = dinv[q];
					if (UPD) {
						pq[v] = p[q];
						uq[v] = u[q];
					}
				}
				const int *pt = tab + (size_t) e * P + (c >> 3);
#pragma unroll
				for (int k = 0; k < 7; ++k)
					ix[v][k] = k < (c & 7) ? pt[k] : -1;
			} else {
#pragma unroll
				for (int k = 0; k < 7; ++k)
					ix[v][k] = -1;
			}
		}
#pragma unroll
		for (int v = 0; v < U; ++v)
#pragma unroll
			for (int k = 0; k < 7; ++k)
				pv[v][k] = ix[v][k] >= 0 ? ap[ix[v][k]] : T(0);
#pragma unroll
		for (int v = 0; v < U; ++v) {
			const unsigned q = q0 + v * stride;
			if (q >= n)
				break;
			// own + partners left to right; unused entries are -1 and
			// come last
			T s = own[v];
			int m = 1;
#pragma unroll
			for (int k = 0; k < 7; ++k)
				if (ix[v][k] >= 0) {
					s = s + pv[v][k];
					++m;
				}
			const T w = T(1) / T(m);
			if (UPD)
				u[q] = a * (HOIST ? pq[v] : p[q]) + (HOIST ? uq[v] : u[q]);
			const T rv = -a * s + (HOIST ? rq[v] : r[q]);
			r[q] = rv;
			const T zv = rv * (HOIST ? dq[v] : dinv[q]);
			z[q] = zv;
			acc[0] += (double) (rv * w) * (double) zv;
			acc[1] += (double) (rv * w) * (double) rv;
			if (pressure) {
				acc[2] += (double) (zv * w);
				acc[3] += (double) (rv * w);
			}
		}
	}
	block_sum4(acc, sh);
	if (threadIdx.x < 4)
		part[(size_t) threadIdx.x * SSX_NPART + lb] = sh[threadIdx.x][0];
}

template <typename T>
static int launch_cg_update_g(hipStream_t st, size_t n, int n3, int P,
		const int *soff, const int *tab, const double *scal, int ia,
		const void *p, const void *ap, void *u, void *r, void *z,
		const void *dinv, int pressure, double *part, const int *done)
{
	if (n >= (1ull << 31) || n % (size_t) n3) {
		ssx_set_error("k_cg_update_g: %zu nodes, %d per element", n, n3);
		return SSX_EINVAL;
	}
	if (u)
		hipLaunchKernelGGL((k_cg_update_g<T, SSX_GATHER_U, true>),
				dim3(SSX_NPART), dim3(256), 0, st, (unsigned) n,
				(unsigned) n3, P, soff, tab, scal, ia, (const T *) p,
				(const T *) ap, (T *) u, (T *) r, (T *) z, (const T *) dinv,
				pressure, part, done);
	else
		hipLaunchKernelGGL((k_cg_update_g<T, SSX_GATHER_U, false>),
				dim3(SSX_NPART), dim3(256), 0, st, (unsigned) n,
				(unsigned) n3, P, soff, tab, scal, ia, (const T *) p,
				(const T *) ap, (T *) u, (T *) r, (T *) z, (const T *) dinv,
				pressure, part, done);
	LAUNCH_CHECK("k_cg_update_g");
	return SSX_OK;
}

extern "C" int ssxk_cg_update_g(void *stream, size_t n, int n3, int P,
		const int *soff, const int *tab, const double *scal, int ia,
		const double *p, const double *ap, double *u, double *r, double *z,
		const double *dinv, int pressure, double *part, const int *done)
{
	return launch_cg_update_g<double>((hipStream_t) stream, n, n3, P, soff,
			tab, scal, ia, p, ap, u, r, z, dinv, pressure, part, done);
}

extern "C" int ssxk_cg_update_g_lp(void *stream, size_t n, int n3, int P,
		const int *soff, const int *tab, const double *scal, int ia,
		const void *p, const void *ap, void *u, void *r, void *z,
		const void *dinv, int pressure, double *part, const int *done)
{
	return launch_cg_update_g<float>((hipStream_t) stream, n, n3, P, soff,
			tab, scal, ia, p, ap, u, r, z, dinv, pressure, part, done);
}

__device__ __forceinline__ double block_sum256(double v, double *sh);

// The residual r = f - DSS(a) with the DSS of the operator's unassembled
// output a done as a gather (k_cg_update_g's partner lists and order, so
// the assembled value is the group kernel's bit for bit).
//   INIT false: r only (the V-cycle's compute_residual, precon.c:302-309,
//               after the unassembled global_opeval).
//   INIT true:  the start of the CG (solver.c:46-57, var != PRESSURE):
//               f = f * mask first (mask nullable, f written back), then
//               r, z = dinv r, p = z, and the partials of <r,z>_w in the
//               grid, per-thread order and form of k_dot (one partial per
//               logical workgroup,
//               weights 1/M), so ssxk_reduce gives k_dot's value.
// f - s is the exact k_axpby(1, f, -1, r) result: both products are exact.
template <typename T, bool INIT>
__global__ void __launch_bounds__(256)
k_resid_g(unsigned n, unsigned n3, int P, const int *__restrict__ soff,
		const int *__restrict__ tab, const T *__restrict__ a,
		T *__restrict__ f, const T *__restrict__ mask,
		T *__restrict__ r, T *__restrict__ z, T *__restrict__ p,
		const T *__restrict__ dinv, double *__restrict__ part)
{
	constexpr int U = 2;
	// FP32: the node's own operands go out with the lists (as in
	// k_cg_update_g)
	constexpr bool HOIST = sizeof(T) == 4;
	__shared__ double sh[256];
	const unsigned stride = gridDim.x * 256u;
	const unsigned lb = xcd_block();
	double acc = 0.0;
	for (unsigned q0 = lb * 256u + threadIdx.x; q0 < n;
			q0 += U * stride) {
		int ix[U][7];
		T own[U], pv[U][7], fq[U], mq[U], dq[U];
#pragma unroll
		for (int v = 0; v < U; ++v) {
			const unsigned q = q0 + v * stride;
			own[v] = T(0);
			fq[v] = mq[v] = dq[v] = T(0);
			if (q < n) {
				const unsigned e = q / n3;
				const int c = soff[q - e * n3];
				own[v] = a[q];
				if (HOIST) {
					fq[v] = f[q];
					if (INIT && mask)
						mq[v] = mask[q];
					if (INIT)
						dq[v] = dinv[q];
				}
				const int *pt = tab + (size_t) e * P + (c >> 3);
#pragma unroll
				for (int k = 0; k < 7; ++k)
					ix[v][k] = k < (c & 7) ? pt[k] : -1;
			} else {
#pragma unroll
				for (int k = 0; k < 7; ++k)
					ix[v][k] = -1;
			}
		}
#pragma unroll
		for (int v = 0; v < U; ++v)
#pragma unroll
			for (int k = 0; k < 7; ++k)
				pv[v][k] = ix[v][k] >= 0 ? a[ix[v][k]] : T(0);
#pragma unroll
		for (int v = 0; v < U; ++v) {
			const unsigned q = q0 + v * stride;
			if (q >= n)
				break;
			T s = own[v];
			int m = 1;
#pragma unroll
			for (int k = 0; k < 7; ++k)
				if (ix[v][k] >= 0) {
					s = s + pv[v][k];
					++m;
				}
			T fv = HOIST ? fq[v] : f[q];
			if (INIT && mask) {
				fv = fv * (HOIST ? mq[v] : mask[q]);
				f[q] = fv;
			}
			const T rv = fv - s;
			r[q] = rv;
			if (INIT) {
				const T zv = rv * (HOIST ? dq[v] : dinv[q]);
				z[q] = zv;
				p[q] = zv;
				const T wr = rv * (T(1) / T(m));
				acc += (double) wr * (double) zv;
			}
		}
	}
	if (INIT) {
		const double t = block_sum256(acc, sh);
		if (threadIdx.x == 0)
			part[lb] = t;
	}
}

template <typename T>
static int launch_resid_g(hipStream_t st, int init, size_t n, int n3, int P,
		const int *soff, const int *tab, const void *a, void *f,
		const void *mask, void *r, void *z, void *p, const void *dinv,
		double *part)
{
	if (n >= (1ull << 31) || n % (size_t) n3) {
		ssx_set_error("k_resid_g: %zu nodes, %d per element", n, n3);
		return SSX_EINVAL;
	}
	if (init)
		hipLaunchKernelGGL((k_resid_g<T, true>), dim3(SSX_NPART), dim3(256),
				0, st, (unsigned) n, (unsigned) n3, P, soff, tab,
				(const T *) a, (T *) f, (const T *) mask, (T *) r, (T *) z,
				(T *) p, (const T *) dinv, part);
	else
		hipLaunchKernelGGL((k_resid_g<T, false>), dim3(SSX_NPART), dim3(256),
				0, st, (unsigned) n, (unsigned) n3, P, soff, tab,
				(const T *) a, (T *) f, (const T *) mask, (T *) r, (T *) z,
				(T *) p, (const T *) dinv, part);
	LAUNCH_CHECK("k_resid_g");
	return SSX_OK;
}

extern "C" int ssxk_resid_g(void *stream, int init, size_t n, int n3, int P,
		const int *soff, const int *tab, const double *a, double *f,
		const double *mask, double *r, double *z, double *p,
		const double *dinv, double *part)
{
	return launch_resid_g<double>((hipStream_t) stream, init, n, n3, P, soff,
			tab, a, f, mask, r, z, p, dinv, part);
}

extern "C" int ssxk_resid_g_lp(void *stream, int init, size_t n, int n3,
		int P, const int *soff, const int *tab, const void *a, void *f,
		const void *mask, void *r, void *z, void *p, const void *dinv,
		double *part)
{
	return launch_resid_g<float>((hipStream_t) stream, init, n, n3, P, soff,
			tab, a, f, mask, r, z, p, dinv, part);
}

// p = (z - c) + beta p, masked (c: pressure null-space constant)
template <typename T>
__global__ void __launch_bounds__(256)
k_cg_pupdate(size_t n, const T *__restrict__ z, T *__restrict__ p,
		const double *__restrict__ scal, int ib, int ic,
		const T *__restrict__ mask, const int *__restrict__ done)
{
	if (done && *done)
		return;
	const T b = (T) scal[ib];
	const T c = ic >= 0 ? (T) -scal[ic] : T(0);
	GRID_STRIDE(q, n) {
		const T zv = ic >= 0 ? c * T(1) + z[q] : z[q];
		T v = b * p[q] + zv;
		p[q] = mask ? v * mask[q] : v;
	}
}

// u += scal[ia] p: the last CG iteration's solution step when the steps
// are taken by the operator launches (runs whatever the done flag says)
template <typename T>
__global__ void __launch_bounds__(256)
k_cg_uflush(size_t n, const double *__restrict__ scal, int ia,
		const T *__restrict__ p, T *__restrict__ u)
{
	const T a = (T) scal[ia];
	GRID_STRIDE(q, n)
		u[q] = a * p[q] + u[q];
}

extern "C" int ssxk_cg_uflush(void *stream, size_t n, const double *scal,
		int ia, const double *p, double *u)
{
	hipLaunchKernelGGL(k_cg_uflush<double>, vgrid(n), dim3(256), 0,
			(hipStream_t) stream, n, scal, ia, p, u);
	LAUNCH_CHECK("k_cg_uflush");
	return SSX_OK;
}

extern "C" int ssxk_cg_uflush_lp(void *stream, size_t n, const double *scal,
		int ia, const void *p, void *u)
{
	hipLaunchKernelGGL(k_cg_uflush<float>, vgrid(n), dim3(256), 0,
			(hipStream_t) stream, n, scal, ia, (const float *) p, (float *) u);
	LAUNCH_CHECK("k_cg_uflush");
	return SSX_OK;
}

// partials a reduction lane loads ahead of its adds
#define RED_PRE 8

// single-block reductions + scalar logic of one CG iteration
//   mode 0: s[o+PAP] = sum part[0..cnt); s[o+ALPHA] = s[o+RDZ] / pap
//   mode 1: rz, rr (, zs, rs) from the 4 update partials; null-space
//           constant c = zs/nunique; rdz = rz - c rs; beta = rdz / rdz_old;
//           iteration count ++, done when sqrt(rr) <= eps or count >= maxit
//           (eps < 0: the tolerance is s[o+TOL], formed on the device)

__global__ void __launch_bounds__(1024)
k_cg_reduce(int mode, const double *__restrict__ part, long long cnt,
		double *__restrict__ scal, int o, int *__restrict__ flags,
		double eps, int maxit, double nunique, int pressure)
{
	if (flags[0])
		return;
	__shared__ double sh[4][1024];
	const int nq = mode == 0 ? 1 : (pressure ? 4 : 2);
	const long long m = (mode == 0 || cnt > 0) ? cnt : SSX_NPART;
	const long long t = threadIdx.x;
	// The lane's partials are added in index order, as one loop would, but
	// loaded RED_PRE at a time (and every channel's first together) so
	// that the loads' latencies overlap.
	double first[4];
#pragma unroll
	for (int c = 0; c < 4; ++c) {
		// branch-free: an unused slot loads part[0] and discards it
		const bool use = c < nq && t < m;
		const double v = part[use ? (size_t) c * SSX_NPART + t : 0];
		first[c] = use ? v : 0.0;
	}
	for (int c = 0; c < 4; ++c) {
		double acc = 0.0;
		if (c < nq && t < m) {
			const double *pp = part + (size_t) c * SSX_NPART;
			acc += first[c];
			// RED_PRE loads in flight; past m a slot loads pp[0] unused
			for (long long q = t + 1024; q < m; q += RED_PRE * 1024) {
				double v[RED_PRE];
#pragma unroll
				for (int i = 0; i < RED_PRE; ++i) {
					const long long qi = q + i * 1024;
					v[i] = pp[qi < m ? qi : 0];
				}
#pragma unroll
				for (int i = 0; i < RED_PRE; ++i)
					if (q + i * 1024 < m)
						acc += v[i];
			}
		}
		sh[c][threadIdx.x] = acc;
	}
	__syncthreads();
	for (int st = 512; st >= 64; st >>= 1) {
		if ((int) threadIdx.x < st)
			for (int c = 0; c < 4; ++c)
				sh[c][threadIdx.x] += sh[c][threadIdx.x + st];
		__syncthreads();
	}
	// the tree's steps below 64 inside wavefront 0, by shuffles: lane t
	// adds lane t + st's value as the LDS step would, without barriers
	if (threadIdx.x >= 64)
		return;
	double v[4];
	for (int c = 0; c < 4; ++c)
		v[c] = sh[c][threadIdx.x];
	for (int st = 32; st > 0; st >>= 1)
		for (int c = 0; c < 4; ++c)
			v[c] += __shfl_down(v[c], st, 64);
	if (threadIdx.x != 0)
		return;
	if (mode == 0)
		cg_scalars0(v[0], scal, o);
	else
		cg_scalars1(v[0], v[1], v[2], v[3], scal, o, flags, eps, maxit,
				nunique, pressure);
}

__global__ void k_cg_tol(double *scal, int dst, int src, double eps,
		double rtol)
{
	if (threadIdx.x == 0)
		scal[dst] = fmax(eps, rtol * sqrt(scal[src]));
}

extern "C" int ssxk_cg_tol(void *stream, double *scal, int dst, int src,
		double eps, double rtol)
{
	hipLaunchKernelGGL(k_cg_tol, dim3(1), dim3(64), 0, (hipStream_t) stream,
			scal, dst, src, eps, rtol);
	LAUNCH_CHECK("k_cg_tol");
	return SSX_OK;
}

template <typename T>
static int launch_cg_update(hipStream_t st, size_t n, const double *scal,
		int ia, const void *p, const void *ap, void *u, void *r, void *z,
		const void *dinv, const void *scale, int pressure, double *part,
		const int *done, const ssx_lastred *lr = nullptr,
		int nblocks = SSX_NPART)
{
	if (lr && !u) {
		ssx_set_error("k_cg_update: the last-workgroup form takes u");
		return SSX_EINVAL;
	}
	if (lr)
		hipLaunchKernelGGL((k_cg_update<T, true, true>), dim3(SSX_NPART),
				dim3(256), 0, st, n, scal, ia, (const T *) p, (const T *) ap,
				(T *) u, (T *) r, (T *) z, (const T *) dinv,
				(const T *) scale, pressure, part, done, *lr);
	else if (u)
		hipLaunchKernelGGL((k_cg_update<T, false, true>), dim3(nblocks),
				dim3(256), 0, st, n, scal, ia, (const T *) p, (const T *) ap,
				(T *) u, (T *) r, (T *) z, (const T *) dinv,
				(const T *) scale, pressure, part, done, ssx_lastred{});
	else
		hipLaunchKernelGGL((k_cg_update<T, false, false>), dim3(nblocks),
				dim3(256), 0, st, n, scal, ia, (const T *) p, (const T *) ap,
				(T *) u, (T *) r, (T *) z, (const T *) dinv,
				(const T *) scale, pressure, part, done, ssx_lastred{});
	LAUNCH_CHECK("k_cg_update");
	return SSX_OK;
}

extern "C" int ssxk_cg_update_nb(void *stream, size_t n, const double *scal,
		int ia, const double *p, const double *ap, double *u, double *r,
		double *z, const double *dinv, const double *scale, int pressure,
		double *part, const int *done, int nblocks)
{
	if (nblocks < 1 || nblocks > SSX_NPART) {
		ssx_set_error("ssxk_cg_update_nb: %d workgroups", nblocks);
		return SSX_EINVAL;
	}
	return launch_cg_update<double>((hipStream_t) stream, n, scal, ia, p, ap,
			u, r, z, dinv, scale, pressure, part, done, nullptr, nblocks);
}

extern "C" int ssxk_cg_update_red(void *stream, size_t n, const double *scal,
		int ia, const double *p, const double *ap, double *u, double *r,
		double *z, const double *dinv, const double *scale, int pressure,
		double *part, const int *done, const ssx_lastred *lr)
{
	return launch_cg_update<double>((hipStream_t) stream, n, scal, ia, p, ap,
			u, r, z, dinv, scale, pressure, part, done, lr);
}

extern "C" int ssxk_cg_update_red_lp(void *stream, size_t n,
		const double *scal, int ia, const void *p, const void *ap, void *u,
		void *r, void *z, const void *dinv, const void *scale, int pressure,
		double *part, const int *done, const ssx_lastred *lr)
{
	return launch_cg_update<float>((hipStream_t) stream, n, scal, ia, p, ap,
			u, r, z, dinv, scale, pressure, part, done, lr);
}

template <typename T>
static int launch_cg_pupdate(hipStream_t st, size_t n, const void *z,
		void *p, const double *scal, int ib, int ic, const void *mask,
		const int *done)
{
	hipLaunchKernelGGL(k_cg_pupdate<T>, vgrid(n), dim3(256), 0, st, n,
			(const T *) z, (T *) p, scal, ib, ic, (const T *) mask, done);
	LAUNCH_CHECK("k_cg_pupdate");
	return SSX_OK;
}

extern "C" int ssxk_cg_update(void *stream, size_t n, const double *scal,
		int ia, const double *p, const double *ap, double *u, double *r,
		double *z, const double *dinv, const double *scale, int pressure,
		double *part, const int *done)
{
	return launch_cg_update<double>((hipStream_t) stream, n, scal, ia, p, ap,
			u, r, z, dinv, scale, pressure, part, done);
}

extern "C" int ssxk_cg_update_lp(void *stream, size_t n, const double *scal,
		int ia, const void *p, const void *ap, void *u, void *r, void *z,
		const void *dinv, const void *scale, int pressure, double *part,
		const int *done)
{
	return launch_cg_update<float>((hipStream_t) stream, n, scal, ia, p, ap,
			u, r, z, dinv, scale, pressure, part, done);
}

extern "C" int ssxk_cg_pupdate(void *stream, size_t n, const double *z,
		double *p, const double *scal, int ib, int ic, const double *mask,
		const int *done)
{
	return launch_cg_pupdate<double>((hipStream_t) stream, n, z, p, scal, ib,
			ic, mask, done);
}

extern "C" int ssxk_cg_pupdate_lp(void *stream, size_t n, const void *z,
		void *p, const double *scal, int ib, int ic, const void *mask,
		const int *done)
{
	return launch_cg_pupdate<float>((hipStream_t) stream, n, z, p, scal, ib,
			ic, mask, done);
}

extern "C" int ssxk_cg_reduce(void *stream, int mode, const double *part,
		long long cnt, double *scal, int o, int *flags, double eps,
		int maxit, double nunique, int pressure)
{
	hipLaunchKernelGGL(k_cg_reduce, dim3(1), dim3(1024), 0,
			(hipStream_t) stream, mode, part, cnt, scal, o, flags, eps, maxit,
			nunique, pressure);
	LAUNCH_CHECK("k_cg_reduce");
	return SSX_OK;
}

// ---------------------------------------------------------------------
// Order transfer (interpolation.c:136-176), one element per block
// prol[p*nf+i] = l_p^{coarse}(xi_i^{fine})
// ---------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(256)
k_restrict(int nf, int nc, const T *__restrict__ prol,
		const T *__restrict__ fine, T *__restrict__ coarse)
{
	extern __shared__ char smraw[];
	T *sm = (T *) smraw;
	const int nf3 = nf * nf * nf, nc3 = nc * nc * nc;
	T *sp = sm;                            // nc*nf
	T *su = sp + nc * nf;                  // nf^3
	T *t1 = su + nf3;                      // nf*nf*nc
	T *t2 = t1 + nf * nf * nc;             // nf*nc*nc
	const int e = blockIdx.x;
	for (int q = threadIdx.x; q < nc * nf; q += blockDim.x)
		sp[q] = prol[q];
	for (int q = threadIdx.x; q < nf3; q += blockDim.x)
		su[q] = fine[(size_t) e * nf3 + q];
	__syncthreads();
	for (int q = threadIdx.x; q < nf * nf * nc; q += blockDim.x) {
		const int p = q % nc, kj = q / nc;
		T s = T(0);
		for (int c = 0; c < nf; ++c)
			s += sp[p * nf + c] * su[kj * nf + c];
		t1[q] = s;
	}
	__syncthreads();
	for (int q = threadIdx.x; q < nf * nc * nc; q += blockDim.x) {
		const int i = q % nc, jq = (q / nc) % nc, k = q / (nc * nc);
		T s = T(0);
		for (int l = 0; l < nf; ++l)
			s += t1[(k * nf + l) * nc + i] * sp[jq * nf + l];
		t2[q] = s;
	}
	__syncthreads();
	for (int q = threadIdx.x; q < nc3; q += blockDim.x) {
		const int ji = q % (nc * nc), k = q / (nc * nc);
		T s = T(0);
		for (int l = 0; l < nf; ++l)
			s += t2[l * nc * nc + ji] * sp[k * nf + l];
		coarse[(size_t) e * nc3 + q] = s;
	}
}

template <typename T>
__global__ void __launch_bounds__(256)
k_prolong(int nf, int nc, const T *__restrict__ prol,
		const T *__restrict__ coarse, T *__restrict__ fine, int accumulate)
{
	extern __shared__ char smraw[];
	T *sm = (T *) smraw;
	const int nf3 = nf * nf * nf, nc3 = nc * nc * nc;
	T *sp = sm;                            // nc*nf
	T *su = sp + nc * nf;                  // nc^3
	T *t1 = su + nc3;                      // nc*nc*nf
	T *t2 = t1 + nc * nc * nf;             // nc*nf*nf
	const int e = blockIdx.x;
	for (int q = threadIdx.x; q < nc * nf; q += blockDim.x)
		sp[q] = prol[q];
	for (int q = threadIdx.x; q < nc3; q += blockDim.x)
		su[q] = coarse[(size_t) e * nc3 + q];
	__syncthreads();
	for (int q = threadIdx.x; q < nc * nc * nf; q += blockDim.x) {
		const int i = q % nf, kj = q / nf;
		T s = T(0);
		for (int p = 0; p < nc; ++p)
			s += sp[p * nf + i] * su[kj * nc + p];
		t1[q] = s;
	}
	__syncthreads();
	for (int q = threadIdx.x; q < nc * nf * nf; q += blockDim.x) {
		const int i = q % nf, jq = (q / nf) % nf, k = q / (nf * nf);
		T s = T(0);
		for (int l = 0; l < nc; ++l)
			s += t1[(k * nc + l) * nf + i] * sp[l * nf + jq];
		t2[q] = s;
	}
	__syncthreads();
	for (int q = threadIdx.x; q < nf3; q += blockDim.x) {
		const int ji = q % (nf * nf), k = q / (nf * nf);
		T s = T(0);
		for (int l = 0; l < nc; ++l)
			s += t2[l * nf * nf + ji] * sp[l * nf + k];
		const size_t o = (size_t) e * nf3 + q;
		fine[o] = accumulate ? s + fine[o] : s;
	}
}

// The same transfers with one wavefront per element and the sizes known at
// compile time, for the p-multigrid's own level pairs (nc = (nf-1)/2 + 1).
// Every load of the element is issued before the first product (the
// accumulated fine values into registers, indexes clamped to the element,
// so no branch separates them), the three stages are separated by wave
// barriers only, and a block holds four elements.  Each output is the same
// sum in the same order as above, so results are identical bit for bit.
#define XW 4                            // elements (wavefronts) per block

template <typename T, int NF, int NC, bool ACC>
__global__ void __launch_bounds__(64 * XW)
k_prolong_w(int nel, const T *__restrict__ prol, const T *__restrict__ coarse,
		T *__restrict__ fine)
{
	constexpr int F3 = NF * NF * NF, C3 = NC * NC * NC;
	constexpr int S1 = NC * NC * NF, S2 = NC * NF * NF;
	constexpr int PC = (C3 + 63) / 64, PF = (F3 + 63) / 64;
	constexpr int WS = C3 + S1 + S2;
	extern __shared__ char smraw[];
	T *sp = (T *) smraw;                  // nc*nf
	const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
	T *su = sp + NC * NF + wv * WS, *t1 = su + C3, *t2 = t1 + S1;
	const int e = blockIdx.x * XW + wv;
	const bool act = e < nel;
	const size_t ec = (size_t) (act ? e : 0);
	static_assert(NC * NF <= 64 * XW, "one matrix entry per lane");
	// the matrix by uniform (scalar) loads, each lane keeping its entry: its
	// LDS store then waits for no vector load of the element
	T pv = T(0);
#pragma unroll
	for (int q = 0; q < NC * NF; ++q) {
		const T m = prol[q];
		pv = (int) threadIdx.x == q ? m : pv;
	}
	T cv[PC], fv[PF];
#pragma unroll
	for (int m = 0; m < PC; ++m) {
		const int q = lane + 64 * m;
		cv[m] = coarse[ec * C3 + (q < C3 ? q : C3 - 1)];
	}
	if constexpr (ACC) {
#pragma unroll
		for (int m = 0; m < PF; ++m) {
			const int q = lane + 64 * m;
			fv[m] = fine[ec * F3 + (q < F3 ? q : F3 - 1)];
		}
	}
	if (threadIdx.x < NC * NF)
		sp[threadIdx.x] = pv;
#pragma unroll
	for (int m = 0; m < PC; ++m) {
		const int q = lane + 64 * m;
		if (q < C3)
			su[q] = cv[m];
	}
	lds_sync();
#pragma unroll
	for (int m = 0; m < (S1 + 63) / 64; ++m) {
		const int q = lane + 64 * m;
		if (q < S1) {
			const int i = q % NF, kj = q / NF;
			T s = T(0);
#pragma unroll
			for (int p = 0; p < NC; ++p)
				s += sp[p * NF + i] * su[kj * NC + p];
			t1[q] = s;
		}
	}
	wave_sync();
#pragma unroll
	for (int m = 0; m < (S2 + 63) / 64; ++m) {
		const int q = lane + 64 * m;
		if (q < S2) {
			const int i = q % NF, jq = (q / NF) % NF, k = q / (NF * NF);
			T s = T(0);
#pragma unroll
			for (int l = 0; l < NC; ++l)
				s += t1[(k * NC + l) * NF + i] * sp[l * NF + jq];
			t2[q] = s;
		}
	}
	wave_sync();
#pragma unroll
	for (int m = 0; m < PF; ++m) {
		const int q = lane + 64 * m;
		if (q < F3) {
			const int ji = q % (NF * NF), k = q / (NF * NF);
			T s = T(0);
#pragma unroll
			for (int l = 0; l < NC; ++l)
				s += t2[l * NF * NF + ji] * sp[l * NF + k];
			if (act)
				fine[ec * F3 + q] = ACC ? s + fv[m] : s;
		}
	}
}

template <typename T, int NF, int NC>
__global__ void __launch_bounds__(64 * XW)
k_restrict_w(int nel, const T *__restrict__ prol, const T *__restrict__ fine,
		T *__restrict__ coarse)
{
	constexpr int F3 = NF * NF * NF, C3 = NC * NC * NC;
	constexpr int S1 = NF * NF * NC, S2 = NF * NC * NC;
	constexpr int PF = (F3 + 63) / 64;
	constexpr int WS = F3 + S1 + S2;
	extern __shared__ char smraw[];
	T *sp = (T *) smraw;                  // nc*nf
	const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
	T *su = sp + NC * NF + wv * WS, *t1 = su + F3, *t2 = t1 + S1;
	const int e = blockIdx.x * XW + wv;
	const bool act = e < nel;
	const size_t ec = (size_t) (act ? e : 0);
	static_assert(NC * NF <= 64 * XW, "one matrix entry per lane");
	// the matrix by uniform (scalar) loads, each lane keeping its entry: its
	// LDS store then waits for no vector load of the element
	T pv = T(0);
#pragma unroll
	for (int q = 0; q < NC * NF; ++q) {
		const T m = prol[q];
		pv = (int) threadIdx.x == q ? m : pv;
	}
	T fv[PF];
#pragma unroll
	for (int m = 0; m < PF; ++m) {
		const int q = lane + 64 * m;
		fv[m] = fine[ec * F3 + (q < F3 ? q : F3 - 1)];
	}
	if (threadIdx.x < NC * NF)
		sp[threadIdx.x] = pv;
#pragma unroll
	for (int m = 0; m < PF; ++m) {
		const int q = lane + 64 * m;
		if (q < F3)
			su[q] = fv[m];
	}
	lds_sync();
#pragma unroll
	for (int m = 0; m < (S1 + 63) / 64; ++m) {
		const int q = lane + 64 * m;
		if (q < S1) {
			const int p = q % NC, kj = q / NC;
			T s = T(0);
#pragma unroll
			for (int c = 0; c < NF; ++c)
				s += sp[p * NF + c] * su[kj * NF + c];
			t1[q] = s;
		}
	}
	wave_sync();
#pragma unroll
	for (int m = 0; m < (S2 + 63) / 64; ++m) {
		const int q = lane + 64 * m;
		if (q < S2) {
			const int i = q % NC, jq = (q / NC) % NC, k = q / (NC * NC);
			T s = T(0);
#pragma unroll
			for (int l = 0; l < NF; ++l)
				s += t1[(k * NF + l) * NC + i] * sp[jq * NF + l];
			t2[q] = s;
		}
	}
	wave_sync();
#pragma unroll
	for (int m = 0; m < (C3 + 63) / 64; ++m) {
		const int q = lane + 64 * m;
		if (q < C3) {
			const int ji = q % (NC * NC), k = q / (NC * NC);
			T s = T(0);
#pragma unroll
			for (int l = 0; l < NF; ++l)
				s += t2[l * NC * NC + ji] * sp[k * NF + l];
			if (act)
				coarse[ec * C3 + q] = s;
		}
	}
}

// the wavefront form for (nf, nc) = (n, (n-1)/2 + 1), n = 3..12; false for
// other pairs (the block-per-element kernels above take them)
template <typename T>
static bool launch_transfer_w(hipStream_t st, int restrict_, int nf, int nc,
		int nel, const void *prol, const void *src, void *dst,
		int accumulate)
{
	if (nf < 3 || nf > 12 || nc != (nf - 1) / 2 + 1)
		return false;
	const int grid = (nel + XW - 1) / XW;
	const size_t ws = restrict_
		? (size_t) nf * nf * nf + (size_t) nf * nf * nc + (size_t) nf * nc * nc
		: (size_t) nc * nc * nc + (size_t) nc * nc * nf + (size_t) nc * nf * nf;
	const size_t sh = sizeof(T) * ((size_t) nc * nf + XW * ws);
	if (sh > 65536)                 // beyond the default dynamic LDS limit
		return false;
	switch (nf) {
#define XCASE(NF) case NF:                                                     \
		if (restrict_)                                                 \
			hipLaunchKernelGGL((k_restrict_w<T, NF, (NF - 1) / 2 + 1>),   \
					dim3(grid), dim3(64 * XW), sh, st, nel,        \
					(const T *) prol, (const T *) src, (T *) dst); \
		else if (accumulate)                                           \
			hipLaunchKernelGGL((k_prolong_w<T, NF, (NF - 1) / 2 + 1, true>), \
					dim3(grid), dim3(64 * XW), sh, st, nel,        \
					(const T *) prol, (const T *) src, (T *) dst); \
		else                                                           \
			hipLaunchKernelGGL((k_prolong_w<T, NF, (NF - 1) / 2 + 1, false>), \
					dim3(grid), dim3(64 * XW), sh, st, nel,        \
					(const T *) prol, (const T *) src, (T *) dst); \
		return true;
	XCASE(3) XCASE(4) XCASE(5) XCASE(6) XCASE(7) XCASE(8) XCASE(9)
	XCASE(10) XCASE(11) XCASE(12)
#undef XCASE
	default:
		return false;
	}
}

template <typename T>
static int launch_restrict(hipStream_t st, int nf, int nc, int nel,
		const void *prol, const void *fine, void *coarse)
{
	if (nel > 0 && launch_transfer_w<T>(st, 1, nf, nc, nel, prol, fine,
			coarse, 0)) {
		LAUNCH_CHECK("k_restrict_w");
		return SSX_OK;
	}
	const size_t sh = sizeof(T) * ((size_t) nc * nf + (size_t) nf * nf * nf
			+ (size_t) nf * nf * nc + (size_t) nf * nc * nc);
	hipLaunchKernelGGL(k_restrict<T>, dim3(nel), dim3(256), sh, st, nf, nc,
			(const T *) prol, (const T *) fine, (T *) coarse);
	LAUNCH_CHECK("k_restrict");
	return SSX_OK;
}

template <typename T>
static int launch_prolong(hipStream_t st, int nf, int nc, int nel,
		const void *prol, const void *coarse, void *fine, int accumulate)
{
	if (nel > 0 && launch_transfer_w<T>(st, 0, nf, nc, nel, prol, coarse,
			fine, accumulate)) {
		LAUNCH_CHECK("k_prolong_w");
		return SSX_OK;
	}
	const size_t sh = sizeof(T) * ((size_t) nc * nf + (size_t) nc * nc * nc
			+ (size_t) nc * nc * nf + (size_t) nc * nf * nf);
	hipLaunchKernelGGL(k_prolong<T>, dim3(nel), dim3(256), sh, st, nf, nc,
			(const T *) prol, (const T *) coarse, (T *) fine, accumulate);
	LAUNCH_CHECK("k_prolong");
	return SSX_OK;
}

extern "C" int ssxk_restrict(void *stream, int nf, int nc, int nel,
		const double *prol, const double *fine, double *coarse)
{
	return launch_restrict<double>((hipStream_t) stream, nf, nc, nel, prol,
			fine, coarse);
}

extern "C" int ssxk_prolong(void *stream, int nf, int nc, int nel,
		const double *prol, const double *coarse, double *fine,
		int accumulate)
{
	return launch_prolong<double>((hipStream_t) stream, nf, nc, nel, prol,
			coarse, fine, accumulate);
}

extern "C" int ssxk_restrict_lp(void *stream, int nf, int nc, int nel,
		const void *prol, const void *fine, void *coarse)
{
	return launch_restrict<float>((hipStream_t) stream, nf, nc, nel, prol,
			fine, coarse);
}

extern "C" int ssxk_prolong_lp(void *stream, int nf, int nc, int nel,
		const void *prol, const void *coarse, void *fine, int accumulate)
{
	return launch_prolong<float>((hipStream_t) stream, nf, nc, nel, prol,
			coarse, fine, accumulate);
}

// ---------------------------------------------------------------------
// Reductions: fixed grid of SSX_NPART blocks, fixed tree -> deterministic
// ---------------------------------------------------------------------

__device__ __forceinline__ double block_sum256(double v, double *sh)
{
	sh[threadIdx.x] = v;
	__syncthreads();
#pragma unroll
	for (int s = 128; s > 0; s >>= 1) {
		if (threadIdx.x < s)
			sh[threadIdx.x] += sh[threadIdx.x + s];
		__syncthreads();
	}
	return sh[0];
}

template <typename T>
__global__ void __launch_bounds__(256)
k_dot(size_t n, const T *__restrict__ x, const T *__restrict__ s,
		const T *__restrict__ y, double *__restrict__ part)
{
	__shared__ double sh[256];
	double acc = 0.0;
	GRID_STRIDE(q, n) {
		T a = x[q] * s[q];
		acc += y ? (double) a * (double) y[q] : (double) a;
	}
	double t = block_sum256(acc, sh);
	if (threadIdx.x == 0)
		part[blockIdx.x] = t;
}

__global__ void __launch_bounds__(256)
k_reduce(const double *__restrict__ part, double *__restrict__ scal, int dst)
{
	__shared__ double sh[256];
	double acc = 0.0;
	for (int q = threadIdx.x; q < SSX_NPART; q += 256)
		acc += part[q];
	double t = block_sum256(acc, sh);
	if (threadIdx.x == 0)
		scal[dst] = t;
}

extern "C" int ssxk_dot(void *stream, size_t n, const double *x,
		const double *s, const double *y, double *part, int slot)
{
	hipLaunchKernelGGL(k_dot<double>, dim3(SSX_NPART), dim3(256), 0,
			(hipStream_t) stream, n, x, s, y, part + (size_t) slot * SSX_NPART);
	LAUNCH_CHECK("k_dot");
	return SSX_OK;
}

extern "C" int ssxk_dot_lp(void *stream, size_t n, const void *x,
		const void *s, const void *y, double *part, int slot)
{
	hipLaunchKernelGGL(k_dot<float>, dim3(SSX_NPART), dim3(256), 0,
			(hipStream_t) stream, n, (const float *) x, (const float *) s,
			(const float *) y, part + (size_t) slot * SSX_NPART);
	LAUNCH_CHECK("k_dot");
	return SSX_OK;
}

extern "C" int ssxk_reduce(void *stream, double *part, int slot, double *scal,
		int dst)
{
	hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, (hipStream_t) stream,
			part + (size_t) slot * SSX_NPART, scal, dst);
	LAUNCH_CHECK("k_reduce");
	return SSX_OK;
}

// ---------------------------------------------------------------------
// Flexible CG (solver.c:83-157) orthogonalisation, fused: each kernel
// reads every vector once.  Per-thread accumulation order, grid and tree
// are those of k_dot / k_reduce, so every inner product is bit-identical
// to the one-dot-per-launch sequence.
// ---------------------------------------------------------------------

struct VecSet {
	const double *v[SSX_FCG_M];
};

// part[k*NPART + blk] = block partial of <x, y_k>_w, k < m
__global__ void __launch_bounds__(256)
k_multidot(size_t n, const double *__restrict__ x,
		const double *__restrict__ s, VecSet ys, int m,
		double *__restrict__ part)
{
	__shared__ double sh[256];
	double acc[SSX_FCG_M];
#pragma unroll
	for (int k = 0; k < SSX_FCG_M; ++k)
		acc[k] = 0.0;
	GRID_STRIDE(q, n) {
		const double a = x[q] * s[q];
#pragma unroll
		for (int k = 0; k < SSX_FCG_M; ++k)
			if (k < m)
				acc[k] += a * ys.v[k][q];
	}
	for (int k = 0; k < m; ++k) {
		const double t = block_sum256(acc[k], sh);
		if (threadIdx.x == 0)
			part[(size_t) k * SSX_NPART + blockIdx.x] = t;
		__syncthreads();
	}
}

// block k: scal[dst + k] = sum of part[k*NPART ..]; with div >= 0 also
// scal[quot + k] = scal[dst + k] / scal[div + k]
__global__ void __launch_bounds__(256)
k_reduce_multi(const double *__restrict__ part, double *__restrict__ scal,
		int dst, int div, int quot)
{
	__shared__ double sh[256];
	const double *pp = part + (size_t) blockIdx.x * SSX_NPART;
	double acc = 0.0;
	for (int q = threadIdx.x; q < SSX_NPART; q += 256)
		acc += pp[q];
	const double t = block_sum256(acc, sh);
	if (threadIdx.x == 0) {
		scal[dst + blockIdx.x] = t;
		if (div >= 0)
			scal[quot + blockIdx.x] = t / scal[div + blockIdx.x];
	}
}

// y = mask * (x - b_0 p_0 - b_1 p_1 - ... - b_{m-1} p_{m-1}), the updates
// applied in that order (one axpy per history vector in the reference)
__global__ void __launch_bounds__(256)
k_multi_axpy(size_t n, const double *__restrict__ x, VecSet ps, int m,
		const double *__restrict__ scal, int ib,
		const double *__restrict__ mask, double *__restrict__ y)
{
	double b[SSX_FCG_M];
#pragma unroll
	for (int k = 0; k < SSX_FCG_M; ++k)
		b[k] = k < m ? -scal[ib + k] : 0.0;
	GRID_STRIDE(q, n) {
		double v = x[q];
#pragma unroll
		for (int k = 0; k < SSX_FCG_M; ++k)
			if (k < m)
				v = b[k] * ps.v[k][q] + v;
		y[q] = mask ? v * mask[q] : v;
	}
}

extern "C" int ssxk_multidot(void *stream, size_t n, const double *x,
		const double *s, const double *const *ys, int m, double *part,
		double *scal, int dst, int div, int quot)
{
	if (m <= 0)
		return SSX_OK;
	if (m > SSX_FCG_M) {
		ssx_set_error("ssxk_multidot: m=%d > %d", m, SSX_FCG_M);
		return SSX_EINVAL;
	}
	VecSet v;
	for (int k = 0; k < SSX_FCG_M; ++k)
		v.v[k] = k < m ? ys[k] : ys[0];
	hipLaunchKernelGGL(k_multidot, dim3(SSX_NPART), dim3(256), 0,
			(hipStream_t) stream, n, x, s, v, m, part);
	LAUNCH_CHECK("k_multidot");
	hipLaunchKernelGGL(k_reduce_multi, dim3(m), dim3(256), 0,
			(hipStream_t) stream, part, scal, dst, div, quot);
	LAUNCH_CHECK("k_reduce_multi");
	return SSX_OK;
}

extern "C" int ssxk_multi_axpy(void *stream, size_t n, const double *x,
		const double *const *ps, int m, const double *scal, int ib,
		const double *mask, double *y)
{
	if (m > SSX_FCG_M) {
		ssx_set_error("ssxk_multi_axpy: m=%d > %d", m, SSX_FCG_M);
		return SSX_EINVAL;
	}
	VecSet v;
	for (int k = 0; k < SSX_FCG_M; ++k)
		v.v[k] = k < m ? ps[k] : x;
	hipLaunchKernelGGL(k_multi_axpy, vgrid(n), dim3(256), 0,
			(hipStream_t) stream, n, x, v, m, scal, ib, mask, y);
	LAUNCH_CHECK("k_multi_axpy");
	return SSX_OK;
}

// u += a p; r -= a ap; partial <r,r>_w (the FCG step, solver.c:139-147)
__global__ void __launch_bounds__(256)
k_fcg_step(size_t n, const double *__restrict__ scal, int ia,
		const double *__restrict__ p, const double *__restrict__ ap,
		double *__restrict__ u, double *__restrict__ r,
		const double *__restrict__ s, double *__restrict__ part)
{
	__shared__ double sh[256];
	const double a = scal[ia], am = -scal[ia];
	double acc = 0.0;
	GRID_STRIDE(q, n) {
		u[q] = a * p[q] + u[q];
		const double rv = am * ap[q] + r[q];
		r[q] = rv;
		acc += (double) (rv * s[q]) * (double) rv;
	}
	const double t = block_sum256(acc, sh);
	if (threadIdx.x == 0)
		part[blockIdx.x] = t;
}

extern "C" int ssxk_fcg_step(void *stream, size_t n, const double *scal,
		int ia, const double *p, const double *ap, double *u, double *r,
		const double *s, double *part)
{
	hipLaunchKernelGGL(k_fcg_step, dim3(SSX_NPART), dim3(256), 0,
			(hipStream_t) stream, n, scal, ia, p, ap, u, r, s, part);
	LAUNCH_CHECK("k_fcg_step");
	return SSX_OK;
}

// ---------------------------------------------------------------------
// Coarsest level over unique nodes (ssx_coarse.c)
// ---------------------------------------------------------------------

__global__ void k_gather_u(int nu, const int *__restrict__ rep,
		const double *__restrict__ e, double *__restrict__ u)
{
	GRID_STRIDE(q, (size_t) nu) u[q] = e[rep[q]];
}

__global__ void k_scatter_u(size_t ntot, const int *__restrict__ cmap,
		const double *__restrict__ u, double *__restrict__ e)
{
	GRID_STRIDE(q, ntot) e[q] = u[cmap[q]];
}

// CSR_LPR lanes per row: pn = first ? pold : mask ((z - c) + beta pold) at
// the row and, recomputed with the same arithmetic, at every column read
// (the k_cg_pupdate formula); ap = mask (A pn); part = block sum of pn ap
#define CSR_LPR 8
__device__ __forceinline__ double csr_pn(const double *pold,
		const double *z, const double *mask, int j, int first, double b,
		double cc, int ic)
{
	double pj = pold[j];
	if (!first) {
		const double zv = ic >= 0 ? cc * 1.0 + z[j] : z[j];
		pj = b * pj + zv;
		if (mask)
			pj = pj * mask[j];
	}
	return pj;
}

// csr_pn on values already loaded (pold, z and mask at the column)
__device__ __forceinline__ double csr_pv(double pj, double zj, double mj,
		bool has_mask, int first, double b, double cc, int ic)
{
	if (!first) {
		const double zv = ic >= 0 ? cc * 1.0 + zj : zj;
		pj = b * pj + zv;
		if (has_mask)
			pj = pj * mj;
	}
	return pj;
}

__global__ void __launch_bounds__(256)
k_csr_pupd(int nu, const int *__restrict__ rowptr,
		const int *__restrict__ col, const double *__restrict__ val,
		const double *__restrict__ pold, double *__restrict__ pnew,
		const double *__restrict__ z, const double *__restrict__ mask,
		const double *__restrict__ scal, int ib, int ic, int first,
		double *__restrict__ ap, double *__restrict__ part,
		const int *__restrict__ done)
{
	if (done && *done)
		return;
	__shared__ double sh[256];
	const double b = first ? 0.0 : scal[ib];
	const double cc = (!first && ic >= 0) ? -scal[ic] : 0.0;
	const int lane = threadIdx.x % CSR_LPR;
	const int u = (blockIdx.x * 256 + threadIdx.x) / CSR_LPR;
	double pd = 0.0;
	double y = 0.0;
	if (u < nu)
		for (int k = rowptr[u] + lane; k < rowptr[u + 1]; k += CSR_LPR)
			y += val[k] * csr_pn(pold, z, mask, col[k], first, b, cc, ic);
	// lanes of a row are adjacent in the wavefront: fixed-order fold
	for (int o = 1; o < CSR_LPR; o <<= 1)
		y += __shfl_xor(y, o, CSR_LPR);
	if (u < nu && lane == 0) {
		const double pu = csr_pn(pold, z, mask, u, first, b, cc, ic);
		pnew[u] = pu;
		if (mask && mask[u] == 0.0)
			y = 0.0;
		ap[u] = y;
		pd = pu * y;
	}
	if (part) {
		const double t = block_sum256(pd, sh);
		if (threadIdx.x == 0)
			part[blockIdx.x] = t;
	}
}

extern "C" int ssxk_csr_blocks(int nu)
{
	return (int) (((long long) nu * CSR_LPR + 255) / 256);
}

extern "C" int ssxk_gather_u(void *stream, int nu, const int *rep,
		const double *e, double *u)
{
	hipLaunchKernelGGL(k_gather_u, vgrid((size_t) nu), dim3(256), 0,
			(hipStream_t) stream, nu, rep, e, u);
	LAUNCH_CHECK("k_gather_u");
	return SSX_OK;
}

extern "C" int ssxk_scatter_u(void *stream, size_t ntot, const int *cmap,
		const double *u, double *e)
{
	hipLaunchKernelGGL(k_scatter_u, vgrid(ntot), dim3(256), 0,
			(hipStream_t) stream, ntot, cmap, u, e);
	LAUNCH_CHECK("k_scatter_u");
	return SSX_OK;
}

// ---------------------------------------------------------------------
// Dense coarse operator and its inverse (the direct coarsest solve)
// ---------------------------------------------------------------------

// A (nu x nu, row i at A + i*nu) = the CSR operator with masked rows and
// columns replaced by the identity; reg > 0 adds reg to every entry (the
// rank-one term that makes a singular pressure operator definite)
__global__ void __launch_bounds__(256)
k_csr_dense(int nu, const int *__restrict__ rowptr,
		const int *__restrict__ col, const double *__restrict__ val,
		const double *__restrict__ mask, double reg, double *__restrict__ A)
{
	const int i = blockIdx.x;
	double *row = A + (size_t) i * nu;
	for (int j = threadIdx.x; j < nu; j += 256)
		row[j] = reg;
	__syncthreads();
	const bool mi = mask && mask[i] == 0.0;
	if (mi) {
		if (threadIdx.x == 0)
			row[i] = 1.0;
		return;
	}
	for (int k = rowptr[i] + threadIdx.x; k < rowptr[i + 1]; k += 256) {
		const int j = col[k];
		if (!(mask && mask[j] == 0.0))
			row[j] = val[k] + reg;
	}
}

// the lower triangle from the upper one: rocSOLVER's potri leaves the
// inverse in the lower triangle of the column-major matrix, which is the
// upper triangle of this row-major one
__global__ void __launch_bounds__(256)
k_dense_sym(int nu, double *__restrict__ A)
{
	const int i = blockIdx.x;
	for (int j = i + 1 + threadIdx.x; j < nu; j += 256)
		A[(size_t) j * nu + i] = A[(size_t) i * nu + j];
}

// y = A x, one workgroup per row, a fixed tree per row (deterministic)
__global__ void __launch_bounds__(256)
k_dense_mv(int nu, const double *__restrict__ A, const double *__restrict__ x,
		double *__restrict__ y)
{
	__shared__ double sh[256];
	const double *row = A + (size_t) blockIdx.x * nu;
	double s = 0.0;
	for (int j = threadIdx.x; j < nu; j += 256)
		s += row[j] * x[j];
	const double t = block_sum256(s, sh);
	if (threadIdx.x == 0)
		y[blockIdx.x] = t;
}

extern "C" int ssxk_csr_dense(void *stream, int nu, const int *rowptr,
		const int *col, const double *val, const double *mask, double reg,
		double *A)
{
	hipLaunchKernelGGL(k_csr_dense, dim3(nu), dim3(256), 0,
			(hipStream_t) stream, nu, rowptr, col, val, mask, reg, A);
	LAUNCH_CHECK("k_csr_dense");
	return SSX_OK;
}

extern "C" int ssxk_dense_sym(void *stream, int nu, double *A)
{
	hipLaunchKernelGGL(k_dense_sym, dim3(nu), dim3(256), 0,
			(hipStream_t) stream, nu, A);
	LAUNCH_CHECK("k_dense_sym");
	return SSX_OK;
}

extern "C" int ssxk_dense_mv(void *stream, int nu, const double *A,
		const double *x, double *y)
{
	hipLaunchKernelGGL(k_dense_mv, dim3(nu), dim3(256), 0,
			(hipStream_t) stream, nu, A, x, y);
	LAUNCH_CHECK("k_dense_mv");
	return SSX_OK;
}

extern "C" int ssxk_csr_pupd(void *stream, int nu, const int *rowptr,
		const int *col, const double *val, const double *pold, double *pnew,
		const double *z, const double *mask, const double *scal, int ib,
		int ic, int first, double *ap, double *part, const int *done)
{
	hipLaunchKernelGGL(k_csr_pupd, dim3(ssxk_csr_blocks(nu)), dim3(256), 0,
			(hipStream_t) stream, nu, rowptr, col, val, pold, pnew, z, mask,
			scal, ib, ic, first, ap, part, done);
	LAUNCH_CHECK("k_csr_pupd");
	return SSX_OK;
}

// ---------------------------------------------------------------------
// Two-launch coarse CG: each kernel replays the single-workgroup
// reduction it depends on (same partials, same order, same scalar
// arithmetic as k_cg_reduce), so the iteration is bit-identical to the
// four-launch one with half the launches.
// ---------------------------------------------------------------------

// the four mode-1 sums (rz, rr, zs, rs) of k_cg_reduce in a 256-lane
// workgroup: lane t folds virtual lanes t, t+256, t+512, t+768 as the first
// two tree levels do, then the same tree
__device__ void sum4_as_1024(const double *__restrict__ part, long long m,
		int nq, double out[4], double (*sh)[256])
{
	// each virtual lane's first partial of every channel is loaded before
	// any add, so the loads' latencies overlap; the adds keep index order
	double x[4][4];
#pragma unroll
	for (int c = 0; c < 4; ++c)
#pragma unroll
		for (int v = 0; v < 4; ++v) {
			// branch-free: an unused slot loads part[0] and discards it
			const long long q = threadIdx.x + 256 * v;
			const bool use = c < nq && q < m;
			const double t = part[use ? (size_t) c * SSX_NPART + q : 0];
			x[c][v] = use ? t : 0.0;
		}
#pragma unroll
	for (int c = 0; c < 4; ++c) {
		double a[4];
#pragma unroll
		for (int v = 0; v < 4; ++v) {
			double acc = 0.0;
			const long long q0 = threadIdx.x + 256 * v;
			if (c < nq && q0 < m) {
				acc += x[c][v];
				for (long long q = q0 + 1024; q < m; q += 1024)
					acc += part[(size_t) c * SSX_NPART + q];
			}
			a[v] = acc;
		}
		a[0] += a[2];
		a[1] += a[3];
		a[0] += a[1];
		sh[c][threadIdx.x] = a[0];
	}
	__syncthreads();
	for (int st = 128; st > 0; st >>= 1) {
		if ((int) threadIdx.x < st)
#pragma unroll
			for (int c = 0; c < 4; ++c)
				sh[c][threadIdx.x] += sh[c][threadIdx.x + st];
		__syncthreads();
	}
#pragma unroll
	for (int c = 0; c < 4; ++c)
		out[c] = sh[c][0];
	__syncthreads();
}

// the mode-0 sum of k_cg_reduce (m partials) in a 256-lane workgroup
__device__ double sum1_as_1024(const double *__restrict__ part, long long m,
		double *sh)
{
	// the first two partials of each virtual lane are loaded before any
	// add (latencies overlap); the adds keep index order
	double x[4][2];
#pragma unroll
	for (int v = 0; v < 4; ++v)
#pragma unroll
		for (int k = 0; k < 2; ++k) {
			// branch-free: an unused slot loads part[0] and discards it
			const long long q = threadIdx.x + 256 * v + 1024 * k;
			const double t = part[q < m ? q : 0];
			x[v][k] = q < m ? t : 0.0;
		}
	double a[4];
#pragma unroll
	for (int v = 0; v < 4; ++v) {
		double acc = 0.0;
		const long long q0 = threadIdx.x + 256 * v;
#pragma unroll
		for (int k = 0; k < 2; ++k)
			if (q0 + 1024 * k < m)
				acc += x[v][k];
		for (long long q = q0 + 2048; q < m; q += 1024)
			acc += part[q];
		a[v] = acc;
	}
	a[0] += a[2];
	a[1] += a[3];
	a[0] += a[1];
	sh[threadIdx.x] = a[0];
	__syncthreads();
	for (int st = 128; st > 0; st >>= 1) {
		if ((int) threadIdx.x < st)
			sh[threadIdx.x] += sh[threadIdx.x + st];
		__syncthreads();
	}
	const double t = sh[0];
	__syncthreads();
	return t;
}

// the row entries a lane holds in registers across the reduction replay
// (rows of up to CSR_LPR * CSR_PRE entries; longer rows load the rest after)
#define CSR_PRE 4

__global__ void __launch_bounds__(256)
k_csr_cg_spmv(int nu, const int *__restrict__ rowptr,
		const int *__restrict__ col, const double *__restrict__ val,
		const double *__restrict__ pold, double *__restrict__ pnew,
		const double *__restrict__ z, const double *__restrict__ mask,
		double *__restrict__ scal, int o, int j, double eps, int maxit,
		double nunique, int pressure, double *__restrict__ ap,
		double *__restrict__ ps, const double *__restrict__ pu, int nbu,
		int *__restrict__ flags, int check)
{
	if (flags[0])
		return;
	__shared__ double s4[4][256];
	const int first = j == 0;
	const int lane = threadIdx.x % CSR_LPR;
	const int u = (blockIdx.x * 256 + threadIdx.x) / CSR_LPR;
	// Every load that does not depend on beta goes out before the
	// reduction replay, so the two latency chains overlap: the lane's
	// first CSR_PRE entries with their column's pold, z and mask, and the
	// row's own.
	int kb = 0, ke = 0;
	double va[CSR_PRE], pa[CSR_PRE], za[CSR_PRE], ma[CSR_PRE];
	double po = 0.0, zo = 0.0, mo = 1.0;
	if (!check && u < nu) {
		kb = rowptr[u] + lane;
		ke = rowptr[u + 1];
#pragma unroll
		for (int i = 0; i < CSR_PRE; ++i) {
			const int k = kb + i * CSR_LPR;
			va[i] = 0.0;
			pa[i] = 0.0;
			za[i] = 0.0;
			ma[i] = 1.0;
			if (k < ke) {
				const int c = col[k];
				va[i] = val[k];
				pa[i] = pold[c];
				if (!first) {
					za[i] = z[c];
					if (mask)
						ma[i] = mask[c];
				}
			}
		}
		if (lane == 0) {
			po = pold[u];
			if (!first)
				zo = z[u];
			if (mask)
				mo = mask[u];
		}
	}
	double b = 0.0, cc = 0.0;
	if (j > 0) {
		// cg_scalars1 of iteration j-1, in every workgroup
		double v[4];
		sum4_as_1024(pu, nbu, pressure ? 4 : 2, v, s4);
		double c = 0.0, rz = v[0];
		if (pressure) {
			c = v[2] / nunique;
			rz = rz - c * v[3];
		}
		const double beta = rz / scal[o + ((j - 1) & 1 ? CG_RDZ1 : CG_RDZ)];
		const double tol = eps >= 0.0 ? eps : scal[o + CG_TOL];
		const bool stop = sqrt(v[1]) <= tol || j >= maxit || !(v[1] == v[1]);
		if (blockIdx.x == 0 && threadIdx.x == 0) {
			scal[o + CG_NULLC] = c;
			scal[o + CG_BETA] = beta;
			scal[o + ((j & 1) ? CG_RDZ1 : CG_RDZ)] = rz;
			scal[o + CG_RR] = v[1];
			flags[1] = j;
			if (stop)
				flags[0] = 1;
		}
		if (stop)
			return;
		b = beta;
		cc = pressure ? -c : 0.0;
	}
	if (check)
		return;
	const int ic = pressure ? 0 : -1;      // csr_pn: null-space shift on
	double pd = 0.0;
	double y = 0.0;
	if (u < nu) {
#pragma unroll
		for (int i = 0; i < CSR_PRE; ++i)
			if (kb + i * CSR_LPR < ke)
				y += va[i] * csr_pv(pa[i], za[i], ma[i], mask != nullptr,
						first, b, cc, ic);
		for (int k = kb + CSR_PRE * CSR_LPR; k < ke; k += CSR_LPR)
			y += val[k] * csr_pn(pold, z, mask, col[k], first, b, cc, ic);
	}
	for (int w = 1; w < CSR_LPR; w <<= 1)
		y += __shfl_xor(y, w, CSR_LPR);
	if (u < nu && lane == 0) {
		const double pv = csr_pv(po, zo, mo, mask != nullptr, first, b, cc,
				ic);
		pnew[u] = pv;
		if (mask && mo == 0.0)
			y = 0.0;
		ap[u] = y;
		pd = pv * y;
	}
	const double t = block_sum256(pd, s4[0]);
	if (threadIdx.x == 0)
		ps[blockIdx.x] = t;
}

__global__ void __launch_bounds__(256)
k_csr_cg_update(size_t n, double *__restrict__ scal, int o, int j,
		const double *__restrict__ p, const double *__restrict__ ap,
		double *__restrict__ u, double *__restrict__ r, double *__restrict__ z,
		const double *__restrict__ dinv, const double *__restrict__ scale,
		int pressure, const double *__restrict__ ps, int nbs,
		double *__restrict__ pu, const int *__restrict__ flags)
{
	if (flags[0])
		return;
	__shared__ double sh[4][256];
	// the lane's first node is loaded before the reduction replay (the
	// loads do not depend on alpha), the rest of its nodes after
	const size_t q0 = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
	const size_t gs = (size_t) gridDim.x * blockDim.x;
	double p0 = 0.0, ap0 = 0.0, u0 = 0.0, r0 = 0.0, d0 = 0.0, w0 = 0.0;
	if (q0 < n) {
		p0 = p[q0];
		ap0 = ap[q0];
		u0 = u[q0];
		r0 = r[q0];
		d0 = dinv[q0];
		w0 = scale[q0];
	}
	// cg_scalars0, in every workgroup
	const double pap = sum1_as_1024(ps, nbs, sh[0]);
	const double a = scal[o + ((j & 1) ? CG_RDZ1 : CG_RDZ)] / pap;
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		scal[o + CG_PAP] = pap;
		scal[o + CG_ALPHA] = a;
	}
	double acc[4] = { 0.0, 0.0, 0.0, 0.0 };
	for (size_t q = q0; q < n; q += gs) {
		const bool pre = q == q0;
		const double pq = pre ? p0 : p[q];
		const double apq = pre ? ap0 : ap[q];
		const double uq = pre ? u0 : u[q];
		const double rq = pre ? r0 : r[q];
		const double dq = pre ? d0 : dinv[q];
		const double w = pre ? w0 : scale[q];
		u[q] = a * pq + uq;
		const double rv = -a * apq + rq;
		r[q] = rv;
		const double zv = rv * dq;
		z[q] = zv;
		acc[0] += (double) (rv * w) * (double) zv;
		acc[1] += (double) (rv * w) * (double) rv;
		if (pressure) {
			acc[2] += (double) (zv * w);
			acc[3] += (double) (rv * w);
		}
	}
	block_sum4(acc, sh);
	if (threadIdx.x < 4)
		pu[(size_t) threadIdx.x * SSX_NPART + blockIdx.x] = sh[threadIdx.x][0];
}

extern "C" int ssxk_csr_cg_spmv(void *stream, int nu, const int *rowptr,
		const int *col, const double *val, const double *pold, double *pnew,
		const double *z, const double *mask, double *scal, int o, int j,
		double eps, int maxit, double nunique, int pressure, double *ap,
		double *ps, const double *pu, int nbu, int *flags, int check)
{
	if (nbu < 1 || nbu > SSX_NPART) {
		ssx_set_error("ssxk_csr_cg_spmv: %d update partials", nbu);
		return SSX_EINVAL;
	}
	hipLaunchKernelGGL(k_csr_cg_spmv,
			dim3(check ? 1 : ssxk_csr_blocks(nu)), dim3(256), 0,
			(hipStream_t) stream, nu, rowptr, col, val, pold, pnew, z, mask,
			scal, o, j, eps, maxit, nunique, pressure, ap, ps, pu, nbu, flags,
			check);
	LAUNCH_CHECK("k_csr_cg_spmv");
	return SSX_OK;
}

extern "C" int ssxk_csr_cg_update(void *stream, int nu, double *scal, int o,
		int j, const double *p, const double *ap, double *u, double *r,
		double *z, const double *dinv, const double *scale, int pressure,
		const double *ps, int nbs, double *pu, int nbu, const int *flags)
{
	if (nbu < 1 || nbu > SSX_NPART) {
		ssx_set_error("ssxk_csr_cg_update: %d workgroups", nbu);
		return SSX_EINVAL;
	}
	hipLaunchKernelGGL(k_csr_cg_update, dim3(nbu), dim3(256), 0,
			(hipStream_t) stream, (size_t) nu, scal, o, j, p, ap, u, r, z,
			dinv, scale, pressure, ps, nbs, pu, flags);
	LAUNCH_CHECK("k_csr_cg_update");
	return SSX_OK;
}

// scalar program, one lane:
//   op 0: s[a] = s[b] / s[c]        op 1: s[a] = v
//   op 2: s[a] = s[b]               op 3: s[a] = -s[b]
__global__ void k_scalar(double *scal, int op, int a, int b, int c, double v)
{
	if (threadIdx.x != 0)
		return;
	switch (op) {
	case 0: scal[a] = scal[b] / scal[c]; break;
	case 1: scal[a] = v; break;
	case 2: scal[a] = scal[b]; break;
	case 3: scal[a] = -scal[b]; break;
	}
}

extern "C" int ssxk_scalar(void *stream, double *scal, int op, int a, int b,
		int c, double v)
{
	hipLaunchKernelGGL(k_scalar, dim3(1), dim3(64), 0, (hipStream_t) stream,
			scal, op, a, b, c, v);
	LAUNCH_CHECK("k_scalar");
	return SSX_OK;
}

// ---------------------------------------------------------------------
// Vector kernels
// ---------------------------------------------------------------------

template <typename T>
__global__ void k_axpby(size_t n, T a, const T *__restrict__ x,
		const T *__restrict__ w, T b, T *__restrict__ y)
{
	GRID_STRIDE(q, n) {
		T xv = w ? x[q] * w[q] : x[q];
		y[q] = a * xv + b * y[q];
	}
}

extern "C" int ssxk_axpby(void *stream, size_t n, double a, const double *x,
		const double *w, double b, double *y)
{
	hipLaunchKernelGGL(k_axpby<double>, vgrid(n), dim3(256), 0,
			(hipStream_t) stream, n, a, x, w, b, y);
	LAUNCH_CHECK("k_axpby");
	return SSX_OK;
}

extern "C" int ssxk_axpby_lp(void *stream, size_t n, double a, const void *x,
		const void *w, double b, void *y)
{
	hipLaunchKernelGGL(k_axpby<float>, vgrid(n), dim3(256), 0,
			(hipStream_t) stream, n, (float) a, (const float *) x,
			(const float *) w, (float) b, (float *) y);
	LAUNCH_CHECK("k_axpby");
	return SSX_OK;
}

template <typename T>
__global__ void k_mul(size_t n, const T *__restrict__ x,
		const T *__restrict__ w, T *__restrict__ z)
{
	GRID_STRIDE(q, n) z[q] = x[q] * w[q];
}

extern "C" int ssxk_mul(void *stream, size_t n, const double *x,
		const double *w, double *z)
{
	hipLaunchKernelGGL(k_mul<double>, vgrid(n), dim3(256), 0,
			(hipStream_t) stream, n, x, w, z);
	LAUNCH_CHECK("k_mul");
	return SSX_OK;
}

extern "C" int ssxk_mul_lp(void *stream, size_t n, const void *x,
		const void *w, void *z)
{
	hipLaunchKernelGGL(k_mul<float>, vgrid(n), dim3(256), 0,
			(hipStream_t) stream, n, (const float *) x, (const float *) w,
			(float *) z);
	LAUNCH_CHECK("k_mul");
	return SSX_OK;
}

__global__ void k_axpy_s(size_t n, const double *__restrict__ scal, int ia,
		double sgn, const double *__restrict__ x, double *__restrict__ y)
{
	const double a = sgn * scal[ia];
	GRID_STRIDE(q, n) y[q] = a * x[q] + y[q];
}

extern "C" int ssxk_axpy_s(void *stream, size_t n, const double *scal,
		int ia, double sgn, const double *x, double *y)
{
	hipLaunchKernelGGL(k_axpy_s, vgrid(n), dim3(256), 0,
			(hipStream_t) stream, n, scal, ia, sgn, x, y);
	LAUNCH_CHECK("k_axpy_s");
	return SSX_OK;
}

__global__ void k_xpby_s(size_t n, const double *__restrict__ x,
		const double *__restrict__ scal, int ib, double *__restrict__ y,
		const double *__restrict__ mask)
{
	const double b = scal[ib];
	GRID_STRIDE(q, n) {
		double v = x[q] + b * y[q];
		y[q] = mask ? v * mask[q] : v;
	}
}

extern "C" int ssxk_xpby_s(void *stream, size_t n, const double *x,
		const double *scal, int ib, double *y, const double *mask)
{
	hipLaunchKernelGGL(k_xpby_s, vgrid(n), dim3(256), 0,
			(hipStream_t) stream, n, x, scal, ib, y, mask);
	LAUNCH_CHECK("k_xpby_s");
	return SSX_OK;
}

template <typename T>
__global__ void k_shift_s(size_t n, const double *__restrict__ scal, int ia,
		double nunique, T *__restrict__ y)
{
	const T c = (T) (-scal[ia] / nunique);
	GRID_STRIDE(q, n) y[q] = c * T(1) + y[q];
}

extern "C" int ssxk_shift_s(void *stream, size_t n, const double *scal,
		int ia, double nunique, double *y)
{
	hipLaunchKernelGGL(k_shift_s<double>, vgrid(n), dim3(256), 0,
			(hipStream_t) stream, n, scal, ia, nunique, y);
	LAUNCH_CHECK("k_shift_s");
	return SSX_OK;
}

extern "C" int ssxk_shift_s_lp(void *stream, size_t n, const double *scal,
		int ia, double nunique, void *y)
{
	hipLaunchKernelGGL(k_shift_s<float>, vgrid(n), dim3(256), 0,
			(hipStream_t) stream, n, scal, ia, nunique, (float *) y);
	LAUNCH_CHECK("k_shift_s");
	return SSX_OK;
}

__global__ void k_set_ints(int *dst, int v, int n)
{
	if ((int) threadIdx.x < n)
		dst[threadIdx.x] = v;
}

extern "C" int ssxk_set_ints(void *stream, int *dst, int v, int n)
{
	hipLaunchKernelGGL(k_set_ints, dim3(1), dim3(64), 0, (hipStream_t) stream,
			dst, v, n);
	LAUNCH_CHECK("k_set_ints");
	return SSX_OK;
}

template <typename T>
__global__ void k_fill(size_t n, T v, T *__restrict__ y)
{
	GRID_STRIDE(q, n) y[q] = v;
}

extern "C" int ssxk_fill(void *stream, double *dst, double v, size_t n)
{
	hipLaunchKernelGGL(k_fill<double>, vgrid(n), dim3(256), 0,
			(hipStream_t) stream, n, v, dst);
	LAUNCH_CHECK("k_fill");
	return SSX_OK;
}

extern "C" int ssxk_fill_lp(void *stream, void *dst, double v, size_t n)
{
	hipLaunchKernelGGL(k_fill<float>, vgrid(n), dim3(256), 0,
			(hipStream_t) stream, n, (float) v, (float *) dst);
	LAUNCH_CHECK("k_fill");
	return SSX_OK;
}

extern "C" int ssxk_copy(void *stream, double *dst, const double *src,
		size_t n)
{
	hipError_t e = hipMemcpyAsync(dst, src, n * sizeof(double),
			hipMemcpyDeviceToDevice, (hipStream_t) stream);
	if (e != hipSuccess) {
		ssx_set_error("copy: %s", hipGetErrorString(e));
		return SSX_EHIP;
	}
	return SSX_OK;
}

extern "C" int ssxk_copy_lp(void *stream, void *dst, const void *src, size_t n)
{
	hipError_t e = hipMemcpyAsync(dst, src, n * sizeof(float),
			hipMemcpyDeviceToDevice, (hipStream_t) stream);
	if (e != hipSuccess) {
		ssx_set_error("copy: %s", hipGetErrorString(e));
		return SSX_EHIP;
	}
	return SSX_OK;
}

// precision conversion: FP64 -> FP32 (round to nearest), FP32 -> FP64
// (exact), FP64 -> bfloat16 (via FP32, round to nearest even)
__global__ void k_d2f(size_t n, const double *__restrict__ x,
		float *__restrict__ y)
{
	GRID_STRIDE(q, n) y[q] = (float) x[q];
}

__global__ void k_f2d(size_t n, const float *__restrict__ x,
		double *__restrict__ y)
{
	GRID_STRIDE(q, n) y[q] = (double) x[q];
}

__global__ void k_d2bf(size_t n, const double *__restrict__ x,
		bf16_t *__restrict__ y)
{
	GRID_STRIDE(q, n) {
		unsigned b = __float_as_uint((float) x[q]);
		b += 0x7fffu + ((b >> 16) & 1u);
		y[q] = (bf16_t) (b >> 16);
	}
}

// The mixed-precision V-cycle's ends in one pass each.
//   k_d2f_jac:  y = (float) x; e = y * dinv  (k_d2f, then k_mul)
//   k_f2d_avg_g: y = DSS-average of (double) e over the copies of each
//               shared node, as a gather over the per-copy partner lists:
//               own + partners in the copy's add order, times 1/M (k_f2d,
//               then the averaging group kernel, bit for bit)
__global__ void k_d2f_jac(size_t n, const double *__restrict__ x,
		float *__restrict__ y, const float *__restrict__ dinv,
		float *__restrict__ e)
{
	GRID_STRIDE(q, n) {
		const float v = (float) x[q];
		y[q] = v;
		e[q] = v * dinv[q];
	}
}

__global__ void __launch_bounds__(256)
k_f2d_avg_g(unsigned n, unsigned n3, int P, const int *__restrict__ soff,
		const int *__restrict__ tab, const float *__restrict__ e,
		double *__restrict__ y)
{
	constexpr int U = 2;
	const unsigned stride = gridDim.x * 256u;
	for (unsigned q0 = xcd_block() * 256u + threadIdx.x; q0 < n;
			q0 += U * stride) {
		int ix[U][7];
		float own[U], pv[U][7];
#pragma unroll
		for (int v = 0; v < U; ++v) {
			const unsigned q = q0 + v * stride;
			own[v] = 0.0f;
			if (q < n) {
				const unsigned el = q / n3;
				const int c = soff[q - el * n3];
				own[v] = e[q];
				const int *pt = tab + (size_t) el * P + (c >> 3);
#pragma unroll
				for (int k = 0; k < 7; ++k)
					ix[v][k] = k < (c & 7) ? pt[k] : -1;
			} else {
#pragma unroll
				for (int k = 0; k < 7; ++k)
					ix[v][k] = -1;
			}
		}
		// branch-free: an unused slot loads e[0] and discards it (a
		// conditional load here was compiled into one wait per load)
#pragma unroll
		for (int v = 0; v < U; ++v)
#pragma unroll
			for (int k = 0; k < 7; ++k) {
				const int j = ix[v][k];
				const float t = e[j >= 0 ? j : 0];
				pv[v][k] = j >= 0 ? t : 0.0f;
			}
#pragma unroll
		for (int v = 0; v < U; ++v) {
			const unsigned q = q0 + v * stride;
			if (q >= n)
				break;
			double s = (double) own[v];
			int m = 1;
#pragma unroll
			for (int k = 0; k < 7; ++k)
				if (ix[v][k] >= 0) {
					s = s + (double) pv[v][k];
					++m;
				}
			y[q] = m > 1 ? s * (1.0 / (double) m) : s;
		}
	}
}

extern "C" int ssxk_d2f_jacobi(void *stream, size_t n, const double *x,
		void *y, const void *dinv, void *e)
{
	hipLaunchKernelGGL(k_d2f_jac, vgrid(n), dim3(256), 0,
			(hipStream_t) stream, n, x, (float *) y, (const float *) dinv,
			(float *) e);
	LAUNCH_CHECK("k_d2f_jac");
	return SSX_OK;
}

extern "C" int ssxk_f2d_average_g(void *stream, size_t n, int n3, int P,
		const int *soff, const int *tab, const void *e, double *y)
{
	if (n >= (1ull << 31) || n % (size_t) n3) {
		ssx_set_error("k_f2d_avg_g: %zu nodes, %d per element", n, n3);
		return SSX_EINVAL;
	}
	hipLaunchKernelGGL(k_f2d_avg_g, dim3(SSX_NPART), dim3(256), 0,
			(hipStream_t) stream, (unsigned) n, (unsigned) n3, P, soff, tab,
			(const float *) e, y);
	LAUNCH_CHECK("k_f2d_avg_g");
	return SSX_OK;
}

extern "C" int ssxk_cvt(void *stream, int from, int to, const void *x,
		void *y, size_t n)
{
	hipStream_t st = (hipStream_t) stream;
	if (from == SSX_PREC_FP64 && to == SSX_PREC_FP32)
		hipLaunchKernelGGL(k_d2f, vgrid(n), dim3(256), 0, st, n,
				(const double *) x, (float *) y);
	else if (from == SSX_PREC_FP32 && to == SSX_PREC_FP64)
		hipLaunchKernelGGL(k_f2d, vgrid(n), dim3(256), 0, st, n,
				(const float *) x, (double *) y);
	else if (from == SSX_PREC_FP64 && to == SSX_PREC_BF16)
		hipLaunchKernelGGL(k_d2bf, vgrid(n), dim3(256), 0, st, n,
				(const double *) x, (bf16_t *) y);
	else {
		ssx_set_error("ssxk_cvt: %d -> %d unsupported", from, to);
		return SSX_EINVAL;
	}
	LAUNCH_CHECK("k_cvt");
	return SSX_OK;
}
