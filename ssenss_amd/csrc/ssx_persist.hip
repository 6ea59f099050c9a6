// ssx_persist.hip -- the whole Jacobi-PCG solve of one small multigrid level
// (conjugate_gradient, solver.c:33-81) in a single cooperative launch.
//
// On the coarse levels of the p-multigrid V-cycle (n = 2: 8 nodes per
// element) each CG iteration moves a few MB that sit in L2 / Infinity
// Cache, so six dependent launches per iteration cost far more in dispatch
// and ramp-up than in work.  Here every block of a co-resident grid runs
// the iteration loop and the phases are separated by grid-wide barriers:
//
//   A  p = (z - c) + beta p (masked), Ap_e = A_e p per element, mask,
//      sum of p Ap_e over the local copies (= <p, DSS(Ap)>_w) -> part[blk]
//   -- barrier
//   B  DSS of Ap over the shared-node groups (the reference's add order)
//   -- barrier
//   C  every block sums the G partials in the same fixed order (so all
//      blocks hold the same alpha), u += alpha p, r -= alpha Ap, z = D^-1 r,
//      partials of <r,z>_w, <r,r>_w (and <z,1>_w, <r,1>_w for the pressure
//      null space)                                      -> part[2G + 4 blk]
//   -- barrier
//   D  every block sums those partials in the same order: null-space
//      constant, rdz, beta, the stopping test sqrt(<r,r>_w) <= tol; all
//      blocks take the same decision.
//
// The p update of A is done per element by the thread that then applies
// the operator to it, so no barrier separates it from the product.
// Barrier waits are bounded: a block that waits too long raises the abort
// flag, every block leaves at its next barrier, and the host reports an
// error.  The launch is cooperative, so the runtime refuses a grid that
// cannot be co-resident instead of letting it wait forever.
#include <hip/hip_runtime.h>
#include "ssx_internal.h"
#include "ssx_device.h"

#define PB 256                      // threads per block
#define SPIN_LIMIT (1u << 22)       // barrier polls before giving up

struct PersistArgs {
	int nel, op, apply_mask, pressure, maxit;
	double eps;                     // < 0: tolerance in scal[o + CG_TOL]
	double lambda, nunique;
	const double *gf, *mass, *dmat, *dinv, *scale, *mask;
	const int *bc;
	const int *g2, *g4, *g8;
	long long G2, G4, G8;
	double *u, *r, *z, *p, *ap;
	double *part, *scal;
	int o;
	int *flags;                     // [0] done, [1] iterations
	unsigned *bar;                  // [0] arrivals, [1] generation, [2] abort
};

// Grid barrier, one thread per block doing the handshake.  __syncthreads
// completes the block's stores into this XCD's L2; the agent-scope release
// fence then writes L2 back so other XCDs can see them, and the acquire
// fence after the wait invalidates this CU's L1 and the stale part of L2.
// The wait itself polls with relaxed loads (an acquire per poll would
// invalidate the caches on every poll).  Returns false when the grid is
// aborting.
__device__ bool grid_sync(unsigned *bar, unsigned nblocks, unsigned &phase)
{
	__shared__ int s_ok;
	__syncthreads();
	if (threadIdx.x == 0) {
		int ok = 1;
		const unsigned my = phase;
		__threadfence();
		const unsigned prev = atomicAdd(&bar[0], 1u);
		if (prev == nblocks - 1) {
			atomicExch(&bar[0], 0u);
			__threadfence();
			atomicAdd(&bar[1], 1u);
		} else {
			unsigned n = 0;
			while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED,
					__HIP_MEMORY_SCOPE_AGENT) == my) {
				if (__hip_atomic_load(&bar[2], __ATOMIC_RELAXED,
						__HIP_MEMORY_SCOPE_AGENT) || ++n > SPIN_LIMIT) {
					atomicExch(&bar[2], 1u);
					ok = 0;
					break;
				}
				__builtin_amdgcn_s_sleep(2);
			}
		}
		__threadfence();
		phase = my + 1;
		s_ok = ok && !__hip_atomic_load(&bar[2], __ATOMIC_RELAXED,
				__HIP_MEMORY_SCOPE_AGENT);
	}
	__syncthreads();
	return s_ok;
}

// sum of v over the block (fixed tree), valid in every thread
__device__ double block_sum(double v, double *sh)
{
	sh[threadIdx.x] = v;
	__syncthreads();
#pragma unroll
	for (int s = PB / 2; s > 0; s >>= 1) {
		if ((int) threadIdx.x < s)
			sh[threadIdx.x] += sh[threadIdx.x + s];
		__syncthreads();
	}
	const double t = sh[0];
	__syncthreads();
	return t;
}

// sum of part[0..cnt) in a fixed order: the same value in every block
__device__ double sum_parts(const double *part, int cnt, double *sh)
{
	double acc = 0.0;
	for (int q = threadIdx.x; q < cnt; q += PB)
		acc += __hip_atomic_load(part + q, __ATOMIC_RELAXED,
				__HIP_MEMORY_SCOPE_AGENT);
	return block_sum(acc, sh);
}

template <int N, int OP>
__global__ void __launch_bounds__(PB)
k_cg_persist(PersistArgs a)
{
	constexpr int N2 = N * N, N3 = N * N * N;
	constexpr int EPB = PB / N3;            // elements per block pass
	__shared__ double sh[PB];
	__shared__ double s_u[EPB * N3], s_t0[EPB * N3], s_t1[EPB * N3],
		s_t2[EPB * N3];
	__shared__ double s_d[2 * N2];
	const unsigned G = gridDim.x;
	unsigned phase = 0;
	for (int q = threadIdx.x; q < 2 * N2; q += PB)
		s_d[q] = a.dmat[q];
	__syncthreads();
	const double *Dn = s_d, *Ds = s_d + (OP == 1 ? N2 : 0);

	const int loc = threadIdx.x % N3, le = threadIdx.x / N3;
	const int i = loc % N, j = (loc / N) % N, k = loc / N2;
	const bool lane_ok = le < EPB;
	const int ngroups = (a.nel + EPB - 1) / EPB;
	const size_t nt = (size_t) a.nel * N3;
	double *partA = a.part, *partC = a.part + 2 * G;

	double rdz = a.scal[a.o + CG_RDZ];
	const double tol = a.eps >= 0.0 ? a.eps : a.scal[a.o + CG_TOL];
	double beta = 0.0, cnull = 0.0, rr = 0.0, pap = 0.0;
	int it = 0;
	bool ok = true;
	for (;;) {
		// ---- A: p update, element operator, unshared part of <p,Ap>
		double pd = 0.0;
		for (int g = blockIdx.x; g < ngroups; g += G) {
			const int e = g * EPB + le;
			const bool act = lane_ok && e < a.nel;
			const size_t q = (size_t) e * N3 + loc;
			const int eb = act ? a.bc[e] : 0;
			const int bits = a.apply_mask ? eb : 0;
			const bool m = ((bits & 1) && i == 0) || ((bits & 2) && i == N - 1)
				|| ((bits & 4) && j == 0) || ((bits & 8) && j == N - 1)
				|| ((bits & 16) && k == 0) || ((bits & 32) && k == N - 1);
			double pv = 0.0;
			if (act) {
				pv = a.p[q];
				if (it > 0) {
					const double zv = a.pressure ? -cnull * 1.0 + a.z[q]
						: a.z[q];
					pv = beta * pv + zv;
					if (a.mask)
						pv = pv * a.mask[q];
					a.p[q] = pv;
				}
			}
			if (lane_ok)
				s_u[le * N3 + loc] = pv;
			__syncthreads();
			double t0 = 0.0, t1 = 0.0, t2 = 0.0;
			if (act) {
				const double *su = s_u + le * N3;
				double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
				for (int l = 0; l < N; ++l) {
					gx += Ds[i * N + l] * su[(k * N + j) * N + l];
					gy += Dn[j * N + l] * su[(k * N + l) * N + i];
					gz += Dn[k * N + l] * su[(l * N + j) * N + i];
				}
				const double *gg = a.gf + (size_t) e * 6 * N3 + loc;
				const double gxx = gg[0], gxy = gg[N3], gxz = gg[2 * N3];
				const double gyy = gg[3 * N3], gyz = gg[4 * N3];
				const double gzz = gg[5 * N3];
				t0 = gxx * gx + gxy * gy + gxz * gz;
				t1 = gxy * gx + gyy * gy + gyz * gz;
				t2 = gxz * gx + gyz * gy + gzz * gz;
			}
			if (lane_ok) {
				s_t0[le * N3 + loc] = t0;
				s_t1[le * N3 + loc] = t1;
				s_t2[le * N3 + loc] = t2;
			}
			__syncthreads();
			if (act) {
				const double *a0 = s_t0 + le * N3, *a1 = s_t1 + le * N3;
				const double *a2 = s_t2 + le * N3;
				double v = 0.0;
#pragma unroll
				for (int l = 0; l < N; ++l)
					v += Dn[l * N + i] * a0[(k * N + j) * N + l]
						+ Ds[l * N + j] * a1[(k * N + l) * N + i];
#pragma unroll
				for (int l = 0; l < N; ++l)
					v += Ds[l * N + k] * a2[(l * N + j) * N + i];
				if (OP == 1)
					v += a.lambda * (a.mass[q] * pv);
				v = m ? 0.0 : v;
				a.ap[q] = v;
				pd += pv * v;       // all local copies: <p, DSS(Ap)>_w
			}
			__syncthreads();
		}
		double bs = block_sum(pd, sh);
		if (threadIdx.x == 0)
			partA[blockIdx.x] = bs;
		if (!(ok = grid_sync(a.bar, G, phase)))
			break;

		// ---- B: DSS of Ap
		const long long ng = a.G2 + a.G4 + a.G8;
		for (long long t = (long long) blockIdx.x * PB + threadIdx.x; t < ng;
				t += (long long) G * PB) {
			if (t < a.G2)
				dss_group<2, double>(a.G2, t, a.g2, a.ap, nullptr, 0);
			else if (t < a.G2 + a.G4)
				dss_group<4, double>(a.G4, t - a.G2, a.g4, a.ap, nullptr, 0);
			else
				dss_group<8, double>(a.G8, t - a.G2 - a.G4, a.g8, a.ap,
						nullptr, 0);
		}
		if (!(ok = grid_sync(a.bar, G, phase)))
			break;

		// ---- C: alpha, vector update, partials of the next inner products
		pap = sum_parts(a.part, G, sh);
		const double alpha = rdz / pap;
		double acc[4] = { 0.0, 0.0, 0.0, 0.0 };
		for (size_t q = (size_t) blockIdx.x * PB + threadIdx.x; q < nt;
				q += (size_t) G * PB) {
			a.u[q] = alpha * a.p[q] + a.u[q];
			const double rv = -alpha * a.ap[q] + a.r[q];
			a.r[q] = rv;
			const double zv = rv * a.dinv[q];
			a.z[q] = zv;
			const double w = a.scale[q];
			acc[0] += (rv * w) * zv;
			acc[1] += (rv * w) * rv;
			if (a.pressure) {
				acc[2] += zv * w;
				acc[3] += rv * w;
			}
		}
		for (int c = 0; c < 4; ++c) {
			bs = block_sum(acc[c], sh);
			if (threadIdx.x == 0)
				partC[(size_t) c * G + blockIdx.x] = bs;
		}
		if (!(ok = grid_sync(a.bar, G, phase)))
			break;

		// ---- D: beta and the stopping test, identical in every block
		double rz = sum_parts(partC, G, sh);
		rr = sum_parts(partC + G, G, sh);
		cnull = 0.0;
		if (a.pressure) {
			cnull = sum_parts(partC + 2 * G, G, sh) / a.nunique;
			rz = rz - cnull * sum_parts(partC + 3 * G, G, sh);
		}
		beta = rz / rdz;
		rdz = rz;
		++it;
		if (sqrt(rr) <= tol || it >= a.maxit || !(rr == rr))
			break;
		// partA/partB are rewritten in the next A/B phases only after every
		// block has passed D's reads: D reads partC, and A writes partA,
		// which C (not D) read -- C is behind a barrier already.
	}
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		a.scal[a.o + CG_RDZ] = rdz;
		a.scal[a.o + CG_RR] = rr;
		a.scal[a.o + CG_PAP] = pap;
		a.scal[a.o + CG_BETA] = beta;
		a.scal[a.o + CG_NULLC] = cnull;
		a.flags[1] = it;
		a.flags[0] = ok ? 1 : -1;
	}
}

// Cooperative launch of the persistent CG on one level.  grid: blocks (at
// most what the device keeps co-resident).  Returns SSX_EHIP if the launch
// is refused.  A barrier wait that gave up makes every block leave at its
// next barrier and leaves flags[0] == -1; the caller reads it and reports
// SSX_EHIP (u, r and p are then those of an unfinished iteration).
extern "C" int ssxk_cg_persist(void *stream, int n, int grid, int nel,
		int op, int apply_mask, int pressure, int maxit, double eps,
		double lambda, double nunique, const double *gf, const double *mass,
		const double *dmat, const double *dinv, const double *scale,
		const double *mask, const int *bc, int *const grp[3],
		const long long ng[3], double *u, double *r, double *z, double *p,
		double *ap, double *part, double *scal, int o, int *flags,
		unsigned *bar)
{
	PersistArgs a;
	a.nel = nel; a.op = op; a.apply_mask = apply_mask; a.pressure = pressure;
	a.maxit = maxit; a.eps = eps; a.lambda = lambda; a.nunique = nunique;
	a.gf = gf; a.mass = mass; a.dmat = dmat; a.dinv = dinv; a.scale = scale;
	a.mask = mask; a.bc = bc;
	a.g2 = grp[0]; a.g4 = grp[1]; a.g8 = grp[2];
	a.G2 = ng[0]; a.G4 = ng[1]; a.G8 = ng[2];
	a.u = u; a.r = r; a.z = z; a.p = p; a.ap = ap;
	a.part = part; a.scal = scal; a.o = o; a.flags = flags; a.bar = bar;
	void *args[] = { &a };
	const void *fn;
	switch (n * 2 + (op == SSX_OP_HELMHOLTZ)) {
#define CASE(NN) \
	case NN * 2: fn = (const void *) k_cg_persist<NN, 0>; break; \
	case NN * 2 + 1: fn = (const void *) k_cg_persist<NN, 1>; break;
	CASE(2) CASE(3) CASE(4)
#undef CASE
	default:
		ssx_set_error("persistent CG: n=%d unsupported", n);
		return SSX_EINVAL;
	}
	hipError_t e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(PB), args,
			0, (hipStream_t) stream);
	if (e != hipSuccess) {
		ssx_set_error("k_cg_persist (grid %d): %s", grid,
				hipGetErrorString(e));
		return SSX_EHIP;
	}
	return SSX_OK;
}

extern "C" int ssxk_cg_persist_max_grid(int n, int *grid)
{
	int dev, ncu = 0, per = 0;
	if (hipGetDevice(&dev) != hipSuccess
			|| hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
					dev) != hipSuccess) {
		ssx_set_error("persistent CG: device query failed");
		return SSX_EHIP;
	}
	const void *fn;
	switch (n) {
	case 2: fn = (const void *) k_cg_persist<2, 0>; break;
	case 3: fn = (const void *) k_cg_persist<3, 0>; break;
	case 4: fn = (const void *) k_cg_persist<4, 0>; break;
	default:
		ssx_set_error("persistent CG: n=%d unsupported", n);
		return SSX_EINVAL;
	}
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, PB, 0)
			!= hipSuccess || per < 1) {
		ssx_set_error("persistent CG: occupancy query failed");
		return SSX_EHIP;
	}
	/* one block per CU: every barrier costs one handshake per block */
	*grid = ncu;
	return SSX_OK;
}
