// ssx_ax.h -- the element operator (discrete_laplace_operator,
// poisson.c:34-56; discrete_helmholtz_operator, helmholtz.c:33-49) as a
// gfx950 kernel template, instantiated once per precision in
// ssx_ax_f64.hip, ssx_ax_f32.hip and ssx_ax_bf16.hip.
//
// One "team" of N*N lanes per element: lane (i,j) owns the column of the
// element at (i,j) and loops over the k planes.  The element and the two
// in-plane fluxes are staged in LDS with rows padded to N+1 (conflict-free),
// the zeta contraction stays in registers, the six premultiplied factors
// G~ = G*w*detJ stream from HBM in coalesced per-element blocks.  HBM-bound:
// FP64 Poisson moves 64 B per local node (u 8, Au 8, factors 48).
//
// Launch variants (chosen per order by the autotuner in ssx_api.c):
//   bit 0  one-wave teams: max(1, 64/N^2) elements per workgroup instead of
//          max(1, 256/N^2), so the plane barriers span one wavefront
//   bit 1  non-temporal streaming: the factors (read once per application)
//          are loaded and Au stored with the non-temporal hint, so they do
//          not displace the element data in the caches
//   bit 2  low-register: the transposed derivative rows are read from LDS
//          in the plane loop instead of being held in registers (2N fewer
//          registers per lane, more waves per SIMD)
#ifndef SSX_AX_H
#define SSX_AX_H
#include <hip/hip_runtime.h>
#include "ssx_internal.h"
#include "ssx_device.h"


template <int N, int V>
struct AxShape {
	static constexpr int N2 = N * N;
	static constexpr int N3 = N * N * N;
	static constexpr int P = N + 1;   // padded LDS row: conflict-free rows
	static constexpr int TEAMS = (V & 1)
		? ((64 / N2) > 0 ? (64 / N2) : 1)
		: ((256 / N2) > 0 ? (256 / N2) : 1);
	static constexpr int BLOCK = TEAMS * N2;
};

static inline int ssx_ax_teams(int n, int v)
{
	const int n2 = n * n, per = (v & 1) ? 64 : 256;
	return per / n2 > 0 ? per / n2 : 1;
}

// the six factors of one node, q = plane * N^2 + lane, as stored (the
// conversion of narrow factors is left to their use)
template <bool NT, int N3, typename G>
__device__ __forceinline__ void ax_factors(const G *__restrict__ g, int q,
		G f[6])
{
#pragma unroll
	for (int a = 0; a < 6; ++a) {
		if constexpr (NT)
			f[a] = __builtin_nontemporal_load(g + a * N3 + q);
		else
			f[a] = g[a * N3 + q];
	}
}

// the CG direction update fused into the operator (PUPD): p = (z - c) +
// beta p (masked), with beta at scal[ib] and c at scal[ic] (ic < 0: none)
struct PUpd {
	const void *z;
	void *p;
	const double *scal;
	int ib, ic;
	void *u;                     // nullable: u += scal[ia] p_old first
	int ia;
};

// PUPD (PU >= 1): the update is done while the element is loaded, p is
// written back and the operator is applied to the new p (the k_cg_pupdate
// arithmetic); u is not read.  PU == 2: the previous iteration's solution
// step uio += scal[ia] p_old is taken from the same load of p_old first
// (the CG update then leaves u alone).
// One workgroup's elements [eblk*TEAMS, eblk*TEAMS + TEAMS): Au (masked)
// written, and each lane's part of sum p (A_e p) over its nodes returned.
template <int N, int OP, typename T, typename G, int PU, int V>
__device__ __forceinline__ double ax_team(int eblk, int nel,
		const T *__restrict__ u, T *__restrict__ au,
		const G *__restrict__ gf, const T *__restrict__ mass,
		const T *__restrict__ dmat, T lambda,
		const int *__restrict__ bc, int apply_mask,
		const T *__restrict__ zin, T *__restrict__ pio,
		const double *__restrict__ scal, int ib, int ic,
		T *__restrict__ uio, int ia)
{
	using S = AxShape<N, V>;
	constexpr int N2 = S::N2, N3 = S::N3, P = S::P, TEAMS = S::TEAMS;
	constexpr bool NT = (V & 2) != 0, LR = (V & 4) != 0;
	__shared__ T s_d[2 * N2];
	__shared__ T s_u[TEAMS][N * N * P];
	__shared__ T s_a[TEAMS][N * P];
	__shared__ T s_b[TEAMS][N * P];

	const int team = threadIdx.x / N2;
	const int tid = threadIdx.x - team * N2;
	const int i = tid % N, j = tid / N;
	const int e = eblk * TEAMS + team;
	const bool act = e < nel;

	// D and the SVV matrix (2 N^2 entries) for LDS, stored once the
	// element's loads are out
	constexpr int DPER = (2 * N2 + S::BLOCK - 1) / S::BLOCK;
	T dv[DPER];
#pragma unroll
	for (int r = 0; r < DPER; ++r) {
		const int q = (int) threadIdx.x + r * S::BLOCK;
		dv[r] = dmat[q < 2 * N2 ? q : 0];
	}
	// D_il at [i*N+l]; the SVV matrix enters the xi-forward and the
	// eta/zeta-transpose products only (helmholtz.c:41-44, deriv.c:29-66)
	const T *gDn = dmat;
	const T *gDs = dmat + (OP == 1 ? N2 : 0);

	const size_t eoff = (size_t) (act ? e : 0) * N3;
	T *su = s_u[team], *sa = s_a[team], *sb = s_b[team];
	// bc[e]: bits 0-5 Dirichlet faces, bits 8-13 faces with a neighbour
	const int ebr = bc[act ? e : 0];
	const int eb = act ? ebr : 0;
	const int bits = apply_mask ? eb : 0;
	const bool mi = ((bits & 1) && i == 0) || ((bits & 2) && i == N - 1)
		|| ((bits & 4) && j == 0) || ((bits & 8) && j == N - 1);
	// Every load of the element is issued before any is used: the address
	// of an idle lane (e >= nel) is clamped to element 0, so the loads need
	// no branch, and a branch around each would make the compiler wait for
	// each load before the next one goes out.  Stores stay behind act.
	T uc[N], acc[N];
	const G *g = gf + (size_t) (act ? e : 0) * 6 * N3;
	G fa[6], fb[6];
	if (PU >= 1) {
		const T bb = (T) scal[ib];
		const T cc = ic >= 0 ? (T) -scal[ic] : T(0);
		const T aa = PU == 2 ? (T) scal[ia] : T(0);
		T po[N], zz[N], uo[PU == 2 ? N : 1];
#pragma unroll
		for (int k = 0; k < N; ++k) {
			const size_t q = eoff + k * N2 + tid;
			po[k] = pio[q];
			zz[k] = zin[q];
			if (PU == 2)
				uo[k] = uio[q];
		}
		ax_factors<NT, N3>(g, tid, fa);       // plane 0, in flight
#pragma unroll
		for (int k = 0; k < N; ++k) {
			const size_t q = eoff + k * N2 + tid;
			const T zv = ic >= 0 ? cc * T(1) + zz[k] : zz[k];
			T v = bb * po[k] + zv;
			const bool m = mi || ((bits & 16) && k == 0)
				|| ((bits & 32) && k == N - 1);
			if (apply_mask)
				v = v * (m ? T(0) : T(1));
			if (act) {
				if (PU == 2)
					uio[q] = aa * po[k] + uo[PU == 2 ? k : 0];
				pio[q] = v;
			}
			v = act ? v : T(0);
			uc[k] = v;
			su[(k * N + j) * P + i] = v;
			acc[k] = T(0);
		}
	} else {
#pragma unroll
		for (int k = 0; k < N; ++k)
			uc[k] = u[eoff + k * N2 + tid];
		ax_factors<NT, N3>(g, tid, fa);       // plane 0, in flight
#pragma unroll
		for (int k = 0; k < N; ++k) {
			uc[k] = act ? uc[k] : T(0);
			su[(k * N + j) * P + i] = uc[k];
			acc[k] = T(0);
		}
	}
#pragma unroll
	for (int r = 0; r < DPER; ++r) {
		const int q = (int) threadIdx.x + r * S::BLOCK;
		if (q < 2 * N2)
			s_d[q] = dv[r];
	}
	// only LDS data crosses the barriers of this function: the factor
	// loads stay in flight across them
	lds_sync();
	const T *Dn = s_d;
	const T *Ds = s_d + (OP == 1 ? N2 : 0);
	T dxf[N], dyf[N], dxt[LR ? 1 : N], dyt[LR ? 1 : N];
#pragma unroll
	for (int l = 0; l < N; ++l) {
		dxf[l] = Ds[i * N + l];
		dyf[l] = Dn[j * N + l];
		if (!LR) {
			dxt[l] = Dn[l * N + i];
			dyt[l] = Ds[l * N + j];
		}
	}

	// One plane of the element: the factors of the plane are the lane's
	// registers fc, loaded one plane ahead (below).
	auto plane_step = [&](int k, const G *fc) {
		T gx = T(0), gy = T(0), gz = T(0);
		const T *plane = su + k * N * P;
#pragma unroll
		for (int l = 0; l < N; ++l) {
			gx += dxf[l] * plane[j * P + l];
			gy += dyf[l] * plane[l * P + i];
			gz += gDn[k * N + l] * uc[l];      // uniform: scalar loads
		}
		const T gxx = ldv(fc[0]), gxy = ldv(fc[1]), gxz = ldv(fc[2]);
		const T gyy = ldv(fc[3]), gyz = ldv(fc[4]), gzz = ldv(fc[5]);
		T t0 = gxx * gx + gxy * gy + gxz * gz;
		T t1 = gxy * gx + gyy * gy + gyz * gz;
		T t2 = gxz * gx + gyz * gy + gzz * gz;
		t0 = act ? t0 : T(0);
		t1 = act ? t1 : T(0);
		t2 = act ? t2 : T(0);
		sa[j * P + i] = t0;
		sb[j * P + i] = t1;
		lds_sync();
		T r = T(0);
#pragma unroll
		for (int l = 0; l < N; ++l) {
			const T xt = LR ? Dn[l * N + i] : dxt[l];
			const T yt = LR ? Ds[l * N + j] : dyt[l];
			r += xt * sa[j * P + l] + yt * sb[l * P + i];
		}
#pragma unroll
		for (int kk = 0; kk < N; ++kk)
			acc[kk] += (kk == k ? r : T(0)) + gDs[k * N + kk] * t2;
		lds_sync();
	};
	// The six factors of plane k + 1 are loaded before plane k is computed,
	// so two planes' loads are in flight per lane instead of one.  The loop
	// takes two planes a turn with two register sets, so no loaded value is
	// copied (a copy would wait for its load); the last one or two planes
	// are taken after it, with no load past the last plane (a second read
	// of a plane streamed with the non-temporal hint went back to HBM).
	int k = 0;
#pragma unroll 1
	for (; k + 2 < N; k += 2) {
		ax_factors<NT, N3>(g, (k + 1) * N2 + tid, fb);
		plane_step(k, fa);
		ax_factors<NT, N3>(g, (k + 2) * N2 + tid, fa);
		plane_step(k + 1, fb);
	}
	if constexpr (N % 2 == 0) {
		ax_factors<NT, N3>(g, (N - 1) * N2 + tid, fb);
		plane_step(N - 2, fa);
		plane_step(N - 1, fb);
	} else {
		plane_step(N - 1, fa);
	}
	T ms[OP == 1 ? N : 1];
	if (OP == 1) {
#pragma unroll
		for (int k = 0; k < N; ++k)
			ms[k] = mass[eoff + k * N2 + tid];
	}
	double pd = 0.0;
	if (act) {
#pragma unroll
		for (int k = 0; k < N; ++k) {
			const int q = k * N2 + tid;
			T v = acc[k];
			if (OP == 1)
				v += lambda * (ms[OP == 1 ? k : 0] * uc[k]);
			const bool m = mi || ((bits & 16) && k == 0)
				|| ((bits & 32) && k == N - 1);
			v = m ? T(0) : v;
			if (NT)
				__builtin_nontemporal_store(v, au + eoff + q);
			else
				au[eoff + q] = v;
			// <p, DSS(Ap)>_w: for p equal on the copies of each shared
			// node, sum_w p DSS(Ap) is the sum over all local copies of
			// p (A_e p), so it is complete before the assembly
			pd += (double) uc[k] * (double) v;
		}
	}
	return pd;
}

// One partial per element: the N*N lanes of each team are summed by a
// fixed tree over a power-of-two segment padded with zeros, and element e's
// sum goes to part[e].  The tree does not depend on how many teams share a
// workgroup, so every launch shape gives the same partials (the reductions
// that follow, and with them the whole solve, do not depend on the shape
// the autotuner picked).
template <int N, int V>
__device__ __forceinline__ void ax_elem_partial(double pd, double *part,
		int eblk, int nel)
{
	using S = AxShape<N, V>;
	constexpr int N2 = S::N2, TEAMS = S::TEAMS, BLOCK = S::BLOCK;
	constexpr int P2 = N2 <= 4 ? 4 : N2 <= 16 ? 16 : N2 <= 32 ? 32
		: N2 <= 64 ? 64 : N2 <= 128 ? 128 : 256;
	static_assert(N2 <= 256, "team larger than the partial tree");
	if constexpr (N2 == P2 && P2 <= 64) {
		// a team fills an aligned segment of one wavefront: the same tree
		// (lane l adds lane l + st) by shuffles, no LDS and no barrier
		double v = pd;
#pragma unroll
		for (int st = P2 / 2; st > 0; st >>= 1)
			v += __shfl_down(v, st, P2);
		const int team = threadIdx.x / N2;
		const int e = eblk * TEAMS + team;
		if (threadIdx.x - team * N2 == 0 && e < nel)
			part[e] = v;
		return;
	}
	__shared__ double s_red[TEAMS * P2];
	const int team = threadIdx.x / N2;
	const int tid = threadIdx.x - team * N2;
	s_red[team * P2 + tid] = pd;
	if (tid < P2 - N2)
		s_red[team * P2 + N2 + tid] = 0.0;
	__syncthreads();
#pragma unroll
	for (int st = P2 / 2; st > 0; st >>= 1) {
		for (int q = threadIdx.x; q < TEAMS * st; q += BLOCK) {
			const int t = q / st, l = q - t * st;
			s_red[t * P2 + l] += s_red[t * P2 + l + st];
		}
		__syncthreads();
	}
	if ((int) threadIdx.x < TEAMS) {
		const int e = eblk * TEAMS + threadIdx.x;
		if (e < nel)
			part[e] = s_red[threadIdx.x * P2];
	}
}

template <int N, int OP, typename T, typename G, int PU, int V>
__global__ void __launch_bounds__((AxShape<N, V>::BLOCK))
k_ax(int nel, const T *__restrict__ u, T *__restrict__ au,
		const G *__restrict__ gf, const T *__restrict__ mass,
		const T *__restrict__ dmat, T lambda,
		const int *__restrict__ bc, int apply_mask,
		const int *__restrict__ done, double *__restrict__ part,
		const T *__restrict__ zin, T *__restrict__ pio,
		const double *__restrict__ scal, int ib, int ic,
		T *__restrict__ uio, int ia)
{
	if (done && *done)          // solver already converged: no-op
		return;
	const double pd = ax_team<N, OP, T, G, PU, V>(blockIdx.x, nel, u, au,
			gf, mass, dmat, lambda, bc, apply_mask, zin, pio, scal, ib, ic,
			uio, ia);
	if (part)
		ax_elem_partial<N, V>(pd, part, blockIdx.x, nel);
}

template <int N, typename T, typename G, int PU, int V>
static void launch_ax1(hipStream_t st, int nel, int op, const void *u,
		void *au, const void *gf, const void *mass, const void *dmat,
		double lambda, const int *bc, int apply_mask, const int *done,
		double *part, const PUpd &pu)
{
	using S = AxShape<N, V>;
	const int grid = (nel + S::TEAMS - 1) / S::TEAMS;
	if (op == SSX_OP_HELMHOLTZ)
		hipLaunchKernelGGL((k_ax<N, 1, T, G, PU, V>), dim3(grid),
				dim3(S::BLOCK), 0, st, nel, (const T *) u, (T *) au,
				(const G *) gf, (const T *) mass, (const T *) dmat, (T) lambda,
				bc, apply_mask, done, part, (const T *) pu.z, (T *) pu.p,
				pu.scal, pu.ib, pu.ic, (T *) pu.u, pu.ia);
	else
		hipLaunchKernelGGL((k_ax<N, 0, T, G, PU, V>), dim3(grid),
				dim3(S::BLOCK), 0, st, nel, (const T *) u, (T *) au,
				(const G *) gf, (const T *) mass, (const T *) dmat, (T) lambda,
				bc, apply_mask, done, part, (const T *) pu.z, (T *) pu.p,
				pu.scal, pu.ib, pu.ic, (T *) pu.u, pu.ia);
}

template <int N, typename T, typename G, int V>
static void launch_ax_v(hipStream_t st, int nel, int op, const void *u,
		void *au, const void *gf, const void *mass, const void *dmat,
		double lambda, const int *bc, int apply_mask, const int *done,
		double *part, const PUpd *pu)
{
	// the deferred solution step (PU = 2) is instantiated for FP64 fields
	// only, where the fused CG takes it
	if (pu && pu->u && sizeof(T) == 8)
		launch_ax1<N, T, G, (sizeof(T) == 8 ? 2 : 1), V>(st, nel, op, u, au,
				gf, mass, dmat, lambda, bc, apply_mask, done, part, *pu);
	else if (pu)
		launch_ax1<N, T, G, 1, V>(st, nel, op, u, au, gf, mass, dmat,
				lambda, bc, apply_mask, done, part, *pu);
	else
		launch_ax1<N, T, G, 0, V>(st, nel, op, u, au, gf, mass, dmat,
				lambda, bc, apply_mask, done, part, PUpd{ nullptr, nullptr,
				nullptr, 0, -1, nullptr, 0 });
}

// dispatcher body shared by the per-precision files: n in [2,16], variant
// v in [0, SSX_AX_NVARIANTS)
template <typename T, typename G>
static int ax_dispatch(hipStream_t st, int n, int v, int nel, int op,
		const void *u, void *au, const void *gf, const void *mass,
		const void *dmat, double lambda, const int *bc, int apply_mask,
		const int *done, double *part, const PUpd *pu)
{
	if (v < 0 || v >= SSX_AX_NVARIANTS) {
		ssx_set_error("ssxk_ax: variant %d unsupported", v);
		return SSX_EINVAL;
	}
	switch (n * SSX_AX_NVARIANTS + v) {
#define CASEV(NN, VV) case NN * SSX_AX_NVARIANTS + VV: \
		launch_ax_v<NN, T, G, VV>(st, nel, op, u, au, gf, mass, dmat, \
				lambda, bc, apply_mask, done, part, pu); break;
#define CASE(NN) CASEV(NN, 0) CASEV(NN, 1) CASEV(NN, 2) CASEV(NN, 3) \
		CASEV(NN, 4) CASEV(NN, 5) CASEV(NN, 6) CASEV(NN, 7)
	CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9)
	CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
#undef CASEV
	default:
		ssx_set_error("ssxk_ax: n=%d unsupported", n);
		return SSX_EINVAL;
	}
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) {
		ssx_set_error("k_ax: %s", hipGetErrorString(e));
		return SSX_EHIP;
	}
	return SSX_OK;
}

#endif
