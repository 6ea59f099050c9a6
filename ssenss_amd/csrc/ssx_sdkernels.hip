// ssx_sdkernels.hip -- the triangular solves of the sparse direct coarse
// solver (ssx_sdirect.c): one forward and one backward launch per level of
// the separator tree, each over (supernode, block of outputs) work items.
//
// Every product is a set of dot products over contiguous rows: the
// forward ones read M and K = L21 M row-major, the backward ones (with M^T
// and K^T) read them column-major.  A group of G lanes takes one output, G a
// power of two chosen from the row length so that each lane issues at most
// SD_U loads, all together, before a fixed shuffle tree adds the group:
// a few memory round trips per product instead of one per entry.  The
// vector operand sits in LDS.  Results do not depend on the launch (G and
// the trees are fixed by the sizes), so every run gives the same bits.
#include <hip/hip_runtime.h>
#include "ssx_internal.h"

#define SD_U 16                         // loads per lane per pass

#define SD_CHECK(name)                                                       \
	do {                                                                 \
		hipError_t e_ = hipGetLastError();                           \
		if (e_ != hipSuccess) {                                      \
			ssx_set_error("%s: %s", name, hipGetErrorString(e_)); \
			return SSX_EHIP;                                     \
		}                                                            \
	} while (0)

// lanes per output for a row of length r: the least power of two with
// r <= SD_U * G, at most 64
__host__ __device__ __forceinline__ int sd_group(int r)
{
	int g = 1;
	while (g < 64 && g * SD_U < r)
		g <<= 1;
	return g;
}

// sum_{k in [k0, k1)} row[k] v[k] by the G lanes of a group (gl: lane in
// group); every lane of the group gets the sum
__device__ __forceinline__ double sd_group_dot(const double *__restrict__ row,
		const double *v, int k0, int k1, int gl, int G)
{
	double s = 0.0;
	if (k1 > k0) {
		for (int k = k0 + gl; k < k1; k += SD_U * G) {
			double a[SD_U], b[SD_U];
#pragma unroll
			for (int u = 0; u < SD_U; ++u) {
				const int kk = k + u * G;
				const int kc = kk < k1 ? kk : k1 - 1;
				a[u] = row[kc];
				b[u] = v[kc];
			}
#pragma unroll
			for (int u = 0; u < SD_U; ++u)
				if (k + u * G < k1)
					s += a[u] * b[u];
		}
	}
	for (int o = G >> 1; o > 0; o >>= 1)
		s += __shfl_xor(s, o, G);
	return s;
}

__global__ void __launch_bounds__(256)
k_sd_gather(int n, const int *__restrict__ perm, const double *__restrict__ src,
		double *__restrict__ dst)
{
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i < n)
		dst[i] = src[perm[i]];
}

__global__ void __launch_bounds__(256)
k_sd_scatter(int n, const int *__restrict__ perm, const double *__restrict__ src,
		double *__restrict__ dst)
{
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i < n)
		dst[perm[i]] = src[i];
}

// dst[q] = src[map[q]], or 0 where map[q] < 0
__global__ void __launch_bounds__(256)
k_sd_scatter_map(int m, const int *__restrict__ map,
		const double *__restrict__ src, double *__restrict__ dst)
{
	const int q = blockIdx.x * 256 + threadIdx.x;
	if (q < m) {
		const int i = map[q];
		dst[q] = i >= 0 ? src[i] : 0.0;
	}
}

// b~_i = b_{c0+i} - (child 0's update) - (child 1's update)
__device__ __forceinline__ double sd_beff(const ssx_sdnode &p,
		const ssx_sdnode *__restrict__ node, const int *__restrict__ idx,
		const double *__restrict__ b, const double *__restrict__ t, int i)
{
	double v = b[p.c0 + i];
	const int k0 = p.ch0 >= 0 ? idx[p.g0 + i] : -1;
	const int k1 = p.ch1 >= 0 ? idx[p.g1 + i] : -1;
	const double t0 = k0 >= 0 ? t[node[p.ch0].toff + k0] : 0.0;
	const double t1 = k1 >= 0 ? t[node[p.ch1].toff + k1] : 0.0;
	if (k0 >= 0)
		v -= t0;
	if (k1 >= 0)
		v -= t1;
	return v;
}

// the children's updates on row r of U (they pass through to the parent)
__device__ __forceinline__ double sd_pass(const ssx_sdnode &p,
		const ssx_sdnode *__restrict__ node, const int *__restrict__ idx,
		const double *__restrict__ t, int r)
{
	const int k0 = p.ch0 >= 0 ? idx[p.g0 + p.nc + r] : -1;
	const int k1 = p.ch1 >= 0 ? idx[p.g1 + p.nc + r] : -1;
	const double t0 = k0 >= 0 ? t[node[p.ch0].toff + k0] : 0.0;
	const double t1 = k1 >= 0 ? t[node[p.ch1].toff + k1] : 0.0;
	double ps = 0.0;
	if (k0 >= 0)
		ps += t0;
	if (k1 >= 0)
		ps += t1;
	return ps;
}

// a workgroup's outputs: block blk of 256 / G outputs
#define SD_BLOCK_OUTPUTS(G) (256 / (G))

// Forward step of one level: a workgroup takes a block of one supernode's
// outputs, y rows (y = M b~) then t rows (t = K b~ + the children's
// updates passing through); b~ is gathered into LDS by every workgroup of
// the supernode.
__global__ void __launch_bounds__(256)
k_sd_fwd(const ssx_sdwork *__restrict__ wk,
		const ssx_sdnode *__restrict__ node, const double *__restrict__ val,
		const int *__restrict__ idx, const double *__restrict__ b,
		double *__restrict__ y, double *__restrict__ t)
{
	extern __shared__ double v[];         // nc entries (the level's largest)
	const ssx_sdwork w = wk[blockIdx.x];
	const ssx_sdnode p = node[w.sn];
	const int nc = p.nc, nu = p.nu;
	const int G = sd_group(nc), gl = threadIdx.x & (G - 1);
	const int o = w.blk * SD_BLOCK_OUTPUTS(G) + threadIdx.x / G;
	const bool trow = o >= nc && o < nc + nu;
	// the pass-through loads go out first, with the gathers of b~
	const double ps = trow && gl == 0 ? sd_pass(p, node, idx, t, o - nc)
		: 0.0;
	for (int k = threadIdx.x; k < nc; k += 256)
		v[k] = sd_beff(p, node, idx, b, t, k);
	__syncthreads();
	if (o < nc) {
		const double s = sd_group_dot(val + p.mroff + (size_t) o * nc, v, 0,
				o + 1, gl, G);
		if (gl == 0)
			y[p.c0 + o] = s;
	} else if (trow) {
		const int r = o - nc;
		const double s = sd_group_dot(val + p.kroff + (size_t) r * nc, v, 0,
				nc, gl, G);
		if (gl == 0)
			t[p.toff + r] = s + ps;
	}
}

// Backward step of one level: x_j = (M^T y)_j - (K^T x_U)_j for a block of
// one supernode's columns j; y of the supernode and x of its rows below
// (computed by the levels above) staged in LDS.
__global__ void __launch_bounds__(256)
k_sd_bwd(const ssx_sdwork *__restrict__ wk,
		const ssx_sdnode *__restrict__ node, const double *__restrict__ val,
		const int *__restrict__ idx, const double *__restrict__ y,
		double *__restrict__ x)
{
	extern __shared__ double v[];         // nc + nu entries
	const ssx_sdwork w = wk[blockIdx.x];
	const ssx_sdnode p = node[w.sn];
	const int nc = p.nc, nu = p.nu;
	for (int k = threadIdx.x; k < nc + nu; k += 256)
		v[k] = k < nc ? y[p.c0 + k] : x[idx[p.uoff + k - nc]];
	__syncthreads();
	const int G = sd_group(nc + nu), gl = threadIdx.x & (G - 1);
	const int j = w.blk * SD_BLOCK_OUTPUTS(G) + threadIdx.x / G;
	if (j >= nc)
		return;
	const double s = sd_group_dot(val + p.moff + (size_t) j * nc, v, j, nc,
			gl, G);
	const double q = sd_group_dot(val + p.koff + (size_t) j * nu, v + nc, 0,
			nu, gl, G);
	if (gl == 0)
		x[p.c0 + j] = s - q;
}

extern "C" int ssxk_sd_blocks(int nout, int rowlen)
{
	const int per = 256 / sd_group(rowlen);
	return (nout + per - 1) / per;
}

extern "C" int ssxk_sd_gather(void *st, int n, const int *perm,
		const double *src, double *dst)
{
	if (n <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_sd_gather, dim3((n + 255) / 256), dim3(256), 0,
			(hipStream_t) st, n, perm, src, dst);
	SD_CHECK("k_sd_gather");
	return SSX_OK;
}

extern "C" int ssxk_sd_scatter(void *st, int n, const int *perm,
		const double *src, double *dst)
{
	if (n <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_sd_scatter, dim3((n + 255) / 256), dim3(256), 0,
			(hipStream_t) st, n, perm, src, dst);
	SD_CHECK("k_sd_scatter");
	return SSX_OK;
}

extern "C" int ssxk_sd_scatter_map(void *st, int m, const int *map,
		const double *src, double *dst)
{
	if (m <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_sd_scatter_map, dim3((m + 255) / 256), dim3(256), 0,
			(hipStream_t) st, m, map, src, dst);
	SD_CHECK("k_sd_scatter_map");
	return SSX_OK;
}

// dynamic LDS beyond 64 KB needs the kernel's attribute raised once
static int sd_lds_attr(const void *fn, int *done)
{
	if (*done)
		return SSX_OK;
	if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
			(int) (sizeof(double) * SD_MAX_VEC)) != hipSuccess) {
		ssx_set_error("sparse direct: %zu bytes of LDS refused",
				sizeof(double) * SD_MAX_VEC);
		return SSX_EHIP;
	}
	*done = 1;
	return SSX_OK;
}

extern "C" int ssxk_sd_fwd(void *st, int nwk, const ssx_sdwork *wk,
		const ssx_sdnode *node, const double *val, const int *idx,
		const double *b, double *y, double *t, int vec)
{
	static int attr = 0;
	if (nwk <= 0)
		return SSX_OK;
	if (vec > SD_MAX_VEC) {
		ssx_set_error("sparse direct: %d-entry vector beyond %d", vec,
				SD_MAX_VEC);
		return SSX_EINVAL;
	}
	if (sd_lds_attr((const void *) k_sd_fwd, &attr))
		return SSX_EHIP;
	hipLaunchKernelGGL(k_sd_fwd, dim3(nwk), dim3(256),
			sizeof(double) * (size_t) (vec > 0 ? vec : 1), (hipStream_t) st,
			wk, node, val, idx, b, y, t);
	SD_CHECK("k_sd_fwd");
	return SSX_OK;
}

extern "C" int ssxk_sd_bwd(void *st, int nwk, const ssx_sdwork *wk,
		const ssx_sdnode *node, const double *val, const int *idx,
		const double *y, double *x, int vec)
{
	static int attr = 0;
	if (nwk <= 0)
		return SSX_OK;
	if (vec > SD_MAX_VEC) {
		ssx_set_error("sparse direct: %d-entry vector beyond %d", vec,
				SD_MAX_VEC);
		return SSX_EINVAL;
	}
	if (sd_lds_attr((const void *) k_sd_bwd, &attr))
		return SSX_EHIP;
	hipLaunchKernelGGL(k_sd_bwd, dim3(nwk), dim3(256),
			sizeof(double) * (size_t) (vec > 0 ? vec : 1), (hipStream_t) st,
			wk, node, val, idx, y, x);
	SD_CHECK("k_sd_bwd");
	return SSX_OK;
}
