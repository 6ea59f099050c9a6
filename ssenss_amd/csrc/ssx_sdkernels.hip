// ssx_sdkernels.hip -- the triangular solves of the sparse direct coarse
// solver (ssx_sdirect.c): one launch per level of the separator tree (one
// workgroup per supernode), or three (forward) / two (backward) on levels
// whose supernodes are split over many workgroups.
//
// Every product is a set of dot products over contiguous rows: the
// forward ones read M and L21 row-major, the backward ones (with M^T and
// L21^T) read them column-major.  A group of G lanes takes one output, G a
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

// ---- levels of small supernodes: one workgroup of 1024 lanes each ----

__global__ void __launch_bounds__(1024)
k_sd_fwd_small(const int *__restrict__ lev_sn,
		const ssx_sdnode *__restrict__ node, const double *__restrict__ val,
		const int *__restrict__ idx, const double *__restrict__ b,
		double *__restrict__ y, double *__restrict__ t)
{
	__shared__ double w[SD_SMALL_NC], ys[SD_SMALL_NC], tp[SD_SMALL_NU];
	const ssx_sdnode p = node[lev_sn[blockIdx.x]];
	const int nc = p.nc, nu = p.nu;
	for (int i = threadIdx.x; i < nc; i += 1024)
		w[i] = sd_beff(p, node, idx, b, t, i);
	for (int r = threadIdx.x; r < nu; r += 1024)
		tp[r] = sd_pass(p, node, idx, t, r);
	__syncthreads();
	const int G = sd_group(nc);
	const int gl = threadIdx.x & (G - 1), grp = threadIdx.x / G;
	const int per = 1024 / G;
	const double *M = val + p.mroff;
	for (int i = grp; i < nc; i += per) {
		const double s = sd_group_dot(M + (size_t) i * nc, w, 0, i + 1, gl, G);
		if (gl == 0) {
			ys[i] = s;
			y[p.c0 + i] = s;
		}
	}
	__syncthreads();
	const double *L = val + p.lroff;
	for (int r = grp; r < nu; r += per) {
		const double s = sd_group_dot(L + (size_t) r * nc, ys, 0, nc, gl, G);
		if (gl == 0)
			t[p.toff + r] = s + tp[r];
	}
}

__global__ void __launch_bounds__(1024)
k_sd_bwd_small(const int *__restrict__ lev_sn,
		const ssx_sdnode *__restrict__ node, const double *__restrict__ val,
		const int *__restrict__ idx, const double *__restrict__ y,
		double *__restrict__ x)
{
	__shared__ double xu[SD_SMALL_NU], g[SD_SMALL_NC];
	const ssx_sdnode p = node[lev_sn[blockIdx.x]];
	const int nc = p.nc, nu = p.nu;
	for (int r = threadIdx.x; r < nu; r += 1024)
		xu[r] = x[idx[p.uoff + r]];
	__syncthreads();
	int G = sd_group(nu);
	int gl = threadIdx.x & (G - 1), grp = threadIdx.x / G;
	const double *L = val + p.loff;
	for (int j = grp; j < nc; j += 1024 / G) {
		const double s = sd_group_dot(L + (size_t) j * nu, xu, 0, nu, gl, G);
		if (gl == 0)
			g[j] = y[p.c0 + j] - s;
	}
	__syncthreads();
	G = sd_group(nc);
	gl = threadIdx.x & (G - 1);
	grp = threadIdx.x / G;
	const double *M = val + p.moff;
	for (int j = grp; j < nc; j += 1024 / G) {
		const double s = sd_group_dot(M + (size_t) j * nc, g, j, nc, gl, G);
		if (gl == 0)
			x[p.c0 + j] = s;
	}
}

// ---- split levels: (supernode, block of outputs) work items ----

// b~ into e and the pass-through part of t, one workgroup per supernode
__global__ void __launch_bounds__(256)
k_sd_fwd_beff(const int *__restrict__ lev_sn,
		const ssx_sdnode *__restrict__ node, const int *__restrict__ idx,
		const double *__restrict__ b, double *__restrict__ e,
		double *__restrict__ t)
{
	const ssx_sdnode p = node[lev_sn[blockIdx.x]];
	for (int i = threadIdx.x; i < p.nc; i += 256)
		e[p.c0 + i] = sd_beff(p, node, idx, b, t, i);
	for (int r = threadIdx.x; r < p.nu; r += 256)
		t[p.toff + r] = sd_pass(p, node, idx, t, r);
}

// the workgroup's outputs: block blk of 256 / G outputs
#define SD_BLOCK_OUTPUTS(G) (256 / (G))

// y = M b~
__global__ void __launch_bounds__(256)
k_sd_fwd_y(const ssx_sdwork *__restrict__ wk,
		const ssx_sdnode *__restrict__ node, const double *__restrict__ val,
		const double *__restrict__ e, double *__restrict__ y)
{
	__shared__ double v[SD_MAX_VEC];
	const ssx_sdwork w = wk[blockIdx.x];
	const ssx_sdnode p = node[w.sn];
	const int nc = p.nc;
	for (int k = threadIdx.x; k < nc; k += 256)
		v[k] = e[p.c0 + k];
	__syncthreads();
	const int G = sd_group(nc), gl = threadIdx.x & (G - 1);
	const int i = w.blk * SD_BLOCK_OUTPUTS(G) + threadIdx.x / G;
	if (i >= nc)
		return;
	const double s = sd_group_dot(val + p.mroff + (size_t) i * nc, v, 0,
			i + 1, gl, G);
	if (gl == 0)
		y[p.c0 + i] = s;
}

// t += L21 y (t holds the pass-through part)
__global__ void __launch_bounds__(256)
k_sd_fwd_t(const ssx_sdwork *__restrict__ wk,
		const ssx_sdnode *__restrict__ node, const double *__restrict__ val,
		const double *__restrict__ y, double *__restrict__ t)
{
	__shared__ double v[SD_MAX_VEC];
	const ssx_sdwork w = wk[blockIdx.x];
	const ssx_sdnode p = node[w.sn];
	const int nc = p.nc;
	for (int k = threadIdx.x; k < nc; k += 256)
		v[k] = y[p.c0 + k];
	__syncthreads();
	const int G = sd_group(nc), gl = threadIdx.x & (G - 1);
	const int r = w.blk * SD_BLOCK_OUTPUTS(G) + threadIdx.x / G;
	if (r >= p.nu)
		return;
	const double s = sd_group_dot(val + p.lroff + (size_t) r * nc, v, 0, nc,
			gl, G);
	if (gl == 0)
		t[p.toff + r] = s + t[p.toff + r];
}

// e = y - L21^T x_U
__global__ void __launch_bounds__(256)
k_sd_bwd_g(const ssx_sdwork *__restrict__ wk,
		const ssx_sdnode *__restrict__ node, const double *__restrict__ val,
		const int *__restrict__ idx, const double *__restrict__ y,
		const double *__restrict__ x, double *__restrict__ e)
{
	__shared__ double v[SD_MAX_VEC];
	const ssx_sdwork w = wk[blockIdx.x];
	const ssx_sdnode p = node[w.sn];
	const int nu = p.nu;
	for (int r = threadIdx.x; r < nu; r += 256)
		v[r] = x[idx[p.uoff + r]];
	__syncthreads();
	const int G = sd_group(nu), gl = threadIdx.x & (G - 1);
	const int j = w.blk * SD_BLOCK_OUTPUTS(G) + threadIdx.x / G;
	if (j >= p.nc)
		return;
	const double s = sd_group_dot(val + p.loff + (size_t) j * nu, v, 0, nu,
			gl, G);
	if (gl == 0)
		e[p.c0 + j] = y[p.c0 + j] - s;
}

// x = M^T e
__global__ void __launch_bounds__(256)
k_sd_bwd_x(const ssx_sdwork *__restrict__ wk,
		const ssx_sdnode *__restrict__ node, const double *__restrict__ val,
		const double *__restrict__ e, double *__restrict__ x)
{
	__shared__ double v[SD_MAX_VEC];
	const ssx_sdwork w = wk[blockIdx.x];
	const ssx_sdnode p = node[w.sn];
	const int nc = p.nc;
	for (int k = threadIdx.x; k < nc; k += 256)
		v[k] = e[p.c0 + k];
	__syncthreads();
	const int G = sd_group(nc), gl = threadIdx.x & (G - 1);
	const int j = w.blk * SD_BLOCK_OUTPUTS(G) + threadIdx.x / G;
	if (j >= nc)
		return;
	const double s = sd_group_dot(val + p.moff + (size_t) j * nc, v, j, nc,
			gl, G);
	if (gl == 0)
		x[p.c0 + j] = s;
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

extern "C" int ssxk_sd_fwd_small(void *st, int nsn, const int *lev_sn,
		const ssx_sdnode *node, const double *val, const int *idx,
		const double *b, double *y, double *t)
{
	if (nsn <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_sd_fwd_small, dim3(nsn), dim3(1024), 0,
			(hipStream_t) st, lev_sn, node, val, idx, b, y, t);
	SD_CHECK("k_sd_fwd_small");
	return SSX_OK;
}

extern "C" int ssxk_sd_fwd_beff(void *st, int nsn, const int *lev_sn,
		const ssx_sdnode *node, const int *idx, const double *b, double *e,
		double *t)
{
	if (nsn <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_sd_fwd_beff, dim3(nsn), dim3(256), 0,
			(hipStream_t) st, lev_sn, node, idx, b, e, t);
	SD_CHECK("k_sd_fwd_beff");
	return SSX_OK;
}

extern "C" int ssxk_sd_fwd_y(void *st, int nwk, const ssx_sdwork *wk,
		const ssx_sdnode *node, const double *val, const double *e,
		double *y)
{
	if (nwk <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_sd_fwd_y, dim3(nwk), dim3(256), 0,
			(hipStream_t) st, wk, node, val, e, y);
	SD_CHECK("k_sd_fwd_y");
	return SSX_OK;
}

extern "C" int ssxk_sd_fwd_t(void *st, int nwk, const ssx_sdwork *wk,
		const ssx_sdnode *node, const double *val, const double *y,
		double *t)
{
	if (nwk <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_sd_fwd_t, dim3(nwk), dim3(256), 0,
			(hipStream_t) st, wk, node, val, y, t);
	SD_CHECK("k_sd_fwd_t");
	return SSX_OK;
}

extern "C" int ssxk_sd_bwd_small(void *st, int nsn, const int *lev_sn,
		const ssx_sdnode *node, const double *val, const int *idx,
		const double *y, double *x)
{
	if (nsn <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_sd_bwd_small, dim3(nsn), dim3(1024), 0,
			(hipStream_t) st, lev_sn, node, val, idx, y, x);
	SD_CHECK("k_sd_bwd_small");
	return SSX_OK;
}

extern "C" int ssxk_sd_bwd_g(void *st, int nwk, const ssx_sdwork *wk,
		const ssx_sdnode *node, const double *val, const int *idx,
		const double *y, const double *x, double *e)
{
	if (nwk <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_sd_bwd_g, dim3(nwk), dim3(256), 0,
			(hipStream_t) st, wk, node, val, idx, y, x, e);
	SD_CHECK("k_sd_bwd_g");
	return SSX_OK;
}

extern "C" int ssxk_sd_bwd_x(void *st, int nwk, const ssx_sdwork *wk,
		const ssx_sdnode *node, const double *val, const double *e,
		double *x)
{
	if (nwk <= 0)
		return SSX_OK;
	hipLaunchKernelGGL(k_sd_bwd_x, dim3(nwk), dim3(256), 0,
			(hipStream_t) st, wk, node, val, e, x);
	SD_CHECK("k_sd_bwd_x");
	return SSX_OK;
}
