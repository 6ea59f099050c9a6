// ssx_device.h -- device helpers shared by the gfx950 kernel files.
#ifndef SSX_DEVICE_H
#define SSX_DEVICE_H
#include <hip/hip_runtime.h>

typedef unsigned short bf16_t;

// factor loads: bf16 is the upper half of an IEEE float
__device__ __forceinline__ double ldg(const double *p, size_t i) { return p[i]; }
__device__ __forceinline__ float ldg(const float *p, size_t i) { return p[i]; }
__device__ __forceinline__ float ldg(const bf16_t *p, size_t i)
{
	return __uint_as_float(((unsigned) p[i]) << 16);
}

// a loaded factor as arithmetic type
__device__ __forceinline__ double ldv(double x) { return x; }
__device__ __forceinline__ float ldv(float x) { return x; }
__device__ __forceinline__ float ldv(bf16_t x)
{
	return __uint_as_float(((unsigned) x) << 16);
}

// barriers among the lanes of one wavefront (LDS written by one lane and
// read by another), and for LDS data only across a block: loads of global
// memory still in flight stay in flight across the latter
__device__ __forceinline__ void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void lds_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
	__builtin_amdgcn_s_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A pointer read from device memory (a batch problem's description) is a
// generic pointer: loads and stores through it compile to flat memory
// instructions, which also count against the LDS counter and make the
// compiler wait more often.  The buffers are all in global memory, and a
// pointer of the global address space gets global loads and stores.
#define SSX_GLOBAL __attribute__((address_space(1)))

// a generic pointer rebuilt from its address in the global address space:
// loads through it (once inlined) can then be global loads rather than
// flat ones
template <typename T>
__device__ __forceinline__ T *gptr(T *p)
{
	return (T *) (SSX_GLOBAL T *) (unsigned long long) p;
}

// Logical workgroup of a gather kernel.  Workgroups go to the 8 XCDs
// round-robin (workgroup b on XCD b % 8), and each XCD has its own L2.
// Renumbering so that XCD x takes a contiguous range of logical workgroups
// (the i-th workgroup it receives is the i-th of its range) gives each XCD
// a contiguous run of elements: the partner copies a node gathers from the
// neighbouring elements are then mostly in the L2 that loaded them.  The
// map is a bijection of [0, nb), and logical workgroup lb covers exactly
// the nodes workgroup lb covered before, so partials stored at lb keep
// every reduction's order.
__device__ __forceinline__ unsigned xcd_block(void)
{
	const unsigned nb = gridDim.x, b = blockIdx.x;
	const unsigned x = b & 7u, q = nb >> 3, rem = nb & 7u;
	return x * q + (x < rem ? x : rem) + (b >> 3);
}

// partner order of copy t: t ^ ord(k), k = 1..M-1 (faces x,y,z, then edges
// along z, y, x, then the corner; global.c:252-256)
__device__ __forceinline__ int dss_ord(int M, int k)
{
	return M == 8 ? ((k == 3) ? 4 : (k == 4) ? 3 : k) : k;
}

// One group of M copies.  Optionally each sum is scaled by 1/M
// (direct_stiffness_averaging, global.c:259-268: every copy in a group of M
// has multiplicity exactly M), and the shared-node part of <p, DSS(u)>_w is
// returned.
template <int M, typename T>
__device__ __forceinline__ double dss_group(long long G, long long g,
		const int *__restrict__ idx, T *__restrict__ u,
		const T *__restrict__ p, int average)
{
	int ix[M];
	T v[M];
#pragma unroll
	for (int t = 0; t < M; ++t)
		ix[t] = idx[(long long) t * G + g];
#pragma unroll
	for (int t = 0; t < M; ++t)
		v[t] = u[ix[t]];
	const T w = T(1) / T(M);
	double pd = 0.0;
#pragma unroll
	for (int t = 0; t < M; ++t) {
		T s = v[t];
#pragma unroll
		for (int k = 1; k < M; ++k)
			s = s + v[t ^ dss_ord(M, k)];
		if (average)
			s = s * w;
		u[ix[t]] = s;
		if (p)
			pd += (double) (p[ix[t]] * w) * (double) s;
	}
	return pd;
}

#endif
