// ssx_nskernels.hip -- element operators of the Navier-Stokes step around
// the elliptic solves (stokes.c, convection.c, poisson.c:58-224, time.c):
// reference-space Jacobians of fields, curl, weak divergence and gradient,
// the dealiased skew-symmetric convection operator at the convective
// resolution, face integrals, and the BDF3/EXT3 and interpolated-velocity
// combinations.  All on element-local E-vectors, node (e,k,j,i) at
// ((e*n+k)*n+j)*n+i; one thread per node unless stated.  The arithmetic
// keeps the reference's order of operations per node (products formed
// and summed in its loop order).
//
// D layout: dm[i*n+l] = D_il = l_l'(xi_i); reference derivatives
//   (D_xi u)(k,j,i)   = sum_l D_il u(k,j,l)
//   (D_eta u)(k,j,i)  = sum_l D_jl u(k,l,i)
//   (D_zeta u)(k,j,i) = sum_l D_kl u(l,j,i)          (deriv.c:29-52)
// Inverse Jacobian: ti[3*a+b] = d xi_b / d x_a (the transposed inverse
// tinv_jac, geometry.c:109-112).
#include <hip/hip_runtime.h>
#include "ssx_internal.h"
#include "ssx_device.h"

#define NS_LAUNCH(name)                                                    \
	do {                                                               \
		hipError_t e_ = hipGetLastError();                         \
		if (e_ != hipSuccess) {                                    \
			ssx_set_error("%s: %s", name, hipGetErrorString(e_)); \
			return SSX_EHIP;                                   \
		}                                                          \
	} while (0)

struct P3 { const double *p[3]; };
struct M3 { double *p[3]; };
struct P9 { const double *p[9]; };
struct M9 { double *p[9]; };

static inline unsigned ns_blocks(size_t n)
{
	return (unsigned) ((n + 255) / 256);
}

// reference derivative of u along direction d at node (k,j,i) of element
// base `u` (n^3 block)
__device__ __forceinline__ double dref(const double *__restrict__ u,
		const double *__restrict__ dm, int n, int d, int k, int j, int i)
{
	double s = 0.0;
	if (d == 0) {
		const double *r = u + (k * n + j) * n;
		for (int l = 0; l < n; ++l)
			s += dm[i * n + l] * r[l];
	} else if (d == 1) {
		const double *r = u + k * n * n + i;
		for (int l = 0; l < n; ++l)
			s += dm[j * n + l] * r[l * n];
	} else {
		const double *r = u + j * n + i;
		for (int l = 0; l < n; ++l)
			s += dm[k * n + l] * r[l * n * n];
	}
	return s;
}

// J[3a+b] = D_b a_a for three fields a_0..a_2 (compute_jacobian,
// deriv.c:68-79)
__global__ void __launch_bounds__(256)
k_ns_jac(long long nt, int n, const double *__restrict__ dm, P3 a, M9 out)
{
	const long long q = (long long) blockIdx.x * 256 + threadIdx.x;
	if (q >= nt)
		return;
	const int n3 = n * n * n;
	const long long e = q / n3;
	const int r = (int) (q - e * n3);
	const int i = r % n, j = (r / n) % n, k = r / (n * n);
	for (int f = 0; f < 3; ++f) {
		const double *u = a.p[f] + e * n3;
		for (int d = 0; d < 3; ++d)
			out.p[3 * f + d][q] = dref(u, dm, n, d, k, j, i);
	}
}

// curl from the Jacobian of u (discrete_curl_operator, stokes.c:71-92):
// omega_c, for c = 0..2 and xi-direction i = 0..2 in that loop order,
// += s_c (D_i u_k)(d xi_i/d x_l) - s_c (D_i u_l)(d xi_i/d x_k) with
// k = (5-c)/2, l = (2-c)/2, s = (+1, -1, +1); with mass, times w detJ
// (the test-function product of double_curl_operator, stokes.c:94-105)
__global__ void __launch_bounds__(256)
k_ns_curl(long long nt, P9 J, P9 ti, const double *__restrict__ mass,
		M3 out)
{
	const long long q = (long long) blockIdx.x * 256 + threadIdx.x;
	if (q >= nt)
		return;
	double jv[9], tv[9];
	for (int a = 0; a < 9; ++a) {
		jv[a] = J.p[a][q];
		tv[a] = ti.p[a][q];
	}
	double om[3];
	for (int i = 0; i < 3; ++i)
		for (int c = 0; c < 3; ++c) {
			const int k = (5 - c) / 2, l = (2 - c) / 2;
			const double s = c == 1 ? -1.0 : 1.0;
			const double t = s * jv[3 * k + i] * tv[3 * l + i];
			om[c] = i == 0 ? t : t + om[c];
			om[c] = -s * jv[3 * l + i] * tv[3 * k + i] + om[c];
		}
	for (int c = 0; c < 3; ++c)
		out.p[c][q] = mass ? om[c] * mass[q] : om[c];
}

// weak divergence (discrete_divergence_operator, stokes.c:52-69):
// sum_a sum_b (D_b u_a)(d xi_b/d x_a) in that loop order, times w detJ,
// times `scale`
__global__ void __launch_bounds__(256)
k_ns_div(long long nt, P9 J, P9 ti, const double *__restrict__ mass,
		double scale, double *__restrict__ out)
{
	const long long q = (long long) blockIdx.x * 256 + threadIdx.x;
	if (q >= nt)
		return;
	double s = 0.0;
	for (int a = 0; a < 3; ++a)
		for (int b = 0; b < 3; ++b) {
			const double t = J.p[3 * a + b][q] * ti.p[3 * a + b][q];
			s = (a + b == 0) ? t : t + s;
		}
	out[q] = (s * mass[q]) * scale;
}

// weak gradient (discrete_gradient_operator, stokes.c:30-50): with
// t = p w detJ and f_b = (d xi_b/d x_c) t, grad_c = -(D_xi^T f_0 +
// D_eta^T f_1 + D_zeta^T f_2) (transpose_divergence, deriv.c:54-66)
__global__ void __launch_bounds__(256)
k_ns_grad(long long nt, int n, const double *__restrict__ dm,
		const double *__restrict__ p, const double *__restrict__ mass, P9 ti,
		M3 out)
{
	const long long q = (long long) blockIdx.x * 256 + threadIdx.x;
	if (q >= nt)
		return;
	const int n3 = n * n * n;
	const long long e = q / n3;
	const int r = (int) (q - e * n3);
	const int i = r % n, j = (r / n) % n, k = r / (n * n);
	const long long b0 = e * n3;
	for (int c = 0; c < 3; ++c) {
		double sx = 0.0, sy = 0.0, sz = 0.0;
		for (int l = 0; l < n; ++l) {
			const long long qx = b0 + (k * n + j) * n + l;
			const long long qy = b0 + (k * n + l) * n + i;
			const long long qz = b0 + (l * n + j) * n + i;
			sx += dm[l * n + i] * (ti.p[3 * c][qx] * (p[qx] * mass[qx]));
			sy += dm[l * n + j] * (ti.p[3 * c + 1][qy] * (p[qy] * mass[qy]));
			sz += dm[l * n + k] * (ti.p[3 * c + 2][qz] * (p[qz] * mass[qz]));
		}
		out.p[c][q] = -((sx + sy) + sz);
	}
}

// skew-symmetric convection at the convective resolution
// (skew_symmetric_convection_operator, convection.c:129-140): half the
// convective form (u . grad) v_c (convective_convection_operator, :71-93)
// plus half the conservative form as convection.c:95-127 forms it,
// sum_a sum_b (d xi_b/d x_a) D_b(u_b v_c), each times w detJ
__global__ void __launch_bounds__(256)
k_ns_conv(long long nt, int n, const double *__restrict__ dm, P3 u, P3 v,
		P9 Jv, P9 ti, const double *__restrict__ mass, M3 out)
{
	const long long q = (long long) blockIdx.x * 256 + threadIdx.x;
	if (q >= nt)
		return;
	const int n3 = n * n * n;
	const long long e = q / n3;
	const int r = (int) (q - e * n3);
	const int i = r % n, j = (r / n) % n, k = r / (n * n);
	const long long b0 = e * n3;
	double tv[9], uq[3];
	for (int a = 0; a < 9; ++a)
		tv[a] = ti.p[a][q];
	for (int a = 0; a < 3; ++a)
		uq[a] = u.p[a][q];
	const double wj = mass[q];
	for (int c = 0; c < 3; ++c) {
		// convective: dv_ci = sum_k tinv[i][k] D_k v_c; sum_i u_i dv_ci
		double cv = 0.0;
		for (int a = 0; a < 3; ++a) {
			double dv = 0.0;
			for (int b = 0; b < 3; ++b) {
				const double t = tv[3 * a + b] * Jv.p[3 * c + b][q];
				dv = b == 0 ? t : t + dv;
			}
			const double t = uq[a] * dv;
			cv = a == 0 ? t : t + cv;
		}
		cv = cv * wj;
		// conservative: tmp_b = D_b(u_b v_c); duvd_a = sum_b tinv[a][b]
		// tmp_b; sum over a
		double tb[3] = { 0.0, 0.0, 0.0 };
		for (int l = 0; l < n; ++l) {
			const long long qx = b0 + (k * n + j) * n + l;
			const long long qy = b0 + (k * n + l) * n + i;
			const long long qz = b0 + (l * n + j) * n + i;
			tb[0] += dm[i * n + l] * (u.p[0][qx] * v.p[c][qx]);
			tb[1] += dm[j * n + l] * (u.p[1][qy] * v.p[c][qy]);
			tb[2] += dm[k * n + l] * (u.p[2][qz] * v.p[c][qz]);
		}
		double cs = 0.0;
		for (int a = 0; a < 3; ++a) {
			double d = 0.0;
			for (int b = 0; b < 3; ++b) {
				const double t = tv[3 * a + b] * tb[b];
				d = b == 0 ? t : t + d;
			}
			cs = a == 0 ? d : cs + d;
		}
		cs = cs * wj;
		out.p[c][q] = 0.5 * cs + 0.5 * cv;
	}
}

// The same operator, one workgroup per element: the three u components
// and one v component at a time are staged in LDS, so the reference
// derivatives of v and of u_b v along every line come from LDS instead of
// per-node global gathers, and the Jacobian of v is never stored.  Per
// node and component the arithmetic is k_ns_conv's (dref's sums, then the
// same combinations).  LDS: 4 n^3 + n^2 doubles (n <= 15).
__global__ void __launch_bounds__(1024)
k_ns_conv_elem(int n, const double *__restrict__ dm, P3 u, P3 v, P9 ti,
		const double *__restrict__ mass, M3 out)
{
	extern __shared__ double sm[];
	const int n2 = n * n, n3 = n2 * n;
	double *su = sm;                    // u_0, u_1, u_2
	double *sv = sm + 3 * n3;           // v_c
	double *sd = sm + 4 * n3;           // D
	const long long b0 = (long long) blockIdx.x * n3;
	for (int q = threadIdx.x; q < n2; q += blockDim.x)
		sd[q] = dm[q];
	for (int a = 0; a < 3; ++a)
		for (int q = threadIdx.x; q < n3; q += blockDim.x)
			su[a * n3 + q] = u.p[a][b0 + q];
	for (int c = 0; c < 3; ++c) {
		__syncthreads();                // sv free (c > 0), su/sd ready
		for (int q = threadIdx.x; q < n3; q += blockDim.x)
			sv[q] = v.p[c][b0 + q];
		__syncthreads();
		for (int r = threadIdx.x; r < n3; r += blockDim.x) {
			const int i = r % n, j = (r / n) % n, k = r / n2;
			const long long q = b0 + r;
			// D_b v_c, b = xi, eta, zeta (dref's order)
			double jv[3] = { 0.0, 0.0, 0.0 };
			double tb[3] = { 0.0, 0.0, 0.0 };
			const double *rx = sv + (k * n + j) * n;
			const double *ry = sv + k * n2 + i;
			const double *rz = sv + j * n + i;
			const double *ux = su + (k * n + j) * n;
			const double *uy = su + n3 + k * n2 + i;
			const double *uz = su + 2 * n3 + j * n + i;
			for (int l = 0; l < n; ++l) {
				jv[0] += sd[i * n + l] * rx[l];
				jv[1] += sd[j * n + l] * ry[l * n];
				jv[2] += sd[k * n + l] * rz[l * n2];
				tb[0] += sd[i * n + l] * (ux[l] * rx[l]);
				tb[1] += sd[j * n + l] * (uy[l * n] * ry[l * n]);
				tb[2] += sd[k * n + l] * (uz[l * n2] * rz[l * n2]);
			}
			double tv[9];
#pragma unroll
			for (int a = 0; a < 9; ++a)
				tv[a] = ti.p[a][q];
			const double wj = mass[q];
			double cv = 0.0;
#pragma unroll
			for (int a = 0; a < 3; ++a) {
				double dv = 0.0;
#pragma unroll
				for (int b = 0; b < 3; ++b) {
					const double t = tv[3 * a + b] * jv[b];
					dv = b == 0 ? t : t + dv;
				}
				const double t = su[a * n3 + r] * dv;
				cv = a == 0 ? t : t + cv;
			}
			cv = cv * wj;
			double cs = 0.0;
#pragma unroll
			for (int a = 0; a < 3; ++a) {
				double d = 0.0;
#pragma unroll
				for (int b = 0; b < 3; ++b) {
					const double t = tv[3 * a + b] * tb[b];
					d = b == 0 ? t : t + d;
				}
				cs = a == 0 ? d : cs + d;
			}
			cs = cs * wj;
			out.p[c][q] = 0.5 * cs + 0.5 * cv;
		}
	}
}

// The whole dealiased operator of one element in one workgroup
// (antialiased_convection_operator, convection.c:37-69): u and v are
// interpolated from n to nc points into LDS (interpolate_to_grid, the three
// contractions of k_prolong), the skew-symmetric operator is formed at the
// nc points as in k_ns_conv_elem, and each component is integrated back to
// n points (integrate_to_grid, the contractions of k_restrict).  Every sum
// runs in the order of those kernels.  prol[p*nc+i] = l_p(xi_i) (p < n
// coarse, i < nc fine).
#define CONV_SCR(n, nc) ((n) * (n) * (n) + (n) * (n) * (nc) + (n) * (nc) * (nc))

// dst[nc^3] = the interpolation to nc points of src[n^3] (global); scr:
// CONV_SCR doubles
__device__ __forceinline__ void conv_prolong(int n, int nc,
		const double *__restrict__ sp, const double *__restrict__ src,
		double *__restrict__ dst, double *__restrict__ scr)
{
	double *in = scr, *t1 = scr + n * n * n, *t2 = t1 + n * n * nc;
	const int n3 = n * n * n;
	for (int q = threadIdx.x; q < n3; q += blockDim.x)
		in[q] = src[q];
	__syncthreads();
	for (int q = threadIdx.x; q < n * n * nc; q += blockDim.x) {
		const int i = q % nc, kj = q / nc;
		double s = 0.0;
		for (int p = 0; p < n; ++p)
			s += sp[p * nc + i] * in[kj * n + p];
		t1[q] = s;
	}
	__syncthreads();
	for (int q = threadIdx.x; q < n * nc * nc; q += blockDim.x) {
		const int i = q % nc, jq = (q / nc) % nc, k = q / (nc * nc);
		double s = 0.0;
		for (int l = 0; l < n; ++l)
			s += t1[(k * n + l) * nc + i] * sp[l * nc + jq];
		t2[q] = s;
	}
	__syncthreads();
	for (int q = threadIdx.x; q < nc * nc * nc; q += blockDim.x) {
		const int ji = q % (nc * nc), k = q / (nc * nc);
		double s = 0.0;
		for (int l = 0; l < n; ++l)
			s += t2[l * nc * nc + ji] * sp[l * nc + k];
		dst[q] = s;
	}
	__syncthreads();
}

// dst[n^3] (global) = the integration to n points of src[nc^3] (LDS)
__device__ __forceinline__ void conv_restrict(int n, int nc,
		const double *__restrict__ sp, const double *__restrict__ src,
		double *__restrict__ dst, double *__restrict__ scr)
{
	double *t1 = scr, *t2 = scr + nc * nc * n;
	for (int q = threadIdx.x; q < nc * nc * n; q += blockDim.x) {
		const int p = q % n, kj = q / n;
		double s = 0.0;
		for (int c = 0; c < nc; ++c)
			s += sp[p * nc + c] * src[kj * nc + c];
		t1[q] = s;
	}
	__syncthreads();
	for (int q = threadIdx.x; q < nc * n * n; q += blockDim.x) {
		const int i = q % n, jq = (q / n) % n, k = q / (n * n);
		double s = 0.0;
		for (int l = 0; l < nc; ++l)
			s += t1[(k * nc + l) * n + i] * sp[jq * nc + l];
		t2[q] = s;
	}
	__syncthreads();
	for (int q = threadIdx.x; q < n * n * n; q += blockDim.x) {
		const int ji = q % (n * n), k = q / (n * n);
		double s = 0.0;
		for (int l = 0; l < nc; ++l)
			s += t2[l * n * n + ji] * sp[k * nc + l];
		dst[q] = s;
	}
	__syncthreads();
}

#define CONV_PTS 4                  // nc^3 / 1024 rounded up (nc <= 15)

// the skew-symmetric operator at node r of the element (k_ns_conv_elem's
// arithmetic): su the three u components, sv v_c, sd D, all at nc points
__device__ __forceinline__ double conv_point(int nc, int r,
		const double *__restrict__ su, const double *__restrict__ sv,
		const double *__restrict__ sd, const P9 &ti,
		const double *__restrict__ mass, long long q)
{
	const int c2 = nc * nc, c3 = c2 * nc;
	const int i = r % nc, j = (r / nc) % nc, k = r / c2;
	double jv[3] = { 0.0, 0.0, 0.0 };
	double tb[3] = { 0.0, 0.0, 0.0 };
	const double *rx = sv + (k * nc + j) * nc;
	const double *ry = sv + k * c2 + i;
	const double *rz = sv + j * nc + i;
	const double *ux = su + (k * nc + j) * nc;
	const double *uy = su + c3 + k * c2 + i;
	const double *uz = su + 2 * c3 + j * nc + i;
	for (int l = 0; l < nc; ++l) {
		jv[0] += sd[i * nc + l] * rx[l];
		jv[1] += sd[j * nc + l] * ry[l * nc];
		jv[2] += sd[k * nc + l] * rz[l * c2];
		tb[0] += sd[i * nc + l] * (ux[l] * rx[l]);
		tb[1] += sd[j * nc + l] * (uy[l * nc] * ry[l * nc]);
		tb[2] += sd[k * nc + l] * (uz[l * c2] * rz[l * c2]);
	}
	double tv[9];
#pragma unroll
	for (int a = 0; a < 9; ++a)
		tv[a] = ti.p[a][q];
	const double wj = mass[q];
	double cv = 0.0;
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		double dv = 0.0;
#pragma unroll
		for (int b = 0; b < 3; ++b) {
			const double t = tv[3 * a + b] * jv[b];
			dv = b == 0 ? t : t + dv;
		}
		const double t = su[a * c3 + r] * dv;
		cv = a == 0 ? t : t + cv;
	}
	cv = cv * wj;
	double cs = 0.0;
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		double d = 0.0;
#pragma unroll
		for (int b = 0; b < 3; ++b) {
			const double t = tv[3 * a + b] * tb[b];
			d = b == 0 ? t : t + d;
		}
		cs = a == 0 ? d : cs + d;
	}
	cs = cs * wj;
	return 0.5 * cs + 0.5 * cv;
}

__global__ void __launch_bounds__(1024)
k_ns_conv_full(int n, int nc, const double *__restrict__ prol,
		const double *__restrict__ dm, P3 u, P3 v, P9 ti,
		const double *__restrict__ mass, M3 out)
{
	extern __shared__ double sm[];
	const int c2 = nc * nc, c3 = c2 * nc, n3 = n * n * n;
	double *sp = sm;                    // n*nc
	double *sd = sp + n * nc;           // nc^2
	double *su = sd + c2;               // u_0, u_1, u_2 at nc
	double *sv = su + 3 * c3;           // v_c at nc, then the result
	double *scr = sv + c3;              // transfer scratch
	const long long e = blockIdx.x;
	for (int q = threadIdx.x; q < n * nc; q += blockDim.x)
		sp[q] = prol[q];
	for (int q = threadIdx.x; q < c2; q += blockDim.x)
		sd[q] = dm[q];
	__syncthreads();
	for (int a = 0; a < 3; ++a)
		conv_prolong(n, nc, sp, u.p[a] + e * n3, su + a * c3, scr);
	const long long b0 = e * c3;
	for (int c = 0; c < 3; ++c) {
		conv_prolong(n, nc, sp, v.p[c] + e * n3, sv, scr);
		// this lane's nodes r = threadIdx.x + 1024 m, results in named
		// registers (no array indexed by m)
		double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
#pragma unroll 1
		for (int m = 0; m < CONV_PTS; ++m) {
			const int r = threadIdx.x + m * 1024;
			if (r >= c3)
				break;
			const double val = conv_point(nc, r, su, sv, sd, ti, mass, b0 + r);
			r0 = m == 0 ? val : r0;
			r1 = m == 1 ? val : r1;
			r2 = m == 2 ? val : r2;
			r3 = m == 3 ? val : r3;
		}
		__syncthreads();                // every read of sv is done
#pragma unroll
		for (int m = 0; m < CONV_PTS; ++m) {
			const int r = threadIdx.x + m * 1024;
			if (r < c3)
				sv[r] = m == 0 ? r0 : m == 1 ? r1 : m == 2 ? r2 : r3;
		}
		__syncthreads();
		conv_restrict(n, nc, sp, sv, out.p[c] + e * n3, scr);
	}
}

// f += (w_b w_a) n_s g on the face nodes of every face s of an element
// whose Dirichlet bit equals `dirichlet` (add_surface_integral,
// poisson.c:145-210: all nodes of face W, then E, S, N, B, T).  One
// workgroup per element; faces in order, so edge and corner nodes take
// their additions in the reference's order.  wn[(e*6+s)*n*n + a*n+b] =
// (w_b w_a) n_s(a,b) for the normal component of the direction.
__global__ void __launch_bounds__(256)
k_ns_surf(int nel, int n, const int *__restrict__ bc, int dirichlet,
		const double *__restrict__ wn, const double *__restrict__ g,
		double *__restrict__ f)
{
	const int e = blockIdx.x;
	const int n2 = n * n, n3 = n2 * n, m = n - 1;
	const int bits = bc[e];
	const long long b0 = (long long) e * n3;
	for (int s = 0; s < 6; ++s) {
		const int isd = (bits >> s) & 1;
		if (isd == dirichlet) {
			for (int t = threadIdx.x; t < n2; t += blockDim.x) {
				const int a = t / n, b = t % n;
				int k, j, i;
				switch (s) {
				case 0: k = a; j = b; i = 0; break;
				case 1: k = a; j = b; i = m; break;
				case 2: k = a; j = 0; i = b; break;
				case 3: k = a; j = m; i = b; break;
				case 4: k = 0; j = a; i = b; break;
				default: k = m; j = a; i = b; break;
				}
				const long long q = b0 + (k * n + j) * n + i;
				f[q] += wn[((long long) e * 6 + s) * n2 + t] * g[q];
			}
		}
		__syncthreads();
	}
}

// u = 0 on the nodes of the element's Dirichlet faces
// (set_dirichlet_boundaries with the reference's zero boundary function,
// poisson.c:101-143, structs.c:67)
__global__ void __launch_bounds__(256)
k_ns_zero_dir(long long nt, int n, const int *__restrict__ bc, M3 u)
{
	const long long q = (long long) blockIdx.x * 256 + threadIdx.x;
	if (q >= nt)
		return;
	const int n3 = n * n * n, m = n - 1;
	const long long e = q / n3;
	const int r = (int) (q - e * n3);
	const int i = r % n, j = (r / n) % n, k = r / (n * n);
	const int bits = bc[e];
	const bool z = ((bits & 1) && i == 0) || ((bits & 2) && i == m)
		|| ((bits & 4) && j == 0) || ((bits & 8) && j == m)
		|| ((bits & 16) && k == 0) || ((bits & 32) && k == m);
	if (z)
		for (int c = 0; c < 3; ++c)
			u.p[c][q] = 0.0;
}

// y_c = ((a0 x0_c + y_c) + a1 x1_c) + a2 x2_c for the three components,
// a term skipped where its flag is clear (the daxpy sequences of
// stiffly_stable, time.c:51-73, and add_interpolated_velocity,
// time.c:132-152)
struct Comb {
	const double *x[3][3];    // [term][component]
	double a[3];
	int use[3];
};

__global__ void __launch_bounds__(256)
k_ns_comb(long long nt, Comb cb, M3 y)
{
	const long long q = (long long) blockIdx.x * 256 + threadIdx.x;
	if (q >= nt)
		return;
	for (int c = 0; c < 3; ++c) {
		double v = y.p[c][q];
		for (int t = 0; t < 3; ++t)
			if (cb.use[t])
				v = cb.a[t] * cb.x[t][c][q] + v;
		y.p[c][q] = v;
	}
}

// z_c = alpha x_c w + beta z_c, w shared by the components or per
// component (multiply, linalg.c:142-152); w NULL: z_c = alpha x_c +
// beta z_c
__global__ void __launch_bounds__(256)
k_ns_mul3(long long nt, double alpha, P3 x, P3 w, double beta, M3 z)
{
	const long long q = (long long) blockIdx.x * 256 + threadIdx.x;
	if (q >= nt)
		return;
	for (int c = 0; c < 3; ++c) {
		if (!z.p[c])
			continue;
		const double wv = w.p[c] ? w.p[c][q] : 1.0;
		const double old = beta != 0.0 ? beta * z.p[c][q] : 0.0;
		z.p[c][q] = alpha * x.p[c][q] * wv + old;
	}
}

extern "C" int ssxk_ns_jac(void *st, size_t nt, int n, const double *dm,
		const double *const a[3], double *const out[9])
{
	P3 pa = { { a[0], a[1], a[2] } };
	M9 po;
	for (int k = 0; k < 9; ++k)
		po.p[k] = out[k];
	hipLaunchKernelGGL(k_ns_jac, dim3(ns_blocks(nt)), dim3(256), 0,
			(hipStream_t) st, (long long) nt, n, dm, pa, po);
	NS_LAUNCH("k_ns_jac");
	return SSX_OK;
}

extern "C" int ssxk_ns_curl(void *st, size_t nt, const double *const J[9],
		const double *const ti[9], const double *mass, double *const out[3])
{
	P9 pj, pt;
	for (int k = 0; k < 9; ++k) {
		pj.p[k] = J[k];
		pt.p[k] = ti[k];
	}
	M3 po = { { out[0], out[1], out[2] } };
	hipLaunchKernelGGL(k_ns_curl, dim3(ns_blocks(nt)), dim3(256), 0,
			(hipStream_t) st, (long long) nt, pj, pt, mass, po);
	NS_LAUNCH("k_ns_curl");
	return SSX_OK;
}

extern "C" int ssxk_ns_div(void *st, size_t nt, const double *const J[9],
		const double *const ti[9], const double *mass, double scale,
		double *out)
{
	P9 pj, pt;
	for (int k = 0; k < 9; ++k) {
		pj.p[k] = J[k];
		pt.p[k] = ti[k];
	}
	hipLaunchKernelGGL(k_ns_div, dim3(ns_blocks(nt)), dim3(256), 0,
			(hipStream_t) st, (long long) nt, pj, pt, mass, scale, out);
	NS_LAUNCH("k_ns_div");
	return SSX_OK;
}

extern "C" int ssxk_ns_grad(void *st, size_t nt, int n, const double *dm,
		const double *p, const double *mass, const double *const ti[9],
		double *const out[3])
{
	P9 pt;
	for (int k = 0; k < 9; ++k)
		pt.p[k] = ti[k];
	M3 po = { { out[0], out[1], out[2] } };
	hipLaunchKernelGGL(k_ns_grad, dim3(ns_blocks(nt)), dim3(256), 0,
			(hipStream_t) st, (long long) nt, n, dm, p, mass, pt, po);
	NS_LAUNCH("k_ns_grad");
	return SSX_OK;
}

extern "C" int ssxk_ns_conv(void *st, size_t nt, int n, const double *dm,
		const double *const u[3], const double *const v[3],
		const double *const Jv[9], const double *const ti[9],
		const double *mass, double *const out[3])
{
	P3 pu = { { u[0], u[1], u[2] } }, pv = { { v[0], v[1], v[2] } };
	P9 pj, pt;
	for (int k = 0; k < 9; ++k) {
		pj.p[k] = Jv[k];
		pt.p[k] = ti[k];
	}
	M3 po = { { out[0], out[1], out[2] } };
	hipLaunchKernelGGL(k_ns_conv, dim3(ns_blocks(nt)), dim3(256), 0,
			(hipStream_t) st, (long long) nt, n, dm, pu, pv, pj, pt, mass,
			po);
	NS_LAUNCH("k_ns_conv");
	return SSX_OK;
}

extern "C" int ssxk_ns_conv_elem(void *st, int nel, int n, const double *dm,
		const double *const u[3], const double *const v[3],
		const double *const ti[9], const double *mass, double *const out[3])
{
	const size_t bytes = sizeof(double) * (4 * (size_t) n * n * n
			+ (size_t) n * n);
	if (n < 2 || bytes > 160 * 1024) {
		ssx_set_error("k_ns_conv_elem: n = %d does not fit LDS", n);
		return SSX_EINVAL;
	}
	P3 pu = { { u[0], u[1], u[2] } }, pv = { { v[0], v[1], v[2] } };
	P9 pt;
	for (int k = 0; k < 9; ++k)
		pt.p[k] = ti[k];
	M3 po = { { out[0], out[1], out[2] } };
	if (hipFuncSetAttribute((const void *) k_ns_conv_elem,
			hipFuncAttributeMaxDynamicSharedMemorySize, (int) bytes)
			!= hipSuccess) {
		ssx_set_error("k_ns_conv_elem: %zu bytes of LDS refused", bytes);
		return SSX_EHIP;
	}
	hipLaunchKernelGGL(k_ns_conv_elem, dim3(nel), dim3(1024), bytes,
			(hipStream_t) st, n, dm, pu, pv, pt, mass, po);
	NS_LAUNCH("k_ns_conv_elem");
	return SSX_OK;
}

// LDS of k_ns_conv_full, bytes
extern "C" size_t ssxk_ns_conv_full_bytes(int n, int nc)
{
	const size_t scr_p = CONV_SCR((size_t) n, (size_t) nc);
	const size_t scr_r = (size_t) nc * nc * n + (size_t) nc * n * n;
	return sizeof(double) * ((size_t) n * nc + (size_t) nc * nc
			+ 4 * (size_t) nc * nc * nc + (scr_p > scr_r ? scr_p : scr_r));
}

extern "C" int ssxk_ns_conv_full(void *st, int nel, int n, int nc,
		const double *prol, const double *dm, const double *const u[3],
		const double *const v[3], const double *const ti[9],
		const double *mass, double *const out[3])
{
	const size_t bytes = ssxk_ns_conv_full_bytes(n, nc);
	if (n < 2 || nc < n || nc * nc * nc > CONV_PTS * 1024
			|| bytes > 160 * 1024) {
		ssx_set_error("k_ns_conv_full: n = %d, nc = %d unsupported", n, nc);
		return SSX_EINVAL;
	}
	P3 pu = { { u[0], u[1], u[2] } }, pv = { { v[0], v[1], v[2] } };
	P9 pt;
	for (int k = 0; k < 9; ++k)
		pt.p[k] = ti[k];
	M3 po = { { out[0], out[1], out[2] } };
	if (hipFuncSetAttribute((const void *) k_ns_conv_full,
			hipFuncAttributeMaxDynamicSharedMemorySize, (int) bytes)
			!= hipSuccess) {
		ssx_set_error("k_ns_conv_full: %zu bytes of LDS refused", bytes);
		return SSX_EHIP;
	}
	hipLaunchKernelGGL(k_ns_conv_full, dim3(nel), dim3(1024), bytes,
			(hipStream_t) st, n, nc, prol, dm, pu, pv, pt, mass, po);
	NS_LAUNCH("k_ns_conv_full");
	return SSX_OK;
}

extern "C" int ssxk_ns_surf(void *st, int nel, int n, const int *bc,
		int dirichlet, const double *wn, const double *g, double *f)
{
	hipLaunchKernelGGL(k_ns_surf, dim3(nel), dim3(256), 0, (hipStream_t) st,
			nel, n, bc, dirichlet, wn, g, f);
	NS_LAUNCH("k_ns_surf");
	return SSX_OK;
}

extern "C" int ssxk_ns_zero_dir(void *st, size_t nt, int n, const int *bc,
		double *const u[3])
{
	M3 pu = { { u[0], u[1], u[2] } };
	hipLaunchKernelGGL(k_ns_zero_dir, dim3(ns_blocks(nt)), dim3(256), 0,
			(hipStream_t) st, (long long) nt, n, bc, pu);
	NS_LAUNCH("k_ns_zero_dir");
	return SSX_OK;
}

extern "C" int ssxk_ns_comb(void *st, size_t nt, const double *const x[3][3],
		const double a[3], const int use[3], double *const y[3])
{
	Comb cb;
	for (int t = 0; t < 3; ++t) {
		cb.a[t] = a[t];
		cb.use[t] = use[t];
		for (int c = 0; c < 3; ++c)
			cb.x[t][c] = x[t][c];
	}
	M3 py = { { y[0], y[1], y[2] } };
	hipLaunchKernelGGL(k_ns_comb, dim3(ns_blocks(nt)), dim3(256), 0,
			(hipStream_t) st, (long long) nt, cb, py);
	NS_LAUNCH("k_ns_comb");
	return SSX_OK;
}

extern "C" int ssxk_ns_mul3(void *st, size_t nt, double alpha,
		const double *const x[3], const double *const w[3], double beta,
		double *const z[3])
{
	P3 px = { { x[0], x[1], x[2] } };
	P3 pw = { { w ? w[0] : nullptr, w ? w[1] : nullptr, w ? w[2] : nullptr } };
	M3 pz = { { z[0], z[1], z[2] } };
	hipLaunchKernelGGL(k_ns_mul3, dim3(ns_blocks(nt)), dim3(256), 0,
			(hipStream_t) st, (long long) nt, alpha, px, pw, beta, pz);
	NS_LAUNCH("k_ns_mul3");
	return SSX_OK;
}
