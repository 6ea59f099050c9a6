/*
 * ssx_mesh.c -- structured hexahedral mesh generator (host C).
 *
 * Connectivity follows the conventions the reference's grid generator
 * writes into grid.h5 (misc/rectangular_grid.py:56-135): element id
 * i + j*mx + k*mx*my, face neighbours W,E,S,N,B,T, the 12 edge neighbours
 * across the diagonals (sw,se,nw,ne | bw,be,tw,te | bs,bn,ts,tn) and the 8
 * corner neighbours, all wrapped periodically; then, face by face in order
 * W..T, a boundary face with code bc < 0 overwrites that face's neighbour,
 * its 4 edges and its 4 corners with the code and sets their sizes to 0
 * (rectangular_grid.py:56-78).
 *
 * Unlike the reference generator, which writes only the 8 vertices (file
 * n = 2, straight-sided elements), nodes are written at all n^3 GLL points so
 * that curved deformations reach the geometric factors; the reference's
 * reader then interpolates file n to run n with the identity
 * (dataio.c:407-452 with interpolation.c:31-39).
 */
#include <math.h>
#include <string.h>
#include "ssx_internal.h"

static int wrap(int i, int n)
{
	return i - (int) floor((double) i / (double) n) * n;
}

/* rectangular_grid.py:70-73 */
static const int edge_of_face[6][4] = {
	{ 0, 2, 4, 6 }, { 1, 3, 5, 7 }, { 0, 1, 8, 10 },
	{ 2, 3, 9, 11 }, { 4, 5, 8, 9 }, { 6, 7, 10, 11 } };
static const int corner_of_face[6][4] = {
	{ 0, 2, 4, 6 }, { 1, 3, 5, 7 }, { 0, 1, 4, 5 },
	{ 2, 3, 6, 7 }, { 0, 1, 2, 3 }, { 4, 5, 6, 7 } };
/* neighbour offsets of the 12 edges (rectangular_grid.py:120-127) */
static const int edge_off[12][3] = {
	{ -1, -1, 0 }, { 1, -1, 0 }, { -1, 1, 0 }, { 1, 1, 0 },
	{ -1, 0, -1 }, { 1, 0, -1 }, { -1, 0, 1 }, { 1, 0, 1 },
	{ 0, -1, -1 }, { 0, 1, -1 }, { 0, -1, 1 }, { 0, 1, 1 } };

int ssx_mesh_box(int n, int mx, int my, int mz, const double lo[3],
		const double hi[3], const int bc[6], int deform, double amp,
		double *x, double *y, double *z, int *neigh, int *edge_neigh,
		int *edge_size, int *corn_neigh, int *corn_size)
{
	if (n < SSX_NMIN || n > SSX_NMAX || mx < 1 || my < 1 || mz < 1) {
		ssx_set_error("ssx_mesh_box: bad size n=%d m=(%d,%d,%d)", n, mx,
				my, mz);
		return SSX_EINVAL;
	}
	ssx_basis b;
	ssx_basis_init(&b, n, 1, 0.0);
	const int me[3] = { mx, my, mz };
	const int n3 = n * n * n;
	for (int k = 0; k < mz; ++k)
	for (int j = 0; j < my; ++j)
	for (int i = 0; i < mx; ++i) {
		const int e = i + j * mx + k * mx * my;
		const int ix[3] = { i, j, k };
		/* coordinates: t = (element index + (1+xi)/2) / m along each
		 * axis, so nodes shared by two elements are bit-identical */
		for (int c = 0; c < n; ++c)
		for (int bb = 0; bb < n; ++bb)
		for (int a = 0; a < n; ++a) {
			const int loc[3] = { a, bb, c };
			double t[3], p[3];
			for (int d = 0; d < 3; ++d) {
				t[d] = ((double) ix[d] + (1.0 + b.xi[loc[d]]) * 0.5)
					/ (double) me[d];
				p[d] = lo[d] + (hi[d] - lo[d]) * t[d];
			}
			if (deform == 1) {
				const double s = amp * sin(M_PI * t[0])
					* sin(M_PI * t[1]) * sin(M_PI * t[2]);
				for (int d = 0; d < 3; ++d)
					p[d] += s * (hi[d] - lo[d]);
			} else if (deform == 2) {
				const double eta = 2.0 * t[1] - 1.0;
				const double ys = tanh(amp * eta) / tanh(amp);
				p[1] = lo[1] + (hi[1] - lo[1]) * 0.5 * (ys + 1.0);
			}
			const size_t o = (size_t) e * n3 + (c * n + bb) * n + a;
			x[o] = p[0];
			y[o] = p[1];
			z[o] = p[2];
		}
		for (int f = 0; f < 6; ++f) {       /* rectangular_grid.py:113-118 */
			int d[3] = { 0, 0, 0 };
			d[f / 2] = 2 * (f % 2) - 1;
			neigh[e * 6 + f] = wrap(i + d[0], mx) + wrap(j + d[1], my) * mx
				+ wrap(k + d[2], mz) * mx * my;
		}
		for (int q = 0; q < 12; ++q) {
			const int *d = edge_off[q];
			edge_neigh[e * 12 + q] = wrap(i + d[0], mx)
				+ wrap(j + d[1], my) * mx + wrap(k + d[2], mz) * mx * my;
			edge_size[e * 12 + q] = 1;
		}
		for (int q = 0; q < 8; ++q) {        /* rectangular_grid.py:129-135 */
			const int d0 = 2 * (q % 2) - 1, d1 = 2 * ((q / 2) % 2) - 1,
				d2 = 2 * ((q / 4) % 2) - 1;
			corn_neigh[e * 8 + q] = wrap(i + d0, mx) + wrap(j + d1, my) * mx
				+ wrap(k + d2, mz) * mx * my;
			corn_size[e * 8 + q] = 1;
		}
	}
	/* boundary faces, in face order (rectangular_grid.py:56-78) */
	for (int s = 0; s < 6; ++s) {
		if (bc[s] >= 0)
			continue;
		const int ax = s / 2, side = s % 2;
		for (int k = 0; k < mz; ++k)
		for (int j = 0; j < my; ++j)
		for (int i = 0; i < mx; ++i) {
			const int ix[3] = { i, j, k };
			if (ix[ax] != side * (me[ax] - 1))
				continue;
			const int e = i + j * mx + k * mx * my;
			neigh[e * 6 + s] = bc[s];
			for (int q = 0; q < 4; ++q) {
				edge_neigh[e * 12 + edge_of_face[s][q]] = bc[s];
				edge_size[e * 12 + edge_of_face[s][q]] = 0;
				corn_neigh[e * 8 + corner_of_face[s][q]] = bc[s];
				corn_size[e * 8 + corner_of_face[s][q]] = 0;
			}
		}
	}
	return SSX_OK;
}
