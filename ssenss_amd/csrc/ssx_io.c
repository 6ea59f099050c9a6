/*
 * ssx_io.c -- SSENSS's grid / velocity HDF5 files and setup input
 * (include/ssx_io.h).  Serial HDF5: the whole "Element array" is read or
 * written by one process (the reference reads its rank's hyperslab,
 * dataio.c:489-529; with one rank that is the whole array).
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>

#include <hdf5.h>

#include "ssx_internal.h"        /* GLL basis, interpolation (libssx.so) */
#include "ssx_io.h"

static __thread char g_err[512];

const char *ssx_io_last_error(void)
{
	return g_err;
}

static int fail(int code, const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof(g_err), fmt, ap);
	va_end(ap);
	return code;
}

#define NINT 52                 /* ints per grid record */

/* byte offsets of a grid record of n3 points: the member order of the
 * format's writer (misc/rectangular_grid.py:8-17) */
static void grid_offsets(size_t n3, size_t off[9], size_t *size)
{
	const size_t I = sizeof(int), D = sizeof(double);
	off[0] = 0;             /* Neighbours [6]          */
	off[1] = 6 * I;         /* Neighbour rotation [6]  */
	off[2] = 12 * I;        /* Corner size [8]         */
	off[3] = 20 * I;        /* Corner neighbours [8]   */
	off[4] = 28 * I;        /* Edge size [12]          */
	off[5] = 40 * I;        /* Edge neighbours [12]    */
	off[6] = NINT * I;      /* X [n3]                  */
	off[7] = NINT * I + n3 * D;
	off[8] = NINT * I + 2 * n3 * D;
	*size = NINT * I + 3 * n3 * D;
}

static const char *gnames[9] = { "Neighbours", "Neighbour rotation",
	"Corner size", "Corner neighbours", "Edge size", "Edge neighbours",
	"X", "Y", "Z" };
static const int glen[6] = { 6, 6, 8, 8, 12, 12 };

static hid_t grid_type(size_t n3)
{
	size_t off[9], size;
	grid_offsets(n3, off, &size);
	hid_t t = H5Tcreate(H5T_COMPOUND, size);
	for (int m = 0; m < 9; ++m) {
		hsize_t dim = m < 6 ? (hsize_t) glen[m] : (hsize_t) n3;
		hid_t a = H5Tarray_create2(m < 6 ? H5T_NATIVE_INT
				: H5T_NATIVE_DOUBLE, 1, &dim);
		H5Tinsert(t, gnames[m], off[m], a);
		H5Tclose(a);
	}
	return t;
}

static hid_t vel_type(size_t n3)
{
	const size_t D = sizeof(double);
	hid_t t = H5Tcreate(H5T_COMPOUND, 3 * n3 * D);
	static const char *nm[3] = { "U", "V", "W" };
	hsize_t dim = (hsize_t) n3;
	hid_t a = H5Tarray_create2(H5T_NATIVE_DOUBLE, 1, &dim);
	for (int m = 0; m < 3; ++m)
		H5Tinsert(t, nm[m], m * n3 * D, a);
	H5Tclose(a);
	return t;
}

typedef struct {
	int n, nel;
} fparams;

static hid_t param_type(void)
{
	hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(fparams));
	H5Tinsert(t, "n", HOFFSET(fparams, n), H5T_NATIVE_INT);
	H5Tinsert(t, "nel", HOFFSET(fparams, nel), H5T_NATIVE_INT);
	return t;
}

static hid_t open_ro(const char *path)
{
	H5E_auto2_t f;
	void *d;
	H5Eget_auto2(H5E_DEFAULT, &f, &d);
	H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
	hid_t id = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
	H5Eset_auto2(H5E_DEFAULT, f, d);
	return id;
}

int ssx_io_file_params(const char *path, int *n, int *nel)
{
	if (!path || !n || !nel)
		return fail(SSX_EINVAL, "ssx_io_file_params: null argument");
	hid_t f = open_ro(path);
	if (f < 0)
		return fail(SSX_EINVAL, "unable to open '%s'", path);
	hid_t d = H5Dopen2(f, "Parameters", H5P_DEFAULT);
	fparams p = { 0, 0 };
	hid_t t = param_type();
	herr_t e = d < 0 ? -1 : H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, &p);
	H5Tclose(t);
	if (d >= 0)
		H5Dclose(d);
	H5Fclose(f);
	if (e < 0)
		return fail(SSX_EINVAL, "'%s': no readable Parameters dataset",
				path);
	*n = p.n;
	*nel = p.nel;
	return SSX_OK;
}

/* the whole Element array of type t into buf (nel records) */
static int read_elements(const char *path, hid_t t, int nel_expect,
		void *buf)
{
	hid_t f = open_ro(path);
	if (f < 0)
		return fail(SSX_EINVAL, "unable to open '%s'", path);
	hid_t d = H5Dopen2(f, "Element array", H5P_DEFAULT);
	int rc = SSX_OK;
	if (d < 0) {
		rc = fail(SSX_EINVAL, "'%s': no Element array", path);
	} else {
		hid_t s = H5Dget_space(d);
		hsize_t dim = 0;
		if (H5Sget_simple_extent_ndims(s) != 1
				|| H5Sget_simple_extent_dims(s, &dim, NULL) != 1
				|| (int) dim != nel_expect)
			rc = fail(SSX_EINVAL, "'%s': Element array holds %llu records,"
					" expected %d", path, (unsigned long long) dim,
					nel_expect);
		else if (H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0)
			rc = fail(SSX_EINVAL, "'%s': Element array not readable as "
					"this record type", path);
		H5Sclose(s);
		H5Dclose(d);
	}
	H5Fclose(f);
	return rc;
}

static int write_file(const char *path, int n, int nel, hid_t t,
		const void *buf)
{
	hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
	if (f < 0)
		return fail(SSX_EINVAL, "unable to create '%s'", path);
	int rc = SSX_OK;
	hsize_t one = 1, dim = (hsize_t) nel;
	fparams p = { n, nel };
	hid_t pt = param_type();
	hid_t ps = H5Screate_simple(1, &one, NULL);
	hid_t pd = H5Dcreate2(f, "Parameters", pt, ps, H5P_DEFAULT, H5P_DEFAULT,
			H5P_DEFAULT);
	if (pd < 0 || H5Dwrite(pd, pt, H5S_ALL, H5S_ALL, H5P_DEFAULT, &p) < 0)
		rc = fail(SSX_EHIP, "'%s': writing Parameters failed", path);
	if (pd >= 0)
		H5Dclose(pd);
	H5Sclose(ps);
	H5Tclose(pt);
	hid_t es = H5Screate_simple(1, &dim, NULL);
	hid_t ed = H5Dcreate2(f, "Element array", t, es, H5P_DEFAULT,
			H5P_DEFAULT, H5P_DEFAULT);
	if (rc == SSX_OK && (ed < 0 || H5Dwrite(ed, t, H5S_ALL, H5S_ALL,
			H5P_DEFAULT, buf) < 0))
		rc = fail(SSX_EHIP, "'%s': writing the Element array failed", path);
	if (ed >= 0)
		H5Dclose(ed);
	H5Sclose(es);
	if (H5Fclose(f) < 0 && rc == SSX_OK)
		rc = fail(SSX_EHIP, "'%s': closing failed", path);
	return rc;
}

/* file points nf -> run points nr of one element's n^3 block
 * (interpolate_to_grid, interpolation.c:136-155) */
static void interp_elem(const double *A, int nf, int nr, const double *u,
		double *o, double *t1, double *t2)
{
	for (int k = 0; k < nf; ++k)
		for (int j = 0; j < nf; ++j)
			for (int i = 0; i < nr; ++i) {
				double s = 0.0;
				for (int p = 0; p < nf; ++p)
					s += A[p * nr + i] * u[(k * nf + j) * nf + p];
				t1[(k * nf + j) * nr + i] = s;
			}
	for (int k = 0; k < nf; ++k)
		for (int j = 0; j < nr; ++j)
			for (int i = 0; i < nr; ++i) {
				double s = 0.0;
				for (int p = 0; p < nf; ++p)
					s += t1[(k * nf + p) * nr + i] * A[p * nr + j];
				t2[(k * nr + j) * nr + i] = s;
			}
	for (int k = 0; k < nr; ++k)
		for (int ji = 0; ji < nr * nr; ++ji) {
			double s = 0.0;
			for (int p = 0; p < nf; ++p)
				s += t2[p * nr * nr + ji] * A[p * nr + k];
			o[k * nr * nr + ji] = s;
		}
}

static int interp_setup(int nf, int nr, double *A)
{
	if (nf < 2 || nf > 16 || nr < 2 || nr > 16)
		return fail(SSX_EINVAL, "points per direction must be in [2,16] "
				"(file %d, run %d)", nf, nr);
	ssx_basis bf, br;
	ssx_basis_init(&bf, nf, (2 * nf) / 3, 0.0);
	ssx_basis_init(&br, nr, (2 * nr) / 3, 0.0);
	ssx_interp(bf.xi, nf, br.xi, nr, A);
	return SSX_OK;
}

int ssx_io_read_grid(const char *path, int n_run, int nel, double *x,
		double *y, double *z, int *neigh, int *neigh_rot, int *edge_neigh,
		int *edge_size, int *corn_neigh, int *corn_size)
{
	if (!path || !x || !y || !z || !neigh || !edge_neigh || !edge_size
			|| !corn_neigh || !corn_size)
		return fail(SSX_EINVAL, "ssx_io_read_grid: null argument");
	int nf, nelf;
	int rc = ssx_io_file_params(path, &nf, &nelf);
	if (rc)
		return rc;
	if (nelf != nel)
		return fail(SSX_EINVAL, "'%s' holds %d elements, expected %d", path,
				nelf, nel);
	double A[256];
	if ((rc = interp_setup(nf, n_run, A)))
		return rc;
	const size_t n3 = (size_t) nf * nf * nf, r3 = (size_t) n_run * n_run
		* n_run;
	size_t off[9], size;
	grid_offsets(n3, off, &size);
	char *buf = malloc(size * (size_t) nel);
	double *t = malloc(sizeof(double) * 2 * 16 * 16 * 16);
	if (!buf || !t) {
		free(buf);
		free(t);
		return fail(SSX_ENOMEM, "out of host memory");
	}
	hid_t gt = grid_type(n3);
	rc = read_elements(path, gt, nel, buf);
	H5Tclose(gt);
	for (int e = 0; e < nel && rc == SSX_OK; ++e) {
		const char *r = buf + (size_t) e * size;
		int *dst[6] = { neigh + 6 * e, neigh_rot ? neigh_rot + 6 * e : NULL,
			corn_size + 8 * e, corn_neigh + 8 * e, edge_size + 12 * e,
			edge_neigh + 12 * e };
		for (int m = 0; m < 6; ++m)
			if (dst[m])
				memcpy(dst[m], r + off[m], sizeof(int) * glen[m]);
		double *out[3] = { x + e * r3, y + e * r3, z + e * r3 };
		for (int a = 0; a < 3; ++a)
			interp_elem(A, nf, n_run, (const double *) (r + off[6 + a]),
					out[a], t, t + 16 * 16 * 16);
	}
	free(buf);
	free(t);
	return rc;
}

int ssx_io_write_grid(const char *path, int n, int nel, const double *x,
		const double *y, const double *z, const int *neigh,
		const int *neigh_rot, const int *edge_neigh, const int *edge_size,
		const int *corn_neigh, const int *corn_size)
{
	if (!path || !x || !y || !z || !neigh || !edge_neigh || !edge_size
			|| !corn_neigh || !corn_size || n < 2 || nel < 1)
		return fail(SSX_EINVAL, "ssx_io_write_grid: bad argument");
	const size_t n3 = (size_t) n * n * n;
	size_t off[9], size;
	grid_offsets(n3, off, &size);
	char *buf = calloc((size_t) nel, size);
	if (!buf)
		return fail(SSX_ENOMEM, "out of host memory");
	for (int e = 0; e < nel; ++e) {
		char *r = buf + (size_t) e * size;
		const int *src[6] = { neigh + 6 * e, neigh_rot ? neigh_rot + 6 * e
			: NULL, corn_size + 8 * e, corn_neigh + 8 * e,
			edge_size + 12 * e, edge_neigh + 12 * e };
		for (int m = 0; m < 6; ++m)
			if (src[m])
				memcpy(r + off[m], src[m], sizeof(int) * glen[m]);
		memcpy(r + off[6], x + e * n3, sizeof(double) * n3);
		memcpy(r + off[7], y + e * n3, sizeof(double) * n3);
		memcpy(r + off[8], z + e * n3, sizeof(double) * n3);
	}
	hid_t gt = grid_type(n3);
	const int rc = write_file(path, n, nel, gt, buf);
	H5Tclose(gt);
	free(buf);
	return rc;
}

int ssx_io_read_velocity(const char *path, int n_run, int nel, double *u,
		double *v, double *w)
{
	if (!path || !u || !v || !w)
		return fail(SSX_EINVAL, "ssx_io_read_velocity: null argument");
	int nf, nelf;
	int rc = ssx_io_file_params(path, &nf, &nelf);
	if (rc)
		return rc;
	if (nelf != nel)                 /* dataio.c:165-166 */
		return fail(SSX_EINVAL, "'%s' holds %d elements, expected %d", path,
				nelf, nel);
	double A[256];
	if ((rc = interp_setup(nf, n_run, A)))
		return rc;
	const size_t n3 = (size_t) nf * nf * nf, r3 = (size_t) n_run * n_run
		* n_run;
	double *buf = malloc(sizeof(double) * 3 * n3 * (size_t) nel);
	double *t = malloc(sizeof(double) * 2 * 16 * 16 * 16);
	if (!buf || !t) {
		free(buf);
		free(t);
		return fail(SSX_ENOMEM, "out of host memory");
	}
	hid_t vt = vel_type(n3);
	rc = read_elements(path, vt, nel, buf);
	H5Tclose(vt);
	for (int e = 0; e < nel && rc == SSX_OK; ++e) {
		double *out[3] = { u + e * r3, v + e * r3, w + e * r3 };
		for (int a = 0; a < 3; ++a)
			interp_elem(A, nf, n_run, buf + (3 * (size_t) e + a) * n3,
					out[a], t, t + 16 * 16 * 16);
	}
	free(buf);
	free(t);
	return rc;
}

int ssx_io_write_velocity(const char *path, int n, int nel, const double *u,
		const double *v, const double *w)
{
	if (!path || !u || !v || !w || n < 2 || nel < 1)
		return fail(SSX_EINVAL, "ssx_io_write_velocity: bad argument");
	const size_t n3 = (size_t) n * n * n;
	double *buf = malloc(sizeof(double) * 3 * n3 * (size_t) nel);
	if (!buf)
		return fail(SSX_ENOMEM, "out of host memory");
	for (int e = 0; e < nel; ++e) {
		memcpy(buf + (3 * (size_t) e + 0) * n3, u + e * n3, n3 * 8);
		memcpy(buf + (3 * (size_t) e + 1) * n3, v + e * n3, n3 * 8);
		memcpy(buf + (3 * (size_t) e + 2) * n3, w + e * n3, n3 * 8);
	}
	hid_t vt = vel_type(n3);
	const int rc = write_file(path, n, nel, vt, buf);
	H5Tclose(vt);
	free(buf);
	return rc;
}

/* ---- setup input ---- */

void ssx_io_setup_defaults(ssx_setup *s)
{
	memset(s, 0, sizeof(*s));
	s->dt = 1.0;
	s->visc = 1.0;
	s->dens = 1.0;
	s->tmax = 1.0;
	s->n = 2;
	s->write_every = 1000;
	strcpy(s->gridfile, "dat/grid.h5");
	strcpy(s->initfile, "dat/init.h5");
	strcpy(s->datdir, "dat");
}

static void copy_str(char *dst, const char *src)
{
	snprintf(dst, 100, "%s", src);
}

/* one "name: value" of the parameters block (set_parameter_value) */
static void set_field(ssx_setup *s, const char *name, const char *val)
{
	if (!strcmp(name, "dt"))
		s->dt = atof(val);
	else if (!strcmp(name, "tmax"))
		s->tmax = atof(val);
	else if (!strcmp(name, "visc"))
		s->visc = atof(val);
	else if (!strcmp(name, "dens"))
		s->dens = atof(val);
	else if (!strcmp(name, "gx"))
		s->gx = atof(val);
	else if (!strcmp(name, "gy"))
		s->gy = atof(val);
	else if (!strcmp(name, "gz"))
		s->gz = atof(val);
	else if (!strcmp(name, "n"))
		s->n = atoi(val);
	else if (!strcmp(name, "write_every"))
		s->write_every = atoi(val);
	else if (!strcmp(name, "datdir"))
		copy_str(s->datdir, val);
	else if (!strcmp(name, "gridfile"))
		copy_str(s->gridfile, val);
	else if (!strcmp(name, "initfile"))
		copy_str(s->initfile, val);
}

/* The file is read with all white space dropped (input.c:88-115 skips it
 * character by character); blocks are "#name{...}", fields "name:value;". */
int ssx_io_read_setup(const char *path, ssx_setup *s)
{
	if (!path || !s)
		return fail(SSX_EINVAL, "ssx_io_read_setup: null argument");
	FILE *fp = fopen(path, "r");
	if (!fp)
		return fail(SSX_EINVAL, "unable to read setup file '%s'", path);
	size_t cap = 4096, len = 0;
	char *txt = malloc(cap);
	int c;
	while (txt && (c = getc(fp)) != EOF) {
		if (isspace(c))
			continue;
		if (len + 1 >= cap) {
			char *t = realloc(txt, cap *= 2);
			if (!t) {
				free(txt);
				txt = NULL;
				break;
			}
			txt = t;
		}
		txt[len++] = (char) c;
	}
	fclose(fp);
	if (!txt)
		return fail(SSX_ENOMEM, "out of host memory");
	txt[len] = '\0';
	int found = 0;
	for (char *p = strchr(txt, '#'); p && !found; p = strchr(p + 1, '#')) {
		char *open = strchr(p, '{');
		if (!open)
			break;
		*open = '\0';
		const int is_params = !strcmp(p + 1, "parameters");
		char *close = strchr(open + 1, '}');
		if (!close)
			break;
		*close = '\0';
		if (is_params) {
			found = 1;
			for (char *f = open + 1; *f;) {
				char *semi = strchr(f, ';');
				if (!semi)
					break;
				*semi = '\0';
				char *colon = strchr(f, ':');
				if (colon) {
					*colon = '\0';
					set_field(s, f, colon + 1);
				}
				f = semi + 1;
			}
		}
		p = close;
	}
	free(txt);
	return SSX_OK;
}

int ssx_io_parse_args(int argc, char *const *argv, ssx_setup *s)
{
	if (!s || (argc > 0 && !argv))
		return fail(SSX_EINVAL, "ssx_io_parse_args: null argument");
	for (int i = 1; i < argc; ++i) {
		const char *a = argv[i];
		const int more = i + 1 < argc;
		if ((!strcmp(a, "-n") || !strcmp(a, "--n")) && more)
			s->n = atoi(argv[++i]);
		else if ((!strcmp(a, "-d") || !strcmp(a, "--dens")) && more)
			s->dens = atof(argv[++i]);
		else if ((!strcmp(a, "-v") || !strcmp(a, "--visc")) && more)
			s->visc = atof(argv[++i]);
		else if ((!strcmp(a, "-dt") || !strcmp(a, "--dt")) && more)
			s->dt = atof(argv[++i]);
		else if ((!strcmp(a, "-t") || !strcmp(a, "--tmax")) && more)
			s->tmax = atof(argv[++i]);
		else if (!strcmp(a, "-rt") || !strcmp(a, "--run-test"))
			s->test = 1;
	}
	return SSX_OK;
}
