/*
 * ssx_batch.c -- independent CG problems solved together: the Jacobi-PCG
 * of conjugate_gradient (solver.c:33-81) on level 0 of several contexts
 * (any element counts and orders; a context may appear more than once,
 * e.g. the three velocity components), with one launch per CG stage for
 * the whole batch and the iterations replayed from a captured HIP graph.
 *
 * Set-up (r = f - A u, z = D^-1 r, p = z, <r,z>_w) runs per problem with
 * the single-problem kernels.  Then every iteration is: the operator (one
 * launch per distinct order, with the previous direction update fused in
 * from the second iteration), the DSS, the alpha reduction, the update and
 * the beta / stopping reduction -- each one launch over all problems.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <hip/hip_runtime_api.h>

#include "ssx_internal.h"
#include "ssx_batch.h"
#include "ssx_batch_host.h"

#define HIPB(x)                                                              \
	do {                                                                 \
		hipError_t e_ = (x);                                         \
		if (e_ != hipSuccess) {                                      \
			ssx_set_error("%s: %s", #x, hipGetErrorString(e_));   \
			return SSX_EHIP;                                     \
		}                                                            \
	} while (0)
#define TRYB(x)                                                              \
	do {                                                                 \
		int rc_ = (x);                                               \
		if (rc_ != SSX_OK)                                           \
			return rc_;                                          \
	} while (0)

int ssx_balloc(void **p, size_t bytes)
{
	if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) {
		*p = NULL;
		ssx_set_error("hipMalloc of %zu bytes failed (batch)", bytes);
		return SSX_ENOMEM;
	}
	return SSX_OK;
}

static const double *jdiag0(const ssx_ctx *c, int var)
{
	int ch = c->diag_choice;
	if (ch < 0)
		ch = var == SSX_PRESSURE ? 0 : 1;      /* precon.c:211-225 */
	return ch == 0 ? c->lev[0].diag_p : c->lev[0].diag_h;
}

void ssx_batch_destroy(ssx_batch *b)
{
	if (!b)
		return;
	if (b->have_graph)
		hipGraphExecDestroy(b->gexec);
	if (b->have_event)
		hipEventDestroy(b->ready);
	ssx_bfcg_free(b);
	for (int i = 0; i < b->np; ++i) {
		if (b->r) hipFree(b->r[i]);
		if (b->z) hipFree(b->z[i]);
		if (b->p) hipFree(b->p[i]);
		if (b->ap) hipFree(b->ap[i]);
	}
	for (int k = 0; k < b->norder; ++k)
		hipFree(b->d_axtab[k]);
	hipFree(b->d_dsstab);
	hipFree(b->d_updtab);
	hipFree(b->part);
	hipFree(b->scal);
	hipFree(b->flags);
	hipFree(b->d_prob);
	hipFree(b->bpart);
	if (b->h_flags)
		hipHostFree(b->h_flags);
	if (b->h_scal)
		hipHostFree(b->h_scal);
	if (b->own_stream)
		hipStreamDestroy((hipStream_t) b->own_stream);
	free(b->r); free(b->z); free(b->p); free(b->ap);
	free(b->poff); free(b->h_prob); free(b->htab); free(b->active);
	free(b->ctx); free(b->op); free(b->var);
	free(b);
}

static int build(ssx_batch *b)
{
	const int np = b->np;
	b->r = calloc(np, sizeof(double *));
	b->z = calloc(np, sizeof(double *));
	b->p = calloc(np, sizeof(double *));
	b->ap = calloc(np, sizeof(double *));
	b->poff = calloc(np + 1, sizeof(size_t));
	b->h_prob = calloc(np, sizeof(ssx_bprob));
	if (!b->r || !b->z || !b->p || !b->ap || !b->poff || !b->h_prob) {
		ssx_set_error("out of host memory (batch)");
		return SSX_ENOMEM;
	}
	/* partial-sum layout and per-problem constants */
	for (int i = 0; i < np; ++i) {
		const ssx_ctx *c = b->ctx[i];
		const ssx_level *L = &c->lev[0];
		ssx_bprob *d = &b->h_prob[i];
		d->n = L->n;
		d->nel = c->nel;
		d->op = b->op[i];
		d->pressure = b->var[i] == SSX_PRESSURE;
		d->apply_mask = !d->pressure;
		d->ntot = (long long) L->ntot;
		d->lambda = c->lambda;
		d->nunique = (double) L->nunique;
		d->gf = L->gf;
		d->mass = L->mass;
		d->dmat = L->dmat;
		d->dinv = jdiag0(c, b->var[i]);
		d->scale = L->scale;
		d->mask = d->pressure ? NULL : L->mask;
		d->bc = L->bc;
		d->g2 = L->grp[0]; d->g4 = L->grp[1]; d->g8 = L->grp[2];
		d->gP = c->gather_upd ? L->gP : 0;
		d->gsoff = L->gsoff;
		d->gtab = L->gtab;
		d->G2 = L->ngrp[0]; d->G4 = L->ngrp[1]; d->G8 = L->ngrp[2];
		const int v = L->axv[SSX_PREC_FP64];
		const int teams = ssxk_bax_teams(L->n, v);
		d->nba = (c->nel + teams - 1) / teams;
		d->nbd = ssxk_dss_blocks(L->ngrp);
		d->nbu = (d->ntot + SSX_BATCH_CHUNK - 1) / SSX_BATCH_CHUNK;
		b->poff[i + 1] = b->poff[i] + (size_t) (d->nel + d->nbd
				+ 4 * d->nbu);
		TRYB(ssx_balloc((void **) &b->r[i], sizeof(double) * L->ntot));
		TRYB(ssx_balloc((void **) &b->z[i], sizeof(double) * L->ntot));
		TRYB(ssx_balloc((void **) &b->p[i], sizeof(double) * L->ntot));
		TRYB(ssx_balloc((void **) &b->ap[i], sizeof(double) * L->ntot));
		d->r = b->r[i]; d->z = b->z[i]; d->p = b->p[i]; d->ap = b->ap[i];
	}
	TRYB(ssx_balloc((void **) &b->part, sizeof(double) * b->poff[np]));
	TRYB(ssx_balloc((void **) &b->scal, sizeof(double) * CG_NSLOT * np));
	TRYB(ssx_balloc((void **) &b->flags, sizeof(int) * 2 * np));
	TRYB(ssx_balloc((void **) &b->d_prob, sizeof(ssx_bprob) * np));
	TRYB(ssx_balloc((void **) &b->bpart, sizeof(double) * SSX_NPART));
	HIPB(hipHostMalloc((void **) &b->h_flags, sizeof(int) * 2 * np, 0));
	HIPB(hipHostMalloc((void **) &b->h_scal, sizeof(double) * CG_NSLOT * np,
			0));
	for (int i = 0; i < np; ++i) {
		ssx_bprob *d = &b->h_prob[i];
		d->part = b->part + b->poff[i];
		d->scal = b->scal + CG_NSLOT * i;
		d->flags = b->flags + 2 * i;
	}
	/* (order, operator) groups and the launch shape the autotuner chose for
	 * the order; a problem's operator block count follows its group's
	 * shape */
	b->norder = 0;
	for (int i = 0; i < np; ++i) {
		const int n = b->h_prob[i].n, op = b->h_prob[i].op;
		int k;
		for (k = 0; k < b->norder; ++k)
			if (b->order_n[k] == n && b->order_op[k] == op)
				break;
		if (k == b->norder) {
			b->order_n[k] = n;
			b->order_op[k] = op;
			b->order_v[k] = b->ctx[i]->lev[0].axv[SSX_PREC_FP64];
			b->norder++;
		}
		const int teams = ssxk_bax_teams(n, b->order_v[k]);
		b->h_prob[i].nba = (b->h_prob[i].nel + teams - 1) / teams;
	}
	/* An order whose problems share a mesh (the velocity components of a
	 * time step) has its operator blocks interleaved so that the factors
	 * read for one problem are found in L2 by the others: its factor loads
	 * and result stores must not be non-temporal (launch-shape bit 1, which
	 * the single-problem autotuner may have chosen).  The other bits keep
	 * the block count, and every shape gives the same bits. */
	for (int k = 0; k < b->norder; ++k) {
		int shared = 0;
		for (int i = 0; i < np && !shared; ++i) {
			const ssx_bprob *d = &b->h_prob[i];
			if (d->n != b->order_n[k] || d->op != b->order_op[k])
				continue;
			for (int j = i + 1; j < np && !shared; ++j) {
				const ssx_bprob *e = &b->h_prob[j];
				shared = e->n == d->n && e->op == d->op && e->gf == d->gf;
			}
		}
		if (shared)
			b->order_v[k] &= ~2;
	}
	int rc = SSX_OK;
	for (int i = 0; i < np; ++i) {
		const ssx_bprob *d = &b->h_prob[i];
		b->poff[i + 1] = b->poff[i] + (size_t) (d->nel + d->nbd + 4 * d->nbu);
	}
	hipFree(b->part);
	rc = ssx_balloc((void **) &b->part, sizeof(double) * b->poff[np]);
	for (int i = 0; i < np && rc == SSX_OK; ++i)
		b->h_prob[i].part = b->part + b->poff[i];
	/* device tables at full size; host staging for rebuilds */
	long long tot = 0;
	for (int k = 0; k < b->norder && rc == SSX_OK; ++k) {
		long long m = 0;
		for (int i = 0; i < np; ++i)
			if (b->h_prob[i].n == b->order_n[k]
					&& b->h_prob[i].op == b->order_op[k])
				m += b->h_prob[i].nba;
		rc = ssx_balloc((void **) &b->d_axtab[k], sizeof(ssx_bwork) * m);
		tot += m;
	}
	long long md = 0, mu = 0;
	for (int i = 0; i < np; ++i) {
		md += b->h_prob[i].nbd;
		mu += b->h_prob[i].nbu;
	}
	if (rc == SSX_OK)
		rc = ssx_balloc((void **) &b->d_dsstab, sizeof(ssx_bwork) * md);
	if (rc == SSX_OK)
		rc = ssx_balloc((void **) &b->d_updtab, sizeof(ssx_bwork) * mu);
	tot += md + mu;
	b->htab = malloc(sizeof(ssx_bwork) * (size_t) (tot ? tot : 1));
	b->htab_n = tot;
	b->active = calloc(np, 1);
	if (rc == SSX_OK && (!b->htab || !b->active)) {
		ssx_set_error("out of host memory (batch)");
		rc = SSX_ENOMEM;
	}
	return rc;
}

/* the work tables of the problems with active[i] != 0, uploaded on the
 * batch stream (ordered after the kernels that read the old ones) */
static int fill_tables(ssx_batch *b, const unsigned char *active)
{
	hipStream_t st = (hipStream_t) b->stream;
	ssx_bwork *t = b->htab;
	long long off = 0;
	/* Problems on one mesh (same factors, same blocks: the velocity
	 * components of a time step) take their operator blocks interleaved
	 * in runs of 8: workgroup w runs on XCD w % 8, so block q of every such
	 * problem lands on one XCD, close in time, and the element's factors
	 * are read from HBM once for all of them. */
	unsigned char *grouped = calloc((size_t) b->np, 1);
	int *mem = malloc(sizeof(int) * (size_t) (b->np ? b->np : 1));
	if (!grouped || !mem) {
		free(grouped);
		free(mem);
		ssx_set_error("out of host memory (batch)");
		return SSX_ENOMEM;
	}
	for (int k = 0; k < b->norder; ++k) {
		long long m = 0;
		memset(grouped, 0, (size_t) b->np);
		for (int i = 0; i < b->np; ++i) {
			const ssx_bprob *d = &b->h_prob[i];
			if (!active[i] || grouped[i] || d->n != b->order_n[k]
					|| d->op != b->order_op[k])
				continue;
			int nm = 0;                 /* i's group: same n, op, mesh */
			for (int j = i; j < b->np; ++j) {
				const ssx_bprob *e = &b->h_prob[j];
				if (active[j] && !grouped[j] && e->n == d->n
						&& e->op == d->op && e->gf == d->gf
						&& e->nba == d->nba) {
					grouped[j] = 1;
					mem[nm++] = j;
				}
			}
			for (long long q0 = 0; q0 < d->nba; q0 += 8)
				for (int g = 0; g < nm; ++g)
					for (long long q = q0; q < q0 + 8 && q < d->nba; ++q) {
						t[off + m].prob = mem[g];
						t[off + m].blk = (int) q;
						++m;
					}
		}
		b->nax[k] = m;
		if (m && hipMemcpyAsync(b->d_axtab[k], t + off, sizeof(ssx_bwork) * m,
				hipMemcpyHostToDevice, st) != hipSuccess) {
			free(grouped);
			free(mem);
			ssx_set_error("hipMemcpyAsync of a batch work table failed");
			return SSX_EHIP;
		}
		off += m;
	}
	free(grouped);
	free(mem);
	for (int pass = 0; pass < 2; ++pass) {
		long long m = 0;
		for (int i = 0; i < b->np; ++i) {
			if (!active[i])
				continue;
			/* problems with the gather form have no DSS launch */
			const long long nb = pass == 0
				? (b->h_prob[i].gP > 0 ? 0 : b->h_prob[i].nbd)
				: b->h_prob[i].nbu;
			for (long long q = 0; q < nb; ++q) {
				t[off + m].prob = i;
				t[off + m].blk = (int) q;
				++m;
			}
		}
		if (m)
			HIPB(hipMemcpyAsync(pass == 0 ? b->d_dsstab : b->d_updtab,
					t + off, sizeof(ssx_bwork) * m, hipMemcpyHostToDevice,
					st));
		if (pass == 0)
			b->ndss = m;
		else
			b->nupd = m;
		off += m;
	}
	/* the staging buffer is reused by the next rebuild: let the copies
	 * finish first */
	HIPB(hipStreamSynchronize(st));
	memcpy(b->active, active, b->np);
	if (b->have_graph) {
		hipGraphExecDestroy(b->gexec);
		b->have_graph = 0;
	}
	return SSX_OK;
}


int ssx_batch_create(ssx_batch **out, ssx_ctx *const *ctx, const int *op,
		const int *var, int count)
{
	if (!out)
		return SSX_EINVAL;
	*out = NULL;
	if (!ctx || !op || !var || count < 1) {
		ssx_set_error("ssx_batch_create: empty batch");
		return SSX_EINVAL;
	}
	for (int i = 0; i < count; ++i) {
		if (!ctx[i] || (op[i] != SSX_OP_POISSON && op[i] != SSX_OP_HELMHOLTZ)
				|| var[i] < SSX_PRESSURE || var[i] > SSX_Z_VELOCITY) {
			ssx_set_error("ssx_batch_create: bad problem %d", i);
			return SSX_EINVAL;
		}
		if (!ctx[i]->lev[0].fused_ok) {
			ssx_set_error("ssx_batch_create: problem %d: mesh needs the "
					"plain CG path", i);
			return SSX_EMESH;
		}
	}
	ssx_batch *b = calloc(1, sizeof(*b));
	if (!b)
		return SSX_ENOMEM;
	b->np = count;
	b->ctx = malloc(sizeof(ssx_ctx *) * count);
	b->op = malloc(sizeof(int) * count);
	b->var = malloc(sizeof(int) * count);
	if (!b->ctx || !b->op || !b->var) {
		ssx_batch_destroy(b);
		return SSX_ENOMEM;
	}
	memcpy(b->ctx, ctx, sizeof(ssx_ctx *) * count);
	memcpy(b->op, op, sizeof(int) * count);
	memcpy(b->var, var, sizeof(int) * count);
	hipStream_t s;
	if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
		ssx_set_error("hipStreamCreate (batch)");
		ssx_batch_destroy(b);
		return SSX_EHIP;
	}
	b->own_stream = b->stream = s;
	if (hipEventCreateWithFlags(&b->ready, hipEventDisableTiming)
			!= hipSuccess) {
		ssx_set_error("hipEventCreate (batch)");
		ssx_batch_destroy(b);
		return SSX_EHIP;
	}
	b->have_event = 1;
	int rc = build(b);
	if (rc == SSX_OK)
		rc = ssx_bfcg_init(b);
	if (rc != SSX_OK) {
		ssx_batch_destroy(b);
		return rc;
	}
	*out = b;
	return SSX_OK;
}

int ssx_batch_set_stream(ssx_batch *b, void *s)
{
	if (!b)
		return SSX_EINVAL;
	b->stream = s;
	if (b->have_graph) {             /* recaptured on the new stream */
		hipGraphExecDestroy(b->gexec);
		b->have_graph = 0;
	}
	return SSX_OK;
}

/* one iteration: operator per order, DSS (problems without the gather
 * form only), alpha, update, beta + stop */
static int iteration(ssx_batch *b, int pupd)
{
	void *st = b->stream;
	for (int k = 0; k < b->norder; ++k)
		TRYB(ssxk_bax(st, b->order_n[k], b->order_op[k], b->order_v[k], pupd,
				b->d_prob, b->d_axtab[k], b->nax[k]));
	TRYB(ssxk_bdss(st, b->d_prob, b->d_dsstab, b->ndss));
	TRYB(ssxk_breduce(st, 0, b->d_prob, b->np));
	TRYB(ssxk_bupdate(st, b->d_prob, b->d_updtab, b->nupd));
	return ssxk_breduce(st, 1, b->d_prob, b->np);
}

static int capture(ssx_batch *b)
{
	hipStream_t st = (hipStream_t) b->stream;
	hipGraph_t g;
	HIPB(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
	int rc = SSX_OK;
	for (int k = 0; k < GRAPH_ITERS && rc == SSX_OK; ++k)
		rc = iteration(b, 1);
	hipError_t e = hipStreamEndCapture(st, &g);
	if (rc != SSX_OK)
		return rc;
	if (e != hipSuccess) {
		ssx_set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
		return SSX_EHIP;
	}
	e = hipGraphInstantiate(&b->gexec, g, NULL, NULL, 0);
	hipGraphDestroy(g);
	if (e != hipSuccess) {
		ssx_set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
		return SSX_EHIP;
	}
	b->have_graph = 1;
	return SSX_OK;
}

/* per-problem set-up: mask f, r = f - A u, null space, z = D^-1 r,
 * p = z, <r,z>_w (solver.c:46-56) */
int ssx_batch_setup_problem(ssx_batch *b, int i, double *f, double *u)
{
	const ssx_ctx *c = b->ctx[i];
	const ssx_level *L = &c->lev[0];
	const ssx_bprob *d = &b->h_prob[i];
	void *st = b->stream;
	const size_t nt = L->ntot;
	double *sc = b->scal + CG_NSLOT * i;
	if (!d->pressure)
		TRYB(ssxk_mul(st, nt, f, L->mask, f));
	TRYB(ssxk_ax(st, L->n, L->axv[SSX_PREC_FP64], c->nel, d->op, u, d->r,
			L->gf, L->mass, L->dmat, c->lambda, L->bc, !d->pressure, NULL,
			NULL));
	TRYB(ssxk_dss_all(st, L->grp, L->ngrp, d->r, NULL, NULL, 0, NULL));
	TRYB(ssxk_axpby(st, nt, 1.0, f, NULL, -1.0, d->r));
	for (int pass = 0; pass < 2; ++pass) {
		double *x = pass == 0 ? d->r : d->z;
		if (pass == 1)
			TRYB(ssxk_mul(st, nt, d->r, d->dinv, d->z));
		if (d->pressure) {      /* solver.c:200-213 */
			TRYB(ssxk_dot(st, nt, x, L->scale, NULL, b->bpart, 0));
			TRYB(ssxk_reduce(st, b->bpart, 0, sc, CG_NULLC));
			TRYB(ssxk_shift_s(st, nt, sc, CG_NULLC, (double) L->nunique,
					x));
		}
	}
	TRYB(ssxk_copy(st, d->p, d->z, nt));
	TRYB(ssxk_dot(st, nt, d->r, L->scale, d->z, b->bpart, 0));
	return ssxk_reduce(st, b->bpart, 0, sc, CG_RDZ);
}

/* Settings the batch takes from its contexts (Jacobi diagonal choice,
 * gather-form DSS, operator launch shape) are re-read at every solve, so a
 * batch follows ssx_set_diag_choice, SSX_OPT_GATHER_DSS and
 * ssx_set_ax_variant made after ssx_batch_create.  A changed launch shape
 * or DSS form rebuilds the work tables (and the captured graph). */
int ssx_batch_sync_settings(ssx_batch *b)
{
	int rebuild = 0;
	for (int i = 0; i < b->np; ++i) {
		const ssx_ctx *c = b->ctx[i];
		ssx_bprob *d = &b->h_prob[i];
		d->dinv = jdiag0(c, b->var[i]);
		const int gP = c->gather_upd ? c->lev[0].gP : 0;
		if (gP != d->gP) {
			d->gP = gP;
			rebuild = 1;
		}
	}
	for (int k = 0; k < b->norder; ++k) {
		int first = -1;
		for (int i = 0; i < b->np && first < 0; ++i)
			if (b->h_prob[i].n == b->order_n[k]
					&& b->h_prob[i].op == b->order_op[k])
				first = i;
		const int v = b->ctx[first]->lev[0].axv[SSX_PREC_FP64];
		if (v == b->order_v[k])
			continue;
		const int t_old = ssxk_bax_teams(b->order_n[k], b->order_v[k]);
		const int t_new = ssxk_bax_teams(b->order_n[k], v);
		b->order_v[k] = v;
		rebuild = 1;
		for (int i = 0; i < b->np; ++i)
			if (b->h_prob[i].n == b->order_n[k]
					&& b->h_prob[i].op == b->order_op[k])
				b->h_prob[i].nba = (b->h_prob[i].nel + t_new - 1) / t_new;
		if (t_new < t_old) {    /* more operator workgroups: larger table */
			long long m = 0;
			for (int i = 0; i < b->np; ++i)
				if (b->h_prob[i].n == b->order_n[k]
						&& b->h_prob[i].op == b->order_op[k])
					m += b->h_prob[i].nba;
			HIPB(hipStreamSynchronize((hipStream_t) b->stream));
			hipFree(b->d_axtab[k]);
			TRYB(ssx_balloc((void **) &b->d_axtab[k], sizeof(ssx_bwork) * m));
			long long tot = m;
			for (int q = 0; q < b->norder; ++q)
				if (q != k)
					for (int i = 0; i < b->np; ++i)
						if (b->h_prob[i].n == b->order_n[q]
								&& b->h_prob[i].op == b->order_op[q])
							tot += b->h_prob[i].nba;
			for (int i = 0; i < b->np; ++i)
				tot += b->h_prob[i].nbd + b->h_prob[i].nbu;
			if (tot > b->htab_n) {
				ssx_bwork *t = realloc(b->htab, sizeof(ssx_bwork) * tot);
				if (!t) {
					ssx_set_error("out of host memory (batch)");
					return SSX_ENOMEM;
				}
				b->htab = t;
				b->htab_n = tot;
			}
		}
	}
	if (rebuild)                    /* the next solve refills the tables */
		memset(b->active, 0, b->np);
	return SSX_OK;
}

/* The batch stream waits for the work already enqueued on every context's
 * stream (which may still be writing f and u) before it reads them. */
int ssx_batch_order_after_contexts(ssx_batch *b)
{
	hipStream_t st = (hipStream_t) b->stream;
	for (int i = 0; i < b->np; ++i) {
		hipStream_t cs = (hipStream_t) b->ctx[i]->stream;
		int seen = cs == st;
		for (int q = 0; q < i && !seen; ++q)
			seen = (hipStream_t) b->ctx[q]->stream == cs;
		if (seen)
			continue;
		HIPB(hipEventRecord(b->ready, cs));
		HIPB(hipStreamWaitEvent(st, b->ready, 0));
	}
	return SSX_OK;
}

int ssx_batch_cg(ssx_batch *b, double *const *f, double *const *u,
		const double *eps, int max_iter, int use_graph, ssx_stats *stats)
{
	if (!b || !f || !u || !eps || max_iter < 1) {
		ssx_set_error("ssx_batch_cg: bad arguments");
		return SSX_EINVAL;
	}
	if (use_graph && !b->stream) {
		ssx_set_error("ssx_batch_cg: graph replay needs a non-null stream");
		return SSX_EINVAL;
	}
	const int np = b->np;
	hipStream_t st = (hipStream_t) b->stream;
	struct timespec t0, t1;
	clock_gettime(CLOCK_MONOTONIC, &t0);
	for (int i = 0; i < np; ++i) {
		if (!f[i] || !u[i]) {
			ssx_set_error("ssx_batch_cg: null vector for problem %d", i);
			return SSX_EINVAL;
		}
		b->h_prob[i].u = u[i];
		b->h_prob[i].eps = eps[i];
		b->h_prob[i].maxit = max_iter;
	}
	TRYB(ssx_batch_sync_settings(b));
	TRYB(ssx_batch_order_after_contexts(b));
	HIPB(hipMemcpyAsync(b->d_prob, b->h_prob, sizeof(ssx_bprob) * np,
			hipMemcpyHostToDevice, st));
	for (int i = 0; i < np; ++i)
		TRYB(ssx_batch_setup_problem(b, i, f[i], u[i]));
	TRYB(ssxk_set_ints(st, b->flags, 0, 2 * np < 64 ? 2 * np : 64));
	for (int q = 64; q < 2 * np; q += 64)
		TRYB(ssxk_set_ints(st, b->flags + q, 0,
				2 * np - q < 64 ? 2 * np - q : 64));
	unsigned char *act = malloc(np);
	if (!act)
		return SSX_ENOMEM;
	memset(act, 1, np);
	int rc0 = SSX_OK;
	if (memcmp(act, b->active, np) != 0)
		rc0 = fill_tables(b, act);
	/* first iteration: p = z already */
	if (rc0 == SSX_OK)
		rc0 = iteration(b, 0);
	int launched = 1;
	/* replays between host checks: doubled (up to 16) while no problem
	 * finishes, back to 1 when one does; a finished problem's kernels are
	 * no-ops until the next check drops it from the tables */
	int reps = 1;
	for (; rc0 == SSX_OK;) {
		int all = 1, changed = 0;
		if ((rc0 = hipMemcpyAsync(b->h_flags, b->flags, sizeof(int) * 2 * np,
				hipMemcpyDeviceToHost, st)) != hipSuccess
				|| (rc0 = hipStreamSynchronize(st)) != hipSuccess) {
			ssx_set_error("batch flags: %s",
					hipGetErrorString((hipError_t) rc0));
			rc0 = SSX_EHIP;
			break;
		}
		for (int i = 0; i < np; ++i) {
			act[i] = b->h_flags[2 * i] == 0;
			all &= !act[i];
			changed |= act[i] != b->active[i];
		}
		if (all || launched >= max_iter)
			break;
		/* drop converged problems from the launches */
		if (changed && (rc0 = fill_tables(b, act)) != SSX_OK)
			break;
		reps = changed ? 1 : (reps < 16 ? 2 * reps : 16);
		if ((max_iter - launched + GRAPH_ITERS - 1) / GRAPH_ITERS < reps)
			reps = (max_iter - launched + GRAPH_ITERS - 1) / GRAPH_ITERS;
		if (use_graph && !b->have_graph && (rc0 = capture(b)) != SSX_OK)
			break;
		for (int r = 0; r < reps && rc0 == SSX_OK; ++r) {
			if (use_graph) {
				if (hipGraphLaunch(b->gexec, st) != hipSuccess) {
					ssx_set_error("hipGraphLaunch failed (batch)");
					rc0 = SSX_EHIP;
				}
			} else {
				for (int k = 0; k < GRAPH_ITERS && rc0 == SSX_OK; ++k)
					rc0 = iteration(b, 1);
			}
			launched += GRAPH_ITERS;
		}
	}
	free(act);
	if (rc0 != SSX_OK)
		return rc0;
	/* the last iteration's solution step, which the update leaves to the
	 * next operator launch (u += alpha p, the arithmetic of the update) */
	for (int i = 0; i < np; ++i)
		TRYB(ssxk_cg_uflush(st, (size_t) b->h_prob[i].ntot,
				b->scal + CG_NSLOT * i, CG_ALPHA, b->h_prob[i].p,
				b->h_prob[i].u));
	HIPB(hipMemcpyAsync(b->h_scal, b->scal, sizeof(double) * CG_NSLOT * np,
			hipMemcpyDeviceToHost, st));
	HIPB(hipStreamSynchronize(st));
	clock_gettime(CLOCK_MONOTONIC, &t1);
	const double secs = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec
			- t0.tv_nsec);
	int rc = SSX_OK;
	for (int i = 0; i < np; ++i) {
		const double err = sqrt(b->h_scal[CG_NSLOT * i + CG_RR]);
		if (stats) {
			memset(&stats[i], 0, sizeof(ssx_stats));
			stats[i].iterations = b->h_flags[2 * i + 1];
			stats[i].residual = err;
			stats[i].converged = err <= eps[i];
			stats[i].fine_ax = b->h_flags[2 * i + 1] + 1;
			stats[i].seconds = secs;
		}
		if (isnan(err) || isinf(err)) {
			ssx_set_error("batch problem %d: residual is %g", i, err);
			rc = SSX_ENUMERIC;
		} else if (rc == SSX_OK && !(err <= eps[i])) {
			rc = SSX_ENOTCONV;
		}
	}
	return rc;
}
