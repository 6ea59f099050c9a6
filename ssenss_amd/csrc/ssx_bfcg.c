/*
 * ssx_bfcg.c -- a batch of independent problems solved together, each with
 * the reference's solver for it: flexible_conjugate_gradient
 * (solver.c:83-157) preconditioned by the p-multigrid V-cycle
 * (precon.c:227-271) for SSX_PC_MG problems (the pressure solve,
 * navier_stokes.c:157), conjugate_gradient (solver.c:33-81) with the
 * Jacobi preconditioner for SSX_PC_JACOBI ones (the velocity solves,
 * navier_stokes.c:219).
 *
 * The multigrid problems advance in lock-step, one FCG iteration at a
 * time.  In each V-cycle the level CGs of all problems -- the smoother of
 * every level (15 iterations, early exit at eps) and the coarsest-level
 * solve (500, to eps) -- run as batched passes: one launch per CG stage for
 * all problems at once (the operator: one launch per (order, operator,
 * precision) group).  The Jacobi problems ride along in every pass until
 * they converge.  The stages between the passes (residual, restriction,
 * prolongation, the FCG orthogonalisation and step) are the single-problem
 * kernels, enqueued per problem on the batch stream, with the problem's
 * smoother precision (SSX_OPT_SMOOTHER_PRECISION of its context).
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

#define SMOOTH_ITERS 15         /* precon.c:265 */
#define COARSE_ITERS 500        /* precon.c:243 */

/* precision of a problem's smoother: a single level is the coarsest,
 * which is FP64 in every mode */
static int mg_prec(const ssx_ctx *c)
{
	return c->nlev > 1 ? c->smooth_prec : SSX_PREC_FP64;
}

static int task_of(const ssx_batch *b, int i, int l)
{
	return b->np + b->toff[i] + l;
}

static long long nbu_of(size_t ntot)
{
	return (long long) ((ntot + SSX_BATCH_CHUNK - 1) / SSX_BATCH_CHUNK);
}

int ssx_bfcg_init(ssx_batch *b)
{
	const int np = b->np;
	b->toff = malloc(sizeof(int) * (np + 1));
	if (!b->toff)
		return SSX_ENOMEM;
	int nt = 0;
	size_t npart = 0;
	for (int i = 0; i < np; ++i) {
		const ssx_ctx *c = b->ctx[i];
		b->toff[i] = nt;
		nt += c->nlev;
		for (int l = 0; l < c->nlev; ++l) {
			const ssx_level *L = &c->lev[l];
			npart += (size_t) c->nel + ssxk_dss_blocks(L->ngrp)
				+ 4 * nbu_of(L->ntot);
		}
	}
	b->toff[np] = nt;
	b->ntask = np + nt;
	ssx_bprob *h = realloc(b->h_prob, sizeof(ssx_bprob) * b->ntask);
	if (!h) {
		ssx_set_error("out of host memory (batch)");
		return SSX_ENOMEM;
	}
	b->h_prob = h;
	memset(h + np, 0, sizeof(ssx_bprob) * nt);
	hipFree(b->d_prob);
	b->d_prob = NULL;
	TRYB(ssx_balloc((void **) &b->d_prob, sizeof(ssx_bprob) * b->ntask));
	TRYB(ssx_balloc((void **) &b->tpart, sizeof(double) * (npart ? npart
			: 1)));
	TRYB(ssx_balloc((void **) &b->tscal, sizeof(double) * CG_NSLOT * nt));
	TRYB(ssx_balloc((void **) &b->tflags, sizeof(int) * 2 * nt));
	HIPB(hipHostMalloc((void **) &b->h_tflags, sizeof(int) * 2 * nt, 0));
	size_t off = 0;
	for (int i = 0; i < np; ++i) {
		const ssx_ctx *c = b->ctx[i];
		for (int l = 0; l < c->nlev; ++l) {
			const ssx_level *L = &c->lev[l];
			const int k = b->toff[i] + l;
			ssx_bprob *d = &h[np + k];
			d->nbd = ssxk_dss_blocks(L->ngrp);
			d->nbu = nbu_of(L->ntot);
			d->part = b->tpart + off;
			d->scal = b->tscal + CG_NSLOT * k;
			d->flags = b->tflags + 2 * k;
			off += (size_t) c->nel + d->nbd + 4 * d->nbu;
		}
	}
	return SSX_OK;
}

void ssx_bfcg_free(ssx_batch *b)
{
	for (int k = 0; k < SSX_NPASS; ++k) {
		hipFree(b->pass[k].d_tab);
		hipFree(b->pass[k].d_list);
		free(b->pass[k].h_list);
	}
	hipFree(b->tpart);
	hipFree(b->tscal);
	hipFree(b->tflags);
	if (b->h_tflags)
		hipHostFree(b->h_tflags);
	free(b->hwork);
	free(b->toff);
}

/* the level-l task of problem i for this solve: the level's arrays of the
 * problem's context, its FP64 vectors r, z, p, ap (and FP32 copies on a
 * reduced-precision smoother level); the solution is fcg_z on level 0 and
 * the level's error mg_e below it */
static void fill_task(ssx_batch *b, int i, int l, int prec, double eps)
{
	ssx_ctx *c = b->ctx[i];
	ssx_level *L = &c->lev[l];
	ssx_bprob *d = &b->h_prob[task_of(b, i, l)];
	const int coarse = l == c->nlev - 1;
	const int pres = b->var[i] == SSX_PRESSURE;
	d->n = L->n;
	d->nel = c->nel;
	d->op = b->op[i];
	d->pressure = pres;
	d->apply_mask = !pres;
	d->maxit = coarse ? COARSE_ITERS : SMOOTH_ITERS;
	d->ntot = (long long) L->ntot;
	d->lambda = c->lambda;
	d->nunique = (double) L->nunique;
	/* a coarsest solve under a reduced-precision smoother stops at the
	 * tolerance its start leaves in scal[CG_TOL] */
	d->eps = (coarse && prec != SSX_PREC_FP64) ? -1.0 : eps;
	d->gf = L->gf;
	d->mass = L->mass;
	d->dmat = L->dmat;
	d->dinv = ssx_jdiag(c, l, b->var[i]);
	d->scale = L->scale;
	d->mask = pres ? NULL : L->mask;
	d->bc = L->bc;
	d->g2 = L->grp[0];
	d->g4 = L->grp[1];
	d->g8 = L->grp[2];
	d->G2 = L->ngrp[0];
	d->G4 = L->ngrp[1];
	d->G8 = L->ngrp[2];
	d->gP = c->gather_upd ? L->gP : 0;
	d->gsoff = L->gsoff;
	d->gtab = L->gtab;
	d->u = l == 0 ? c->fcg_z : L->mg_e;
	d->r = L->r;
	d->z = L->z;
	d->p = L->p;
	d->ap = L->ap;
	d->prec = (!coarse && prec != SSX_PREC_FP64) ? prec : SSX_PREC_FP64;
	if (d->prec != SSX_PREC_FP64) {
		d->gf_lp = L->gf_lp;
		d->mass_f = L->mass_lp;
		d->dmat_f = L->dmat_lp;
		d->dinv_f = ssx_jdiag_lp(c, l, b->var[i]);
		d->scale_f = L->scale_lp;
		d->u_f = L->e_lp;
		d->r_f = L->r_lp;
		d->z_f = L->z_lp;
		d->p_f = L->p_lp;
		d->ap_f = L->ap_lp;
	}
}

/* context and level of task t (riders: level 0 of problem t) */
static void task_where(const ssx_batch *b, int t, const ssx_ctx **c,
		int *l)
{
	if (t < b->np) {
		*c = b->ctx[t];
		*l = 0;
		return;
	}
	int i = 0;
	while (i + 1 < b->np && b->toff[i + 1] <= t - b->np)
		++i;
	*c = b->ctx[i];
	*l = t - b->np - b->toff[i];
}

/* the pass's tables for tasks own[0..nown) (started by the pass) and the
 * riding Jacobi problems rid[0..nrid), uploaded on the batch stream */
static int build_pass(ssx_batch *b, ssx_bpass *ps, const int *own, int nown,
		const int *rid, int nrid)
{
	hipStream_t st = (hipStream_t) b->stream;
	const int nl = nown + nrid;
	if (nl + nown > ps->list_cap) {
		hipFree(ps->d_list);
		ps->d_list = NULL;
		free(ps->h_list);
		ps->list_cap = nl + nown;
		ps->h_list = malloc(sizeof(int) * ps->list_cap);
		if (!ps->h_list) {
			ps->list_cap = 0;
			ssx_set_error("out of host memory (batch)");
			return SSX_ENOMEM;
		}
		TRYB(ssx_balloc((void **) &ps->d_list, sizeof(int) * ps->list_cap));
	}
	int *lst = ps->h_list;
	for (int q = 0; q < nown; ++q)
		lst[q] = lst[nl + q] = own[q];
	for (int q = 0; q < nrid; ++q)
		lst[nown + q] = rid[q];
	ps->nlist = nl;
	ps->nown = nown;
	/* groups and table sizes */
	ps->ngrp = 0;
	long long nax[SSX_BGROUPS] = { 0 };
	long long ndss = 0, ndsslp = 0, nupd = 0, nupdlp = 0, nfl = 0;
	for (int q = 0; q < nl; ++q) {
		const int t = lst[q];
		const ssx_bprob *d = &b->h_prob[t];
		const ssx_ctx *c;
		int l;
		task_where(b, t, &c, &l);
		int g;
		for (g = 0; g < ps->ngrp; ++g)
			if (ps->g_n[g] == d->n && ps->g_op[g] == d->op
					&& ps->g_prec[g] == d->prec)
				break;
		if (g == ps->ngrp) {
			if (g == SSX_BGROUPS) {
				ssx_set_error("batch: too many operator groups");
				return SSX_EINVAL;
			}
			ps->g_n[g] = d->n;
			ps->g_op[g] = d->op;
			ps->g_prec[g] = d->prec;
			ps->g_v[g] = c->lev[l].axv[d->prec];
			ps->ngrp++;
		}
		const int teams = ssxk_bax_teams(d->n, ps->g_v[g]);
		nax[g] += (d->nel + teams - 1) / teams;
		if (d->gP == 0) {
			if (d->prec == SSX_PREC_FP64)
				ndss += d->nbd;
			else
				ndsslp += d->nbd;
		}
		if (d->prec == SSX_PREC_FP64) {
			nupd += d->nbu;
			if (q < nown)
				nfl += d->nbu;
		} else {
			nupdlp += d->nbu;
		}
	}
	long long tot = 0;
	for (int g = 0; g < ps->ngrp; ++g) {
		ps->g_off[g] = tot;
		ps->g_cnt[g] = nax[g];
		tot += nax[g];
	}
	ps->dss_off = tot; ps->ndss = ndss; tot += ndss;
	ps->dsslp_off = tot; ps->ndsslp = ndsslp; tot += ndsslp;
	ps->upd_off = tot; ps->nupd = nupd; tot += nupd;
	ps->updlp_off = tot; ps->nupdlp = nupdlp; tot += nupdlp;
	ps->fl_off = tot; ps->nfl = nfl; tot += nfl;
	if (tot > b->hwork_cap) {
		ssx_bwork *w = realloc(b->hwork, sizeof(ssx_bwork) * tot);
		if (!w) {
			ssx_set_error("out of host memory (batch)");
			return SSX_ENOMEM;
		}
		b->hwork = w;
		b->hwork_cap = tot;
	}
	if (tot > ps->tab_cap) {
		HIPB(hipStreamSynchronize(st));
		hipFree(ps->d_tab);
		ps->d_tab = NULL;
		ps->tab_cap = 0;
		TRYB(ssx_balloc((void **) &ps->d_tab, sizeof(ssx_bwork) * tot));
		ps->tab_cap = tot;
	}
	/* fill: operator tables group by group, then DSS, update, flush */
	ssx_bwork *w = b->hwork;
	long long pos[SSX_BGROUPS];
	for (int g = 0; g < ps->ngrp; ++g)
		pos[g] = ps->g_off[g];
	long long pd = ps->dss_off, pdl = ps->dsslp_off, pu = ps->upd_off;
	long long pul = ps->updlp_off, pf = ps->fl_off;
	for (int q = 0; q < nl; ++q) {
		const int t = lst[q];
		const ssx_bprob *d = &b->h_prob[t];
		int g;
		for (g = 0; g < ps->ngrp; ++g)
			if (ps->g_n[g] == d->n && ps->g_op[g] == d->op
					&& ps->g_prec[g] == d->prec)
				break;
		const int teams = ssxk_bax_teams(d->n, ps->g_v[g]);
		const long long nba = (d->nel + teams - 1) / teams;
		for (long long k = 0; k < nba; ++k)
			w[pos[g]++] = (ssx_bwork) { t, (int) k };
		if (d->gP == 0)
			for (long long k = 0; k < d->nbd; ++k)
				w[d->prec == SSX_PREC_FP64 ? pd++ : pdl++] =
					(ssx_bwork) { t, (int) k };
		for (long long k = 0; k < d->nbu; ++k) {
			w[d->prec == SSX_PREC_FP64 ? pu++ : pul++] =
				(ssx_bwork) { t, (int) k };
			if (d->prec == SSX_PREC_FP64 && q < nown)
				w[pf++] = (ssx_bwork) { t, (int) k };
		}
	}
	if (tot)
		HIPB(hipMemcpyAsync(ps->d_tab, w, sizeof(ssx_bwork) * tot,
				hipMemcpyHostToDevice, st));
	if (nl + nown)
		HIPB(hipMemcpyAsync(ps->d_list, lst, sizeof(int) * (nl + nown),
				hipMemcpyHostToDevice, st));
	/* the staging buffers are reused by the next build */
	HIPB(hipStreamSynchronize(st));
	ps->valid = 1;
	return SSX_OK;
}

/* one CG iteration of every task of the pass */
static int pass_iter(ssx_batch *b, const ssx_bpass *ps)
{
	void *st = b->stream;
	for (int g = 0; g < ps->ngrp; ++g) {
		const ssx_bwork *t = ps->d_tab + ps->g_off[g];
		if (ps->g_prec[g] == SSX_PREC_FP64)
			TRYB(ssxk_bax(st, ps->g_n[g], ps->g_op[g], ps->g_v[g], 1,
					b->d_prob, t, ps->g_cnt[g]));
		else
			TRYB(ssxk_bax_lp(st, ps->g_prec[g], ps->g_n[g], ps->g_op[g],
					ps->g_v[g], b->d_prob, t, ps->g_cnt[g]));
	}
	TRYB(ssxk_bdss(st, b->d_prob, ps->d_tab + ps->dss_off, ps->ndss));
	TRYB(ssxk_bdss_lp(st, b->d_prob, ps->d_tab + ps->dsslp_off,
			ps->ndsslp));
	TRYB(ssxk_breduce_list(st, 0, b->d_prob, ps->d_list, ps->nlist));
	TRYB(ssxk_bupdate(st, b->d_prob, ps->d_tab + ps->upd_off, ps->nupd));
	TRYB(ssxk_bupdate_lp(st, b->d_prob, ps->d_tab + ps->updlp_off,
			ps->nupdlp));
	return ssxk_breduce_list(st, 1, b->d_prob, ps->d_list, ps->nlist);
}

/* read the flags of every task back (multigrid tasks and riders) */
static int read_flags(ssx_batch *b)
{
	hipStream_t st = (hipStream_t) b->stream;
	HIPB(hipMemcpyAsync(b->h_flags, b->flags, sizeof(int) * 2 * b->np,
			hipMemcpyDeviceToHost, st));
	if (b->ntask > b->np)
		HIPB(hipMemcpyAsync(b->h_tflags, b->tflags, sizeof(int) * 2
				* (b->ntask - b->np), hipMemcpyDeviceToHost, st));
	HIPB(hipStreamSynchronize(st));
	return SSX_OK;
}

static const int *task_flags(const ssx_batch *b, int t)
{
	return t < b->np ? b->h_flags + 2 * t : b->h_tflags + 2 * (t - b->np);
}

/* A pass over its tasks (started, i.e. set up, by the caller): flags and
 * first-iteration scalars reset, `iters` iterations (checked: in chunks,
 * stopping when every own task, or with none every rider, is done), and
 * the own FP64 tasks' last solution step. */
static int run_pass(ssx_batch *b, const ssx_bpass *ps, int iters,
		int checked, int hint)
{
	void *st = b->stream;
	TRYB(ssxk_btask_reset(st, b->d_prob, ps->d_list + ps->nlist, ps->nown));
	if (!checked) {
		for (int k = 0; k < iters; ++k)
			TRYB(pass_iter(b, ps));
	} else {
		int launched = 0, chunk = hint > 0 ? hint + 1 : 8;
		for (;;) {
			if (chunk > iters - launched)
				chunk = iters - launched;
			for (int k = 0; k < chunk; ++k)
				TRYB(pass_iter(b, ps));
			launched += chunk;
			TRYB(read_flags(b));
			int done = 1;
			for (int q = 0; q < (ps->nown ? ps->nown : ps->nlist); ++q)
				done &= task_flags(b, ps->h_list[q])[0] != 0;
			if (done || launched >= iters)
				break;
			chunk = chunk < 64 ? 2 * chunk : 64;
			if (ps->nown)
				chunk = 8;
		}
	}
	if (ps->nfl)
		TRYB(ssxk_bflush(st, b->d_prob, ps->d_tab + ps->fl_off, ps->nfl));
	return SSX_OK;
}

/* the riding Jacobi problems still iterating */
static int riders(const ssx_batch *b, const int *pc, const unsigned char *ract,
		int *rid)
{
	int n = 0;
	for (int i = 0; i < b->np; ++i)
		if (pc[i] == SSX_PC_JACOBI && ract[i])
			rid[n++] = i;
	return n;
}

/* tables of every pass for the multigrid problems in mact and the riders
 * in ract */
static int build_passes(ssx_batch *b, const int *pc,
		const unsigned char *mact, const unsigned char *ract, int *own,
		int *rid)
{
	const int nr = riders(b, pc, ract, rid);
	int lmax = 0;
	for (int i = 0; i < b->np; ++i)
		if (mact[i] && b->ctx[i]->nlev > lmax)
			lmax = b->ctx[i]->nlev;
	for (int l = 0; l + 1 < lmax; ++l) {
		int no = 0;
		for (int i = 0; i < b->np; ++i)
			if (mact[i] && l + 1 < b->ctx[i]->nlev)
				own[no++] = task_of(b, i, l);
		TRYB(build_pass(b, &b->pass[l], own, no, rid, nr));
	}
	int no = 0;
	for (int i = 0; i < b->np; ++i)
		if (mact[i])
			own[no++] = task_of(b, i, b->ctx[i]->nlev - 1);
	TRYB(build_pass(b, &b->pass[SSX_PASS_COARSE], own, no, rid, nr));
	return SSX_OK;
}

/* one V-cycle (precon.c:227-271) of every multigrid problem in mact:
 * z = M r on each problem's fcg_r / fcg_z */
static int vcycle(ssx_batch *b, const unsigned char *mact, const int *prec,
		const double *eps)
{
	const int np = b->np;
	int lmax = 0;
	for (int i = 0; i < np; ++i)
		if (mact[i]) {
			ssx_ctx *c = b->ctx[i];
			TRYB(ssx_vc_begin(c, prec[i], b->var[i], c->fcg_r, c->fcg_z));
			if (c->nlev > lmax)
				lmax = c->nlev;
		}
	for (int l = 0; l + 1 < lmax; ++l)
		for (int i = 0; i < np; ++i) {
			ssx_ctx *c = b->ctx[i];
			if (mact[i] && l + 1 < c->nlev)
				TRYB(ssx_vc_down(c, prec[i], l, b->op[i], b->var[i],
						c->fcg_r, c->fcg_z));
		}
	/* the coarsest levels (precon.c:242-244): Jacobi-PCG from a zero
	 * guess below level 0 */
	for (int i = 0; i < np; ++i) {
		if (!mact[i])
			continue;
		ssx_ctx *c = b->ctx[i];
		const int l = c->nlev - 1;
		ssx_level *L = &c->lev[l];
		double *ts = b->tscal + CG_NSLOT * b->toff[i] + CG_NSLOT * l;
		double *f = l == 0 ? c->fcg_r : L->mg_fn;
		double *u = l == 0 ? c->fcg_z : L->mg_e;
		TRYB(ssx_cg_start(c, l, b->op[i], b->var[i], f, u, l > 0, b->bpart,
				ts, 0, eps[i], prec[i] != SSX_PREC_FP64
				? c->lp_coarse_rtol : 0.0));
	}
	TRYB(run_pass(b, &b->pass[SSX_PASS_COARSE], COARSE_ITERS, 1,
			b->coarse_hint));
	int most = 0;
	for (int i = 0; i < np; ++i)
		if (mact[i]) {
			ssx_ctx *c = b->ctx[i];
			const int it = task_flags(b, task_of(b, i, c->nlev - 1))[1];
			c->cnt[1] += it;
			if (it > most)
				most = it;
			TRYB(ssx_vc_coarse_done(c, prec[i]));
		}
	b->coarse_hint = most;
	/* ascent: prolongation and post-smoother (precon.c:262-266) */
	for (int l = lmax - 2; l >= 0; --l) {
		int any = 0;
		for (int i = 0; i < np; ++i) {
			ssx_ctx *c = b->ctx[i];
			if (!mact[i] || l + 1 >= c->nlev)
				continue;
			ssx_level *L = &c->lev[l];
			double *ts = b->tscal + CG_NSLOT * b->toff[i] + CG_NSLOT * l;
			TRYB(ssx_vc_up(c, prec[i], l, c->fcg_z));
			if (prec[i] == SSX_PREC_FP64)
				TRYB(ssx_cg_start(c, l, b->op[i], b->var[i],
						l == 0 ? c->fcg_r : L->mg_fn,
						l == 0 ? c->fcg_z : L->mg_e, 0, b->bpart, ts, 0,
						eps[i], 0.0));
			else
				TRYB(ssx_cg_start_lp(c, l, b->op[i], b->var[i], L->fn_lp,
						L->e_lp, b->bpart, ts, 0));
			c->cnt[2] += SMOOTH_ITERS;
			any = 1;
		}
		if (any)
			TRYB(run_pass(b, &b->pass[l], SMOOTH_ITERS, 0, 0));
	}
	for (int i = 0; i < np; ++i)
		if (mact[i])
			TRYB(ssx_vc_end(b->ctx[i], prec[i], b->ctx[i]->fcg_z));
	return SSX_OK;
}

typedef struct {
	int *prec, *own, *rid, *it;
	unsigned char *mact, *ract;
	double *err;
	ssx_fcg_state *fs;
	void **saved;
	long long (*c0)[4];
} bfcg_work;

static int solve_body(ssx_batch *b, const int *pc, double *const *f,
		double *const *u, const double *eps, int max_iter, bfcg_work *w)
{
	const int np = b->np;
	hipStream_t st = (hipStream_t) b->stream;
	for (int i = 0; i < np; ++i) {
		ssx_bprob *d = &b->h_prob[i];
		if (pc[i] == SSX_PC_MG) {
			ssx_ctx *c = b->ctx[i];
			w->prec[i] = mg_prec(c);
			if (w->prec[i] != SSX_PREC_FP64)
				TRYB(ssx_lp_ready(c, w->prec[i]));
			for (int l = 0; l < c->nlev; ++l)
				fill_task(b, i, l, w->prec[i], eps[i]);
			w->mact[i] = 1;
		} else {
			d->u = u[i];
			d->eps = eps[i];
			d->maxit = max_iter;
			d->prec = SSX_PREC_FP64;
			w->ract[i] = 1;
		}
	}
	HIPB(hipMemcpyAsync(b->d_prob, b->h_prob, sizeof(ssx_bprob) * b->ntask,
			hipMemcpyHostToDevice, st));
	/* Jacobi problems: CG start, then their flags and first-iteration
	 * scalars reset (every pass applies the fused direction update) */
	const int nr = riders(b, pc, w->ract, w->rid);
	for (int q = 0; q < nr; ++q)
		TRYB(ssx_batch_setup_problem(b, w->rid[q], f[w->rid[q]],
				u[w->rid[q]]));
	if (nr) {
		TRYB(build_pass(b, &b->pass[SSX_PASS_RIDERS], NULL, 0, w->rid, nr));
		TRYB(ssxk_btask_reset(st, b->d_prob, b->pass[SSX_PASS_RIDERS].d_list,
				nr));
	}
	for (int i = 0; i < np; ++i)
		if (w->mact[i])
			TRYB(ssx_fcg_begin(b->ctx[i], b->op[i], b->var[i], SSX_PC_MG,
					f[i], u[i], &w->fs[i]));
	int rc = SSX_OK, dirty = 1, any_m = 1;
	while (any_m) {
		if (dirty) {
			TRYB(build_passes(b, pc, w->mact, w->ract, w->own, w->rid));
			dirty = 0;
		}
		TRYB(vcycle(b, w->mact, w->prec, eps));
		for (int i = 0; i < np; ++i) {
			if (!w->mact[i])
				continue;
			ssx_ctx *c = b->ctx[i];
			TRYB(ssx_fcg_iter(c, b->op[i], b->var[i], u[i], &w->fs[i]));
			HIPB(hipMemcpyAsync(c->h_scal + SC_RR, c->scal + SC_RR,
					sizeof(double), hipMemcpyDeviceToHost, st));
		}
		TRYB(read_flags(b));
		any_m = 0;
		for (int i = 0; i < np; ++i) {
			if (w->mact[i]) {
				const double err = sqrt(b->ctx[i]->h_scal[SC_RR]);
				w->err[i] = err;
				w->it[i]++;
				if (isnan(err) || isinf(err)) {
					ssx_set_error("batch problem %d: FCG residual is %g", i,
							err);
					rc = SSX_ENUMERIC;
				}
				if (!(err > eps[i]) || w->fs[i].j >= max_iter) {
					w->mact[i] = 0;
					dirty = 1;
				}
				any_m |= w->mact[i];
			}
			if (w->ract[i] && b->h_flags[2 * i]) {
				w->ract[i] = 0;
				dirty = 1;
			}
		}
		if (rc != SSX_OK)
			return rc;
	}
	/* Jacobi problems still iterating go on alone */
	const int nr2 = riders(b, pc, w->ract, w->rid);
	if (nr2) {
		ssx_bpass *ps = &b->pass[SSX_PASS_RIDERS];
		TRYB(build_pass(b, ps, NULL, 0, w->rid, nr2));
		int launched = 0;
		for (int q = 0; q < nr2; ++q) {
			const int li = b->h_flags[2 * w->rid[q] + 1];
			if (li > launched)
				launched = li;
		}
		TRYB(run_pass(b, ps, max_iter > launched ? max_iter - launched : 1,
				1, 0));
	}
	/* the Jacobi problems' last solution step, then their scalars */
	for (int q = 0; q < nr; ++q) {
		const ssx_bprob *d = &b->h_prob[w->rid[q]];
		TRYB(ssxk_cg_uflush(st, (size_t) d->ntot, b->scal + CG_NSLOT
				* w->rid[q], CG_ALPHA, d->p, d->u));
	}
	HIPB(hipMemcpyAsync(b->h_scal, b->scal, sizeof(double) * CG_NSLOT * np,
			hipMemcpyDeviceToHost, st));
	TRYB(read_flags(b));
	return SSX_OK;
}

int ssx_batch_solve(ssx_batch *b, const int *precon, double *const *f,
		double *const *u, const double *eps, int max_iter,
		int use_graph, ssx_stats *stats)
{
	if (!b || !precon || !f || !u || !eps || max_iter < 1) {
		ssx_set_error("ssx_batch_solve: bad arguments");
		return SSX_EINVAL;
	}
	const int np = b->np;
	int nmg = 0;
	for (int i = 0; i < np; ++i) {
		if (precon[i] != SSX_PC_JACOBI && precon[i] != SSX_PC_MG) {
			ssx_set_error("ssx_batch_solve: problem %d: bad preconditioner",
					i);
			return SSX_EINVAL;
		}
		if (!f[i] || !u[i]) {
			ssx_set_error("ssx_batch_solve: null vector for problem %d", i);
			return SSX_EINVAL;
		}
		if (precon[i] != SSX_PC_MG)
			continue;
		++nmg;
		for (int k = 0; k < i; ++k)
			if (precon[k] == SSX_PC_MG && b->ctx[k] == b->ctx[i]) {
				ssx_set_error("ssx_batch_solve: problems %d and %d are both "
						"multigrid solves on one context (they would share "
						"its level vectors)", k, i);
				return SSX_EINVAL;
			}
		for (int l = 0; l < b->ctx[i]->nlev; ++l)
			if (!b->ctx[i]->lev[l].fused_ok) {
				ssx_set_error("ssx_batch_solve: problem %d: level %d needs "
						"the plain CG path", i, l);
				return SSX_EMESH;
			}
	}
	if (nmg == 0)
		return ssx_batch_cg(b, f, u, eps, max_iter, use_graph, stats);
	struct timespec t0, t1;
	clock_gettime(CLOCK_MONOTONIC, &t0);
	TRYB(ssx_batch_sync_settings(b));
	TRYB(ssx_batch_order_after_contexts(b));
	bfcg_work w;
	memset(&w, 0, sizeof(w));
	w.prec = calloc(np, sizeof(int));
	w.own = calloc(np, sizeof(int));
	w.rid = calloc(np, sizeof(int));
	w.it = calloc(np, sizeof(int));
	w.mact = calloc(np, 1);
	w.ract = calloc(np, 1);
	w.err = calloc(np, sizeof(double));
	w.fs = calloc(np, sizeof(ssx_fcg_state));
	w.saved = calloc(np, sizeof(void *));
	w.c0 = calloc(np, sizeof(long long[4]));
	int rc = SSX_OK;
	if (!w.prec || !w.own || !w.rid || !w.it || !w.mact || !w.ract || !w.err
			|| !w.fs || !w.saved || !w.c0) {
		ssx_set_error("out of host memory (batch)");
		rc = SSX_ENOMEM;
	}
	/* the multigrid problems' contexts run on the batch stream meanwhile */
	for (int i = 0; i < np && rc == SSX_OK; ++i)
		if (precon[i] == SSX_PC_MG) {
			w.saved[i] = b->ctx[i]->stream;
			b->ctx[i]->stream = b->stream;
			memcpy(w.c0[i], b->ctx[i]->cnt, sizeof(w.c0[i]));
		}
	if (rc == SSX_OK)
		rc = solve_body(b, precon, f, u, eps, max_iter, &w);
	for (int i = 0; i < np && w.saved && w.fs; ++i)
		if (precon[i] == SSX_PC_MG) {
			ssx_fcg_end(b->ctx[i], &w.fs[i]);
			b->ctx[i]->stream = w.saved[i];
		}
	if (rc == SSX_OK && hipStreamSynchronize((hipStream_t) b->stream)
			!= hipSuccess) {
		ssx_set_error("batch: stream synchronisation failed");
		rc = SSX_EHIP;
	}
	clock_gettime(CLOCK_MONOTONIC, &t1);
	const double secs = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec
			- t0.tv_nsec);
	for (int i = 0; i < np && rc == SSX_OK; ++i) {
		ssx_stats s;
		memset(&s, 0, sizeof(s));
		s.seconds = secs;
		if (precon[i] == SSX_PC_MG) {
			const ssx_ctx *c = b->ctx[i];
			s.iterations = w.it[i];
			s.residual = w.err[i];
			s.vcycles = c->cnt[3] - w.c0[i][3];
			/* FCG operator applications plus the level-0 smoother's */
			s.fine_ax = c->cnt[0] - w.c0[i][0]
				+ (c->nlev > 1 ? SMOOTH_ITERS * s.vcycles : 0);
			s.coarse_iters = c->cnt[1] - w.c0[i][1];
			s.smooth_iters = c->cnt[2] - w.c0[i][2];
		} else {
			s.iterations = b->h_flags[2 * i + 1];
			s.residual = sqrt(b->h_scal[CG_NSLOT * i + CG_RR]);
			s.fine_ax = s.iterations + 1;
		}
		s.converged = s.residual <= eps[i];
		if (isnan(s.residual) || isinf(s.residual)) {
			ssx_set_error("batch problem %d: residual is %g", i, s.residual);
			rc = SSX_ENUMERIC;
		} else if (!s.converged && rc == SSX_OK) {
			rc = SSX_ENOTCONV;
		}
		if (stats)
			stats[i] = s;
	}
	free(w.prec); free(w.own); free(w.rid); free(w.it); free(w.mact);
	free(w.ract); free(w.err); free(w.fs); free(w.saved); free(w.c0);
	return rc;
}
