/* ssx_batch_host.h -- host-side state of a batch of independent problems
 * (ssx_batch.c: Jacobi-PCG; ssx_bfcg.c: flexible CG with the p-multigrid
 * V-cycle, whose smoother and coarsest-level CG run as batched passes). */
#ifndef SSX_BATCH_HOST_H
#define SSX_BATCH_HOST_H

#include <hip/hip_runtime_api.h>

#include "ssx_internal.h"
#include "ssx_batch.h"

#define GRAPH_ITERS 4           /* iterations per captured graph */
/* (order, operator, precision) groups of a pass */
#define SSX_BGROUPS (3 * 2 * (SSX_NMAX + 1))
/* pass kinds: smoother at level l < SSX_LMAX, the coarsest levels, and
 * the Jacobi problems alone */
#define SSX_PASS_COARSE SSX_LMAX
#define SSX_PASS_RIDERS (SSX_LMAX + 1)
#define SSX_NPASS (SSX_LMAX + 2)

/* launch tables of one pass: batched CG iterations over a set of tasks.
 * d_tab holds, in order: one operator table per group, the DSS tables
 * (FP64, FP32) of tasks without partner lists, the update tables (FP64,
 * FP32) and the flush table of the pass's own FP64 tasks.  d_list holds
 * the iterated tasks (own, then riding Jacobi problems) followed by the
 * own tasks again (the ones a pass starts). */
typedef struct ssx_bpass {
	int nlist, nown;
	int ngrp;
	int g_n[SSX_BGROUPS], g_op[SSX_BGROUPS], g_prec[SSX_BGROUPS];
	int g_v[SSX_BGROUPS];
	long long g_off[SSX_BGROUPS], g_cnt[SSX_BGROUPS];
	long long dss_off, ndss, dsslp_off, ndsslp;
	long long upd_off, nupd, updlp_off, nupdlp, fl_off, nfl;
	ssx_bwork *d_tab;
	long long tab_cap;
	int *d_list, *h_list;
	int list_cap;
	int valid;
} ssx_bpass;

struct ssx_batch {
	int np;
	ssx_ctx **ctx;
	int *op, *var;
	void *stream, *own_stream;
	/* per-problem work vectors and scratch (device) */
	double **r, **z, **p, **ap;
	double *part, *scal;
	int *flags;
	size_t *poff;              /* partial offsets per problem */
	ssx_bprob *h_prob, *d_prob;
	/* work tables: one operator table per (order, operator) group, rebuilt
	 * for the problems still iterating whenever that set changes */
	int norder, order_n[2 * (SSX_NMAX + 1)], order_op[2 * (SSX_NMAX + 1)];
	int order_v[2 * (SSX_NMAX + 1)];
	ssx_bwork *d_axtab[2 * (SSX_NMAX + 1)];
	long long nax[2 * (SSX_NMAX + 1)];
	ssx_bwork *d_dsstab, *d_updtab;
	long long ndss, nupd;
	ssx_bwork *htab;           /* host staging of the tables */
	long long htab_n;
	unsigned char *active;     /* problems in the current tables */
	double *bpart;             /* set-up reductions [SSX_NPART] */
	int *h_flags;
	double *h_scal;
	hipGraphExec_t gexec;
	int have_graph;
	hipEvent_t ready;          /* orders the batch after the contexts */
	int have_event;
	/* multigrid tasks (ssx_batch_solve): the level-l CG of problem i is
	 * task np + toff[i] + l of h_prob / d_prob; the level's own vectors
	 * of the problem's context are its work vectors */
	int ntask;                 /* np + multigrid tasks */
	int *toff;
	double *tpart, *tscal;
	int *tflags, *h_tflags;
	ssx_bpass pass[SSX_NPASS];
	ssx_bwork *hwork;          /* host staging of pass tables */
	long long hwork_cap;
	int coarse_hint;           /* coarse pass iterations of the last cycle */
};

int ssx_balloc(void **p, size_t bytes);
/* ssx_batch.c: a Jacobi problem's CG start, the contexts' settings read
 * again, the batch stream ordered after the contexts' streams */
int ssx_batch_setup_problem(ssx_batch *b, int i, double *f, double *u);
int ssx_batch_sync_settings(ssx_batch *b);
int ssx_batch_order_after_contexts(ssx_batch *b);
int ssx_bfcg_init(ssx_batch *b);
void ssx_bfcg_free(ssx_batch *b);

#endif
