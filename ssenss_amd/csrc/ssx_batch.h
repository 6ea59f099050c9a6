/* ssx_batch.h -- device-side description of a batch of independent CG
 * problems (shared by ssx_batch.c and ssx_bkernels.hip). */
#ifndef SSX_BATCH_H
#define SSX_BATCH_H

/* nodes per update workgroup: 256 lanes x 8 */
#define SSX_BATCH_CHUNK 2048

/* one problem of the batch, in device memory */
typedef struct ssx_bprob {
	int n, nel, op, pressure, apply_mask, maxit;
	long long ntot;
	double lambda, nunique;
	double eps;                /* < 0: stop at scal[CG_TOL] instead */
	const double *gf, *mass, *dmat, *dinv, *scale, *mask;
	const int *bc;
	const int *g2, *g4, *g8;
	long long G2, G4, G8;
	/* gather form of the DSS (gP > 0): per-copy partner lists in add
	 * order, [nel][gP], and per-node offset << 3 | capacity [n^3]; the
	 * problem then has no DSS launch and its update assembles Ap */
	int gP;
	const int *gsoff, *gtab;
	long long nba, nbd, nbu;   /* operator, DSS, update workgroups */
	double *u, *r, *z, *p, *ap;
	double *part;              /* [nel][nbd][4][nbu] partial sums */
	double *scal;              /* CG_NSLOT scalars (CG_RDZ, ...)   */
	int *flags;                /* [0] done, [1] iterations          */
	/* reduced-precision task (prec SSX_PREC_FP32 / SSX_PREC_BF16): FP32
	 * fields, factors FP32 or BF16; the solution step stays in the
	 * update (no deferred step) */
	int prec;
	const void *gf_lp;
	const float *mass_f, *dmat_f, *dinv_f, *scale_f;
	float *u_f, *r_f, *z_f, *p_f, *ap_f;
} ssx_bprob;

/* work-table entry: problem index and that problem's local workgroup */
typedef struct ssx_bwork {
	int prob, blk;
} ssx_bwork;

#ifdef __cplusplus
extern "C" {
#endif
/* launchers (ssx_bkernels.hip); all return SSX_OK or SSX_EHIP */
int ssxk_bax(void *stream, int n, int op, int v, int pupd,
		const ssx_bprob *pr, const ssx_bwork *tab, long long ntab);
int ssxk_bdss(void *stream, const ssx_bprob *pr, const ssx_bwork *tab,
		long long ntab);
int ssxk_bupdate(void *stream, const ssx_bprob *pr, const ssx_bwork *tab,
		long long ntab);
int ssxk_breduce(void *stream, int mode, const ssx_bprob *pr, int np);
int ssxk_bax_teams(int n, int v);
/* reduced-precision tasks (FP32 fields): operator with the direction
 * update fused (always; beta = 0 on the first iteration), DSS and update
 * with the solution step */
int ssxk_bax_lp(void *stream, int prec, int n, int op, int v,
		const ssx_bprob *pr, const ssx_bwork *tab, long long ntab);
int ssxk_bdss_lp(void *stream, const ssx_bprob *pr, const ssx_bwork *tab,
		long long ntab);
int ssxk_bupdate_lp(void *stream, const ssx_bprob *pr, const ssx_bwork *tab,
		long long ntab);
/* reductions over the tasks list[0..nt) (one workgroup each) */
int ssxk_breduce_list(void *stream, int mode, const ssx_bprob *pr,
		const int *list, int nt);
/* start of a pass for tasks list[0..nt): done flag and count cleared,
 * alpha = beta = null-space constant = 0 (the first iteration's fused
 * direction update then gives p = z and no solution step) */
int ssxk_btask_reset(void *stream, const ssx_bprob *pr, const int *list,
		int nt);
/* FP64 tasks: the solution step u += alpha p the update left to the next
 * operator launch, over the update table */
int ssxk_bflush(void *stream, const ssx_bprob *pr, const ssx_bwork *tab,
		long long ntab);
#ifdef __cplusplus
}
#endif

#endif
