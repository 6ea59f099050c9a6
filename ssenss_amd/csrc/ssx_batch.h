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
	double lambda, nunique, eps;
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
#ifdef __cplusplus
}
#endif

#endif
