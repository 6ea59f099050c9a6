/*
 * ssx_internal.h -- shared declarations of the ssx host code (C) and the
 * gfx950 kernel launchers (ssx_kernels.hip).
 */
#ifndef SSX_INTERNAL_H
#define SSX_INTERNAL_H

#include <stddef.h>
#include <stdint.h>
#include "ssx.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SSX_NMIN 2
#define SSX_NMAX 16          /* GLL points per direction, kernels 2..16    */
#define SSX_LMAX 10          /* reference structs.c:58 new_preconditioner(n,10) */
#define SSX_NPART 1024       /* blocks of the reduction kernels            */
#define SSX_NSCAL 256        /* device scalar slots                        */
#define SSX_FCG_M 10         /* FCG history length (solver.c:90-92)        */
/* mixed-precision V-cycle: the FP64 coarsest solve stops at
 * max(eps, 10^-k * |f|_w), default k = 3 */
#define SSX_LP_COARSE_RTOL_EXP 3

/* device scalar slots */
enum {
	SC_RDZ = 0, SC_RDZ_OLD, SC_PAP, SC_ALPHA, SC_BETA, SC_RR, SC_TMP,
	SC_NULLC, SC_DONE, SC_PR, SC_ZAP0 /* .. SC_ZAP0+9 */ ,
	SC_PAP0 = SC_ZAP0 + 10 /* .. SC_PAP0+9 */,
	SC_BETA0 = SC_PAP0 + 10 /* .. SC_BETA0+9 */,
	SC_PAPN = SC_BETA0 + 10 /* <p0,Ap0>, <p0,r> */,
	SC_END = SC_PAPN + 2
};

/* per-level scalar block of the fused CG (solver.c:33-81) */
enum { CG_RDZ = 0, CG_PAP, CG_ALPHA, CG_BETA, CG_RR, CG_NULLC, CG_TOL,
	CG_RDZ1, CG_NSLOT = 8 };     /* CG_RDZ1: <r,z> of odd iterations in
	                                the two-launch coarse CG          */
#define SC_CG(l) (64 + CG_NSLOT * (l))

/* host-side basis of one level */
typedef struct {
	int n;
	double xi[SSX_NMAX], rho[SSX_NMAX];
	double d[SSX_NMAX * SSX_NMAX];     /* d[i*n+l] = D_il = l_l'(xi_i)  */
	double dsvv[SSX_NMAX * SSX_NMAX];  /* same layout                   */
} ssx_basis;

/* the coarsest multigrid level's operators assembled over unique nodes
 * (CSR, FP64) and the work vectors of its CG in that space */
typedef struct ssx_coarse {
	int nu;                      /* unique nodes                          */
	long long nnz;
	int *rowptr, *col;           /* device CSR structure                  */
	double *valp, *valh;         /* Poisson / Helmholtz values            */
	int *rep;                    /* [nu] one local copy of each node      */
	int *cmap;                   /* [ntot] unique node of each copy       */
	double *dinvp, *dinvh, *mask, *ones;
	double *f, *x, *r, *z, *p0, *p1, *ap;
} ssx_coarse;

/* one multigrid level on the device */
typedef struct {
	int n, n3;
	size_t ntot;                 /* nel*n^3 */
	ssx_basis b;
	/* device arrays */
	double *gf;                  /* [e][6][n3] G*w*detJ                 */
	double *mass;                /* [e][n3] w*detJ                       */
	double *diag_p, *diag_h;     /* inverse assembled diagonals          */
	double *scale;               /* 1/mult                               */
	double *mask;                /* Dirichlet mask                       */
	double *dmat;                /* [2][n*n]: D, Dsvv, layout [i*n+l]    */
	double *prol;                /* [nc][n]: l_p^{l+1}(xi_i^l) at [p*n+i] */
	int *grp[3];                 /* shared-node groups, m = 2,4,8: [m][G] */
	long long ngrp[3];
	int *bc;                     /* per element Dirichlet bits           */
	double nunique;              /* <1,1>_w                              */
	ssx_coarse cs;               /* coarsest level (l > 0) only: nu > 0   */
	int axv[3];                  /* operator launch variant per precision
	                                (SSX_PREC_*), set by the autotuner   */
	int fused_ok;                /* multiplicity > 1 exactly on faces with
	                                a neighbour: fused inner products OK */
	int *gtab, *gsoff;           /* gather DSS: partners [nel][gP] in add
	                                order, per-node offset<<3|capacity   */
	int gP;                      /* 0: no gather tables on this level    */
	/* solver scratch */
	double *r, *z, *p, *ap, *mg_r, *mg_fn, *mg_e;
	/* host copies kept for ssx_level_data */
	double *h_x, *h_y, *h_z;
	/* low-precision smoother copies (FP32 fields; factors FP32 or BF16),
	 * built on first use of SSX_OPT_SMOOTHER_PRECISION != FP64 */
	int lp_prec;                 /* precision of gf_lp, 0 = not built     */
	void *gf_lp;
	float *mass_lp, *diagp_lp, *diagh_lp, *scale_lp, *mask_lp, *dmat_lp;
	float *prol_lp;
	float *r_lp, *z_lp, *p_lp, *ap_lp, *mgr_lp, *fn_lp, *e_lp;
} ssx_level;

struct ssx_ctx {
	int nel, nlev, n;
	int device;                  /* HIP device current at ssx_create      */
	int nmg[SSX_LMAX];
	ssx_level lev[SSX_LMAX];
	double lambda;
	int diag_choice;
	int fused;                   /* SSX_OPT_FUSED_CG                     */
	int smooth_prec;             /* SSX_OPT_SMOOTHER_PRECISION           */
	double lp_coarse_rtol;       /* SSX_OPT_LP_COARSE_RTOL               */
	double *inv_mass;            /* device, level 0 */
	void *stream;                /* hipStream_t in use */
	void *own_stream;
	double *scal;                /* device scalars [SSX_NSCAL]           */
	double *part;                /* device partial sums [8][SSX_NPART]   */
	double *h_scal;              /* pinned host mirror                   */
	int *cg_flags;               /* device: [2*l] done, [2*l+1] count    */
	int *h_flags;                /* pinned host mirror                   */
	int cg_hint[SSX_LMAX];       /* last iteration count per level       */
	int persist;                 /* SSX_OPT_PERSISTENT_CG                */
	int fused_red;               /* SSX_OPT_FUSED_REDUCE                 */
	int coarse_csr;              /* SSX_OPT_COARSE_ASSEMBLED             */
	int gather_upd;              /* SSX_OPT_GATHER_UPDATE                */
	unsigned *redctr;            /* device: last-workgroup counter       */
	int persist_grid;            /* blocks of the cooperative CG launch  */
	int persist_abort_test;      /* SSX_OPT_PERSIST_ABORT_TEST           */
	unsigned *pbar;              /* device: grid-barrier words [4]       */
	size_t npart;                /* doubles in part                      */
	/* FCG history (level 0) */
	double *fcg_r, *fcg_z, *fcg_p[11], *fcg_ap[10];
	/* host staging for the host-array entry points */
	double *d_f, *d_u;
	/* particles (level 0): J, J^-1, reference-space velocity history */
	int *h_neigh, *d_neigh;
	double *p_jac[9], *p_jinv[9], *p_ur[3][3], *p_xi;
	int *p_maxpass;
	int p_oriented;              /* every face pair's neighbours point back */
	void *p_stage;               /* device staging of host particle lists */
	size_t p_stage_bytes;
	int *p_perm, *p_bucket;      /* processing order by element, buckets  */
	int *p_work, *p_pcnt;        /* pass lists and state (5 x cap), counts */
	long p_perm_cap;
	int p_elem_walk;             /* SSX_OPT_PARTICLE_WALK                 */
	int coarse_direct;           /* SSX_OPT_COARSE_DIRECT: max unknowns   */
	double *cd_inv;              /* dense inverse of the coarsest operator */
	int cd_n, cd_op, cd_pres, cd_bad, cd_sing;
	double cd_lambda;
	void *cd_blas;               /* rocblas handle of the factorisation   */
	int coarse_sparse;           /* SSX_OPT_COARSE_SPARSE: max unknowns   */
	struct ssx_sdfactor *sd;     /* sparse factor of the coarsest operator */
	int sd_op, sd_pres, sd_bad, sd_sing, sd_nu;
	double sd_lambda;
	int *sd_gmap, *sd_smap;      /* coarsest E-vector copy of each new index,
	                                new index of each E-vector copy (-1) */
	long p_pass_np;              /* particles of the last pass-by-pass step */
	long long cnt[4];
	double ax_us[SSX_LMAX][3][SSX_AX_NVARIANTS];  /* autotuner timings  */
};

/* ---- sparse direct solver (ssx_sdirect.c, ssx_sdkernels.hip) ---- */
/* the device solve stages a supernode's vectors (at most SD_MAX_VEC
 * entries: its columns and the rows below them) in LDS */
#define SD_MAX_VEC 16384

/* one supernode of the factor: columns [c0, c0+nc) of the new order, nu
 * rows below them (their new indices at idx[uoff..]), children ch0/ch1
 * (-1: none) with gather maps at idx[g0..] / idx[g1..] (nc + nu entries:
 * for each row of this front, the child's update entry landing there or
 * -1); M = L11^-1 (nc x nc) and K = L21 M (nu x nc), column-major at
 * val[moff] / val[koff] and row-major at val[mroff] / val[kroff]; the
 * update vector t at T[toff] (nu) */
typedef struct {
	int c0, nc, nu, ch0, ch1, g0, g1, uoff;
	long long moff, koff, toff, mroff, kroff;
} ssx_sdnode;

typedef struct {
	int sn, blk;
} ssx_sdwork;

typedef struct ssx_sdfactor {
	int n, nsn, nlev;
	ssx_sdnode *node;
	int *perm;                   /* [n] system index of new index i       */
	int *lev_off, *lev_sn;       /* supernodes level by level, leaves first */
	long long nval, nidx, nt;
	double *val;
	int *idx;
	double flops;
	int ndev;                    /* fronts factored on the device */
	int *wk_off;                 /* [2 nlev + 1] work-table ranges: per
	                                level, forward then backward         */
	int *wk_vec;                 /* [2 nlev] LDS entries of those launches */
	ssx_sdwork *wk;
	/* device */
	ssx_sdnode *d_node;
	int *d_idx, *d_perm;
	double *d_val;
	ssx_sdwork *d_wk;
	double *d_b, *d_y, *d_x, *d_t;
} ssx_sdfactor;

/* blas, stream: a rocblas handle and its stream to factor the large
 * fronts on the device, or NULL (host only) */
int ssx_sd_factor(int n, const int *rp, const int *ci, const double *a,
		const double *xyz, void *blas, void *stream, ssx_sdfactor **out);
int ssx_sd_solve_host(const ssx_sdfactor *f, const double *b, double *x);
int ssx_sd_upload(ssx_sdfactor *f);
/* x = A^-1 b on the device, b and x in the system's numbering (n) */
int ssx_sd_solve_dev(ssx_sdfactor *f, void *stream, const double *b,
		double *x);
/* the same with b read as src[gmap[i]] for new index i (gmap: n entries)
 * and x written to dst[q] = x_new[smap[q]] (0 where smap[q] < 0), q < m */
int ssx_sd_solve_dev_map(ssx_sdfactor *f, void *stream, const double *src,
		const int *gmap, double *dst, const int *smap, int m);
void ssx_sd_free(ssx_sdfactor *f);
/* workgroups of a product: nout outputs of rows of length rowlen */
int ssxk_sd_blocks(int nout, int rowlen);
int ssxk_sd_gather(void *stream, int n, const int *perm, const double *src,
		double *dst);
int ssxk_sd_scatter(void *stream, int n, const int *perm, const double *src,
		double *dst);
/* dst[q] = src[map[q]] (0 where map[q] < 0), q < m */
int ssxk_sd_scatter_map(void *stream, int m, const int *map,
		const double *src, double *dst);
/* vec: LDS entries of the launch (the level's largest nc for the forward
 * step, nc + nu for the backward one) */
int ssxk_sd_fwd(void *stream, int nwk, const ssx_sdwork *wk,
		const ssx_sdnode *node, const double *val, const int *idx,
		const double *b, double *y, double *t, int vec);
int ssxk_sd_bwd(void *stream, int nwk, const ssx_sdwork *wk,
		const ssx_sdnode *node, const double *val, const int *idx,
		const double *y, double *x, int vec);

/* ---- assembled coarsest level (ssx_coarse.c) ---- */
int ssx_coarse_build(ssx_ctx *c, int l, const double *gf, const double *mass,
		const double *dp, const double *dh, const double *mask);
void ssx_coarse_free(ssx_coarse *cs);
/* frees the direct coarse solve's inverse and rocblas handle */
void ssx_coarse_direct_free(ssx_ctx *c);
/* Jacobi-PCG (solver.c:33-81) of level l in unique-node space: f, x
 * E-vectors of the level (f averaged over copies; x_zero: x is zero on
 * entry); stops at max(eps, rtol |r0|) or max_iter */
int ssx_coarse_cg(ssx_ctx *c, int l, int op, int var, const double *f,
		double *x, double eps, double rtol, int max_iter, int x_zero,
		int *iters, double *err);
int ssxk_gather_u(void *stream, int nu, const int *rep, const double *e,
		double *u);
int ssxk_scatter_u(void *stream, size_t ntot, const int *cmap,
		const double *u, double *e);
/* ap = mask (A pn), pn = first ? pold : mask ((z - c) + beta pold) written
 * to pnew; part[blk] = partial sum of pn ap (unique nodes: weight 1) */
int ssxk_csr_pupd(void *stream, int nu, const int *rowptr, const int *col,
		const double *val, const double *pold, double *pnew,
		const double *z, const double *mask, const double *scal, int ib,
		int ic, int first, double *ap, double *part, const int *done);
int ssxk_csr_blocks(int nu);
int ssxk_csr_dense(void *stream, int nu, const int *rowptr, const int *col,
		const double *val, const double *mask, double reg, double *A);
int ssxk_dense_sym(void *stream, int nu, double *A);
int ssxk_dense_mv(void *stream, int nu, const double *A, const double *x,
		double *y);
/* Two-launch coarse CG iteration j (0-based) over unique nodes.
 * ssxk_csr_cg_spmv: every workgroup first replays the stopping test and
 * beta of iteration j-1 from the update partials pu (nbu per quantity,
 * stride SSX_NPART) exactly as ssxk_cg_reduce mode 1 does, workgroup 0
 * storing the scalars (<r,z>_j in slot CG_RDZ or CG_RDZ1 by parity of j,
 * count j, done flag); then, unless stopped, pnew = mask((z - c) + beta
 * pold) (j = 0: pnew = pold) and ap = mask(A pnew), partials of <pnew,ap>
 * to ps.  check != 0: the prologue only, one workgroup (the host's
 * stopping test after a chunk; idempotent).
 * ssxk_csr_cg_update: every workgroup forms alpha = <r,z>_j / sum(ps) as
 * ssxk_cg_reduce mode 0 does (workgroup 0 stores PAP, ALPHA), then the
 * ssxk_cg_update_nb step with its partials to pu.  ps and pu must not
 * overlap. */
int ssxk_csr_cg_spmv(void *stream, int nu, const int *rowptr, const int *col,
		const double *val, const double *pold, double *pnew,
		const double *z, const double *mask, double *scal, int o, int j,
		double eps, int maxit, double nunique, int pressure, double *ap,
		double *ps, const double *pu, int nbu, int *flags, int check);
int ssxk_csr_cg_update(void *stream, int nu, double *scal, int o, int j,
		const double *p, const double *ap, double *u, double *r, double *z,
		const double *dinv, const double *scale, int pressure,
		const double *ps, int nbs, double *pu, int nbu, const int *flags);

/* ---- host setup (ssx_setup.c) ---- */
void ssx_basis_init(ssx_basis *b, int n, int svvn, double svva);
void ssx_interp(const double *xfrom, int nfrom, const double *xto, int nto,
		double *out);
/* build the shared-node groups of one element size from the connectivity,
 * reproducing the reference's add order; returns SSX_OK or SSX_EMESH.
 * grp[m] receives malloc'ed [mult][G] arrays of local node indices. */
int ssx_build_groups(const ssx_mesh *m, int n, int *grp[3], long long ng[3]);
/* host DSS with the groups (same order as the device kernel) */
void ssx_host_dss(int *grp[3], const long long ng[3], double *u);
/* per-copy partner lists in add order for the gather form of the DSS:
 * tab [nel][P], soff [n^3] = offset << 3 | capacity (host arrays) */
int ssx_build_gather(const ssx_mesh *m, int n, int **tab, int **soff,
		int *P);
void ssx_set_error(const char *fmt, ...);

/* ---- kernel launchers (ssx_kernels.hip); all return 0 or SSX_EHIP ---- */
/* bc[e]: bits 0-5 Dirichlet faces (W,E,S,N,B,T), bits 8-13 faces with a
 * neighbour; done (nullable): skip when *done; part (nullable): part[e] =
 * sum of u (A_e u) over element e's nodes, one partial per element
 * whatever the launch shape v */
int ssxk_ax(void *stream, int n, int v, int nel, int op, const double *u,
		double *au, const double *gf, const double *mass,
		const double *dmat, double lambda, const int *bc, int apply_mask,
		const int *done, double *part);
/* ssxk_ax of the new CG direction: p = (z - scal[ic]) + scal[ib] p
 * (ic < 0: no null-space constant), masked where apply_mask, written back
 * to p, and ap = A p -- the k_cg_pupdate of the previous iteration fused
 * into the operator; u (nullable): first u += scal[ia] p (old p), the
 * previous iteration's solution step */
int ssxk_ax_pupd(void *stream, int n, int v, int nel, int op, const double *z,
		double *p, const double *scal, int ib, int ic, double *ap,
		const double *gf, const double *mass, const double *dmat,
		double lambda, const int *bc, int apply_mask, const int *done,
		double *part, double *u, int ia);
int ssxk_ax_pupd_lp(void *stream, int prec, int n, int v, int nel, int op,
		const void *z, void *p, const double *scal, int ib, int ic, void *ap,
		const void *gf, const void *mass, const void *dmat, double lambda,
		const int *bc, int apply_mask, const int *done, double *part,
		void *u, int ia);
/* u += scal[ia] p whatever the done flags (the last solution step of a CG
 * whose steps the operator launches take: ssxk_ax_pupd with u, updates
 * with u = NULL) */
int ssxk_cg_uflush(void *stream, size_t n, const double *scal, int ia,
		const double *p, double *u);
int ssxk_cg_uflush_lp(void *stream, size_t n, const double *scal, int ia,
		const void *p, void *u);
long long ssxk_dss_blocks(const long long ng[3]);
/* DSS of all groups in one launch; average: scale by 1/M; p, part
 * (nullable): per block shared-node part of <p, DSS(u)>_w */
int ssxk_dss_all(void *stream, int *const grp[3], const long long ng[3],
		double *u, const double *p, double *part, int average,
		const int *done);
/* u += a p (skipped with u = NULL); r -= a ap; z = dinv r; partials */
int ssxk_cg_update(void *stream, size_t n, const double *scal, int ia,
		const double *p, const double *ap, double *u, double *r, double *z,
		const double *dinv, const double *scale, int pressure, double *part,
		const int *done);
int ssxk_cg_pupdate(void *stream, size_t n, const double *z, double *p,
		const double *scal, int ib, int ic, const double *mask,
		const int *done);
/* The two reductions of a fused CG iteration done by the last workgroup
 * of the DSS kernel (mode 0: <p,Ap>_w -> alpha) and of the update kernel
 * (mode 1: beta, null-space constant, stopping test) instead of two
 * single-workgroup launches; sums in the order of ssxk_cg_reduce, so the
 * results are the same bit for bit.  counter: zero between launches. */
typedef struct ssx_lastred {
	unsigned *counter;
	const double *part;      /* mode 0: [0, cnt); mode 1: c*SSX_NPART + q */
	long long cnt;
	double *scal;
	int o, maxit, pressure;
	int *flags;
	double eps, nunique;
} ssx_lastred;
int ssxk_dss_all_red(void *stream, int *const grp[3], const long long ng[3],
		double *u, const double *p, double *part, const int *done,
		const ssx_lastred *lr);
int ssxk_dss_all_red_lp(void *stream, int *const grp[3],
		const long long ng[3], void *u, const void *p, double *part,
		const int *done, const ssx_lastred *lr);
int ssxk_cg_update_red(void *stream, size_t n, const double *scal, int ia,
		const double *p, const double *ap, double *u, double *r, double *z,
		const double *dinv, const double *scale, int pressure, double *part,
		const int *done, const ssx_lastred *lr);
int ssxk_cg_update_red_lp(void *stream, size_t n, const double *scal, int ia,
		const void *p, const void *ap, void *u, void *r, void *z,
		const void *dinv, const void *scale, int pressure, double *part,
		const int *done, const ssx_lastred *lr);
/* ssxk_cg_update with the DSS of ap as a gather over per-copy partner lists
 * (tab, soff from ssx_build_gather): ap is read unassembled and left so;
 * weights 1/M from the lists (no scale vector).  Bit-identical to
 * ssxk_dss_all followed by ssxk_cg_update. */
int ssxk_cg_update_g(void *stream, size_t n, int n3, int P, const int *soff,
		const int *tab, const double *scal, int ia, const double *p,
		const double *ap, double *u, double *r, double *z, const double *dinv,
		int pressure, double *part, const int *done);
int ssxk_cg_update_g_lp(void *stream, size_t n, int n3, int P,
		const int *soff, const int *tab, const double *scal, int ia,
		const void *p, const void *ap, void *u, void *r, void *z,
		const void *dinv, int pressure, double *part, const int *done);
/* r = f - DSS(a) with the DSS of the unassembled a as a gather (lists as
 * ssxk_cg_update_g); init != 0: the CG start -- f = f * mask first (mask
 * nullable, f written back), then r, z = dinv r, p = z and the k_dot
 * partials of <r,z>_w in part[0..SSX_NPART) (ssxk_reduce slot 0) */
int ssxk_resid_g(void *stream, int init, size_t n, int n3, int P,
		const int *soff, const int *tab, const double *a, double *f,
		const double *mask, double *r, double *z, double *p,
		const double *dinv, double *part);
int ssxk_resid_g_lp(void *stream, int init, size_t n, int n3, int P,
		const int *soff, const int *tab, const void *a, void *f,
		const void *mask, void *r, void *z, void *p, const void *dinv,
		double *part);
/* the mixed V-cycle's ends: y = (float) x and e = y * dinv (FP32 dinv);
 * y = the FP64 DSS-average of the FP32 E-vector e as a gather (lists as
 * ssxk_cg_update_g): bit-identical to ssxk_cvt + averaging ssxk_dss_all */
int ssxk_d2f_jacobi(void *stream, size_t n, const double *x, void *y,
		const void *dinv, void *e);
int ssxk_f2d_average_g(void *stream, size_t n, int n3, int P,
		const int *soff, const int *tab, const void *e, double *y);
/* ssxk_cg_update on a grid of nblocks <= SSX_NPART workgroups (the mode-1
 * reduce then takes cnt = nblocks) */
int ssxk_cg_update_nb(void *stream, size_t n, const double *scal, int ia,
		const double *p, const double *ap, double *u, double *r, double *z,
		const double *dinv, const double *scale, int pressure, double *part,
		const int *done, int nblocks);
/* eps < 0: the stopping tolerance is read from scal[o + CG_TOL]; mode 1
 * sums cnt partials per quantity (0: SSX_NPART) */
int ssxk_cg_reduce(void *stream, int mode, const double *part, long long cnt,
		double *scal, int o, int *flags, double eps, int maxit,
		double nunique, int pressure);
/* scal[dst] = max(eps, rtol * sqrt(scal[src])) */
int ssxk_cg_tol(void *stream, double *scal, int dst, int src, double eps,
		double rtol);
int ssxk_set_ints(void *stream, int *dst, int v, int n);
/* FCG: scal[dst+k] = <x, ys[k]>_w for k < m (each bit-identical to
 * ssxk_dot + ssxk_reduce); div >= 0: scal[quot+k] = scal[dst+k] /
 * scal[div+k] */
int ssxk_multidot(void *stream, size_t n, const double *x, const double *s,
		const double *const *ys, int m, double *part, double *scal, int dst,
		int div, int quot);
/* y = mask * (x - sum_k scal[ib+k] ps[k]), k = 0..m-1 in order */
int ssxk_multi_axpy(void *stream, size_t n, const double *x,
		const double *const *ps, int m, const double *scal, int ib,
		const double *mask, double *y);
/* u += a p; r -= a ap; part[blk] = partial <r,r>_w (a = scal[ia]) */
int ssxk_fcg_step(void *stream, size_t n, const double *scal, int ia,
		const double *p, const double *ap, double *u, double *r,
		const double *s, double *part);
/* low-precision (FP32 fields) variants; prec selects the factor type of
 * ssxk_ax_lp (SSX_PREC_FP32 or SSX_PREC_BF16) */
int ssxk_ax_lp(void *stream, int prec, int n, int v, int nel, int op,
		const void *u, void *au, const void *gf, const void *mass,
		const void *dmat, double lambda, const int *bc, int apply_mask,
		const int *done, double *part);
int ssxk_dss_all_lp(void *stream, int *const grp[3], const long long ng[3],
		void *u, const void *p, double *part, int average, const int *done);
int ssxk_cg_update_lp(void *stream, size_t n, const double *scal, int ia,
		const void *p, const void *ap, void *u, void *r, void *z,
		const void *dinv, const void *scale, int pressure, double *part,
		const int *done);
int ssxk_cg_pupdate_lp(void *stream, size_t n, const void *z, void *p,
		const double *scal, int ib, int ic, const void *mask,
		const int *done);
int ssxk_restrict_lp(void *stream, int nf, int nc, int nel, const void *prol,
		const void *fine, void *coarse);
int ssxk_prolong_lp(void *stream, int nf, int nc, int nel, const void *prol,
		const void *coarse, void *fine, int accumulate);
int ssxk_axpby_lp(void *stream, size_t n, double a, const void *x,
		const void *w, double b, void *y);
int ssxk_mul_lp(void *stream, size_t n, const void *x, const void *w,
		void *z);
int ssxk_dot_lp(void *stream, size_t n, const void *x, const void *s,
		const void *y, double *part, int slot);
int ssxk_shift_s_lp(void *stream, size_t n, const double *scal, int ia,
		double nunique, void *y);
int ssxk_fill_lp(void *stream, void *dst, double v, size_t n);
int ssxk_copy_lp(void *stream, void *dst, const void *src, size_t n);
int ssxk_particles_ur(void *stream, size_t nt, const double *const jinv[9],
		const double *u, const double *v, const double *w, double *u0,
		double *u1, double *u2);
/* perm: the order the particles are taken in (NULL: as stored) */
int ssxk_tracers_step(void *stream, int n_pts, int nel, const double *xi,
		const double *const ur[9], const int *neigh, double T, double dt,
		long np, const int *perm, int *el, int *old_el, double *x, double *y,
		double *z, double *u, double *v, double *w, double *t, int *reinit,
		int *flags);
/* perm = particle indices sorted by owning element and, for n > 5, the
 * z plane their interpolation stencil starts at (counting sort; cnt:
 * ssxk_particle_order_ints ints; z: the particles' reference z) */
long ssxk_particle_order_buckets(int nel, int n_pts);
long ssxk_particle_order_ints(int nel, int n_pts);
int ssxk_particle_order(void *stream, long np, const int *el,
		const double *z, const double *xi, int nel, int n_pts, int *cnt,
		int *perm);
/* The relocation loop of a particle or tracer runs at most 12 iterations
 * (after the 11th the particle is re-seeded and stops). */
#define SSX_WALK_PASSES SSX_PASS_COUNTS
/* The walk one loop iteration per pass: pass 0 over perm, passes 1 .. 11
 * over the particles still moving, compacted (the particles' pass 1 sorted
 * again, with bucket, the order's scratch).  work: 5 np ints, cnt:
 * 2 (SSX_WALK_PASSES + 1) ints, the first SSX_WALK_PASSES + 1 the number
 * of particles entering each pass. */
int ssxk_particles_passes(void *stream, int n_pts, int nel, const double *xi,
		const double *const ur[9], const double *const jac[9],
		const double *const jinv[9], const int *neigh, double T, double dt,
		long np, const int *perm, int *bucket, int *work, int *cnt, int *el,
		int *old_el, double *x, double *y, double *z, double *u, double *v,
		double *w, double *t, const double *tau, int *reinit, int *flags);
int ssxk_tracers_passes(void *stream, int n_pts, int nel, const double *xi,
		const double *const ur[9], const int *neigh, double T, double dt,
		long np, const int *perm, int *work, int *cnt, int *el, int *old_el,
		double *x, double *y, double *z, double *u, double *v, double *w,
		double *t, int *reinit, int *flags);
int ssxk_particles_step(void *stream, int n_pts, int nel, const double *xi,
		const double *const ur[9], const double *const jac[9],
		const double *const jinv[9], const int *neigh, double T, double dt,
		long np, const int *perm, int *el, int *old_el, double *x, double *y,
		double *z, double *u, double *v, double *w, double *t,
		const double *tau, int *reinit, int *flags);
/* whole Jacobi-PCG solve of one level with n <= 4 in one cooperative
 * launch (ssx_persist.hip); setup (r, z, p, <r,z>_w, flags) done by the
 * caller, bar[0..2] zeroed */
int ssxk_cg_persist(void *stream, int n, int grid, int nel, int op,
		int apply_mask, int pressure, int maxit, double eps, double lambda,
		double nunique, const double *gf, const double *mass,
		const double *dmat, const double *dinv, const double *scale,
		const double *mask, const int *bc, int *const grp[3],
		const long long ng[3], double *u, double *r, double *z, double *p,
		double *ap, double *part, double *scal, int o, int *flags,
		unsigned *bar);
int ssxk_cg_persist_max_grid(int n, int *grid);
/* precision conversion (SSX_PREC_FP64 -> FP32 / BF16, FP32 -> FP64) */
int ssxk_cvt(void *stream, int from, int to, const void *x, void *y,
		size_t n);
int ssxk_restrict(void *stream, int nf, int nc, int nel, const double *prol,
		const double *fine, double *coarse);
int ssxk_prolong(void *stream, int nf, int nc, int nel, const double *prol,
		const double *coarse, double *fine, int accumulate);
/* y[i] = a*x[i]*w[i] + b*y[i]  (w may be NULL = 1) */
int ssxk_axpby(void *stream, size_t n, double a, const double *x,
		const double *w, double b, double *y);
/* z = x * w */
int ssxk_mul(void *stream, size_t n, const double *x, const double *w,
		double *z);
/* per-block partials of sum(x*s*y) into part[slot*NPART + blk] */
/* flexible_conjugate_gradient in phases (ssx_fcg and the batch) */
typedef struct ssx_fcg_state {
	double *p[11], *ap[10];      /* direction and A-direction history    */
	long j;                      /* iterations taken                      */
	int mmax;                    /* history length (10 with p-MG)         */
} ssx_fcg_state;
int ssx_fcg_begin(ssx_ctx *c, int op, int var, int precon, double *f,
		const double *u, ssx_fcg_state *s);
int ssx_fcg_iter(ssx_ctx *c, int op, int var, double *u, ssx_fcg_state *s);
void ssx_fcg_end(ssx_ctx *c, const ssx_fcg_state *s);

/* J, J^-1 and det J of a mesh at the nodes of basis b (ssx_api.c) */
void ssx_jacobians(const ssx_basis *b, int nel, const double *x,
		const double *y, const double *z, double *const jac[9],
		double *const jinv[9], double *det_out);

/* level CG starts (ssx_api.c), shared by the single-problem and batched
 * solvers */
int ssx_cg_start(ssx_ctx *c, int l, int op, int var, double *f,
		const double *u, int x_zero, double *part, double *scal, int o,
		double eps, double rtol);
int ssx_cg_start_lp(ssx_ctx *c, int l, int op, int var, float *f,
		const float *u, double *part, double *scal, int o);
/* V-cycle stages (prec: the smoother precision SSX_PREC_*) */
int ssx_lp_ready(ssx_ctx *c, int prec);
const double *ssx_jdiag(ssx_ctx *c, int l, int var);
const float *ssx_jdiag_lp(ssx_ctx *c, int l, int var);
int ssx_vc_begin(ssx_ctx *c, int prec, int var, const double *x, double *px);
int ssx_vc_down(ssx_ctx *c, int prec, int l, int op, int var,
		const double *x, const double *px);
int ssx_vc_coarse_done(ssx_ctx *c, int prec);
int ssx_vc_up(ssx_ctx *c, int prec, int l, double *px);
int ssx_vc_end(ssx_ctx *c, int prec, double *px);

/* one step of the particle walk (inertial particles, or RK2 tracers) on
 * device particle arrays; counters: two device ints (largest pass count,
 * particles naming no element), NULL for the context's own, which then
 * also allows the particle tables to be set up on first use */
int ssx_walk_dev(ssx_ctx *c, const ssx_particles *p, double t_now, double dt,
		int *max_passes, int tracers, int *counters);
/* the same on host particle arrays staged through `stage` (stage_cap
 * bytes; NULL: the context's own staging, grown on demand) */
int ssx_walk_host(ssx_ctx *c, const ssx_particles *h, double t_now,
		double dt, int *max_passes, int tracers, void *stage,
		size_t stage_cap, int *counters);
/* device bytes ssx_walk_host stages for n particles */
size_t ssx_walk_stage_bytes(size_t n);
int ssxk_dot(void *stream, size_t n, const double *x, const double *s,
		const double *y, double *part, int slot);
/* scal[dst] = sum of part[slot*NPART ...] (fixed order) */
int ssxk_reduce(void *stream, double *part, int slot, double *scal, int dst);
/* scalar program: see ssx_kernels.hip */
int ssxk_scalar(void *stream, double *scal, int op, int a, int b, int c,
		double v);
/* y += scal[ia]*sgn*x  */
int ssxk_axpy_s(void *stream, size_t n, const double *scal, int ia,
		double sgn, const double *x, double *y);
/* y = x + scal[ib]*y, then mask (mask may be NULL) */
int ssxk_xpby_s(void *stream, size_t n, const double *x, const double *scal,
		int ib, double *y, const double *mask);
/* y += c*1 with c = -scal[ia]/nunique */
int ssxk_shift_s(void *stream, size_t n, const double *scal, int ia,
		double nunique, double *y);
int ssxk_copy(void *stream, double *dst, const double *src, size_t n);
int ssxk_fill(void *stream, double *dst, double v, size_t n);

#ifdef __cplusplus
}
#endif
#endif
