/*
 * oracle.h -- CPU restatement of openQCD-2.4.2's Wilson-clover hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code.  It is used by tests/, by __graft_entry__.smoke() as the checker
 * and by bench.py's cpu_baseline leg.  Each function states the reference
 * file:line whose behaviour it restates.  The restatement keeps the
 * reference's data formats (spinor = 24 doubles [spin][colour][re,im],
 * su3 = 18 doubles row-major, pauli = 36 reals, see include/su3.h:33-121 of
 * the reference) and its site order `ipt` (lattice/geometry.c), so that
 * fields can be compared element by element with the reference build under
 * oracle/_ref and with the device library after export.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   geometry      - bijection/neighbour checks + reference build
 *   ranlux/gauss  - known-answer vectors of devel/nompi/random/check1.c
 *   random_ud/sd  - reference build (same seed, same NTHREAD layout)
 *   Dw_dble, Dw   - free-field plane-wave closed form, gamma5-hermiticity,
 *                   gauge covariance, reference build
 *   sw_term       - Abelian constant field closed form, reference build
 *   inv_pauli     - numpy inverse on random Hermitian matrices
 *   qsum          - exact rational arithmetic (fractions)
 *   sap/fgcr      - SAP invariant rho = eta - D psi, FP64 true residual
 */
#ifndef ORACLE_H
#define ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oc_ctx oc_ctx;

/* ---- context / geometry (lattice/geometry.c, lattice/geogen.c) ---- */
oc_ctx *oc_create(const int *L, const int *Ltrd);
void oc_destroy(oc_ctx *c);
int oc_volume(const oc_ctx *c);
int oc_nthread(const oc_ctx *c);
void oc_get_ipt(const oc_ctx *c, int *ipt);          /* lex -> index   */
void oc_get_iupdn(const oc_ctx *c, int *iup, int *idn); /* [V][4] each */

/* ---- parameters (flags/lat_parms.c, flags/sap_parms.c) ---- */
void oc_set_bc(oc_ctx *c, int type, double cF0, double cF1);
void oc_set_sw(oc_ctx *c, double m0, double csw);
void oc_set_tm(oc_ctx *c, int eoflg);
void oc_set_sap(oc_ctx *c, const int *bs, int isolv, int nmr, int ncy);

/* ---- random numbers (random/ranlux*.c, random/gauss.c) ---- */
void oc_start_ranlux(oc_ctx *c, int level, int seed);
void oc_ranlxd(oc_ctx *c, int thread, double *r, int n);
void oc_ranlxs(oc_ctx *c, int thread, float *r, int n);
void oc_gauss_dble(oc_ctx *c, int thread, double *r, int n);
/* standalone KAT helpers: single-stream ranlxs/ranlxd with rlxs_init(1,..) */
void oc_kat_ranlxs(int level, int seed, float *r, int n);
void oc_kat_ranlxd(int level, int seed, double *r, int n);

/* ---- gauge field (uflds/uflds.c, lattice/bcnds.c) ---- */
void oc_random_ud(oc_ctx *c);
void oc_set_ud_phase(oc_ctx *c);
double *oc_udfld(oc_ctx *c);            /* 4V su3_dble, reference layout */
void oc_set_ud(oc_ctx *c, const double *ud);
void oc_assign_ud2u(oc_ctx *c);         /* FP64 -> FP32 copy (ufld)     */
float *oc_ufld(oc_ctx *c);

/* ---- Wilson flow (wflow/wflow.c, forces/force0.c, su3fcts/chexp.c) ---- */
void oc_plaq_frc(oc_ctx *c, double *frc);            /* 4V su3_alg_dble (8 doubles) */
void oc_expXsu3(double eps, const double *X, double *u);
void oc_fwd_euler(oc_ctx *c, int n, double eps);     /* acts on the FP64 field  */
double oc_plaq_sum(const oc_ctx *c, int iw);        /* plaq_sum_dble (iw=0), plaq_wsum_dble (iw=1) */

/* ---- spinor fields (sflds/sflds.c, sflds/unorm.c, lattice/bcnds.c) ---- */
void oc_random_sd(oc_ctx *c, double *sd, double sigma);
void oc_random_s(oc_ctx *c, float *s, float sigma);
void oc_random_sd_master(oc_ctx *c, int vol, double *sd, double sigma);
void oc_random_source(oc_ctx *c, int t0, double *sd);   /* main/ms4.c:383-406 */
void oc_bnd_sd2zero(oc_ctx *c, int set, double *sd);
void oc_bnd_s2zero(oc_ctx *c, int set, float *s);
double oc_unorm_dble(int vol, const double *sd);
float oc_unorm(int vol, const float *s);

/* ---- SW term (sw_term, tcharge/ftensor.c) ---- */
/* set: 0 ALL_PTS, 1 EVEN_PTS, 2 ODD_PTS, 3 NO_PTS (flags.h ptset_t order) */
void oc_ftensor(oc_ctx *c, double *ft);          /* 6*V*9 doubles */
int oc_sw_term(oc_ctx *c, int set);
double *oc_swdfld(oc_ctx *c);                    /* 2V pauli_dble */
void oc_assign_swd2sw(oc_ctx *c);
float *oc_swfld(oc_ctx *c);
void oc_mul_pauli_dble(double mu, const double *m, const double *s, double *r);
void oc_mul_pauli(float mu, const float *m, const float *s, float *r);
int oc_inv_pauli_dble(double mu, const double *m, double *im);
void oc_apply_sw_dble(int vol, double mu, const double *m, const double *s, double *r);

/* ---- Dirac operator (dirac/Dw_dble.c, dirac/Dw.c) ---- */
void oc_Dw_dble(oc_ctx *c, double mu, double *s, double *r);
void oc_Dwee_dble(oc_ctx *c, double mu, double *s, double *r);
void oc_Dwoo_dble(oc_ctx *c, double mu, double *s, double *r);
void oc_Dwoe_dble(oc_ctx *c, double *s, double *r);
void oc_Dweo_dble(oc_ctx *c, double *s, double *r);
void oc_Dwhat_dble(oc_ctx *c, double mu, double *s, double *r);
void oc_Dw(oc_ctx *c, float mu, float *s, float *r);
void oc_Dwee(oc_ctx *c, float mu, float *s, float *r);
void oc_Dwoo(oc_ctx *c, float mu, float *s, float *r);
void oc_Dwoe(oc_ctx *c, float *s, float *r);
void oc_Dweo(oc_ctx *c, float *s, float *r);
void oc_Dwhat(oc_ctx *c, float mu, float *s, float *r);

/* ---- linear algebra (linalg/salg*.c, utils/qsum.c) ---- */
void oc_acc_qflt(double u, double *qr);
void oc_add_qflt(const double *qu, const double *qv, double *qr);
void oc_spinor_prod_dble(int vol, const double *s, const double *r, double *qre, double *qim);
void oc_norm_square_dble(int vol, const double *s, double *q);
void oc_spinor_prod(int vol, const float *s, const float *r, float *z);
float oc_norm_square(int vol, const float *s);
void oc_mulc_spinor_add_dble(int vol, double *s, const double *r, double zre, double zim);
void oc_mulc_spinor_add(int vol, float *s, const float *r, float zre, float zim);

/* ---- SAP and solvers (sap, linsolv/fgcr.c, block) ---- */
int oc_sap_nblocks(oc_ctx *c);
void oc_sap_block_info(oc_ctx *c, int n, int *bo, int *vol);
void oc_sap_block_imb(oc_ctx *c, int n, int *imb);   /* block index -> global index */
int oc_sap_prepare(oc_ctx *c);                       /* FP32 links/sw on block grid */
void oc_Dw_blk(oc_ctx *c, int n, float mu, const float *s, float *r); /* block-local fields, vol+1 */
void oc_Dwhat_blk(oc_ctx *c, int n, float mu, const float *s, float *r);
void oc_blk_eo_op(oc_ctx *c, int n, int op, float mu, float *s, float *r); /* 0 Dwee 1 Dwoo 2 Dwoe 3 Dweo _blk */
void oc_blk_mres(oc_ctx *c, int n, float mu, int nmr, float *s0, float *s1);
void oc_blk_eo_mres(oc_ctx *c, int n, float mu, int nmr, float *s0, float *s1);
/* FP64 block operators (Dw_dble.c:1891-2335) and the FP64 Schwarz sweep on
   the FP64 links and the FP64 block SW field */
void oc_Dw_blk_dble(oc_ctx *c, int n, double mu, const double *s, double *r);
void oc_Dwhat_blk_dble(oc_ctx *c, int n, double mu, const double *s, double *r);
void oc_blk_eo_op_dble(oc_ctx *c, int n, int op, double mu, double *s, double *r);
void oc_blk_mres_dble(oc_ctx *c, int n, double mu, int nmr, double *s0, double *s1);
void oc_blk_eo_mres_dble(oc_ctx *c, int n, double mu, int nmr, double *s0, double *s1);
void oc_sap_dble(oc_ctx *c, double mu, int isolv, int nmr, double *psi, double *rho);
void oc_sap(oc_ctx *c, float mu, int isolv, int nmr, float *psi, float *rho);
double oc_sap_gcr(oc_ctx *c, int nkv, int nmx, int istop, double res, double mu,
                  double *eta, double *psi, int *ifail, int *status);

#ifdef __cplusplus
}
#endif
#endif
