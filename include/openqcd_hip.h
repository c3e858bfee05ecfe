/*
 * openqcd_hip.h -- C ABI of libopenqcd_hip.so, an MI355X (gfx950) build of
 * openQCD-2.4.2's Wilson-clover Dirac operator, SW term and SAP+GCR solver.
 *
 * The functions in part B keep the names, argument meaning and error
 * behaviour of the reference routines they replace (file:line of the
 * reference prototype given with each group).  Global state (lattice
 * geometry, gauge field, SW field, parameters) is held by the library, as in
 * the reference; part A sets it, replacing the compile-time geometry of
 * include/global.h and the parameter data base of modules/flags.
 *
 * Field pointers.  Spinor fields passed to part B are either
 *   (1) device fields obtained from oqh_alloc_s/oqh_alloc_sd or the workspace
 *       calls alloc_ws/reserve_ws (alloc_wsd/reserve_wsd); pointer arithmetic
 *       in units of spinor(_dble) (e.g. psd[0]+VOLUME/2) addresses the sites
 *       in the reference's order, exactly as in host memory; or
 *   (2) ordinary host arrays in the reference layout (NSPIN elements, site
 *       order ipt); these are copied to the device and back around the call.
 * Device fields are stored as structure-of-arrays in 64-site blocks; their
 * contents are read and written with oqh_import_* / oqh_export_*.
 *
 * Errors (argument, state, allocation) print a message and abort the
 * process, as error() does in the reference (utils/error.c), unless
 * oqh_set_error_mode(0) was called; then the call returns and
 * oqh_last_error() reports the failure.
 */
#ifndef OPENQCD_HIP_H
#define OPENQCD_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- data types, layout-identical to include/su3.h:33-121 ---- */
#ifndef SU3_H
typedef struct { float re, im; } complex;
typedef struct { complex c1, c2, c3; } su3_vector;
typedef struct { complex c11, c12, c13, c21, c22, c23, c31, c32, c33; } su3;
typedef struct { su3_vector c1, c2; } weyl;
typedef struct { su3_vector c1, c2, c3, c4; } spinor;
typedef struct { float u[36]; } pauli;
typedef struct { double re, im; } complex_dble;
typedef struct { complex_dble c1, c2, c3; } su3_vector_dble;
typedef struct { complex_dble c11, c12, c13, c21, c22, c23, c31, c32, c33; } su3_dble;
typedef struct { su3_vector_dble c1, c2; } weyl_dble;
typedef struct { su3_vector_dble c1, c2, c3, c4; } spinor_dble;
typedef struct { double u[36]; } pauli_dble;
typedef struct { double q[2]; } qflt;
typedef struct { qflt re, im; } complex_qflt;
#endif
#ifndef UTILS_H
/* include/utils.h:65-68 */
typedef enum { ALL_PTS, EVEN_PTS, ODD_PTS, NO_PTS, PT_SETS } ptset_t;
#endif
#ifndef BLOCK_H
/* include/block.h:48-52 */
typedef enum { SAP_BLOCKS, DFL_BLOCKS, TEST_BLOCKS, BLK_GRIDS } blk_grid_t;
#endif
#ifndef FLAGS_H
/* include/flags.h:102-118 */
typedef struct { int bs[4]; int Ns; } dfl_parms_t;
typedef struct { int nkv, nmx, nmx_gcr; double res, res_gcr; } dfl_pro_parms_t;
typedef struct { int ninv, nmr, ncy; double kappa, m0, mu; } dfl_gen_parms_t;
#endif

/* ======================= A. library state (no reference counterpart
   as functions; they replace include/global.h and modules/flags) ====== */

/* select the HIP device and create the library stream */
int oqh_init(int device);
void oqh_finalize(void);
void oqh_set_error_mode(int abort_on_error);
int oqh_last_error(char *msg, int len);   /* returns error count since last call */

/* lattice and thread-block sizes of include/global.h:18-36 (NPROC=1);
   builds the site order ipt of lattice/geometry.c */
int oqh_geometry(const int *L, const int *Ltrd);
int oqh_volume(void);
void oqh_get_ipt(int *ipt);

/* parameter data base, flags/lat_parms.c:516 (set_bc_parms: type, cF, cF'),
   lat_parms.c:920 (set_sw_parms m0, with csw of set_lat_parms),
   lat_parms.c:963 (set_tm_parms), flags/sap_parms.c:112 (set_sap_parms) */
void oqh_set_bc_parms(int type, double cF, double cF_prime);
void oqh_set_sw_parms(double m0, double csw);
void oqh_set_tm_parms(int eoflg);
void oqh_set_sap_parms(const int *bs, int isolv, int nmr, int ncy);
/* flags/sap_parms.c:161 (sap_parms): the current SAP parameters (ncy = 0:
   not set) */
void oqh_get_sap_parms(int *bs, int *isolv, int *nmr, int *ncy);
/* storage precision of the SAP block solver (sap, sap_gcr; no counterpart in
   the reference): 0 = FP32 fields as in sap/sap.c (default), 1 = FP16 links,
   SW matrices and spinor fields (psi, rho and the block corrections, with a
   per-site scale), FP32 arithmetic, for blocks of 256 points (other block
   sizes use FP32).  With 1, sap_gcr runs its ncy Schwarz cycles on FP16
   fields and computes chi = Dw(phi) in FP32; the outer FGCR keeps its
   precision. */
void oqh_set_sap_prec(int prec);
/* the ncy Schwarz cycles of the FP16 preconditioner of sap_gcr are captured
   as a HIP graph when the same arguments come twice in a row (within a
   solve) and replayed for every further Krylov vector (1, default), or
   launched kernel by kernel (0); no counterpart in the reference, the
   results are the same */
void oqh_set_sap_graph(int on);
/* SAP block-size autotuner (no counterpart in the reference): among the
   legal block sizes of sap_parms.c (even, >= 4, an even number of blocks per
   axis; here with equal spatial sizes and at most maxvol points, 4096 if
   maxvol <= 0) picks the one with the least time per e-fold of residual
   reduction in ncy Schwarz cycles (the given isolv and nmr, the current
   precision variant) and returns that time in seconds (bs_out: the block
   size); 0 on failure.  On success only the block size of the SAP
   parameters changes: isolv, nmr and ncy keep the values set before (if no
   SAP parameters were set, they become the given ones).  On failure the
   SAP parameters are left as they were. */
double oqh_sap_autotune(int isolv, int nmr, int ncy, int maxvol, int *bs_out);
/* several right-hand sides (no counterpart in the reference): solves
   D psi[s] = eta[s] for s < n as sap_gcr would (per source: rho_out[s],
   ifail[s], status[s] as its return value and arguments); with FP32 SAP on
   256-point blocks and istop = 0 the FGCR iterations of all sources advance
   together and the Schwarz cycles of all active sources (up to 16 per
   launch) run in the same kernel launches, consecutive work-groups taking
   the same block for different sources; otherwise the sources are solved one
   after the other.  Each source's result equals that of sap_gcr. */
/* the propagator loop of main/ms4.c (propagator, ms4.c:496-531, with
   random_source 383-406, solve_dirac's SAP_GCR branch 409-459 and
   save_prop's export_sfld branch 470-481) for the current gauge field:
   nsrc random sources on time slice x0, solved by SAP+GCR with the current
   SAP parameters (up to 16 together, sap_gcr_multi), written as
   "<dir>/<nbase>n<icnfg>.s<isrc>".  status[isrc]: Krylov vectors of source
   isrc.  Returns the number of failed solves (their files are not
   written), -1 for bad arguments. */
int oqh_ms4_propagators(int x0, int nsrc, int nkv, int nmx, int istop, double res, double mu,
                        const char *dir, const char *nbase, int icnfg, int *status);
void sap_gcr_multi(int n, int nkv, int nmx, int istop, double res, double mu, spinor_dble **eta,
                   spinor_dble **psi, double *rho_out, int *ifail, int *status);
/* orthogonalisation of the Krylov vectors in fgcr on device fields (no
   counterpart in the reference): 1 = modified Gram-Schmidt in its inverse
   compact WY form, one pass for all scalar products (with the Gram entries
   of the newest vector) and one for all updates (default); 0 = one pass per
   term in the order of linsolv/fgcr.c:157-178.  Both give the coefficients
   of the modified process in exact arithmetic. */
void oqh_set_gcr_orth(int mode);

/* random numbers and random fields on the device: the streams of
   random/ranlux.c:136-171 (ranlxd, one per reference thread of the
   geometry's thread layout), random_ud (uflds/uflds.c:159-222, uflds.h),
   random_sd (sflds/sflds.c:1254-1268, sflds.h: with icom & 2 every
   reference thread k fills vol points from sd + k vol, e.g.
   random_sd(VOLUME_TRD, 2, sd, sigma) for a full field; otherwise the
   calling thread, stream 0, fills vol points from sd), ranlxd
   (random/ranlxd.c; the master thread's stream, n numbers into the host
   array r).  The fields are the reference's for the same seed and layout. */
/* lattice/bcnds.c:750, 709: zero a field on the points of the set at the
   boundary time slices of the boundary conditions (time 0 and N0-1 for open
   bc; nothing for periodic bc) */
void bnd_sd2zero(ptset_t set, spinor_dble *sd);
void bnd_s2zero(ptset_t set, spinor *s);
void start_ranlux(int level, int seed);
void random_ud(void);
void random_sd(int vol, int icom, spinor_dble *sd, double sigma);
/* the random source of main/ms4.c:383-406 (random_source): Gaussian spinors
   (random_sd(1,0,..) by each reference thread in turn over its VOLUME_TRD
   ipt indices) on the points of time slice t0, zero elsewhere; full field */
void oqh_random_source(int t0, spinor_dble *sd);
void ranlxd(double *r, int n);
/* one start field of dfl_modes (replaces the ranlxs loop of
   dfl/dfl_modes.c:248-271, random_fields): component 4 j + i of point x
   (ipt index) is word i of Philox4x32-10 (Salmon et al., SC11 2011) at
   counter (x, j, 0, 0), j < 6, as (w >> 8) 2^-24 - 1/2, keyed by two ranlxd
   numbers of stream 0 times 2^32; exported as test access */
void oqh_random_start_field(spinor *s);

/* learned preconditioner (no counterpart in the reference): a linear
   gauge-equivariant network of parallel-transport layers,
      out(x)[o] = sum_{p=0..8} sum_i W_l[p][o][i] (T_p in)(x)[i],
   T_0 = 1, T_{2mu+1} in(x) = U(x,mu) in(x+mu), T_{2mu+2} in(x) =
   U(x-mu,mu)^+ in(x-mu), with C_l colour-vector channels per point (a spinor
   has its 4 spin components as channels).  Channel mixing runs as BF16
   matrix-core GEMMs with FP32 accumulation.
   oqh_nn_create: nlayer layers, widths[0..nlayer] in {4,8,12,16} with
   widths[0] = widths[nlayer] = 4; returns 0 on success.  Weights are
   complex, per layer [9][C_out][C_in], layers in order, as (re, im)
   doubles (2 * oqh_nn_nweights() of them).  oqh_nn_apply: out = NN(in)
   (single-precision full fields, in != out).  oqh_nn_correct: phi +=
   NN(r) as the preconditioner of sap_gcr applies it, with the activations
   between the layers rounded to BF16 (device fields).  oqh_nn_grad: mean over the n
   pairs of the loss |NN(r_k) - e_k|^2 / |e_k|^2, its gradient with respect
   to the weights (dL/dRe w + i dL/dIm w) in dw.  oqh_nn_train_step: one Adam
   step (beta 0.9, 0.999) of step size lr on that loss; returns the loss
   before the step.  oqh_set_nn_prec(1): sap_gcr adds NN(residual of the
   Schwarz cycles) to the preconditioner's output and takes chi = Dw(phi);
   plain SAP is the fallback: the learned solve may use nmx/2 Krylov
   vectors (at least 1), and if it has not converged by then, plain SAP+GCR
   continues from its solution (from psi = 0 if that is no better than the
   start) with the remaining nmx - nmx/2 vectors to the same tolerance.  A
   call therefore never uses more than nmx vectors, status counts both
   parts, and ifail = -1 only if the continuation fails as well. */
int oqh_nn_create(int nlayer, const int *widths);
int oqh_nn_nweights(void);
void oqh_nn_set_weights(const double *w);
void oqh_nn_get_weights(double *w);
void oqh_nn_apply(spinor *in, spinor *out);
void oqh_nn_correct(spinor *r, spinor *phi);
double oqh_nn_grad(int n, spinor **r, spinor **e, double *dw);
double oqh_nn_train_step(int n, spinor **r, spinor **e, double lr);
void oqh_set_nn_prec(int on);
void oqh_nn_free(void);

/* gauge field in the layout of udfld() (uflds/uflds.c:113-156; 8 links per
   odd point, README.global).  Import invalidates the SW term and the
   single-precision copy; the imported field is taken as given, with the
   library's phase flag of set_ud_phase() cleared. */
void oqh_import_ud(const su3_dble *ud);
void oqh_export_ud(su3_dble *ud);
/* no counterpart: the boundary field l of face ifc of block n after Dw_bnd
   (bb[ifc].w[l] of the reference, block/block.c; vol/bs[ifc/2] Weyl
   spinors in the point order of lattice/blk_geometry.c:241-313) */
void oqh_export_bnd(int n, int ifc, int l, weyl *w);
/* SW field in the layout of swdfld() (2 pauli_dble per point) */
void oqh_export_swd(pauli_dble *swd);
void oqh_import_swd(const pauli_dble *swd);

/* device-resident spinor fields (n sites each, n = volume if n <= 0) */
spinor *oqh_alloc_s(int n);
spinor_dble *oqh_alloc_sd(int n);
void oqh_free(void *field);
void oqh_import_s(spinor *dev, const spinor *host, int n);
void oqh_export_s(const spinor *dev, spinor *host, int n);
void oqh_import_sd(spinor_dble *dev, const spinor_dble *host, int n);
void oqh_export_sd(const spinor_dble *dev, spinor_dble *host, int n);
void oqh_sync(void);

/* No counterpart in the reference: for callers that cannot allocate device
   memory (code under HIP graph capture, tight loops, harnesses that forbid
   hipMalloc).  spinor_prod_dble(vol, icom, s, r) of two double-precision
   device fields of the library (alloc_wsd/reserve_wsd, oqh_alloc_sd; host
   arrays are refused), written as complex_qflt to the device memory z and
   enqueued on the library stream without synchronisation (oqh_sync waits).
   Same kernels and summation order as spinor_prod_dble, so the result is
   bitwise the same.  The reduction's partial sums use a fixed 32 KiB device
   array per GPU (at most 1024 work-group partials, enough for every vol);
   calls that overlap on one GPU share it and race.  Allocates nothing
   (no hipMalloc, hipFree or pinned host memory).  Returns 0, or 1 after
   reporting an error (library not set up, current device not the
   library's, host or too short fields, z not device memory). */
int oqh_spinor_prod_dble_noalloc(int vol, int icom, spinor_dble *s, spinor_dble *r, complex_qflt *z);

/* the operators read rows 0 and 1 of every FP64 link and rebuild row 2 =
   +-conj(row0 x row1) when every stored row 2 deviates from that by less
   than 64 ulp (1.4e-14); otherwise they read all nine entries.  Returns 1
   if the current gauge field is rebuilt, and the largest deviation in *dev
   (no counterpart in the reference). */
int oqh_link_rebuild(double *dev);

/* timing of ncall applications of an operator on device fields, measured
   with HIP events on the library stream; op: 0 Dw_dble, 1 Dw, 2 Dwhat_dble,
   3 Dwhat.  Returns the average milliseconds per call. */
double oqh_time_dirac(int op, double mu, void *s, void *r, int ncall);
/* Dw_dble / Dw applied to n fields at once, r[k] = D s[k] (no counterpart in
   the reference: a batched form of dirac/Dw_dble.c:1377 and dirac/Dw.c:973
   that loads every link and Pauli matrix once for up to 8 fields).  The
   fields must be full device fields of the library and s[k] != r[k]; unlike
   Dw_dble, the inputs are left unchanged on the boundary points of open bc. */
void oqh_Dw_dble_multi(double mu, int n, spinor_dble **s, spinor_dble **r);
void oqh_Dw_multi(float mu, int n, spinor **s, spinor **r);
/* average milliseconds of ncall batched applications (prec 1: FP64, 0: FP32) */
double oqh_time_dirac_multi(int prec, double mu, int n, void **s, void **r, int ncall);
/* bandwidth (GB/s, bytes read + written) of a plain device copy of gbytes
   in total, averaged over nrep launches */
double oqh_copy_bandwidth(double gbytes, int nrep);

/* force field of the last plaq_frc() (4V su3_alg_dble = 8 doubles per
   link, layout of udfld(); the MD force field mdflds()->frc of the
   reference, mdflds/mdflds.c:269) */
void oqh_export_frc(double *frc);

/* SAP block grid access for tests: block-local field k (0..2) of block n in
   the block's own point order (vol spinors), and the embedding of the block
   points into the lattice (b.imb of block/block.c) */
void oqh_export_sblk(int n, int k, spinor *host);
void oqh_get_imb(int n, int *imb);

/* ======================= B. reference entry points ====================== */

/* include/dirac.h:28-53 (dirac/Dw.c, dirac/Dw_dble.c) */
void Dw(float mu, spinor *s, spinor *r);
void Dwee(float mu, spinor *s, spinor *r);
void Dwoo(float mu, spinor *s, spinor *r);
void Dweo(spinor *s, spinor *r);
void Dwoe(spinor *s, spinor *r);
void Dwhat(float mu, spinor *s, spinor *r);
void Dw_dble(double mu, spinor_dble *s, spinor_dble *r);
void Dwee_dble(double mu, spinor_dble *s, spinor_dble *r);
void Dwoo_dble(double mu, spinor_dble *s, spinor_dble *r);
void Dweo_dble(spinor_dble *s, spinor_dble *r);
void Dwoe_dble(spinor_dble *s, spinor_dble *r);
void Dwhat_dble(double mu, spinor_dble *s, spinor_dble *r);

/* include/sw_term.h:77-82 (sw_term/sw_term.c:399, swflds.c:129) */
int sw_term(ptset_t set);
void assign_swd2sw(void);
/* include/sw_term.h:75-76 (sw_term/swflds.c:79-126): host copies of the
   device SW fields (2V matrices, the reference's layout) taken at the call:
   the current term (with the parts sw_term() inverted), or unit matrices
   before any SW term; writes to them do not reach the device */
pauli *swfld(void);
pauli_dble *swdfld(void);
/* include/sw_term.h:43,52 (sw_term/pauli.c:1107-1119,
   pauli_dble.c:1840-1857): r = (A + i mu gamma5) s on vol points with the
   matrices m[2k], m[2k+1] (any host array of that layout); s and r are
   device fields or host arrays and may coincide */
void apply_sw(int vol, float mu, pauli *m, spinor *s, spinor *r);
void apply_sw_dble(int vol, double mu, pauli_dble *m, spinor_dble *s, spinor_dble *r);
/* include/uflds.h:59 (uflds/uflds.c:479) */
void assign_ud2u(void);

/* include/uflds.h:56-57 (uflds/uflds.c:338-400): sign of the time-like
   links at x0 = N0-1 for periodic bc (theta = 0) */
void set_ud_phase(void);
void unset_ud_phase(void);
/* include/forces.h:27 (forces/force0.c:281): force of the Wilson plaquette
   action, into the library's force field (oqh_export_frc) */
void plaq_frc(void);
/* include/wflow.h:16 (wflow/wflow.c:217): n forward-Euler steps of the
   Wilson flow with step eps on the double-precision gauge field */
void fwd_euler(int n, double eps);

/* include/uflds.h:33-34 (uflds/plaq_sum.c:155-191): sum of Re tr U(p) over
   all unoriented plaquettes of the double-precision field; plaq_wsum_dble
   gives the space-like plaquettes on the boundary slices of open bc weight
   1/2.  Double-double sums on the device; icom has no effect (one process
   holds the whole lattice). */
double plaq_sum_dble(int icom);
double plaq_wsum_dble(int icom);

/* include/archive.h:34-36 (archive/archive.c:769-977, main/README.io):
   gauge configurations in the universal little-endian format (lattice
   sizes, average plaquette, then the 8 links of each odd point in
   lexicographic order).  export_cnfg refuses a phase-set field; import_cnfg
   requires the sizes on the file to be the current ones (the whole lattice
   is local), checks the boundary links and the average plaquette, and
   leaves the field without phase. */
void export_cnfg(char *out);
void import_cnfg(char *in, int mask);
void lat_sizes(char *in, int *nl);
/* include/archive.h:64-65 (archive/sarchive.c:424-580): a double-precision
   quark field in the universal format (sizes, eo, square norm, then the
   spinors of all points (eo=0) or of the even points (eo=1) in
   lexicographic order).  sd is a device field of the library or a host
   array; import_sfld checks the sizes, eo and the square norm and, with
   eo=1, leaves the odd points of sd unchanged. */
void export_sfld(char *out, int eo, spinor_dble *sd);
void import_sfld(char *in, int eo, spinor_dble *sd);

/* include/linalg.h:70-79 (linalg/salg.c:1223-1429) */
complex spinor_prod(int vol, int icom, spinor *s, spinor *r);
float spinor_prod_re(int vol, int icom, spinor *s, spinor *r);
float norm_square(int vol, int icom, spinor *s);
void mulc_spinor_add(int vol, int icom, spinor *s, spinor *r, complex z);
void mulr_spinor_add(int vol, int icom, spinor *s, spinor *r, float c);
void scale(int vol, int icom, float c, spinor *s);
void project(int vol, int icom, spinor *s, spinor *r);
float normalize(int vol, int icom, spinor *s);
void mulg5(int vol, int icom, spinor *s);
void mulmg5(int vol, int icom, spinor *s);

/* include/linalg.h:82-99 (linalg/salg_dble.c:2611-2891) */
complex_qflt spinor_prod_dble(int vol, int icom, spinor_dble *s, spinor_dble *r);
qflt spinor_prod_re_dble(int vol, int icom, spinor_dble *s, spinor_dble *r);
complex_qflt spinor_prod5_dble(int vol, int icom, spinor_dble *s, spinor_dble *r);
qflt norm_square_dble(int vol, int icom, spinor_dble *s);
void mulc_spinor_add_dble(int vol, int icom, spinor_dble *s, spinor_dble *r, complex_dble z);
void mulr_spinor_add_dble(int vol, int icom, spinor_dble *s, spinor_dble *r, double c);
void combine_spinor_dble(int vol, int icom, spinor_dble *s, spinor_dble *r, double cs, double cr);
void scale_dble(int vol, int icom, double c, spinor_dble *s);
void project_dble(int vol, int icom, spinor_dble *s, spinor_dble *r);
double normalize_dble(int vol, int icom, spinor_dble *s);
void mulg5_dble(int vol, int icom, spinor_dble *s);
void mulmg5_dble(int vol, int icom, spinor_dble *s);

/* include/utils.h:122-126 (utils/qsum.c:262-424); host arithmetic */
void acc_qflt(double u, double *qr);
void add_qflt(double *qu, double *qv, double *qr);
void scl_qflt(double u, double *qr);
void mul_qflt(double *qu, double *qv, double *qr);
void global_qsum(int n, double **qu, double **qr);

/* include/sflds.h:37-61 (sflds/sflds.c:1203-1405, sflds/unorm.c:261-314) */
void set_s2zero(int vol, int icom, spinor *s);
void set_sd2zero(int vol, int icom, spinor_dble *sd);
void assign_s2s(int vol, int icom, spinor *s, spinor *r);
void assign_s2sd(int vol, int icom, spinor *s, spinor_dble *rd);
void assign_sd2s(int vol, int icom, spinor_dble *sd, spinor *r);
void assign_sd2sd(int vol, int icom, spinor_dble *sd, spinor_dble *rd);
void add_s2s(int vol, int icom, spinor *s, spinor *r);
void add_s2sd(int vol, int icom, spinor *s, spinor_dble *rd);
void diff_s2s(int vol, int icom, spinor *s, spinor *r);
void diff_sd2s(int vol, int icom, spinor_dble *sd, spinor_dble *rd, spinor *r);
float unorm(int vol, int icom, spinor *s);
double unorm_dble(int vol, int icom, spinor_dble *sd);

/* include/utils.h:143-152 (utils/wspace.c:394-510): workspace fields are
   device fields */
void alloc_ws(int n);
spinor **reserve_ws(int n);
int release_ws(void);
void alloc_wsd(int n);
spinor_dble **reserve_wsd(int n);
int release_wsd(void);

/* include/block.h:62-75 (block/blk_grid.c:139, map_u2blk.c:308,
   map_sw2blk.c:102, map_s2blk.c:551-737), SAP_BLOCKS only */
void alloc_bgr(blk_grid_t grid);
void assign_ud2ubgr(blk_grid_t grid);
int assign_swd2swbgr(blk_grid_t grid, ptset_t set);
void assign_s2sblk(blk_grid_t grid, int n, ptset_t set, spinor *s, int k);
void assign_sblk2s(blk_grid_t grid, int n, ptset_t set, int k, spinor *s);
/* include/dirac.h:34-39 (dirac/Dw.c:1475-1907); the even/odd pieces accept
   k == l */
void Dw_blk(blk_grid_t grid, int n, float mu, int k, int l);
void Dwee_blk(blk_grid_t grid, int n, float mu, int k, int l);
void Dwoo_blk(blk_grid_t grid, int n, float mu, int k, int l);
void Dwoe_blk(blk_grid_t grid, int n, int k, int l);
void Dweo_blk(blk_grid_t grid, int n, int k, int l);
void Dwhat_blk(blk_grid_t grid, int n, float mu, int k, int l);
/* include/block.h:77-80 and include/dirac.h:48-53 (block/map_s2blk.c,
   dirac/Dw_dble.c:1891-2290): double-precision block fields sd[0..2] of the
   library's block grid and the block operators on them, with the FP64
   links and the FP64 block SW term (inverted on the point set of
   assign_swd2swbgr / the SAP solver, as the single-precision one) */
void assign_sd2sdblk(blk_grid_t grid, int n, ptset_t set, spinor_dble *sd, int k);
void assign_sdblk2sd(blk_grid_t grid, int n, ptset_t set, int k, spinor_dble *sd);
void Dw_blk_dble(blk_grid_t grid, int n, double mu, int k, int l);
void Dwee_blk_dble(blk_grid_t grid, int n, double mu, int k, int l);
void Dwoo_blk_dble(blk_grid_t grid, int n, double mu, int k, int l);
void Dwoe_blk_dble(blk_grid_t grid, int n, int k, int l);
void Dweo_blk_dble(blk_grid_t grid, int n, int k, int l);
void Dwhat_blk_dble(blk_grid_t grid, int n, double mu, int k, int l);
/* test access: FP64 block field k of block n in block order */
void oqh_export_sdblk(int n, int k, spinor_dble *host);
/* include/dirac.h:25 (dirac/Dw_bnd.c:801-1011): the Weyl fields of the
   hopping terms from block n to its exterior, face by face, into the
   block's boundary field l (SAP_BLOCKS has one, l = 0); read them with
   oqh_export_bnd */
void Dw_bnd(blk_grid_t grid, int n, int k, int l);

/* include/sap.h:21-35 (sap/blk_solv.c:1117-1186, sap/sap.c:720,
   sap/sap_gcr.c:140) */
void blk_mres(int n, float mu, int nmr);
void blk_eo_mres(int n, float mu, int nmr);
void sap(float mu, int isolv, int nmr, spinor *psi, spinor *eta);
double sap_gcr(int nkv, int nmx, int istop, double res, double mu,
               spinor_dble *eta, spinor_dble *psi, int *ifail, int *status);

/* include/linsolv.h:34-37 (linsolv/fgcr.c:215).  Dop and Mop are host
   callbacks that receive the library's device fields (the workspace ws/wsd
   must be device fields). */
double fgcr(int vol, void (*Dop)(spinor_dble *s, spinor_dble *r),
            void (*Mop)(int k, spinor *rho, spinor *phi, spinor *chi),
            spinor **ws, spinor_dble **wsd, int nkv, int nmx, int istop,
            double res, spinor_dble *eta, spinor_dble *psi, int *status);

/* Deflation (Luescher's local deflation; include/flags.h:250-260,
   include/dfl.h:34-55, include/little.h:48-82).  Parameters as
   flags/dfl_parms.c:209, 277, 351 (set once; the DFL block grid follows the
   rules of check_block_size).  dfl_subspace (dfl/dfl_subspace.c): block
   Gram-Schmidt of the Ns modes into the basis; dfl_s2vd / dfl_vd2s: a spinor
   field's components along the subspace and back (little vectors: Ns
   complex numbers per DFL block, even blocks first, host arrays of
   nb * Ns complex_dble); dfl_renormalize_modes: Gram-Schmidt of the global
   modes.  set_Aw / set_Awhat (little/Aw_ops.c): the little Dirac operator
   and its even-odd form (set_Awhat returns 1 if an inversion of a diagonal
   block was not safe, Frobenius condition >= 1e6); Awhat_dble: the even-odd
   operator on host arrays of nb/2 * Ns entries.  ltl_gcr (dfl/ltl_gcr.c):
   the little Dirac equation, FGCR in double precision preconditioned by
   single-pass GCR in single precision, both deflated by the little modes
   (little/ltl_modes.c, set by set_Awhat).  dfl_modes / dfl_modes2
   (dfl/dfl_modes.c:423, 583): the subspace by inverse iteration from random
   fields; dfl_update / dfl_update2 (dfl_modes.c:504, 604): nsm further
   deflated steps on the modes restored by dfl_restore_modes
   (dfl_subspace.c:279).  dfl_sap_gcr /
   dfl_sap_gcr2 (dfl/dfl_sap_gcr.c): the deflated SAP+GCR solver, status[0]
   the Krylov count, status[1], status[2] the average outer and inner counts
   of the little solver (dfl_sap_gcr2: ifail[2], status[6]).  With
   oqh_set_sap_prec(1) the Schwarz cycles of the preconditioner run on FP16
   fields. */
dfl_parms_t set_dfl_parms(int *bs, int Ns);
dfl_parms_t dfl_parms(void);
dfl_pro_parms_t set_dfl_pro_parms(int nkv, int nmx, double res, int nmx_gcr, double res_gcr);
dfl_pro_parms_t dfl_pro_parms(void);
dfl_gen_parms_t set_dfl_gen_parms(double kappa, double mu, int ninv, int nmr, int ncy);
dfl_gen_parms_t dfl_gen_parms(void);
void dfl_subspace(spinor **mds);
void dfl_s2vd(spinor *s, complex_dble *vd);
void dfl_vd2s(complex_dble *vd, spinor *s);
void dfl_renormalize_modes(spinor **mds);
void set_Aw(double mu);
int set_Awhat(double mu);
void Awhat_dble(complex_dble *v, complex_dble *w);
double ltl_gcr(double mu, complex_dble *eta, complex_dble *psi, int *ifail, int *status);
void dfl_modes(int *ifail, int *status);
void dfl_modes2(int *ifail, int *status);
void dfl_restore_modes(spinor **mds);
void dfl_update(int nsm, int *ifail, int *status);
void dfl_update2(int nsm, int *ifail, int *status);
double dfl_sap_gcr(int nkv, int nmx, int istop, double res, double mu,
                   spinor_dble *eta, spinor_dble *psi, int *ifail, int *status);
double dfl_sap_gcr2(int nkv, int nmx, int istop, double res, double mu,
                    spinor_dble *eta, spinor_dble *psi, int *ifail, int *status);
/* test access (no reference counterpart): basis vector k as a host array,
   the block structure (inn[nb][8] with ifc = 2 mu for -mu and 2 mu + 1 for
   +mu, the points of every block bpts[nb][bvol]; returns nb), the little
   operator (Ablk[nb][Ns][Ns], Ahop[nb][8][Ns][Ns], complex, row-major) and
   the release of the deflation state with its parameters (also done by
   oqh_finalize) */
void oqh_dfl_export_basis(int k, spinor *host);
int oqh_dfl_blocks(int *inn, int *bpts);
void oqh_dfl_export_Aw(complex_dble *ablk, complex_dble *ahop);
void oqh_dfl_release(void);

#ifdef __cplusplus
}
#endif
#endif
