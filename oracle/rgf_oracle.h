/* CPU oracle for the recursive Green's function (RGF) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP engine
 * (liteqt_amd/csrc) is compared against, and the CPU baseline that bench.py
 * times.  Product code never links, loads or calls it.
 *
 * It restates, in plain C, the algorithm of the reference's
 *   green_RGF_RS  src/green_rgf.f90:3950-4164   (one energy, full RGF)
 *   sancho        src/green_rgf.f90:4451-4537   (Sancho-Rubio lead decimation)
 *   invert        src/green_rgf.f90:4421-4441   (zgetrf + zgetri, zero on failure)
 *   ferm          src/green_rgf.f90:4444-4447
 * The LAPACK routines zgetrf/zgetri (an absent third-party dependency, Intel
 * MKL 2020.0.166 per Makefile:7) are restated from LAPACK's published
 * unblocked algorithms (zgetf2, ztrti2, unblocked zgetri).
 *
 * Arrays: complex128 stored as interleaved (re, im) doubles, n x n blocks in
 * Fortran (column-major) order, blocks contiguous along the slab index, i.e.
 * the memory layout of a Fortran complex(8) array (nm, nm, nx).
 *
 * Parity pins (tests/): orc_zinv against scipy LAPACK zgetrf/zgetri and the
 * residual |A A^-1 - I|; orc_sancho against the closed-form surface Green's
 * function of a 1D chain, the Dyson fixed-point residual and a dense
 * truncated-chain inverse; orc_rgf (physical mode, and compat mode with uniform
 * couplings) against a dense inverse of the whole open device; compat-only
 * deviations against closed-form relations to dense quantities.  The compat
 * backward sweep with NON-uniform couplings (reference quirk: H1i(:,:,l) used
 * as the l,l+1 coupling) has no independent pin: parity unpinned for that
 * combination only.
 */
#ifndef RGF_ORACLE_H
#define RGF_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_MODE_COMPAT = 0, ORC_MODE_PHYSICAL = 1 };

typedef struct orc_params {
  double eta_dev;   /* device broadening; reference: 0 (z=dcmplx(E,0.0d-6), green_rgf.f90:3974) */
  double eta_lead;  /* decimation broadening; reference: 1e-3 (green_rgf.f90:4474) */
  double dec_tol;   /* decimation tolerance; reference: 1e-100 (green_rgf.f90:4457) */
  int dec_nmax;     /* decimation iteration cap; reference: 100 (green_rgf.f90:4473) */
  double boltz;     /* eV/K; reference: 8.61734e-5 (green_rgf.f90:3968) */
  int mode;         /* ORC_MODE_COMPAT or ORC_MODE_PHYSICAL */
} orc_params;

/* Per-energy diagnostics. */
typedef struct orc_diag {
  int iters_left, iters_right;   /* decimation iterations run per lead */
  int nmax_hit;                  /* bit0 left lead, bit1 right lead hit dec_nmax */
  int inv_failures;              /* number of singular inversions (zeroed, as the reference) */
} orc_diag;

void orc_params_reference(orc_params *p);

/* A <- A^-1 (LAPACK zgetrf + zgetri semantics).  On a zero pivot A is zeroed.
 * Returns LAPACK info (0 on success). */
int orc_zinv(int n, double *A);

/* C = op(A) op(B), op in {'n','c'} (no transpose / conjugate transpose). */
void orc_zgemm(char ta, char tb, int n, const double *A, const double *B, double *C);

/* Sancho-Rubio decimation of green_rgf.f90:4451-4537 (S00 = I).
 * Returns the number of iterations run; *nmax_hit set when the cap was reached
 * (then the reference falls back to the bare H00 surface). */
int orc_sancho(int n, double E, double eta_lead, double tol, int nmax,
               const double *H00, const double *H10, double *G00, double *GBB,
               int *nmax_hit, int *inv_failures);

double orc_ferm(double a);

/* One energy.  Argument order and meaning follow green_RGF_RS
 * (green_rgf.f90:3950-3958).  sigL/sigG/sigR may be NULL (zero).  GRi1, GLi1,
 * GGi1 may be NULL (absent optional arguments).  diag may be NULL. */
void orc_rgf(const orc_params *p, const double temp[2], int nm, int nx, double E,
             const double mu[2], const double *Hii, const double *H1i,
             const double *sigL, const double *sigG, const double *sigR,
             double *ndens, double *pdens, double *ldos, double *tr, double *tre,
             double *cur, double *GRi1, double *GLi1, double *GGi1, orc_diag *diag);

/* orc_rgf with the lead decimations fed by explicit onsite blocks (the
 * reference feeds them Hii(:,:,1)+sigR(:,:,1) and Hii(:,:,nx)+sigR(:,:,nx),
 * green_rgf.f90:3992, 4032); used to differentiate the device with the
 * contacts held fixed. */
void orc_rgf_fixed_leads(const orc_params *p, const double temp[2], int nm, int nx, double E,
                         const double mu[2], const double *Hii, const double *H1i,
                         const double *sigL, const double *sigR, const double *leadL_H00,
                         const double *leadR_H00, double *ndens, double *ldos, double *tr,
                         double *tre, orc_diag *diag);

/* orc_sancho with the overlap S00 of the lead cells (green_rgf.f90:4481,
 * 4521, 4525: z*S00 - H); orc_sancho is this with S00 = I. */
int orc_sancho_s(int n, double E, double eta_lead, double tol, int nmax, const double *S00,
                 const double *H00, const double *H10, double *G00, double *GBB,
                 int *nmax_hit, int *inv_failures);

/* green_RGF_CMS (green_rgf.f90:4167-4383): the compat RGF with overlap blocks
 * Sii (z S_l, Sigma^R_l S_l, Sigma^<_l S_l, decimation with S00) and the
 * reference's G^> recursion with the never-filled left-connected G^> (Glp = 0):
 * pdens_l = g_l H1i_l^H pdens_{l+1} H1i_l g_l^H for l < nx, GGi1 from it.
 * Same argument order as the reference (Sii before Hii). */
void orc_rgf_cms(const orc_params *p, const double temp[2], int nm, int nx, double E,
                 const double mu[2], const double *Sii, const double *Hii, const double *H1i,
                 const double *sigL, const double *sigG, const double *sigR,
                 double *ndens, double *pdens, double *ldos, double *tr, double *tre,
                 double *cur, double *GRi1, double *GLi1, double *GGi1, orc_diag *diag);

/* Physical-mode RGF with block-tridiagonal scattering self-energies (an
 * extension beyond green_RGF_RS, whose sigma arrays are diagonal blocks only,
 * src/green_rgf.f90:3952).  sigR, sigL: diagonal blocks as in orc_rgf;
 * sigR_lo[l] = Sigma^R_{l,l-1}, sigR_up[l] = Sigma^R_{l-1,l},
 * sigL_up[l] = Sigma^<_{l-1,l} for l = 1..nx-1 (0-based; block 0 unused),
 * Sigma^<_{l,l-1} = -(Sigma^<_{l-1,l})^H.  Any of them may be NULL (zero).
 * The leads are built from Hii + sigR at the edge slabs and H1i(:,:,1),
 * H1i(:,:,nx) as in orc_rgf.  Outputs as orc_rgf (physical mode) plus
 * GRi1_lo[l] = G^R_{l+1,l}; every pointer may be NULL.  Product order is free
 * (no compat arithmetic to reproduce).  Pinned by the dense inverse of the
 * whole open device with the same Sigma. */
void orc_rgf_offdiag(const orc_params *p, const double temp[2], int nm, int nx, double E,
                     const double mu[2], const double *Hii, const double *H1i,
                     const double *sigL, const double *sigR, const double *sigR_lo,
                     const double *sigR_up, const double *sigL_up, double *ndens,
                     double *pdens, double *ldos, double *tr, double *tre, double *cur,
                     double *GRi1, double *GLi1, double *GGi1, double *GRi1_lo, orc_diag *diag);

/* orc_rgf_offdiag with the lead decimations fed by explicit onsite blocks
 * (NULL: the edge slabs plus sigR), for derivatives with the contacts held
 * fixed; ndens and tr only. */
void orc_rgf_offdiag_fixed_leads(const orc_params *p, const double temp[2], int nm, int nx, double E,
                                 const double mu[2], const double *Hii, const double *H1i,
                                 const double *sigL, const double *sigR, const double *sigR_lo,
                                 const double *sigR_up, const double *sigL_up, const double *leadL_H00,
                                 const double *leadR_H00, double *ndens, double *tr);

/* Energy loop of the reference drivers (green_rgf.f90:1305-1316): OpenMP over
 * energies, one energy per thread.  Per-energy arrays have a trailing energy
 * index (stride nm*nm*nx complex); sig_stride_e is the energy stride of the
 * sig* arrays in complex elements, 0 when one set is shared by
 * all energies.  Any output pointer may be NULL.  nthreads <= 0: OpenMP
 * default. */
void orc_rgf_energies(const orc_params *p, const double temp[2], int nm, int nx,
                      int nE, const double *E, const double mu[2],
                      const double *Hii, const double *H1i,
                      const double *sigL, const double *sigG, const double *sigR,
                      long long sig_stride_e,
                      double *ndens, double *pdens, double *ldos, double *tr,
                      double *tre, double *cur, double *GRi1, double *GLi1,
                      double *GGi1, orc_diag *diag, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
