/* rgf.h -- C ABI of the MI355X recursive Green's function (RGF) engine.
 *
 * The engine solves, per energy E, the open block-tridiagonal device of the
 * reference's green_RGF_RS (src/green_rgf.f90:3950-4164): Sancho-Rubio lead
 * decimation (src/green_rgf.f90:4451-4537), forward left-connected sweep,
 * backward sweep for the diagonal and first off-diagonal blocks of G^R, G^<,
 * G^>, the bond-current blocks and the Meir-Wingreen spectra tr/tre.
 *
 * Conventions shared by every entry point:
 *   - complex arrays are complex128 as interleaved (re, im) doubles in Fortran
 *     order: a (nm, nm, nx) array holds the nx blocks of nm x nm column-major
 *     matrices back to back (the layout of a Fortran complex(8) array);
 *   - Hii(:,:,l) is the onsite block of slab l, H1i(:,:,l) the coupling
 *     H_{l,l-1} (H1i(:,:,1) also couples the left lead, H1i(:,:,nx) the right);
 *   - optional arrays are passed as NULL;
 *   - functions return 0 on success and a nonzero RGF_E* code on error, with a
 *     message from rgf_last_error().  Numerical failures follow the reference:
 *     a singular slab matrix is replaced by zero (src/green_rgf.f90:4429-4437)
 *     and an unconverged decimation falls back to the bare surface
 *     (src/green_rgf.f90:4513-4518); both are counted in the per-energy info.
 * All arithmetic is complex FP64 on the GPU (HIP, gfx950).
 */
#ifndef LITEQT_RGF_H
#define LITEQT_RGF_H

#ifdef __cplusplus
extern "C" {
#endif

#define RGF_MODE_COMPAT 0   /* bit-for-bit semantics of green_RGF_RS, quirks included */
#define RGF_MODE_PHYSICAL 1 /* consistent coupling indexing, true G^> off-diagonal, T_L/T_R not swapped */

#define RGF_OK 0
#define RGF_EINVAL 1   /* invalid argument (sizes, pointers, mode) */
#define RGF_EHIP 2     /* HIP runtime error */
#define RGF_ENOMEM 3   /* device memory exhausted */
#define RGF_EUNSUP 4   /* unsupported size for this entry point */

/* Numerical parameters.  The reference hard-codes them; rgf_params_reference
 * fills its values:
 *   eta_dev  = 0       (z = dcmplx(E, 0.0d-6), src/green_rgf.f90:3974)
 *   eta_lead = 1e-3    (src/green_rgf.f90:4474)
 *   dec_tol  = 1e-100  (src/green_rgf.f90:4457)
 *   dec_nmax = 100     (src/green_rgf.f90:4473)
 *   boltz    = 8.61734e-5 eV/K (src/green_rgf.f90:3968)
 *   mode     = RGF_MODE_COMPAT */
typedef struct rgf_params {
  double eta_dev;
  double eta_lead;
  double dec_tol;
  int dec_nmax;
  double boltz;
  int mode;
} rgf_params;

void rgf_params_reference(rgf_params *p);

/* Message for the last error on the calling thread (empty when none). */
const char *rgf_last_error(void);

/* Library version string. */
const char *rgf_version(void);

/* Drop-in for green_RGF_RS (src/green_rgf.f90:3950-3958), same argument order
 * and meaning, host arrays, reference parameters.  Scalars are passed by
 * value.  sigG is accepted and unused, as in the reference.  GRi1/GLi1/GGi1
 * (the absent-able optional arguments) are written for l = 1..nx-1, block nx
 * untouched.  Thread-safe: concurrent callers (the reference's OpenMP energy
 * loops, src/green_rgf.f90:1305-1316) are serialised on the device. */
int rgf_rs(const double temp[2], int nm, int nx, double E, const double mu[2],
           const double *Hii, const double *H1i, const double *sigL,
           const double *sigG, const double *sigR, double *ndens, double *pdens,
           double *ldos, double *tr, double *tre, double *cur, double *GRi1,
           double *GLi1, double *GGi1);

/* rgf_rs with explicit parameters. */
int rgf_rs_ex(const rgf_params *p, const double temp[2], int nm, int nx, double E,
              const double mu[2], const double *Hii, const double *H1i,
              const double *sigL, const double *sigG, const double *sigR,
              double *ndens, double *pdens, double *ldos, double *tr, double *tre,
              double *cur, double *GRi1, double *GLi1, double *GGi1);

/* Drop-in for green_RGF_CMS (src/green_rgf.f90:4167-4175): the RGF with the
 * overlap blocks Sii (argument before Hii, as in the reference): z S_l,
 * Sigma^R_l S_l and Sigma^<_l S_l, the lead decimation with S00 = Sii(:,:,1)
 * and Sii(:,:,nx), and pdens / GGi1 from the reference's G^> recursion whose
 * left-connected G^> (Glp) is never filled:
 * pdens_l = g_l H1i_l^H pdens_{l+1} H1i_l g_l^H below the last slab.  Compat
 * arithmetic and quirks as rgf_rs.  The reference never calls it. */
int rgf_cms(const double temp[2], int nm, int nx, double E, const double mu[2],
            const double *Sii, const double *Hii, const double *H1i, const double *sigL,
            const double *sigG, const double *sigR, double *ndens, double *pdens,
            double *ldos, double *tr, double *tre, double *cur, double *GRi1,
            double *GLi1, double *GGi1);

/* rgf_cms with explicit parameters (mode must be RGF_MODE_COMPAT). */
int rgf_cms_ex(const rgf_params *p, const double temp[2], int nm, int nx, double E,
               const double mu[2], const double *Sii, const double *Hii, const double *H1i,
               const double *sigL, const double *sigG, const double *sigR,
               double *ndens, double *pdens, double *ldos, double *tr, double *tre,
               double *cur, double *GRi1, double *GLi1, double *GGi1);

/* Energy loop of rgf_cms, as rgf_rs_energies (Sii shared by all energies). */
int rgf_cms_energies(const rgf_params *p, const double temp[2], int nm, int nx,
                     int nE, const double *E, const double mu[2], const double *Sii,
                     const double *Hii, const double *H1i, const double *sigL,
                     const double *sigG, const double *sigR, int sig_per_energy,
                     double *ndens, double *pdens, double *ldos, double *tr,
                     double *tre, double *cur, double *GRi1, double *GLi1,
                     double *GGi1, int *info);

/* Replacement of the drivers' energy loops `do ie ... call green_RGF_RS(...)`
 * (src/green_rgf.f90:3174-3180, 2914-2927, 1301-1318): all nE energies in one
 * batched GPU solve.  Host arrays with a trailing energy index, i.e. the
 * caller's (nm, nm, nx, nen) slices.  sig* have energy stride nm*nm*nx when
 * sig_per_energy != 0, else one (nm, nm, nx) set is shared.  tr/tre have nE
 * entries.  info (optional, 4*nE ints) receives per energy: decimation
 * iterations left, right, nmax-hit bits (1 left, 2 right), singular-matrix
 * count. */
int rgf_rs_energies(const rgf_params *p, const double temp[2], int nm, int nx,
                    int nE, const double *E, const double mu[2],
                    const double *Hii, const double *H1i, const double *sigL,
                    const double *sigG, const double *sigR, int sig_per_energy,
                    double *ndens, double *pdens, double *ldos, double *tr,
                    double *tre, double *cur, double *GRi1, double *GLi1,
                    double *GGi1, int *info);

/* Device-resident batched solve (the engine's native entry point).  Every
 * pointer is device memory; the call is asynchronous on `stream` (a
 * hipStream_t, NULL for the default stream).  nm must be a multiple of 32 up
 * to 256 or a multiple of 64 up to 512 (with off-diagonal Sigma or overlaps:
 * a multiple of 32 up to 256); the host-array entry points pad other sizes.  Per-energy
 * arrays have a trailing energy index with stride nm*nm*nx (complex for the G
 * blocks, real for cur).  dens_n / dens_p (nx per energy) receive
 * Im Tr G^<_ll and -Im Tr G^>_ll, the integrands of the electron and hole
 * densities of src/green_rgf.f90:3919-3945.  max_slots > 0 caps the number of
 * energies in flight (each holds 2*nx stored left-connected blocks in HBM);
 * 0 lets the engine size it from free memory and the CU count. */
typedef struct rgf_batch {
  int nm, nx, nE;
  int max_slots;
  const double *E;
  double temp[2];
  double mu[2];
  const double *Hii, *H1i;
  const double *sigL, *sigG, *sigR;
  long long sig_stride; /* complex elements between energies; 0 = shared */
  double *ldos, *ndens, *pdens, *GRi1, *GLi1, *GGi1;
  double *cur;
  double *tr, *tre;
  double *dens_n, *dens_p;
  int *info;
  /* First off-diagonal blocks of a block-tridiagonal scattering self-energy
   * (an extension: the reference's sigma arrays are diagonal blocks only,
   * src/green_rgf.f90:3952).  Block l (0-based, l = 1..nx-1; block 0 unused)
   * of sigR_lo / sigR_up / sigL_up is Sigma^R_{l,l-1} / Sigma^R_{l-1,l} /
   * Sigma^<_{l-1,l}; Sigma^<_{l,l-1} = -(Sigma^<_{l-1,l})^H.  Same energy
   * stride as sigL/sigR.  Any may be NULL (zero); a non-NULL one needs
   * RGF_MODE_PHYSICAL.  Zero-initialise the struct so unused fields are NULL. */
  const double *sigR_lo, *sigR_up, *sigL_up;
  /* Overlap blocks S_l (nx blocks, shared by all energies) for the
   * green_RGF_CMS solve (src/green_rgf.f90:4167-4383); NULL for green_RGF_RS.
   * Needs RGF_MODE_COMPAT and no off-diagonal Sigma. */
  const double *Sii;
  /* Storage of the left-connected blocks g_l, g^<_l between the forward and the
   * backward sweep (diagonal-Sigma solve): -1 keeps every slab's blocks (2*nx
   * blocks per energy in flight); K > 0 keeps them every K slabs and recomputes
   * each K-slab segment from its checkpoint before the backward sweep enters it
   * (2*((nx-2)/K + 1) + 2*K blocks; one more forward sweep, results bit-identical);
   * 0 (zero-initialised struct) keeps every slab unless the slots that fit in
   * free device memory leave CUs idle and checkpoints every ceil(sqrt(nx)) slabs
   * run more energies per unit of time. */
  int ckpt;
  /* Lower off-diagonal blocks G^R_{l+1,l} (l = 1..nx-1 at block l-1, as GRi1),
   * written by the block-tridiagonal-Sigma solve only (sigR_lo, sigR_up or sigL_up
   * given); NULL elsewhere.  With a non-Hermitian effective A = z - H - Sigma^R
   * (Sigma^R_{l-1,l} unrelated to Sigma^R_{l,l-1}) it is not fixed by GRi1. */
  double *GRi1_lo;
} rgf_batch;

int rgf_solve_batched(const rgf_params *p, const rgf_batch *b, void *stream);

/* rgf_solve_batched without any allocation: no counterpart in the reference.
 * It is for callers that cannot allocate device memory (code under HIP-graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, semantics and results as rgf_solve_batched, but the per-energy
 * workspace (the 2*nx stored left-connected blocks plus temporaries of each
 * energy slot) comes from a fixed device array of 1 GiB per GPU that the
 * library holds until the process ends; the number of energies in flight is
 * bounded by how many slots fit in it.  A batch whose single slot exceeds it
 * ((2*nx + 30) * nm*nm*16 bytes > 1 GiB, or (2*nx + 36) blocks with
 * off-diagonal Sigma) is rejected with RGF_EUNSUP before anything runs.  The
 * workspace needs no initialisation.  Every call on one GPU shares that GPU's
 * array: calls that overlap in time on one GPU (different streams or host
 * threads without ordering) race and must be serialised by the caller. */
int rgf_solve_batched_noalloc(const rgf_params *p, const rgf_batch *b, void *stream);

/* One elastic self-consistent Born (SCBA) iteration: the batched solve of b
 * (RGF_MODE_PHYSICAL, per-energy Sigma with sig_stride = nm*nm*nx) that also
 * refreshes, in place and per energy, every Sigma block it was given:
 *   Sigma <- (1 - mix) * Sigma + mix * D * G
 * Sigma^R (sigR, sigR_lo, sigR_up) from G^R, Sigma^< (sigL, sigL_up) from
 * G^< on the same blocks.  sigL/sigR are written through despite their const
 * qualifier.  The outputs of b are those of the solve with the old Sigma.
 * No counterpart in the reference (whose SCBA loops build Sigma on the host,
 * src/green_rgf.f90:3174-3290); it keeps the self-energy in HBM between
 * iterations. */
typedef struct rgf_scba {
  double D;   /* scattering strength (eV^2 for deformation-potential models) */
  double mix; /* linear mixing of the new self-energy, 0 < mix <= 1 */
} rgf_scba;

int rgf_scba_step(const rgf_params *p, const rgf_batch *b, const rgf_scba *scba, void *stream);

/* HIP-graph form of rgf_solve_batched / rgf_scba_step: rgf_graph_create
 * captures one launch on `stream` (a created stream, not NULL) into an
 * executable graph that owns its workspace; rgf_graph_launch replays it
 * `replays` times on `stream` (with scba given, each replay is one SCBA
 * iteration on the same device arrays).  The graph keeps the pointers of b.
 * No counterpart in the reference; for drivers that repeat the solve on
 * refreshed inputs without host launch overhead. */
typedef struct rgf_graph_s *rgf_graph;
int rgf_graph_create(const rgf_params *p, const rgf_batch *b, const rgf_scba *scba, void *stream,
                     rgf_graph *out);
int rgf_graph_launch(rgf_graph g, int replays, void *stream);
void rgf_graph_destroy(rgf_graph g);

/* Many independent devices in one persistent launch (the reference's
 * drivers call green_RGF_RS once per device and energy, e.g. the k-point and
 * device loops around src/green_rgf.f90:1301-1318).  devs[j] describes device
 * j exactly as rgf_solve_batched's batch does (its own nm, multiple of 32 and
 * at most 256; nx; nE and energy grid; Hii, H1i, diagonal sigma and outputs;
 * max_slots and the off-diagonal sigma fields must be 0/NULL).  All
 * (device, energy) pairs form one work list ordered by predicted cost,
 * longest first; the workgroups of one launch take items from it
 * dynamically, each with the tile variant of its device's block size.
 * autotune != 0: before the first solve at a block size, time its candidate
 * tile variants on the device (results kept per process) and use the fastest
 * and its measured cost for the ordering.  The number of workgroups is
 * bounded by the CU count and by free HBM for the largest device's stored
 * blocks.  tile_T (optional, ndev ints) receives the tile parameter chosen
 * per device; rep (optional) the launch shape.  Synchronises the device
 * before it starts (its descriptor buffer is reused), then runs
 * asynchronously on `stream`.  No counterpart in the reference. */
typedef struct rgf_multi_report {
  int slots;          /* workgroups of the persistent launch */
  int tuned;          /* block sizes timed by the autotuner in this call */
  long long items;    /* (device, energy) pairs */
  long long ws_bytes; /* device workspace of the launch */
} rgf_multi_report;

int rgf_solve_devices(const rgf_params *p, const rgf_batch *devs, int ndev, int autotune,
                      int *tile_T, rgf_multi_report *rep, void *stream);

/* rgf_solve_devices captured as one HIP graph (counter reset + the persistent
 * launch; the graph owns its workspace and descriptors); replay it with
 * rgf_graph_launch, release it with rgf_graph_destroy. */
int rgf_graph_create_devices(const rgf_params *p, const rgf_batch *devs, int ndev, int autotune,
                             void *stream, rgf_graph *out);

/* Bytes of device workspace rgf_solve_batched uses for `slots` energies in
 * flight at block size nm and nx slabs. */
long long rgf_workspace_bytes(int nm, int nx, int slots);

/* Adjoint (reverse-mode) RGF of the physical-mode device with the scattering
 * self-energy held fixed (diagonal sigL/sigR and, when given, the off-diagonal
 * sigR_lo/sigR_up/sigL_up of b).  Per energy e the objective is
 *   J_e = w_tr[e] * tr_e + sum_l w_n[e*nx + l] * Im Tr G^<_ll
 * (tr: current spectrum into the right contact, src/green_rgf.f90:4058-4066;
 * Im Tr G^<_ll: the density integrand of src/green_rgf.f90:3919-3945); w_tr
 * or w_n may be NULL (zero).  Outputs (device arrays): J[e]; dU[e*nx + l] =
 * dJ_e/dU_l for a potential U_l added to the diagonal of slab l; dH1i (nx
 * blocks, may be NULL) = sum_e (dJ_e/dRe H1i_l + i dJ_e/dIm H1i_l).
 * Contacts: lead_H00 NULL builds the leads from the edge slabs, Hii + sigR at
 * l = 1 and nx with the couplings H1i_1 and H1i_nx, as the reference does
 * (src/green_rgf.f90:3995, 4035), and differentiates through them: a potential
 * on an edge slab moves its lead too, and dH1i_1, dH1i_nx include the contact
 * couplings' role in the decimation (the surface function's fixed point,
 * differentiated implicitly: a Stein equation summed by doubling; info bit 2 /
 * bit 3 of info[4e+2] flag a left / right sum that did not converge within
 * dec_nmax doublings).  lead_H00 given (2 blocks: left and right lead onsite)
 * holds the contacts fixed at those blocks: dH1i is then zero for l = 1 and nx
 * and dU_1, dU_nx move only the device slabs.  b supplies E, Hii, H1i, sigL,
 * sigR, sig_stride, temp, mu, nE, nm (multiple of 32, <= 256), nx, max_slots
 * and optionally tr and info.  Asynchronous on `stream`.  No counterpart in
 * the reference (which has no derivatives); it serves optimisation and
 * inverse-design drivers around the energy loop. */
int rgf_adjoint_batched(const rgf_params *p, const rgf_batch *b, const double *w_tr,
                        const double *w_n, const double *lead_H00, double *J, double *dU,
                        double *dH1i, void *stream);

/* Device workspace of rgf_adjoint_batched for `slots` energies in flight with the
 * small slot (7*nx + 48 blocks per energy, plus nx coupling-adjoint blocks).  When
 * every slot fits in 90% of free device memory the engine takes the larger slot
 * (10*nx + 48 blocks), which keeps three more products per slab and saves 5 GEMMs per
 * slab; results are bit-identical.  The environment variable LITEQT_ADJ_RECOMPUTE=1
 * forces the small slot. */
long long rgf_adjoint_workspace_bytes(int nm, int nx, int slots);

/* Batched block operations of the engine, one workgroup per matrix, device
 * arrays of nbatch contiguous n x n column-major complex matrices, asynchronous
 * on `stream`; n a multiple of 32 in [32, 256].  They are the building blocks
 * every slab step is made of (the reference's zgemm and invert call sites,
 * src/green_rgf.f90:4002-4155 and 4421-4441), exposed for validation and
 * for callers that assemble their own sweeps.
 *   rgf_zgemm_batched: C = op(A) op(B), op 0 = none, 1 = conjugate transpose;
 *                      C row-major when c_rowmajor.
 *   rgf_zinv_batched:  Ainv = A^-1 by Gauss-Jordan with partial row pivoting;
 *                      info[b] = 0 or k+1 for a zero pivot in column k (then
 *                      Ainv = 0, the reference's convention). */
int rgf_zgemm_batched(int n, int nbatch, int opA, const double *A, int opB, const double *B,
                      double *C, int c_rowmajor, void *stream);
int rgf_zinv_batched(int n, int nbatch, const double *A, double *Ainv, int *info, void *stream);

/* Energy and k-point integrals of the batched solve's per-energy reductions,
 * replacing calc_n_electron (src/green_rgf.f90:3920-3945; called at 752) and
 * the terminal current of the SCBA drivers (src/green_rgf.f90:3189, 788).
 * Device arrays, asynchronous on `stream`:
 *   nelec[l] = dE/(2 pi) * dK * sum_{ik, ie: E[ie] >  midgap[l]} dens_n[(ik*nE + ie)*nx + l]
 *   pelec[l] = dE/(2 pi) * dK * sum_{ik, ie: E[ie] <= midgap[l]} dens_p[(ik*nE + ie)*nx + l]
 *   current[0] = scale * dE * sum_{ik, ie} tr[ik*nE + ie]
 * with dE = E[1] - E[0] and dK = 2 pi / nk (pi = 3.14159265359, the
 * reference's constant).  dens_n / dens_p are what rgf_solve_batched writes
 * (Im Tr G^<_ll and -Im Tr G^>_ll per energy and slab), so the integral reads
 * 2*nx + 1 doubles per energy and k-point instead of the full Green's
 * functions; it equals the reference's sum of Im G^<_jj / -Im G^>_jj over the
 * orbitals j of each slab.  The reference's current is scale =
 * e0^2 / (2 pi hbar) * spindeg / nk (e0 = 1.6022e-19 C, hbar = 1.0546e-34 J s).
 * dens_n / dens_p / tr may be NULL (their outputs are then not written);
 * densities need midgap, nelec and pelec, tr needs current.  Sums run in a
 * fixed order (deterministic).  nE >= 2. */
int rgf_integrate(int nk, int nE, int nx, const double *E, const double *midgap,
                  const double *dens_n, const double *dens_p, const double *tr, double scale,
                  double *nelec, double *pelec, double *current, void *stream);

/* Screened interaction W of the GW loop: green_rgf_calc_w
 * (src/green_rgf.f90:3399-3831) for a batch of energies (frequencies), one
 * persistent workgroup per energy on the solve's GEMM and inverse blocks.  For
 * each energy, M = I - V P^r is block-tridiagonal (P^r block-diagonal) and
 *   W^r_ii  = diagonal blocks of M^-1 V by the reference's right-to-left /
 *             left-to-right block recursion, whose off-diagonal step uses plain
 *             transposes (src/green_rgf.f90:3771-3791): it gives the blocks of
 *             (I - V P^r)^-1 V when V and P^r are symmetric, as in the reference;
 *   W^<_ii  = W^r_ii P^<_ii W^r_ii^H - X_a P^<_b X_a^T - X_m^T P^<_m X_m with X the
 *             imaginary part (times i) of W^r_{i,i+1}, a = i, b = i + 1 (a = nx - 1,
 *             b = nx at the last slab), m = max(i - 1, 1) (1-based; 3798-3826);
 *   W^>_ii  likewise with P^>.
 * Every driver calls it with NBC = 0 and without off-diagonal P blocks
 * (src/green_rgf.f90:846, 1551, 3008); those are the supported inputs.
 * Device arrays of complex double in the reference layout: Vii, V1i (nm, nm, nx)
 * with V1i block ix = V_{ix+1,ix}; PL, PG, PR, WL, WG, WR (nm, nm, nx, nE).
 * Asynchronous on `stream`; any nm up to 512 (zero-padded internally to a
 * multiple of 32, or of 64 above 256), nx >= 2.  Outputs must not alias inputs. */
typedef struct {
  int nm, nx, nE;
  int nbc;             /* open-boundary blocks: 0 (others: RGF_EUNSUP) */
  const double *Vii;   /* (nm, nm, nx) */
  const double *V1i;   /* (nm, nm, nx) */
  const double *PL;    /* (nm, nm, nx, nE) diagonal blocks of P^<, P^>, P^r */
  const double *PG;
  const double *PR;
  double *WL;          /* (nm, nm, nx, nE) diagonal blocks of W^<, W^>, W^r */
  double *WG;
  double *WR;
  double *WRi1;        /* optional: W^r_{ix,ix+1} (the reference's WRi1; slab nx zero) */
  int *info;           /* optional device array [nE]: singular blocks met per energy */
} rgf_w_batch;

int rgf_w_solve_batched(const rgf_w_batch *b, void *stream);

/* Release cached device buffers of this library on the current device. */
void rgf_release(void);

#ifdef __cplusplus
}
#endif

#endif /* LITEQT_RGF_H */
