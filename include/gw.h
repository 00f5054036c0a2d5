/* GW self-energy step of the one-dimensional SCBA driver, on MI355X (gfx950).
 *
 * Replaces the loop bodies between two RGF solves of green_rgf_solve_gw_1D
 * (src/green_rgf.f90:3186-3290): the polarization
 *   P^<_ij(w)  = dE_P sum_E G^<_ij(E) G^>_ji(E - w)
 *   P^>_ij(w)  = dE_P sum_E G^>_ij(E) G^<_ji(E - w)
 *   P^r_ij(w)  = dE_P sum_E [G^<_ij(E) conj(G^r_ij(E - w)) + G^r_ij(E) G^<_ji(E - w)]
 * (src/green_rgf.f90:3190-3221) and the self-energy
 *   Sigma^x_ij(E) = dE_S sum_w G^x_ij(E - w) W^x_ij(w)   (x = <, >)
 *   Sigma^r_ij(E) = dE_S sum_w [G^<_ij W^r_ij + G^r_ij W^<_ij + G^r_ij W^r_ij]
 * with Im Sigma^r replaced by Im(Sigma^> - Sigma^<)/2, the linear mixing into the
 * running self-energy and the copy of slab 3 into slabs 1-2 and of slab nx-2 into
 * slabs nx-1..nx (src/green_rgf.f90:3233-3290).
 *
 * All arrays are device pointers to complex double (interleaved re, im) in the
 * reference's Fortran layout (nm, nm, nx, nen).  As the reference does before this
 * step (src/green_rgf.f90:3186-3188), only the imaginary parts of G^<, G^>, G^r are
 * used (their real parts are taken as zero).  Frequencies w = nop * (E2 - E1) with
 * |nop| <= nopmax are stored at energy index nop + nen/2 (1-based); P is zero at the
 * other indices and for |i - j| > ndiag.  Sigma sums use the reference's index
 * condition (E > max(w, E_1) and E - w < E_nen: the last energy of G never enters and
 * Sigma_new(E_1) = 0).
 *
 * Two forms compute the same sums.  The direct form (an element's energy series staged in
 * LDS, O(nen^2) FMAs per element) serves gw_polarization up to nen = 3072 and
 * gw_selfenergy up to nen = 2040; above that, and always in the _fft entry points, the
 * FFT form of the reference's corr1d2 / conv1d2 with method 'fft' (src/fft_mod.f90:56-212)
 * runs: every band element's series zero-padded to N = 2^m >= nen + nopmax points
 * (N >= 64, N <= 2^24), transformed, multiplied and transformed back (O(N log N) per
 * element).  Its spectra live in a per-device workspace of K N * 16 bytes per band slot
 * (K = 3 for P, 5 for Sigma; slots nm min(2 ndiag + 1, nm) per slab), allocated for a
 * chunk of slabs of up to 4 GiB (one slab at least) and kept like the gather workspace.
 * The two forms agree to rounding: the FFT's error is relative to the norms of the
 * series, not to each output value.
 *
 * Calls on one stream run in order; the arrays must not overlap each other.  For dense
 * bands (ndiag >= 8) gw_polarization gathers the band's Im G series into a per-device
 * workspace it allocates (nm min(2 ndiag + 1, nm) nx 3 nen doubles, kept until
 * gw_release or the end of the process): polarization calls that overlap on one GPU
 * (different streams) race on it.  When that allocation fails the call still runs,
 * staging from the arrays directly.
 * Return codes are those of rgf.h (0 ok, 1 invalid argument, 2 HIP error, 3 device
 * memory exhausted, 4 unsupported size); gw_last_error() gives the message (per thread).
 */
#ifndef LITEQT_GW_H
#define LITEQT_GW_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nm;           /* orbitals per slab */
  int nx;           /* slabs (>= 5 for the self-energy step: the edge copy needs slabs 3 and nx-2) */
  int nen;          /* energies (uniform grid) */
  int ndiag;        /* band half-width: pairs with |i - j| <= ndiag (reference: 1 or nb*ns) */
  int nopmax;       /* largest |nop| (reference 1D driver: nen/2 - 10) */
  double dE_P[2];   /* complex prefactor of P (reference: (0, -(E2-E1)/(2 pi) * spindeg)) */
  double dE_S[2];   /* complex prefactor of Sigma (reference: (0, (E2-E1)/(2 pi))) */
  double alpha_mix; /* Sigma_gw += alpha_mix * (Sigma_new - Sigma_gw) */
} gw_params;

/* gw_params of green_rgf_solve_gw_1D for an energy grid of nen points spaced dE. */
void gw_params_reference(gw_params* p, int nm, int nx, int nen, int ndiag, double dE, double spindeg,
                         double alpha_mix);

/* P^<, P^>, P^r from G^<, G^>, G^r (src/green_rgf.f90:3190-3221): the direct form for
 * nen <= 3072, the FFT form above. */
int gw_polarization(const gw_params* p, const double* Gl, const double* Gg, const double* Gr, double* Pl,
                    double* Pg, double* Pr, void* stream);

/* Sigma^<, Sigma^>, Sigma^r (the running self-energy, updated in place) from G and
 * W^<, W^>, W^r (src/green_rgf.f90:3233-3290): the direct form for nen <= 2040, the FFT
 * form above.  nx >= 5. */
int gw_selfenergy(const gw_params* p, const double* Gl, const double* Gg, const double* Gr, const double* Wl,
                  const double* Wg, const double* Wr, double* Sl, double* Sg, double* Sr, void* stream);

/* The same steps always in the FFT form: the polarization's corr1d2(..., method='fft')
 * calls of green_rgf_solve_gw_ephoton_3d_ijs (src/green_rgf.f90:1397-1468,
 * src/fft_mod.f90:56-90) and the self-energy's conv1d2 (src/green_rgf.f90:1600-1672,
 * src/fft_mod.f90:182-212), applied to the one-dimensional driver's index ranges. */
int gw_polarization_fft(const gw_params* p, const double* Gl, const double* Gg, const double* Gr, double* Pl,
                        double* Pg, double* Pr, void* stream);
int gw_selfenergy_fft(const gw_params* p, const double* Gl, const double* Gg, const double* Gr, const double* Wl,
                      const double* Wg, const double* Wr, double* Sl, double* Sg, double* Sr, void* stream);

/* Dense screened interaction of the one-dimensional driver, green_calc_w_full
 * (src/green_rgf.f90:3347-3396, called per energy at 3235), for nE energies at once:
 * V the block-tridiagonal bare interaction (Vii diagonal blocks, V1i block ix =
 * V_{ix+1,ix}, its conjugate transpose above the diagonal), P^{<,>,r} block-diagonal
 * (nm, nm, nx, nE), M = I - V P^r, W^r = M^-1 V (solved from the LU factors),
 * W^{<,>} = W^r P^{<,>} W^r^H; the diagonal blocks (nm, nm, nx, nE) are written.  Dense
 * NT x NT algebra (NT = nm nx <= 16384) by rocBLAS / rocSOLVER over batches of energies;
 * workspace 3 NT^2 complex per energy of a batch (up to 64), kept per device. */
int gw_w_full(int nm, int nx, int nE, const double* Vii, const double* V1i, const double* PL, const double* PG,
              const double* PR, double* WL, double* WG, double* WR, void* stream);

/* Valid (ie, nop) pairs summed per band element (algorithmic work): for P and for Sigma. */
long long gw_pairs_polarization(int nen, int nopmax);
long long gw_pairs_selfenergy(int nen, int nopmax);

const char* gw_last_error(void);

/* Free the polarization's gather workspace on the current device. */
void gw_release(void);

#ifdef __cplusplus
}
#endif

#endif
