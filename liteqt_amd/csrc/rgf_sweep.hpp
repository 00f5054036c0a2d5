// rgf_sweep.hpp -- the per-energy RGF solve as a persistent workgroup program.
//
// Kernel rgf_sweep<T, NMAX>: workgroup b takes energies b, b + grid, ... and
// for each runs the whole solve of green_RGF_RS (src/green_rgf.f90:3950-4164):
//   left lead decimation (4451-4537) -> Sigma_L, g_1, g^<_1   (3992-4012)
//   forward sweep over slabs 2..nx-1, g_l and g^<_l kept in HBM (4013-4030)
//   right lead decimation, G_nn, G^<_nn, G^>_nn, tr          (4032-4066)
//   backward sweep: cur, G^<_{l,l+1}, G^>_{l,l+1}, G_{l,l+1}, G_ll, G^<_ll (4069-4149)
//   tre                                                       (4151-4160)
// Every product is a workgroup GEMM on FP64 MFMA and every inverse the blocked
// Gauss-Jordan of rgf_blocks.hpp; elementwise terms are folded into GEMM
// epilogues.  Per energy the slot's workspace holds g_l, g^<_l for all slabs
// (2*nx blocks) plus kTmpBlocks temporaries.
#pragma once
#include "rgf_blocks.hpp"

namespace rgf {

constexpr int kTmpBlocks = 30;

struct SweepArgs {
  int n, nx, nE, mode, dec_nmax;
  double eta_dev, eta_lead, dec_tol, boltz;
  double temp[2], mu[2];
  const double* E;
  const double2* Hii;
  const double2* H1i;
  const double2* sigL;
  const double2* sigR;
  long long sig_stride;  // complex elements between energies (0: shared)
  double2* ldos;
  double2* ndens;
  double2* pdens;
  double2* GRi1;
  double2* GLi1;
  double2* GGi1;
  double* cur;
  double* tr;
  double* tre;
  double* dens_n;
  double* dens_p;
  int* info;
  double2* ws;
  long long ws_slot;  // complex elements of workspace per slot
  // block-tridiagonal scattering self-energy (rgf_offdiag.hpp only): block l
  // of sigR_lo / sigR_up / sigL_up is Sigma^R_{l,l-1} / Sigma^R_{l-1,l} /
  // Sigma^<_{l-1,l}, l = 1..nx-1; any may be null
  const double2* sigR_lo;
  const double2* sigR_up;
  const double2* sigL_up;
  // in-place elastic SCBA refresh (rgf_offdiag.hpp only):
  // Sigma <- (1 - scba_mix) Sigma + scba_mix * scba_D * G on every block given
  int scba;
  double scba_D, scba_mix;
  // overlap blocks (green_RGF_CMS semantics, solve_energy<..., true> only)
  const double2* Sii;
  // segment recompute: 0 keeps g_l, g^<_l of every slab; K > 0 keeps them at the
  // checkpoint slabs 0, K, 2K, ... and one K-slab segment, recomputing a segment's
  // left-connected blocks from its checkpoint before the backward sweep enters it
  int ckpt;
  // G^R_{l+1,l} per bond (rgf_offdiag.hpp only), or null
  double2* GRi1_lo;
};

// Workspace blocks of one energy slot: the stored left-connected blocks (all slabs,
// or the checkpoints and one segment) and the temporaries.
__host__ __device__ inline long long sweep_slot_blocks(int nx, int ckpt, int tmp) {
  if (ckpt <= 0) return 2LL * nx + tmp;
  const long long ncp = (nx - 2) / ckpt + 1;  // checkpoint slabs 0, K, .. <= nx - 2
  return 2 * ncp + 2LL * ckpt + tmp;
}

__device__ __forceinline__ double ferm_dev(double a) { return 1.0 / (1.0 + exp(a)); }

// Sancho-Rubio decimation (src/green_rgf.f90:4451-4537) with the lead-cell
// overlap S00 (null: identity, as green_RGF_RS passes).  The lead couples
// through Hc: left lead (left=1) H_10 = Hc^H, H_01 = Hc (the call at 3995);
// right lead H_10 = Hc, H_01 = Hc^H (4035).  Writes the surface function
// G00 = (z S00 - H_SS)^-1 and returns the iterations run.
template <int T, int NMAX>
__device__ int wg_sancho(int n, double E, double eta, double tol, int nmax, const double2* H00,
                         const double2* Hc, int left, double2* G00, double2* tb, ldsd* lds,
                         ldsd* red, int& hit, int& fails, const double2* S00 = nullptr) {
  const size_t nn = (size_t)n * n;
  const double2 one = make_double2(1, 0), mone = make_double2(-1, 0), zero = make_double2(0, 0);
  const double2 zl = make_double2(E, eta);
  double2* HBB = tb;
  double2* HSS = tb + nn;
  double2* h10 = tb + 2 * nn;
  double2* h10n = tb + 3 * nn;
  double2* h01 = tb + 4 * nn;
  double2* h01n = tb + 5 * nn;
  double2* A = tb + 6 * nn;
  double2* B = tb + 7 * nn;
  double2* C = tb + 8 * nn;
  double2* B2 = tb + 9 * nn;
  double2* W = tb + 10 * nn;
  if (left) {
    wg_ew(n, h10, 0, zero, nullptr, one, nullptr, one, Hc, one);
    wg_ew(n, h01, 0, zero, Hc, one, nullptr, one, nullptr, one);
  } else {
    wg_ew(n, h10, 0, zero, Hc, one, nullptr, one, nullptr, one);
    wg_ew(n, h01, 0, zero, nullptr, one, nullptr, one, Hc, one);
  }
  wg_ew(n, HBB, 0, zero, H00, one, nullptr, one, nullptr, one);
  wg_ew(n, HSS, 0, zero, H00, one, nullptr, one, nullptr, one);
  int it_run = 0;
  for (int it = 1; it <= nmax; ++it) {
    it_run = it;
    if (S00)
      wg_ew(n, W, 1, zero, S00, zl, HBB, mone, nullptr, one);
    else
      wg_ew(n, W, 1, zl, HBB, mone, nullptr, one, nullptr, one);
    if (wg_inverse<T, NMAX>(n, W, A, lds)) ++fails;
    Epi st = epi_store();
    wg_gemm<T>(n, OP_N, A, OP_N, h10, B, 0, st, lds);
    wg_gemm<T>(n, OP_N, h01, OP_N, B, C, 0, st, lds);
    wg_ew(n, HSS, 0, zero, HSS, one, C, one, nullptr, one);
    wg_ew(n, HBB, 0, zero, HBB, one, C, one, nullptr, one);
    wg_gemm<T>(n, OP_N, h10, OP_N, B, h10n, 0, st, lds);
    wg_gemm<T>(n, OP_N, A, OP_N, h01, B2, 0, st, lds);
    Epi acc = epi_store();
    acc.D1 = HBB;
    wg_gemm<T>(n, OP_N, h10, OP_N, B2, HBB, 0, acc, lds);
    Epi es = epi_store();
    es.abs_sum = 1;
    const double err = wg_gemm<T>(n, OP_N, h01, OP_N, B2, h01n, 0, es, lds);
    double2* t = h10; h10 = h10n; h10n = t;
    t = h01; h01 = h01n; h01n = t;
    if (fabs(err) < tol) break;
    if (it == nmax) {
      hit = 1;
      wg_ew(n, HSS, 0, zero, H00, one, nullptr, one, nullptr, one);
      wg_ew(n, HBB, 0, zero, H00, one, nullptr, one, nullptr, one);
    }
  }
  if (S00)
    wg_ew(n, W, 1, zero, S00, zl, HSS, mone, nullptr, one);
  else
    wg_ew(n, W, 1, zl, HSS, mone, nullptr, one, nullptr, one);
  if (wg_inverse<T, NMAX>(n, W, G00, lds)) ++fails;
  (void)red;
  return it_run;
}

// CMS = true: green_RGF_CMS (src/green_rgf.f90:4167-4383), the compat solve
// with the overlap blocks Sii: z S_l, Sigma^R_l S_l, Sigma^<_l S_l, the
// decimation with S00, and pdens / GGi1 from the reference's G^> recursion
// whose left-connected G^> is never filled: G^>_ll = g H^H G^>_{l+1} H g^H.
template <int T, int NMAX, bool CMS = false>
__device__ void solve_energy(const SweepArgs& a, int e, double2* ws, ldsd* lds, ldsd* red) {
  const int n = a.n, nx = a.nx;
  const size_t nn = (size_t)n * n;
  const bool phys = (a.mode == 1);
  const double E = a.E[e];
  const double2 z = make_double2(E, a.eta_dev);
  const double2 one = make_double2(1, 0), mone = make_double2(-1, 0), zero = make_double2(0, 0);
  // stored left-connected blocks: slab l at position l (keep-all), or, with
  // checkpoints every K slabs, at its checkpoint slot (l % K == 0) or in the segment
  // buffer (slot l % K), which holds the segment the sweeps are in
  const int K = CMS ? 0 : a.ckpt;
  const long long nstore = K > 0 ? (long long)((nx - 2) / K + 1) + K : nx;
  double2* Gl = ws;
  double2* Gln = ws + (size_t)nstore * nn;
  double2* tb = ws + 2 * (size_t)nstore * nn;
  double2* G = tb + 0 * nn;
  double2* Gn = tb + 1 * nn;
  double2* Gp = tb + 2 * nn;
  double2* SL = tb + 3 * nn;
  double2* SR = tb + 4 * nn;
  double2* W = tb + 5 * nn;
  double2* H00 = tb + 6 * nn;
  double2* T0 = tb + 7 * nn;
  double2* T1 = tb + 8 * nn;
  double2* T2 = tb + 9 * nn;
  double2* T3 = tb + 10 * nn;
  double2* X = tb + 11 * nn;
  double2* P7 = tb + 12 * nn;
  double2* Wm = tb + 13 * nn;
  double2* Y2 = tb + 14 * nn;
  double2* GN0 = tb + 15 * nn;
  double2* V = tb + 16 * nn;
  double2* S = tb + 17 * nn;
  double2* sig = tb + 18 * nn;
  double2* stmp = tb + 19 * nn;  // 11 blocks for the decimation
  auto Hb = [&](int l) { return a.Hii + (size_t)l * nn; };
  auto Cb = [&](int l) { return a.H1i + (size_t)l * nn; };
  auto sR = [&](int l) -> const double2* { return a.sigR ? a.sigR + (size_t)e * a.sig_stride + (size_t)l * nn : nullptr; };
  auto sL = [&](int l) -> const double2* { return a.sigL ? a.sigL + (size_t)e * a.sig_stride + (size_t)l * nn : nullptr; };
  auto ob = [&](double2* arr, int l) -> double2* { return arr ? arr + ((size_t)e * nx + l) * nn : nullptr; };
  auto gpos = [&](int l) -> size_t {
    if (K <= 0) return (size_t)l;
    const int ncp = (nx - 2) / K + 1;
    return (l % K == 0) ? (size_t)(l / K) : (size_t)(ncp + l % K);
  };
  auto gl = [&](int l) { return Gl + gpos(l) * nn; };
  auto gln = [&](int l) { return Gln + gpos(l) * nn; };
  const Epi st = epi_store();
  int fails = 0, hitL = 0, hitR = 0;

  auto Sb = [&](int l) -> const double2* { return CMS ? a.Sii + (size_t)l * nn : nullptr; };
  // onsite block of slab l: H + Sigma^R (CMS: H + Sigma^R S)
  auto onsite = [&](int l, double2* out) {
    if (CMS && sR(l)) {
      Epi ea = epi_store();
      ea.D1 = Hb(l);
      wg_gemm<T>(n, OP_N, sR(l), OP_N, Sb(l), out, 0, ea, lds);
    } else {
      wg_ew(n, out, 0, zero, Hb(l), one, CMS ? nullptr : sR(l), one, nullptr, one);
    }
  };
  // lesser scattering source of slab l: Sigma^< (CMS: Sigma^< S, formed in T2)
  auto lesser = [&](int l) -> const double2* {
    if (CMS && sL(l)) {
      wg_gemm<T>(n, OP_N, sL(l), OP_N, Sb(l), T2, 0, epi_store(), lds);
      return T2;
    }
    return sL(l);
  };

  // Without a scattering Sigma^< every left-connected g^<_l = g S g^H is anti-Hermitian
  // (S: contact terms -(Sigma - Sigma^H) f, and T g^<_{l-1} T^H); the GEMM result is
  // projected onto the anti-Hermitian matrices, (X - X^H) / 2, which only removes
  // rounding error, and the backward sweep then uses the identity
  // g< H^H G_{l,l+1}^H = -(G_{l,l+1} H g<)^H (one GEMM less per slab).
  const bool ahp = !CMS && !a.sigL;
  auto lesser_store = [&](int l, const double2* Tm) {
    wg_gemm<T>(n, OP_N, Tm, OP_C, gl(l), ahp ? T2 : gln(l), 0, st, lds);
    if (ahp) wg_antiherm(n, T2, gln(l), lds);
  };

  // ---- left contact
  onsite(0, H00);
  const int itL = wg_sancho<T, NMAX>(n, E, a.eta_lead, a.dec_tol, a.dec_nmax, H00, Cb(0), 1, G, stmp,
                                     lds, red, hitL, fails, Sb(0));
  wg_gemm<T>(n, OP_N, Cb(0), OP_N, G, T0, 0, st, lds);
  wg_gemm<T>(n, OP_N, T0, OP_C, Cb(0), SL, 0, st, lds);
  const double fL = ferm_dev((E - a.mu[0]) / (a.boltz * a.temp[0]));
  wg_ew(n, sig, 0, zero, SL, make_double2(-fL, 0), lesser(0), one, SL, make_double2(fL, 0));
  if (CMS) {
    wg_ew(n, T1, 0, zero, H00, one, SL, one, nullptr, one);
    wg_ew(n, W, 1, zero, Sb(0), z, T1, mone, nullptr, one);
  } else {
    wg_ew(n, W, 1, z, H00, mone, SL, mone, nullptr, one);
  }
  if (wg_inverse<T, NMAX>(n, W, gl(0), lds)) ++fails;
  wg_gemm<T>(n, OP_N, gl(0), OP_N, sig, T0, 0, st, lds);
  lesser_store(0, T0);

  // ---- forward sweep
  // one step: g_l, g^<_l from g_{l-1}, g^<_{l-1} (also the segment recompute); a
  // singular slab is counted on the first pass only
  auto forward_step = [&](int l, bool count) {
    Epi ew = epi_store();
    if (CMS) {
      onsite(l, T1);
      ew.sacc = -1.0; ew.D1 = Sb(l); ew.s1 = z; ew.D2 = T1; ew.s2 = mone;
    } else {
      ew.diag = z; ew.sacc = -1.0; ew.D1 = Hb(l); ew.s1 = mone; ew.D2 = sR(l); ew.s2 = mone;
    }
    wg_gemm<T>(n, OP_N, Cb(l), OP_N, gl(l - 1), T0, 0, st, lds);
    wg_gemm<T>(n, OP_N, T0, OP_C, Cb(l), W, 1, ew, lds);
    if (wg_inverse<T, NMAX>(n, W, gl(l), lds) && count) ++fails;
    Epi es = epi_store();
    es.D1 = lesser(l);
    wg_gemm<T>(n, OP_N, Cb(l), OP_N, gln(l - 1), T0, 0, st, lds);
    wg_gemm<T>(n, OP_N, T0, OP_C, Cb(l), T1, 0, es, lds);
    wg_gemm<T>(n, OP_N, gl(l), OP_N, T1, T0, 0, st, lds);
    lesser_store(l, T0);
  };
  for (int l = 1; l < nx - 1; ++l) forward_step(l, true);

  // ---- right contact
  const int L1 = nx - 1;
  onsite(L1, H00);
  const int itR = wg_sancho<T, NMAX>(n, E, a.eta_lead, a.dec_tol, a.dec_nmax, H00, Cb(L1), 0, G, stmp,
                                     lds, red, hitR, fails, Sb(L1));
  wg_gemm<T>(n, OP_C, Cb(L1), OP_N, G, T0, 0, st, lds);
  wg_gemm<T>(n, OP_N, T0, OP_N, Cb(L1), SR, 0, st, lds);
  wg_gemm<T>(n, OP_N, Cb(L1), OP_N, gl(L1 - 1), T0, 0, st, lds);
  {
    Epi ew = epi_store();
    if (CMS) {
      wg_ew(n, T1, 0, zero, H00, one, SR, one, nullptr, one);
      ew.sacc = -1.0; ew.D1 = Sb(L1); ew.s1 = z; ew.D2 = T1; ew.s2 = mone;
    } else {
      ew.diag = z; ew.sacc = -1.0; ew.D1 = H00; ew.s1 = mone; ew.D2 = SR; ew.s2 = mone;
    }
    wg_gemm<T>(n, OP_N, T0, OP_C, Cb(L1), W, 1, ew, lds);
  }
  if (wg_inverse<T, NMAX>(n, W, G, lds)) ++fails;
  const double fR = ferm_dev((E - a.mu[1]) / (a.boltz * (phys ? a.temp[1] : a.temp[0])));
  wg_gemm<T>(n, OP_N, Cb(L1), OP_N, gln(L1 - 1), T0, 0, st, lds);
  {
    Epi es = epi_store();
    es.D1 = SR; es.s1 = make_double2(-fR, 0); es.D2 = lesser(L1); es.s2 = one; es.D3 = SR; es.s3 = make_double2(fR, 0);
    wg_gemm<T>(n, OP_N, T0, OP_C, Cb(L1), sig, 0, es, lds);
  }
  wg_gemm<T>(n, OP_N, G, OP_N, sig, T0, 0, st, lds);
  wg_gemm<T>(n, OP_N, T0, OP_C, G, Gn, 0, st, lds);
  wg_ew(n, Gp, 0, zero, Gn, one, G, one, G, mone);
  // tr = -Re Tr[Sigma^<_R G^>_nn - Sigma^>_R G^<_nn], Sigma^<_R = iGam fR, Sigma^>_R = iGam (fR-1)
  wg_ew(n, T0, 0, zero, SR, mone, nullptr, one, SR, one);  // iGam_R = -(SR - SR^H)
  {
    const double a1 = wg_retrace_prod(n, T0, Gp, red);
    const double a2 = wg_retrace_prod(n, T0, Gn, red);
    if (a.tr && threadIdx.x == 0) a.tr[e] = -(fR * a1 - (fR - 1.0) * a2);
  }
  if (a.ldos) wg_ew(n, ob(a.ldos, L1), 0, zero, G, one, nullptr, one, nullptr, one);
  if (a.ndens) wg_ew(n, ob(a.ndens, L1), 0, zero, Gn, one, nullptr, one, nullptr, one);
  if (a.pdens) wg_ew(n, ob(a.pdens, L1), 0, zero, Gp, one, nullptr, one, nullptr, one);
  if (a.dens_n) { double v = wg_imtrace(n, Gn, red); if (threadIdx.x == 0) a.dens_n[(size_t)e * nx + L1] = v; }
  if (a.dens_p) { double v = wg_imtrace(n, Gp, red); if (threadIdx.x == 0) a.dens_p[(size_t)e * nx + L1] = -v; }

  // ---- backward sweep
  const bool want_cur = a.cur != nullptr;
  for (int l = nx - 2; l >= 0; --l) {
    // entering a segment from above: recompute its left-connected blocks from its
    // checkpoint (the last segment is still in the buffer from the forward sweep)
    if (K > 0 && l % K == K - 1 && l < nx - 2)
      for (int m = l - K + 2; m <= l; ++m) forward_step(m, false);
    const double2* H = Cb(phys ? l + 1 : l);
    if (want_cur) {
      wg_gemm<T>(n, OP_N, H, OP_C, gl(l), T0, 0, st, lds);    // H g^H
      wg_gemm<T>(n, OP_N, Gn, OP_N, T0, T1, 0, st, lds);      // G<_{l+1} H g^H
      wg_gemm<T>(n, OP_N, H, OP_N, gln(l), T0, 0, st, lds);   // H g<
      Epi ea = epi_store();
      ea.D1 = T1;
      wg_gemm<T>(n, OP_N, G, OP_N, T0, T2, 0, ea, lds);       // G<_{l+1,l}
      wg_gemm<T>(n, OP_C, H, OP_N, T2, T3, 0, st, lds);       // H^H G<_{l+1,l}
    }
    wg_gemm<T>(n, OP_N, gl(l), OP_C, H, X, 0, st, lds);       // X = g H^H
    wg_gemm<T>(n, OP_N, X, OP_N, Gn, P7, 0, st, lds);         // X G<_{l+1}
    wg_gemm<T>(n, OP_N, gln(l), OP_C, H, Wm, 0, st, lds);     // g< H^H
    {
      Epi ea = epi_store();
      ea.D1 = P7;
      wg_gemm<T>(n, OP_N, Wm, OP_C, G, Y2, 0, ea, lds);       // G<_{l,l+1}
    }
    if (want_cur) {
      Epi ec = epi_store();
      ec.sacc = -1.0; ec.D1 = T3; ec.real_out = 1;
      wg_gemm<T>(n, OP_N, H, OP_N, Y2, a.cur + ((size_t)e * nx + l) * nn, 0, ec, lds);
    }
    if (a.GGi1) {
      if (phys) {
        wg_gemm<T>(n, OP_N, G, OP_N, H, T0, 0, st, lds);
        wg_gemm<T>(n, OP_N, T0, OP_N, gl(l), S, 0, st, lds);  // G_{l+1,l}
      } else {
        // reference: g^H H^H G^>_{nx,nx} + (Glp == 0) terms
        wg_gemm<T>(n, OP_C, gl(l), OP_C, H, T0, 0, st, lds);
        wg_gemm<T>(n, OP_N, T0, OP_N, Gp, ob(a.GGi1, l), 0, st, lds);
      }
    }
    wg_gemm<T>(n, OP_N, X, OP_N, G, GN0, 0, st, lds);         // G_{l,l+1}
    wg_gemm<T>(n, OP_N, GN0, OP_N, H, V, 0, st, lds);
    if (a.GGi1 && phys) wg_ew(n, ob(a.GGi1, l), 0, zero, Y2, one, GN0, one, S, mone);
    {
      Epi ea = epi_store();
      ea.D1 = gl(l);
      wg_gemm<T>(n, OP_N, V, OP_N, gl(l), G, 0, ea, lds);     // G_ll
      if (!ahp) {
        ea.D1 = gln(l);
        wg_gemm<T>(n, OP_N, P7, OP_C, X, T0, 0, ea, lds);       // g< + X G< X^H
        ea.D1 = T0;
        wg_gemm<T>(n, OP_N, V, OP_N, gln(l), T1, 0, ea, lds);   // + G_{l,l+1} H g<
        ea.D1 = T1;
        wg_gemm<T>(n, OP_N, Wm, OP_C, GN0, Gn, 0, ea, lds);     // + g< H^H G_{l,l+1}^H
      } else {
        // g^<_l anti-Hermitian: g< H^H G_{l,l+1}^H = -(G_{l,l+1} H g<)^H, the last
        // term is the negated conjugate transpose of the third, read in the epilogue
        wg_gemm<T>(n, OP_N, V, OP_N, gln(l), T1, 0, st, lds);   // Q = G_{l,l+1} H g<
        Epi eq = epi_store();
        eq.D1 = gln(l);
        eq.D2 = T1;
        eq.D3 = T1; eq.s3 = mone;
        wg_gemm<T>(n, OP_N, P7, OP_C, X, Gn, 0, eq, lds);       // g< + X G< X^H + Q - Q^H
      }
    }
    if (CMS) {
      // G^>_ll = g H^H G^>_{l+1} H g^H  (the left-connected G^> of the reference is zero)
      wg_gemm<T>(n, OP_N, X, OP_N, Gp, T0, 0, st, lds);
      wg_gemm<T>(n, OP_N, T0, OP_C, X, Gp, 0, st, lds);
    }
    if (a.ldos) wg_ew(n, ob(a.ldos, l), 0, zero, G, one, nullptr, one, nullptr, one);
    if (a.ndens) wg_ew(n, ob(a.ndens, l), 0, zero, Gn, one, nullptr, one, nullptr, one);
    if (a.pdens) {
      if (CMS) wg_ew(n, ob(a.pdens, l), 0, zero, Gp, one, nullptr, one, nullptr, one);
      else wg_ew(n, ob(a.pdens, l), 0, zero, Gn, one, G, one, G, mone);
    }
    if (a.GRi1) wg_ew(n, ob(a.GRi1, l), 0, zero, GN0, one, nullptr, one, nullptr, one);
    if (a.GLi1) wg_ew(n, ob(a.GLi1, l), 0, zero, Y2, one, nullptr, one, nullptr, one);
    if (a.dens_n) { double v = wg_imtrace(n, Gn, red); if (threadIdx.x == 0) a.dens_n[(size_t)e * nx + l] = v; }
    if (a.dens_p) {
      // -Im Tr G^>_ll with G^> = G^< + G - G^H (CMS: the recursion's G^>)
      double v = CMS ? wg_imtrace(n, Gp, red) : wg_imtrace(n, Gn, red) + 2.0 * wg_imtrace(n, G, red);
      if (threadIdx.x == 0) a.dens_p[(size_t)e * nx + l] = -v;
    }
  }

  // ---- tre (left contact)
  wg_ew(n, T0, 0, zero, Gn, one, G, one, G, mone);   // G^>_11
  wg_ew(n, T1, 0, zero, SL, mone, nullptr, one, SL, one);  // iGam_L
  {
    const double fLt = ferm_dev((E - a.mu[0]) / (a.boltz * (phys ? a.temp[0] : a.temp[1])));
    const double a1 = wg_retrace_prod(n, T1, T0, red);
    const double a2 = wg_retrace_prod(n, T1, Gn, red);
    if (a.tre && threadIdx.x == 0) a.tre[e] = fLt * a1 - (fLt - 1.0) * a2;
  }
  if (a.info && threadIdx.x == 0) {
    a.info[4 * (size_t)e + 0] = itL;
    a.info[4 * (size_t)e + 1] = itR;
    a.info[4 * (size_t)e + 2] = hitL | (hitR << 1);
    a.info[4 * (size_t)e + 3] = fails;
  }
  __syncthreads();
}

template <int T, int NMAX>
struct SweepLds {
  static constexpr int A = GemmLds<T>::DOUBLES;
  static constexpr int B = InvLds<T, NMAX>::DOUBLES;
  static constexpr int C = kAntihermLds;
  static constexpr int AB = (A > B ? A : B);
  static constexpr int DOUBLES = (AB > C ? AB : C);
};

template <int T, int NMAX>
__global__ __launch_bounds__(Cfg<T>::NTH, Cfg<T>::OCC) void rgf_sweep(SweepArgs a) {
  __shared__ double lds[SweepLds<T, NMAX>::DOUBLES];
  double2* ws = a.ws + (size_t)blockIdx.x * (size_t)a.ws_slot;
  for (int e = blockIdx.x; e < a.nE; e += gridDim.x) solve_energy<T, NMAX>(a, e, ws, (ldsd*)lds, (ldsd*)lds);
}

// green_RGF_CMS: the same persistent loop with the overlap variant of the solve.
template <int T, int NMAX>
__global__ __launch_bounds__(Cfg<T>::NTH, Cfg<T>::OCC) void rgf_sweep_cms(SweepArgs a) {
  __shared__ double lds[SweepLds<T, NMAX>::DOUBLES];
  double2* ws = a.ws + (size_t)blockIdx.x * (size_t)a.ws_slot;
  for (int e = blockIdx.x; e < a.nE; e += gridDim.x)
    solve_energy<T, NMAX, true>(a, e, ws, (ldsd*)lds, (ldsd*)lds);
}

// Batched building blocks: one workgroup per matrix (used by the block-level
// entry points rgf_zgemm_batched / rgf_zinv_batched).
template <int T>
__global__ __launch_bounds__(Cfg<T>::NTH, Cfg<T>::OCC) void k_gemm_batched(int n, int nb, int opA, const double2* A, int opB,
                                                        const double2* B, double2* C, int c_rowmajor) {
  __shared__ double lds[GemmLds<T>::DOUBLES];
  const size_t nn = (size_t)n * n;
  for (int b = blockIdx.x; b < nb; b += gridDim.x)
    wg_gemm<T>(n, opA, A + b * nn, opB, B + b * nn, C + b * nn, c_rowmajor, epi_store(), (ldsd*)lds);
}

// Inverse of column-major matrices: transposed copy into the row-major work
// buffer W, then the Gauss-Jordan sweep.
template <int T, int NMAX>
__global__ __launch_bounds__(Cfg<T>::NTH, Cfg<T>::OCC) void k_inv_batched(int n, int nb, const double2* A, double2* Ainv, double2* W,
                                                       int* info) {
  __shared__ double lds[SweepLds<T, NMAX>::DOUBLES];
  const size_t nn = (size_t)n * n;
  const double2 one = make_double2(1, 0), zero = make_double2(0, 0);
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    double2* w = W + (size_t)blockIdx.x * nn;
    wg_ew(n, w, 1, zero, A + b * nn, one, nullptr, one, nullptr, one);
    const int r = wg_inverse<T, NMAX>(n, w, Ainv + b * nn, (ldsd*)lds);
    if (info && threadIdx.x == 0) info[b] = r;
  }
}

}  // namespace rgf
