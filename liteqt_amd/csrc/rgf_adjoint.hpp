// rgf_adjoint.hpp -- reverse-mode (adjoint) RGF: gradients of
//   J(E) = w_tr(E) * tr(E) + sum_l w_n(E,l) * Im Tr G^<_ll(E)
// with respect to a potential U_l on the diagonal of every slab and to every
// coupling block H1i(:,:,l), for the consistent (physical) device with the
// scattering self-energy (diagonal and, optionally, first off-diagonal blocks)
// held fixed.  The contacts either repeat the edge slabs (the reference's
// device, differentiated through the decimation: lead_adjoint) or are held
// fixed at given onsite blocks (AdjArgs.lead; the edge couplings then get no
// gradient).
// tr is the Meir-Wingreen current spectrum into the right contact
// (src/green_rgf.f90:4058-4066): with f_L = 1, f_R = 0 it is T(E); with Fermi
// factors and weights dE*e/h it gives dI/dU_l.  Densities are the integrands of
// src/green_rgf.f90:3919-3945.
//
// Per energy a workgroup runs the forward sweep, the backward sweep, the
// reverse of the backward sweep and the reverse of the forward sweep (the
// recursion is written out above adjoint_energy).  The adjoint convention is
// Mbar = dJ/dRe M + i dJ/dIm M, so C = A B gives Abar += Cbar B^H,
// Bbar += A^H Cbar and C = A^-1 gives Abar -= C^H Cbar C^H.
#pragma once
#include "rgf_sweep.hpp"

namespace rgf {

constexpr int kAdjTmp = 48;

struct AdjArgs {
  int n, nx, nE, dec_nmax;
  double eta_dev, eta_lead, dec_tol, boltz;
  double temp[2], mu[2];
  const double* E;
  const double2* Hii;
  const double2* H1i;
  const double2* sigL;
  const double2* sigR;
  const double2* sigR_lo;  // Sigma^R_{l,l-1} (block l), or null
  const double2* sigR_up;  // Sigma^R_{l-1,l}, or null
  const double2* sigL_up;  // Sigma^<_{l-1,l}, or null
  long long sig_stride;
  const double* w_tr;  // nE or null (0)
  const double* w_n;   // nE x nx or null (0)
  const double2* lead; // 2 blocks: left and right lead onsite, or null (edge slabs + sigR)
  double* J;           // nE
  double* tr;          // nE or null
  double* dU;          // nE x nx
  double2* dH;         // slots x nx blocks: per-slot coupling adjoints, summed over its energies
  int* info;
  double2* ws;
  long long ws_slot;
  int keep;  // 1: the products the reverse passes reuse are kept per slab (larger slot)
  // segment recompute of the forward state (g, g<, S): 0 keeps it for every slab; K > 0
  // keeps it at the slabs 0, K, 2K, ... <= nx-2 and nx-1 plus one K-slab segment, which
  // the backward sweep and the two reverse passes refill from its checkpoint (keep = 0)
  int ckpt;
};

// Forward-state slabs of a slot (g, g<, S each): nx, or the checkpoints, the last slab
// and one segment.
__host__ __device__ inline long long adj_state_slabs(int nx, int ckpt) {
  if (ckpt <= 0) return nx;
  return (long long)((nx - 2) / ckpt + 1) + 1 + ckpt;
}

// Blocks of a workspace slot: g, g<, S per state slab, G, G<, gb, glb per slab (+ R,
// R d, g S per slab when keep), then kAdjTmp temporaries.
__host__ __device__ inline long long adj_slot_blocks(int nx, int keep, int ckpt = 0) {
  return 3 * adj_state_slabs(nx, ckpt) + (long long)(keep ? 7 : 4) * nx + kAdjTmp;
}

// Per energy, with d_l = H1i_l + Sigma^R_{l,l-1}, u_l = H1i_l^H + Sigma^R_{l-1,l},
// s_l = Sigma^<_{l-1,l} (the plain couplings when the off-diagonal blocks are absent):
//   forward   g_l = (z - H_l - Sigma^R_l - B_l u_l)^-1, B_l = d_l g_{l-1} (contacts on the
//             edge slabs), g<_l = g_l S_l g_l^H, S_l = Sigma^<_l + Y_l d_l^H + V_l - V_l^H,
//             Y_l = d_l g<_{l-1}, V_l = B_l s_l            (g, g<, S stored; slab nx-1 is G, G<)
//   backward  Q = g_l u, R = Q G_{l+1}, G_l = g_l + R d g_l,
//             G<_l = g<_l + Q G<_{l+1} Q^H + W - W^H, W = g_l s R^H + R d g<_l   (G, G< stored)
//   pass A    reverse of the backward sweep, l = 0..nx-2 (Q, R, R d, Q G< kept from the
//             backward sweep when keep, else recomputed; g s recomputed)
//   pass B    reverse of the forward sweep, l = nx-1..0: dJ/dU_l = -Re Tr Abar_l (M = g S
//             kept from the forward sweep when keep)
// With keep, Q_l and (Q G<)_l sit in the slab-l blocks of glb and gb, which pass A
// writes only after their last use in its step l.
// Coupling adjoints d-bar + u-bar^H accumulate per slab in dH.
//
// Contacts that repeat the edge slabs (a.lead null, the reference's device): after pass
// B the adjoints of the contact self-energies, Sigma_L-bar = -Abar_0 + f_L (Sbar_0^H -
// Sbar_0) and Sigma_R-bar = -Abar_N + f_R (Sbar_N^H - Sbar_N) + w_tr ((G<_N)^H - G<_N)
// + 2 w_tr f_R (G_N^H - G_N) (the last two from tr's own K_R), are carried through
// Sigma_L = C g C^H (C = H1i_1) and Sigma_R = C^H g C (C = H1i_nx) and through the
// decimation's fixed point g = (z - H00 - Sigma)^-1 (lead_adjoint): dU_1, dU_nx gain
// Re Tr H00-bar, dH_1 and dH_nx the contact couplings' adjoints.

// Sum of |P_ij| below which the doubling of lead_adjoint stops: P is dimensionless
// (a coupling times a surface function), and the next term is then below 1e-16 of X.
constexpr double kSteinTol = 1e-8;

// Adjoint of a contact given Sigma-bar Sb (left: Sigma = C g C^H with g = (z - H00 -
// C g C^H)^-1; right: Sigma = C^H g C with g = (z - H00 - C^H g C)^-1).  C-bar is added
// to Cbar; returns Re Tr H00-bar (uniform).  The total g-bar X solves the Stein
// equation X = g-bar + P X P2 (left: P = C^H g^H, P2 = g^H C; right: P = C g^H,
// P2 = g^H C^H), summed by doubling (X += P X P2, P <- P^2, P2 <- P2^2); steps is the
// number of doublings, or -1 when kSteinTol is not reached within nmax.
// t: 8 temporary blocks.
template <int T>
__device__ double lead_adjoint(int n, bool left, const double2* C, const double2* g, const double2* Sb,
                               double2* Cbar, double2* t, int nmax, int& steps, ldsd* lds, ldsd* red) {
  const size_t nn = (size_t)n * n;
  double2 *X = t, *P = t + nn, *P2 = t + 2 * nn, *Pn = t + 3 * nn, *P2n = t + 4 * nn, *Tt = t + 5 * nn;
  double2 *U = t + 6 * nn, *Hb = t + 7 * nn;
  const Epi st = epi_store();
  Epi ac = epi_store();
  ac.D1 = Cbar;
  if (left) {
    wg_gemm<T>(n, OP_N, Sb, OP_N, C, U, 0, st, lds);      // Sb C
    wg_gemm<T>(n, OP_C, C, OP_N, U, X, 0, st, lds);       // g-bar = C^H Sb C
    wg_gemm<T>(n, OP_N, U, OP_C, g, Cbar, 0, ac, lds);    // C-bar += Sb C g^H
    wg_gemm<T>(n, OP_C, Sb, OP_N, C, U, 0, st, lds);
    wg_gemm<T>(n, OP_N, U, OP_N, g, Cbar, 0, ac, lds);    // C-bar += Sb^H C g
    wg_gemm<T>(n, OP_C, C, OP_C, g, P, 0, st, lds);       // P = C^H g^H
    wg_gemm<T>(n, OP_C, g, OP_N, C, P2, 0, st, lds);      // P2 = g^H C
  } else {
    wg_gemm<T>(n, OP_N, C, OP_N, Sb, U, 0, st, lds);      // C Sb
    wg_gemm<T>(n, OP_N, U, OP_C, C, X, 0, st, lds);       // g-bar = C Sb C^H
    wg_gemm<T>(n, OP_C, g, OP_N, C, U, 0, st, lds);
    wg_gemm<T>(n, OP_N, U, OP_N, Sb, Cbar, 0, ac, lds);   // C-bar += g^H C Sb
    wg_gemm<T>(n, OP_N, g, OP_N, C, U, 0, st, lds);
    wg_gemm<T>(n, OP_N, U, OP_C, Sb, Cbar, 0, ac, lds);   // C-bar += g C Sb^H
    wg_gemm<T>(n, OP_N, C, OP_C, g, P, 0, st, lds);       // P = C g^H
    wg_gemm<T>(n, OP_C, g, OP_C, C, P2, 0, st, lds);      // P2 = g^H C^H
  }
  steps = -1;
  Epi ax = epi_store();
  ax.D1 = X;
  Epi as = epi_store();
  as.abs_sum = 1;
  for (int k = 1; k <= nmax; ++k) {
    wg_gemm<T>(n, OP_N, P, OP_N, X, Tt, 0, st, lds);
    wg_gemm<T>(n, OP_N, Tt, OP_N, P2, X, 0, ax, lds);     // X += P X P2
    const double sp = wg_gemm<T>(n, OP_N, P, OP_N, P, Pn, 0, as, lds);
    wg_gemm<T>(n, OP_N, P2, OP_N, P2, P2n, 0, st, lds);
    double2* q = P; P = Pn; Pn = q;
    q = P2; P2 = P2n; P2n = q;
    if (sp <= kSteinTol) { steps = k; break; }
  }
  wg_gemm<T>(n, OP_C, g, OP_N, X, U, 0, st, lds);
  wg_gemm<T>(n, OP_N, U, OP_C, g, Hb, 0, st, lds);        // H00-bar = g^H X g^H
  double trh = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) trh += gc(Hb)[i + (size_t)i * n].x;
  trh = block_sum(trh, red);
  if (left) {
    wg_gemm<T>(n, OP_N, Hb, OP_N, C, U, 0, st, lds);
    wg_gemm<T>(n, OP_N, U, OP_C, g, Cbar, 0, ac, lds);    // C-bar += H00b C g^H
    wg_gemm<T>(n, OP_C, Hb, OP_N, C, U, 0, st, lds);
    wg_gemm<T>(n, OP_N, U, OP_N, g, Cbar, 0, ac, lds);    // C-bar += H00b^H C g
  } else {
    wg_gemm<T>(n, OP_C, g, OP_N, C, U, 0, st, lds);
    wg_gemm<T>(n, OP_N, U, OP_N, Hb, Cbar, 0, ac, lds);   // C-bar += g^H C H00b
    wg_gemm<T>(n, OP_N, g, OP_N, C, U, 0, st, lds);
    wg_gemm<T>(n, OP_N, U, OP_C, Hb, Cbar, 0, ac, lds);   // C-bar += g C H00b^H
  }
  return trh;
}

template <int T, int NMAX>
__device__ void adjoint_energy(const AdjArgs& a, int e, double2* ws, double2* dH, ldsd* lds, ldsd* red) {
  const int n = a.n, N = a.nx;
  const size_t nn = (size_t)n * n;
  const double E = a.E[e];
  const double2 z = make_double2(E, a.eta_dev);
  const double2 one = make_double2(1, 0), mone = make_double2(-1, 0), zero = make_double2(0, 0);
  const int KS = a.ckpt > 0 ? a.ckpt : 0;  // segment length (0: every slab kept)
  const size_t NS = (size_t)adj_state_slabs(N, KS);
  const int ncp = KS ? (N - 2) / KS + 1 : 0;  // checkpoint slabs 0, KS, .. <= N-2
  double2* gv = ws;
  double2* glv = ws + NS * nn;
  double2* Sv = ws + 2 * NS * nn;
  double2* Gv = ws + 3 * NS * nn;
  double2* Gnv = Gv + (size_t)N * nn;
  double2* gbv = Gnv + (size_t)N * nn;
  double2* glbv = gbv + (size_t)N * nn;
  const bool keep = a.keep != 0 && KS == 0;
  double2* Rv = glbv + (size_t)N * nn;   // keep only
  double2* Rdv = Rv + (size_t)N * nn;    // keep only
  double2* Mv = Rdv + (size_t)N * nn;    // keep only
  double2* tb = glbv + (size_t)(keep ? 4 : 1) * N * nn;
  auto B_ = [&](double2* base, int l) { return base + (size_t)l * nn; };
  auto Tm = [&](int k) { return tb + (size_t)k * nn; };
  double2 *H00 = Tm(0), *W = Tm(1), *T0 = Tm(2), *T1 = Tm(3), *T2 = Tm(4), *SL = Tm(5), *SR = Tm(6);
  double2 *Dd = Tm(7), *Du = Tm(8), *V = Tm(9), *Q = Tm(10), *R = Tm(11), *Rd = Tm(12), *QGn = Tm(13);  // 10-13: Q, R, R d, Q G< when not kept
  double2 *gs = Tm(14), *Wb = Tm(15), *gsb = Tm(16), *Rb = Tm(17), *Rdb = Tm(18), *QGnb = Tm(19), *Qb = Tm(20);
  double2 *Gb = Tm(21), *Lb = Tm(22), *Gb2 = Tm(23), *Lb2 = Tm(24), *Mb = Tm(26), *Sb = Tm(27);  // 25: g S when not kept
  double2 *Ab = Tm(28), *Bb = Tm(29), *Yb = Tm(30), *Bm = Tm(31), *Y = Tm(32), *K = Tm(33), *Gs = Tm(34);
  double2 *GsL = Tm(35), *SRb = Tm(0);  // left surface function; Sigma_R-bar (after the contacts)
  double2* stmp = Tm(36);  // 11 blocks for the decimation
  const size_t so = (size_t)e * a.sig_stride;
  auto Hb = [&](int l) { return a.Hii + (size_t)l * nn; };
  auto Cb = [&](int l) { return a.H1i + (size_t)l * nn; };
  auto sR = [&](int l) -> const double2* { return a.sigR ? a.sigR + so + (size_t)l * nn : nullptr; };
  auto sL = [&](int l) -> const double2* { return a.sigL ? a.sigL + so + (size_t)l * nn : nullptr; };
  auto sRlo = [&](int l) -> const double2* { return a.sigR_lo ? a.sigR_lo + so + (size_t)l * nn : nullptr; };
  auto sRup = [&](int l) -> const double2* { return a.sigR_up ? a.sigR_up + so + (size_t)l * nn : nullptr; };
  auto sLup = [&](int l) -> const double2* { return a.sigL_up ? a.sigL_up + so + (size_t)l * nn : nullptr; };
  // state slot of slab l: itself (K = 0), its checkpoint, the last-slab slot or the segment
  auto st_ix = [&](int l) -> int {
    if (!KS) return l;
    if (l == N - 1) return ncp;
    if (l % KS == 0) return l / KS;
    return ncp + 1 + l % KS;
  };
  auto g = [&](int l) { return B_(gv, st_ix(l)); };
  auto gl = [&](int l) { return B_(glv, st_ix(l)); };
  auto Sst = [&](int l) { return B_(Sv, st_ix(l)); };
  auto Gf = [&](int l) { return l == N - 1 ? g(l) : B_(Gv, l); };    // G_ll
  auto Gnf = [&](int l) { return l == N - 1 ? gl(l) : B_(Gnv, l); };  // G<_ll
  const double wtr = a.w_tr ? a.w_tr[e] : 0.0;
  auto wn = [&](int l) { return a.w_n ? a.w_n[(size_t)e * N + l] : 0.0; };
  const Epi st = epi_store();
  auto acc = [&](const double2* D) { Epi x = epi_store(); x.D1 = D; return x; };
  Epi neg = epi_store();
  neg.sacc = -1.0;
  // Coupling operands of slab l: d = H1i_l + Sigma^R_{l,l-1} in dP, u = H1i_l^H +
  // Sigma^R_{l-1,l} in uP.  Without the off-diagonal blocks the GEMMs read H1i_l
  // itself: dP = H1i_l, and uP = H1i_l holding u^H (uC = 1 flips the GEMM op).
  const double2* dP = Dd;
  const double2* uP = Du;
  int uC = 0;
  auto couple = [&](int l) {
    if (a.sigR_lo) {
      wg_ew(n, Dd, 0, zero, Cb(l), one, sRlo(l), one, nullptr, one);
      dP = Dd;
    } else {
      dP = Cb(l);
    }
    if (a.sigR_up) {
      wg_ew(n, Du, 0, zero, sRup(l), one, nullptr, one, Cb(l), one);
      uP = Du;
      uC = 0;
    } else {
      uP = Cb(l);
      uC = 1;
    }
  };
  auto uo = [&](int op) { return op ^ uC; };
  int fails = 0, hitL = 0, hitR = 0;
  double Jv = 0.0;

  // ---- contacts (fixed)
  if (a.lead)
    wg_ew(n, H00, 0, zero, a.lead, one, nullptr, one, nullptr, one);
  else
    wg_ew(n, H00, 0, zero, Hb(0), one, sR(0), one, nullptr, one);
  const int itL = wg_sancho<T, NMAX>(n, E, a.eta_lead, a.dec_tol, a.dec_nmax, H00, Cb(0), 1, GsL, stmp, lds, red,
                                     hitL, fails);
  wg_gemm<T>(n, OP_N, Cb(0), OP_N, GsL, T0, 0, st, lds);
  wg_gemm<T>(n, OP_N, T0, OP_C, Cb(0), SL, 0, st, lds);
  if (a.lead)
    wg_ew(n, H00, 0, zero, a.lead + nn, one, nullptr, one, nullptr, one);
  else
    wg_ew(n, H00, 0, zero, Hb(N - 1), one, sR(N - 1), one, nullptr, one);
  const int itR = wg_sancho<T, NMAX>(n, E, a.eta_lead, a.dec_tol, a.dec_nmax, H00, Cb(N - 1), 0, Gs, stmp, lds,
                                     red, hitR, fails);
  wg_gemm<T>(n, OP_C, Cb(N - 1), OP_N, Gs, T0, 0, st, lds);
  wg_gemm<T>(n, OP_N, T0, OP_N, Cb(N - 1), SR, 0, st, lds);
  const double fL = ferm_dev((E - a.mu[0]) / (a.boltz * a.temp[0]));
  const double fR = ferm_dev((E - a.mu[1]) / (a.boltz * a.temp[1]));

  // ---- forward sweep, storing g, g<, S (slab N-1: G, G<)
  wg_ew(n, T1, 0, zero, Hb(0), one, sR(0), one, nullptr, one);
  wg_ew(n, W, 1, z, T1, mone, SL, mone, nullptr, one);  // z - (H_0 + sigR_0) - Sigma_L
  if (wg_inverse<T, NMAX>(n, W, g(0), lds)) ++fails;
  wg_ew(n, Sst(0), 0, zero, SL, make_double2(-fL, 0), sL(0), one, SL, make_double2(fL, 0));
  {
    double2* Mp = keep ? B_(Mv, 0) : T0;  // g S
    wg_gemm<T>(n, OP_N, g(0), OP_N, Sst(0), Mp, 0, st, lds);
    wg_gemm<T>(n, OP_N, Mp, OP_C, g(0), gl(0), 0, st, lds);
  }
  // one forward step: g_l, g<_l, S_l from g_{l-1}, g<_{l-1} (also the segment refill;
  // a singular slab is counted on the first pass only)
  auto fwd_step = [&](int l, bool count) {
    const bool last = (l == N - 1);
    couple(l);
    wg_gemm<T>(n, OP_N, dP, OP_N, g(l - 1), T0, 0, st, lds);  // B = d g_{l-1}
    Epi ew = epi_store();
    ew.diag = z; ew.sacc = -1.0; ew.D1 = Hb(l); ew.s1 = mone; ew.D2 = sR(l); ew.s2 = mone;
    if (last) {
      wg_ew(n, T1, 0, zero, sR(l), one, SR, one, nullptr, one);  // sigR_N + Sigma_R
      ew.D2 = T1;
    }
    wg_gemm<T>(n, OP_N, T0, uo(OP_N), uP, W, 1, ew, lds);         // z - H - Sigma^R - B u
    if (wg_inverse<T, NMAX>(n, W, g(l), lds) && count) ++fails;
    Epi es = epi_store();
    es.D1 = sL(l);
    if (sLup(l)) {
      wg_gemm<T>(n, OP_N, T0, OP_N, sLup(l), V, 0, st, lds);  // V = B s
      es.D2 = V; es.D3 = V; es.s3 = mone;
    }
    if (last) {
      wg_ew(n, T2, 0, zero, SR, make_double2(-fR, 0), sL(l), one, SR, make_double2(fR, 0));
      es.D1 = T2;
    }
    wg_gemm<T>(n, OP_N, dP, OP_N, gl(l - 1), T1, 0, st, lds);  // Y = d g<_{l-1}
    wg_gemm<T>(n, OP_N, T1, OP_C, dP, Sst(l), 0, es, lds);     // S
    double2* Mp = keep ? B_(Mv, l) : T0;  // g S
    wg_gemm<T>(n, OP_N, g(l), OP_N, Sst(l), Mp, 0, st, lds);
    wg_gemm<T>(n, OP_N, Mp, OP_C, g(l), gl(l), 0, st, lds);
  };
  for (int l = 1; l < N; ++l) fwd_step(l, true);
  // segment refill: the forward state of slab l available (K > 0); the segment buffer
  // holds one segment at a time, refilled from its checkpoint
  int loaded = -1;
  auto ensure = [&](int l) {
    if (!KS || l == N - 1 || l % KS == 0) return;
    const int c = l / KS;
    if (c == loaded) return;
    const int hi = c * KS + KS - 1 < N - 2 ? c * KS + KS - 1 : N - 2;
    for (int m = c * KS + 1; m <= hi; ++m) fwd_step(m, false);
    loaded = c;
  };
  if (KS) {  // the forward sweep leaves the segment of its last non-checkpoint slab <= N-2
    const int lm = (N - 2) % KS == 0 ? N - 3 : N - 2;
    loaded = lm >= 1 ? lm / KS : -1;
  }
  // tr = -Re Tr(KR Gn_N) - fR Re Tr(KR (G_N - G_N^H)),  KR = -(SR - SR^H)
  wg_ew(n, K, 0, zero, SR, mone, nullptr, one, SR, one);
  wg_ew(n, T1, 0, zero, g(N - 1), one, nullptr, one, g(N - 1), mone);
  {
    const double a1 = wg_retrace_prod(n, K, gl(N - 1), red);
    const double a2 = wg_retrace_prod(n, K, T1, red);
    const double trv = -a1 - fR * a2;
    Jv += wtr * trv;
    if (a.tr && threadIdx.x == 0) a.tr[e] = trv;
  }
  Jv += wn(N - 1) * wg_imtrace(n, gl(N - 1), red);

  // ---- backward sweep, storing G_ll and G<_ll
  for (int l = N - 2; l >= 0; --l) {
    ensure(l);
    const double2* s = sLup(l + 1);
    couple(l + 1);
    double2* Qp = keep ? B_(glbv, l) : Q;
    double2* Rp = keep ? B_(Rv, l) : R;
    double2* Rdp = keep ? B_(Rdv, l) : Rd;
    double2* QGnp = keep ? B_(gbv, l) : QGn;
    wg_gemm<T>(n, OP_N, g(l), uo(OP_N), uP, Qp, 0, st, lds);             // Q = g u
    wg_gemm<T>(n, OP_N, Qp, OP_N, Gf(l + 1), Rp, 0, st, lds);        // R = Q G
    wg_gemm<T>(n, OP_N, Rp, OP_N, dP, Rdp, 0, st, lds);              // R d
    wg_gemm<T>(n, OP_N, Rdp, OP_N, g(l), Gf(l), 0, acc(g(l)), lds);  // G_l = g + R d g
    wg_gemm<T>(n, OP_N, Qp, OP_N, Gnf(l + 1), QGnp, 0, st, lds);     // Q G<
    wg_gemm<T>(n, OP_N, QGnp, OP_C, Qp, T0, 0, acc(gl(l)), lds);     // g< + Q G< Q^H
    wg_gemm<T>(n, OP_N, Rdp, OP_N, gl(l), T1, 0, st, lds);           // R d g<
    if (s) {
      wg_gemm<T>(n, OP_N, g(l), OP_N, s, gs, 0, st, lds);            // g s
      wg_gemm<T>(n, OP_N, gs, OP_C, Rp, T2, 0, acc(T1), lds);        // W = g s R^H + R d g<
      wg_ew(n, Gnf(l), 0, zero, T0, one, T2, one, T2, mone);
    } else {
      wg_ew(n, Gnf(l), 0, zero, T0, one, T1, one, T1, mone);
    }
    Jv += wn(l) * wg_imtrace(n, Gnf(l), red);
  }

  // ---- pass A: reverse of the backward sweep, l = 0 .. N-2
  wg_ew(n, Gb, 0, zero, nullptr, one, nullptr, one, nullptr, one);
  wg_ew(n, Lb, 0, make_double2(0, wn(0)), nullptr, one, nullptr, one, nullptr, one);
  for (int l = 0; l < N - 1; ++l) {
    ensure(l);
    const double2* s = sLup(l + 1);
    double2* Tbar = dH + (size_t)(l + 1) * nn;
    couple(l + 1);
    const double2* Q = keep ? B_(glbv, l) : Tm(10);
    const double2* R = keep ? B_(Rv, l) : Tm(11);
    const double2* Rd = keep ? B_(Rdv, l) : Tm(12);
    const double2* QGn = keep ? B_(gbv, l) : Tm(13);
    if (!keep) {
      wg_gemm<T>(n, OP_N, g(l), uo(OP_N), uP, Tm(10), 0, st, lds);
      wg_gemm<T>(n, OP_N, Q, OP_N, Gf(l + 1), Tm(11), 0, st, lds);
      wg_gemm<T>(n, OP_N, R, OP_N, dP, Tm(12), 0, st, lds);
      wg_gemm<T>(n, OP_N, Q, OP_N, Gnf(l + 1), Tm(13), 0, st, lds);
    }
    wg_ew(n, Wb, 0, zero, Lb, one, nullptr, one, Lb, mone);          // Wb = Lb - Lb^H
    if (s) {
      wg_gemm<T>(n, OP_N, g(l), OP_N, s, gs, 0, st, lds);
      wg_gemm<T>(n, OP_N, Wb, OP_N, R, gsb, 0, st, lds);             // gsb = Wb R
      wg_gemm<T>(n, OP_C, Wb, OP_N, gs, Rb, 0, st, lds);             // Rb = Wb^H g s
    }
    wg_gemm<T>(n, OP_N, Wb, OP_C, gl(l), Rdb, 0, st, lds);           // Rdb = Wb g<^H
    wg_gemm<T>(n, OP_N, Lb, OP_N, Q, QGnb, 0, st, lds);              // QGnb = Lb Q
    wg_gemm<T>(n, OP_C, Lb, OP_N, QGn, Qb, 0, st, lds);              // Qb = Lb^H Q G< (last use of Q G<)
    wg_gemm<T>(n, OP_C, Rd, OP_N, Gb, B_(gbv, l), 0, acc(Gb), lds);  // gb_l = Gb + (R d)^H Gb
    wg_gemm<T>(n, OP_N, Gb, OP_C, g(l), Rdb, 0, acc(Rdb), lds);      // Rdb += Gb g^H
    if (s)
      wg_gemm<T>(n, OP_N, Rdb, OP_C, dP, Rb, 0, acc(Rb), lds);       // Rb += Rdb d^H
    else
      wg_gemm<T>(n, OP_N, Rdb, OP_C, dP, Rb, 0, st, lds);
    wg_gemm<T>(n, OP_C, R, OP_N, Rdb, Tbar, 0, acc(Tbar), lds);      // d-bar += R^H Rdb
    if (s) wg_gemm<T>(n, OP_N, gsb, OP_C, s, B_(gbv, l), 0, acc(B_(gbv, l)), lds);  // gb_l += gsb s^H
    wg_gemm<T>(n, OP_N, QGnb, OP_C, Gnf(l + 1), Qb, 0, acc(Qb), lds);  // Qb += QGnb G<^H
    wg_gemm<T>(n, OP_N, Rb, OP_C, Gf(l + 1), Qb, 0, acc(Qb), lds);     // Qb += Rb G^H
    {
      Epi dg = epi_store();
      dg.diag = make_double2(0, wn(l + 1));
      wg_gemm<T>(n, OP_C, Q, OP_N, QGnb, Lb2, 0, dg, lds);           // Lb_{l+1} = Q^H QGnb + i w I
    }
    wg_gemm<T>(n, OP_C, Q, OP_N, Rb, Gb2, 0, st, lds);               // Gb_{l+1} = Q^H Rb
    wg_gemm<T>(n, OP_N, Qb, uo(OP_C), uP, B_(gbv, l), 0, acc(B_(gbv, l)), lds);  // gb_l += Qb u^H
    wg_gemm<T>(n, OP_C, Qb, OP_N, g(l), Tbar, 0, acc(Tbar), lds);   // u-bar^H = Qb^H g
    wg_gemm<T>(n, OP_C, Rd, OP_N, Wb, B_(glbv, l), 0, acc(Lb), lds); // glb_l = Lb + (R d)^H Wb (after the last use of Q)
    double2* t = Gb; Gb = Gb2; Gb2 = t;
    t = Lb; Lb = Lb2; Lb2 = t;
  }
  // objective terms on the last slab: Lb += -w KR^H, Gb += -w fR (KR^H - KR)
  wg_ew(n, Lb, 0, zero, Lb, one, nullptr, one, K, make_double2(-wtr, 0));
  wg_ew(n, Gb, 0, zero, Gb, one, K, make_double2(wtr * fR, 0), K, make_double2(-wtr * fR, 0));

  // ---- pass B: reverse of the forward sweep, l = N-1 .. 0
  for (int l = N - 1; l >= 0; --l) {
    ensure(l);
    double2* gbl = (l == N - 1) ? Gb : B_(gbv, l);
    const double2* glbl = (l == N - 1) ? Lb : B_(glbv, l);
    const double2* Sl = Sst(l);
    const double2* M = keep ? B_(Mv, l) : Tm(25);
    if (!keep) wg_gemm<T>(n, OP_N, g(l), OP_N, Sl, Tm(25), 0, st, lds);  // M = g S
    wg_gemm<T>(n, OP_N, glbl, OP_N, g(l), Mb, 0, st, lds);           // Mb = glb g
    wg_gemm<T>(n, OP_C, glbl, OP_N, M, gbl, 0, acc(gbl), lds);       // gb += glb^H M
    wg_gemm<T>(n, OP_N, Mb, OP_C, Sl, gbl, 0, acc(gbl), lds);        // gb += Mb S^H
    wg_gemm<T>(n, OP_C, g(l), OP_N, Mb, Sb, 0, st, lds);             // Sb = g^H Mb
    wg_gemm<T>(n, OP_C, g(l), OP_N, gbl, T0, 0, st, lds);
    wg_gemm<T>(n, OP_N, T0, OP_C, g(l), Ab, 0, neg, lds);            // Ab = -g^H gb g^H
    {
      double tr_ab = 0.0;
      for (int i = threadIdx.x; i < n; i += blockDim.x) tr_ab += gc(Ab)[i + (size_t)i * n].x;
      tr_ab = block_sum(tr_ab, red);
      if (threadIdx.x == 0) a.dU[(size_t)e * N + l] = -tr_ab;
    }
    if (!a.lead && l == N - 1) {
      // Sigma_R-bar = -Abar + fR (Sbar^H - Sbar) + w_tr (G<^H - G<) + 2 w_tr fR (G^H - G)
      wg_ew(n, SRb, 0, zero, Ab, mone, Sb, make_double2(-fR, 0), Sb, make_double2(fR, 0));
      wg_ew(n, SRb, 0, zero, SRb, one, gl(N - 1), make_double2(-wtr, 0), gl(N - 1), make_double2(wtr, 0));
      wg_ew(n, SRb, 0, zero, SRb, one, g(N - 1), make_double2(-2.0 * wtr * fR, 0), g(N - 1),
            make_double2(2.0 * wtr * fR, 0));
    }
    if (l >= 1) {
      ensure(l - 1);  // may replace slab l's segment: slab l's own blocks are no longer read
      const double2* s = sLup(l);
      double2* Tbar = dH + (size_t)l * nn;
      couple(l);
      wg_gemm<T>(n, OP_N, dP, OP_N, g(l - 1), Bm, 0, st, lds);       // B = d g_{l-1}
      wg_gemm<T>(n, OP_N, Ab, uo(OP_C), uP, Bb, 0, neg, lds);            // Bb = -Ab u^H
      Epi na = epi_store();
      na.sacc = -1.0; na.D1 = Tbar;
      wg_gemm<T>(n, OP_C, Ab, OP_N, Bm, Tbar, 0, na, lds);           // u-bar^H = -Ab^H B
      wg_gemm<T>(n, OP_N, dP, OP_N, gl(l - 1), Y, 0, st, lds);       // Y = d g<_{l-1}
      wg_gemm<T>(n, OP_N, Sb, OP_N, dP, Yb, 0, st, lds);             // Yb = Sb d
      wg_gemm<T>(n, OP_C, Sb, OP_N, Y, Tbar, 0, acc(Tbar), lds);     // d-bar += Sb^H Y
      wg_gemm<T>(n, OP_N, Yb, OP_C, gl(l - 1), Tbar, 0, acc(Tbar), lds);  // d-bar += Yb g<^H
      wg_gemm<T>(n, OP_C, dP, OP_N, Yb, B_(glbv, l - 1), 0, acc(B_(glbv, l - 1)), lds);
      if (s) {
        wg_ew(n, V, 0, zero, Sb, one, nullptr, one, Sb, mone);       // Vb = Sb - Sb^H
        wg_gemm<T>(n, OP_N, V, OP_C, s, Bb, 0, acc(Bb), lds);        // Bb += Vb s^H
      }
      wg_gemm<T>(n, OP_N, Bb, OP_C, g(l - 1), Tbar, 0, acc(Tbar), lds);  // d-bar += Bb g^H
      wg_gemm<T>(n, OP_C, dP, OP_N, Bb, B_(gbv, l - 1), 0, acc(B_(gbv, l - 1)), lds);
    }
  }
  int lead_flags = 0;
  if (!a.lead) {
    // contacts that repeat the edge slabs: Ab, Sb hold slab 0's adjoints
    double2* SLb = Tm(1);
    wg_ew(n, SLb, 0, zero, Ab, mone, Sb, make_double2(-fL, 0), Sb, make_double2(fL, 0));
    int stL = 0, stR = 0;
    const double hL = lead_adjoint<T>(n, true, Cb(0), GsL, SLb, dH, Tm(2), a.dec_nmax, stL, lds, red);
    const double hR = lead_adjoint<T>(n, false, Cb(N - 1), Gs, SRb, dH + (size_t)(N - 1) * nn, Tm(2), a.dec_nmax,
                                      stR, lds, red);
    if (threadIdx.x == 0) {
      a.dU[(size_t)e * N] += hL;
      a.dU[(size_t)e * N + N - 1] += hR;
    }
    lead_flags = (stL < 0 ? 4 : 0) | (stR < 0 ? 8 : 0);
  }
  if (threadIdx.x == 0) {
    a.J[e] = Jv;
    if (a.info) {
      a.info[4 * (size_t)e + 0] = itL;
      a.info[4 * (size_t)e + 1] = itR;
      a.info[4 * (size_t)e + 2] = hitL | (hitR << 1) | lead_flags;
      a.info[4 * (size_t)e + 3] = fails;
    }
  }
  __syncthreads();
}

template <int T, int NMAX>
__global__ __launch_bounds__(Cfg<T>::NTH, Cfg<T>::OCC) void rgf_adjoint(AdjArgs a) {
  __shared__ double lds[SweepLds<T, NMAX>::DOUBLES];
  double2* ws = a.ws + (size_t)blockIdx.x * (size_t)a.ws_slot;
  double2* dH = a.dH + (size_t)blockIdx.x * a.nx * (size_t)a.n * a.n;
  for (int e = blockIdx.x; e < a.nE; e += gridDim.x) adjoint_energy<T, NMAX>(a, e, ws, dH, (ldsd*)lds, (ldsd*)lds);
}

// Sum the per-slot coupling adjoints: out[l] = sum_s dH[s][l] for l = 1..nx-2, and at
// the edges when `edges` (contacts that repeat the edge slabs), else zero there.
__global__ void k_sum_slots(const double2* __restrict__ dH, int slots, int nx, int n, int edges,
                            double2* __restrict__ out) {
  const size_t nn = (size_t)n * n, tot = (size_t)nx * nn;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
    const size_t l = i / nn;
    double2 s = make_double2(0, 0);
    if (edges || (l >= 1 && l + 1 < (size_t)nx))
      for (int k = 0; k < slots; ++k) {
        const double2 v = dH[(size_t)k * tot + i];
        s.x += v.x;
        s.y += v.y;
      }
    out[i] = s;
  }
}

}  // namespace rgf
