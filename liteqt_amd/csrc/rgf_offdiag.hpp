// rgf_offdiag.hpp -- the per-energy RGF solve of the consistent (physical)
// device with a block-tridiagonal scattering self-energy: diagonal blocks
// Sigma^R_ll, Sigma^<_ll (as green_RGF_RS, src/green_rgf.f90:3952) plus the
// first off-diagonal blocks Sigma^R_{l,l-1}, Sigma^R_{l-1,l}, Sigma^<_{l-1,l}
// (Sigma^<_{l,l-1} = -(Sigma^<_{l-1,l})^H).  With A = z - H - Sigma^R and
//   d_l = H_{l,l-1} + Sigma^R_{l,l-1},  u_l = H_{l-1,l} + Sigma^R_{l-1,l},
//   s_l = Sigma^<_{l-1,l}
// the recursion is
//   forward   g_l  = (z - H_ll - Sigma^R_ll - d_l g_{l-1} u_l)^-1
//             g<_l = g_l [Sigma^<_ll + d_l g<_{l-1} d_l^H + V - V^H] g_l^H,  V = d_l g_{l-1} s_l
//   backward  (pair l, l+1; Q = g_l u, R = Q G_{l+1} = G_{l,l+1})
//             G_{l+1,l}  = G_{l+1} d g_l
//             G<_{l,l+1} = (g<_l d^H + g_l s) G_{l+1}^H + Q G<_{l+1}
//             G_ll       = g_l + R d g_l
//             G<_ll      = g<_l + Q G<_{l+1} Q^H + (Z - Z^H) + (Y - Y^H),
//                          Z = R d g<_l,  Y = g_l s R^H
// which reduces to the physical-mode sweep of rgf_sweep.hpp when the
// off-diagonal blocks vanish.  The contacts, tr, tre, cur and the output set
// follow the physical mode.
//
// With a.scba set, each block of Sigma that was given is refreshed in place
// once its Green's function block is final:
//   Sigma <- (1 - mix) Sigma + mix * D * G     (elastic scattering, same energy)
// Sigma^R from G^R, Sigma^< from G^<, diagonal and off-diagonal alike.  A block
// is only rewritten after the last read of its old value for that energy, and
// energies are independent, so one launch is one SCBA iteration.
#pragma once
#include "rgf_sweep.hpp"

namespace rgf {

constexpr int kTmpBlocksOd = 36;

template <int T, int NMAX>
__device__ void solve_energy_od(const SweepArgs& a, int e, double2* ws, ldsd* lds, ldsd* red) {
  const int n = a.n, nx = a.nx, N = nx;
  const size_t nn = (size_t)n * n;
  const double E = a.E[e];
  const double2 z = make_double2(E, a.eta_dev);
  const double2 one = make_double2(1, 0), mone = make_double2(-1, 0), zero = make_double2(0, 0);
  double2* gv = ws;
  double2* gnv = ws + (size_t)nx * nn;
  double2* tb = ws + 2 * (size_t)nx * nn;
  auto Tm = [&](int k) { return tb + (size_t)k * nn; };
  double2 *H00 = Tm(0), *W = Tm(1), *SL = Tm(2), *SR = Tm(3), *T0 = Tm(4), *T1 = Tm(5), *T2 = Tm(6);
  double2 *Dd = Tm(7), *Du = Tm(8), *V = Tm(9), *G = Tm(10), *Gn = Tm(11), *Q = Tm(12), *R = Tm(13);
  double2 *Glo = Tm(14), *QGn = Tm(15), *gs = Tm(16), *M = Tm(17), *Gnn = Tm(18), *Rd = Tm(19), *Z = Tm(20);
  double2 *W1 = Tm(21), *K1 = Tm(22), *Gs = Tm(23);
  double2* stmp = Tm(25);  // 11 blocks for the decimation
  const size_t so = (size_t)e * a.sig_stride;
  auto Hb = [&](int l) { return a.Hii + (size_t)l * nn; };
  auto Cb = [&](int l) { return a.H1i + (size_t)l * nn; };
  auto sR = [&](int l) -> const double2* { return a.sigR ? a.sigR + so + (size_t)l * nn : nullptr; };
  auto sL = [&](int l) -> const double2* { return a.sigL ? a.sigL + so + (size_t)l * nn : nullptr; };
  auto sRlo = [&](int l) -> const double2* { return a.sigR_lo ? a.sigR_lo + so + (size_t)l * nn : nullptr; };
  auto sRup = [&](int l) -> const double2* { return a.sigR_up ? a.sigR_up + so + (size_t)l * nn : nullptr; };
  auto sLup = [&](int l) -> const double2* { return a.sigL_up ? a.sigL_up + so + (size_t)l * nn : nullptr; };
  auto ob = [&](double2* arr, int l) -> double2* { return arr ? arr + ((size_t)e * nx + l) * nn : nullptr; };
  auto g = [&](int l) { return gv + (size_t)l * nn; };
  auto gn = [&](int l) { return gnv + (size_t)l * nn; };
  const Epi st = epi_store();
  auto acc = [&](const double2* D) { Epi x = epi_store(); x.D1 = D; return x; };
  const double2 keep = make_double2(1.0 - a.scba_mix, 0), take = make_double2(a.scba_mix * a.scba_D, 0);
  // in-place SCBA refresh of one Sigma block from its Green's function block
  auto refresh = [&](const double2* sig, const double2* Gblk) {
    if (a.scba && sig) wg_ew(n, const_cast<double2*>(sig), 0, zero, sig, keep, Gblk, take, nullptr, one);
  };
  // d = H_{l,l-1} + Sigma^R_{l,l-1},  u = H_{l,l-1}^H + Sigma^R_{l-1,l}
  auto couple = [&](int l) {
    wg_ew(n, Dd, 0, zero, Cb(l), one, sRlo(l), one, nullptr, one);
    wg_ew(n, Du, 0, zero, sRup(l), one, nullptr, one, Cb(l), one);
  };
  int fails = 0, hitL = 0, hitR = 0;

  // ---- contacts (leads repeat the edge slabs with their diagonal Sigma^R)
  wg_ew(n, H00, 0, zero, Hb(0), one, sR(0), one, nullptr, one);
  const int itL = wg_sancho<T, NMAX>(n, E, a.eta_lead, a.dec_tol, a.dec_nmax, H00, Cb(0), 1, Gs, stmp, lds, red,
                                     hitL, fails);
  wg_gemm<T>(n, OP_N, Cb(0), OP_N, Gs, T0, 0, st, lds);
  wg_gemm<T>(n, OP_N, T0, OP_C, Cb(0), SL, 0, st, lds);
  wg_ew(n, H00, 0, zero, Hb(N - 1), one, sR(N - 1), one, nullptr, one);
  const int itR = wg_sancho<T, NMAX>(n, E, a.eta_lead, a.dec_tol, a.dec_nmax, H00, Cb(N - 1), 0, Gs, stmp, lds,
                                     red, hitR, fails);
  wg_gemm<T>(n, OP_C, Cb(N - 1), OP_N, Gs, T0, 0, st, lds);
  wg_gemm<T>(n, OP_N, T0, OP_N, Cb(N - 1), SR, 0, st, lds);
  const double fL = ferm_dev((E - a.mu[0]) / (a.boltz * a.temp[0]));
  const double fR = ferm_dev((E - a.mu[1]) / (a.boltz * a.temp[1]));

  // ---- forward sweep
  wg_ew(n, T1, 0, zero, Hb(0), one, sR(0), one, nullptr, one);
  wg_ew(n, W, 1, z, T1, mone, SL, mone, nullptr, one);
  if (wg_inverse<T, NMAX>(n, W, g(0), lds)) ++fails;
  wg_ew(n, T2, 0, zero, SL, make_double2(-fL, 0), sL(0), one, SL, make_double2(fL, 0));
  wg_gemm<T>(n, OP_N, g(0), OP_N, T2, T0, 0, st, lds);
  wg_gemm<T>(n, OP_N, T0, OP_C, g(0), gn(0), 0, st, lds);
  for (int l = 1; l < N; ++l) {
    const bool last = (l == N - 1);
    double2* gg = last ? G : g(l);
    double2* gng = last ? Gn : gn(l);
    couple(l);
    wg_gemm<T>(n, OP_N, Dd, OP_N, g(l - 1), T0, 0, st, lds);   // d g_{l-1}
    Epi ew = epi_store();
    ew.diag = z; ew.sacc = -1.0; ew.D1 = Hb(l); ew.s1 = mone; ew.D2 = sR(l); ew.s2 = mone;
    if (last) {
      wg_ew(n, T1, 0, zero, sR(l), one, SR, one, nullptr, one);
      ew.D2 = T1;
    }
    wg_gemm<T>(n, OP_N, T0, OP_N, Du, W, 1, ew, lds);          // z - H - Sigma^R - d g u
    if (wg_inverse<T, NMAX>(n, W, gg, lds)) ++fails;
    Epi es = epi_store();
    es.D1 = sL(l);
    if (sLup(l)) {
      wg_gemm<T>(n, OP_N, T0, OP_N, sLup(l), V, 0, st, lds);   // V = d g s
      es.D2 = V; es.D3 = V; es.s3 = mone;
    }
    if (last) {
      wg_ew(n, T2, 0, zero, SR, make_double2(-fR, 0), sL(l), one, SR, make_double2(fR, 0));
      es.D1 = T2;
    }
    wg_gemm<T>(n, OP_N, Dd, OP_N, gn(l - 1), T1, 0, st, lds);  // d g<_{l-1}
    wg_gemm<T>(n, OP_N, T1, OP_C, Dd, T2, 0, es, lds);         // + Sigma^< + V - V^H
    wg_gemm<T>(n, OP_N, gg, OP_N, T2, T0, 0, st, lds);
    wg_gemm<T>(n, OP_N, T0, OP_C, gg, gng, 0, st, lds);
  }

  // ---- last slab outputs and tr
  const int L1 = N - 1;
  double2* Gp = W;  // G^> (W is free outside the forward sweep)
  wg_ew(n, Gp, 0, zero, Gn, one, G, one, G, mone);
  wg_ew(n, T0, 0, zero, SR, mone, nullptr, one, SR, one);  // i Gamma_R
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
  refresh(sR(L1), G);
  refresh(sL(L1), Gn);

  // ---- backward sweep
  const bool want_lo = a.GGi1 || a.GRi1_lo || (a.scba && a.sigR_lo);
  for (int l = N - 2; l >= 0; --l) {
    const double2* s = sLup(l + 1);
    couple(l + 1);
    wg_gemm<T>(n, OP_N, g(l), OP_N, Du, Q, 0, st, lds);         // Q = g u
    wg_gemm<T>(n, OP_N, Q, OP_N, G, R, 0, st, lds);             // R = G_{l,l+1}
    if (want_lo) {
      wg_gemm<T>(n, OP_N, G, OP_N, Dd, T0, 0, st, lds);
      wg_gemm<T>(n, OP_N, T0, OP_N, g(l), Glo, 0, st, lds);     // G_{l+1,l}
    }
    wg_gemm<T>(n, OP_N, Q, OP_N, Gn, QGn, 0, st, lds);          // Q G<
    Epi em = epi_store();
    if (s) {
      wg_gemm<T>(n, OP_N, g(l), OP_N, s, gs, 0, st, lds);       // g s
      em.D1 = gs;
    }
    wg_gemm<T>(n, OP_N, gn(l), OP_C, Dd, M, 0, em, lds);        // g< d^H + g s
    wg_gemm<T>(n, OP_N, M, OP_C, G, Gnn, 0, acc(QGn), lds);     // G<_{l,l+1}
    if (a.GRi1) wg_ew(n, ob(a.GRi1, l), 0, zero, R, one, nullptr, one, nullptr, one);
    if (a.GLi1) wg_ew(n, ob(a.GLi1, l), 0, zero, Gnn, one, nullptr, one, nullptr, one);
    if (a.GGi1) wg_ew(n, ob(a.GGi1, l), 0, zero, Gnn, one, R, one, Glo, mone);
    if (a.GRi1_lo) wg_ew(n, ob(a.GRi1_lo, l), 0, zero, Glo, one, nullptr, one, nullptr, one);
    if (a.cur) {
      // Re(H_{l,l+1} G<_{l+1,l}) - Re(H_{l+1,l} G<_{l,l+1}),  G<_{l+1,l} = -(G<_{l,l+1})^H
      wg_gemm<T>(n, OP_N, Gnn, OP_N, Cb(l + 1), K1, 0, st, lds);
      Epi ec = epi_store();
      ec.sacc = -1.0; ec.D3 = K1; ec.s3 = mone; ec.real_out = 1;
      wg_gemm<T>(n, OP_N, Cb(l + 1), OP_N, Gnn, a.cur + ((size_t)e * nx + l) * nn, 0, ec, lds);
    }
    refresh(sRup(l + 1), R);
    refresh(sRlo(l + 1), Glo);
    refresh(s, Gnn);
    wg_gemm<T>(n, OP_N, R, OP_N, Dd, Rd, 0, st, lds);           // R d
    wg_gemm<T>(n, OP_N, Rd, OP_N, g(l), G, 0, acc(g(l)), lds);  // G_ll = g + R d g
    wg_gemm<T>(n, OP_N, QGn, OP_C, Q, T0, 0, acc(gn(l)), lds);  // g< + Q G< Q^H
    wg_gemm<T>(n, OP_N, Rd, OP_N, gn(l), Z, 0, st, lds);        // Z = R d g<
    if (s) wg_gemm<T>(n, OP_N, gs, OP_C, R, W1, 0, acc(Z), lds);  // Z + g s R^H
    const double2* ZY = s ? W1 : Z;
    wg_ew(n, Gn, 0, zero, T0, one, ZY, one, ZY, mone);          // G<_ll
    if (a.ldos) wg_ew(n, ob(a.ldos, l), 0, zero, G, one, nullptr, one, nullptr, one);
    if (a.ndens) wg_ew(n, ob(a.ndens, l), 0, zero, Gn, one, nullptr, one, nullptr, one);
    if (a.pdens) wg_ew(n, ob(a.pdens, l), 0, zero, Gn, one, G, one, G, mone);
    if (a.dens_n) { double v = wg_imtrace(n, Gn, red); if (threadIdx.x == 0) a.dens_n[(size_t)e * nx + l] = v; }
    if (a.dens_p) {
      double v = wg_imtrace(n, Gn, red) + 2.0 * wg_imtrace(n, G, red);
      if (threadIdx.x == 0) a.dens_p[(size_t)e * nx + l] = -v;
    }
    refresh(sR(l), G);
    refresh(sL(l), Gn);
  }

  // ---- tre (left contact)
  wg_ew(n, T0, 0, zero, Gn, one, G, one, G, mone);         // G^>_11
  wg_ew(n, T1, 0, zero, SL, mone, nullptr, one, SL, one);  // i Gamma_L
  {
    const double a1 = wg_retrace_prod(n, T1, T0, red);
    const double a2 = wg_retrace_prod(n, T1, Gn, red);
    if (a.tre && threadIdx.x == 0) a.tre[e] = fL * a1 - (fL - 1.0) * a2;
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
__global__ __launch_bounds__(Cfg<T>::NTH, Cfg<T>::OCC) void rgf_sweep_od(SweepArgs a) {
  __shared__ double lds[SweepLds<T, NMAX>::DOUBLES];
  double2* ws = a.ws + (size_t)blockIdx.x * (size_t)a.ws_slot;
  for (int e = blockIdx.x; e < a.nE; e += gridDim.x) solve_energy_od<T, NMAX>(a, e, ws, (ldsd*)lds, (ldsd*)lds);
}

}  // namespace rgf
