// rgf_adjoint.hpp -- reverse-mode (adjoint) RGF: gradients of
//   J(E) = w_tr(E) * tr(E) + sum_l w_n(E,l) * Im Tr G^<_ll(E)
// with respect to a potential U_l on the diagonal of every slab and to the
// device-internal coupling blocks H1i(:,:,l), l = 2..nx-1 (1-based), for the
// consistent (physical) device with the contact self-energies held fixed.
// tr is the Meir-Wingreen current spectrum into the right contact
// (src/green_rgf.f90:4058-4066): with f_L = 1, f_R = 0 it is T(E); with Fermi
// factors and weights dE*e/h it gives dI/dU_l.  Densities are the integrands of
// src/green_rgf.f90:3919-3945.
//
// Per energy a workgroup runs the forward sweep (stores g_l, g^<_l, s_l), the
// backward sweep (stores G_ll, G^<_ll), the reverse of the backward sweep
// (l = 1..nx-1, accumulating adjoints of g_l and g^<_l), the reverse of the
// right boundary and the reverse of the forward sweep (l = nx-1..1), where
// dJ/dU_l = -Re Tr(Abar_l) and the coupling adjoints are accumulated.  The
// adjoint convention is Mbar = dJ/dRe M + i dJ/dIm M.
#pragma once
#include "rgf_sweep.hpp"

namespace rgf {

constexpr int kAdjTmp = 40;

struct AdjArgs {
  int n, nx, nE, dec_nmax;
  double eta_dev, eta_lead, dec_tol, boltz;
  double temp[2], mu[2];
  const double* E;
  const double2* Hii;
  const double2* H1i;
  const double2* sigL;
  const double2* sigR;
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
};

template <int T, int NMAX>
__device__ void adjoint_energy(const AdjArgs& a, int e, double2* ws, double2* dH, ldsd* lds, ldsd* red) {
  const int n = a.n, N = a.nx;
  const size_t nn = (size_t)n * n;
  const double E = a.E[e];
  const double2 z = make_double2(E, a.eta_dev);
  const double2 one = make_double2(1, 0), mone = make_double2(-1, 0), zero = make_double2(0, 0);
  double2* g = ws;
  double2* gl = ws + (size_t)N * nn;
  double2* sv = ws + 2 * (size_t)N * nn;
  double2* G = ws + 3 * (size_t)N * nn;
  double2* Gn = ws + 4 * (size_t)N * nn;
  double2* gb = ws + 5 * (size_t)N * nn;
  double2* glb = ws + 6 * (size_t)N * nn;
  double2* tb = ws + 7 * (size_t)N * nn;
  auto B = [&](double2* base, int l) { return base + (size_t)l * nn; };
  auto Tm = [&](int k) { return tb + (size_t)k * nn; };
  double2 *H00 = Tm(0), *W = Tm(1), *T0 = Tm(2), *T1 = Tm(3), *SL = Tm(4), *SR = Tm(5);
  double2 *X = Tm(6), *P = Tm(7), *R = Tm(8), *Q = Tm(9), *Wm = Tm(10);
  double2 *Gb = Tm(11), *Lb = Tm(12), *Gb2 = Tm(13), *Lb2 = Tm(14), *Rb = Tm(15), *Qb = Tm(16);
  double2 *Xb = Tm(17), *Wb = Tm(18), *Pb = Tm(19), *M = Tm(20), *Mb = Tm(21), *sb = Tm(22), *Ab = Tm(23);
  double2 *K = Tm(24), *Kb = Tm(25), *Nn = Tm(26), *Nb = Tm(27), *Gs = Tm(28), *stmp = Tm(29);
  auto Hb = [&](int l) { return a.Hii + (size_t)l * nn; };
  auto Cb = [&](int l) { return a.H1i + (size_t)l * nn; };
  auto sR = [&](int l) -> const double2* { return a.sigR ? a.sigR + (size_t)e * a.sig_stride + (size_t)l * nn : nullptr; };
  auto sL = [&](int l) -> const double2* { return a.sigL ? a.sigL + (size_t)e * a.sig_stride + (size_t)l * nn : nullptr; };
  const double wtr = a.w_tr ? a.w_tr[e] : 0.0;
  auto wn = [&](int l) { return a.w_n ? a.w_n[(size_t)e * N + l] : 0.0; };
  const Epi st = epi_store();
  auto acc = [&](const double2* D) { Epi x = epi_store(); x.D1 = D; return x; };
  int fails = 0, hitL = 0, hitR = 0;
  double Jv = 0.0;

  // ---- contacts (fixed)
  if (a.lead)
    wg_ew(n, H00, 0, zero, a.lead, one, nullptr, one, nullptr, one);
  else
    wg_ew(n, H00, 0, zero, Hb(0), one, sR(0), one, nullptr, one);
  const int itL = wg_sancho<T, NMAX>(n, E, a.eta_lead, a.dec_tol, a.dec_nmax, H00, Cb(0), 1, Gs, stmp, lds, red,
                                     hitL, fails);
  wg_gemm<T>(n, OP_N, Cb(0), OP_N, Gs, T0, 0, st, lds);
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

  // ---- forward sweep, storing g, g^<, s
  wg_ew(n, T1, 0, zero, Hb(0), one, sR(0), one, nullptr, one);
  wg_ew(n, W, 1, z, T1, mone, SL, mone, nullptr, one);  // z - (H_0 + sigR_0) - Sigma_L
  if (wg_inverse<T, NMAX>(n, W, B(g, 0), lds)) ++fails;
  wg_ew(n, B(sv, 0), 0, zero, SL, make_double2(-fL, 0), sL(0), one, SL, make_double2(fL, 0));
  wg_gemm<T>(n, OP_N, B(g, 0), OP_N, B(sv, 0), T0, 0, st, lds);
  wg_gemm<T>(n, OP_N, T0, OP_C, B(g, 0), B(gl, 0), 0, st, lds);
  for (int l = 1; l < N; ++l) {
    const bool last = (l == N - 1);
    wg_gemm<T>(n, OP_N, Cb(l), OP_N, B(g, l - 1), T0, 0, st, lds);
    Epi ew = epi_store();
    ew.diag = z; ew.sacc = -1.0; ew.D1 = Hb(l); ew.s1 = mone; ew.D2 = sR(l); ew.s2 = mone;
    if (last) {
      wg_ew(n, T1, 0, zero, sR(l), one, SR, one, nullptr, one);  // sigR_N + Sigma_R
      ew.D2 = T1;
    }
    wg_gemm<T>(n, OP_N, T0, OP_C, Cb(l), W, 1, ew, lds);
    if (wg_inverse<T, NMAX>(n, W, last ? B(G, l) : B(g, l), lds)) ++fails;
    wg_gemm<T>(n, OP_N, Cb(l), OP_N, B(gl, l - 1), T0, 0, st, lds);
    Epi es = epi_store();
    es.D1 = sL(l);
    if (last) { es.D2 = SR; es.s2 = make_double2(-fR, 0); es.D3 = SR; es.s3 = make_double2(fR, 0); }
    wg_gemm<T>(n, OP_N, T0, OP_C, Cb(l), B(sv, l), 0, es, lds);
    double2* gg = last ? B(G, l) : B(g, l);
    wg_gemm<T>(n, OP_N, gg, OP_N, B(sv, l), T0, 0, st, lds);
    wg_gemm<T>(n, OP_N, T0, OP_C, gg, last ? B(Gn, l) : B(gl, l), 0, st, lds);
  }
  // tr = -Re Tr(KR Gn_N) - fR Re Tr(KR (G_N - G_N^H)),  KR = -(SR - SR^H)
  wg_ew(n, K, 0, zero, SR, mone, nullptr, one, SR, one);
  wg_ew(n, T1, 0, zero, B(G, N - 1), one, nullptr, one, B(G, N - 1), mone);
  {
    const double a1 = wg_retrace_prod(n, K, B(Gn, N - 1), red);
    const double a2 = wg_retrace_prod(n, K, T1, red);
    const double trv = -a1 - fR * a2;
    Jv += wtr * trv;
    if (a.tr && threadIdx.x == 0) a.tr[e] = trv;
  }
  Jv += wn(N - 1) * wg_imtrace(n, B(Gn, N - 1), red);

  // ---- backward sweep, storing G_ll and G^<_ll
  for (int l = N - 2; l >= 0; --l) {
    const double2* Tc = Cb(l + 1);
    wg_gemm<T>(n, OP_N, B(g, l), OP_C, Tc, X, 0, st, lds);
    wg_gemm<T>(n, OP_N, X, OP_N, B(G, l + 1), P, 0, st, lds);
    wg_gemm<T>(n, OP_N, P, OP_N, Tc, R, 0, st, lds);
    wg_gemm<T>(n, OP_N, R, OP_N, B(g, l), B(G, l), 0, acc(B(g, l)), lds);
    wg_gemm<T>(n, OP_N, X, OP_N, B(Gn, l + 1), Q, 0, st, lds);
    wg_gemm<T>(n, OP_N, Q, OP_C, X, T0, 0, acc(B(gl, l)), lds);
    wg_gemm<T>(n, OP_N, R, OP_N, B(gl, l), T1, 0, acc(T0), lds);
    wg_gemm<T>(n, OP_N, B(gl, l), OP_C, Tc, Wm, 0, st, lds);
    wg_gemm<T>(n, OP_N, Wm, OP_C, P, B(Gn, l), 0, acc(T1), lds);
    Jv += wn(l) * wg_imtrace(n, B(Gn, l), red);
  }

  // ---- reverse of the backward sweep, l = 0 .. N-2
  wg_ew(n, Gb, 0, zero, nullptr, one, nullptr, one, nullptr, one);
  wg_ew(n, Lb, 0, make_double2(0, wn(0)), nullptr, one, nullptr, one, nullptr, one);
  for (int l = 0; l < N - 1; ++l) {
    const double2* Tc = Cb(l + 1);
    double2* Tbar = dH + (size_t)(l + 1) * nn;
    wg_gemm<T>(n, OP_N, B(g, l), OP_C, Tc, X, 0, st, lds);
    wg_gemm<T>(n, OP_N, X, OP_N, B(G, l + 1), P, 0, st, lds);
    wg_gemm<T>(n, OP_N, P, OP_N, Tc, R, 0, st, lds);
    wg_gemm<T>(n, OP_N, X, OP_N, B(Gn, l + 1), Q, 0, st, lds);
    wg_gemm<T>(n, OP_N, B(gl, l), OP_C, Tc, Wm, 0, st, lds);
    wg_gemm<T>(n, OP_C, R, OP_N, Gb, B(gb, l), 0, acc(Gb), lds);      // gb_l = Gb + R^H Gb
    wg_gemm<T>(n, OP_N, Gb, OP_C, B(g, l), Rb, 0, st, lds);            // Rb = Gb g^H
    wg_gemm<T>(n, OP_N, Lb, OP_C, B(gl, l), Rb, 0, acc(Rb), lds);      //    + Lb gl^H
    wg_gemm<T>(n, OP_C, R, OP_N, Lb, B(glb, l), 0, acc(Lb), lds);     // glb_l = Lb + R^H Lb
    wg_gemm<T>(n, OP_N, Lb, OP_N, X, Qb, 0, st, lds);                  // Qb = Lb X
    wg_gemm<T>(n, OP_C, Lb, OP_N, Q, Xb, 0, st, lds);                  // Xb = Lb^H Q
    wg_gemm<T>(n, OP_N, Lb, OP_N, P, Wb, 0, st, lds);                  // Wb = Lb P
    wg_gemm<T>(n, OP_C, Lb, OP_N, Wm, Pb, 0, st, lds);                 // Pb = Lb^H W
    wg_gemm<T>(n, OP_N, Wb, OP_N, Tc, B(glb, l), 0, acc(B(glb, l)), lds);
    wg_gemm<T>(n, OP_C, Wb, OP_N, B(gl, l), Tbar, 0, acc(Tbar), lds);
    wg_gemm<T>(n, OP_N, Qb, OP_C, B(Gn, l + 1), Xb, 0, acc(Xb), lds);
    {
      Epi d = epi_store();
      d.diag = make_double2(0, wn(l + 1));
      wg_gemm<T>(n, OP_C, X, OP_N, Qb, Lb2, 0, d, lds);                // Lb_{l+1} = X^H Qb + i w I
    }
    wg_gemm<T>(n, OP_N, Rb, OP_C, Tc, Pb, 0, acc(Pb), lds);
    wg_gemm<T>(n, OP_C, P, OP_N, Rb, Tbar, 0, acc(Tbar), lds);
    wg_gemm<T>(n, OP_N, Pb, OP_C, B(G, l + 1), Xb, 0, acc(Xb), lds);
    wg_gemm<T>(n, OP_C, X, OP_N, Pb, Gb2, 0, st, lds);                 // Gb_{l+1} = X^H Pb
    wg_gemm<T>(n, OP_N, Xb, OP_N, Tc, B(gb, l), 0, acc(B(gb, l)), lds);
    wg_gemm<T>(n, OP_C, Xb, OP_N, B(g, l), Tbar, 0, acc(Tbar), lds);
    double2* t = Gb; Gb = Gb2; Gb2 = t;
    t = Lb; Lb = Lb2; Lb2 = t;
  }

  // ---- objective on the last slab: Lb += -w KR^H, Gb += -w fR (KR^H - KR)
  wg_ew(n, Lb, 0, zero, Lb, one, nullptr, one, K, make_double2(-wtr, 0));
  wg_ew(n, Gb, 0, zero, Gb, one, K, make_double2(wtr * fR, 0), K, make_double2(-wtr * fR, 0));

  // ---- reverse of the right boundary
  {
    const double2* GN = B(G, N - 1);
    const double2* sN = B(sv, N - 1);
    const double2* Tc = Cb(N - 1);
    wg_gemm<T>(n, OP_N, GN, OP_N, sN, M, 0, st, lds);                  // M = G s
    wg_gemm<T>(n, OP_N, Lb, OP_N, GN, Mb, 0, st, lds);                 // Mb = Lb G
    wg_gemm<T>(n, OP_C, Lb, OP_N, M, Gb, 0, acc(Gb), lds);             // Gb += Lb^H M
    wg_gemm<T>(n, OP_N, Mb, OP_C, sN, Gb, 0, acc(Gb), lds);            // Gb += Mb s^H
    wg_gemm<T>(n, OP_C, GN, OP_N, Mb, sb, 0, st, lds);                 // sb = G^H Mb
    wg_gemm<T>(n, OP_C, GN, OP_N, Gb, T0, 0, st, lds);
    Epi ng = epi_store();
    ng.sacc = -1.0;
    wg_gemm<T>(n, OP_N, T0, OP_C, GN, Ab, 0, ng, lds);                 // Ab = -G^H Gb G^H
    {
      double tr_ab = 0.0;
      for (int i = threadIdx.x; i < n; i += blockDim.x) tr_ab += gc(Ab)[i + (size_t)i * n].x;
      tr_ab = block_sum(tr_ab, red);
      if (threadIdx.x == 0) a.dU[(size_t)e * N + N - 1] = -tr_ab;
    }
    wg_gemm<T>(n, OP_N, Ab, OP_N, Tc, Kb, 0, ng, lds);                 // Kb = -Ab T
    wg_gemm<T>(n, OP_C, Tc, OP_N, Kb, B(gb, N - 2), 0, acc(B(gb, N - 2)), lds);
    wg_gemm<T>(n, OP_N, sb, OP_N, Tc, Nb, 0, st, lds);                 // Nb = sb T
    wg_gemm<T>(n, OP_C, Tc, OP_N, Nb, B(glb, N - 2), 0, acc(B(glb, N - 2)), lds);
  }

  // ---- reverse of the forward sweep, l = N-2 .. 0
  for (int l = N - 2; l >= 0; --l) {
    wg_gemm<T>(n, OP_N, B(g, l), OP_N, B(sv, l), M, 0, st, lds);       // M = g s
    wg_gemm<T>(n, OP_N, B(glb, l), OP_N, B(g, l), Mb, 0, st, lds);     // Mb = glb g
    wg_gemm<T>(n, OP_C, B(glb, l), OP_N, M, B(gb, l), 0, acc(B(gb, l)), lds);
    wg_gemm<T>(n, OP_N, Mb, OP_C, B(sv, l), B(gb, l), 0, acc(B(gb, l)), lds);
    wg_gemm<T>(n, OP_C, B(g, l), OP_N, Mb, sb, 0, st, lds);            // sb = g^H Mb
    wg_gemm<T>(n, OP_C, B(g, l), OP_N, B(gb, l), T0, 0, st, lds);
    Epi ng = epi_store();
    ng.sacc = -1.0;
    wg_gemm<T>(n, OP_N, T0, OP_C, B(g, l), Ab, 0, ng, lds);            // Ab = -g^H gb g^H
    {
      double tr_ab = 0.0;
      for (int i = threadIdx.x; i < n; i += blockDim.x) tr_ab += gc(Ab)[i + (size_t)i * n].x;
      tr_ab = block_sum(tr_ab, red);
      if (threadIdx.x == 0) a.dU[(size_t)e * N + l] = -tr_ab;
    }
    if (l >= 1) {
      const double2* Tc = Cb(l);
      double2* Tbar = dH + (size_t)l * nn;
      wg_gemm<T>(n, OP_N, Tc, OP_N, B(g, l - 1), K, 0, st, lds);       // K = T g_{l-1}
      wg_gemm<T>(n, OP_N, Ab, OP_N, Tc, Kb, 0, ng, lds);               // Kb = -Ab T
      Epi na = epi_store();
      na.sacc = -1.0; na.D1 = Tbar;
      wg_gemm<T>(n, OP_C, Ab, OP_N, K, Tbar, 0, na, lds);              // Tbar -= Ab^H K
      wg_gemm<T>(n, OP_N, Kb, OP_C, B(g, l - 1), Tbar, 0, acc(Tbar), lds);
      wg_gemm<T>(n, OP_C, Tc, OP_N, Kb, B(gb, l - 1), 0, acc(B(gb, l - 1)), lds);
      wg_gemm<T>(n, OP_N, Tc, OP_N, B(gl, l - 1), Nn, 0, st, lds);     // N = T gl_{l-1}
      wg_gemm<T>(n, OP_N, sb, OP_N, Tc, Nb, 0, st, lds);               // Nb = sb T
      wg_gemm<T>(n, OP_C, sb, OP_N, Nn, Tbar, 0, acc(Tbar), lds);
      wg_gemm<T>(n, OP_N, Nb, OP_C, B(gl, l - 1), Tbar, 0, acc(Tbar), lds);
      wg_gemm<T>(n, OP_C, Tc, OP_N, Nb, B(glb, l - 1), 0, acc(B(glb, l - 1)), lds);
    }
  }
  if (threadIdx.x == 0) {
    a.J[e] = Jv;
    if (a.info) {
      a.info[4 * (size_t)e + 0] = itL;
      a.info[4 * (size_t)e + 1] = itR;
      a.info[4 * (size_t)e + 2] = hitL | (hitR << 1);
      a.info[4 * (size_t)e + 3] = fails;
    }
  }
  __syncthreads();
}

template <int T, int NMAX>
__global__ __launch_bounds__(Cfg<T>::NTH, 1) void rgf_adjoint(AdjArgs a) {
  __shared__ double lds[SweepLds<T, NMAX>::DOUBLES];
  __shared__ double red[16];
  double2* ws = a.ws + (size_t)blockIdx.x * (size_t)a.ws_slot;
  double2* dH = a.dH + (size_t)blockIdx.x * a.nx * (size_t)a.n * a.n;
  for (int e = blockIdx.x; e < a.nE; e += gridDim.x) adjoint_energy<T, NMAX>(a, e, ws, dH, (ldsd*)lds, (ldsd*)red);
}

// Sum the per-slot coupling adjoints: out[l] = sum_s dH[s][l] for l = 1..nx-2, zero at the edges.
__global__ void k_sum_slots(const double2* __restrict__ dH, int slots, int nx, int n, double2* __restrict__ out) {
  const size_t nn = (size_t)n * n, tot = (size_t)nx * nn;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < tot; i += (size_t)gridDim.x * blockDim.x) {
    const size_t l = i / nn;
    double2 s = make_double2(0, 0);
    if (l >= 1 && l + 1 < (size_t)nx)
      for (int k = 0; k < slots; ++k) {
        const double2 v = dH[(size_t)k * tot + i];
        s.x += v.x;
        s.y += v.y;
      }
    out[i] = s;
  }
}

}  // namespace rgf
