// rgf_wsweep.hpp -- the screened interaction W of green_rgf_calc_w
// (src/green_rgf.f90:3399-3831) as a persistent workgroup program, one energy
// (frequency) per workgroup at a time, on the GEMM and inverse blocks of the solve.
//
// Per energy, with M = I - V P^r block-tridiagonal (P^r block-diagonal, as every
// caller passes it: src/green_rgf.f90:846, 1551, 3008; no open-boundary blocks, NBC = 0):
//   M_ii = I - V_ii P_ii, M_{i+1,i} = -V1i_i P_ii, M_{i,i+1} = -V1i_i^H P_{i+1,i+1}
//   (src/green_rgf.f90:3456-3492)
//   right-to-left:  xlr_nx = M_nx^-1, Wlr_nx = xlr_nx V_nx;
//                   xlr_i = (M_ii - M_{i,i+1} xlr_{i+1} M_{i+1,i})^-1,
//                   Wlr_i = xlr_i (V_ii - M_{i,i+1} xlr_{i+1} V1i_i)       (3700-3738)
//   left-to-right:  W^r_1 = Wlr_1, W^r_i = Wlr_i - xlr_i M_{i,i-1} W^r_{i-1,i},
//                   W^r_{i,i+1} = (V1i_i^T - W^r_i M_{i+1,i}^T) xlr_{i+1}^T   (3739-3791)
//   W^{<,>}_i = W^r_i P^{<,>}_i W^r_i^H - X_a P^{<,>}_b X_a^T - X_m^T P^{<,>}_m X_m
//   with X = i Im W^r_{i,i+1} (the real part zeroed, 3798), a = i, b = i + 1 for i < nx
//   and a = nx - 1, b = nx at i = nx, m = max(i - 1, 1)                       (3798-3826)
// The transposes (not conjugate transposes) are the reference's: they give the blocks
// of (I - V P^r)^-1 V when V and P^r are symmetric.  Storage: the outputs hold the
// intermediate blocks until their own values are written -- W^r_i holds Wlr_i,
// W^<_i holds xlr_i, W^>_{i+1} holds M_{i+1,i} -- so an energy slot needs only
// kWTmpBlocks temporaries.
#pragma once
#include "rgf_sweep.hpp"

namespace rgf {

constexpr int kWTmpBlocks = 12;

struct WArgs {
  int n, nx, nE;
  const double2* Vii;  // (n, n, nx), energy independent
  const double2* V1i;  // (n, n, nx): block (i+1, i) of V
  const double2* Vc;   // conj(V1i): conj(Vc)^T = V1i^T enters through the epilogue's D3
  const double2* PL;   // (n, n, nx, nE) diagonal blocks of P^<, P^>, P^r
  const double2* PG;
  const double2* PR;
  double2* WL;  // (n, n, nx, nE) diagonal blocks of W^<, W^>, W^r
  double2* WG;
  double2* WR;
  double2* WRi1;  // optional: W^r_{i,i+1} (slab nx: zero)
  int* info;      // optional, per energy: the number of singular blocks met (0: none)
  double2* ws;
  long long ws_slot;
};

template <int T, int NMAX>
__device__ void w_energy(const WArgs& a, int e, double2* ws, ldsd* lds) {
  const int n = a.n, nx = a.nx;
  const size_t nn = (size_t)n * n;
  const double2 one = make_double2(1, 0), zero = make_double2(0, 0);
  auto eb = [&](const double2* base, int i) { return base + ((size_t)e * nx + i) * nn; };
  auto ebw = [&](double2* base, int i) { return base + ((size_t)e * nx + i) * nn; };
  auto vb = [&](const double2* base, int i) { return base + (size_t)i * nn; };
  double2* W = ws;
  double2* Mc = ws + nn;
  double2* Mi = ws + 2 * nn;
  double2* B = ws + 3 * nn;
  double2* D = ws + 4 * nn;
  double2* C2 = ws + 5 * nn;
  double2* Wr[2] = {ws + 6 * nn, ws + 7 * nn};  // W^r_{i,i+1} of even / odd i
  double2* Wi[2] = {ws + 8 * nn, ws + 9 * nn};  // i Im of the same blocks
  double2* Ab = ws + 10 * nn;
  double2* Bb = ws + 11 * nn;
  // output slots holding intermediates
  auto xlr = [&](int i) { return ebw(a.WL, i); };
  auto wlr = [&](int i) { return ebw(a.WR, i); };
  auto m1i = [&](int i) { return ebw(a.WG, i + 1); };  // M_{i+1,i}, i <= nx - 2
  int bad = 0;  // singular inverses
  Epi st = epi_store();
  Epi neg = epi_store();
  neg.sacc = -1.0;
  Epi idm = epi_store();  // I - acc
  idm.sacc = -1.0;
  idm.diag = one;

  // ---- right-to-left pass
  const int L = nx - 1;
  wg_gemm<T>(n, OP_N, vb(a.Vii, L), OP_N, eb(a.PR, L), W, 1, idm, lds);
  if (wg_inverse<T, NMAX>(n, W, xlr(L), lds)) ++bad;
  wg_gemm<T>(n, OP_N, xlr(L), OP_N, vb(a.Vii, L), wlr(L), 0, st, lds);
  for (int i = L - 1; i >= 0; --i) {
    wg_gemm<T>(n, OP_N, vb(a.V1i, i), OP_N, eb(a.PR, i), m1i(i), 0, neg, lds);
    wg_gemm<T>(n, OP_C, vb(a.V1i, i), OP_N, eb(a.PR, i + 1), Mi, 0, neg, lds);
    wg_gemm<T>(n, OP_N, Mi, OP_N, xlr(i + 1), B, 0, st, lds);
    wg_gemm<T>(n, OP_N, vb(a.Vii, i), OP_N, eb(a.PR, i), Mc, 0, idm, lds);
    Epi em = neg;
    em.D1 = Mc;
    wg_gemm<T>(n, OP_N, B, OP_N, m1i(i), W, 1, em, lds);
    if (wg_inverse<T, NMAX>(n, W, xlr(i), lds)) ++bad;
    Epi ed = neg;
    ed.D1 = vb(a.Vii, i);
    wg_gemm<T>(n, OP_N, B, OP_N, vb(a.V1i, i), D, 0, ed, lds);
    wg_gemm<T>(n, OP_N, xlr(i), OP_N, D, wlr(i), 0, st, lds);
  }

  // ---- left-to-right pass, with W^< and W^> of each slab once its blocks exist.
  // W^r_{i,i+1} lives in Wr[i & 1] (its imaginary part in Wi[i & 1]) while slabs i and
  // i + 1 need it.
  auto offdiag = [&](int i) {  // i <= nx - 2
    Epi ec = neg;
    ec.D3 = vb(a.Vc, i);
    wg_gemm_k<T, OP_N, OP_T, 0>(n, wlr(i), m1i(i), C2, ec, lds);  // V1i_i^T - W^r_i M_{i+1,i}^T
    wg_gemm_k<T, OP_N, OP_T, 0>(n, C2, xlr(i + 1), Wr[i & 1], st, lds);
    wg_imag_part(n, Wr[i & 1], Wi[i & 1], a.WRi1 ? ebw(a.WRi1, i) : nullptr);
  };
  // W^<_i or W^>_i (P = P^< or P^>) into Wout
  auto lesser = [&](int i, const double2* P, double2* Wout) {
    const int ia = i < L ? i : L - 1, ib = i < L ? i + 1 : L;
    const int im = i >= 1 ? i - 1 : 0;
    Epi eacc = neg;
    eacc.D1 = Bb;
    wg_gemm<T>(n, OP_N, wlr(i), OP_N, eb(P, i), Ab, 0, st, lds);
    wg_gemm<T>(n, OP_N, Ab, OP_C, wlr(i), Bb, 0, st, lds);
    wg_gemm<T>(n, OP_N, Wi[ia & 1], OP_N, eb(P, ib), Ab, 0, st, lds);
    wg_gemm_k<T, OP_N, OP_T, 0>(n, Ab, Wi[ia & 1], Bb, eacc, lds);
    wg_gemm_k<T, OP_T, OP_N, 0>(n, Wi[im & 1], eb(P, im), Ab, st, lds);
    wg_gemm<T>(n, OP_N, Ab, OP_N, Wi[im & 1], Wout, 0, eacc, lds);
  };
  offdiag(0);
  lesser(0, a.PL, xlr(0));  // xlr_0 is not needed any more
  lesser(0, a.PG, ebw(a.WG, 0));
  for (int i = 1; i < nx; ++i) {
    wg_gemm<T>(n, OP_N, xlr(i), OP_N, m1i(i - 1), B, 0, st, lds);
    Epi ew = neg;
    ew.D1 = wlr(i);
    wg_gemm<T>(n, OP_N, B, OP_N, Wr[(i - 1) & 1], wlr(i), 0, ew, lds);
    if (i < L) offdiag(i);
    lesser(i, a.PL, xlr(i));        // xlr_i and M_{i,i-1} (held in W^>_i) are spent
    lesser(i, a.PG, ebw(a.WG, i));
  }
  if (a.WRi1) wg_ew(n, ebw(a.WRi1, L), 0, zero, nullptr, one, nullptr, one, nullptr, one);
  if (a.info && threadIdx.x == 0) a.info[e] = bad;
  __syncthreads();
}

template <int T, int NMAX>
__global__ __launch_bounds__(Cfg<T>::NTH, Cfg<T>::OCC) void rgf_wsweep(WArgs a) {
  __shared__ double lds[SweepLds<T, NMAX>::DOUBLES];
  double2* ws = a.ws + (size_t)blockIdx.x * (size_t)a.ws_slot;
  for (int e = blockIdx.x; e < a.nE; e += gridDim.x) w_energy<T, NMAX>(a, e, ws, (ldsd*)lds);
}

}  // namespace rgf
