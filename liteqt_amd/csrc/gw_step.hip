// GW self-energy step of the one-dimensional SCBA driver (green_rgf_solve_gw_1D,
// src/green_rgf.f90:3186-3290) for gfx950: polarization P^{<,>,r} and self-energy
// Sigma^{<,>,r} as energy correlations of the band elements (i, j, slab), mixing and
// edge copy.  C ABI in include/gw.h.
//
// Work per band element is a 1D correlation of length nen: O(nen^2) FP64 FMAs against
// O(nen) data, so one workgroup takes one element, stages its energy series in LDS
// and runs the correlation from LDS at the FP64 FMA rate.  The reference zeroes the
// real parts of G before this step (src/green_rgf.f90:3186-3188), so the G series are
// real (their imaginary parts): P^< = dE_P * (-sum a d) etc. are sums of real
// products, Sigma sums real G times complex W.
//
// Lane mapping: each lane owns R = 4 consecutive outputs and slides an R-value register window along the shifted ("Y") series,
// so one LDS read per series and step feeds R outputs; the other series ("X") is read
// at the same address by the whole wave (broadcast).  The Y series are stored
// zero-padded and R-way de-interleaved (padded index p at (p % R) * Q + p / R), so the
// lanes' reads, R elements apart, fall on consecutive words.  Zero padding also takes
// the place of every index bound of the reference's loops: a term outside them reads
// a zero.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "../../include/gw.h"
#include "gw_fft.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(2, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// polarization: 8 waves x 64 lanes x 4 outputs = 2048 outputs per pass (two waves per SIMD)
constexpr int NTH_P = 512, RP = 4, SP = 64 * RP;
constexpr int PADP = SP;   // a wave's window reaches SP - 1 below its first output's series start
// self-energy: 8 waves x 64 lanes x 4 outputs = 2048 outputs per pass (two waves per SIMD)
constexpr int NTH_S = 512, RS = 4, SS_ = 64 * RS;
constexpr int PADY = 260;  // zero padding of a self-energy Y series on each side (reach 258)
// largest nen of the direct forms (an element's series live in LDS)
constexpr int kPolDirectMax = 3072, kSigDirectMax = 2040;

struct GwArgs {
  int nm, nx, nen, nd, nopmax, half;  // nd = min(ndiag, nm - 1), half = nen / 2
  int slab0;                          // first slab of the grid
  long long nel;                      // logical blocks (band positions or tile slots x slabs)
  int tiled;                          // 8 x 4 (i, j) tiles (dense bands) instead of column order
  double dr, di;                      // complex prefactor
  double alpha;
  const double* Gl;
  const double* Gg;
  const double* Gr;
  const double* Wl;
  const double* Wg;
  const double* Wr;
  double* Ol;  // P or Sigma outputs
  double* Og;
  double* Or;
  double* X;   // polarization: Im G^<, Im G^>, Im G^r of every band element, energy-contiguous
};

// Index (in series of nen doubles) of series s (0: G^<, 1: G^>, 2: G^r) of band element
// (i, j, slab ix) in the polarization's gathered copy: column j holds its band rows
// max(0, j - nd) .. min(nm - 1, j + nd), at most min(2 nd + 1, nm) of them.
__device__ __forceinline__ size_t xseries(const GwArgs& a, int i, int j, int ix, int s) {
  const int rows = min(2 * a.nd + 1, a.nm);
  return ((((size_t)ix * a.nm + j) * rows + (i - max(0, j - a.nd))) * 3 + s);
}

// Gather for the polarization: the imaginary parts of G^<, G^>, G^r of the band elements
// into energy-contiguous series, through a 64 x 64 LDS tile (reads along i, writes along
// the energy: both coalesced).  blockIdx.x: 64-energy tile; blockIdx.y strides over the
// (column j, slab) pairs.
__global__ __launch_bounds__(256) void k_pol_gather(GwArgs a) {
  __shared__ double tile[64][65];
  const int e0 = blockIdx.x * 64, tid = threadIdx.x;
  const size_t ES = (size_t)a.nm * a.nm * a.nx;
  for (long long cj = blockIdx.y; cj < (long long)a.nm * a.nx; cj += gridDim.y) {
    const int j = (int)(cj % a.nm), ix = (int)(cj / a.nm);
    const int lo = max(0, j - a.nd), hi = min(a.nm - 1, j + a.nd);
    for (int r0 = lo; r0 <= hi; r0 += 64) {
      const int nr = min(64, hi - r0 + 1);
      for (int s = 0; s < 3; ++s) {
        const double* G = s == 0 ? a.Gl : (s == 1 ? a.Gg : a.Gr);
        for (int k = tid; k < 64 * 64; k += 256) {
          const int r = k % 64, e = k / 64;
          double v = 0.0;
          if (r < nr && e0 + e < a.nen)
            v = G[2 * ((size_t)(r0 + r) + (size_t)a.nm * j + (size_t)a.nm * a.nm * ix + ES * (size_t)(e0 + e)) + 1];
          tile[r][e] = v;
        }
        __syncthreads();
        for (int k = tid; k < 64 * 64; k += 256) {
          const int e = k % 64, r = k / 64;
          if (r < nr && e0 + e < a.nen) a.X[xseries(a, r0 + r, j, ix, s) * a.nen + e0 + e] = tile[r][e];
        }
        __syncthreads();
      }
    }
  }
}

template <int R>
__device__ __forceinline__ int dpos(int p, int q) { return (p % R) * q + p / R; }  // p >= 0, R a power of 2

// Element (i, j, slab) of a workgroup.  Logical block b = (slab' * nm + j) * (2 nd + 1) +
// (i - j + nd): consecutive blocks are consecutive rows i of one column j, whose energy
// series are adjacent in memory.  Workgroups are dealt round-robin to the 8 XCDs (each
// with its own L2), so the grid (a multiple of 8) gives XCD x the contiguous logical range
// [x * per, (x + 1) * per): the ~32 workgroups an XCD runs at once read neighbouring
// elements and share the cache lines of their strided series.
// Dense bands (nd >= 8) use 8 x 4 tiles of (i, j) instead, i fastest: the 32 workgroups
// an XCD runs at once share the lines of the ij series 8 ways and, in the polarization,
// those of the transposed ji series 4 ways.
__device__ __forceinline__ bool element(const GwArgs& a, int& i, int& j, int& ix) {
  const int band = 2 * a.nd + 1;
  const long long per = gridDim.x / 8;
  long long b = (long long)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (b >= a.nel) return false;
  if (a.tiled) {
    const int ti = (a.nm + 7) / 8;
    const int w = (int)(b % 32);
    b /= 32;
    const int t = (int)(b % ((long long)ti * ((a.nm + 3) / 4)));
    ix = a.slab0 + (int)(b / ((long long)ti * ((a.nm + 3) / 4)));
    i = (t % ti) * 8 + (w % 8);
    j = (t / ti) * 4 + (w / 8);
    return i < a.nm && j < a.nm && abs(i - j) <= a.nd;
  }
  const int di = (int)(b % band);
  b /= band;
  j = (int)(b % a.nm);
  ix = a.slab0 + (int)(b / a.nm);
  i = j + di - a.nd;
  return i >= 0 && i < a.nm;
}

template <int NENMAX>
struct PolLds {
  static constexpr int XL = NENMAX + RP;                              // X read up to nen + RP - 2
  static constexpr int YL = (NENMAX + PADP + SP + RP + RP - 1) / RP * RP;  // Y index up to nen + SP + RP - 3
  static constexpr int DOUBLES = 3 * XL + 3 * YL;
};

// P^<_ij(w) = dE_P sum_E G^<_ij(E) G^>_ji(E-w), P^>_ij(w) = dE_P sum_E G^>_ij(E) G^<_ji(E-w),
// P^r_ij(w) = dE_P sum_E [G^<_ij(E) conj(G^r_ij(E-w)) + G^r_ij(E) G^<_ji(E-w)]
// (src/green_rgf.f90:3199-3218).  With G = i g: G^< G^> = -a d, G^< conj(G^r) +
// G^r G^<_ji = a f - c e (a, b, c = g^<_ij, g^>_ij, g^r_ij; d, e = g^>_ji, g^<_ji; f = c).
template <int NENMAX>
__global__ __launch_bounds__(NTH_P) void k_polarization(GwArgs a) {
  using L = PolLds<NENMAX>;
  __shared__ double lds[L::DOUBLES];
  int i, j, ix;
  if (!element(a, i, j, ix)) return;  // uniform over the workgroup, before any barrier
  const int nen = a.nen, tid = threadIdx.x;
  const size_t ES = (size_t)a.nm * a.nm * a.nx;  // energy stride (complex elements)
  const size_t eij = (size_t)i + (size_t)a.nm * j + (size_t)a.nm * a.nm * ix;
  const size_t eji = (size_t)j + (size_t)a.nm * i + (size_t)a.nm * a.nm * ix;
  // dense bands: the series of ij and ji come gathered (energy-contiguous); else strided
  const double* xij = a.X ? a.X + xseries(a, i, j, ix, 0) * nen : nullptr;
  const double* xji = a.X ? a.X + xseries(a, j, i, ix, 0) * nen : nullptr;
  double* xa = lds;
  double* xb = xa + L::XL;
  double* xc = xb + L::XL;
  double* yd = xc + L::XL;
  double* ye = yd + L::YL;
  double* yf = ye + L::YL;
  constexpr int Q = L::YL / RP;
  // staging: every load of a thread is issued before its LDS stores (one HBM latency
  // per group instead of one per energy block)
  {
    constexpr int NI = (L::XL + NTH_P - 1) / NTH_P;
    double va[NI], vb[NI], vc[NI];
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int p = tid + u * NTH_P;
      va[u] = vb[u] = vc[u] = 0.0;
      if (p < nen) {
        if (xij) {
          va[u] = xij[p];
          vb[u] = xij[nen + p];
          vc[u] = xij[2 * nen + p];
        } else {
          va[u] = a.Gl[2 * (eij + ES * p) + 1];
          vb[u] = a.Gg[2 * (eij + ES * p) + 1];
          vc[u] = a.Gr[2 * (eij + ES * p) + 1];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int p = tid + u * NTH_P;
      if (p < L::XL) {
        xa[p] = va[u];
        xb[p] = vb[u];
        xc[p] = vc[u];
      }
    }
  }
  {
    constexpr int NI = (L::YL + NTH_P - 1) / NTH_P;
    double vd[NI], ve[NI];
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int k = tid + u * NTH_P - PADP;
      vd[u] = ve[u] = 0.0;
      if (k >= 0 && k < nen) {
        if (xji) {
          vd[u] = xji[nen + k];
          ve[u] = xji[k];
        } else {
          vd[u] = a.Gg[2 * (eji + ES * k) + 1];
          ve[u] = a.Gl[2 * (eji + ES * k) + 1];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int p = tid + u * NTH_P;
      if (p < L::YL) {
        const int q = dpos<RP>(p, Q);
        yd[q] = vd[u];
        ye[q] = ve[u];
      }
    }
  }
  __syncthreads();
  for (int p = tid; p < L::YL; p += NTH_P) {  // the shifted copy of g^r_ij comes from LDS
    const int k = p - PADP;
    yf[dpos<RP>(p, Q)] = (k >= 0 && k < nen) ? xc[k] : 0.0;
  }
  __syncthreads();
  const int nout = 2 * a.nopmax + 1;
  const int lo_iop = a.half - 1 - a.nopmax;  // 0-based energy index of nop = -nopmax
  // indices outside the frequency window hold zero (the reference zeroes P first)
  for (int p = tid; p < nen; p += NTH_P) {
    if (p < lo_iop || p >= lo_iop + nout) {
      const size_t o = 2 * (eij + ES * p);
      a.Ol[o] = 0.0; a.Ol[o + 1] = 0.0;
      a.Og[o] = 0.0; a.Og[o + 1] = 0.0;
      a.Or[o] = 0.0; a.Or[o + 1] = 0.0;
    }
  }
  const int wave = tid >> 6, lane = tid & 63;
  for (int ow = SP * wave; ow < nout; ow += (NTH_P / 64) * SP) {
    const int nop_lo = ow - a.nopmax;
    const int nop_hi = min(ow + SP - 1, nout - 1) - a.nopmax;
    const int elo = max(nop_lo, 0);                 // union over the wave of the
    const int ehi = min(nen - 1, nen - 1 + nop_hi); // valid energies of its outputs
    const int nb = nop_lo + RP * lane;              // this lane's outputs: nop = nb .. nb+RP-1
    double rd[RP], re[RP], rf[RP];
#pragma unroll
    for (int s = 1; s < RP; ++s) {  // slots of energies elo-1 .. elo-RP+1
      const int q = dpos<RP>(elo - s - nb + PADP, Q);
      rd[RP - s] = yd[q];
      re[RP - s] = ye[q];
      rf[RP - s] = yf[q];
    }
    // Y(e + s - nb) for e = elo + RP it sits at off[s] + it (the residue mod RP of the
    // padded index is fixed per s, its quotient grows by one per iteration)
    int off[RP];
#pragma unroll
    for (int s = 0; s < RP; ++s) off[s] = dpos<RP>(elo + s - nb + PADP, Q);
    double sl[RP], sg[RP], sr[RP];
#pragma unroll
    for (int r = 0; r < RP; ++r) sl[r] = sg[r] = sr[r] = 0.0;
    for (int e = elo, it = 0; e <= ehi; e += RP, ++it) {
#pragma unroll
      for (int s = 0; s < RP; ++s) {
        const int m = e + s;
        rd[s] = yd[off[s] + it];
        re[s] = ye[off[s] + it];
        rf[s] = yf[off[s] + it];
        const double va = xa[m], vb = xb[m], vc = xc[m];
#pragma unroll
        for (int r = 0; r < RP; ++r) {  // output nb + r reads Y(m - nb - r), slot (s - r) mod RP
          const int q = (s - r) & (RP - 1);
          sl[r] = fma(va, rd[q], sl[r]);
          sg[r] = fma(vb, re[q], sg[r]);
          sr[r] = fma(va, rf[q], sr[r]);
          sr[r] = fma(-vc, re[q], sr[r]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < RP; ++r) {
      const int o = ow + RP * lane + r;
      if (o < nout) {
        const size_t w = 2 * (eij + ES * (size_t)(lo_iop + o));
        a.Ol[w] = -sl[r] * a.dr; a.Ol[w + 1] = -sl[r] * a.di;
        a.Og[w] = -sg[r] * a.dr; a.Og[w + 1] = -sg[r] * a.di;
        a.Or[w] = sr[r] * a.dr;  a.Or[w + 1] = sr[r] * a.di;
      }
    }
  }
}

template <int NENMAX>
struct SigLds {
  static constexpr int WL = NENMAX + 4;  // W over o = nop + nopmax (nout <= nen - 1), zero tail
  static constexpr int YL = (NENMAX + 2 * PADY + 3) / 4 * 4;
  static constexpr int DOUBLES = 3 * 2 * WL + 3 * YL;
};

// Sigma_new^x(E) = dE_S sum_w G^x(E-w) W^x(w) over the reference's terms (E_ie with
// ie > max(nop, 1) and ie < nen + nop: G's index ie - nop in 1..nen-1, output ie = 1
// empty), Im Sigma_new^r = Im(Sigma_new^> - Sigma_new^<)/2, mixing and edge copy
// (src/green_rgf.f90:3247-3290).  With G = i g: (i g) W = (-g Im W, g Re W).
template <int NENMAX>
__global__ __launch_bounds__(NTH_S) void k_selfenergy(GwArgs a) {
  using L = SigLds<NENMAX>;
  __shared__ double lds[L::DOUBLES];
  int i, j, ix;
  if (!element(a, i, j, ix)) return;
  const int nen = a.nen, tid = threadIdx.x;
  const size_t ES = (size_t)a.nm * a.nm * a.nx;
  const size_t SS = (size_t)a.nm * a.nm;  // slab stride
  const size_t eij = (size_t)i + (size_t)a.nm * j + SS * ix;
  double* wl = lds;
  double* wg = wl + 2 * L::WL;
  double* wr = wg + 2 * L::WL;
  double* yl = wr + 2 * L::WL;
  double* yg = yl + L::YL;
  double* yr = yg + L::YL;
  constexpr int Q = L::YL / 4;
  const int nout = 2 * a.nopmax + 1;
  const int lo_iop = a.half - 1 - a.nopmax;
  {  // staging: every load of a thread is issued before its LDS stores
    constexpr int NI = (L::WL + NTH_S - 1) / NTH_S;
    double2 vl[NI], vg[NI], vr[NI];
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int p = tid + u * NTH_S;
      vl[u] = vg[u] = vr[u] = make_double2(0, 0);
      if (p < nout) {
        const size_t o = eij + ES * (size_t)(lo_iop + p);
        vl[u] = reinterpret_cast<const double2*>(a.Wl)[o];
        vg[u] = reinterpret_cast<const double2*>(a.Wg)[o];
        vr[u] = reinterpret_cast<const double2*>(a.Wr)[o];
      }
    }
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int p = tid + u * NTH_S;
      if (p < L::WL) {
        wl[2 * p] = vl[u].x; wl[2 * p + 1] = vl[u].y;
        wg[2 * p] = vg[u].x; wg[2 * p + 1] = vg[u].y;
        wr[2 * p] = vr[u].x; wr[2 * p + 1] = vr[u].y;
      }
    }
  }
  {
    constexpr int NI = (L::YL + NTH_S - 1) / NTH_S;
    double vl[NI], vg[NI], vr[NI];
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int k = tid + u * NTH_S - PADY;
      vl[u] = vg[u] = vr[u] = 0.0;
      if (k >= 0 && k < nen - 1) {  // the last energy of G never enters
        vl[u] = a.Gl[2 * (eij + ES * k) + 1];
        vg[u] = a.Gg[2 * (eij + ES * k) + 1];
        vr[u] = a.Gr[2 * (eij + ES * k) + 1];
      }
    }
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int p = tid + u * NTH_S;
      if (p < L::YL) {
        const int q = dpos<RS>(p, Q);
        yl[q] = vl[u];
        yg[q] = vg[u];
        yr[q] = vr[u];
      }
    }
  }
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63;
  const int last = a.nx - 3;
  for (int ew = SS_ * wave; ew < nen; ew += 8 * SS_) {
    const int e_hi = min(ew + SS_ - 1, nen - 1);
    const int nlo = max(-a.nopmax, ew - (nen - 2));
    const int nhi = min(a.nopmax, e_hi);
    const int eb = ew + 4 * lane;  // this lane's outputs: e = eb .. eb+3
    const int c = eb - nlo;        // step t reads G(c - t)
    double gl[4], gg[4], gr[4];
#pragma unroll
    for (int s = 1; s < 4; ++s) {  // slots of steps -1, -2, -3: G(c+1), G(c+2), G(c+3)
      const int q = dpos<RS>(c + s + PADY, Q);
      gl[4 - s] = yl[q];
      gg[4 - s] = yg[q];
      gr[4 - s] = yr[q];
    }
    // G(c - t) for t = 4 it + s sits at off[s] - it
    int off[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) off[s] = dpos<RS>(c - s + PADY, Q);
    double lr[4] = {0, 0, 0, 0}, li[4] = {0, 0, 0, 0}, gre[4] = {0, 0, 0, 0}, gim[4] = {0, 0, 0, 0};
    double rr[4] = {0, 0, 0, 0}, ri[4] = {0, 0, 0, 0};
    for (int t0 = 0, it = 0; t0 <= nhi - nlo; t0 += 4, ++it) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int o = nlo + t0 + s + a.nopmax;  // W index of nop = nlo + t0 + s
        gl[s] = yl[off[s] - it];
        gg[s] = yg[off[s] - it];
        gr[s] = yr[off[s] - it];
        const double wlr = wl[2 * o], wli = wl[2 * o + 1];
        const double wgr = wg[2 * o], wgi = wg[2 * o + 1];
        const double wrr = wr[2 * o], wri = wr[2 * o + 1];
        const double wsr = wlr + wrr, wsi = wli + wri;
#pragma unroll
        for (int r = 0; r < 4; ++r) {  // output eb + r reads G(c + r - t), slot (s - r) & 3
          const int q = (s - r) & 3;
          lr[r] = fma(-gl[q], wli, lr[r]);
          li[r] = fma(gl[q], wlr, li[r]);
          gre[r] = fma(-gg[q], wgi, gre[r]);
          gim[r] = fma(gg[q], wgr, gim[r]);
          rr[r] = fma(-gl[q], wri, rr[r]);
          rr[r] = fma(-gr[q], wsi, rr[r]);
          ri[r] = fma(gl[q], wrr, ri[r]);
          ri[r] = fma(gr[q], wsr, ri[r]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = eb + r;
      if (e >= nen) continue;
      const bool first = (e == 0);  // ie = 1 has no term
      const double2 al = first ? make_double2(0, 0) : make_double2(lr[r], li[r]);
      const double2 ag = first ? make_double2(0, 0) : make_double2(gre[r], gim[r]);
      const double2 ar = first ? make_double2(0, 0) : make_double2(rr[r], ri[r]);
      const double2 nl = make_double2(al.x * a.dr - al.y * a.di, al.x * a.di + al.y * a.dr);
      const double2 ng = make_double2(ag.x * a.dr - ag.y * a.di, ag.x * a.di + ag.y * a.dr);
      const double2 nr = make_double2(ar.x * a.dr - ar.y * a.di, (ng.y - nl.y) / 2.0);
      const size_t o = eij + ES * (size_t)e;
      double2* Sl = reinterpret_cast<double2*>(a.Ol);
      double2* Sg = reinterpret_cast<double2*>(a.Og);
      double2* Sr = reinterpret_cast<double2*>(a.Or);
      const double2 ol = Sl[o], og = Sg[o], orr = Sr[o];
      const double2 ml = make_double2(ol.x + a.alpha * (nl.x - ol.x), ol.y + a.alpha * (nl.y - ol.y));
      const double2 mg = make_double2(og.x + a.alpha * (ng.x - og.x), og.y + a.alpha * (ng.y - og.y));
      const double2 mr = make_double2(orr.x + a.alpha * (nr.x - orr.x), orr.y + a.alpha * (nr.y - orr.y));
      Sl[o] = ml; Sg[o] = mg; Sr[o] = mr;
      if (ix == 2) {  // slabs 1-2 (1-based) take slab 3
        Sl[o - 2 * SS] = ml; Sg[o - 2 * SS] = mg; Sr[o - 2 * SS] = mr;
        Sl[o - SS] = ml; Sg[o - SS] = mg; Sr[o - SS] = mr;
      }
      if (ix == last) {  // slabs nx-1..nx take slab nx-2
        Sl[o + SS] = ml; Sg[o + SS] = mg; Sr[o + SS] = mr;
        Sl[o + 2 * SS] = ml; Sg[o + 2 * SS] = mg; Sr[o + 2 * SS] = mr;
      }
    }
  }
}

// P outside the band (|i - j| > nd): zero.  blockIdx.x covers the nm^2 pairs of a slab
// (32-bit index math), blockIdx.y strides over the nx * nen (slab, energy) rows.
__global__ __launch_bounds__(256) void k_pol_offband(GwArgs a) {
  const int nn = a.nm * a.nm;
  const int ij = blockIdx.x * 256 + threadIdx.x;
  if (ij >= nn) return;
  const int i = ij % a.nm, j = ij / a.nm;
  if (abs(i - j) <= a.nd) return;
  const long long rows = (long long)a.nx * a.nen;
  for (long long r = blockIdx.y; r < rows; r += gridDim.y) {
    const size_t o = (size_t)r * nn + ij;
    reinterpret_cast<double2*>(a.Ol)[o] = make_double2(0, 0);
    reinterpret_cast<double2*>(a.Og)[o] = make_double2(0, 0);
    reinterpret_cast<double2*>(a.Or)[o] = make_double2(0, 0);
  }
}

// Sigma outside the band: Sigma_new = 0, so Sigma_gw += alpha (0 - Sigma_gw), on slabs
// 3..nx-2 (1-based), with the same edge copy.  Rows are (energy, interior slab) pairs.
__global__ __launch_bounds__(256) void k_sig_offband(GwArgs a) {
  const int nn = a.nm * a.nm;
  const int ij = blockIdx.x * 256 + threadIdx.x;
  if (ij >= nn) return;
  const int i = ij % a.nm, j = ij / a.nm;
  if (abs(i - j) <= a.nd) return;
  const size_t SS = (size_t)nn;
  const int ns = a.nx - 4;
  const int last = a.nx - 3;
  const long long rows = (long long)ns * a.nen;
  double2* S[3] = {reinterpret_cast<double2*>(a.Ol), reinterpret_cast<double2*>(a.Og),
                   reinterpret_cast<double2*>(a.Or)};
  for (long long r = blockIdx.y; r < rows; r += gridDim.y) {
    const int ix = 2 + (int)(r % ns);
    const long long e = r / ns;
    const size_t o = (size_t)ij + SS * ix + SS * a.nx * (size_t)e;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double2 v = S[k][o];
      const double2 m = make_double2(v.x + a.alpha * (0.0 - v.x), v.y + a.alpha * (0.0 - v.y));
      S[k][o] = m;
      if (ix == 2) { S[k][o - 2 * SS] = m; S[k][o - SS] = m; }
      if (ix == last) { S[k][o + SS] = m; S[k][o + 2 * SS] = m; }
    }
  }
}

// ---------------------------------------------------------------------------------
// FFT form (src/fft_mod.f90:56-212): every band element's series are zero-padded to
// N = 2^logN >= nen + nopmax points (no wrap-around reaches the frequency window or the
// energy range), transformed, multiplied and transformed back.  A chunk of slabs holds
// K series of N points per slot, slot = (slab - slab0) * nm + j) * R + r for row
// i = max(0, j - nd) + r (R = min(2 nd + 1, nm)), series (slot * K + kind) * N.
//   polarization, K = 3: 0 = Im G^< + i Im G^> of ij, 1 = Im G^r of ij (then
//     F_l + i F_g), 2 = F_r.  With a, b, c the ij series and d, e = g^>_ji, g^<_ji:
//     F_l = A conj(D), F_g = B conj(E), F_r = A conj(C) - C conj(E) (the correlations
//     sum_E x(E) y(E - w) of the direct form are IDFT(X conj(Y)) at w mod N).
//   self-energy, K = 5: 0 = Im G^< + i Im G^>, 1 = Im G^r (energies 0 .. nen-2; the
//     last energy of G never enters), 2, 3, 4 = W^<, W^>, W^r with W(w) at w mod N,
//     then Y^< = G^< W^<, Y^> = G^> W^>, Y^r = G^< W^r + G^r (W^< + W^r) (the sums
//     sum_w g(E - w) W(w) are IDFT(G W) at E).
// Real series are transformed two to a complex series and separated with
// X_k = (Z_k + conj Z_{N-k}) / 2, Y_k = (Z_k - conj Z_{N-k}) / (2 i).
struct FftGw {
  int nm, nx, nen, nd, nopmax, half, R, K;
  int lrb;  // log2 of the rows of a gather / scatter tile
  int slab0, nsl;
  int logN, logN1, logN2;
  double dr, di, alpha, inv_n;
  const double* Gl;
  const double* Gg;
  const double* Gr;
  const double* Wl;
  const double* Wg;
  const double* Wr;
  double* Ol;
  double* Og;
  double* Or;
  double2* Z;
};

__device__ __forceinline__ long long fslot(const FftGw& a, int i, int j, int ixl) {
  return ((long long)ixl * a.nm + j) * a.R + (i - max(0, j - a.nd));
}

// Column run (ixl, j) and row block rb (rows r0 .. r0 + nr - 1 of the band rows of
// column j, blocks of RB = 2^lrb rows) of blockIdx-strided index cy; false when the block
// holds no row (uniform over the workgroup).
__device__ __forceinline__ bool fft_run(const FftGw& a, long long cy, int& j, int& ixl, int& r0, int& nr, int& lo) {
  const int RB = 1 << a.lrb;
  const int nrb = (a.R + RB - 1) / RB;
  const int rb = (int)(cy % nrb);
  const long long cj = cy / nrb;
  j = (int)(cj % a.nm);
  ixl = (int)(cj / a.nm);
  lo = max(0, j - a.nd);
  const int hi = min(a.nm - 1, j + a.nd);
  r0 = rb * RB;
  nr = min(RB, hi - lo + 1 - r0);
  return nr > 0;
}

// Series gather through an LDS tile of RB rows x EB = 4096 / RB positions (reads along
// i, writes along the energy; RB = the band's rows per column rounded up to a power of
// two, at most 64, so narrow bands fill the tile with energies instead of empty rows).
// POL: kinds 0, 1 of the polarization; else kinds 0 .. 4 of the self-energy.
template <bool POL>
__global__ __launch_bounds__(256) void k_fft_gather(FftGw a) {
  __shared__ double2 tile[4096 + 64];
  const int lrb = a.lrb, RB = 1 << lrb, EB = min(4096 >> lrb, 1 << a.logN), TS = EB + 1;
  const int p0 = blockIdx.x * EB, tid = threadIdx.x, N = 1 << a.logN;
  const size_t ES = (size_t)a.nm * a.nm * a.nx;
  const long long runs = (long long)a.nsl * a.nm * ((a.R + RB - 1) / RB);
  const int glim = POL ? a.nen : a.nen - 1;  // G energies that enter
  for (long long cy = blockIdx.y; cy < runs; cy += gridDim.y) {
    int j, ixl, r0, nr, lo;
    if (!fft_run(a, cy, j, ixl, r0, nr, lo)) continue;
    const int ix = a.slab0 + ixl;
    const size_t col = (size_t)a.nm * j + (size_t)a.nm * a.nm * ix + lo + r0;
    for (int s = 0; s < (POL ? 2 : 5); ++s) {
      for (int k = tid; k < RB * EB; k += 256) {
        const int r = k & (RB - 1), p = p0 + (k >> lrb);
        double2 v = make_double2(0.0, 0.0);
        if (r < nr) {
          if (s < 2) {
            if (p < glim) {
              const size_t o = 2 * (col + r + ES * (size_t)p) + 1;
              v = s == 0 ? make_double2(a.Gl[o], a.Gg[o]) : make_double2(a.Gr[o], 0.0);
            }
          } else {
            int nop = -1 - a.nopmax;  // none
            if (p <= a.nopmax) nop = p;
            else if (p >= N - a.nopmax) nop = p - N;
            if (nop >= -a.nopmax) {
              const double* W = s == 2 ? a.Wl : (s == 3 ? a.Wg : a.Wr);
              v = reinterpret_cast<const double2*>(W)[col + r + ES * (size_t)(a.half - 1 + nop)];
            }
          }
        }
        tile[r * TS + (k >> lrb)] = v;
      }
      __syncthreads();
      for (int k = tid; k < RB * EB; k += 256) {
        const int e = k % EB, r = k / EB;
        if (r < nr)
          a.Z[((size_t)fslot(a, lo + r0 + r, j, ixl) * a.K + s) * N + p0 + e] = tile[r * TS + e];
      }
      __syncthreads();
    }
  }
}

// One slot per blockIdx.x / pblocks; each frequency pair {k, N - k} is handled by the
// thread at the storage position of its smaller member.
__device__ __forceinline__ bool fft_pair(const FftGw& a, long long& slot, int& i, int& j, int& ixl, long long& p,
                                         long long& pp) {
  // 32-bit index math: slots and positions of one launch stay below 2^31
  const int N = 1 << a.logN;
  const int lpb = a.logN > 8 ? a.logN - 8 : 0;  // log2 of the 256-position blocks per slot
  const unsigned sl = blockIdx.x >> lpb;
  slot = sl;
  p = (long long)((blockIdx.x & ((1u << lpb) - 1)) * 256 + threadIdx.x);
  if (p >= N) return false;
  const unsigned col = sl / (unsigned)a.R;
  const int r = (int)(sl - col * (unsigned)a.R);
  ixl = (int)(col / (unsigned)a.nm);
  j = (int)(col - (unsigned)ixl * (unsigned)a.nm);
  i = max(0, j - a.nd) + r;
  if (i > min(a.nm - 1, j + a.nd)) return false;
  const long long k = gwfft::freq(p, a.logN, a.logN1, a.logN2);
  const long long kk = (N - k) & (N - 1);
  if (k > kk) return false;
  pp = gwfft::pos(kk, a.logN, a.logN1, a.logN2);
  return true;
}

__device__ __forceinline__ double2 sep_x(double2 zk, double2 zm) {  // spectrum of Re z at k
  return make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
}
__device__ __forceinline__ double2 sep_y(double2 zk, double2 zm) {  // spectrum of Im z at k
  return make_double2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x));
}

__global__ __launch_bounds__(256) void k_pol_fft_product(FftGw a) {
  long long slot, p, pp;
  int i, j, ixl;
  if (!fft_pair(a, slot, i, j, ixl, p, pp)) return;
  const size_t N = (size_t)1 << a.logN;
  const double2* Z1 = a.Z + (size_t)slot * 3 * N;
  double2* Z2 = a.Z + ((size_t)slot * 3 + 1) * N;
  double2* O2 = a.Z + ((size_t)slot * 3 + 2) * N;
  const double2* T1 = a.Z + (size_t)fslot(a, j, i, ixl) * 3 * N;
  const double2 zk = Z1[p], zm = Z1[pp], tk = T1[p], tm = T1[pp], C = Z2[p];
  const double2 A = sep_x(zk, zm), B = sep_y(zk, zm);
  const double2 E = sep_x(tk, tm), D = sep_y(tk, tm);
  using gwfft::cmul;
  using gwfft::conj2;
  const double2 Fl = cmul(A, conj2(D)), Fg = cmul(B, conj2(E));
  const double2 Fr = gwfft::csub(cmul(A, conj2(C)), cmul(C, conj2(E)));
  Z2[p] = make_double2(Fl.x - Fg.y, Fl.y + Fg.x);  // F_l + i F_g
  O2[p] = Fr;
  if (pp != p) {
    Z2[pp] = make_double2(Fl.x + Fg.y, -Fl.y + Fg.x);  // conj(F_l) + i conj(F_g)
    O2[pp] = conj2(Fr);
  }
}

__global__ __launch_bounds__(256) void k_sig_fft_product(FftGw a) {
  long long slot, p, pp;
  int i, j, ixl;
  if (!fft_pair(a, slot, i, j, ixl, p, pp)) return;
  const size_t N = (size_t)1 << a.logN;
  const double2* G1 = a.Z + (size_t)slot * 5 * N;
  const double2* G2 = G1 + N;
  double2* W1 = a.Z + ((size_t)slot * 5 + 2) * N;
  double2* W2 = W1 + N;
  double2* W3 = W2 + N;
  using gwfft::cadd;
  using gwfft::cmul;
  using gwfft::conj2;
  const double2 zk = G1[p], zm = G1[pp];
  const double2 Gl = sep_x(zk, zm), Gg = sep_y(zk, zm), Gr = G2[p];
  const double2 w1 = W1[p], w2 = W2[p], w3 = W3[p];
  const double2 w1m = W1[pp], w2m = W2[pp], w3m = W3[pp];
  W1[p] = cmul(Gl, w1);
  W2[p] = cmul(Gg, w2);
  W3[p] = cadd(cmul(Gl, w3), cmul(Gr, cadd(w1, w3)));
  if (pp != p) {
    W1[pp] = cmul(conj2(Gl), w1m);
    W2[pp] = cmul(conj2(Gg), w2m);
    W3[pp] = cadd(cmul(conj2(Gl), w3m), cmul(conj2(Gr), cadd(w1m, w3m)));
  }
}

// P from the transformed-back series: S_l + i S_g (kind 1) and S_r (kind 2) at
// w mod N, times 1/N; P^< = -dE_P S_l, P^> = -dE_P S_g, P^r = dE_P S_r in the window,
// zero at the other energies.  Tiles of RB rows x 2048 / RB energies through LDS.
__global__ __launch_bounds__(256) void k_pol_fft_scatter(FftGw a) {
  __shared__ double tl[2048 + 64], tg[2048 + 64], tr[2048 + 64];
  const int lrb = a.lrb, RB = 1 << lrb, EB = 2048 >> lrb, TS = EB + 1;
  const int e0 = blockIdx.x * EB, tid = threadIdx.x, N = 1 << a.logN;
  const size_t ES = (size_t)a.nm * a.nm * a.nx;
  const long long runs = (long long)a.nsl * a.nm * ((a.R + RB - 1) / RB);
  for (long long cy = blockIdx.y; cy < runs; cy += gridDim.y) {
    int j, ixl, r0, nr, lo;
    if (!fft_run(a, cy, j, ixl, r0, nr, lo)) continue;
    for (int k = tid; k < RB * EB; k += 256) {
      const int e = k % EB, r = k / EB, eo = e0 + e;
      const int nop = eo - (a.half - 1);
      double sl = 0.0, sg = 0.0, sr = 0.0;
      if (r < nr && eo < a.nen && nop >= -a.nopmax && nop <= a.nopmax) {
        const size_t b = (size_t)fslot(a, lo + r0 + r, j, ixl) * 3 * N + (size_t)(nop & (N - 1));
        const double2 v1 = a.Z[b + N], v2 = a.Z[b + 2 * N];
        sl = v1.x * a.inv_n;
        sg = v1.y * a.inv_n;
        sr = v2.x * a.inv_n;
      }
      tl[r * TS + e] = sl;
      tg[r * TS + e] = sg;
      tr[r * TS + e] = sr;
    }
    __syncthreads();
    const size_t col = (size_t)a.nm * j + (size_t)a.nm * a.nm * (a.slab0 + ixl) + lo + r0;
    for (int k = tid; k < RB * EB; k += 256) {
      const int r = k & (RB - 1), e = k >> lrb, eo = e0 + e;
      if (r < nr && eo < a.nen) {
        const size_t w = 2 * (col + r + ES * (size_t)eo);
        const double vl = tl[r * TS + e], vg = tg[r * TS + e], vr = tr[r * TS + e];
        a.Ol[w] = -vl * a.dr; a.Ol[w + 1] = -vl * a.di;
        a.Og[w] = -vg * a.dr; a.Og[w + 1] = -vg * a.di;
        a.Or[w] = vr * a.dr;  a.Or[w + 1] = vr * a.di;
      }
    }
    __syncthreads();
  }
}

// Sigma from the transformed-back series (kinds 2, 3, 4 at E, times 1/N): with
// G = i g, Sigma_new^x = dE_S i y^x (Sigma_new(E_1) = 0), Im Sigma_new^r =
// Im(Sigma_new^> - Sigma_new^<) / 2, mixing and the edge copy as in the direct form.
// Tiles of RB rows x 1024 / RB energies.
__global__ __launch_bounds__(256) void k_sig_fft_scatter(FftGw a) {
  __shared__ double2 tl[1024 + 64], tg[1024 + 64], tr[1024 + 64];
  const int lrb = a.lrb, RB = 1 << lrb, EB = 1024 >> lrb, TS = EB + 1;
  const int e0 = blockIdx.x * EB, tid = threadIdx.x, N = 1 << a.logN;
  const size_t ES = (size_t)a.nm * a.nm * a.nx;
  const size_t SS = (size_t)a.nm * a.nm;
  const int last = a.nx - 3;
  const long long runs = (long long)a.nsl * a.nm * ((a.R + RB - 1) / RB);
  for (long long cy = blockIdx.y; cy < runs; cy += gridDim.y) {
    int j, ixl, r0, nr, lo;
    if (!fft_run(a, cy, j, ixl, r0, nr, lo)) continue;
    for (int k = tid; k < RB * EB; k += 256) {
      const int e = k % EB, r = k / EB, eo = e0 + e;
      double2 nl = make_double2(0, 0), ng = nl, nrr = nl;
      if (r < nr && eo < a.nen && eo > 0) {
        const size_t b = ((size_t)fslot(a, lo + r0 + r, j, ixl) * 5 + 2) * N + eo;
        const double2 y1 = a.Z[b], y2 = a.Z[b + N], y3 = a.Z[b + 2 * N];
        const double2 al = make_double2(-y1.y * a.inv_n, y1.x * a.inv_n);  // i y / N
        const double2 ag = make_double2(-y2.y * a.inv_n, y2.x * a.inv_n);
        const double2 ar = make_double2(-y3.y * a.inv_n, y3.x * a.inv_n);
        nl = make_double2(al.x * a.dr - al.y * a.di, al.x * a.di + al.y * a.dr);
        ng = make_double2(ag.x * a.dr - ag.y * a.di, ag.x * a.di + ag.y * a.dr);
        nrr = make_double2(ar.x * a.dr - ar.y * a.di, (ng.y - nl.y) / 2.0);
      }
      tl[r * TS + e] = nl;
      tg[r * TS + e] = ng;
      tr[r * TS + e] = nrr;
    }
    __syncthreads();
    const int ix = a.slab0 + ixl;
    const size_t col = (size_t)a.nm * j + SS * ix + lo + r0;
    double2* S[3] = {reinterpret_cast<double2*>(a.Ol), reinterpret_cast<double2*>(a.Og),
                     reinterpret_cast<double2*>(a.Or)};
    for (int k = tid; k < RB * EB; k += 256) {
      const int r = k & (RB - 1), e = k >> lrb, eo = e0 + e;
      if (r < nr && eo < a.nen) {
        const size_t o = col + r + ES * (size_t)eo;
        const double2 nv[3] = {tl[r * TS + e], tg[r * TS + e], tr[r * TS + e]};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const double2 old = S[q][o];
          const double2 m = make_double2(old.x + a.alpha * (nv[q].x - old.x), old.y + a.alpha * (nv[q].y - old.y));
          S[q][o] = m;
          if (ix == 2) { S[q][o - 2 * SS] = m; S[q][o - SS] = m; }
          if (ix == last) { S[q][o + SS] = m; S[q][o + 2 * SS] = m; }
        }
      }
    }
    __syncthreads();
  }
}

// Per-device workspace of the polarization's gathered series (grow-only).
struct Workspace {
  void* p = nullptr;
  size_t bytes = 0;
};
constexpr int kMaxDev = 64;
Workspace g_ws[kMaxDev];
std::mutex g_ws_mu;

int check(const gw_params* p, int min_nx) {
  if (!p) return fail(1, "gw: params are required");
  if (p->nm < 1 || p->nx < min_nx || p->nen < 2)
    return fail(1, "gw: need nm >= 1, nx >= " + std::to_string(min_nx) + ", nen >= 2");
  if (p->ndiag < 0) return fail(1, "gw: ndiag must be >= 0");
  if (p->nm > 4096) return fail(4, "gw: nm > 4096");
  if (p->nopmax < 0 || p->nopmax > p->nen / 2 - 1)
    return fail(1, "gw: need 0 <= nopmax <= nen/2 - 1 (frequency index nop + nen/2 inside 1..nen)");
  const long long band = 2LL * std::min(p->ndiag, p->nm - 1) + 1;
  const long long tiles = 32LL * ((p->nm + 7) / 8) * ((p->nm + 3) / 4);
  if (std::max((long long)p->nm * band, tiles) * p->nx > 2147483640LL)
    return fail(4, "gw: too many band elements for one launch");
  return 0;
}

GwArgs args(const gw_params* p) {
  GwArgs a{};
  a.nm = p->nm;
  a.nx = p->nx;
  a.nen = p->nen;
  a.nd = std::min(p->ndiag, p->nm - 1);
  a.nopmax = p->nopmax;
  a.half = p->nen / 2;
  a.alpha = p->alpha_mix;
  return a;
}

unsigned grid8(long long nel) { return (unsigned)((nel + 7) / 8 * 8); }

// (pair blocks, row stride) of the off-band kernels: ~16k workgroups in flight
dim3 offband_grid(int nm, long long rows) {
  const unsigned gx = (unsigned)((nm * nm + 255) / 256);
  const long long gy = std::max(1LL, std::min<long long>(rows, std::max(1LL, 16384LL / gx)));
  return dim3(gx, (unsigned)std::min<long long>(gy, 65535));
}


// FFT form: workspace bytes per slab of series (K kinds per slot) and the launches.
constexpr size_t kFftChunkBytes = size_t(4) << 30;  // spectra of one chunk of slabs

int fft_log2(const gw_params* p, int& logN) {
  const long long need = (long long)p->nen + p->nopmax;
  logN = 6;
  while ((1LL << logN) < need) ++logN;
  if (logN > 24) return fail(4, "gw fft: nen + nopmax > 2^24");
  return 0;
}

void fft_pass(hipStream_t s, const FftGw& g, int k0, int kc, bool inverse) {
  gwfft::FftArgs f{};
  f.base = g.Z;
  f.nslots = (long long)g.nsl * g.nm * g.R;
  f.K = g.K;
  f.k0 = k0;
  f.kc = kc;
  f.logN = g.logN;
  f.logN1 = g.logN1;
  f.logN2 = g.logN2;
  f.nm = g.nm;
  f.nd = g.nd;
  f.R = g.R;
  const long long nser = f.nslots * kc;
  if (g.logN <= gwfft::kLog) {
    const long long B = gwfft::kLds >> g.logN;
    const unsigned grid = (unsigned)((nser + B - 1) / B);
    if (inverse) hipLaunchKernelGGL((gwfft::k_fft_single<true, 256, 4096>), dim3(grid), dim3(256), 0, s, f);
    else hipLaunchKernelGGL((gwfft::k_fft_single<false, 256, 4096>), dim3(grid), dim3(256), 0, s, f);
    return;
  }
  if (g.logN == gwfft::kLog1) {  // 8192 points: one series per 512-thread workgroup
    if (inverse) hipLaunchKernelGGL((gwfft::k_fft_single<true, 512, 8192>), dim3((unsigned)nser), dim3(512), 0, s, f);
    else hipLaunchKernelGGL((gwfft::k_fft_single<false, 512, 8192>), dim3((unsigned)nser), dim3(512), 0, s, f);
    return;
  }
  const long long N1 = 1LL << g.logN1, N2 = 1LL << g.logN2;
  const unsigned gc = (unsigned)(nser * (N1 / (gwfft::kLds / N2)));
  const unsigned gr = (unsigned)(nser * (N2 / (gwfft::kLds / N1)));
  if (!inverse) {
    hipLaunchKernelGGL((gwfft::k_fft_cols<false, true>), dim3(gc), dim3(gwfft::kNth), 0, s, f);
    hipLaunchKernelGGL((gwfft::k_fft_rows<false, false>), dim3(gr), dim3(gwfft::kNth), 0, s, f);
  } else {
    hipLaunchKernelGGL((gwfft::k_fft_rows<true, true>), dim3(gr), dim3(gwfft::kNth), 0, s, f);
    hipLaunchKernelGGL((gwfft::k_fft_cols<true, false>), dim3(gc), dim3(gwfft::kNth), 0, s, f);
  }
}

// The polarization (pol) or self-energy step in the FFT form, chunk by chunk of slabs.
int fft_step(const gw_params* p, bool pol, const double* Gl, const double* Gg, const double* Gr, const double* Wl,
             const double* Wg, const double* Wr, double* Ol, double* Og, double* Or, hipStream_t s) {
  FftGw g{};
  g.nm = p->nm;
  g.nx = p->nx;
  g.nen = p->nen;
  g.nd = std::min(p->ndiag, p->nm - 1);
  g.nopmax = p->nopmax;
  g.half = p->nen / 2;
  g.R = std::min(2 * g.nd + 1, g.nm);
  g.K = pol ? 3 : 5;
  if (int rc = fft_log2(p, g.logN)) return rc;
  gwfft::split(g.logN, g.logN1, g.logN2);
  g.lrb = 0;
  while ((1 << g.lrb) < std::min(g.R, 64)) ++g.lrb;
  g.dr = pol ? p->dE_P[0] : p->dE_S[0];
  g.di = pol ? p->dE_P[1] : p->dE_S[1];
  g.alpha = p->alpha_mix;
  g.inv_n = 1.0 / (double)(1LL << g.logN);
  g.Gl = Gl; g.Gg = Gg; g.Gr = Gr;
  g.Wl = Wl; g.Wg = Wg; g.Wr = Wr;
  g.Ol = Ol; g.Og = Og; g.Or = Or;
  const int slab_lo = pol ? 0 : 2, nslab = pol ? p->nx : p->nx - 4;
  const size_t per_slab = (size_t)g.K * g.nm * g.R * ((size_t)1 << g.logN) * sizeof(double2);
  const int chunk = (int)std::max<size_t>(1, std::min<size_t>(nslab, kFftChunkBytes / per_slab));
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDev) return fail(1, "gw fft: device index out of range");
  std::lock_guard<std::mutex> lk(g_ws_mu);
  Workspace& ws = g_ws[dev];
  const size_t need = per_slab * chunk;
  if (ws.bytes < need) {
    if (ws.p) HIPCHK(hipFree(ws.p));
    ws.p = nullptr;
    ws.bytes = 0;
    if (hipMalloc(&ws.p, need) != hipSuccess) {
      (void)hipGetLastError();
      ws.p = nullptr;
      return fail(3, "gw fft: cannot allocate " + std::to_string(need) + " bytes of spectra");
    }
    ws.bytes = need;
  }
  g.Z = reinterpret_cast<double2*>(ws.p);
  hipLaunchKernelGGL(gwfft::k_fft_twiddles, dim3(gwfft::kTw / 256), dim3(256), 0, s);
  HIPCHK(hipGetLastError());
  GwArgs ob = args(p);  // off-band elements: as in the direct form
  ob.Ol = Ol; ob.Og = Og; ob.Or = Or;
  if (ob.nd < ob.nm - 1) {
    if (pol) hipLaunchKernelGGL(k_pol_offband, offband_grid(ob.nm, (long long)ob.nx * ob.nen), dim3(256), 0, s, ob);
    else hipLaunchKernelGGL(k_sig_offband, offband_grid(ob.nm, (long long)(ob.nx - 4) * ob.nen), dim3(256), 0, s, ob);
    HIPCHK(hipGetLastError());
  }
  const int N = 1 << g.logN;
  for (int s0 = 0; s0 < nslab; s0 += chunk) {
    g.slab0 = slab_lo + s0;
    g.nsl = std::min(chunk, nslab - s0);
    const int RB = 1 << g.lrb;
    const long long runs = (long long)g.nsl * g.nm * ((g.R + RB - 1) / RB);
    const unsigned gy = (unsigned)std::min<long long>(runs, 65535);
    const long long slots = (long long)g.nsl * g.nm * g.R;
    const unsigned gp = (unsigned)(slots * std::max(1, N / 256));
    const int EBg = std::min(4096 / RB, N), EBp = 2048 / RB, EBs = 1024 / RB;
    if (pol) {
      hipLaunchKernelGGL(k_fft_gather<true>, dim3(N / EBg, gy), dim3(256), 0, s, g);
      fft_pass(s, g, 0, 2, false);
      hipLaunchKernelGGL(k_pol_fft_product, dim3(gp), dim3(256), 0, s, g);
      fft_pass(s, g, 1, 2, true);
      hipLaunchKernelGGL(k_pol_fft_scatter, dim3((g.nen + EBp - 1) / EBp, gy), dim3(256), 0, s, g);
    } else {
      hipLaunchKernelGGL(k_fft_gather<false>, dim3(N / EBg, gy), dim3(256), 0, s, g);
      fft_pass(s, g, 0, 5, false);
      hipLaunchKernelGGL(k_sig_fft_product, dim3(gp), dim3(256), 0, s, g);
      fft_pass(s, g, 2, 3, true);
      hipLaunchKernelGGL(k_sig_fft_scatter, dim3((g.nen + EBs - 1) / EBs, gy), dim3(256), 0, s, g);
    }
    HIPCHK(hipGetLastError());
  }
  return 0;
}

}  // namespace

// The error state shared with the other sources of this library (gw_wfull.hip).
namespace gwerr {
int fail(int code, const std::string& msg) { return ::fail(code, msg); }
void clear() { g_err.clear(); }
}  // namespace gwerr

extern "C" {

void gw_params_reference(gw_params* p, int nm, int nx, int nen, int ndiag, double dE, double spindeg,
                         double alpha_mix) {
  const double pi = 3.14159265358979323846;
  p->nm = nm;
  p->nx = nx;
  p->nen = nen;
  p->ndiag = ndiag;
  p->nopmax = nen / 2 - 10;
  p->dE_P[0] = 0.0;
  p->dE_P[1] = -1.0 * dE / 2.0 / pi * spindeg;  // dcmplx(0, -1*(En(2)-En(1))/2/pi) * spindeg
  p->dE_S[0] = 0.0;
  p->dE_S[1] = dE / 2.0 / pi;                   // dcmplx(0, (En(2)-En(1))/2/pi)
  p->alpha_mix = alpha_mix;
}

int gw_polarization(const gw_params* p, const double* Gl, const double* Gg, const double* Gr, double* Pl,
                    double* Pg, double* Pr, void* stream) {
  g_err.clear();
  if (int rc = check(p, 1)) return rc;
  int logN = 0;
  if (p->nen > kPolDirectMax)
    if (int rc = fft_log2(p, logN)) return rc;
  if (!Gl || !Gg || !Gr || !Pl || !Pg || !Pr) return fail(1, "gw_polarization: null array");
  // the direct form keeps an element's series in LDS: longer grids take the FFT form
  if (p->nen > kPolDirectMax)
    return fft_step(p, true, Gl, Gg, Gr, nullptr, nullptr, nullptr, Pl, Pg, Pr, reinterpret_cast<hipStream_t>(stream));
  GwArgs a = args(p);
  a.slab0 = 0;
  a.dr = p->dE_P[0];
  a.di = p->dE_P[1];
  a.Gl = Gl; a.Gg = Gg; a.Gr = Gr;
  a.Ol = Pl; a.Og = Pg; a.Or = Pr;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  a.tiled = a.nd >= 8;
  const long long per_slab = a.tiled ? 32LL * ((a.nm + 7) / 8) * ((a.nm + 3) / 4) : (long long)a.nm * (2 * a.nd + 1);
  const long long nel = per_slab * a.nx;
  a.nel = nel;
  // Dense bands (the tiled mapping) stage from a gathered, energy-contiguous copy of the
  // band's Im G series: the strided reads of the dense arrays cost a cache line per value,
  // the gather reads whole lines.  Narrow bands read the arrays directly (a gather would
  // touch as many lines and add a pass).
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDev) return fail(1, "gw_polarization: device index out of range");
  const size_t rows = (size_t)std::min(2 * a.nd + 1, a.nm);
  size_t need = a.tiled ? (size_t)a.nm * rows * a.nx * 3 * (size_t)a.nen * sizeof(double) : 0;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  Workspace& ws = g_ws[dev];
  if (need && ws.bytes < need) {
    if (ws.p) HIPCHK(hipFree(ws.p));  // synchronizes with the work that used it
    ws.p = nullptr;
    ws.bytes = 0;
    if (hipMalloc(&ws.p, need) != hipSuccess) {
      // not fatal: the kernels stage from the arrays directly (slower for dense bands);
      // the failed allocation's error is cleared so later launch checks do not see it
      (void)hipGetLastError();
      ws.p = nullptr;
      need = 0;
    } else {
      ws.bytes = need;
    }
  }
  a.X = need ? reinterpret_cast<double*>(ws.p) : nullptr;
  if (a.nd < a.nm - 1) {
    hipLaunchKernelGGL(k_pol_offband, offband_grid(a.nm, (long long)a.nx * a.nen), dim3(256), 0, s, a);
    HIPCHK(hipGetLastError());
  }
  if (a.X) {
    const unsigned gy = (unsigned)std::min<long long>((long long)a.nm * a.nx, 65535);
    hipLaunchKernelGGL(k_pol_gather, dim3((unsigned)((a.nen + 63) / 64), gy), dim3(256), 0, s, a);
    HIPCHK(hipGetLastError());
  }
  if (a.nen <= 1024)
    hipLaunchKernelGGL(k_polarization<1024>, dim3(grid8(nel)), dim3(NTH_P), 0, s, a);
  else if (a.nen <= 2048)
    hipLaunchKernelGGL(k_polarization<2048>, dim3(grid8(nel)), dim3(NTH_P), 0, s, a);
  else
    hipLaunchKernelGGL(k_polarization<3072>, dim3(grid8(nel)), dim3(NTH_P), 0, s, a);
  HIPCHK(hipGetLastError());
  return 0;
}

int gw_selfenergy(const gw_params* p, const double* Gl, const double* Gg, const double* Gr, const double* Wl,
                  const double* Wg, const double* Wr, double* Sl, double* Sg, double* Sr, void* stream) {
  g_err.clear();
  if (int rc = check(p, 5)) return rc;
  int logN = 0;
  if (p->nen > kSigDirectMax)
    if (int rc = fft_log2(p, logN)) return rc;
  if (!Gl || !Gg || !Gr || !Wl || !Wg || !Wr || !Sl || !Sg || !Sr) return fail(1, "gw_selfenergy: null array");
  if (p->nen > kSigDirectMax)
    return fft_step(p, false, Gl, Gg, Gr, Wl, Wg, Wr, Sl, Sg, Sr, reinterpret_cast<hipStream_t>(stream));
  GwArgs a = args(p);
  a.slab0 = 2;
  a.dr = p->dE_S[0];
  a.di = p->dE_S[1];
  a.Gl = Gl; a.Gg = Gg; a.Gr = Gr;
  a.Wl = Wl; a.Wg = Wg; a.Wr = Wr;
  a.Ol = Sl; a.Og = Sg; a.Or = Sr;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  a.tiled = a.nd >= 8;
  const long long per_slab = a.tiled ? 32LL * ((a.nm + 7) / 8) * ((a.nm + 3) / 4) : (long long)a.nm * (2 * a.nd + 1);
  const long long nel = per_slab * (a.nx - 4);
  a.nel = nel;
  if (a.nd < a.nm - 1) {
    hipLaunchKernelGGL(k_sig_offband, offband_grid(a.nm, (long long)(a.nx - 4) * a.nen), dim3(256), 0, s, a);
    HIPCHK(hipGetLastError());
  }
  if (a.nen <= 1024)
    hipLaunchKernelGGL(k_selfenergy<1024>, dim3(grid8(nel)), dim3(NTH_S), 0, s, a);
  else
    hipLaunchKernelGGL(k_selfenergy<2040>, dim3(grid8(nel)), dim3(NTH_S), 0, s, a);
  HIPCHK(hipGetLastError());
  return 0;
}

int gw_polarization_fft(const gw_params* p, const double* Gl, const double* Gg, const double* Gr, double* Pl,
                        double* Pg, double* Pr, void* stream) {
  g_err.clear();
  if (int rc = check(p, 1)) return rc;
  int logN = 0;
  if (int rc = fft_log2(p, logN)) return rc;
  if (!Gl || !Gg || !Gr || !Pl || !Pg || !Pr) return fail(1, "gw_polarization_fft: null array");
  return fft_step(p, true, Gl, Gg, Gr, nullptr, nullptr, nullptr, Pl, Pg, Pr, reinterpret_cast<hipStream_t>(stream));
}

int gw_selfenergy_fft(const gw_params* p, const double* Gl, const double* Gg, const double* Gr, const double* Wl,
                      const double* Wg, const double* Wr, double* Sl, double* Sg, double* Sr, void* stream) {
  g_err.clear();
  if (int rc = check(p, 5)) return rc;
  int logN = 0;
  if (int rc = fft_log2(p, logN)) return rc;
  if (!Gl || !Gg || !Gr || !Wl || !Wg || !Wr || !Sl || !Sg || !Sr) return fail(1, "gw_selfenergy_fft: null array");
  return fft_step(p, false, Gl, Gg, Gr, Wl, Wg, Wr, Sl, Sg, Sr, reinterpret_cast<hipStream_t>(stream));
}

long long gw_pairs_polarization(int nen, int nopmax) {
  long long n = 0;
  for (int nop = -nopmax; nop <= nopmax; ++nop) n += nen - (nop < 0 ? -nop : nop);
  return n;
}

long long gw_pairs_selfenergy(int nen, int nopmax) {
  long long n = 0;
  for (int e = 1; e < nen; ++e) {
    const int lo = std::max(-nopmax, e - (nen - 2)), hi = std::min(nopmax, e);
    if (hi >= lo) n += hi - lo + 1;
  }
  return n;
}

const char* gw_last_error(void) { return g_err.c_str(); }

void gw_release(void) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  Workspace& ws = g_ws[dev];
  if (ws.p) (void)hipFree(ws.p);
  ws.p = nullptr;
  ws.bytes = 0;
}

}  // extern "C"
