// rgf_blocks.hpp -- workgroup-level complex-FP64 building blocks for gfx950.
//
// One workgroup (256 threads = 4 wave64) owns one energy and runs every matrix
// operation of the RGF on it: complex GEMMs on FP64 MFMA
// (v_mfma_f64_16x16x4_f64) with LDS-staged K panels, a blocked Gauss-Jordan
// inverse with partial (row) pivoting, elementwise combinations and
// reductions.  Matrices live in HBM/L2 (workspace or caller arrays); all
// operations end with a workgroup barrier so the next one sees their results.
//
// MFMA f64 16x16x4 lane layout (measured on gfx950): lane l supplies
// X[l%16][l/16] of the 16x4 operand and Y[l/16][l%16] of the 4x16 operand and
// holds D[4r + l/16][l%16] in accumulator element r.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rgf {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int PAD = 16;   // LDS row padding (doubles): row stride = 16 mod 32 doubles
constexpr int KP = 16;    // K panel depth (complex)
// Row stride (doubles) of an LDS panel stored k-contiguous ([row][k]): 2 mod 32, so
// the fragment reads of 16 rows x 2 k-steps per half-wave hit distinct 8-byte bank pairs.
constexpr int LDK = KP + 2;

enum { OP_N = 0, OP_C = 1, OP_T = 2 };  // op(X) = X, X^H, X^T

// Explicit address spaces: global (1) and LDS (3).  Pointers that reach a
// non-inlined function as generic pointers would be accessed with FLAT
// instructions, which count against both the vector-memory and the LDS wait
// counters and so serialise the global prefetch with the LDS fragment reads.
// The host pass only parses device code, so it sees the unqualified types.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) double ldsd;
typedef __attribute__((address_space(3))) int ldsi;
typedef __attribute__((address_space(1))) double2 gd2;
typedef __attribute__((address_space(1))) const double2 gcd2;
typedef __attribute__((address_space(1))) double gd;
#else
typedef double ldsd;
typedef int ldsi;
typedef double2 gd2;
typedef const double2 gcd2;
typedef double gd;
#endif

__device__ __forceinline__ gcd2* gc(const double2* p) { return (gcd2*)p; }
__device__ __forceinline__ gd2* gw(double2* p) { return (gd2*)p; }

// 16-byte complex load through a native vector type (one global_load_dwordx4).
typedef double dvec2 __attribute__((ext_vector_type(2)));
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(1))) const dvec2 gcv2;
#else
typedef const dvec2 gcv2;
#endif
__device__ __forceinline__ double2 ldg(const double2* p, size_t i) {
  const dvec2 v = ((gcv2*)p)[i];
  return make_double2(v.x, v.y);
}

// Loads and stores at a wave-uniform base plus a 32-bit byte offset: the base
// stays in scalar registers and the address is one VGPR (global_* saddr form).
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(1))) const char gcb;
typedef __attribute__((address_space(1))) char gwb;
typedef __attribute__((address_space(1))) dvec2 gv2;
#else
typedef const char gcb;
typedef char gwb;
typedef dvec2 gv2;
#endif
__device__ __forceinline__ double2 ldgb(const double2* base, uint32_t boff) {
  const dvec2 v = *(gcv2*)((gcb*)base + boff);
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void stgb(double2* base, uint32_t boff, double2 v) {
  dvec2 w;
  w.x = v.x;
  w.y = v.y;
  *(gv2*)((gwb*)base + boff) = w;
}
__device__ __forceinline__ void stgb_re(double* base, uint32_t boff, double v) { *(gd*)((gwb*)base + boff) = v; }

// Uniform values arrive in VGPRs through the calling convention of a
// non-inlined function; moving them to SGPRs frees those VGPRs.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uni(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <typename P>
__device__ __forceinline__ P* uni(P* p) {
  const unsigned long long b = (unsigned long long)(uintptr_t)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return (P*)(uintptr_t)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double2 uni(double2 v) { return make_double2(uni(v.x), uni(v.y)); }

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// Smith's algorithm for 1/b.
__device__ __forceinline__ double2 crecip(double2 b) {
  if (fabs(b.x) >= fabs(b.y)) {
    double t = b.y / b.x, d = b.x + b.y * t;
    return make_double2(1.0 / d, -t / d);
  } else {
    double t = b.x / b.y, d = b.x * t + b.y;
    return make_double2(t / d, -1.0 / d);
  }
}

// out(i,j) = sacc*acc(i,j) + diag*[i==j] + s1*D1(i,j) + s2*D2(i,j) + s3*conj(D3(j,i))
// D1 may alias the output (each element is read and written by one thread);
// D3 must not.  real_out stores Re(value) into a double array.
struct Epi {
  double2 diag;
  double sacc;
  const double2* D1;
  double2 s1;
  const double2* D2;
  double2 s2;
  const double2* D3;
  double2 s3;
  int real_out;
  int abs_sum;  // accumulate sum_ij |value| (returned by the GEMM)
};

__device__ __forceinline__ Epi epi_store() {
  Epi e;
  e.diag = make_double2(0, 0); e.sacc = 1.0;
  e.D1 = nullptr; e.s1 = make_double2(1, 0);
  e.D2 = nullptr; e.s2 = make_double2(1, 0);
  e.D3 = nullptr; e.s3 = make_double2(1, 0);
  e.real_out = 0; e.abs_sum = 0;
  return e;
}

__device__ __forceinline__ double2 epi_value(const Epi& ep, int n, int i, int j, double ar, double ai) {
  double2 v = make_double2(ep.sacc * ar, ep.sacc * ai);
  if (i == j) v = cadd(v, ep.diag);
  if (ep.D1) v = cadd(v, cmul(ep.s1, gc(ep.D1)[i + (size_t)j * n]));
  if (ep.D2) v = cadd(v, cmul(ep.s2, gc(ep.D2)[i + (size_t)j * n]));
  if (ep.D3) v = cadd(v, cmul(ep.s3, cconj(gc(ep.D3)[j + (size_t)i * n])));
  return v;
}

// ---------------------------------------------------------------- reductions
// DPP lane exchange of a double within a row of 16 lanes (CTRL: a DPP pattern that
// reads only lanes of the same row).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double lane_f64(double v, int l) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, l);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), l);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Maximum over the wave (uniform result): quad swaps, half-row and row mirrors,
// one 16-lane swap, then the two 32-lane halves.
__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp_f64<0xB1>(v));   // quad_perm [1,0,3,2]
  v = fmax(v, dpp_f64<0x4E>(v));   // quad_perm [2,3,0,1]
  v = fmax(v, dpp_f64<0x141>(v));  // row_half_mirror
  v = fmax(v, dpp_f64<0x140>(v));  // row_mirror
  v = fmax(v, __shfl_xor(v, 16, 64));
  return fmax(lane_f64(v, 0), lane_f64(v, 32));
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over the workgroup; result valid in every thread.  red: >= 8 doubles of LDS.
__device__ __forceinline__ double block_sum(double v, ldsd* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
  __syncthreads();
  return s;
}

// Argmax of v over the workgroup, ties to the smallest index.
__device__ __forceinline__ void block_argmax(double& v, int& idx, ldsd* redv, ldsi* redi) {
  for (int off = 32; off > 0; off >>= 1) {
    double ov = __shfl_xor(v, off, 64);
    int oi = __shfl_xor(idx, off, 64);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { redv[wave] = v; redi[wave] = idx; }
  __syncthreads();
  v = redv[0]; idx = redi[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
    double ov = redv[w]; int oi = redi[w];
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
  __syncthreads();
}

// ------------------------------------------------------------ tile configs
// Per block-size variant V = T + 8 (OCC - 1): a pass computes an MR x MC macro
// tile of C with NW waves in a WR x WC grid, each wave owning TA x TB MFMA tiles
// of 16x16 complex (2*TA*TB accumulators of 4 doubles).  T = 4 (n = 128, 256)
// uses a 128 x 64 macro tile: 64 x 32 per wave keeps a wave at 128 accumulator
// registers, so nothing spills with one wave per SIMD.  Smaller T use square
// (32T)^2 tiles.  OCC is the number of workgroups per CU the variant's kernels
// are compiled for (their launch bounds): the same code at OCC = 2 is a
// separate instantiation of every block function, so the register budget of
// two workgroups per CU applies to it alone.
template <int V>
struct Cfg {
  static constexpr int T = V & 7;
  static constexpr int OCC = 1 + (V >> 3);
  static constexpr int MR = 32 * T;
  static constexpr int MC = (T == 4) ? 64 : 32 * T;
  static constexpr int NW = 4;
  static constexpr int WR = 2;
  static constexpr int WC = 2;
  static constexpr int TA = MR / (16 * WR);
  static constexpr int TB = MC / (16 * WC);
  static constexpr int NTH = 64 * NW;
  static constexpr int LDA = MR + PAD;  // LDS row stride of the op(A) panel
  static constexpr int LDB = MC + PAD;  // LDS row stride of the op(B) panel
  static constexpr int STAGE = 2 * KP * LDA + 2 * KP * LDB;  // re/im planes of both panels
  static constexpr int PERA = (MR * KP) / NTH;
  static constexpr int PERB = (MC * KP) / NTH;
  static_assert(TA * 16 * WR == MR && TB * 16 * WC == MC, "wave grid must tile the macro tile");
  static_assert((MR * KP) % NTH == 0 && (MC * KP) % NTH == 0, "panels must split evenly");
  static_assert(MR * LDK <= KP * LDA && MC * LDK <= KP * LDB, "k-contiguous panels fit the planes");
};

// ----------------------------------------------------------- MFMA micro-tile
// Accumulate NKS k-steps (4 complex k each) of a wave's TA x TB grid of 16x16
// complex tiles.  A values at Ar/Ai[k*lda + arow + m], B at Br/Bi[k*ldb + bcol + m].
// ROWMAJ=0: lane holds C[arow + ti*16 + l%16][bcol + tj*16 + 4r + l/16];
// ROWMAJ=1: lane holds C[arow + ti*16 + 4r + l/16][bcol + tj*16 + l%16].
// AKC / BKC: the panel is stored k-contiguous (A at Ar[(arow + m) * LDK + k], B at
// Br[(bcol + m) * LDK + k]) instead of row-contiguous (A at Ar[k * lda + arow + m]).
struct NoMid {
  __device__ void operator()(int) const {}
};

// mid(g), g = 0..3, runs after the g-th group of MFMAs of the last k-step, pinned
// there, so work of the caller (the LDS stores of the next panel) issues while
// the MFMA pipe is still busy.
template <int TA, int TB, int ROWMAJ, int NKS, bool AKC = false, bool BKC = false, typename Mid = NoMid>
__device__ __forceinline__ void mma_steps(d4 (&cr)[TA][TB], d4 (&ci)[TA][TB],
                                          const ldsd* Ar, const ldsd* Ai, int lda, int arow,
                                          const ldsd* Br, const ldsd* Bi, int ldb, int bcol,
                                          const Mid& mid = Mid()) {
  constexpr bool kMid = !__is_same(Mid, NoMid);
  const int lane = threadIdx.x & 63;
  const int m = lane & 15, kq = lane >> 4;
  // fragments of k-step ks + 1 are read while the MFMAs of k-step ks run
  double ar[2][TA], ai[2][TA], br[2][TB], bi[2][TB];
  auto frag = [&](int ks, int s) {
    const int k = ks * 4 + kq;
#pragma unroll
    for (int t = 0; t < TA; ++t) {
      const int ia = AKC ? (arow + t * 16 + m) * LDK + k : k * lda + arow + t * 16 + m;
      ar[s][t] = Ar[ia];
      ai[s][t] = Ai[ia];
    }
#pragma unroll
    for (int t = 0; t < TB; ++t) {
      const int ib = BKC ? (bcol + t * 16 + m) * LDK + k : k * ldb + bcol + t * 16 + m;
      br[s][t] = Br[ib];
      bi[s][t] = Bi[ib];
    }
  };
  frag(0, 0);
#pragma unroll
  for (int ks = 0; ks < NKS; ++ks) {
    const int s = ks & 1;
    if (ks + 1 < NKS) frag(ks + 1, s ^ 1);
    double nai[TA];
#pragma unroll
    for (int t = 0; t < TA; ++t) nai[t] = -ai[s][t];
    const bool hook = kMid && ks == NKS - 1;
#pragma unroll
    for (int ti = 0; ti < TA; ++ti)
#pragma unroll
      for (int tj = 0; tj < TB; ++tj)
        cr[ti][tj] = ROWMAJ ? __builtin_amdgcn_mfma_f64_16x16x4f64(ar[s][ti], br[s][tj], cr[ti][tj], 0, 0, 0)
                            : __builtin_amdgcn_mfma_f64_16x16x4f64(br[s][tj], ar[s][ti], cr[ti][tj], 0, 0, 0);
    if (hook) { __builtin_amdgcn_sched_barrier(0); mid(0); __builtin_amdgcn_sched_barrier(0); }
#pragma unroll
    for (int ti = 0; ti < TA; ++ti)
#pragma unroll
      for (int tj = 0; tj < TB; ++tj)
        ci[ti][tj] = ROWMAJ ? __builtin_amdgcn_mfma_f64_16x16x4f64(ar[s][ti], bi[s][tj], ci[ti][tj], 0, 0, 0)
                            : __builtin_amdgcn_mfma_f64_16x16x4f64(bi[s][tj], ar[s][ti], ci[ti][tj], 0, 0, 0);
    if (hook) { __builtin_amdgcn_sched_barrier(0); mid(1); __builtin_amdgcn_sched_barrier(0); }
#pragma unroll
    for (int ti = 0; ti < TA; ++ti)
#pragma unroll
      for (int tj = 0; tj < TB; ++tj)
        cr[ti][tj] = ROWMAJ ? __builtin_amdgcn_mfma_f64_16x16x4f64(nai[ti], bi[s][tj], cr[ti][tj], 0, 0, 0)
                            : __builtin_amdgcn_mfma_f64_16x16x4f64(bi[s][tj], nai[ti], cr[ti][tj], 0, 0, 0);
    if (hook) { __builtin_amdgcn_sched_barrier(0); mid(2); __builtin_amdgcn_sched_barrier(0); }
#pragma unroll
    for (int ti = 0; ti < TA; ++ti)
#pragma unroll
      for (int tj = 0; tj < TB; ++tj)
        ci[ti][tj] = ROWMAJ ? __builtin_amdgcn_mfma_f64_16x16x4f64(ai[s][ti], br[s][tj], ci[ti][tj], 0, 0, 0)
                            : __builtin_amdgcn_mfma_f64_16x16x4f64(br[s][tj], ai[s][ti], ci[ti][tj], 0, 0, 0);
    if (hook) { __builtin_amdgcn_sched_barrier(0); mid(3); __builtin_amdgcn_sched_barrier(0); }
  }
}

// ------------------------------------------------------------------- GEMM
// C = op(A) op(B) (+ epilogue), n x n column-major (or row-major C when CRM),
// n a multiple of MR, MC and 32.  K panels of 16 are double-buffered in LDS; the
// next panel is fetched into registers while the current one is multiplied,
// then stored to the other stage before the single barrier of the panel.  The
// last panel of a tile fetches panel 0 of the next tile, so only the first tile
// has an exposed prologue.  Global addresses are a wave-uniform origin (scalar
// registers, advanced per panel and tile) plus a 32-bit byte offset per thread
// that is fixed for the whole GEMM.  Returns sum |C_ij| when ep.abs_sum (same
// value in every thread).
template <int T>
struct GemmLds {
  static constexpr int DOUBLES = 2 * Cfg<T>::STAGE;  // two stages (the abs-sum reduction reuses stage 0)
};

// Byte offset of each of a thread's panel elements from the panel origin.
template <int T, int OPA, int OPB>
__device__ __forceinline__ void gemm_panel_offsets(int n, uint32_t (&oa)[Cfg<T>::PERA],
                                                   uint32_t (&ob)[Cfg<T>::PERB]) {
  constexpr int MR = Cfg<T>::MR, MC = Cfg<T>::MC, NTH = Cfg<T>::NTH;
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < Cfg<T>::PERA; ++q) {
    const int idx = q * NTH + tid;
    if (OPA == OP_N) {
      const int i = idx % MR, k = idx / MR;  // A(row0 + i, k0 + k)
      oa[q] = (uint32_t)(i + k * n) * 16u;
    } else {
      const int k = idx % KP, i = idx / KP;  // conj A(k0 + k, row0 + i)
      oa[q] = (uint32_t)(k + i * n) * 16u;
    }
  }
#pragma unroll
  for (int q = 0; q < Cfg<T>::PERB; ++q) {
    const int idx = q * NTH + tid;
    if (OPB == OP_N) {
      const int k = idx % KP, j = idx / KP;  // B(k0 + k, col0 + j)
      ob[q] = (uint32_t)(k + j * n) * 16u;
    } else {
      const int j = idx % MC, k = idx / MC;  // conj B(col0 + j, k0 + k)
      ob[q] = (uint32_t)(j + k * n) * 16u;
    }
  }
}

template <int T, int OPA, int OPB>
__device__ __forceinline__ void gemm_load_panel(const double2* Ao, const double2* Bo,
                                                const uint32_t (&oa)[Cfg<T>::PERA], const uint32_t (&ob)[Cfg<T>::PERB],
                                                double2 (&ra)[Cfg<T>::PERA], double2 (&rb)[Cfg<T>::PERB]) {
#pragma unroll
  for (int q = 0; q < Cfg<T>::PERA; ++q) {
    ra[q] = ldgb(Ao, oa[q]);
    if (OPA == OP_C) ra[q] = cconj(ra[q]);
  }
#pragma unroll
  for (int q = 0; q < Cfg<T>::PERB; ++q) {
    rb[q] = ldgb(Bo, ob[q]);
    if (OPB == OP_C) rb[q] = cconj(rb[q]);
  }
}

// part = -1: the whole panel; 0..3: the elements q with q % 4 == part
template <int T, int OPA, int OPB>
__device__ __forceinline__ void gemm_store_panel(ldsd* st, const double2 (&ra)[Cfg<T>::PERA],
                                                 const double2 (&rb)[Cfg<T>::PERB], int part = -1) {
  constexpr int MR = Cfg<T>::MR, MC = Cfg<T>::MC, LDA = Cfg<T>::LDA, LDB = Cfg<T>::LDB, NTH = Cfg<T>::NTH;
  ldsd* Ar = st;
  ldsd* Ai = Ar + KP * LDA;
  ldsd* Br = Ai + KP * LDA;
  ldsd* Bi = Br + KP * LDB;
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < Cfg<T>::PERA; ++q) {
    if (part >= 0 && q % 4 != part) continue;
    const int idx = q * NTH + tid;
    int i, k;
    if (OPA == OP_N) { i = idx % MR; k = idx / MR; } else { k = idx % KP; i = idx / KP; }
    // k-contiguous layout when the global read runs along k (OP_C): conflict-free stores
    const int o = (OPA == OP_N) ? k * LDA + i : i * LDK + k;
    Ar[o] = ra[q].x;
    Ai[o] = ra[q].y;
  }
#pragma unroll
  for (int q = 0; q < Cfg<T>::PERB; ++q) {
    if (part >= 0 && q % 4 != part) continue;
    const int idx = q * NTH + tid;
    int j, k;
    if (OPB == OP_N) { k = idx % KP; j = idx / KP; } else { j = idx % MC; k = idx / MC; }
    const int o = (OPB == OP_N) ? j * LDK + k : k * LDB + j;
    Br[o] = rb[q].x;
    Bi[o] = rb[q].y;
  }
}

template <int M>
struct EpiMode {
  static constexpr int value = M;
};

template <int T, int OPA, int OPB, int CRM>
__device__ __noinline__ double wg_gemm_k(int n_, const double2* __restrict__ A_, const double2* __restrict__ B_,
                                         void* Cout_, Epi ep_, ldsd* lds) {
  using C = Cfg<T>;
  constexpr int MR = C::MR, MC = C::MC, LDA = C::LDA, LDB = C::LDB, STAGE = C::STAGE, TA = C::TA, TB = C::TB;
  constexpr bool AKC = (OPA != OP_N), BKC = (OPB == OP_N);
  const int n = uni(n_);
  const double2* A = uni(A_);
  const double2* B = uni(B_);
  void* Cout = uni(Cout_);
  Epi ep;
  ep.diag = uni(ep_.diag); ep.sacc = uni(ep_.sacc);
  ep.D1 = uni(ep_.D1); ep.s1 = uni(ep_.s1);
  ep.D2 = uni(ep_.D2); ep.s2 = uni(ep_.s2);
  ep.D3 = uni(ep_.D3); ep.s3 = uni(ep_.s3);
  ep.real_out = uni(ep_.real_out); ep.abs_sum = uni(ep_.abs_sum);
  const int emode = (ep.D2 || ep.D3 || ep.real_out || ep.abs_sum) ? 2 : (ep.D1 ? 1 : 0);
  // the abs-sum reduction runs after the last panel's barrier, when stage 0 is free
  ldsd* red = lds;
  const ldsd* s0Ar = lds;
  const ldsd* s0Ai = s0Ar + KP * LDA;
  const ldsd* s0Br = s0Ai + KP * LDA;
  const ldsd* s0Bi = s0Br + KP * LDB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int arow = (wave % C::WR) * 16 * TA, bcol = (wave / C::WR) * 16 * TB;
  const int ntr = n / MR, ntiles = ntr * (n / MC), npan = n / KP;  // npan is even
  uint32_t oa[C::PERA], ob[C::PERB];
  gemm_panel_offsets<T, OPA, OPB>(n, oa, ob);
  // panel origins: A rows row0.. / B columns col0.., k from k0
  auto a_org = [&](int row0, int k0) -> const double2* {
    return OPA == OP_N ? A + row0 + (size_t)k0 * n : A + k0 + (size_t)row0 * n;
  };
  auto b_org = [&](int col0, int k0) -> const double2* {
    return OPB == OP_N ? B + k0 + (size_t)col0 * n : B + col0 + (size_t)k0 * n;
  };
  // epilogue: the thread's element of (ti, tj, r) is (row0 + it + di, col0 + jt + dj)
  const int m = lane & 15, kq = lane >> 4;
  const int it = CRM ? arow + kq : arow + m;
  const int jt = CRM ? bcol + m : bcol + kq;
  const uint32_t t_cm = (uint32_t)(it + jt * n), t_tr = (uint32_t)(jt + it * n);
  const uint32_t t_out = CRM ? t_tr : t_cm;
  double asum = 0.0;
  double2 ra[C::PERA], rb[C::PERB];
  gemm_load_panel<T, OPA, OPB>(a_org(0, 0), b_org(0, 0), oa, ob, ra, rb);
  gemm_store_panel<T, OPA, OPB>(lds, ra, rb);
  __syncthreads();
  for (int tile = 0; tile < ntiles; ++tile) {
    const int row0 = (tile % ntr) * MR, col0 = (tile / ntr) * MC;
    const int tn = tile + 1 < ntiles ? tile + 1 : tile;  // the last tile refetches its own panel 0
    const int nrow0 = (tn % ntr) * MR, ncol0 = (tn / ntr) * MC;
    d4 cr[TA][TB], ci[TA][TB];
#pragma unroll
    for (int a = 0; a < TA; ++a)
#pragma unroll
      for (int b = 0; b < TB; ++b) { cr[a][b] = d4{0, 0, 0, 0}; ci[a][b] = d4{0, 0, 0, 0}; }
    for (int p = 0; p < npan; p += 2) {
      // stage 0 holds panel p; panel p + 1 goes to stage 1
      gemm_load_panel<T, OPA, OPB>(a_org(row0, (p + 1) * KP), b_org(col0, (p + 1) * KP), oa, ob, ra, rb);
      mma_steps<TA, TB, CRM, KP / 4, AKC, BKC>(cr, ci, s0Ar, s0Ai, LDA, arow, s0Br, s0Bi, LDB, bcol,
                                              [&](int g) { gemm_store_panel<T, OPA, OPB>(lds + STAGE, ra, rb, g); });
      __syncthreads();
      // stage 1 holds panel p + 1; panel p + 2 (or panel 0 of the next tile) goes to stage 0
      const bool last = p + 2 == npan;
      gemm_load_panel<T, OPA, OPB>(last ? a_org(nrow0, 0) : a_org(row0, (p + 2) * KP),
                                   last ? b_org(ncol0, 0) : b_org(col0, (p + 2) * KP), oa, ob, ra, rb);
      mma_steps<TA, TB, CRM, KP / 4, AKC, BKC>(cr, ci, s0Ar + STAGE, s0Ai + STAGE, LDA, arow, s0Br + STAGE,
                                              s0Bi + STAGE, LDB, bcol,
                                              [&](int g) { gemm_store_panel<T, OPA, OPB>(lds, ra, rb, g); });
      __syncthreads();
    }
    // epilogue, in batches of one 16x16 tile (4 elements per thread); the D-term
    // loads of the next batch are issued before the current batch is combined.
    // The variant is chosen once per tile (no per-element tests of the epilogue
    // fields): 0 = scale (+ diagonal) and store, 1 = plus s1*D1, 2 = general.
    const size_t cm0 = row0 + (size_t)col0 * n, tr0 = col0 + (size_t)row0 * n;
    auto di = [&](int ti, int r) { return CRM ? ti * 16 + 4 * r : ti * 16; };
    auto dj = [&](int tj, int r) { return CRM ? tj * 16 : tj * 16 + 4 * r; };
    auto epilogue = [&](auto mode) {
      constexpr int MODE = decltype(mode)::value;
      auto load_d = [&](int b, double2 (&d1)[4], double2 (&d2)[4], double2 (&d3)[4]) {
        const int ti = b / TB, tj = b % TB;
        if (MODE == 1 || (MODE == 2 && ep.D1)) {
#pragma unroll
          for (int r = 0; r < 4; ++r) d1[r] = ldgb(ep.D1 + cm0 + di(ti, r) + (size_t)dj(tj, r) * n, 16u * t_cm);
        }
        if (MODE == 2 && ep.D2) {
#pragma unroll
          for (int r = 0; r < 4; ++r) d2[r] = ldgb(ep.D2 + cm0 + di(ti, r) + (size_t)dj(tj, r) * n, 16u * t_cm);
        }
        if (MODE == 2 && ep.D3) {
#pragma unroll
          for (int r = 0; r < 4; ++r) d3[r] = ldgb(ep.D3 + tr0 + dj(tj, r) + (size_t)di(ti, r) * n, 16u * t_tr);
        }
      };
      double2 n1[4], n2[4], n3[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) n1[r] = n2[r] = n3[r] = make_double2(0, 0);
      if (MODE > 0) load_d(0, n1, n2, n3);
#pragma unroll
      for (int b = 0; b < TA * TB; ++b) {
        const int ti = b / TB, tj = b % TB;
        double2 c1[4], c2[4], c3[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) { c1[r] = n1[r]; c2[r] = n2[r]; c3[r] = n3[r]; }
        if (MODE > 0 && b + 1 < TA * TB) load_d(b + 1, n1, n2, n3);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = row0 + it + di(ti, r), j = col0 + jt + dj(tj, r);
          double2 v = make_double2(ep.sacc * cr[ti][tj][r], ep.sacc * ci[ti][tj][r]);
          const double2 vd = cadd(v, ep.diag);
          v.x = (i == j) ? vd.x : v.x;
          v.y = (i == j) ? vd.y : v.y;
          if (MODE == 1) v = cadd(v, cmul(ep.s1, c1[r]));
          if (MODE == 2) {
            if (ep.D1) v = cadd(v, cmul(ep.s1, c1[r]));
            if (ep.D2) v = cadd(v, cmul(ep.s2, c2[r]));
            if (ep.D3) v = cadd(v, cmul(ep.s3, cconj(c3[r])));
            if (ep.abs_sum) asum += sqrt(v.y * v.y + v.x * v.x);
          }
          const size_t u = CRM ? tr0 + dj(tj, r) + (size_t)di(ti, r) * n : cm0 + di(ti, r) + (size_t)dj(tj, r) * n;
          if (MODE == 2 && ep.real_out)
            stgb_re((double*)Cout + u, 8u * t_out, v.x);
          else
            stgb((double2*)Cout + u, 16u * t_out, v);
        }
      }
    };
    if (emode == 0)
      epilogue(EpiMode<0>());
    else if (emode == 1)
      epilogue(EpiMode<1>());
    else
      epilogue(EpiMode<2>());
  }
  double s = 0.0;
  if (ep.abs_sum) s = block_sum(asum, red);
  __syncthreads();
  return s;
}

template <int T>
__device__ __forceinline__ double wg_gemm(int n, int opA, const double2* A, int opB, const double2* B, void* Cout,
                                          int c_rowmajor, Epi ep, ldsd* lds) {
  if (c_rowmajor) {
    if (opA == OP_N) {
      if (opB == OP_N) return wg_gemm_k<T, OP_N, OP_N, 1>(n, A, B, Cout, ep, lds);
      return wg_gemm_k<T, OP_N, OP_C, 1>(n, A, B, Cout, ep, lds);
    }
    if (opB == OP_N) return wg_gemm_k<T, OP_C, OP_N, 1>(n, A, B, Cout, ep, lds);
    return wg_gemm_k<T, OP_C, OP_C, 1>(n, A, B, Cout, ep, lds);
  }
  if (opA == OP_N) {
    if (opB == OP_N) return wg_gemm_k<T, OP_N, OP_N, 0>(n, A, B, Cout, ep, lds);
    return wg_gemm_k<T, OP_N, OP_C, 0>(n, A, B, Cout, ep, lds);
  }
  if (opB == OP_N) return wg_gemm_k<T, OP_C, OP_N, 0>(n, A, B, Cout, ep, lds);
  return wg_gemm_k<T, OP_C, OP_C, 0>(n, A, B, Cout, ep, lds);
}

// --------------------------------------------------------------- elementwise
// out(i,j) = diag*[i==j] + s1*D1(i,j) + s2*D2(i,j) + s3*conj(D3(j,i)); row-major
// output when out_rowmajor.  D1/D2 may alias out (not D3).
__device__ __noinline__ void wg_ew(int n, double2* out, int out_rowmajor, double2 diag,
                                   const double2* D1, double2 s1, const double2* D2, double2 s2,
                                   const double2* D3, double2 s3) {
  const int nn = n * n;
  for (int idx = threadIdx.x; idx < nn; idx += blockDim.x) {
    int i, j;
    if (out_rowmajor) { j = idx % n; i = idx / n; } else { i = idx % n; j = idx / n; }
    double2 v = (i == j) ? diag : make_double2(0, 0);
    if (D1) v = cadd(v, cmul(s1, gc(D1)[i + (size_t)j * n]));
    if (D2) v = cadd(v, cmul(s2, gc(D2)[i + (size_t)j * n]));
    if (D3) v = cadd(v, cmul(s3, cconj(gc(D3)[j + (size_t)i * n])));
    gw(out)[out_rowmajor ? (size_t)i * n + j : (size_t)i + (size_t)j * n] = v;
  }
  __syncthreads();
}

// out = (X - X^H) / 2, the anti-Hermitian part of X (n a multiple of 32, out must not
// alias X).  The 32 x 32 tile pairs (I, J), (J, I), I <= J, are read down their
// columns into LDS, the pair's two output tiles are written down their columns; the
// next pair's elements are loaded into registers while the current pair is combined.
// LDS: kAntihermLds doubles; 256 threads.
constexpr int kAntihermLds = 4 * 32 * 33;
__device__ __noinline__ void wg_antiherm(int n, const double2* X_, double2* out_, ldsd* lds) {
  constexpr int TS = 32, LD = TS + 1, PER = TS * TS / 256;
  static_assert(4 * TS * LD == kAntihermLds, "LDS size");
  const gcd2* X = gc(X_);
  gd2* out = gw(out_);
  ldsd* Ar = lds;
  ldsd* Ai = Ar + TS * LD;
  ldsd* Br = Ai + TS * LD;
  ldsd* Bi = Br + TS * LD;
  const int nt = n / TS, npair = nt * (nt + 1) / 2;
  const int tid = threadIdx.x;  // element k of a tile: row tid % 32, column tid / 32 + 8k
  const int r = tid % TS, c0 = tid / TS;
  auto pair_of = [&](int q, int& I, int& J) {  // q-th pair (I <= J), row by row
    I = 0;
    while (q >= nt - I) { q -= nt - I; ++I; }
    J = I + q;
  };
  double2 ra[PER], rb[PER];
  auto load = [&](int q) {
    int I, J;
    pair_of(q, I, J);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int c = c0 + 8 * k;
      ra[k] = X[(size_t)(I * TS + r) + (size_t)(J * TS + c) * n];
      rb[k] = X[(size_t)(J * TS + r) + (size_t)(I * TS + c) * n];
    }
  };
  load(0);
  for (int q = 0; q < npair; ++q) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int c = c0 + 8 * k;
      Ar[c * LD + r] = ra[k].x; Ai[c * LD + r] = ra[k].y;
      Br[c * LD + r] = rb[k].x; Bi[c * LD + r] = rb[k].y;
    }
    __syncthreads();
    if (q + 1 < npair) load(q + 1);
    int I, J;
    pair_of(q, I, J);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int c = c0 + 8 * k;
      // A = X[I-tile rows, J-tile cols] at (c, r) of the LDS tile, B likewise
      out[(size_t)(I * TS + r) + (size_t)(J * TS + c) * n] =
          make_double2(0.5 * (Ar[c * LD + r] - Br[r * LD + c]), 0.5 * (Ai[c * LD + r] + Bi[r * LD + c]));
      if (I != J)
        out[(size_t)(J * TS + r) + (size_t)(I * TS + c) * n] =
            make_double2(0.5 * (Br[c * LD + r] - Ar[r * LD + c]), 0.5 * (Bi[c * LD + r] + Ai[r * LD + c]));
    }
    __syncthreads();
  }
}

// Y = i Im X (real parts zeroed); Z = X as well when Z is given.  Y, Z must not alias X.
__device__ __noinline__ void wg_imag_part(int n, const double2* X, double2* Y, double2* Z) {
  const int nn = n * n;
  for (int idx = threadIdx.x; idx < nn; idx += blockDim.x) {
    const double2 v = gc(X)[idx];
    gw(Y)[idx] = make_double2(0.0, v.y);
    if (Z) gw(Z)[idx] = v;
  }
  __syncthreads();
}

// Re sum_ij X(i,j) Y(j,i) = Re Tr(X Y).
__device__ __noinline__ double wg_retrace_prod(int n, const double2* X, const double2* Y, ldsd* red) {
  double s = 0.0;
  const int nn = n * n;
  for (int idx = threadIdx.x; idx < nn; idx += blockDim.x) {
    const int i = idx % n, j = idx / n;
    double2 x = gc(X)[i + (size_t)j * n], y = gc(Y)[j + (size_t)i * n];
    s += x.x * y.x - x.y * y.y;
  }
  return block_sum(s, red);
}

// Im Tr(X).
__device__ __forceinline__ double wg_imtrace(int n, const double2* X, ldsd* red) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += gc(X)[i + (size_t)i * n].y;
  return block_sum(s, red);
}

// ------------------------------------------------------------------ inverse
// In-place blocked Gauss-Jordan inverse with partial row pivoting (pivot =
// first maximum of |Re|+|Im| among the rows not yet used, the izamax choice of
// zgetrf) and logical row exchanges.  W is row-major n x n and is destroyed;
// the inverse is written column-major to out.  Returns 0, or k+1 when column k
// has no nonzero pivot, in which case out is zeroed (src/green_rgf.f90:4429-4437).
//
// Panel step (16 columns K, pivot rows P): the panel is swept with one row per
// thread in registers, giving F_P = A_PK^-1 and F_R = -A_RK A_PK^-1; the other
// columns J are updated by one MFMA GEMM W[:,J] += (F - E) W[P,J] with E the pivot
// incidence (accumulating onto the loaded tiles of W), and columns K take F.
// out (a separate n x n buffer) receives inv(i, q) = W[piv[i]][pinv[q]] after the
// last panel.
template <int T, int NMAX>
struct InvSmallLds {
  static constexpr int LDI = NMAX + PAD;
  static constexpr int DOUBLES = 4 * KP * LDI + 2 * NMAX + 16;
};

template <int T, int NMAX>
__device__ __noinline__ int wg_inverse_small(int n, double2* W_, double2* out_, ldsd* lds) {
  using C = Cfg<T>;
  constexpr int LDI = NMAX + PAD;
  constexpr int MR = C::MR, MC = C::MC, NTH = C::NTH, TA = C::TA, TB = C::TB;
  gd2* W = gw(W_);
  gd2* out = gw(out_);
  ldsd* Mr = lds;
  ldsd* Mi = Mr + KP * LDI;
  ldsd* Br = Mi + KP * LDI;
  ldsd* Bi = Br + KP * LDI;
  ldsi* piv = (ldsi*)(Bi + KP * LDI);
  ldsi* used = piv + NMAX;
  ldsd* redv = (ldsd*)(used + NMAX);
  ldsi* redi = (ldsi*)(redv + 4);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int arow = (wave % C::WR) * 16 * TA, bcol = (wave / C::WR) * 16 * TB;
  static_assert(NMAX <= NTH, "one panel row per thread");
  for (int i = tid; i < n; i += NTH) used[i] = 0;
  __syncthreads();
  int info = 0;
  const bool own = tid < n;  // thread tid holds row tid of the current panel in registers
  bool myused = false;
  for (int k0 = 0; k0 < n && !info; k0 += KP) {
    double2 rv[KP];
#pragma unroll
    for (int c2 = 0; c2 < KP; ++c2) rv[c2] = own ? W[(size_t)tid * n + k0 + c2] : make_double2(0, 0);
    if (n <= 64) {
      // all rows are in wave 0: the column steps run in that wave alone, the pivot
      // row broadcast by readlane, no LDS exchange and no barrier per column
      if (wave == 0) {
        for (int c = 0; c < KP; ++c) {
          double2 cand = make_double2(0, 0);
#pragma unroll
          for (int c2 = 0; c2 < KP; ++c2)
            if (c2 == c) cand = rv[c2];
          const bool valid = own && !myused;
          const double myv = valid ? fabs(cand.x) + fabs(cand.y) : -1.0;
          const double best = wave_max(myv);
          if (!(best > 0.0)) { info = k0 + c + 1; break; }
          const int p = uni(__builtin_ctzll(__ballot(valid && myv == best)));  // lowest row
          double2 pr[KP];
#pragma unroll
          for (int c2 = 0; c2 < KP; ++c2) pr[c2] = make_double2(lane_f64(rv[c2].x, p), lane_f64(rv[c2].y, p));
          double2 pc = pr[0];
#pragma unroll
          for (int c2 = 1; c2 < KP; ++c2)
            if (c2 == c) pc = pr[c2];
          const double2 ainv = crecip(pc);
          double2 rp[KP];
#pragma unroll
          for (int c2 = 0; c2 < KP; ++c2) rp[c2] = (c2 == c) ? ainv : cmul(pr[c2], ainv);
          if (tid == p) {
            myused = true;
            used[p] = 1;
            piv[k0 + c] = p;
#pragma unroll
            for (int c2 = 0; c2 < KP; ++c2) rv[c2] = rp[c2];
          } else if (own) {
            const double2 f = cand;
#pragma unroll
            for (int c2 = 0; c2 < KP; ++c2) rv[c2] = csub((c2 == c) ? make_double2(0, 0) : rv[c2], cmul(f, rp[c2]));
          }
        }
        if (tid == 0) redi[0] = info;
      }
      __syncthreads();
      info = redi[0];
    } else {
      for (int c = 0; c < KP; ++c) {
        // pivot: first maximum of |Re|+|Im| over unused rows (wave reduction, then one
        // exchange of the 4 wave winners and their rows through LDS; Br/Bi hold the
        // exchange area until the snapshot below, double-buffered by column parity)
        double2 cand = make_double2(0, 0);
  #pragma unroll
        for (int c2 = 0; c2 < KP; ++c2)
          if (c2 == c) cand = rv[c2];
        const bool valid = own && !myused;
        const double myv = valid ? fabs(cand.x) + fabs(cand.y) : -1.0;
        double best = wave_max(myv);
        const unsigned long long hit = __ballot(valid && myv == best);  // lowest lane = lowest row
        int bi = hit ? wave * 64 + __builtin_ctzll(hit) : 0x7fffffff;
        ldsd* xrow = Br + (c & 1) * (4 * 2 * KP);  // [wave][c2] re, im
        ldsd* xval = Bi + (c & 1) * 8;               // [wave] value, [4 + wave] row index
        if (tid == bi) {
  #pragma unroll
          for (int c2 = 0; c2 < KP; ++c2) {
            xrow[wave * 2 * KP + 2 * c2] = rv[c2].x;
            xrow[wave * 2 * KP + 2 * c2 + 1] = rv[c2].y;
          }
        }
        if (lane == 0) { xval[wave] = best; xval[4 + wave] = (double)bi; }
        __syncthreads();
        best = xval[0];
        bi = (int)xval[4];
        for (int w = 1; w < NTH / 64; ++w) {
          const double ov = xval[w];
          const int oi = (int)xval[4 + w];
          if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
        }
        if (!(best > 0.0)) { info = k0 + c + 1; break; }
        const int p = bi;
        const ldsd* prow = xrow + (p >> 6) * 2 * KP;
        const double2 ainv = crecip(make_double2(prow[2 * c], prow[2 * c + 1]));
        double2 rp[KP];
  #pragma unroll
        for (int c2 = 0; c2 < KP; ++c2)
          rp[c2] = (c2 == c) ? ainv : cmul(make_double2(prow[2 * c2], prow[2 * c2 + 1]), ainv);
        if (tid == p) {
          myused = true;
          used[p] = 1;
          piv[k0 + c] = p;
  #pragma unroll
          for (int c2 = 0; c2 < KP; ++c2) rv[c2] = rp[c2];
        } else if (own) {
          const double2 f = cand;
  #pragma unroll
          for (int c2 = 0; c2 < KP; ++c2) rv[c2] = csub((c2 == c) ? make_double2(0, 0) : rv[c2], cmul(f, rp[c2]));
        }
      }
    }
    if (!info && own) {
#pragma unroll
      for (int c2 = 0; c2 < KP; ++c2) { Mr[c2 * LDI + tid] = rv[c2].x; Mi[c2 * LDI + tid] = rv[c2].y; }
    }
    __syncthreads();
    if (info) break;
    // snapshot of the pivot rows before they are updated: thread j copies column j
    // of the KP rows, all its loads in flight at once
    if (own) {
      double2 v[KP];
#pragma unroll
      for (int c = 0; c < KP; ++c) v[c] = W[(size_t)piv[k0 + c] * n + tid];
#pragma unroll
      for (int c = 0; c < KP; ++c) {
        Br[c * LDI + tid] = v[c].x;
        Bi[c * LDI + tid] = v[c].y;
      }
    }
    __syncthreads();
    if (tid < KP) Mr[tid * LDI + piv[k0 + tid]] -= 1.0;  // M = F - E
    __syncthreads();
    const int ntr = n / MR, ntc = n / MC;
    const int m = lane & 15, kq = lane >> 4;
    // a thread's element (ti, tj, r) of a tile is W[row0 + arow + kq + ti*16 + 4r][col0 + bcol + m + tj*16]
    const uint32_t t_off = 16u * (uint32_t)((arow + kq) * n + bcol + m);
    for (int tile = 0; tile < ntr * ntc; ++tile) {
      const int row0 = (tile % ntr) * MR, col0 = (tile / ntr) * MC;
      // W += M W[P, :]: the MFMAs accumulate onto the tile's old values, which are
      // all requested before the first MFMA waits for them
      d4 cr[TA][TB], ci[TA][TB];
#pragma unroll
      for (int ti = 0; ti < TA; ++ti)
#pragma unroll
        for (int tj = 0; tj < TB; ++tj)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double2 o = ldgb(W_ + (size_t)(row0 + ti * 16 + 4 * r) * n + col0 + tj * 16, t_off);
            cr[ti][tj][r] = o.x;
            ci[ti][tj][r] = o.y;
          }
      mma_steps<TA, TB, 1, KP / 4>(cr, ci, Mr, Mi, LDI, row0 + arow, Br, Bi, LDI, col0 + bcol);
      const bool panel = col0 <= k0 && k0 < col0 + MC;  // the panel's columns take F
      if (!panel) {
#pragma unroll
        for (int ti = 0; ti < TA; ++ti)
#pragma unroll
          for (int tj = 0; tj < TB; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              stgb(W_ + (size_t)(row0 + ti * 16 + 4 * r) * n + col0 + tj * 16, t_off,
                   make_double2(cr[ti][tj][r], ci[ti][tj][r]));
      } else {
#pragma unroll
        for (int ti = 0; ti < TA; ++ti)
#pragma unroll
          for (int tj = 0; tj < TB; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int i = row0 + arow + ti * 16 + 4 * r + kq;
              const int j = col0 + bcol + tj * 16 + m;
              const bool inp = j >= k0 && j < k0 + KP;
              const int c = inp ? j - k0 : 0;
              double2 f = make_double2(Mr[c * LDI + i], Mi[c * LDI + i]);
              f.x += (piv[k0 + c] == i) ? 1.0 : 0.0;
              const double2 v = inp ? f : make_double2(cr[ti][tj][r], ci[ti][tj][r]);
              stgb(W_ + (size_t)(row0 + ti * 16 + 4 * r) * n + col0 + tj * 16, t_off, v);
            }
      }
    }
    __syncthreads();
  }
  if (info) {
    for (int idx = tid; idx < n * n; idx += NTH) out[idx] = make_double2(0, 0);
    __syncthreads();
    return info;
  }
  // inverse permutation into `used`
  for (int q = tid; q < n; q += NTH) used[piv[q]] = q;
  __syncthreads();
  // out(i, q) = W[piv[i]][used[q]], RB rows of out at a time: the source rows are
  // read whole into LDS and out is written in column segments of RB elements
  constexpr int RB = 16;
  static_assert(2 * RB * (NMAX + 1) <= 4 * KP * LDI, "row block fits the panel area");
  const int ldr = n + 1;  // odd stride: the RB rows of one column sit in distinct banks
  ldsd* Lr = Mr;
  ldsd* Li = Mr + RB * ldr;
  for (int i0 = 0; i0 < n; i0 += RB) {
    if (own) {  // thread c copies column c of the RB source rows, all loads in flight
      double2 v[RB];
#pragma unroll
      for (int rr = 0; rr < RB; ++rr) v[rr] = W[(size_t)piv[i0 + rr] * n + tid];
#pragma unroll
      for (int rr = 0; rr < RB; ++rr) {
        Lr[rr * ldr + tid] = v[rr].x;
        Li[rr * ldr + tid] = v[rr].y;
      }
    }
    __syncthreads();
    for (int idx = tid; idx < RB * n; idx += NTH) {
      const int rr = idx % RB, q = idx / RB;
      const int c = used[q];
      out[(size_t)(i0 + rr) + (size_t)q * n] = make_double2(Lr[rr * ldr + c], Li[rr * ldr + c]);
    }
    __syncthreads();
  }
  return 0;
}

// ------------------------------------------------------ inverse, n > NTH
// The same Gauss-Jordan sweep for blocks larger than the workgroup (NMAX rows,
// RPT = NMAX / NTH panel rows per thread).  The swept panel M = F - E (n x KP,
// row-major) and the snapshot of the KP pivot rows (KP x n) do not fit in LDS at
// these sizes; they are kept in `out`, which is free until the final permutation
// copy, and staged per update tile through LDS (one K = KP panel of the tile's
// rows of M and columns of the snapshot), so the update of W is a K = 16 MFMA GEMM
// per tile accumulating onto the tile's old values as for n <= NTH.
template <int T, int NMAX>
struct InvBigLds {
  static constexpr int LDA = Cfg<T>::MR + PAD, LDB = Cfg<T>::MC + PAD;
  static constexpr int RB = 8;                                 // rows per step of the final copy
  static constexpr int PANELS = 2 * KP * LDA + 2 * KP * LDB;  // staged M rows and snapshot columns
  static constexpr int COPY = 2 * RB * (NMAX + 1);
  static constexpr int OFF_IDX = PANELS > COPY ? PANELS : COPY;  // piv, used after the larger area
  static constexpr int XCH = 2 * (4 * 2 * KP) + 2 * 8;          // pivot-row exchange, two parities
  static constexpr int DOUBLES = OFF_IDX + NMAX + XCH + 16;      // piv + used = 2 NMAX ints
};

__device__ __forceinline__ int wave_min_int(int v) {
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  return v;
}

template <int T, int NMAX>
__device__ __noinline__ int wg_inverse_big(int n, double2* W_, double2* out_, ldsd* lds) {
  using C = Cfg<T>;
  using L = InvBigLds<T, NMAX>;
  constexpr int NTH = C::NTH, RPT = (NMAX + NTH - 1) / NTH;
  constexpr int MR = C::MR, MC = C::MC, TA = C::TA, TB = C::TB, LDA = L::LDA, LDB = L::LDB;
  static_assert(RPT <= 8, "panel rows per thread");
  gd2* W = gw(W_);
  gd2* out = gw(out_);
  gd2* Mg = out;                       // M = F - E, row-major n x KP
  gd2* Sg = out + (size_t)n * KP;      // pivot-row snapshot, KP x n (row c at c * n)
  ldsd* Ar = lds;
  ldsd* Ai = Ar + KP * LDA;
  ldsd* Br = Ai + KP * LDA;
  ldsd* Bi = Br + KP * LDB;
  ldsi* piv = (ldsi*)(lds + L::OFF_IDX);
  ldsi* used = piv + NMAX;
  ldsd* xch = (ldsd*)(used + NMAX);  // [parity][wave][2 KP] rows, then [parity][8] values / rows
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int arow = (wave % C::WR) * 16 * TA, bcol = (wave / C::WR) * 16 * TB;
  for (int i = tid; i < n; i += NTH) used[i] = 0;
  __syncthreads();
  int info = 0;
  unsigned myused = 0;  // bit q: row tid + q * NTH has been a pivot
  for (int k0 = 0; k0 < n && !info; k0 += KP) {
    double2 rv[RPT][KP];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = tid + q * NTH;
#pragma unroll
      for (int c2 = 0; c2 < KP; ++c2) rv[q][c2] = r < n ? W[(size_t)r * n + k0 + c2] : make_double2(0, 0);
    }
    for (int c = 0; c < KP; ++c) {
      // the thread's candidate: first maximum of |Re| + |Im| over its unused rows
      double myv = -1.0;
      int myq = 0;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        double2 cand = make_double2(0, 0);
#pragma unroll
        for (int c2 = 0; c2 < KP; ++c2)
          if (c2 == c) cand = rv[q][c2];
        const bool valid = (tid + q * NTH) < n && !((myused >> q) & 1u);
        const double v = valid ? fabs(cand.x) + fabs(cand.y) : -1.0;
        if (v > myv) { myv = v; myq = q; }
      }
      const int myrow = tid + myq * NTH;
      double best = wave_max(myv);
      int brow = wave_min_int((myv == best && myv >= 0.0) ? myrow : 0x7fffffff);
      ldsd* xrow = xch + (c & 1) * (4 * 2 * KP);
      ldsd* xval = xch + 2 * (4 * 2 * KP) + (c & 1) * 8;
      if (myrow == brow) {
#pragma unroll
        for (int q = 0; q < RPT; ++q)
          if (q == myq) {
#pragma unroll
            for (int c2 = 0; c2 < KP; ++c2) {
              xrow[wave * 2 * KP + 2 * c2] = rv[q][c2].x;
              xrow[wave * 2 * KP + 2 * c2 + 1] = rv[q][c2].y;
            }
          }
      }
      if (lane == 0) { xval[wave] = best; xval[4 + wave] = (double)brow; }
      __syncthreads();
      best = xval[0];
      int bw = 0;
      int bi = (int)xval[4];
      for (int w = 1; w < NTH / 64; ++w) {
        const double ov = xval[w];
        const int oi = (int)xval[4 + w];
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; bw = w; }
      }
      if (!(best > 0.0)) { info = k0 + c + 1; break; }
      const int p = bi;
      const ldsd* prow = xrow + bw * 2 * KP;
      const double2 ainv = crecip(make_double2(prow[2 * c], prow[2 * c + 1]));
      double2 rp[KP];
#pragma unroll
      for (int c2 = 0; c2 < KP; ++c2)
        rp[c2] = (c2 == c) ? ainv : cmul(make_double2(prow[2 * c2], prow[2 * c2 + 1]), ainv);
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int r = tid + q * NTH;
        if (r == p) {
          myused |= 1u << q;
          used[p] = 1;
          piv[k0 + c] = p;
#pragma unroll
          for (int c2 = 0; c2 < KP; ++c2) rv[q][c2] = rp[c2];
        } else if (r < n) {
          double2 f = make_double2(0, 0);
#pragma unroll
          for (int c2 = 0; c2 < KP; ++c2)
            if (c2 == c) f = rv[q][c2];
#pragma unroll
          for (int c2 = 0; c2 < KP; ++c2) rv[q][c2] = csub((c2 == c) ? make_double2(0, 0) : rv[q][c2], cmul(f, rp[c2]));
        }
      }
    }
    if (info) break;
    __syncthreads();  // piv of this panel visible to every thread
    // M = F - E into out (row-major n x KP): a pivot row p = piv[k0 + c] has 1 subtracted at column c
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = tid + q * NTH;
      if (r < n) {
#pragma unroll
        for (int c2 = 0; c2 < KP; ++c2) {
          double2 v = rv[q][c2];
          v.x -= (piv[k0 + c2] == r) ? 1.0 : 0.0;
          Mg[(size_t)r * KP + c2] = v;
        }
      }
    }
    // snapshot of the pivot rows before the update changes them, all loads of a thread in flight
    for (int j0 = 0; j0 < n; j0 += NTH) {
      const int j = j0 + tid;
      if (j < n) {
        double2 v[KP];
#pragma unroll
        for (int c2 = 0; c2 < KP; ++c2) v[c2] = W[(size_t)piv[k0 + c2] * n + j];
#pragma unroll
        for (int c2 = 0; c2 < KP; ++c2) Sg[(size_t)c2 * n + j] = v[c2];
      }
    }
    __syncthreads();
    const int ntr = n / MR, ntc = n / MC;
    const int m = lane & 15, kq = lane >> 4;
    const uint32_t t_off = 16u * (uint32_t)((arow + kq) * n + bcol + m);
    for (int tile = 0; tile < ntr * ntc; ++tile) {
      const int row0 = (tile % ntr) * MR, col0 = (tile / ntr) * MC;
      d4 cr[TA][TB], ci[TA][TB];
#pragma unroll
      for (int ti = 0; ti < TA; ++ti)
#pragma unroll
        for (int tj = 0; tj < TB; ++tj)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double2 o = ldgb(W_ + (size_t)(row0 + ti * 16 + 4 * r) * n + col0 + tj * 16, t_off);
            cr[ti][tj][r] = o.x;
            ci[ti][tj][r] = o.y;
          }
      // stage M rows row0.. (k-major, [c][i]) and snapshot columns col0.. ([c][j])
      for (int idx = tid; idx < MR * KP; idx += NTH) {
        const int c2 = idx % KP, i = idx / KP;
        const double2 v = Mg[(size_t)(row0 + i) * KP + c2];
        Ar[c2 * LDA + i] = v.x;
        Ai[c2 * LDA + i] = v.y;
      }
      for (int idx = tid; idx < MC * KP; idx += NTH) {
        const int j = idx % MC, c2 = idx / MC;
        const double2 v = Sg[(size_t)c2 * n + col0 + j];
        Br[c2 * LDB + j] = v.x;
        Bi[c2 * LDB + j] = v.y;
      }
      __syncthreads();
      mma_steps<TA, TB, 1, KP / 4>(cr, ci, Ar, Ai, LDA, arow, Br, Bi, LDB, bcol);
      const bool panel = col0 <= k0 && k0 < col0 + MC;  // the panel's columns take F = M + E
      if (!panel) {
#pragma unroll
        for (int ti = 0; ti < TA; ++ti)
#pragma unroll
          for (int tj = 0; tj < TB; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              stgb(W_ + (size_t)(row0 + ti * 16 + 4 * r) * n + col0 + tj * 16, t_off,
                   make_double2(cr[ti][tj][r], ci[ti][tj][r]));
      } else {
#pragma unroll
        for (int ti = 0; ti < TA; ++ti)
#pragma unroll
          for (int tj = 0; tj < TB; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int il = arow + ti * 16 + 4 * r + kq;  // row within the tile
              const int j = col0 + bcol + tj * 16 + m;
              const bool inp = j >= k0 && j < k0 + KP;
              const int c2 = inp ? j - k0 : 0;
              double2 f = make_double2(Ar[c2 * LDA + il], Ai[c2 * LDA + il]);
              f.x += (piv[k0 + c2] == row0 + il) ? 1.0 : 0.0;
              const double2 v = inp ? f : make_double2(cr[ti][tj][r], ci[ti][tj][r]);
              stgb(W_ + (size_t)(row0 + ti * 16 + 4 * r) * n + col0 + tj * 16, t_off, v);
            }
      }
      __syncthreads();  // the staged panels are overwritten by the next tile
    }
  }
  if (info) {
    __syncthreads();
    for (int idx = tid; idx < n * n; idx += NTH) out[idx] = make_double2(0, 0);
    __syncthreads();
    return info;
  }
  for (int q = tid; q < n; q += NTH) used[piv[q]] = q;  // inverse permutation
  __syncthreads();
  // out(i, q) = W[piv[i]][used[q]], RB rows at a time staged through LDS
  constexpr int RB = L::RB;
  const int ldr = n + 1;
  ldsd* Lr = lds;
  ldsd* Li = lds + RB * ldr;
  for (int i0 = 0; i0 < n; i0 += RB) {
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int j = tid + q * NTH;
      if (j < n) {
        double2 v[RB];
#pragma unroll
        for (int rr = 0; rr < RB; ++rr) v[rr] = W[(size_t)piv[i0 + rr] * n + j];
#pragma unroll
        for (int rr = 0; rr < RB; ++rr) {
          Lr[rr * ldr + j] = v[rr].x;
          Li[rr * ldr + j] = v[rr].y;
        }
      }
    }
    __syncthreads();
    for (int idx = tid; idx < RB * n; idx += NTH) {
      const int rr = idx % RB, q = idx / RB;
      const int c = used[q];
      out[(size_t)(i0 + rr) + (size_t)q * n] = make_double2(Lr[rr * ldr + c], Li[rr * ldr + c]);
    }
    __syncthreads();
  }
  return 0;
}

// ------------------------------------------------ inverse, n = 32 or 64
// The whole matrix held in registers: thread t owns row i = t % N and the C = N*N / NTH
// columns cb*C .. cb*C + C - 1, cb = t / N (16 columns at N = 64, 4 at N = 32).  The
// same Gauss-Jordan sweep with the same pivot rule and logical row exchanges as the
// blocked inverse, unblocked: per column k the threads holding column k (one wave at
// N = 64, half a wave at N = 32) find the pivot row p and publish column k, the
// threads of row p publish row p, and every thread updates its elements in registers.
// Two barriers per column and no global memory between the load of W and the store
// of the inverse, where the blocked inverse at these sizes spends most of its time
// waiting on the loads and stores of each panel update.  Column and row exchange
// buffers alternate by column parity, so a column's reads never meet the next
// column's writes.
template <int T, int N>
struct InvRegLds {
  static constexpr int LDO = N + 1;           // odd row stride of the output staging tile
  static constexpr int XCH = 2 * 4 * N;       // [parity][col re, col im, row re, row im][N]
  static constexpr int OUT = 2 * N * LDO;     // re / im planes of the matrix for the final copy
  static constexpr int AREA = XCH > OUT ? XCH : OUT;
  static constexpr int DOUBLES = AREA + N + 8;  // + piv, pinv (2 N ints) + pivot slots
};

template <int T, int N>
__device__ __noinline__ int wg_inverse_reg(double2* W_, double2* out_, ldsd* lds) {
  using L = InvRegLds<T, N>;
  constexpr int NTH = Cfg<T>::NTH, C = N * N / NTH, LDO = L::LDO;
  static_assert(NTH == 256 && (N == 32 || N == 64), "register inverse: 256 threads, N = 32 or 64");
  gd2* W = gw(W_);
  gd2* out = gw(out_);
  ldsi* piv = (ldsi*)(lds + L::AREA);
  ldsi* pinv = piv + N;
  ldsi* pvs = pinv + N;  // [parity] pivot row of the current column (-1: none)
  const int tid = threadIdx.x;
  const int i = tid % N, cb = tid / N;
  double2 a[C];
#pragma unroll
  for (int c = 0; c < C; ++c) a[c] = W[(size_t)i * N + cb * C + c];
  bool myused = false;
  int info = 0;
  const int wave = tid >> 6;
  for (int k = 0; k < N; ++k) {
    ldsd* X = lds + (k & 1) * (4 * N);
    const int kb = k / C, kc = k % C;
    const bool mine = (cb == kb);
    // (a) the wave holding column k finds the pivot row and publishes the column
    if (wave == (kb * N) / 64) {
      double2 v = a[0];
#pragma unroll
      for (int c = 1; c < C; ++c)
        if (c == kc) v = a[c];
      const double m = (mine && !myused) ? fabs(v.x) + fabs(v.y) : -1.0;
      const double best = wave_max(m);
      const unsigned long long hit = __ballot(mine && !myused && m == best);
      if (mine) {
        X[i] = v.x;
        X[N + i] = v.y;
        if (i == 0) {
          const int p = hit ? (int)(__builtin_ctzll(hit) % N) : -1;  // lowest row among equal maxima
          pvs[k & 1] = (best > 0.0 && hit) ? p : -1;
          piv[k] = p;
        }
      }
    }
    __syncthreads();
    const int p = pvs[k & 1];
    if (p < 0) {
      info = k + 1;
      break;
    }
    // (b) the threads of row p normalise their part of the pivot row and publish it
    if (i == p) {
      const double2 ainv = crecip(make_double2(X[p], X[N + p]));
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const double2 r = (cb * C + c == k) ? ainv : cmul(a[c], ainv);
        a[c] = r;
        X[2 * N + cb * C + c] = r.x;
        X[3 * N + cb * C + c] = r.y;
      }
      myused = true;
    }
    __syncthreads();
    // (c) every other row: a -= a_ik * (pivot row), column k starting from zero
    if (i != p) {
      const double2 f = make_double2(X[i], X[N + i]);
      double2 rp[C];
#pragma unroll
      for (int c = 0; c < C; ++c) rp[c] = make_double2(X[2 * N + cb * C + c], X[3 * N + cb * C + c]);
      if (mine) {
#pragma unroll
        for (int c = 0; c < C; ++c)
          if (c == kc) a[c] = make_double2(0, 0);
      }
#pragma unroll
      for (int c = 0; c < C; ++c) a[c] = csub(a[c], cmul(f, rp[c]));
    }
  }
  if (info) {
    for (int idx = tid; idx < N * N; idx += NTH) out[idx] = make_double2(0, 0);
    __syncthreads();
    return info;
  }
  __syncthreads();  // the last column's exchange reads are done before the staging tile is written
  ldsd* Lr = lds;
  ldsd* Li = lds + N * LDO;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    Lr[i * LDO + cb * C + c] = a[c].x;
    Li[i * LDO + cb * C + c] = a[c].y;
  }
  for (int q = tid; q < N; q += NTH) pinv[piv[q]] = q;
  __syncthreads();
  // out(r, q) = W[piv[r]][pinv[q]], r fastest across the lanes: coalesced column-major stores
  for (int idx = tid; idx < N * N; idx += NTH) {
    const int r = idx % N, q = idx / N;
    const int src = piv[r] * LDO + pinv[q];
    out[idx] = make_double2(Lr[src], Li[src]);
  }
  __syncthreads();
  return 0;
}

template <int T, int NMAX>
struct InvLds {
  static constexpr int A =
      NMAX <= Cfg<T>::NTH ? InvSmallLds<T, NMAX>::DOUBLES : InvBigLds<T, NMAX>::DOUBLES;
  static constexpr int B = (NMAX == 32 || NMAX == 64) ? InvRegLds<T, NMAX>::DOUBLES : 0;
  static constexpr int DOUBLES = A > B ? A : B;
};

// Inverse of an n x n block (n <= NMAX): the whole matrix in registers when n = NMAX is
// 32 or 64, one panel row per thread up to NTH rows, several rows per thread with the
// panels staged through `out` above that.
template <int T, int NMAX>
__device__ __forceinline__ int wg_inverse(int n, double2* W, double2* out, ldsd* lds) {
  if constexpr (NMAX == 32 || NMAX == 64) {
    if (n == NMAX) return wg_inverse_reg<T, NMAX>(W, out, lds);
    return wg_inverse_small<T, NMAX>(n, W, out, lds);
  } else if constexpr (NMAX <= Cfg<T>::NTH) {
    return wg_inverse_small<T, NMAX>(n, W, out, lds);
  } else {
    return wg_inverse_big<T, NMAX>(n, W, out, lds);
  }
}

}  // namespace rgf
