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

enum { OP_N = 0, OP_C = 1 };

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
// Per block-size variant T: a pass computes an MR x MC macro tile of C with NW
// waves in a WR x WC grid, each wave owning TA x TB MFMA tiles of 16x16
// complex (2*TA*TB accumulators of 4 doubles).  T = 4 (n = 128, 256) uses a
// 128 x 64 macro tile: 64 x 32 per wave keeps a wave at 128 accumulator
// registers, so nothing spills with one wave per SIMD.  Smaller T use square
// (32T)^2 tiles.
template <int T>
struct Cfg {
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
template <int TA, int TB, int ROWMAJ, int NKS, bool AKC = false, bool BKC = false>
__device__ __forceinline__ void mma_steps(d4 (&cr)[TA][TB], d4 (&ci)[TA][TB],
                                          const ldsd* Ar, const ldsd* Ai, int lda, int arow,
                                          const ldsd* Br, const ldsd* Bi, int ldb, int bcol) {
  const int lane = threadIdx.x & 63;
  const int m = lane & 15, kq = lane >> 4;
  // not unrolled: unrolling lets the scheduler hoist every k-step's fragment
  // loads, which overflows the 256 registers a wave has at 2 waves per SIMD
#pragma unroll 1
  for (int ks = 0; ks < NKS; ++ks) {
    const int k = ks * 4 + kq;
    double ar[TA], ai[TA], nai[TA], br[TB], bi[TB];
#pragma unroll
    for (int t = 0; t < TA; ++t) {
      const int ia = AKC ? (arow + t * 16 + m) * LDK + k : k * lda + arow + t * 16 + m;
      ar[t] = Ar[ia];
      ai[t] = Ai[ia];
      nai[t] = -ai[t];
    }
#pragma unroll
    for (int t = 0; t < TB; ++t) {
      const int ib = BKC ? (bcol + t * 16 + m) * LDK + k : k * ldb + bcol + t * 16 + m;
      br[t] = Br[ib];
      bi[t] = Bi[ib];
    }
#pragma unroll
    for (int ti = 0; ti < TA; ++ti)
#pragma unroll
      for (int tj = 0; tj < TB; ++tj)
        cr[ti][tj] = ROWMAJ ? __builtin_amdgcn_mfma_f64_16x16x4f64(ar[ti], br[tj], cr[ti][tj], 0, 0, 0)
                            : __builtin_amdgcn_mfma_f64_16x16x4f64(br[tj], ar[ti], cr[ti][tj], 0, 0, 0);
#pragma unroll
    for (int ti = 0; ti < TA; ++ti)
#pragma unroll
      for (int tj = 0; tj < TB; ++tj)
        ci[ti][tj] = ROWMAJ ? __builtin_amdgcn_mfma_f64_16x16x4f64(ar[ti], bi[tj], ci[ti][tj], 0, 0, 0)
                            : __builtin_amdgcn_mfma_f64_16x16x4f64(bi[tj], ar[ti], ci[ti][tj], 0, 0, 0);
#pragma unroll
    for (int ti = 0; ti < TA; ++ti)
#pragma unroll
      for (int tj = 0; tj < TB; ++tj)
        cr[ti][tj] = ROWMAJ ? __builtin_amdgcn_mfma_f64_16x16x4f64(nai[ti], bi[tj], cr[ti][tj], 0, 0, 0)
                            : __builtin_amdgcn_mfma_f64_16x16x4f64(bi[tj], nai[ti], cr[ti][tj], 0, 0, 0);
#pragma unroll
    for (int ti = 0; ti < TA; ++ti)
#pragma unroll
      for (int tj = 0; tj < TB; ++tj)
        ci[ti][tj] = ROWMAJ ? __builtin_amdgcn_mfma_f64_16x16x4f64(ai[ti], br[tj], ci[ti][tj], 0, 0, 0)
                            : __builtin_amdgcn_mfma_f64_16x16x4f64(br[tj], ai[ti], ci[ti][tj], 0, 0, 0);
  }
}

// ------------------------------------------------------------------- GEMM
// C = op(A) op(B) (+ epilogue), n x n column-major (or row-major C when
// c_rowmajor), n a multiple of MR and MC.  K panels of 16 are double-buffered in LDS;
// the next panel is fetched into registers while the current one is
// multiplied, then stored to the other stage before the single barrier of the
// panel.  Returns sum |C_ij| when ep.abs_sum (same value in every thread).
template <int T>
struct GemmLds {
  static constexpr int DOUBLES = 2 * Cfg<T>::STAGE + 8;  // two stages + reduction scratch
};

template <int T, int OPA, int OPB>
__device__ __forceinline__ void gemm_load_panel(int n, const double2* __restrict__ A, const double2* __restrict__ B,
                                                int row0, int col0, int k0,
                                                double2 (&ra)[Cfg<T>::PERA], double2 (&rb)[Cfg<T>::PERB]) {
  constexpr int MR = Cfg<T>::MR, MC = Cfg<T>::MC, NTH = Cfg<T>::NTH;
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < Cfg<T>::PERA; ++q) {
    const int idx = q * NTH + tid;
    if (OPA == OP_N) {
      const int i = idx % MR, k = idx / MR;
      ra[q] = ldg(A, (size_t)(row0 + i) + (size_t)(k0 + k) * n);
    } else {
      const int k = idx % KP, i = idx / KP;
      ra[q] = cconj(ldg(A, (size_t)(k0 + k) + (size_t)(row0 + i) * n));
    }
  }
#pragma unroll
  for (int q = 0; q < Cfg<T>::PERB; ++q) {
    const int idx = q * NTH + tid;
    if (OPB == OP_N) {
      const int k = idx % KP, j = idx / KP;
      rb[q] = ldg(B, (size_t)(k0 + k) + (size_t)(col0 + j) * n);
    } else {
      const int j = idx % MC, k = idx / MC;
      rb[q] = cconj(ldg(B, (size_t)(col0 + j) + (size_t)(k0 + k) * n));
    }
  }
}

template <int T, int OPA, int OPB>
__device__ __forceinline__ void gemm_store_panel(ldsd* st, const double2 (&ra)[Cfg<T>::PERA],
                                                 const double2 (&rb)[Cfg<T>::PERB]) {
  constexpr int MR = Cfg<T>::MR, MC = Cfg<T>::MC, LDA = Cfg<T>::LDA, LDB = Cfg<T>::LDB, NTH = Cfg<T>::NTH;
  ldsd* Ar = st;
  ldsd* Ai = Ar + KP * LDA;
  ldsd* Br = Ai + KP * LDA;
  ldsd* Bi = Br + KP * LDB;
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < Cfg<T>::PERA; ++q) {
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
    const int idx = q * NTH + tid;
    int j, k;
    if (OPB == OP_N) { k = idx % KP; j = idx / KP; } else { j = idx % MC; k = idx / MC; }
    const int o = (OPB == OP_N) ? j * LDK + k : k * LDB + j;
    Br[o] = rb[q].x;
    Bi[o] = rb[q].y;
  }
}

template <int T, int OPA, int OPB>
__device__ __noinline__ double wg_gemm_k(int n_, const double2* __restrict__ A_, const double2* __restrict__ B_,
                                         void* Cout_, int c_rowmajor_, Epi ep_, ldsd* lds) {
  using C = Cfg<T>;
  constexpr int MR = C::MR, MC = C::MC, LDA = C::LDA, LDB = C::LDB, STAGE = C::STAGE, TA = C::TA, TB = C::TB;
  const int n = uni(n_);
  const double2* A = uni(A_);
  const double2* B = uni(B_);
  void* Cout = uni(Cout_);
  const int c_rowmajor = uni(c_rowmajor_);
  Epi ep;
  ep.diag = uni(ep_.diag); ep.sacc = uni(ep_.sacc);
  ep.D1 = uni(ep_.D1); ep.s1 = uni(ep_.s1);
  ep.D2 = uni(ep_.D2); ep.s2 = uni(ep_.s2);
  ep.D3 = uni(ep_.D3); ep.s3 = uni(ep_.s3);
  ep.real_out = uni(ep_.real_out); ep.abs_sum = uni(ep_.abs_sum);
  ldsd* red = lds + 2 * STAGE;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int arow = (wave % C::WR) * 16 * TA, bcol = (wave / C::WR) * 16 * TB;
  const int ntr = n / MR, ntc = n / MC, npan = n / KP;
  double asum = 0.0;
  for (int tile = 0; tile < ntr * ntc; ++tile) {
    const int row0 = (tile % ntr) * MR, col0 = (tile / ntr) * MC;
    d4 cr[TA][TB], ci[TA][TB];
#pragma unroll
    for (int a = 0; a < TA; ++a)
#pragma unroll
      for (int b = 0; b < TB; ++b) { cr[a][b] = d4{0, 0, 0, 0}; ci[a][b] = d4{0, 0, 0, 0}; }
    double2 ra[C::PERA], rb[C::PERB];
    gemm_load_panel<T, OPA, OPB>(n, A, B, row0, col0, 0, ra, rb);
    gemm_store_panel<T, OPA, OPB>(lds, ra, rb);
    __syncthreads();
    for (int p = 0; p < npan; ++p) {
      // branch-free pipeline: the last pass refetches its own panel into the idle stage
      const int kn = (p + 1 < npan ? p + 1 : p) * KP;
      gemm_load_panel<T, OPA, OPB>(n, A, B, row0, col0, kn, ra, rb);
      const ldsd* st = lds + (p & 1) * STAGE;
      const ldsd* sAr = st;
      const ldsd* sAi = sAr + KP * LDA;
      const ldsd* sBr = sAi + KP * LDA;
      const ldsd* sBi = sBr + KP * LDB;
      constexpr bool AKC = (OPA == OP_C), BKC = (OPB == OP_N);
      if (c_rowmajor)
        mma_steps<TA, TB, 1, KP / 4, AKC, BKC>(cr, ci, sAr, sAi, LDA, arow, sBr, sBi, LDB, bcol);
      else
        mma_steps<TA, TB, 0, KP / 4, AKC, BKC>(cr, ci, sAr, sAi, LDA, arow, sBr, sBi, LDB, bcol);
      gemm_store_panel<T, OPA, OPB>(lds + ((p + 1) & 1) * STAGE, ra, rb);
      __syncthreads();
    }
    // epilogue
    const int m = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < TA; ++ti)
#pragma unroll
      for (int tj = 0; tj < TB; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int i, j;
          if (c_rowmajor) {
            i = row0 + arow + ti * 16 + 4 * r + kq;
            j = col0 + bcol + tj * 16 + m;
          } else {
            i = row0 + arow + ti * 16 + m;
            j = col0 + bcol + tj * 16 + 4 * r + kq;
          }
          double2 v = epi_value(ep, n, i, j, cr[ti][tj][r], ci[ti][tj][r]);
          if (ep.abs_sum) asum += sqrt(v.y * v.y + v.x * v.x);
          size_t off = c_rowmajor ? (size_t)i * n + j : (size_t)i + (size_t)j * n;
          if (ep.real_out)
            ((gd*)Cout)[off] = v.x;
          else
            ((gd2*)Cout)[off] = v;
        }
  }
  double s = 0.0;
  if (ep.abs_sum) s = block_sum(asum, red);
  __syncthreads();
  return s;
}

template <int T>
__device__ __forceinline__ double wg_gemm(int n, int opA, const double2* A, int opB, const double2* B, void* Cout,
                                          int c_rowmajor, Epi ep, ldsd* lds) {
  if (opA == OP_N) {
    if (opB == OP_N) return wg_gemm_k<T, OP_N, OP_N>(n, A, B, Cout, c_rowmajor, ep, lds);
    return wg_gemm_k<T, OP_N, OP_C>(n, A, B, Cout, c_rowmajor, ep, lds);
  }
  if (opB == OP_N) return wg_gemm_k<T, OP_C, OP_N>(n, A, B, Cout, c_rowmajor, ep, lds);
  return wg_gemm_k<T, OP_C, OP_C>(n, A, B, Cout, c_rowmajor, ep, lds);
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
// Panel step (16 columns K, pivot rows P): the panel is swept in LDS, giving
// F_P = A_PK^-1 and F_R = -A_RK A_PK^-1; the other columns J are updated by one
// MFMA GEMM W[:,J] += (F - E) W[P,J] with E the pivot incidence, and columns K
// take F.  After all panels, inv(i, q) = W[piv[i]][pinv[q]].
template <int T, int NMAX>
struct InvLds {
  static constexpr int LDI = NMAX + PAD;
  static constexpr int DOUBLES = 4 * KP * LDI + 2 * NMAX + 16;
};

template <int T, int NMAX>
__device__ __noinline__ int wg_inverse(int n, double2* W_, double2* out_, ldsd* lds) {
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
    for (int c = 0; c < KP; ++c) {
      // pivot: first maximum of |Re|+|Im| over unused rows (wave reduction, then one
      // exchange of the 4 wave winners and their rows through LDS; Br/Bi hold the
      // exchange area until the snapshot below, double-buffered by column parity)
      double best = -1.0;
      int bi = 0x7fffffff;
      double2 cand = make_double2(0, 0);
#pragma unroll
      for (int c2 = 0; c2 < KP; ++c2)
        if (c2 == c) cand = rv[c2];
      if (own && !myused) { best = fabs(cand.x) + fabs(cand.y); bi = tid; }
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
      }
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
    if (!info && own) {
#pragma unroll
      for (int c2 = 0; c2 < KP; ++c2) { Mr[c2 * LDI + tid] = rv[c2].x; Mi[c2 * LDI + tid] = rv[c2].y; }
    }
    __syncthreads();
    if (info) break;
    // snapshot of the pivot rows (all columns) before they are updated
    for (int idx = tid; idx < KP * n; idx += NTH) {
      const int c = idx / n, j = idx % n;
      const double2 v = W[(size_t)piv[k0 + c] * n + j];
      Br[c * LDI + j] = v.x;
      Bi[c * LDI + j] = v.y;
    }
    __syncthreads();
    if (tid < KP) Mr[tid * LDI + piv[k0 + tid]] -= 1.0;  // M = F - E
    __syncthreads();
    const int ntr = n / MR, ntc = n / MC;
    const int m = lane & 15, kq = lane >> 4;
    for (int tile = 0; tile < ntr * ntc; ++tile) {
      const int row0 = (tile % ntr) * MR, col0 = (tile / ntr) * MC;
      d4 cr[TA][TB], ci[TA][TB];
#pragma unroll
      for (int a = 0; a < TA; ++a)
#pragma unroll
        for (int b = 0; b < TB; ++b) { cr[a][b] = d4{0, 0, 0, 0}; ci[a][b] = d4{0, 0, 0, 0}; }
      mma_steps<TA, TB, 1, KP / 4>(cr, ci, Mr, Mi, LDI, row0 + arow, Br, Bi, LDI, col0 + bcol);
#pragma unroll
      for (int ti = 0; ti < TA; ++ti)
#pragma unroll
        for (int tj = 0; tj < TB; ++tj)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = row0 + arow + ti * 16 + 4 * r + kq;
            const int j = col0 + bcol + tj * 16 + m;
            gd2* w = &W[(size_t)i * n + j];
            if (j >= k0 && j < k0 + KP) {
              const int c = j - k0;
              double2 v = make_double2(Mr[c * LDI + i], Mi[c * LDI + i]);
              if (piv[j] == i) v.x += 1.0;
              *w = v;
            } else {
              const double2 o = *w;
              *w = make_double2(o.x + cr[ti][tj][r], o.y + ci[ti][tj][r]);
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
  for (int idx = tid; idx < n * n; idx += NTH) {
    const int i = idx % n, q = idx / n;
    out[idx] = W[(size_t)piv[i] * n + used[q]];
  }
  __syncthreads();
  return 0;
}

}  // namespace rgf
