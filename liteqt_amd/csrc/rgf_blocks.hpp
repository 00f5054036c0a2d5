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

namespace rgf {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;   // threads per workgroup
constexpr int PAD = 16;   // LDS row padding (doubles): row stride = 16 mod 32 doubles
constexpr int KP = 16;    // K panel depth (complex)

enum { OP_N = 0, OP_C = 1 };

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
  if (ep.D1) v = cadd(v, cmul(ep.s1, ep.D1[i + (size_t)j * n]));
  if (ep.D2) v = cadd(v, cmul(ep.s2, ep.D2[i + (size_t)j * n]));
  if (ep.D3) v = cadd(v, cmul(ep.s3, cconj(ep.D3[j + (size_t)i * n])));
  return v;
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over the workgroup; result valid in every thread.  red: >= 4 doubles of LDS.
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return s;
}

// Argmax of v over the workgroup, ties to the smallest index.
__device__ __forceinline__ void block_argmax(double& v, int& idx, double* redv, int* redi) {
  for (int off = 32; off > 0; off >>= 1) {
    double ov = __shfl_xor(v, off, 64);
    int oi = __shfl_xor(idx, off, 64);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { redv[wave] = v; redi[wave] = idx; }
  __syncthreads();
  v = redv[0]; idx = redi[0];
  for (int w = 1; w < NT / 64; ++w) {
    double ov = redv[w]; int oi = redi[w];
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
  __syncthreads();
}

// ----------------------------------------------------------- MFMA micro-tile
// Accumulate nks k-steps (4 complex k each) of a wave's T x T grid of 16x16
// complex tiles.  A values at Ar/Ai[k*lda + arow + m], B at Br/Bi[k*ldb + bcol + m].
// ROWMAJ=0: lane holds C[arow + ti*16 + l%16][bcol + tj*16 + 4r + l/16];
// ROWMAJ=1: lane holds C[arow + ti*16 + 4r + l/16][bcol + tj*16 + l%16].
template <int T, int ROWMAJ>
__device__ __forceinline__ void mma_steps(d4 (&cr)[T][T], d4 (&ci)[T][T],
                                          const double* Ar, const double* Ai, int lda, int arow,
                                          const double* Br, const double* Bi, int ldb, int bcol,
                                          int nks) {
  const int lane = threadIdx.x & 63;
  const int m = lane & 15, kq = lane >> 4;
  for (int ks = 0; ks < nks; ++ks) {
    const int k = ks * 4 + kq;
    double ar[T], ai[T], nai[T], br[T], bi[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
      ar[t] = Ar[k * lda + arow + t * 16 + m];
      ai[t] = Ai[k * lda + arow + t * 16 + m];
      nai[t] = -ai[t];
      br[t] = Br[k * ldb + bcol + t * 16 + m];
      bi[t] = Bi[k * ldb + bcol + t * 16 + m];
    }
    if (ROWMAJ) {
#pragma unroll
      for (int ti = 0; ti < T; ++ti)
#pragma unroll
        for (int tj = 0; tj < T; ++tj) cr[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[ti], br[tj], cr[ti][tj], 0, 0, 0);
#pragma unroll
      for (int ti = 0; ti < T; ++ti)
#pragma unroll
        for (int tj = 0; tj < T; ++tj) ci[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[ti], bi[tj], ci[ti][tj], 0, 0, 0);
#pragma unroll
      for (int ti = 0; ti < T; ++ti)
#pragma unroll
        for (int tj = 0; tj < T; ++tj) cr[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[ti], bi[tj], cr[ti][tj], 0, 0, 0);
#pragma unroll
      for (int ti = 0; ti < T; ++ti)
#pragma unroll
        for (int tj = 0; tj < T; ++tj) ci[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[ti], br[tj], ci[ti][tj], 0, 0, 0);
    } else {
#pragma unroll
      for (int ti = 0; ti < T; ++ti)
#pragma unroll
        for (int tj = 0; tj < T; ++tj) cr[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(br[tj], ar[ti], cr[ti][tj], 0, 0, 0);
#pragma unroll
      for (int ti = 0; ti < T; ++ti)
#pragma unroll
        for (int tj = 0; tj < T; ++tj) ci[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(bi[tj], ar[ti], ci[ti][tj], 0, 0, 0);
#pragma unroll
      for (int ti = 0; ti < T; ++ti)
#pragma unroll
        for (int tj = 0; tj < T; ++tj) cr[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(bi[tj], nai[ti], cr[ti][tj], 0, 0, 0);
#pragma unroll
      for (int ti = 0; ti < T; ++ti)
#pragma unroll
        for (int tj = 0; tj < T; ++tj) ci[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(br[tj], ai[ti], ci[ti][tj], 0, 0, 0);
    }
  }
}

// ------------------------------------------------------------------- GEMM
// C = op(A) op(B) (+ epilogue), n x n column-major (or row-major C when
// c_rowmajor), n a multiple of 32*T.  Macro tile (32T)^2 per pass, 4 waves in
// a 2x2 grid of (16T)^2 wave tiles, K panels of 16 staged in LDS with the next
// panel prefetched into registers while the current one is multiplied.
// Returns sum |C_ij| when ep.abs_sum (same value in every thread).
template <int T>
struct GemmLds {
  static constexpr int MT = 32 * T;
  static constexpr int LD = MT + PAD;
  static constexpr int DOUBLES = 4 * KP * LD + 8;
};

template <int T>
__device__ __forceinline__ void gemm_load_panel(int n, int opA, const double2* __restrict__ A, int opB,
                                                const double2* __restrict__ B, int row0, int col0, int k0,
                                                double2 (&ra)[(32 * T * KP) / NT], double2 (&rb)[(32 * T * KP) / NT]) {
  constexpr int MT = 32 * T;
  constexpr int PER = (MT * KP) / NT;
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int idx = q * NT + tid;
    if (opA == OP_N) {
      const int i = idx % MT, k = idx / MT;
      ra[q] = A[(size_t)(row0 + i) + (size_t)(k0 + k) * n];
    } else {
      const int k = idx % KP, i = idx / KP;
      ra[q] = cconj(A[(size_t)(k0 + k) + (size_t)(row0 + i) * n]);
    }
    if (opB == OP_N) {
      const int k = idx % KP, j = idx / KP;
      rb[q] = B[(size_t)(k0 + k) + (size_t)(col0 + j) * n];
    } else {
      const int j = idx % MT, k = idx / MT;
      rb[q] = cconj(B[(size_t)(col0 + j) + (size_t)(k0 + k) * n]);
    }
  }
}

template <int T>
__device__ __forceinline__ void gemm_store_panel(int opA, int opB, double* lds,
                                                 const double2 (&ra)[(32 * T * KP) / NT],
                                                 const double2 (&rb)[(32 * T * KP) / NT]) {
  constexpr int MT = 32 * T;
  constexpr int LD = MT + PAD;
  constexpr int PER = (MT * KP) / NT;
  double* Ar = lds;
  double* Ai = Ar + KP * LD;
  double* Br = Ai + KP * LD;
  double* Bi = Br + KP * LD;
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int idx = q * NT + tid;
    int i, k;
    if (opA == OP_N) { i = idx % MT; k = idx / MT; } else { k = idx % KP; i = idx / KP; }
    Ar[k * LD + i] = ra[q].x;
    Ai[k * LD + i] = ra[q].y;
    int j, kb;
    if (opB == OP_N) { kb = idx % KP; j = idx / KP; } else { j = idx % MT; kb = idx / MT; }
    Br[kb * LD + j] = rb[q].x;
    Bi[kb * LD + j] = rb[q].y;
  }
}

template <int T>
__device__ __noinline__ double wg_gemm(int n, int opA, const double2* __restrict__ A, int opB,
                                       const double2* __restrict__ B, void* Cout, int c_rowmajor,
                                       Epi ep, double* lds) {
  constexpr int MT = 32 * T;
  constexpr int LD = MT + PAD;
  constexpr int PER = (MT * KP) / NT;
  static_assert((MT * KP) % NT == 0, "panel must split evenly over the workgroup");
  double* Ar = lds;
  double* Ai = Ar + KP * LD;
  double* Br = Ai + KP * LD;
  double* Bi = Br + KP * LD;
  double* red = Bi + KP * LD;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave & 1, wc = wave >> 1;
  const int nt = n / MT, npan = n / KP;
  double asum = 0.0;
  for (int tile = 0; tile < nt * nt; ++tile) {
    const int row0 = (tile % nt) * MT, col0 = (tile / nt) * MT;
    d4 cr[T][T], ci[T][T];
#pragma unroll
    for (int a = 0; a < T; ++a)
#pragma unroll
      for (int b = 0; b < T; ++b) { cr[a][b] = d4{0, 0, 0, 0}; ci[a][b] = d4{0, 0, 0, 0}; }
    double2 ra[PER], rb[PER];
    gemm_load_panel<T>(n, opA, A, opB, B, row0, col0, 0, ra, rb);
    gemm_store_panel<T>(opA, opB, lds, ra, rb);
    __syncthreads();
    for (int p = 0; p < npan; ++p) {
      const bool more = (p + 1 < npan);
      if (more) gemm_load_panel<T>(n, opA, A, opB, B, row0, col0, (p + 1) * KP, ra, rb);
      if (c_rowmajor)
        mma_steps<T, 1>(cr, ci, Ar, Ai, LD, wr * 16 * T, Br, Bi, LD, wc * 16 * T, KP / 4);
      else
        mma_steps<T, 0>(cr, ci, Ar, Ai, LD, wr * 16 * T, Br, Bi, LD, wc * 16 * T, KP / 4);
      __syncthreads();
      if (more) {
        gemm_store_panel<T>(opA, opB, lds, ra, rb);
        __syncthreads();
      }
    }
    // epilogue
    const int m = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < T; ++ti)
#pragma unroll
      for (int tj = 0; tj < T; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int i, j;
          if (c_rowmajor) {
            i = row0 + wr * 16 * T + ti * 16 + 4 * r + kq;
            j = col0 + wc * 16 * T + tj * 16 + m;
          } else {
            i = row0 + wr * 16 * T + ti * 16 + m;
            j = col0 + wc * 16 * T + tj * 16 + 4 * r + kq;
          }
          double2 v = epi_value(ep, n, i, j, cr[ti][tj][r], ci[ti][tj][r]);
          if (ep.abs_sum) asum += sqrt(v.y * v.y + v.x * v.x);
          size_t off = c_rowmajor ? (size_t)i * n + j : (size_t)i + (size_t)j * n;
          if (ep.real_out)
            ((double*)Cout)[off] = v.x;
          else
            ((double2*)Cout)[off] = v;
        }
  }
  double s = 0.0;
  if (ep.abs_sum) s = block_sum(asum, red);
  __syncthreads();
  return s;
}

// --------------------------------------------------------------- elementwise
// out(i,j) = diag*[i==j] + s1*D1(i,j) + s2*D2(i,j) + s3*conj(D3(j,i)); row-major
// output when out_rowmajor.  D1/D2 may alias out (not D3).
__device__ __noinline__ void wg_ew(int n, double2* out, int out_rowmajor, double2 diag,
                                   const double2* D1, double2 s1, const double2* D2, double2 s2,
                                   const double2* D3, double2 s3) {
  const int nn = n * n;
  for (int idx = threadIdx.x; idx < nn; idx += NT) {
    int i, j;
    if (out_rowmajor) { j = idx % n; i = idx / n; } else { i = idx % n; j = idx / n; }
    double2 v = (i == j) ? diag : make_double2(0, 0);
    if (D1) v = cadd(v, cmul(s1, D1[i + (size_t)j * n]));
    if (D2) v = cadd(v, cmul(s2, D2[i + (size_t)j * n]));
    if (D3) v = cadd(v, cmul(s3, cconj(D3[j + (size_t)i * n])));
    out[out_rowmajor ? (size_t)i * n + j : (size_t)i + (size_t)j * n] = v;
  }
  __syncthreads();
}

// Re sum_ij X(i,j) Y(j,i) = Re Tr(X Y).
__device__ __noinline__ double wg_retrace_prod(int n, const double2* X, const double2* Y, double* red) {
  double s = 0.0;
  const int nn = n * n;
  for (int idx = threadIdx.x; idx < nn; idx += NT) {
    const int i = idx % n, j = idx / n;
    double2 x = X[i + (size_t)j * n], y = Y[j + (size_t)i * n];
    s += x.x * y.x - x.y * y.y;
  }
  return block_sum(s, red);
}

// Im Tr(X).
__device__ __forceinline__ double wg_imtrace(int n, const double2* X, double* red) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) s += X[i + (size_t)i * n].y;
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
__device__ __noinline__ int wg_inverse(int n, double2* W, double2* out, double* lds) {
  constexpr int LDI = NMAX + PAD;
  constexpr int MT = 32 * T;
  double* Mr = lds;
  double* Mi = Mr + KP * LDI;
  double* Br = Mi + KP * LDI;
  double* Bi = Br + KP * LDI;
  int* piv = (int*)(Bi + KP * LDI);
  int* used = piv + NMAX;
  double* redv = (double*)(used + NMAX);
  int* redi = (int*)(redv + 4);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave & 1, wc = wave >> 1;
  for (int i = tid; i < n; i += NT) used[i] = 0;
  __syncthreads();
  int info = 0;
  for (int k0 = 0; k0 < n && !info; k0 += KP) {
    for (int idx = tid; idx < n * KP; idx += NT) {
      const int c = idx % KP, i = idx / KP;
      const double2 v = W[(size_t)i * n + k0 + c];
      Mr[c * LDI + i] = v.x;
      Mi[c * LDI + i] = v.y;
    }
    __syncthreads();
    for (int c = 0; c < KP; ++c) {
      double best = -1.0;
      int bi = 0x7fffffff;
      for (int i = tid; i < n; i += NT) {
        if (!used[i]) {
          const double v = fabs(Mr[c * LDI + i]) + fabs(Mi[c * LDI + i]);
          if (v > best) { best = v; bi = i; }
        }
      }
      block_argmax(best, bi, redv, redi);
      if (!(best > 0.0)) { info = k0 + c + 1; break; }
      const int p = bi;
      const double2 a = make_double2(Mr[c * LDI + p], Mi[c * LDI + p]);
      const double2 ainv = crecip(a);
      double2 rp[KP];
#pragma unroll
      for (int c2 = 0; c2 < KP; ++c2)
        rp[c2] = (c2 == c) ? ainv : cmul(make_double2(Mr[c2 * LDI + p], Mi[c2 * LDI + p]), ainv);
      __syncthreads();
      if (tid == 0) { used[p] = 1; piv[k0 + c] = p; }
      for (int i = tid; i < n; i += NT) {
        if (i == p) {
#pragma unroll
          for (int c2 = 0; c2 < KP; ++c2) { Mr[c2 * LDI + i] = rp[c2].x; Mi[c2 * LDI + i] = rp[c2].y; }
        } else {
          const double2 f = make_double2(Mr[c * LDI + i], Mi[c * LDI + i]);
#pragma unroll
          for (int c2 = 0; c2 < KP; ++c2) {
            double2 v = (c2 == c) ? make_double2(0, 0) : make_double2(Mr[c2 * LDI + i], Mi[c2 * LDI + i]);
            v = csub(v, cmul(f, rp[c2]));
            Mr[c2 * LDI + i] = v.x;
            Mi[c2 * LDI + i] = v.y;
          }
        }
      }
      __syncthreads();
    }
    if (info) break;
    // snapshot of the pivot rows (all columns) before they are updated
    for (int idx = tid; idx < KP * n; idx += NT) {
      const int c = idx / n, j = idx % n;
      const double2 v = W[(size_t)piv[k0 + c] * n + j];
      Br[c * LDI + j] = v.x;
      Bi[c * LDI + j] = v.y;
    }
    __syncthreads();
    if (tid < KP) Mr[tid * LDI + piv[k0 + tid]] -= 1.0;  // M = F - E
    __syncthreads();
    const int nt = n / MT;
    const int m = lane & 15, kq = lane >> 4;
    for (int tile = 0; tile < nt * nt; ++tile) {
      const int row0 = (tile % nt) * MT, col0 = (tile / nt) * MT;
      d4 cr[T][T], ci[T][T];
#pragma unroll
      for (int a = 0; a < T; ++a)
#pragma unroll
        for (int b = 0; b < T; ++b) { cr[a][b] = d4{0, 0, 0, 0}; ci[a][b] = d4{0, 0, 0, 0}; }
      mma_steps<T, 1>(cr, ci, Mr, Mi, LDI, row0 + wr * 16 * T, Br, Bi, LDI, col0 + wc * 16 * T, KP / 4);
#pragma unroll
      for (int ti = 0; ti < T; ++ti)
#pragma unroll
        for (int tj = 0; tj < T; ++tj)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = row0 + wr * 16 * T + ti * 16 + 4 * r + kq;
            const int j = col0 + wc * 16 * T + tj * 16 + m;
            double2* w = &W[(size_t)i * n + j];
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
    for (int idx = tid; idx < n * n; idx += NT) out[idx] = make_double2(0, 0);
    __syncthreads();
    return info;
  }
  // inverse permutation into `used`
  for (int q = tid; q < n; q += NT) used[piv[q]] = q;
  __syncthreads();
  for (int idx = tid; idx < n * n; idx += NT) {
    const int i = idx % n, q = idx / n;
    out[idx] = W[(size_t)piv[i] * n + used[q]];
  }
  __syncthreads();
  return 0;
}

}  // namespace rgf
