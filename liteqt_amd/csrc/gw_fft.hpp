// Batched complex FP64 FFTs over series in device memory, for the FFT form of the GW
// energy correlations (the reference's corr1d2 / conv1d2 with method 'fft',
// src/fft_mod.f90:56-212, which transforms each zero-padded series with MKL DFTI,
// src/fft_mod.f90:214-234).
//
// Lengths are powers of two.  Up to 8192 points a workgroup holds whole series in LDS
// (4096 points = 64 KiB with two 256-thread workgroups per CU; 8192 points = 128 KiB with
// one 512-thread workgroup) and runs radix-4 Stockham passes (one radix-2 pass first for
// odd log2) in place: every thread loads its butterflies' inputs into registers, the
// workgroup synchronizes, then every thread writes its outputs.  Longer series use the
// four-step split N = N1 * N2 (n = n1 + N1 n2,
// k = k2 + N2 k1) with both factors <= 4096: length-N2 transforms down the columns
// (stride N1 in memory, C = 4096 / N2 columns per workgroup so that every load is a run
// of C consecutive points), the twiddle W_N^(n1 k2), then length-N1 transforms along the
// contiguous rows.  The spectrum is left in the transposed order pos(k) = k1 + N1 k2;
// the inverse undoes the steps in reverse order and returns the series in natural order,
// so callers only need pos() to pair frequency k with N - k.  No 1/N scaling.
//
// Twiddles of the LDS passes come from a table exp(-2 pi i q / 8192) written by
// k_fft_twiddles (sincospi of exact binary fractions); the four-step twiddle is
// computed with sincospi per point.
#pragma once
#include <hip/hip_runtime.h>

namespace gwfft {

constexpr int kLds = 4096;   // points per workgroup (series up to 4096 and the four-step)
constexpr int kNth = 256;    // threads per such workgroup
constexpr int kLog = 12;
constexpr int kLog1 = 13;    // longest series of one LDS pass: 8192 points, 512 threads

constexpr int kTw = 8192;    // twiddle table: exp(-2 pi i q / 8192)

__device__ double2 g_tw[kTw];

__global__ __launch_bounds__(256) void k_fft_twiddles() {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= kTw) return;
  double s, c;
  sincospi(-2.0 * q / kTw, &s, &c);
  g_tw[q] = make_double2(c, s);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 conj2(double2 a) { return make_double2(a.x, -a.y); }

// W_{4 Ns}^m = exp(-+ 2 pi i m / (4 Ns)) for 0 <= m < 3 Ns, 4 Ns <= 8192: INV selects the
// sign of the inverse transform.
template <bool INV>
__device__ __forceinline__ double2 twiddle(int m, int Ns) {
  const double2 w = g_tw[m * (kTw / 4 / Ns)];
  return INV ? conj2(w) : w;
}

// B series of length L = 2^logL, series b at buf[b * S ...] (S = L + 1: the pad keeps
// the column loads of the four-step free of bank conflicts), B * L <= PTS points on NTH
// threads.
// Stockham: pass with sub-length Ns and radix R reads x[j + r L / R] and writes
// y[(j - k) R + k + r Ns] (k = j mod Ns) after the twiddle W_{R Ns}^(k r).
template <bool INV, int NTH = kNth, int PTS = kLds>
__device__ void lds_fft(double2* buf, int logL, int B, int S) {
  const int L = 1 << logL, tid = threadIdx.x;
  int Ns = 1;
  if (logL & 1) {  // radix-2 pass at Ns = 1 (no twiddles)
    const int nb = B * (L >> 1);
    constexpr int U = PTS / 2 / NTH;
    double2 v0[U], v1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int g = tid + u * NTH;
      if (g < nb) {
        const int b = g >> (logL - 1), j = g & ((L >> 1) - 1);
        v0[u] = buf[b * S + j];
        v1[u] = buf[b * S + j + (L >> 1)];
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int g = tid + u * NTH;
      if (g < nb) {
        const int b = g >> (logL - 1), j = g & ((L >> 1) - 1);
        buf[b * S + 2 * j] = cadd(v0[u], v1[u]);
        buf[b * S + 2 * j + 1] = csub(v0[u], v1[u]);
      }
    }
    __syncthreads();
    Ns = 2;
  }
  const int nb = B * (L >> 2);
  for (; Ns < L; Ns <<= 2) {
    constexpr int U = PTS / 4 / NTH;
    double2 v[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int g = tid + u * NTH;
      if (g < nb) {
        const int b = g >> (logL - 2), j = g & ((L >> 2) - 1);
#pragma unroll
        for (int r = 0; r < 4; ++r) v[u][r] = buf[b * S + j + r * (L >> 2)];
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int g = tid + u * NTH;
      if (g < nb) {
        const int b = g >> (logL - 2), j = g & ((L >> 2) - 1);
        const int k = j & (Ns - 1);
        double2 x0 = v[u][0];
        double2 x1 = v[u][1], x2 = v[u][2], x3 = v[u][3];
        if (Ns > 1) {
          x1 = cmul(x1, twiddle<INV>(k, Ns));
          x2 = cmul(x2, twiddle<INV>(2 * k, Ns));
          x3 = cmul(x3, twiddle<INV>(3 * k, Ns));
        }
        // DFT-4 with W_4 = -i (forward) or +i (inverse)
        const double2 s02 = cadd(x0, x2), d02 = csub(x0, x2);
        const double2 s13 = cadd(x1, x3), d13 = csub(x1, x3);
        // -i * d13 (forward) or +i * d13 (inverse)
        const double2 rd13 = INV ? make_double2(-d13.y, d13.x) : make_double2(d13.y, -d13.x);
        double2* out = buf + b * S + (j - k) * 4 + k;
        out[0] = cadd(s02, s13);
        out[Ns] = cadd(d02, rd13);
        out[2 * Ns] = csub(s02, s13);
        out[3 * Ns] = csub(d02, rd13);
      }
    }
    __syncthreads();
  }
}

// Series enumeration shared by the batched kernels: series t of a launch is
// base + (first + (t / kc) * K + t % kc) * N (slot t / kc, kinds k0 .. k0 + kc - 1 of
// K per slot; first = k0).  valid(slot) lets padding slots be skipped.
struct FftArgs {
  double2* base;
  long long nslots;  // slots in the chunk
  int K, k0, kc;     // kinds per slot, first kind, kinds transformed
  int logN, logN1, logN2;
  // slot validity (band geometry): slot = (slab * nm + j) * R + r, row i = lo(j) + r
  int nm, nd, R;
};

__device__ __forceinline__ bool slot_valid(const FftArgs& a, long long slot) {
  const int r = (int)(slot % a.R);
  const int j = (int)((slot / a.R) % a.nm);
  const int lo = max(0, j - a.nd), hi = min(a.nm - 1, j + a.nd);
  return lo + r <= hi;
}

__device__ __forceinline__ double2* series_ptr(const FftArgs& a, long long t) {
  const long long slot = t / a.kc;
  const int kind = a.k0 + (int)(t % a.kc);
  return a.base + ((size_t)slot * a.K + kind) * ((size_t)1 << a.logN);
}

// N <= PTS: B = PTS / N series per workgroup, blockIdx.x = group of B series.
// <256, 4096> for N <= 4096 (two workgroups per CU), <512, 8192> for N = 8192.
template <bool INV, int NTH, int PTS>
__global__ __launch_bounds__(NTH, PTS > kLds ? 1 : 2) void k_fft_single(FftArgs a) {
  __shared__ double2 buf[PTS + PTS / 64 + 64];
  __shared__ double2* ser[PTS / 64];  // the workgroup's series (null: padding slot or past the end)
  const int logN = a.logN, N = 1 << logN, B = PTS >> logN, S = N + 1;
  const long long nser = a.nslots * a.kc;
  const long long t0 = (long long)blockIdx.x * B;
  if (threadIdx.x < B) {
    const long long t = t0 + threadIdx.x;
    ser[threadIdx.x] = (t < nser && slot_valid(a, t / a.kc)) ? series_ptr(a, t) : nullptr;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < B * N; q += NTH) {
    const int b = q >> logN, n = q & (N - 1);
    const double2* x = ser[b];
    buf[b * S + n] = x ? x[n] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  lds_fft<INV, NTH, PTS>(buf, logN, B, S);
  for (int q = threadIdx.x; q < B * N; q += NTH) {
    const int b = q >> logN, n = q & (N - 1);
    double2* x = ser[b];
    if (x) x[n] = buf[b * S + n];
  }
}

// Four-step column pass: C = 4096 / N2 columns n1 of one series per workgroup
// (blockIdx.x = series * (N1 / C) + column group).  TW: multiply output (n1, k2) by
// W_N^(+-n1 k2) (forward pass, after the column transforms).
template <bool INV, bool TW>
__global__ __launch_bounds__(kNth, 2) void k_fft_cols(FftArgs a) {
  __shared__ double2 buf[kLds + kLds / 64 + 64];
  const int N1 = 1 << a.logN1, logN2 = a.logN2, N2 = 1 << logN2;
  const int C = kLds >> logN2, S = N2 + 1;
  const int groups = N1 / C;
  const long long t = blockIdx.x / groups;
  const int c0 = (int)(blockIdx.x % groups) * C;
  if (!slot_valid(a, t / a.kc)) return;  // uniform over the workgroup
  double2* x = series_ptr(a, t);
  for (int q = threadIdx.x; q < C * N2; q += kNth) {
    const int c = q % C, n2 = q / C;
    buf[c * S + n2] = x[(size_t)(c0 + c) + (size_t)N1 * n2];
  }
  __syncthreads();
  lds_fft<INV>(buf, logN2, C, S);
  const double inv_n = 2.0 / (double)((size_t)1 << a.logN);
  for (int q = threadIdx.x; q < C * N2; q += kNth) {
    const int c = q % C, k2 = q / C;
    double2 v = buf[c * S + k2];
    if (TW) {
      const size_t e = ((size_t)(c0 + c) * (size_t)k2) & (((size_t)1 << a.logN) - 1);
      double s, co;
      sincospi((INV ? 1.0 : -1.0) * inv_n * (double)e, &s, &co);
      v = cmul(v, make_double2(co, s));
    }
    x[(size_t)(c0 + c) + (size_t)N1 * k2] = v;
  }
}

// Four-step row pass: B = 4096 / N1 rows k2 of one series per workgroup, contiguous.
// TW: multiply output (n1, k2) by W_N^(+-n1 k2) (inverse pass, after the row transforms).
template <bool INV, bool TW>
__global__ __launch_bounds__(kNth, 2) void k_fft_rows(FftArgs a) {
  __shared__ double2 buf[kLds + kLds / 64 + 64];
  const int logN1 = a.logN1, N1 = 1 << logN1, N2 = 1 << a.logN2;
  const int B = kLds >> logN1, S = N1 + 1;
  const int groups = N2 / B;
  const long long t = blockIdx.x / groups;
  const int r0 = (int)(blockIdx.x % groups) * B;
  if (!slot_valid(a, t / a.kc)) return;
  double2* x = series_ptr(a, t) + (size_t)r0 * N1;
  for (int q = threadIdx.x; q < B * N1; q += kNth) buf[(q >> logN1) * S + (q & (N1 - 1))] = x[q];
  __syncthreads();
  lds_fft<INV>(buf, logN1, B, S);
  const double inv_n = 2.0 / (double)((size_t)1 << a.logN);
  for (int q = threadIdx.x; q < B * N1; q += kNth) {
    const int b = q >> logN1, n1 = q & (N1 - 1);
    double2 v = buf[b * S + n1];
    if (TW) {
      const size_t e = ((size_t)n1 * (size_t)(r0 + b)) & (((size_t)1 << a.logN) - 1);
      double s, co;
      sincospi((INV ? 1.0 : -1.0) * inv_n * (double)e, &s, &co);
      v = cmul(v, make_double2(co, s));
    }
    x[q] = v;
  }
}

// (N1, N2) of the four-step split (N > 8192): N2 >= 64 columns keep the column loads in
// runs of >= 4096 / N2 points; both factors <= 4096 (N <= 2^24).
__host__ __device__ inline void split(int logN, int& logN1, int& logN2) {
  if (logN <= kLog1) {
    logN1 = logN;
    logN2 = 0;
    return;
  }
  logN2 = logN - kLog > 6 ? logN - kLog : 6;
  logN1 = logN - logN2;
}

// Storage position of frequency k after a forward transform of length 2^logN.
__host__ __device__ inline long long pos(long long k, int logN, int logN1, int logN2) {
  if (logN <= kLog1) return k;
  const long long k2 = k & ((1LL << logN2) - 1), k1 = k >> logN2;
  return k1 + (k2 << logN1);
}

// Inverse of pos(): the frequency held at storage position p.
__host__ __device__ inline long long freq(long long p, int logN, int logN1, int logN2) {
  if (logN <= kLog1) return p;
  const long long k1 = p & ((1LL << logN1) - 1), k2 = p >> logN1;
  return k2 + (k1 << logN2);
}

}  // namespace gwfft
