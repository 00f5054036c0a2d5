// rgf_blocks_host.hpp -- host (CPU) stand-in for liteqt_amd/csrc/rgf_blocks.hpp,
// test infrastructure.  Copied next to a copy of rgf_sweep.hpp as rgf_blocks.hpp,
// it lets g++ compile the sweep kernels' block-level program (solve_energy and the
// persistent loops) for one host "workgroup": every block operation runs as plain
// complex arithmetic, and every block pointer it receives is checked against the
// registered arrays (inputs, outputs, workspace slot) so that any block offset
// outside its array, or an operand aliasing that the device functions do not
// allow, is recorded.  Thread-level indexing inside the device functions is not
// modelled (it is shared by every kernel variant).
#pragma once
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __noinline__
#define __launch_bounds__(...)
#define __shared__ static

struct HostDim {
  unsigned x, y, z;
};
static HostDim threadIdx = {0, 0, 0};
static HostDim blockIdx = {0, 0, 0};
static HostDim gridDim = {1, 1, 1};
static HostDim blockDim = {1, 1, 1};  // one host "thread" covers every element of a loop
inline void __syncthreads() {}

struct double2 {
  double x, y;
};
inline double2 make_double2(double x, double y) { return double2{x, y}; }

namespace hostcheck {
struct Region {
  const void* base;
  size_t bytes;
  std::string name;
};
inline std::vector<Region>& regions() {
  static std::vector<Region> r;
  return r;
}
inline std::vector<std::string>& violations() {
  static std::vector<std::string> v;
  return v;
}
inline long long& ops() {
  static long long n = 0;
  return n;
}
inline void add(const void* p, size_t bytes, const char* name) {
  if (p) regions().push_back(Region{p, bytes, name});
}
inline void reset() {
  regions().clear();
  violations().clear();
  ops() = 0;
}
// [p, p + bytes) inside one registered array
inline const Region* find(const void* p, size_t bytes) {
  const char* c = (const char*)p;
  for (auto& r : regions()) {
    const char* b = (const char*)r.base;
    if (c >= b && c + bytes <= b + r.bytes) return &r;
  }
  return nullptr;
}
inline void need(const char* op, const char* arg, const void* p, size_t bytes) {
  if (!p) {
    violations().push_back(std::string(op) + ": " + arg + " is NULL");
    return;
  }
  if (!find(p, bytes)) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "%s: %s [%p, +%zu bytes) outside every registered array", op, arg, p, bytes);
    violations().push_back(buf);
  }
}
inline bool overlap(const void* a, size_t na, const void* b, size_t nb) {
  const char* x = (const char*)a;
  const char* y = (const char*)b;
  return a && b && x < y + nb && y < x + na;
}
inline void no_alias(const char* op, const char* what, const void* a, const void* b, size_t bytes) {
  if (overlap(a, bytes, b, bytes)) violations().push_back(std::string(op) + ": " + what);
}
}  // namespace hostcheck

namespace rgf {

typedef double ldsd;
typedef int ldsi;
typedef std::complex<double> cd;

enum { OP_N = 0, OP_C = 1, OP_T = 2 };

template <int T>
struct Cfg {
  static constexpr int NTH = 256;
};
template <int T>
struct GemmLds {
  static constexpr int DOUBLES = 16;
};
template <int T, int NMAX>
struct InvLds {
  static constexpr int DOUBLES = 16;
};

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
  int abs_sum;
};

inline Epi epi_store() {
  Epi e;
  e.diag = make_double2(0, 0); e.sacc = 1.0;
  e.D1 = nullptr; e.s1 = make_double2(1, 0);
  e.D2 = nullptr; e.s2 = make_double2(1, 0);
  e.D3 = nullptr; e.s3 = make_double2(1, 0);
  e.real_out = 0; e.abs_sum = 0;
  return e;
}

inline const double2* gc(const double2* p) { return p; }
inline double block_sum(double v, ldsd*) { return v; }

inline cd C(double2 v) { return cd(v.x, v.y); }
inline double2 D(cd v) { return make_double2(v.real(), v.imag()); }

template <int T>
inline double wg_gemm(int n, int opA, const double2* A, int opB, const double2* B, void* Cout, int c_rowmajor,
                      Epi ep, ldsd*) {
  using namespace hostcheck;
  ++ops();
  const size_t nn = (size_t)n * n, bytes = nn * 16;
  need("wg_gemm", "A", A, bytes);
  need("wg_gemm", "B", B, bytes);
  need("wg_gemm", "C", Cout, ep.real_out ? nn * 8 : bytes);
  if (ep.D1) need("wg_gemm", "D1", ep.D1, bytes);
  if (ep.D2) need("wg_gemm", "D2", ep.D2, bytes);
  if (ep.D3) need("wg_gemm", "D3", ep.D3, bytes);
  const size_t cb = ep.real_out ? nn * 8 : bytes;
  if (overlap(Cout, cb, A, bytes)) violations().push_back("wg_gemm: output aliases A");
  if (overlap(Cout, cb, B, bytes)) violations().push_back("wg_gemm: output aliases B");
  if (ep.D3 && overlap(Cout, cb, ep.D3, bytes)) violations().push_back("wg_gemm: output aliases D3");
  if (c_rowmajor && ep.D1 && overlap(Cout, cb, ep.D1, bytes)) violations().push_back("wg_gemm: row-major output aliases D1");
  if (c_rowmajor && ep.D2 && overlap(Cout, cb, ep.D2, bytes)) violations().push_back("wg_gemm: row-major output aliases D2");
  if (!violations().empty() && (!A || !B || !Cout)) return 0.0;
  std::vector<cd> a(nn), b(nn), c(nn, cd(0, 0));
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      a[i + (size_t)k * n] = opA == OP_N ? C(A[i + (size_t)k * n])
                                         : (opA == OP_T ? C(A[k + (size_t)i * n]) : std::conj(C(A[k + (size_t)i * n])));
      b[i + (size_t)k * n] = opB == OP_N ? C(B[i + (size_t)k * n])
                                         : (opB == OP_T ? C(B[k + (size_t)i * n]) : std::conj(C(B[k + (size_t)i * n])));
    }
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      const cd bkj = b[k + (size_t)j * n];
      for (int i = 0; i < n; ++i) c[i + (size_t)j * n] += a[i + (size_t)k * n] * bkj;
    }
  std::vector<cd> v(nn);
  double asum = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      cd x = ep.sacc * c[i + (size_t)j * n];
      if (i == j) x += C(ep.diag);
      if (ep.D1) x += C(ep.s1) * C(ep.D1[i + (size_t)j * n]);
      if (ep.D2) x += C(ep.s2) * C(ep.D2[i + (size_t)j * n]);
      if (ep.D3) x += C(ep.s3) * std::conj(C(ep.D3[j + (size_t)i * n]));
      v[i + (size_t)j * n] = x;
      asum += std::abs(x);
    }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const size_t u = c_rowmajor ? (size_t)i * n + j : i + (size_t)j * n;
      if (ep.real_out)
        ((double*)Cout)[u] = v[i + (size_t)j * n].real();
      else
        ((double2*)Cout)[u] = D(v[i + (size_t)j * n]);
    }
  return ep.abs_sum ? asum : 0.0;
}

template <int T, int OPA, int OPB, int CRM>
inline double wg_gemm_k(int n, const double2* A, const double2* B, void* Cout, Epi ep, ldsd* lds) {
  return wg_gemm<T>(n, OPA, A, OPB, B, Cout, CRM, ep, lds);
}

inline void wg_imag_part(int n, const double2* X, double2* Y, double2* Z) {
  using namespace hostcheck;
  ++ops();
  const size_t nn = (size_t)n * n, bytes = nn * 16;
  need("wg_imag_part", "X", X, bytes);
  need("wg_imag_part", "Y", Y, bytes);
  no_alias("wg_imag_part", "Y aliases X", X, Y, bytes);
  if (Z) {
    need("wg_imag_part", "Z", Z, bytes);
    no_alias("wg_imag_part", "Z aliases X", X, Z, bytes);
  }
  if (!X || !Y) return;
  for (size_t q = 0; q < nn; ++q) {
    Y[q] = make_double2(0.0, X[q].y);
    if (Z) Z[q] = X[q];
  }
}

inline void wg_ew(int n, double2* out, int out_rowmajor, double2 diag, const double2* D1, double2 s1,
                  const double2* D2, double2 s2, const double2* D3, double2 s3) {
  using namespace hostcheck;
  ++ops();
  const size_t nn = (size_t)n * n, bytes = nn * 16;
  need("wg_ew", "out", out, bytes);
  if (D1) need("wg_ew", "D1", D1, bytes);
  if (D2) need("wg_ew", "D2", D2, bytes);
  if (D3) need("wg_ew", "D3", D3, bytes);
  if (D3) no_alias("wg_ew", "output aliases D3", out, D3, bytes);
  if (out_rowmajor) {
    if (D1) no_alias("wg_ew", "row-major output aliases D1", out, D1, bytes);
    if (D2) no_alias("wg_ew", "row-major output aliases D2", out, D2, bytes);
  }
  if (!out) return;
  std::vector<cd> v(nn);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      cd x = (i == j) ? C(diag) : cd(0, 0);
      if (D1) x += C(s1) * C(D1[i + (size_t)j * n]);
      if (D2) x += C(s2) * C(D2[i + (size_t)j * n]);
      if (D3) x += C(s3) * std::conj(C(D3[j + (size_t)i * n]));
      v[i + (size_t)j * n] = x;
    }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) out[out_rowmajor ? (size_t)i * n + j : i + (size_t)j * n] = D(v[i + (size_t)j * n]);
}

constexpr int kAntihermLds = 4 * 32 * 33;

inline void wg_antiherm(int n, const double2* X, double2* out, ldsd*) {
  using namespace hostcheck;
  ++ops();
  const size_t bytes = (size_t)n * n * 16;
  need("wg_antiherm", "X", X, bytes);
  need("wg_antiherm", "out", out, bytes);
  no_alias("wg_antiherm", "output aliases X", out, X, bytes);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
      out[i + (size_t)j * n] = D(0.5 * (C(X[i + (size_t)j * n]) - std::conj(C(X[j + (size_t)i * n]))));
}

inline double wg_retrace_prod(int n, const double2* X, const double2* Y, ldsd*) {
  using namespace hostcheck;
  ++ops();
  need("wg_retrace_prod", "X", X, (size_t)n * n * 16);
  need("wg_retrace_prod", "Y", Y, (size_t)n * n * 16);
  double s = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) s += (C(X[i + (size_t)j * n]) * C(Y[j + (size_t)i * n])).real();
  return s;
}

inline double wg_imtrace(int n, const double2* X, ldsd*) {
  hostcheck::need("wg_imtrace", "X", X, (size_t)n * n * 16);
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += X[i + (size_t)i * n].y;
  return s;
}

// W row-major (destroyed), inverse written column-major to out; 0 or k+1 with
// out zeroed when column k has no nonzero pivot (partial row pivoting on |Re|+|Im|).
template <int T, int NMAX>
inline int wg_inverse(int n, double2* W, double2* out, ldsd*) {
  using namespace hostcheck;
  ++ops();
  const size_t nn = (size_t)n * n, bytes = nn * 16;
  need("wg_inverse", "W", W, bytes);
  need("wg_inverse", "out", out, bytes);
  no_alias("wg_inverse", "output aliases W", W, out, bytes);
  if (n > NMAX) violations().push_back("wg_inverse: n > NMAX");
  if (!W || !out) return 0;
  std::vector<cd> a(nn), inv(nn, cd(0, 0));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) a[(size_t)i * n + j] = C(W[(size_t)i * n + j]);  // row-major
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  int info = 0;
  for (int k = 0; k < n && !info; ++k) {
    int p = -1;
    double best = 0.0;
    for (int i = k; i < n; ++i) {
      const double v = std::fabs(a[(size_t)i * n + k].real()) + std::fabs(a[(size_t)i * n + k].imag());
      if (v > best) { best = v; p = i; }
    }
    if (p < 0) { info = k + 1; break; }
    if (p != k)
      for (int j = 0; j < n; ++j) {
        std::swap(a[(size_t)p * n + j], a[(size_t)k * n + j]);
        std::swap(inv[(size_t)p * n + j], inv[(size_t)k * n + j]);
      }
    const cd r = 1.0 / a[(size_t)k * n + k];
    for (int j = 0; j < n; ++j) { a[(size_t)k * n + j] *= r; inv[(size_t)k * n + j] *= r; }
    for (int i = 0; i < n; ++i) {
      if (i == k) continue;
      const cd f = a[(size_t)i * n + k];
      if (f == cd(0, 0)) continue;
      for (int j = 0; j < n; ++j) {
        a[(size_t)i * n + j] -= f * a[(size_t)k * n + j];
        inv[(size_t)i * n + j] -= f * inv[(size_t)k * n + j];
      }
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) out[i + (size_t)j * n] = info ? make_double2(0, 0) : D(inv[(size_t)i * n + j]);
  return info;
}

}  // namespace rgf
