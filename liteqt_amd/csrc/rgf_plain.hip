// rgf_plain.hip -- the plain per-energy GPU form of green_RGF_RS, the
// comparison baseline of the batched engine.
//
// One energy at a time, one launch chain per energy: every product of the
// reference (src/green_rgf.f90:3992-4160, the same zgemm operand order) is a
// rocBLAS zgemm call, every inverse rocSOLVER zgetrf + zgetri (the reference's
// invert, 4421-4441, with A <- 0 on failure), and the decimation
// (4451-4537) checks its convergence on the host after every iteration, as a
// direct port of the CPU code to vendor GPU libraries does.  Device arrays,
// the same rgf_batch contract as rgf_solve_batched.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rgf.h"
#include "../../include/rgf_plain.h"

namespace {

thread_local std::string g_err;
std::mutex g_mu;

int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}

#define HIPCHK(x)                                                                                     \
  do {                                                                                                \
    hipError_t _e = (x);                                                                              \
    if (_e != hipSuccess) return fail(RGF_EHIP, std::string(#x) + ": " + hipGetErrorString(_e));     \
  } while (0)
#define RBCHK(x)                                                                                      \
  do {                                                                                                \
    rocblas_status _s = (x);                                                                          \
    if (_s != rocblas_status_success) return fail(RGF_EHIP, std::string(#x) + ": rocblas status " + std::to_string((int)_s)); \
  } while (0)

typedef rocblas_double_complex zc;

// out = diag*I + s1*D1 + s2*D2 + s3*D3^H (D1/D2 may alias out, D3 must not)
__global__ void k_ew(int n, double2* out, double2 diag, const double2* D1, double2 s1, const double2* D2,
                     double2 s2, const double2* D3, double2 s3) {
  const int nn = n * n;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += gridDim.x * blockDim.x) {
    const int i = idx % n, j = idx / n;
    double2 v = (i == j) ? diag : make_double2(0, 0);
    if (D1) { double2 d = D1[idx]; v.x += s1.x * d.x - s1.y * d.y; v.y += s1.x * d.y + s1.y * d.x; }
    if (D2) { double2 d = D2[idx]; v.x += s2.x * d.x - s2.y * d.y; v.y += s2.x * d.y + s2.y * d.x; }
    if (D3) {
      double2 d = D3[j + (size_t)i * n];
      d.y = -d.y;
      v.x += s3.x * d.x - s3.y * d.y;
      v.y += s3.x * d.y + s3.y * d.x;
    }
    out[idx] = v;
  }
}

// out[slot] = sum_i Re(A_ii)  (one block)
__global__ void k_retrace(int n, const double2* A, double* out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += A[i + (size_t)i * n].x;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// *out = sum_ij |A_ij|
__global__ void k_abssum(int n, const double2* A, double* out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n * n; i += blockDim.x) s += sqrt(A[i].y * A[i].y + A[i].x * A[i].x);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

__global__ void k_zero_if(int n, double2* A, const int* info, int* fails) {
  if (*info == 0) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n * n; i += gridDim.x * blockDim.x) A[i] = make_double2(0, 0);
  if (blockIdx.x == 0 && threadIdx.x == 0 && fails) atomicAdd(fails, 1);
}

// real part copy: out_real(i,j) (+)= sign * Re(A(i,j))
__global__ void k_re(int n, double* out, const double2* A, double sign, int accumulate) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n * n; i += gridDim.x * blockDim.x)
    out[i] = (accumulate ? out[i] : 0.0) + sign * A[i].x;
}

__global__ void k_imtrace(int n, const double2* A, double sign, double2 extra_scale, const double2* B, double* out) {
  // out = sign * (Im Tr A + extra_scale.x * Im Tr B)
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    s += A[i + (size_t)i * n].y;
    if (B) s += extra_scale.x * B[i + (size_t)i * n].y;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sign * red[0];
}

__global__ void k_tr_combine(const double* b, const double* c, double* out, double sign) {
  *out = sign * (*b - *c);
}

__global__ void k_set_info(int* dst, int a, int b, int c, const int* fails) {
  dst[0] = a;
  dst[1] = b;
  dst[2] = c;
  dst[3] = *fails;
}

struct Ctx {
  rocblas_handle h = nullptr;
  hipStream_t s = nullptr;
  int n = 0;
  int grid = 0;
  void* ws = nullptr;
  size_t ws_bytes = 0;
};
Ctx g_ctx;

struct Plain {
  Ctx& c;
  int n;
  size_t nn;
  double2 *ipiv_info;  // unused placeholder
  rocblas_int *ipiv, *info;
  int* fails;
  double* scal;  // device scalars
  const zc one = {1.0, 0.0}, zero = {0.0, 0.0};

  int gemm(char ta, char tb, const double2* A, const double2* B, double2* C) {
    auto op = [](char t) { return (t == 'c') ? rocblas_operation_conjugate_transpose : rocblas_operation_none; };
    RBCHK(rocblas_zgemm(c.h, op(ta), op(tb), n, n, n, &one, (const zc*)A, n, (const zc*)B, n, &zero, (zc*)C, n));
    return 0;
  }
  int ew(double2* out, double2 diag, const double2* D1, double2 s1, const double2* D2, double2 s2,
         const double2* D3, double2 s3) {
    hipLaunchKernelGGL(k_ew, dim3(c.grid), dim3(256), 0, c.s, n, out, diag, D1, s1, D2, s2, D3, s3);
    HIPCHK(hipGetLastError());
    return 0;
  }
  int copy(double2* d, const double2* s) { return ew(d, make_double2(0, 0), s, make_double2(1, 0), nullptr, {}, nullptr, {}); }
  int inv(double2* A) {
    RBCHK(rocsolver_zgetrf(c.h, n, n, (zc*)A, n, ipiv, info));
    hipLaunchKernelGGL(k_zero_if, dim3(c.grid), dim3(256), 0, c.s, n, A, info, fails);
    RBCHK(rocsolver_zgetri(c.h, n, (zc*)A, n, ipiv, info));
    // a zero pivot is reported by zgetrf already; count each failure once
    hipLaunchKernelGGL(k_zero_if, dim3(c.grid), dim3(256), 0, c.s, n, A, info, (int*)nullptr);
    HIPCHK(hipGetLastError());
    return 0;
  }
};

}  // namespace

extern "C" {

const char* rgf_plain_last_error(void) { return g_err.c_str(); }

int rgf_plain_solve(const rgf_params* p, const rgf_batch* b, void* stream) {
  g_err.clear();
  if (!p || !b || !b->E || !b->Hii || !b->H1i) return fail(RGF_EINVAL, "missing arguments");
  if (b->nx < 2 || b->nm < 1) return fail(RGF_EINVAL, "need nm >= 1, nx >= 2");
  if (b->sigR_lo || b->sigR_up || b->sigL_up)
    return fail(RGF_EUNSUP, "the plain path takes diagonal Sigma blocks only");
  std::lock_guard<std::mutex> lk(g_mu);
  Ctx& c = g_ctx;
  c.s = reinterpret_cast<hipStream_t>(stream);
  if (!c.h) RBCHK(rocblas_create_handle(&c.h));
  RBCHK(rocblas_set_stream(c.h, c.s));
  RBCHK(rocblas_set_pointer_mode(c.h, rocblas_pointer_mode_host));
  const int n = b->nm, nx = b->nx;
  const size_t nn = (size_t)n * n;
  c.n = n;
  c.grid = (int)std::min<size_t>(1024, (nn + 255) / 256);
  const int ntmp = 24;
  const size_t need = (2 * (size_t)nx + ntmp) * nn * 16 + 4096 + (size_t)n * 4;
  if (need > c.ws_bytes) {
    if (c.ws) (void)hipFree(c.ws);
    c.ws = nullptr;
    c.ws_bytes = 0;
    HIPCHK(hipMalloc(&c.ws, need));
    c.ws_bytes = need;
  }
  double2* Gl = (double2*)c.ws;
  double2* Gln = Gl + (size_t)nx * nn;
  double2* tb = Gln + (size_t)nx * nn;
  auto T = [&](int k) { return tb + (size_t)k * nn; };
  char* tail = (char*)(tb + (size_t)ntmp * nn);
  Plain P{c, n, nn, nullptr, (rocblas_int*)(tail + 1024), (rocblas_int*)tail, (int*)(tail + 64), (double*)(tail + 128)};
  double2 *H00 = T(0), *H10 = T(1), *A = T(2), *B = T(3), *C = T(4), *D = T(5), *G00 = T(6), *GN0 = T(7);
  double2 *Gn = T(8), *Gp = T(9), *sig = T(10), *sigmal = T(11), *sigmar = T(12), *H10h = T(13), *T1 = T(14);
  double2 *HBB = T(15), *HSS = T(16), *H01 = T(17), *Hx10 = T(18), *SA = T(19), *SB = T(20), *SC = T(21);
  const double2 c1 = make_double2(1, 0), cm1 = make_double2(-1, 0), c0 = make_double2(0, 0);
  const bool phys = (p->mode == RGF_MODE_PHYSICAL);
  const double2* Hii = (const double2*)b->Hii;
  const double2* H1i = (const double2*)b->H1i;
  std::vector<double> Eh(b->nE);
  HIPCHK(hipMemcpyAsync(Eh.data(), b->E, 8 * (size_t)b->nE, hipMemcpyDeviceToHost, c.s));
  HIPCHK(hipStreamSynchronize(c.s));
  auto ferm = [](double a) { return 1.0 / (1.0 + std::exp(a)); };

  // decimation with host-side convergence test (src/green_rgf.f90:4451-4537)
  auto sancho = [&](double E, const double2* h00, const double2* h10, double2* g00, int* iters, int* hit) -> int {
    const double2 zl = make_double2(E, p->eta_lead);
    if (int rc = P.copy(HBB, h00)) return rc;
    if (int rc = P.copy(HSS, h00)) return rc;
    if (int rc = P.copy(Hx10, h10)) return rc;
    if (int rc = P.ew(H01, c0, nullptr, {}, nullptr, {}, h10, c1)) return rc;
    *hit = 0;
    for (int it = 1; it <= p->dec_nmax; ++it) {
      *iters = it;
      if (int rc = P.ew(SA, zl, HBB, cm1, nullptr, {}, nullptr, {})) return rc;
      if (int rc = P.inv(SA)) return rc;
      if (int rc = P.gemm('n', 'n', SA, Hx10, SB)) return rc;
      if (int rc = P.gemm('n', 'n', H01, SB, SC)) return rc;
      if (int rc = P.ew(HSS, c0, HSS, c1, SC, c1, nullptr, {})) return rc;
      if (int rc = P.ew(HBB, c0, HBB, c1, SC, c1, nullptr, {})) return rc;
      if (int rc = P.gemm('n', 'n', Hx10, SB, SC)) return rc;      // new H_10
      if (int rc = P.gemm('n', 'n', SA, H01, SB)) return rc;
      if (int rc = P.gemm('n', 'n', Hx10, SB, SA)) return rc;
      if (int rc = P.copy(Hx10, SC)) return rc;
      if (int rc = P.ew(HBB, c0, HBB, c1, SA, c1, nullptr, {})) return rc;
      if (int rc = P.gemm('n', 'n', H01, SB, SC)) return rc;
      if (int rc = P.copy(H01, SC)) return rc;
      hipLaunchKernelGGL(k_abssum, dim3(1), dim3(256), 0, c.s, n, SC, P.scal);
      double err = 0.0;
      HIPCHK(hipMemcpyAsync(&err, P.scal, 8, hipMemcpyDeviceToHost, c.s));
      HIPCHK(hipStreamSynchronize(c.s));
      if (std::fabs(err) < p->dec_tol) break;
      if (it == p->dec_nmax) {
        *hit = 1;
        if (int rc = P.copy(HSS, h00)) return rc;
        if (int rc = P.copy(HBB, h00)) return rc;
      }
    }
    if (int rc = P.ew(g00, zl, HSS, cm1, nullptr, {}, nullptr, {})) return rc;
    return P.inv(g00);
  };

  for (int e = 0; e < b->nE; ++e) {
    const double E = Eh[e];
    const double2 z = make_double2(E, p->eta_dev);
    const double2* sLb = b->sigL ? (const double2*)b->sigL + (size_t)e * b->sig_stride : nullptr;
    const double2* sRb = b->sigR ? (const double2*)b->sigR + (size_t)e * b->sig_stride : nullptr;
    auto sL = [&](int l) -> const double2* { return sLb ? sLb + (size_t)l * nn : nullptr; };
    auto sR = [&](int l) -> const double2* { return sRb ? sRb + (size_t)l * nn : nullptr; };
    auto ob = [&](double* arr, int l) -> double2* { return arr ? (double2*)arr + ((size_t)e * nx + l) * nn : nullptr; };
    HIPCHK(hipMemsetAsync(P.fails, 0, 4, c.s));
    int itL = 0, itR = 0, hitL = 0, hitR = 0;
    // left contact
    if (int rc = P.ew(H00, c0, Hii, c1, sR(0), c1, nullptr, {})) return rc;
    if (int rc = P.copy(H10, H1i)) return rc;
    if (int rc = P.ew(H10h, c0, nullptr, {}, nullptr, {}, H10, c1)) return rc;
    if (int rc = sancho(E, H00, H10h, G00, &itL, &hitL)) return rc;
    if (int rc = P.gemm('n', 'n', H10, G00, A)) return rc;
    if (int rc = P.gemm('n', 'c', A, H10, sigmal)) return rc;
    const double fL = ferm((E - b->mu[0]) / (p->boltz * b->temp[0]));
    if (int rc = P.ew(sig, c0, sigmal, make_double2(-fL, 0), sL(0), c1, sigmal, make_double2(fL, 0))) return rc;
    if (int rc = P.ew(A, z, H00, cm1, sigmal, cm1, nullptr, {})) return rc;
    if (int rc = P.inv(A)) return rc;
    if (int rc = P.copy(Gl, A)) return rc;
    if (int rc = P.gemm('n', 'n', A, sig, B)) return rc;
    if (int rc = P.gemm('n', 'c', B, A, Gln)) return rc;
    // forward
    for (int l = 1; l < nx - 1; ++l) {
      const double2* H0 = Hii + (size_t)l * nn;
      const double2* Hc = H1i + (size_t)l * nn;
      double2* gl = Gl + (size_t)l * nn;
      double2* gln = Gln + (size_t)l * nn;
      if (int rc = P.gemm('n', 'n', Hc, gl - nn, B)) return rc;
      if (int rc = P.gemm('n', 'c', B, Hc, C)) return rc;
      if (int rc = P.ew(A, z, H0, cm1, sR(l), cm1, nullptr, {})) return rc;
      if (int rc = P.ew(A, c0, A, c1, C, cm1, nullptr, {})) return rc;
      if (int rc = P.inv(A)) return rc;
      if (int rc = P.copy(gl, A)) return rc;
      if (int rc = P.gemm('n', 'n', Hc, gln - nn, B)) return rc;
      if (int rc = P.gemm('n', 'c', B, Hc, C)) return rc;
      if (sL(l)) if (int rc = P.ew(C, c0, C, c1, sL(l), c1, nullptr, {})) return rc;
      if (int rc = P.gemm('n', 'n', A, C, B)) return rc;
      if (int rc = P.gemm('n', 'c', B, A, gln)) return rc;
    }
    // right contact
    const int L1 = nx - 1;
    const double2* HcR = H1i + (size_t)L1 * nn;
    if (int rc = P.ew(H00, c0, Hii + (size_t)L1 * nn, c1, sR(L1), c1, nullptr, {})) return rc;
    if (int rc = sancho(E, H00, HcR, G00, &itR, &hitR)) return rc;
    if (int rc = P.gemm('c', 'n', HcR, G00, A)) return rc;
    if (int rc = P.gemm('n', 'n', A, HcR, sigmar)) return rc;
    if (int rc = P.gemm('n', 'n', HcR, Gl + (size_t)(L1 - 1) * nn, B)) return rc;
    if (int rc = P.gemm('n', 'c', B, HcR, C)) return rc;
    if (int rc = P.ew(G00, z, H00, cm1, sigmar, cm1, nullptr, {})) return rc;
    if (int rc = P.ew(G00, c0, G00, c1, C, cm1, nullptr, {})) return rc;
    if (int rc = P.inv(G00)) return rc;
    if (b->ldos) if (int rc = P.copy(ob(b->ldos, L1), G00)) return rc;
    if (int rc = P.gemm('n', 'n', HcR, Gln + (size_t)(L1 - 1) * nn, B)) return rc;
    if (int rc = P.gemm('n', 'c', B, HcR, C)) return rc;
    const double fR = ferm((E - b->mu[1]) / (p->boltz * (phys ? b->temp[1] : b->temp[0])));
    if (int rc = P.ew(sig, c0, sigmar, make_double2(-fR, 0), C, c1, sigmar, make_double2(fR, 0))) return rc;
    if (sL(L1)) if (int rc = P.ew(sig, c0, sig, c1, sL(L1), c1, nullptr, {})) return rc;
    if (int rc = P.gemm('n', 'n', G00, sig, B)) return rc;
    if (int rc = P.gemm('n', 'c', B, G00, Gn)) return rc;
    if (b->ndens) if (int rc = P.copy(ob(b->ndens, L1), Gn)) return rc;
    if (int rc = P.ew(Gp, c0, Gn, c1, G00, c1, G00, cm1)) return rc;
    if (b->pdens) if (int rc = P.copy(ob(b->pdens, L1), Gp)) return rc;
    if (int rc = P.ew(A, c0, sigmar, make_double2(-fR, 0), nullptr, {}, sigmar, make_double2(fR, 0))) return rc;
    if (int rc = P.gemm('n', 'n', A, Gp, B)) return rc;
    if (int rc = P.ew(A, c0, sigmar, make_double2(-(fR - 1.0), 0), nullptr, {}, sigmar, make_double2(fR - 1.0, 0))) return rc;
    if (int rc = P.gemm('n', 'n', A, Gn, C)) return rc;
    hipLaunchKernelGGL(k_retrace, dim3(1), dim3(256), 0, c.s, n, B, P.scal + 1);
    hipLaunchKernelGGL(k_retrace, dim3(1), dim3(256), 0, c.s, n, C, P.scal + 2);
    if (b->tr) hipLaunchKernelGGL(k_tr_combine, dim3(1), dim3(1), 0, c.s, P.scal + 1, P.scal + 2, b->tr + e, -1.0);
    if (b->dens_n) hipLaunchKernelGGL(k_imtrace, dim3(1), dim3(256), 0, c.s, n, Gn, 1.0, c0, nullptr, b->dens_n + (size_t)e * nx + L1);
    if (b->dens_p) hipLaunchKernelGGL(k_imtrace, dim3(1), dim3(256), 0, c.s, n, Gp, -1.0, c0, nullptr, b->dens_p + (size_t)e * nx + L1);
    // backward (the reference's 22+ products per slab)
    for (int l = nx - 2; l >= 0; --l) {
      const double2* Hc = H1i + (size_t)(phys ? l + 1 : l) * nn;
      const double2* gl = Gl + (size_t)l * nn;
      const double2* gln = Gln + (size_t)l * nn;
      double* curl = b->cur ? b->cur + ((size_t)e * nx + l) * nn : nullptr;
      if (int rc = P.gemm('n', 'c', Hc, gl, B)) return rc;
      if (int rc = P.gemm('n', 'n', Gn, B, C)) return rc;
      if (int rc = P.gemm('n', 'n', Hc, gln, B)) return rc;
      if (int rc = P.gemm('n', 'n', G00, B, A)) return rc;
      if (int rc = P.ew(B, c0, C, c1, A, c1, nullptr, {})) return rc;
      if (int rc = P.gemm('c', 'n', Hc, B, A)) return rc;
      if (curl) hipLaunchKernelGGL(k_re, dim3(c.grid), dim3(256), 0, c.s, n, curl, A, 1.0, 0);
      if (int rc = P.gemm('n', 'c', gl, Hc, B)) return rc;
      if (int rc = P.gemm('n', 'n', B, Gn, C)) return rc;
      if (int rc = P.gemm('n', 'c', gln, Hc, B)) return rc;
      if (int rc = P.gemm('n', 'c', B, G00, A)) return rc;
      if (int rc = P.ew(B, c0, C, c1, A, c1, nullptr, {})) return rc;
      if (b->GLi1) if (int rc = P.copy(ob(b->GLi1, l), B)) return rc;
      if (b->GGi1 && phys) if (int rc = P.copy(T1, B)) return rc;
      if (int rc = P.gemm('n', 'n', Hc, B, A)) return rc;
      if (curl) hipLaunchKernelGGL(k_re, dim3(c.grid), dim3(256), 0, c.s, n, curl, A, -1.0, 1);
      if (b->GGi1 && !phys) {
        if (int rc = P.gemm('c', 'c', gl, Hc, B)) return rc;
        if (int rc = P.gemm('n', 'n', B, Gp, ob(b->GGi1, l))) return rc;
      }
      if (b->GGi1 && phys) {
        if (int rc = P.gemm('n', 'n', G00, Hc, A)) return rc;
        if (int rc = P.gemm('n', 'n', A, gl, C)) return rc;
      }
      if (int rc = P.copy(D, gl)) return rc;
      if (int rc = P.gemm('n', 'c', D, Hc, B)) return rc;
      if (int rc = P.gemm('n', 'n', B, G00, GN0)) return rc;
      if (b->GRi1) if (int rc = P.copy(ob(b->GRi1, l), GN0)) return rc;
      if (b->GGi1 && phys) if (int rc = P.ew(ob(b->GGi1, l), c0, T1, c1, GN0, c1, C, cm1)) return rc;
      if (int rc = P.gemm('n', 'n', GN0, Hc, A)) return rc;
      if (int rc = P.gemm('n', 'n', A, D, C)) return rc;
      if (int rc = P.ew(G00, c0, gl, c1, C, c1, nullptr, {})) return rc;
      if (b->ldos) if (int rc = P.copy(ob(b->ldos, l), G00)) return rc;
      if (int rc = P.gemm('n', 'c', D, Hc, B)) return rc;
      if (int rc = P.gemm('n', 'n', B, Gn, C)) return rc;
      if (int rc = P.gemm('n', 'n', C, Hc, A)) return rc;
      if (int rc = P.gemm('n', 'c', A, D, C)) return rc;
      if (int rc = P.ew(Gn, c0, gln, c1, C, c1, nullptr, {})) return rc;
      if (int rc = P.gemm('n', 'n', GN0, Hc, B)) return rc;
      if (int rc = P.gemm('n', 'n', B, gln, C)) return rc;
      if (int rc = P.ew(Gn, c0, Gn, c1, C, c1, nullptr, {})) return rc;
      if (int rc = P.gemm('n', 'c', gln, Hc, B)) return rc;
      if (int rc = P.gemm('n', 'c', B, GN0, C)) return rc;
      if (int rc = P.ew(Gn, c0, Gn, c1, C, c1, nullptr, {})) return rc;
      if (b->ndens) if (int rc = P.copy(ob(b->ndens, l), Gn)) return rc;
      if (b->pdens) if (int rc = P.ew(ob(b->pdens, l), c0, Gn, c1, G00, c1, G00, cm1)) return rc;
      if (b->dens_n) hipLaunchKernelGGL(k_imtrace, dim3(1), dim3(256), 0, c.s, n, Gn, 1.0, c0, nullptr, b->dens_n + (size_t)e * nx + l);
      if (b->dens_p) hipLaunchKernelGGL(k_imtrace, dim3(1), dim3(256), 0, c.s, n, Gn, -1.0, make_double2(2.0, 0), G00, b->dens_p + (size_t)e * nx + l);
    }
    // tre
    if (int rc = P.ew(Gp, c0, Gn, c1, G00, c1, G00, cm1)) return rc;
    const double fLt = ferm((E - b->mu[0]) / (p->boltz * (phys ? b->temp[0] : b->temp[1])));
    if (int rc = P.ew(A, c0, sigmal, make_double2(-fLt, 0), nullptr, {}, sigmal, make_double2(fLt, 0))) return rc;
    if (int rc = P.gemm('n', 'n', A, Gp, B)) return rc;
    if (int rc = P.ew(A, c0, sigmal, make_double2(-(fLt - 1.0), 0), nullptr, {}, sigmal, make_double2(fLt - 1.0, 0))) return rc;
    if (int rc = P.gemm('n', 'n', A, Gn, C)) return rc;
    hipLaunchKernelGGL(k_retrace, dim3(1), dim3(256), 0, c.s, n, B, P.scal + 1);
    hipLaunchKernelGGL(k_retrace, dim3(1), dim3(256), 0, c.s, n, C, P.scal + 2);
    if (b->tre) hipLaunchKernelGGL(k_tr_combine, dim3(1), dim3(1), 0, c.s, P.scal + 1, P.scal + 2, b->tre + e, 1.0);
    if (b->info)
      hipLaunchKernelGGL(k_set_info, dim3(1), dim3(1), 0, c.s, b->info + 4 * (size_t)e, itL, itR, hitL | (hitR << 1),
                         P.fails);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(c.s));
  return 0;
}

}  // extern "C"
