// Dense screened interaction of the one-dimensional GW driver: green_calc_w_full
// (src/green_rgf.f90:3347-3396), called once per energy at src/green_rgf.f90:3235.
// For every energy e, with V the block-tridiagonal bare interaction (diagonal blocks
// Vii, block (ix+1, ix) = V1i_ix and block (ix, ix+1) = V1i_ix^H: block2full,
// src/green_rgf.f90:3326-3345) and P^{<,>,r} block-diagonal:
//   M = I - V P^r,  W^r = M^-1 V,  W^{<,>} = W^r P^{<,>} W^r^H,
// of which the diagonal blocks are returned.  The reference inverts M (zgetrf +
// zgetri) and multiplies; here M W^r = V is solved with the LU factors (zgetrs), the
// same matrix at 4/3 instead of 2 NT^3 complex multiply-adds.  The products with the
// block-diagonal P are block-column scalings, and only the diagonal blocks of
// W^r P W^r^H are formed: (W^r P)[rows of i, :] times W^r[rows of i, :]^H.
//
// These are plain dense library operations on NT x NT matrices (NT = nm nx): rocBLAS
// strided-batched ZGEMM and rocSOLVER strided-batched ZGETRF / ZGETRS over a batch of
// energies, with hand-written kernels only for the block copies.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/gw.h"

namespace gwerr {
int fail(int code, const std::string& msg);  // gw_step.hip: sets gw_last_error()
void clear();
}  // namespace gwerr

namespace {

int failw(int code, const std::string& msg) { return gwerr::fail(code, msg); }

#define HIPW(expr)                                                                            \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return failw(2, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define RBW(expr)                                                                           \
  do {                                                                                      \
    rocblas_status s_ = (expr);                                                             \
    if (s_ != rocblas_status_success)                                                       \
      return failw(2, std::string(#expr) + ": " + rocblas_status_to_string(s_));            \
  } while (0)

typedef rocblas_double_complex zc;

// Dense V (NT x NT, column-major) from the blocks: diagonal Vii_ix, (ix+1, ix) = V1i_ix,
// (ix, ix+1) = V1i_ix^H, zero elsewhere.
__global__ void k_block2full(int nm, int nx, const double2* __restrict__ Vii, const double2* __restrict__ V1i,
                             double2* __restrict__ V) {
  const size_t NT = (size_t)nm * nx;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < NT * NT;
       idx += (size_t)gridDim.x * blockDim.x) {
    const size_t r = idx % NT, c = idx / NT;
    const int br = (int)(r / nm), bc = (int)(c / nm);
    const int i = (int)(r % nm), j = (int)(c % nm);
    const size_t nn = (size_t)nm * nm;
    double2 v = make_double2(0.0, 0.0);
    if (br == bc) v = Vii[(size_t)br * nn + i + (size_t)j * nm];
    else if (br == bc + 1) v = V1i[(size_t)bc * nn + i + (size_t)j * nm];
    else if (bc == br + 1) {
      const double2 t = V1i[(size_t)br * nn + j + (size_t)i * nm];
      v = make_double2(t.x, -t.y);
    }
    V[idx] = v;
  }
}

// B matrices: copies of V (the right-hand sides of M W = V) and M = I + M (the GEMM left
// -V P^r in M), for nb energies.
__global__ void k_rhs_and_identity(size_t NT, int nb, const double2* __restrict__ V, double2* __restrict__ W,
                                   double2* __restrict__ M) {
  const size_t tot = NT * NT * (size_t)nb;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < tot; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t k = idx % (NT * NT);
    W[idx] = V[k];
    if (k % NT == k / NT) M[idx].x += 1.0;
  }
}

// Diagonal blocks of W^r (one energy slab block per (ix, e)) into the output.
__global__ void k_diag_blocks(int nm, int nx, int nb, const double2* __restrict__ W, double2* __restrict__ out) {
  const size_t nn = (size_t)nm * nm, NT = (size_t)nm * nx;
  const size_t tot = nn * nx * (size_t)nb;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < tot; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t e = idx / (nn * nx), rem = idx % (nn * nx);
    const size_t ix = rem / nn, q = rem % nn;
    const size_t i = q % nm, j = q / nm;
    out[idx] = W[e * NT * NT + (ix * nm + i) + (ix * nm + j) * NT];
  }
}

struct WFullCtx {
  rocblas_handle h = nullptr;
  void* buf = nullptr;
  size_t bytes = 0;
};
constexpr int kMaxDevW = 64;
WFullCtx g_wctx[kMaxDevW];
std::mutex g_wctx_mu;

}  // namespace

extern "C" {

int gw_w_full(int nm, int nx, int nE, const double* Vii, const double* V1i, const double* PL, const double* PG,
              const double* PR, double* WL, double* WG, double* WR, void* stream) {
  gwerr::clear();
  if (nm < 1 || nx < 1 || nE < 0) return failw(1, "gw_w_full: need nm >= 1, nx >= 1, nE >= 0");
  if (!Vii || !V1i || !PL || !PG || !PR || !WL || !WG || !WR) return failw(1, "gw_w_full: null array");
  if ((long long)nm * nx > 16384) return failw(4, "gw_w_full: nm * nx > 16384 (dense NT x NT solve)");
  if (nE == 0) return 0;
  const int NT = nm * nx;
  const size_t nn = (size_t)nm * nm, NT2 = (size_t)NT * NT;
  int dev = 0;
  HIPW(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevW) return failw(1, "gw_w_full: device index out of range");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g_wctx_mu);
  WFullCtx& c = g_wctx[dev];
  if (!c.h) RBW(rocblas_create_handle(&c.h));
  RBW(rocblas_set_stream(c.h, s));
  // per energy of a batch: M (LU), W^r, W^r P, pivots and info; plus V once
  const size_t per = 3 * NT2 * sizeof(double2) + (size_t)NT * sizeof(rocblas_int) + sizeof(rocblas_int);
  size_t fr = 0, tot = 0;
  HIPW(hipMemGetInfo(&fr, &tot));
  const size_t avail = (size_t)((double)(fr + c.bytes) * 0.8);
  const size_t fixed = NT2 * sizeof(double2);
  if (avail < fixed + per) return failw(3, "gw_w_full: not enough device memory for one energy");
  const int nb = (int)std::min<size_t>((size_t)nE, std::min<size_t>(64, (avail - fixed) / per));
  const size_t need = fixed + per * nb;
  if (c.bytes < need) {
    if (c.buf) HIPW(hipFree(c.buf));
    c.buf = nullptr;
    c.bytes = 0;
    if (hipMalloc(&c.buf, need) != hipSuccess) {
      (void)hipGetLastError();
      return failw(3, "gw_w_full: cannot allocate " + std::to_string(need) + " bytes");
    }
    c.bytes = need;
  }
  double2* V = reinterpret_cast<double2*>(c.buf);
  double2* M = V + NT2;
  double2* W = M + NT2 * nb;
  double2* X = W + NT2 * nb;
  rocblas_int* ipiv = reinterpret_cast<rocblas_int*>(X + NT2 * nb);
  rocblas_int* info = ipiv + (size_t)NT * nb;
  const unsigned g = 4096;
  hipLaunchKernelGGL(k_block2full, dim3(g), dim3(256), 0, s, nm, nx, reinterpret_cast<const double2*>(Vii),
                     reinterpret_cast<const double2*>(V1i), V);
  HIPW(hipGetLastError());
  const zc mone = {-1.0, 0.0}, one = {1.0, 0.0}, zero = {0.0, 0.0};
  const size_t eb = nn * nx;  // elements of one energy's blocks
  for (int e0 = 0; e0 < nE; e0 += nb) {
    const int b = std::min(nb, nE - e0);
    // M = -V P^r (block column j of M = -V[:, j] P^r_j), then M += I and W = V
    for (int k = 0; k < b; ++k) {
      const zc* P = reinterpret_cast<const zc*>(PR) + (size_t)(e0 + k) * eb;
      RBW(rocblas_zgemm_strided_batched(c.h, rocblas_operation_none, rocblas_operation_none, NT, nm, nm, &mone,
                                        reinterpret_cast<const zc*>(V), NT, (rocblas_stride)nm * NT, P, nm,
                                        (rocblas_stride)nn, &zero, reinterpret_cast<zc*>(M + k * NT2), NT,
                                        (rocblas_stride)nm * NT, nx));
    }
    hipLaunchKernelGGL(k_rhs_and_identity, dim3(g), dim3(256), 0, s, (size_t)NT, b, V, W, M);
    HIPW(hipGetLastError());
    RBW(rocsolver_zgetrf_strided_batched(c.h, NT, NT, reinterpret_cast<zc*>(M), NT, (rocblas_stride)NT2, ipiv,
                                         (rocblas_stride)NT, info, b));
    RBW(rocsolver_zgetrs_strided_batched(c.h, rocblas_operation_none, NT, NT, reinterpret_cast<zc*>(M), NT,
                                         (rocblas_stride)NT2, ipiv, (rocblas_stride)NT, reinterpret_cast<zc*>(W),
                                         NT, (rocblas_stride)NT2, b));
    hipLaunchKernelGGL(k_diag_blocks, dim3(g), dim3(256), 0, s, nm, nx, b, W,
                       reinterpret_cast<double2*>(WR) + (size_t)e0 * eb);
    HIPW(hipGetLastError());
    // W^{<,>} diagonal blocks: X = W^r P (block columns), then X[rows i, :] W^r[rows i, :]^H
    const double* Ps[2] = {PL, PG};
    double* Outs[2] = {WL, WG};
    for (int q = 0; q < 2; ++q) {
      for (int k = 0; k < b; ++k) {
        const zc* P = reinterpret_cast<const zc*>(Ps[q]) + (size_t)(e0 + k) * eb;
        RBW(rocblas_zgemm_strided_batched(c.h, rocblas_operation_none, rocblas_operation_none, NT, nm, nm, &one,
                                          reinterpret_cast<const zc*>(W + k * NT2), NT, (rocblas_stride)nm * NT, P,
                                          nm, (rocblas_stride)nn, &zero, reinterpret_cast<zc*>(X + k * NT2), NT,
                                          (rocblas_stride)nm * NT, nx));
        zc* out = reinterpret_cast<zc*>(Outs[q]) + (size_t)(e0 + k) * eb;
        RBW(rocblas_zgemm_strided_batched(c.h, rocblas_operation_none, rocblas_operation_conjugate_transpose, nm, nm,
                                          NT, &one, reinterpret_cast<const zc*>(X + k * NT2), NT, (rocblas_stride)nm,
                                          reinterpret_cast<const zc*>(W + k * NT2), NT, (rocblas_stride)nm, &zero,
                                          out, nm, (rocblas_stride)nn, nx));
      }
    }
  }
  return 0;
}

}  // extern "C"
