// rgf_engine.hip -- host side of the MI355X RGF engine: the C ABI of
// include/rgf.h, kernel-variant selection, workspace admission and the
// host-array entry points that stage data through HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rgf.h"
#include "rgf_adjoint.hpp"
#include "rgf_integrate.hpp"
#include "rgf_multi.hpp"
#include "rgf_offdiag.hpp"
#include "rgf_sweep.hpp"
#include "rgf_wsweep.hpp"

namespace {

thread_local std::string g_err;
std::mutex g_mu;  // serialises the host-array entry points (shared staging buffers)

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      return fail(_e == hipErrorOutOfMemory ? RGF_ENOMEM : RGF_EHIP,                      \
                  std::string(#expr) + ": " + hipGetErrorString(_e));                     \
  } while (0)

// Device buffer that only grows; one set per device for the engine workspace
// and one for the host-array staging area.
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t want) {
    if (want <= bytes) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(e == hipErrorOutOfMemory ? RGF_ENOMEM : RGF_EHIP,
                  std::string("hipMalloc workspace: ") + hipGetErrorString(e));
    }
    bytes = want;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

constexpr int kMaxDev = 16;
Buf g_ws[kMaxDev];
std::mutex g_ws_mu;

// Copy nblocks n x n blocks into np x np blocks (np >= n): zero padding and
// `fill` on the padded diagonal, which keeps the padded orbitals decoupled and
// every padded slab matrix invertible.
__global__ void pad_blocks(const double2* __restrict__ src, int n, double2* __restrict__ dst, int np,
                           long long nblocks, double2 fill) {
  const size_t tot = (size_t)nblocks * np * np;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < tot;
       idx += (size_t)gridDim.x * blockDim.x) {
    const size_t b = idx / ((size_t)np * np);
    const int r = (int)(idx % ((size_t)np * np));
    const int i = r % np, j = r / np;
    double2 v = make_double2(0, 0);
    if (i < n && j < n) v = src[b * n * n + i + (size_t)j * n];
    else if (i == j) v = fill;
    dst[idx] = v;
  }
}

template <typename V>
__global__ void unpad_blocks(const V* __restrict__ src, int np, V* __restrict__ dst, int n, long long nblocks) {
  const size_t tot = (size_t)nblocks * n * n;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < tot;
       idx += (size_t)gridDim.x * blockDim.x) {
    const size_t b = idx / ((size_t)n * n);
    const int r = (int)(idx % ((size_t)n * n));
    const int i = r % n, j = r / n;
    dst[idx] = src[b * np * np + i + (size_t)j * np];
  }
}

Buf g_stage[kMaxDev];
Buf g_adj_dh[kMaxDev];

typedef void (*AdjLaunch)(const rgf::AdjArgs&, int grid, hipStream_t);
template <int T, int NMAX>
void launch_adj(const rgf::AdjArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL((rgf::rgf_adjoint<T, NMAX>), dim3(grid), dim3(rgf::Cfg<T>::NTH), 0, s, a);
}
struct AdjVariant {
  AdjLaunch launch;
  const void* fn;
  int nth;
};
bool pick_adj(int n, AdjVariant* v) {
  if (n % 32 != 0 || n < 32 || n > 256) return false;
#define ADJV(T, NM) {launch_adj<T, NM>, reinterpret_cast<const void*>(&rgf::rgf_adjoint<T, NM>), rgf::Cfg<T>::NTH}
  switch (n) {
    case 32: *v = ADJV(1, 32); return true;
    case 64: *v = ADJV(2, 64); return true;
    case 96: *v = ADJV(3, 96); return true;
    case 128: *v = ADJV(4, 128); return true;
    case 192: *v = ADJV(3, 192); return true;
    case 256: *v = ADJV(4, 256); return true;
    default: *v = ADJV(1, 256); return true;
  }
#undef ADJV
}

typedef void (*SweepLaunch)(const rgf::SweepArgs&, int grid, hipStream_t);

template <int T, int NMAX>
void launch_sweep(const rgf::SweepArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL((rgf::rgf_sweep<T, NMAX>), dim3(grid), dim3(rgf::Cfg<T>::NTH), 0, s, a);
}

template <int T, int NMAX>
const void* sweep_fn() {
  return reinterpret_cast<const void*>(&rgf::rgf_sweep<T, NMAX>);
}

template <int T, int NMAX>
void launch_sweep_cms(const rgf::SweepArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL((rgf::rgf_sweep_cms<T, NMAX>), dim3(grid), dim3(rgf::Cfg<T>::NTH), 0, s, a);
}

template <int T, int NMAX>
const void* sweep_cms_fn() {
  return reinterpret_cast<const void*>(&rgf::rgf_sweep_cms<T, NMAX>);
}

template <int T, int NMAX>
void launch_sweep_od(const rgf::SweepArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL((rgf::rgf_sweep_od<T, NMAX>), dim3(grid), dim3(rgf::Cfg<T>::NTH), 0, s, a);
}

template <int T, int NMAX>
const void* sweep_od_fn() {
  return reinterpret_cast<const void*>(&rgf::rgf_sweep_od<T, NMAX>);
}

struct Variant {
  SweepLaunch launch;
  const void* fn;
  int T, nmax, nth;
};

// Largest block size of the plain sweep, and the granule above 256 (blocks
// above 256 run with several inverse panel rows per thread).
constexpr int kNmaxBig = 512;
// Tile-variant offset of the kernels compiled for two workgroups per CU (the
// 32- and 64-orbital blocks, whose single 256-thread workgroup leaves the MFMA
// pipes idle during the latency-bound column steps of the inverse and the short
// K loops of the GEMM).
constexpr int kOcc2 = 8;

// Tile variant per block size: wave tile (16T)^2, macro tile (32T)^2; the
// inverse's LDS panel is sized for NMAX rows.  Above 256: n a multiple of 64,
// 128 x 64 macro tiles when n is a multiple of 128, else 64 x 64.
bool pick_variant(int n, Variant* v) {
  if (n > 256 && n <= kNmaxBig && n % 64 == 0) {
    if (n % 128 == 0) *v = {launch_sweep<4, kNmaxBig>, sweep_fn<4, kNmaxBig>(), 4, kNmaxBig, rgf::Cfg<4>::NTH};
    else *v = {launch_sweep<2, kNmaxBig>, sweep_fn<2, kNmaxBig>(), 2, kNmaxBig, rgf::Cfg<2>::NTH};
    return true;
  }
  if (n % 32 != 0 || n < 32 || n > 256) return false;
  switch (n) {
    case 32: *v = {launch_sweep<kOcc2 + 1, 32>, sweep_fn<kOcc2 + 1, 32>(), 1, 32, rgf::Cfg<1>::NTH}; return true;
    case 64: *v = {launch_sweep<kOcc2 + 2, 64>, sweep_fn<kOcc2 + 2, 64>(), 2, 64, rgf::Cfg<2>::NTH}; return true;
    case 96: *v = {launch_sweep<3, 96>, sweep_fn<3, 96>(), 3, 96, rgf::Cfg<3>::NTH}; return true;
    case 128: *v = {launch_sweep<4, 128>, sweep_fn<4, 128>(), 4, 128, rgf::Cfg<4>::NTH}; return true;
    case 192: *v = {launch_sweep<3, 192>, sweep_fn<3, 192>(), 3, 192, rgf::Cfg<3>::NTH}; return true;
    case 256: *v = {launch_sweep<4, 256>, sweep_fn<4, 256>(), 4, 256, rgf::Cfg<4>::NTH}; return true;
    default: *v = {launch_sweep<1, 256>, sweep_fn<1, 256>(), 1, 256, rgf::Cfg<1>::NTH}; return true;
  }
}

// The same tile variants for the block-tridiagonal-Sigma kernel.
bool pick_variant_od(int n, Variant* v) {
  if (n % 32 != 0 || n < 32 || n > 256) return false;
  switch (n) {
    case 32: *v = {launch_sweep_od<kOcc2 + 1, 32>, sweep_od_fn<kOcc2 + 1, 32>(), 1, 32, rgf::Cfg<1>::NTH}; return true;
    case 64: *v = {launch_sweep_od<kOcc2 + 2, 64>, sweep_od_fn<kOcc2 + 2, 64>(), 2, 64, rgf::Cfg<2>::NTH}; return true;
    case 96: *v = {launch_sweep_od<3, 96>, sweep_od_fn<3, 96>(), 3, 96, rgf::Cfg<3>::NTH}; return true;
    case 128: *v = {launch_sweep_od<4, 128>, sweep_od_fn<4, 128>(), 4, 128, rgf::Cfg<4>::NTH}; return true;
    case 192: *v = {launch_sweep_od<3, 192>, sweep_od_fn<3, 192>(), 3, 192, rgf::Cfg<3>::NTH}; return true;
    case 256: *v = {launch_sweep_od<4, 256>, sweep_od_fn<4, 256>(), 4, 256, rgf::Cfg<4>::NTH}; return true;
    default: *v = {launch_sweep_od<1, 256>, sweep_od_fn<1, 256>(), 1, 256, rgf::Cfg<1>::NTH}; return true;
  }
}

// The same tile variants for the overlap (green_RGF_CMS) kernel.
bool pick_variant_cms(int n, Variant* v) {
  if (n % 32 != 0 || n < 32 || n > 256) return false;
  switch (n) {
    case 32: *v = {launch_sweep_cms<1, 32>, sweep_cms_fn<1, 32>(), 1, 32, rgf::Cfg<1>::NTH}; return true;
    case 64: *v = {launch_sweep_cms<2, 64>, sweep_cms_fn<2, 64>(), 2, 64, rgf::Cfg<2>::NTH}; return true;
    case 96: *v = {launch_sweep_cms<3, 96>, sweep_cms_fn<3, 96>(), 3, 96, rgf::Cfg<3>::NTH}; return true;
    case 128: *v = {launch_sweep_cms<4, 128>, sweep_cms_fn<4, 128>(), 4, 128, rgf::Cfg<4>::NTH}; return true;
    case 192: *v = {launch_sweep_cms<3, 192>, sweep_cms_fn<3, 192>(), 3, 192, rgf::Cfg<3>::NTH}; return true;
    case 256: *v = {launch_sweep_cms<4, 256>, sweep_cms_fn<4, 256>(), 4, 256, rgf::Cfg<4>::NTH}; return true;
    default: *v = {launch_sweep_cms<1, 256>, sweep_cms_fn<1, 256>(), 1, 256, rgf::Cfg<1>::NTH}; return true;
  }
}

// The screened-interaction sweep (green_rgf_calc_w) on the same tile variants.
typedef void (*WLaunch)(const rgf::WArgs&, int grid, hipStream_t);
template <int T, int NMAX>
void launch_w(const rgf::WArgs& a, int grid, hipStream_t s) {
  hipLaunchKernelGGL((rgf::rgf_wsweep<T, NMAX>), dim3(grid), dim3(rgf::Cfg<T>::NTH), 0, s, a);
}
struct WVariant {
  WLaunch launch;
  const void* fn;
  int nth;
};
// Supported block size at or above n (the blocks are zero-padded up to it).
int w_block(int n) {
  if (n <= 256) return (n + 31) / 32 * 32;
  return (n + 63) / 64 * 64;
}
bool pick_w(int n, WVariant* v) {
#define WV(T, NM) {launch_w<T, NM>, reinterpret_cast<const void*>(&rgf::rgf_wsweep<T, NM>), rgf::Cfg<T>::NTH}
  if (n > 256 && n <= kNmaxBig && n % 64 == 0) {
    if (n % 128 == 0) *v = WV(4, kNmaxBig);
    else *v = WV(2, kNmaxBig);
    return true;
  }
  if (n % 32 != 0 || n < 32 || n > 256) return false;
  switch (n) {
    case 32: *v = WV(kOcc2 + 1, 32); return true;
    case 64: *v = WV(kOcc2 + 2, 64); return true;
    case 96: *v = WV(3, 96); return true;
    case 128: *v = WV(4, 128); return true;
    case 192: *v = WV(3, 192); return true;
    case 256: *v = WV(4, 256); return true;
    default: *v = WV(1, 256); return true;
  }
#undef WV
}

__global__ void conj_blocks(const double2* __restrict__ src, double2* __restrict__ dst, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    const double2 v = src[i];
    dst[i] = make_double2(v.x, -v.y);
  }
}

long long slot_elems(int n, int nx, int ckpt = 0) {
  return rgf::sweep_slot_blocks(nx, ckpt, rgf::kTmpBlocks) * n * n;
}
long long slot_elems_od(int n, int nx) { return (long long)(2 * nx + rgf::kTmpBlocksOd) * n * n; }

bool has_offdiag(const rgf_batch* b) { return b->sigR_lo || b->sigR_up || b->sigL_up; }

int check_params(const rgf_params* p) {
  if (!p) return fail(RGF_EINVAL, "params is NULL");
  if (p->mode != RGF_MODE_COMPAT && p->mode != RGF_MODE_PHYSICAL) return fail(RGF_EINVAL, "unknown mode");
  if (p->dec_nmax < 1) return fail(RGF_EINVAL, "dec_nmax must be >= 1");
  return 0;
}

// Everything a sweep launch needs: variant, slot count, kernel arguments.
struct Prepared {
  Variant v;
  rgf::SweepArgs a;
  long long slots = 0, slot_elems = 0;
  int dev = 0;
  bool nothing = false;
};

void fill_args(const rgf_params* p, const rgf_batch* b, long long se, rgf::SweepArgs* out);

// Validate a batch, pick the kernel and size the slots (HBM admission).  The
// workspace pointer is left null; the caller binds the shared or an owned one.
int prepare(const rgf_params* p, const rgf_batch* b, const rgf_scba* scba, Prepared* out) {
  if (int rc = check_params(p)) return rc;
  if (!b) return fail(RGF_EINVAL, "batch is NULL");
  if (b->nx < 2) return fail(RGF_EINVAL, "nx must be >= 2 (the reference reads Gl(:,:,nx-1))");
  if (b->nE < 0) return fail(RGF_EINVAL, "nE < 0");
  if (b->nE == 0) {
    out->nothing = true;
    return 0;
  }
  if (!b->E || !b->Hii || !b->H1i) return fail(RGF_EINVAL, "E, Hii and H1i are required");
  const bool od = has_offdiag(b) || scba;
  const bool cms = b->Sii != nullptr;
  if (od && p->mode != RGF_MODE_PHYSICAL)
    return fail(RGF_EINVAL, "off-diagonal Sigma and the SCBA refresh need RGF_MODE_PHYSICAL");
  if (cms && (od || p->mode != RGF_MODE_COMPAT))
    return fail(RGF_EINVAL, "the overlap (green_RGF_CMS) solve has compat semantics only, without off-diagonal Sigma");
  // The overlap kernel (rgf_sweep_cms) is switched off: its first run at nm = 64
  // ended in a GPU memory fault whose cause has not been found.
  if (cms) return fail(RGF_EUNSUP, "the overlap (green_RGF_CMS) GPU solve is switched off (unresolved memory fault)");
  if (b->ckpt < -1) return fail(RGF_EINVAL, "ckpt must be -1 (keep every slab), 0 (automatic) or a segment length");
  if (b->ckpt > 0 && (od || cms))
    return fail(RGF_EUNSUP, "segment recompute (ckpt > 0) is for the diagonal-Sigma solve only");
  if (b->GRi1_lo && !od)
    return fail(RGF_EUNSUP, "GRi1_lo is written by the block-tridiagonal-Sigma solve only");
  if (scba) {
    if (!b->sigR || !b->sigL) return fail(RGF_EINVAL, "SCBA refresh: sigR and sigL are required");
    if (b->sig_stride != (long long)b->nm * b->nm * b->nx)
      return fail(RGF_EINVAL, "SCBA refresh: Sigma must be per energy (sig_stride = nm*nm*nx)");
  }
  Variant& v = out->v;
  if (!(od ? pick_variant_od(b->nm, &v) : cms ? pick_variant_cms(b->nm, &v) : pick_variant(b->nm, &v)))
    return fail(RGF_EUNSUP, od || cms ? "rgf_solve_batched: with off-diagonal Sigma or overlaps nm must be a multiple of 32 in [32, 256]"
                                      : "rgf_solve_batched: nm must be a multiple of 32 in [32, 256] or of 64 in (256, 512]");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDev) return fail(RGF_EINVAL, "device index out of range");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  int per_cu = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, v.fn, v.nth, 0));
  if (per_cu < 1) return fail(RGF_EHIP, "sweep kernel does not fit on a CU");
  const long long se = od ? slot_elems_od(b->nm, b->nx) : slot_elems(b->nm, b->nx, std::max(b->ckpt, 0));
  long long slots = std::min<long long>(b->nE, (long long)prop.multiProcessorCount * per_cu);
  if (b->max_slots > 0) slots = std::min<long long>(slots, b->max_slots);
  out->slots = slots;
  out->slot_elems = se;
  out->dev = dev;
  fill_args(p, b, se, &out->a);
  if (scba) {
    out->a.scba = 1;
    out->a.scba_D = scba->D;
    out->a.scba_mix = scba->mix;
  }
  return 0;
}

// Kernel arguments of one batch (workspace pointer left null).
void fill_args(const rgf_params* p, const rgf_batch* b, long long se, rgf::SweepArgs* out) {
  rgf::SweepArgs& a = *out;
  std::memset(&a, 0, sizeof(a));
  a.n = b->nm;
  a.nx = b->nx;
  a.nE = b->nE;
  a.mode = p->mode;
  a.dec_nmax = p->dec_nmax;
  a.eta_dev = p->eta_dev;
  a.eta_lead = p->eta_lead;
  a.dec_tol = p->dec_tol;
  a.boltz = p->boltz;
  a.temp[0] = b->temp[0];
  a.temp[1] = b->temp[1];
  a.mu[0] = b->mu[0];
  a.mu[1] = b->mu[1];
  a.E = b->E;
  a.Hii = reinterpret_cast<const double2*>(b->Hii);
  a.H1i = reinterpret_cast<const double2*>(b->H1i);
  a.sigL = reinterpret_cast<const double2*>(b->sigL);
  a.sigR = reinterpret_cast<const double2*>(b->sigR);
  a.sig_stride = b->sig_stride;
  a.ldos = reinterpret_cast<double2*>(b->ldos);
  a.ndens = reinterpret_cast<double2*>(b->ndens);
  a.pdens = reinterpret_cast<double2*>(b->pdens);
  a.GRi1 = reinterpret_cast<double2*>(b->GRi1);
  a.GLi1 = reinterpret_cast<double2*>(b->GLi1);
  a.GGi1 = reinterpret_cast<double2*>(b->GGi1);
  a.GRi1_lo = reinterpret_cast<double2*>(b->GRi1_lo);
  a.cur = b->cur;
  a.tr = b->tr;
  a.tre = b->tre;
  a.dens_n = b->dens_n;
  a.dens_p = b->dens_p;
  a.info = b->info;
  a.ws_slot = se;
  a.sigR_lo = reinterpret_cast<const double2*>(b->sigR_lo);
  a.sigR_up = reinterpret_cast<const double2*>(b->sigR_up);
  a.sigL_up = reinterpret_cast<const double2*>(b->sigL_up);
  a.Sii = reinterpret_cast<const double2*>(b->Sii);
  a.ckpt = std::max(b->ckpt, 0);
}

int solve_impl(const rgf_params* p, const rgf_batch* b, const rgf_scba* scba, void* stream) {
  g_err.clear();
  Prepared pr;
  if (int rc = prepare(p, b, scba, &pr)) return rc;
  if (pr.nothing) return 0;
  long long slot_bytes = pr.slot_elems * 16LL;
  long long slots = pr.slots;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  Buf& ws = g_ws[pr.dev];
  if ((size_t)(slots * slot_bytes) > ws.bytes) {
    // HBM admission: fit the stored left-connected blocks into free memory
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    const size_t avail = (size_t)((double)(fr + ws.bytes) * 0.9);
    long long fit = (long long)(avail / (size_t)slot_bytes);
    // automatic segment recompute: when the full slots leave CUs idle, keep the
    // blocks only at checkpoints every ceil(sqrt(nx)) slabs and one segment, at the
    // cost of one more forward sweep (7 of the 17 GEMM-equivalents of a slab), if
    // that runs more energies per unit of time
    const bool od = has_offdiag(b) || scba;
    if (b->ckpt == 0 && !od && !b->Sii && fit < slots) {
      const int K = std::max(2, (int)std::ceil(std::sqrt((double)b->nx)));
      const long long se_k = slot_elems(b->nm, b->nx, K);
      const long long fit_k = std::min<long long>(slots, (long long)(avail / (size_t)(se_k * 16LL)));
      if (fit_k * 17 > fit * 24) {
        pr.a.ckpt = K;
        pr.a.ws_slot = se_k;
        slot_bytes = se_k * 16LL;
        fit = fit_k;
      }
    }
    if (fit < 1) return fail(RGF_ENOMEM, "not enough device memory for one energy slot");
    slots = std::min(slots, fit);
    if ((size_t)(slots * slot_bytes) > ws.bytes) {
      ws.release();
      if (int rc = ws.ensure((size_t)(slots * slot_bytes))) return rc;
    }
  }
  pr.a.ws = reinterpret_cast<double2*>(ws.p);
  pr.v.launch(pr.a, (int)slots, reinterpret_cast<hipStream_t>(stream));
  HIPCHK(hipGetLastError());
  return 0;
}

// ------------------------------------------------- allocation-free workspace
// Workspace of rgf_solve_batched_noalloc: a fixed device array, so the entry
// point allocates nothing.  Every GPU the library runs on holds its own copy
// (kScratchBytes) until the process ends.  In host code `scratch` names a
// host-side stand-in; its device address is resolved once per device with
// hipGetSymbolAddress and only that address is handed to kernels.
constexpr size_t kScratchBytes = size_t(1) << 30;
static __device__ __align__(256) double2 scratch[kScratchBytes / sizeof(double2)];
void* g_scratch[kMaxDev];
std::mutex g_scratch_mu;

int scratch_address(int dev, double2** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(RGF_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = ptr;
  }
  *out = reinterpret_cast<double2*>(g_scratch[dev]);
  return 0;
}

int solve_noalloc_impl(const rgf_params* p, const rgf_batch* b, void* stream) {
  g_err.clear();
  Prepared pr;
  if (int rc = prepare(p, b, nullptr, &pr)) return rc;
  if (pr.nothing) return 0;
  const long long slot_bytes = pr.slot_elems * 16LL;
  const long long fit = (long long)(kScratchBytes / (size_t)slot_bytes);
  if (fit < 1)
    return fail(RGF_EUNSUP, "rgf_solve_batched_noalloc: one energy slot needs " + std::to_string(slot_bytes) +
                                " bytes ((2*nx + " + std::to_string(pr.slot_elems / ((long long)b->nm * b->nm) - 2 * b->nx) +
                                ") blocks of nm*nm*16 bytes), more than the fixed workspace of " +
                                std::to_string((long long)kScratchBytes) + " bytes");
  double2* ptr = nullptr;
  if (int rc = scratch_address(pr.dev, &ptr)) return rc;
  pr.a.ws = ptr;
  pr.v.launch(pr.a, (int)std::min(pr.slots, fit), reinterpret_cast<hipStream_t>(stream));
  HIPCHK(hipGetLastError());
  return 0;
}


// ------------------------------------------------------------ many devices
int mv_T(int v) {
  static const int t[rgf::MV_COUNT] = {1, 2, 1, 3, 1, 4, 2, 3, 1, 4, 2, 1};
  return (v >= 0 && v < rgf::MV_COUNT) ? t[v] : 0;
}

int default_mv(int n) {
  switch (n) {
    case 32: return rgf::MV_1_32;
    case 64: return rgf::MV_2_64;
    case 96: return rgf::MV_3_96;
    case 128: return rgf::MV_4_128;
    case 192: return rgf::MV_3_192;
    case 256: return rgf::MV_4_256;
    default: return rgf::MV_1_256;
  }
}

std::vector<int> candidate_mv(int n) {
  switch (n) {
    case 64: return {rgf::MV_2_64, rgf::MV_1_64};
    case 96: return {rgf::MV_3_96, rgf::MV_1_96};
    case 128: return {rgf::MV_4_128, rgf::MV_2_128};
    case 192: return {rgf::MV_3_192, rgf::MV_1_192};
    case 256: return {rgf::MV_4_256, rgf::MV_2_256};
    default: return {default_mv(n)};
  }
}

// Autotuner memory per device: block size -> (variant, seconds per slab-energy
// measured with the machine full).
struct Tuned {
  int mv;
  double t_slab;
};
std::map<int, Tuned> g_tuned[kMaxDev];
std::mutex g_multi_mu;
Buf g_multi_desc[kMaxDev];
int g_multi_cu[kMaxDev];

int multi_per_cu(int* per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, reinterpret_cast<const void*>(&rgf::rgf_sweep_multi),
                                                        256, 0) == hipSuccess ? 0 : 1;
}

// Device buffers of one multi-device launch, laid out in one allocation.
struct MultiDesc {
  size_t off_variant, off_items, off_counter, bytes;
};
MultiDesc multi_layout(int ndev, long long nitems) {
  MultiDesc d;
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  d.off_variant = al(sizeof(rgf::SweepArgs) * (size_t)ndev);
  d.off_items = d.off_variant + al(sizeof(int) * (size_t)ndev);
  d.off_counter = d.off_items + al(sizeof(int2) * (size_t)std::max<long long>(nitems, 1));
  d.bytes = d.off_counter + 256;
  return d;
}

// Copy the descriptors of one multi-device launch into `desc` (synchronous).
int multi_upload(const std::vector<rgf::SweepArgs>& jobs, const std::vector<int>& mv, const std::vector<int2>& items,
                 void* desc) {
  const MultiDesc L = multi_layout((int)jobs.size(), (long long)items.size());
  char* base = reinterpret_cast<char*>(desc);
  HIPCHK(hipMemcpy(base, jobs.data(), sizeof(rgf::SweepArgs) * jobs.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(base + L.off_variant, mv.data(), sizeof(int) * mv.size(), hipMemcpyHostToDevice));
  if (!items.empty())
    HIPCHK(hipMemcpy(base + L.off_items, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice));
  return 0;
}

// Counter reset (a fill kernel, so the pair can be captured in a graph) and the
// persistent launch over the uploaded descriptors.
int multi_launch(int ndev, long long nitems, void* desc, double2* ws, long long ws_slot, int slots, hipStream_t s) {
  const MultiDesc L = multi_layout(ndev, nitems);
  char* base = reinterpret_cast<char*>(desc);
  rgf::MultiArgs m;
  m.jobs = reinterpret_cast<const rgf::SweepArgs*>(base);
  m.variant = reinterpret_cast<const int*>(base + L.off_variant);
  m.items = reinterpret_cast<const int2*>(base + L.off_items);
  m.nitems = (int)nitems;
  m.counter = reinterpret_cast<int*>(base + L.off_counter);
  m.ws = ws;
  m.ws_slot = ws_slot;
  hipLaunchKernelGGL(rgf::k_fill_int, dim3(1), dim3(64), 0, s, m.counter, 1, 0);
  hipLaunchKernelGGL(rgf::rgf_sweep_multi, dim3(slots), dim3(256), 0, s, m);
  HIPCHK(hipGetLastError());
  return 0;
}

// Validate the devices and build their kernel arguments.
int multi_jobs(const rgf_params* p, const rgf_batch* devs, int ndev, std::vector<rgf::SweepArgs>* jobs) {
  if (int rc = check_params(p)) return rc;
  if (!devs || ndev < 1) return fail(RGF_EINVAL, "rgf_solve_devices: no devices");
  jobs->resize(ndev);
  for (int j = 0; j < ndev; ++j) {
    const rgf_batch* b = &devs[j];
    if (b->nm % 32 != 0 || b->nm < 32 || b->nm > 256)
      return fail(RGF_EUNSUP, "rgf_solve_devices: device " + std::to_string(j) + ": nm must be a multiple of 32 in [32, 256]");
    if (b->nx < 2 || b->nE < 0) return fail(RGF_EINVAL, "rgf_solve_devices: device " + std::to_string(j) + ": need nx >= 2, nE >= 0");
    if (b->nE > 0 && (!b->E || !b->Hii || !b->H1i))
      return fail(RGF_EINVAL, "rgf_solve_devices: device " + std::to_string(j) + ": E, Hii and H1i are required");
    if (has_offdiag(b) || b->Sii)
      return fail(RGF_EUNSUP, "rgf_solve_devices: off-diagonal Sigma and overlaps are not supported here");
    fill_args(p, b, 0, &(*jobs)[j]);
  }
  return 0;
}

// Autotuner: for every block size with more than one candidate variant, time
// each candidate on a slice of the first device of that size (its first
// energy, at most 12 slabs, 2 decimation iterations), replicated on every
// workgroup so the machine is as loaded as in the real launch; keep the fastest.
int multi_autotune(const std::vector<rgf::SweepArgs>& jobs, int dev, int slots, hipStream_t s, int* ntuned) {
  std::map<int, int> first;
  for (int j = 0; j < (int)jobs.size(); ++j)
    if (jobs[j].nE > 0 && !first.count(jobs[j].n)) first[jobs[j].n] = j;
  *ntuned = 0;
  for (auto& kv : first) {
    const int n = kv.first;
    if (g_tuned[dev].count(n)) continue;
    rgf::SweepArgs a = jobs[kv.second];
    a.nx = std::min(a.nx, 12);
    a.nE = 1;
    a.dec_nmax = 2;
    a.ldos = a.ndens = a.pdens = a.GRi1 = a.GLi1 = a.GGi1 = nullptr;
    a.cur = a.tr = a.tre = a.dens_n = a.dens_p = nullptr;
    a.info = nullptr;
    const long long se = (long long)(2 * a.nx + rgf::kTmpBlocks) * n * n;
    const int cs = std::max(1, std::min(slots, (int)std::min<long long>(slots, (long long)(8LL << 30) / (se * 16))));
    Buf ws, desc;
    if (int rc = ws.ensure((size_t)(se * 16 * cs))) return rc;
    std::vector<rgf::SweepArgs> one(1, a);
    std::vector<int2> items((size_t)cs, make_int2(0, 0));
    const MultiDesc L = multi_layout(1, cs);
    if (int rc = desc.ensure(L.bytes)) return rc;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    Tuned best{default_mv(n), 1e30};
    for (int v : candidate_mv(n)) {
      std::vector<int> mv(1, v);
      if (int rc = multi_upload(one, mv, items, desc.p)) return rc;
      if (int rc = multi_launch(1, cs, desc.p, reinterpret_cast<double2*>(ws.p), se, cs, s)) return rc;  // warm-up
      HIPCHK(hipEventRecord(e0, s));
      if (int rc = multi_launch(1, cs, desc.p, reinterpret_cast<double2*>(ws.p), se, cs, s)) return rc;
      HIPCHK(hipEventRecord(e1, s));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      // per slab-energy: 12 slabs plus 2 x 2 decimation iterations (~7/17 slab each)
      const double t = ms * 1e-3 / (a.nx + 4.0 * 7.0 / 17.0);
      if (t < best.t_slab) best = Tuned{v, t};
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ws.release();
    desc.release();
    g_tuned[dev][n] = best;
    ++*ntuned;
  }
  return 0;
}

// Plan: variants, work list (longest first), workspace per slot, slot count.
struct MultiPlan {
  std::vector<rgf::SweepArgs> jobs;
  std::vector<int> mv;
  std::vector<int2> items;
  long long ws_slot = 0;
  int slots = 0;
  int dev = 0;
  int ntuned = 0;
};

int multi_plan(const rgf_params* p, const rgf_batch* devs, int ndev, int autotune, int max_slots, hipStream_t s,
               MultiPlan* pl, size_t reserved) {
  if (int rc = multi_jobs(p, devs, ndev, &pl->jobs)) return rc;
  HIPCHK(hipGetDevice(&pl->dev));
  const int dev = pl->dev;
  if (dev < 0 || dev >= kMaxDev) return fail(RGF_EINVAL, "device index out of range");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  int per_cu = 0;
  if (multi_per_cu(&per_cu) || per_cu < 1) return fail(RGF_EHIP, "multi-device kernel does not fit on a CU");
  const int full = prop.multiProcessorCount * per_cu;
  long long nitems = 0, se_max = 0;
  for (auto& a : pl->jobs) {
    nitems += a.nE;
    if (a.nE > 0) se_max = std::max<long long>(se_max, (long long)(2 * a.nx + rgf::kTmpBlocks) * a.n * a.n);
  }
  if (autotune) {
    if (int rc = multi_autotune(pl->jobs, dev, full, s, &pl->ntuned)) return rc;
  }
  pl->mv.resize(ndev);
  std::vector<double> key(ndev);
  for (int j = 0; j < ndev; ++j) {
    const int n = pl->jobs[j].n;
    auto it = g_tuned[dev].find(n);
    const bool tuned = autotune && it != g_tuned[dev].end();
    pl->mv[j] = tuned ? it->second.mv : default_mv(n);
    // predicted seconds per energy: slabs plus ~36 decimation iterations at 7/17 slab each
    const double per_slab = tuned ? it->second.t_slab : (double)n * n * n;
    key[j] = per_slab * (pl->jobs[j].nx + 36.0 * 7.0 / 17.0);
  }
  pl->items.clear();
  pl->items.reserve((size_t)nitems);
  for (int j = 0; j < ndev; ++j)
    for (int e = 0; e < pl->jobs[j].nE; ++e) pl->items.push_back(make_int2(j, e));
  std::stable_sort(pl->items.begin(), pl->items.end(),
                   [&](const int2& x, const int2& y) { return key[x.x] > key[y.x]; });
  pl->ws_slot = se_max;
  long long slots = std::min<long long>(nitems, full);
  if (max_slots > 0) slots = std::min<long long>(slots, max_slots);
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  const size_t avail = (size_t)((double)(fr + reserved) * 0.9);
  if (se_max > 0) {
    const long long fit = (long long)(avail / (size_t)(se_max * 16));
    if (fit < 1) return fail(RGF_ENOMEM, "not enough device memory for one slot of the largest device");
    slots = std::min(slots, fit);
  }
  pl->slots = (int)std::max<long long>(slots, 1);
  return 0;
}
}  // namespace

extern "C" {

void rgf_params_reference(rgf_params* p) {
  p->eta_dev = 0.0;
  p->eta_lead = 1.0e-3;
  p->dec_tol = 1.0e-100;
  p->dec_nmax = 100;
  p->boltz = 8.61734e-05;
  p->mode = RGF_MODE_COMPAT;
}

const char* rgf_last_error(void) { return g_err.c_str(); }

const char* rgf_version(void) { return "liteqt_amd rgf 0.1 (gfx950, complex FP64 MFMA)"; }

long long rgf_workspace_bytes(int nm, int nx, int slots) {
  return slot_elems(nm, nx) * 16LL * (long long)slots;
}

int rgf_solve_batched(const rgf_params* p, const rgf_batch* b, void* stream) {
  return solve_impl(p, b, nullptr, stream);
}

int rgf_solve_batched_noalloc(const rgf_params* p, const rgf_batch* b, void* stream) {
  return solve_noalloc_impl(p, b, stream);
}

int rgf_scba_step(const rgf_params* p, const rgf_batch* b, const rgf_scba* scba, void* stream) {
  if (!scba) {
    g_err.clear();
    return fail(RGF_EINVAL, "rgf_scba_step: scba is NULL");
  }
  return solve_impl(p, b, scba, stream);
}

struct rgf_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void* ws = nullptr;
  void* desc = nullptr;
  int dev = 0;
};

int rgf_graph_create(const rgf_params* p, const rgf_batch* b, const rgf_scba* scba, void* stream,
                     rgf_graph* out) {
  g_err.clear();
  if (!out) return fail(RGF_EINVAL, "rgf_graph_create: out is NULL");
  *out = nullptr;
  if (!stream) return fail(RGF_EINVAL, "rgf_graph_create: capture needs a created stream (not the NULL stream)");
  Prepared pr;
  if (int rc = prepare(p, b, scba, &pr)) return rc;
  if (pr.nothing) return fail(RGF_EINVAL, "rgf_graph_create: nothing to solve (nE == 0)");
  // the graph owns its workspace, so later calls that grow the shared one cannot move it
  rgf_graph_s* g = new rgf_graph_s();
  g->dev = pr.dev;
  hipError_t e = hipMalloc(&g->ws, (size_t)(pr.slots * pr.slot_elems * 16LL));
  if (e != hipSuccess) {
    delete g;
    return fail(e == hipErrorOutOfMemory ? RGF_ENOMEM : RGF_EHIP, std::string("hipMalloc graph workspace: ") +
                                                                     hipGetErrorString(e));
  }
  pr.a.ws = reinterpret_cast<double2*>(g->ws);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    pr.v.launch(pr.a, (int)pr.slots, s);
    hipError_t le = hipGetLastError();
    e = hipStreamEndCapture(s, &g->graph);
    if (e == hipSuccess && le != hipSuccess) e = le;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    rgf_graph_destroy(g);
    return fail(RGF_EHIP, std::string("rgf_graph_create: ") + hipGetErrorString(e));
  }
  *out = g;
  return 0;
}

int rgf_graph_launch(rgf_graph g, int replays, void* stream) {
  g_err.clear();
  if (!g || !g->exec) return fail(RGF_EINVAL, "rgf_graph_launch: invalid graph");
  if (replays < 0) return fail(RGF_EINVAL, "rgf_graph_launch: replays < 0");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  for (int i = 0; i < replays; ++i) HIPCHK(hipGraphLaunch(g->exec, s));
  return 0;
}

void rgf_graph_destroy(rgf_graph g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->ws) (void)hipFree(g->ws);
  if (g->desc) (void)hipFree(g->desc);
  delete g;
}

int rgf_solve_devices(const rgf_params* p, const rgf_batch* devs, int ndev, int autotune, int* tile_T,
                      rgf_multi_report* rep, void* stream) {
  g_err.clear();
  {
    std::vector<rgf::SweepArgs> check;  // argument errors before any device call
    if (int rc = multi_jobs(p, devs, ndev, &check)) return rc;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g_multi_mu);
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDev) return fail(RGF_EINVAL, "device index out of range");
  // the descriptor buffer and workspace are reused: let earlier launches finish
  HIPCHK(hipDeviceSynchronize());
  std::lock_guard<std::mutex> lw(g_ws_mu);
  Buf& ws = g_ws[dev];
  MultiPlan pl;
  if (int rc = multi_plan(p, devs, ndev, autotune, 0, s, &pl, ws.bytes)) return rc;
  if (tile_T)
    for (int j = 0; j < ndev; ++j) tile_T[j] = mv_T(pl.mv[j]);
  if (rep) {
    rep->slots = pl.slots;
    rep->items = (long long)pl.items.size();
    rep->ws_bytes = pl.ws_slot * 16LL * pl.slots;
    rep->tuned = pl.ntuned;
  }
  if (pl.items.empty()) return 0;
  if ((size_t)(pl.ws_slot * 16LL * pl.slots) > ws.bytes) {
    ws.release();
    if (int rc = ws.ensure((size_t)(pl.ws_slot * 16LL * pl.slots))) return rc;
  }
  Buf& desc = g_multi_desc[dev];
  const MultiDesc L = multi_layout(ndev, (long long)pl.items.size());
  if (int rc = desc.ensure(L.bytes)) return rc;
  if (int rc = multi_upload(pl.jobs, pl.mv, pl.items, desc.p)) return rc;
  return multi_launch(ndev, (long long)pl.items.size(), desc.p, reinterpret_cast<double2*>(ws.p), pl.ws_slot,
                      pl.slots, s);
}

int rgf_graph_create_devices(const rgf_params* p, const rgf_batch* devs, int ndev, int autotune, void* stream,
                             rgf_graph* out) {
  g_err.clear();
  if (!out) return fail(RGF_EINVAL, "rgf_graph_create_devices: out is NULL");
  *out = nullptr;
  if (!stream) return fail(RGF_EINVAL, "rgf_graph_create_devices: capture needs a created stream (not the NULL stream)");
  {
    std::vector<rgf::SweepArgs> check;
    if (int rc = multi_jobs(p, devs, ndev, &check)) return rc;
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g_multi_mu);
  MultiPlan pl;
  if (int rc = multi_plan(p, devs, ndev, autotune, 0, s, &pl, 0)) return rc;
  if (pl.items.empty()) return fail(RGF_EINVAL, "rgf_graph_create_devices: nothing to solve");
  rgf_graph_s* g = new rgf_graph_s();
  g->dev = pl.dev;
  const MultiDesc L = multi_layout(ndev, (long long)pl.items.size());
  hipError_t e = hipMalloc(&g->ws, (size_t)(pl.ws_slot * 16LL * pl.slots));
  if (e == hipSuccess) e = hipMalloc(&g->desc, L.bytes);
  if (e != hipSuccess) {
    rgf_graph_destroy(g);
    return fail(e == hipErrorOutOfMemory ? RGF_ENOMEM : RGF_EHIP, std::string("hipMalloc graph buffers: ") +
                                                                     hipGetErrorString(e));
  }
  // upload outside the capture, then capture counter reset + persistent launch
  if (int rc = multi_upload(pl.jobs, pl.mv, pl.items, g->desc)) {
    rgf_graph_destroy(g);
    return rc;
  }
  e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    int rc = multi_launch(ndev, (long long)pl.items.size(), g->desc, reinterpret_cast<double2*>(g->ws), pl.ws_slot,
                          pl.slots, s);
    e = hipStreamEndCapture(s, &g->graph);
    if (rc && e == hipSuccess) e = hipErrorLaunchFailure;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    rgf_graph_destroy(g);
    return fail(RGF_EHIP, std::string("rgf_graph_create_devices: ") + hipGetErrorString(e));
  }
  *out = g;
  return 0;
}

// Host-array solve of nE energies: stage (and pad) inputs in HBM, run the
// batched engine chunk by chunk, copy (and unpad) the outputs back.
static int rs_energies_impl(const rgf_params* p, const double temp[2], int nm, int nx, int nE, const double* E,
                            const double mu[2], const double* Sii, const double* Hii, const double* H1i,
                            const double* sigL, const double* sigG, const double* sigR, int sig_per_energy,
                            double* ndens, double* pdens, double* ldos, double* tr, double* tre, double* cur,
                            double* GRi1, double* GLi1, double* GGi1, int* info);

int rgf_rs_energies(const rgf_params* p, const double temp[2], int nm, int nx, int nE, const double* E,
                    const double mu[2], const double* Hii, const double* H1i, const double* sigL,
                    const double* sigG, const double* sigR, int sig_per_energy, double* ndens,
                    double* pdens, double* ldos, double* tr, double* tre, double* cur, double* GRi1,
                    double* GLi1, double* GGi1, int* info) {
  return rs_energies_impl(p, temp, nm, nx, nE, E, mu, nullptr, Hii, H1i, sigL, sigG, sigR, sig_per_energy, ndens,
                          pdens, ldos, tr, tre, cur, GRi1, GLi1, GGi1, info);
}

int rgf_cms_energies(const rgf_params* p, const double temp[2], int nm, int nx, int nE, const double* E,
                     const double mu[2], const double* Sii, const double* Hii, const double* H1i,
                     const double* sigL, const double* sigG, const double* sigR, int sig_per_energy,
                     double* ndens, double* pdens, double* ldos, double* tr, double* tre, double* cur,
                     double* GRi1, double* GLi1, double* GGi1, int* info) {
  g_err.clear();
  if (!Sii) return fail(RGF_EINVAL, "rgf_cms_energies: Sii is required");
  return rs_energies_impl(p, temp, nm, nx, nE, E, mu, Sii, Hii, H1i, sigL, sigG, sigR, sig_per_energy, ndens,
                          pdens, ldos, tr, tre, cur, GRi1, GLi1, GGi1, info);
}

int rgf_cms_ex(const rgf_params* p, const double temp[2], int nm, int nx, double E, const double mu[2],
               const double* Sii, const double* Hii, const double* H1i, const double* sigL, const double* sigG,
               const double* sigR, double* ndens, double* pdens, double* ldos, double* tr, double* tre, double* cur,
               double* GRi1, double* GLi1, double* GGi1) {
  return rgf_cms_energies(p, temp, nm, nx, 1, &E, mu, Sii, Hii, H1i, sigL, sigG, sigR, 0, ndens, pdens, ldos, tr,
                          tre, cur, GRi1, GLi1, GGi1, nullptr);
}

int rgf_cms(const double temp[2], int nm, int nx, double E, const double mu[2], const double* Sii,
            const double* Hii, const double* H1i, const double* sigL, const double* sigG, const double* sigR,
            double* ndens, double* pdens, double* ldos, double* tr, double* tre, double* cur, double* GRi1,
            double* GLi1, double* GGi1) {
  rgf_params p;
  rgf_params_reference(&p);
  return rgf_cms_ex(&p, temp, nm, nx, E, mu, Sii, Hii, H1i, sigL, sigG, sigR, ndens, pdens, ldos, tr, tre, cur,
                    GRi1, GLi1, GGi1);
}

}  // extern "C"

namespace {

// A per-energy host array of a host-array solve: energy e at base + e * stride (the
// drivers' (nm, nm, nx, nen) arrays), or at list[e] (separate per-energy calls
// gathered into one batch).
struct EArr {
  const char* base = nullptr;
  size_t stride = 0;
  std::vector<const char*> list;
  bool present() const { return base || !list.empty(); }
  bool strided() const { return list.empty(); }
  char* at(long long e) const { return const_cast<char*>(strided() ? base + stride * (size_t)e : list[(size_t)e]); }
};

inline EArr arr1(const void* p) {
  EArr a;
  a.base = (const char*)p;
  return a;
}

// The host arrays of one solve.  sigL/sigR either per energy or one shared slab set.
struct HostIO {
  int nE = 0;
  const double* E = nullptr;
  EArr sigL, sigR;
  bool sig_per_energy = false;
  EArr out[6];  // ldos, ndens, pdens, GRi1, GLi1, GGi1
  EArr cur, tr, tre, info;
};

}  // namespace

extern "C" {

static int rs_core(const rgf_params* p, const double temp[2], int nm, int nx, const double mu[2], const double* Sii,
                   const double* Hii, const double* H1i, const HostIO& io);

static int rs_energies_impl(const rgf_params* p, const double temp[2], int nm, int nx, int nE, const double* E,
                            const double mu[2], const double* Sii, const double* Hii, const double* H1i,
                            const double* sigL, const double* sigG, const double* sigR, int sig_per_energy,
                            double* ndens, double* pdens, double* ldos, double* tr, double* tre, double* cur,
                            double* GRi1, double* GLi1, double* GGi1, int* info) {
  (void)sigG;  // accepted and unused, as green_RGF_RS (src/green_rgf.f90:3953)
  g_err.clear();
  if (int rc = check_params(p)) return rc;
  if (nm < 1 || nx < 2 || nE < 0) return fail(RGF_EINVAL, "need nm >= 1, nx >= 2, nE >= 0");
  if (!temp || !mu || !Hii || !H1i || (nE > 0 && !E)) return fail(RGF_EINVAL, "missing required array");
  if (nE == 0) return 0;
  const size_t cslab_u = (size_t)nm * nm * nx * 16, rslab_u = (size_t)nm * nm * nx * 8;
  HostIO io;
  io.nE = nE;
  io.E = E;
  io.sig_per_energy = sig_per_energy != 0;
  auto arr = [](const void* b, size_t stride) {
    EArr a;
    a.base = (const char*)b;
    a.stride = stride;
    return a;
  };
  io.sigL = arr(sigL, sig_per_energy ? cslab_u : 0);
  io.sigR = arr(sigR, sig_per_energy ? cslab_u : 0);
  double* outs[6] = {ldos, ndens, pdens, GRi1, GLi1, GGi1};
  for (int k = 0; k < 6; ++k) io.out[k] = arr(outs[k], cslab_u);
  io.cur = arr(cur, rslab_u);
  io.tr = arr(tr, 8);
  io.tre = arr(tre, 8);
  io.info = arr(info, 16);
  return rs_core(p, temp, nm, nx, mu, Sii, Hii, H1i, io);
}

// Stage the inputs in HBM (padded to the engine's block granule), run the batched
// engine chunk by chunk over the energies, copy (and unpad) the outputs back.
static int rs_core(const rgf_params* p, const double temp[2], int nm, int nx, const double mu[2], const double* Sii,
                   const double* Hii, const double* H1i, const HostIO& io) {
  const int nE = io.nE;
  // blocks are padded to the engine's granule: 32 up to 256, 64 above (overlaps: up to 256)
  const int np = nm <= 256 ? (nm + 31) / 32 * 32 : (nm + 63) / 64 * 64;
  if (np > (Sii ? 256 : kNmaxBig))
    return fail(RGF_EUNSUP, Sii ? "the overlap solve supports nm <= 256" : "nm > 512 is not supported");
  std::lock_guard<std::mutex> lk(g_mu);
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDev) return fail(RGF_EINVAL, "device index out of range");
  const size_t bn = (size_t)nm * nm, bp = (size_t)np * np;
  const size_t cslab = bp * nx * 16, rslab = bp * nx * 8;       // padded, per energy
  const size_t cslab_u = bn * nx * 16, rslab_u = bn * nx * 8;   // caller layout, per energy
  const bool padded = (np != nm);
  const bool sig_pe = io.sig_per_energy;
  int nout = 0;
  for (int k = 0; k < 6; ++k) nout += io.out[k].present();
  const bool has_cur = io.cur.present();
  const int nsig = (io.sigL.present() ? 1 : 0) + (io.sigR.present() ? 1 : 0);
  // bytes per energy in a chunk, then chunk size from free memory
  const size_t per_e = (size_t)nout * cslab + (has_cur ? rslab : 0) + (sig_pe ? nsig * cslab : 0) +
                       (padded ? cslab_u : 0) + 64;
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  const size_t fixed = (Sii ? 3 : 2) * cslab + (sig_pe ? 0 : nsig * cslab) + (padded ? cslab_u : 0);
  const size_t ws_reserve = (size_t)rgf_workspace_bytes(np, nx, 256);
  const size_t budget = (size_t)((double)(fr + g_stage[dev].bytes) * 0.45);
  long long C = nE;
  if (budget > fixed + ws_reserve / 2) C = std::min<long long>(C, (long long)((budget - fixed - ws_reserve / 2) / per_e));
  else C = 1;
  C = std::max<long long>(1, std::min<long long>(C, 8192));
  // staging layout
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) / 256 * 256; return o; };
  const size_t oH = take(cslab), oC = take(cslab), oS = Sii ? take(cslab) : 0;
  const size_t sig_n = sig_pe ? (size_t)C : 1;
  const size_t oSL = io.sigL.present() ? take(cslab * sig_n) : 0, oSR = io.sigR.present() ? take(cslab * sig_n) : 0;
  size_t oOut[6];
  for (int k = 0; k < 6; ++k) oOut[k] = io.out[k].present() ? take(cslab * C) : 0;
  const size_t oCur = has_cur ? take(rslab * C) : 0;
  const size_t oE = take(8 * C), oTr = take(8 * C), oTre = take(8 * C), oInfo = take(16 * C);
  const size_t oU = padded ? take(cslab_u * std::max<size_t>(C, 1)) : 0;
  if (int rc = g_stage[dev].ensure(off)) return rc;
  char* base = (char*)g_stage[dev].p;
  auto D = [&](size_t o) { return (void*)(base + o); };
  hipStream_t s = nullptr;
  const double2 fill = make_double2(1.0e3, 0.0);
  // (count) energies' (nm, nm, nx) complex slabs, energy e0 + i of `a`, into a padded
  // device area; `diag` fills the padded diagonal
  auto h2d = [&](const EArr& a, long long e0, long long count, size_t dst_off, double2 diag) -> int {
    char* stage = padded ? (char*)D(oU) : (char*)D(dst_off);
    if (a.strided()) {
      HIPCHK(hipMemcpyAsync(stage, a.at(e0), cslab_u * count, hipMemcpyHostToDevice, s));
    } else {
      for (long long i = 0; i < count; ++i)
        HIPCHK(hipMemcpyAsync(stage + cslab_u * i, a.at(e0 + i), cslab_u, hipMemcpyHostToDevice, s));
    }
    if (padded) {
      hipLaunchKernelGGL(pad_blocks, dim3(1024), dim3(256), 0, s, (const double2*)D(oU), nm, (double2*)D(dst_off),
                         np, (long long)(nx * count), diag);
      HIPCHK(hipGetLastError());
    }
    return 0;
  };
  if (int rc = h2d(arr1(Hii), 0, 1, oH, fill)) return rc;
  // couplings pad with zeros (no diagonal fill)
  if (int rc = h2d(arr1(H1i), 0, 1, oC, make_double2(0, 0))) return rc;
  // overlaps pad with the identity (padded orbitals stay decoupled: z - 1e3 on their diagonal)
  if (Sii)
    if (int rc = h2d(arr1(Sii), 0, 1, oS, make_double2(1, 0))) return rc;
  if (!sig_pe) {
    if (io.sigL.present()) if (int rc = h2d(io.sigL, 0, 1, oSL, make_double2(0, 0))) return rc;
    if (io.sigR.present()) if (int rc = h2d(io.sigR, 0, 1, oSR, make_double2(0, 0))) return rc;
  }
  std::vector<double> trh((size_t)C), treh((size_t)C);
  std::vector<int> infoh((size_t)(4 * C));
  for (long long e0 = 0; e0 < nE; e0 += C) {
    const long long ce = std::min<long long>(C, nE - e0);
    if (sig_pe) {
      if (io.sigL.present()) if (int rc = h2d(io.sigL, e0, ce, oSL, make_double2(0, 0))) return rc;
      if (io.sigR.present()) if (int rc = h2d(io.sigR, e0, ce, oSR, make_double2(0, 0))) return rc;
    }
    HIPCHK(hipMemcpyAsync(D(oE), io.E + e0, 8 * ce, hipMemcpyHostToDevice, s));
    rgf_batch b;
    std::memset(&b, 0, sizeof(b));
    b.nm = np; b.nx = nx; b.nE = (int)ce;
    b.E = (const double*)D(oE);
    b.temp[0] = temp[0]; b.temp[1] = temp[1]; b.mu[0] = mu[0]; b.mu[1] = mu[1];
    b.Hii = (const double*)D(oH); b.H1i = (const double*)D(oC);
    b.Sii = Sii ? (const double*)D(oS) : nullptr;
    b.sigL = io.sigL.present() ? (const double*)D(oSL) : nullptr;
    b.sigR = io.sigR.present() ? (const double*)D(oSR) : nullptr;
    b.sig_stride = sig_pe ? (long long)(bp * nx) : 0;
    double** bo[6] = {&b.ldos, &b.ndens, &b.pdens, &b.GRi1, &b.GLi1, &b.GGi1};
    for (int k = 0; k < 6; ++k) *bo[k] = io.out[k].present() ? (double*)D(oOut[k]) : nullptr;
    b.cur = has_cur ? (double*)D(oCur) : nullptr;
    b.tr = (double*)D(oTr); b.tre = (double*)D(oTre);
    b.info = (int*)D(oInfo);
    if (int rc = rgf_solve_batched(p, &b, s)) return rc;
    // outputs: diagonal blocks for l = 1..nx, off-diagonal and cur for l = 1..nx-1
    auto d2h = [&](const EArr& a, const void* src, size_t slab, size_t bytes) -> int {
      if (a.strided()) {
        HIPCHK(hipMemcpy2DAsync(a.at(e0), slab, src, slab, bytes, ce, hipMemcpyDeviceToHost, s));
      } else {
        for (long long i = 0; i < ce; ++i)
          HIPCHK(hipMemcpyAsync(a.at(e0 + i), (const char*)src + slab * i, bytes, hipMemcpyDeviceToHost, s));
      }
      return 0;
    };
    for (int k = 0; k < 6; ++k) {
      if (!io.out[k].present()) continue;
      const size_t nb_copy = (k >= 3) ? (size_t)(nx - 1) : (size_t)nx;
      const void* src = D(oOut[k]);
      if (padded) {
        hipLaunchKernelGGL(unpad_blocks<double2>, dim3(1024), dim3(256), 0, s, (const double2*)src, np,
                           (double2*)D(oU), nm, (long long)(nx * ce));
        HIPCHK(hipGetLastError());
        src = D(oU);
      }
      if (int rc = d2h(io.out[k], src, cslab_u, nb_copy * bn * 16)) return rc;
    }
    if (has_cur) {
      const void* src = D(oCur);
      if (padded) {
        hipLaunchKernelGGL(unpad_blocks<double>, dim3(1024), dim3(256), 0, s, (const double*)src, np,
                           (double*)D(oU), nm, (long long)(nx * ce));
        HIPCHK(hipGetLastError());
        src = D(oU);
      }
      if (int rc = d2h(io.cur, src, rslab_u, (size_t)(nx - 1) * bn * 8)) return rc;
    }
    HIPCHK(hipMemcpyAsync(trh.data(), D(oTr), 8 * ce, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(treh.data(), D(oTre), 8 * ce, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(infoh.data(), D(oInfo), 16 * ce, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (long long i = 0; i < ce; ++i) {
      if (io.tr.present()) *(double*)io.tr.at(e0 + i) = trh[(size_t)i];
      if (io.tre.present()) *(double*)io.tre.at(e0 + i) = treh[(size_t)i];
      if (io.info.present()) std::memcpy(io.info.at(e0 + i), &infoh[(size_t)(4 * i)], 16);
      // the reference zeroes cur on entry and never writes block nx
      if (has_cur) std::memset(io.cur.at(e0 + i) + bn * 8 * (nx - 1), 0, bn * 8);
    }
  }
  return 0;
}

// ---------------------------------------------- concurrent per-energy callers
// The reference's drivers call green_RGF_RS from every OpenMP thread of an energy
// loop (src/green_rgf.f90:1305-1316).  Concurrent rgf_rs / rgf_rs_ex calls are
// gathered into one batched launch: a caller that finds no batch being formed
// becomes the leader, waits while further callers keep arriving (a gap of 200 us
// ends the wait; at most 5 ms, and no wait once as many callers are queued as the
// last batch had), takes every queued call that shares its device, nm, nx,
// parameters, temperatures, chemical potentials, Hii / H1i arrays and the set of
// given arrays, and solves them as one energy batch; the others sleep until their
// call is done and return its status.  Results are those of a separate call.
struct RsCall {
  rgf_params p;
  double temp[2], mu[2], E;
  int nm, nx, dev;
  const double *Hii, *H1i, *sigL, *sigR;
  double* out[6];  // ldos, ndens, pdens, GRi1, GLi1, GGi1
  double *cur, *tr, *tre;
  int rc = 0;
  bool done = false;
  std::string err;
};

std::mutex g_rs_mu;
std::condition_variable g_rs_cv;
std::vector<RsCall*> g_rs_queue;
bool g_rs_forming = false;
size_t g_rs_last_batch = 1;

bool rs_same_batch(const RsCall& a, const RsCall& b) {
  if (a.dev != b.dev || a.nm != b.nm || a.nx != b.nx || a.Hii != b.Hii || a.H1i != b.H1i) return false;
  if (std::memcmp(&a.p, &b.p, sizeof(rgf_params)) != 0) return false;
  if (a.temp[0] != b.temp[0] || a.temp[1] != b.temp[1] || a.mu[0] != b.mu[0] || a.mu[1] != b.mu[1]) return false;
  if (!a.sigL != !b.sigL || !a.sigR != !b.sigR || !a.cur != !b.cur || !a.tr != !b.tr || !a.tre != !b.tre) return false;
  for (int k = 0; k < 6; ++k)
    if (!a.out[k] != !b.out[k]) return false;
  return true;
}

int rs_solve_calls(const std::vector<RsCall*>& v) {
  const RsCall& h = *v[0];
  const int n = (int)v.size();
  std::vector<double> E(n);
  HostIO io;
  io.nE = n;
  auto list = [&](auto get) {
    EArr a;
    if (get(h)) {
      a.list.resize(n);
      for (int i = 0; i < n; ++i) a.list[i] = (const char*)get(*v[i]);
    }
    return a;
  };
  for (int i = 0; i < n; ++i) E[i] = v[i]->E;
  io.E = E.data();
  // one Sigma array passed by every call (a ballistic or energy-independent
  // scattering self-energy) is staged once
  bool shared = true;
  for (int i = 1; i < n; ++i) shared = shared && v[i]->sigL == h.sigL && v[i]->sigR == h.sigR;
  io.sig_per_energy = !shared;
  if (shared) {
    io.sigL = arr1(h.sigL);
    io.sigR = arr1(h.sigR);
  } else {
    io.sigL = list([](const RsCall& c) { return (const void*)c.sigL; });
    io.sigR = list([](const RsCall& c) { return (const void*)c.sigR; });
  }
  for (int k = 0; k < 6; ++k) io.out[k] = list([k](const RsCall& c) { return (const void*)c.out[k]; });
  io.cur = list([](const RsCall& c) { return (const void*)c.cur; });
  io.tr = list([](const RsCall& c) { return (const void*)c.tr; });
  io.tre = list([](const RsCall& c) { return (const void*)c.tre; });
  return rs_core(&h.p, h.temp, h.nm, h.nx, h.mu, nullptr, h.Hii, h.H1i, io);
}

int rgf_rs_ex(const rgf_params* p, const double temp[2], int nm, int nx, double E, const double mu[2],
              const double* Hii, const double* H1i, const double* sigL, const double* sigG, const double* sigR,
              double* ndens, double* pdens, double* ldos, double* tr, double* tre, double* cur, double* GRi1,
              double* GLi1, double* GGi1) {
  (void)sigG;  // accepted and unused, as green_RGF_RS (src/green_rgf.f90:3953)
  g_err.clear();
  if (int rc = check_params(p)) return rc;
  if (nm < 1 || nx < 2) return fail(RGF_EINVAL, "need nm >= 1, nx >= 2");
  if (!temp || !mu || !Hii || !H1i) return fail(RGF_EINVAL, "missing required array");
  RsCall c;
  c.p = *p;
  c.temp[0] = temp[0]; c.temp[1] = temp[1]; c.mu[0] = mu[0]; c.mu[1] = mu[1];
  c.E = E; c.nm = nm; c.nx = nx;
  HIPCHK(hipGetDevice(&c.dev));
  c.Hii = Hii; c.H1i = H1i; c.sigL = sigL; c.sigR = sigR;
  double* outs[6] = {ldos, ndens, pdens, GRi1, GLi1, GGi1};
  for (int k = 0; k < 6; ++k) c.out[k] = outs[k];
  c.cur = cur; c.tr = tr; c.tre = tre;
  std::unique_lock<std::mutex> lk(g_rs_mu);
  g_rs_queue.push_back(&c);
  g_rs_cv.notify_all();
  while (!c.done) {
    if (g_rs_forming) {
      g_rs_cv.wait(lk);
      continue;
    }
    g_rs_forming = true;
    // gather the callers that are arriving: wait while they keep coming (a 200 us
    // gap ends it), and up to the size of the previous batch (5 ms at most)
    const auto t0 = std::chrono::steady_clock::now();
    size_t seen = g_rs_queue.size();
    while (std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(5)) {
      g_rs_cv.wait_for(lk, std::chrono::microseconds(200));
      const bool grew = g_rs_queue.size() > seen;
      seen = g_rs_queue.size();
      if (!grew && seen >= g_rs_last_batch) break;
    }
    std::vector<RsCall*> batch;
    RsCall* head = g_rs_queue.front();
    for (size_t i = 0; i < g_rs_queue.size();) {
      if (rs_same_batch(*head, *g_rs_queue[i])) {
        batch.push_back(g_rs_queue[i]);
        g_rs_queue.erase(g_rs_queue.begin() + (long)i);
      } else {
        ++i;
      }
    }
    lk.unlock();
    const int rc = rs_solve_calls(batch);
    const std::string msg = rc ? g_err : std::string();
    lk.lock();
    for (RsCall* b : batch) {
      b->rc = rc;
      b->err = msg;
      b->done = true;
    }
    g_rs_last_batch = std::max<size_t>(1, batch.size());
    g_rs_forming = false;
    g_rs_cv.notify_all();
  }
  lk.unlock();
  g_err = c.err;
  return c.rc;
}

int rgf_rs(const double temp[2], int nm, int nx, double E, const double mu[2], const double* Hii,
           const double* H1i, const double* sigL, const double* sigG, const double* sigR, double* ndens,
           double* pdens, double* ldos, double* tr, double* tre, double* cur, double* GRi1, double* GLi1,
           double* GGi1) {
  rgf_params p;
  rgf_params_reference(&p);
  return rgf_rs_ex(&p, temp, nm, nx, E, mu, Hii, H1i, sigL, sigG, sigR, ndens, pdens, ldos, tr, tre, cur, GRi1,
                   GLi1, GGi1);
}

long long rgf_adjoint_workspace_bytes(int nm, int nx, int slots) {
  return rgf::adj_slot_blocks(nx, 0) * nm * nm * 16LL * slots + (long long)nx * nm * nm * 16LL * slots;
}

int rgf_adjoint_batched(const rgf_params* p, const rgf_batch* b, const double* w_tr, const double* w_n,
                        const double* lead_H00, double* J, double* dU, double* dH1i, void* stream) {
  g_err.clear();
  if (int rc = check_params(p)) return rc;
  if (p->mode != RGF_MODE_PHYSICAL) return fail(RGF_EINVAL, "the adjoint differentiates the physical mode");
  if (b && b->Sii) return fail(RGF_EUNSUP, "the adjoint does not take overlaps");
  if (!b || !b->E || !b->Hii || !b->H1i || !J || !dU) return fail(RGF_EINVAL, "missing arguments");
  if (b->nx < 2 || b->nE < 0) return fail(RGF_EINVAL, "need nx >= 2, nE >= 0");
  if (b->ckpt < -1 || b->ckpt == 1)
    return fail(RGF_EINVAL, "ckpt must be -1 (keep every slab), 0 (automatic) or a segment length >= 2");
  if (b->nE == 0) return 0;
  AdjVariant v;
  if (!pick_adj(b->nm, &v)) return fail(RGF_EUNSUP, "rgf_adjoint_batched: nm must be a multiple of 32 in [32, 256]");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDev) return fail(RGF_EINVAL, "device index out of range");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  int per_cu = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, v.fn, v.nth, 0));
  if (per_cu < 1) return fail(RGF_EHIP, "adjoint kernel does not fit on a CU");
  const long long nn = (long long)b->nm * b->nm;
  const long long dh_bytes = (long long)b->nx * nn * 16LL;
  long long slots = std::min<long long>(b->nE, (long long)prop.multiProcessorCount * per_cu);
  if (b->max_slots > 0) slots = std::min<long long>(slots, b->max_slots);
  // The larger slot (R, R d and g S kept per slab: 5 fewer GEMMs per slab) is used
  // when every slot still fits; otherwise the small slot, recomputing them.
  const long long keep_bytes = rgf::adj_slot_blocks(b->nx, 1) * nn * 16LL;
  const long long small_bytes = rgf::adj_slot_blocks(b->nx, 0) * nn * 16LL;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  Buf& ws = g_ws[dev];
  Buf& dh = g_adj_dh[dev];
  // LITEQT_ADJ_RECOMPUTE=1 forces the small slot (the path taken under memory pressure)
  const char* rc_env = std::getenv("LITEQT_ADJ_RECOMPUTE");
  int keep = (rc_env && rc_env[0] == '1') ? 0 : 1;
  // segment recompute of the forward state: forced (ckpt > 0), or chosen below when
  // even the small slot leaves CUs idle
  int K = b->ckpt > 0 ? b->ckpt : 0;
  if (K) keep = 0;
  long long slot_bytes = K ? rgf::adj_slot_blocks(b->nx, 0, K) * nn * 16LL : keep ? keep_bytes : small_bytes;
  if ((size_t)(slots * slot_bytes) > ws.bytes || (size_t)(slots * dh_bytes) > dh.bytes) {
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    const size_t avail = (size_t)((double)(fr + ws.bytes + dh.bytes) * 0.9);
    if (keep && (long long)(avail / (size_t)(keep_bytes + dh_bytes)) < slots) {
      keep = 0;
      slot_bytes = small_bytes;
    }
    long long fit = (long long)(avail / (size_t)(slot_bytes + dh_bytes));
    if (!K && b->ckpt == 0 && fit < slots) {
      const int Ka = std::max(2, (int)std::ceil(std::sqrt((double)b->nx)));
      const long long kb = rgf::adj_slot_blocks(b->nx, 0, Ka) * nn * 16LL;
      const long long fit_k = (long long)(avail / (size_t)(kb + dh_bytes));
      if (fit_k > fit) {
        K = Ka;
        slot_bytes = kb;
        fit = fit_k;
      }
    }
    if (fit < 1) return fail(RGF_ENOMEM, "not enough device memory for one adjoint slot");
    slots = std::min(slots, fit);
    if ((size_t)(slots * slot_bytes) > ws.bytes || (size_t)(slots * dh_bytes) > dh.bytes) {
      ws.release();
      dh.release();
      if (int rc = ws.ensure((size_t)(slots * slot_bytes))) return rc;
      if (int rc = dh.ensure((size_t)(slots * dh_bytes))) return rc;
    }
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  HIPCHK(hipMemsetAsync(dh.p, 0, (size_t)(slots * dh_bytes), s));
  rgf::AdjArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = b->nm; a.nx = b->nx; a.nE = b->nE; a.dec_nmax = p->dec_nmax;
  a.eta_dev = p->eta_dev; a.eta_lead = p->eta_lead; a.dec_tol = p->dec_tol; a.boltz = p->boltz;
  a.temp[0] = b->temp[0]; a.temp[1] = b->temp[1]; a.mu[0] = b->mu[0]; a.mu[1] = b->mu[1];
  a.E = b->E;
  a.Hii = reinterpret_cast<const double2*>(b->Hii);
  a.H1i = reinterpret_cast<const double2*>(b->H1i);
  a.sigL = reinterpret_cast<const double2*>(b->sigL);
  a.sigR = reinterpret_cast<const double2*>(b->sigR);
  a.sigR_lo = reinterpret_cast<const double2*>(b->sigR_lo);
  a.sigR_up = reinterpret_cast<const double2*>(b->sigR_up);
  a.sigL_up = reinterpret_cast<const double2*>(b->sigL_up);
  a.sig_stride = b->sig_stride;
  a.w_tr = w_tr; a.w_n = w_n; a.lead = reinterpret_cast<const double2*>(lead_H00); a.J = J; a.tr = b->tr; a.dU = dU;
  a.dH = reinterpret_cast<double2*>(dh.p);
  a.info = b->info;
  a.ws = reinterpret_cast<double2*>(ws.p);
  a.ws_slot = slot_bytes / 16LL;
  a.keep = keep;
  a.ckpt = K;
  v.launch(a, (int)slots, s);
  HIPCHK(hipGetLastError());
  if (dH1i) {
    hipLaunchKernelGGL(rgf::k_sum_slots, dim3(1024), dim3(256), 0, s, reinterpret_cast<const double2*>(dh.p),
                       (int)slots, b->nx, b->nm, lead_H00 ? 0 : 1, reinterpret_cast<double2*>(dH1i));
    HIPCHK(hipGetLastError());
  }
  return 0;
}

int rgf_zgemm_batched(int n, int nbatch, int opA, const double* A, int opB, const double* B, double* C,
                      int c_rowmajor, void* stream) {
  g_err.clear();
  if (nbatch < 0 || (opA != 0 && opA != 1) || (opB != 0 && opB != 1)) return fail(RGF_EINVAL, "bad arguments");
  if (nbatch == 0) return 0;
  if (!A || !B || !C) return fail(RGF_EINVAL, "NULL matrix");
  Variant v;
  if (!pick_variant(n, &v)) return fail(RGF_EUNSUP, "n must be a multiple of 32 in [32, 256] or of 64 in (256, 512]");
  const int grid = std::min(nbatch, 4096);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const double2* a = reinterpret_cast<const double2*>(A);
  const double2* b = reinterpret_cast<const double2*>(B);
  double2* c = reinterpret_cast<double2*>(C);
  switch (v.T) {
    case 1: hipLaunchKernelGGL(rgf::k_gemm_batched<1>, dim3(grid), dim3(rgf::Cfg<1>::NTH), 0, s, n, nbatch, opA, a, opB, b, c, c_rowmajor); break;
    case 2: hipLaunchKernelGGL(rgf::k_gemm_batched<2>, dim3(grid), dim3(rgf::Cfg<2>::NTH), 0, s, n, nbatch, opA, a, opB, b, c, c_rowmajor); break;
    case 3: hipLaunchKernelGGL(rgf::k_gemm_batched<3>, dim3(grid), dim3(rgf::Cfg<3>::NTH), 0, s, n, nbatch, opA, a, opB, b, c, c_rowmajor); break;
    default: hipLaunchKernelGGL(rgf::k_gemm_batched<4>, dim3(grid), dim3(rgf::Cfg<4>::NTH), 0, s, n, nbatch, opA, a, opB, b, c, c_rowmajor); break;
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int rgf_zinv_batched(int n, int nbatch, const double* A, double* Ainv, int* info, void* stream) {
  g_err.clear();
  if (nbatch < 0) return fail(RGF_EINVAL, "nbatch < 0");
  if (nbatch == 0) return 0;
  if (!A || !Ainv) return fail(RGF_EINVAL, "NULL matrix");
  Variant v;
  if (!pick_variant(n, &v)) return fail(RGF_EUNSUP, "n must be a multiple of 32 in [32, 256] or of 64 in (256, 512]");
  const int grid = std::min(nbatch, 1024);
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mu);
  // the row-major work buffers live in the engine workspace
  if (int rc = g_ws[dev].ensure((size_t)grid * n * n * 16)) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const double2* a = reinterpret_cast<const double2*>(A);
  double2* o = reinterpret_cast<double2*>(Ainv);
  double2* w = reinterpret_cast<double2*>(g_ws[dev].p);
  switch (n) {
    case 32: hipLaunchKernelGGL((rgf::k_inv_batched<1, 32>), dim3(grid), dim3(rgf::Cfg<1>::NTH), 0, s, n, nbatch, a, o, w, info); break;
    case 64: hipLaunchKernelGGL((rgf::k_inv_batched<2, 64>), dim3(grid), dim3(rgf::Cfg<2>::NTH), 0, s, n, nbatch, a, o, w, info); break;
    case 96: hipLaunchKernelGGL((rgf::k_inv_batched<3, 96>), dim3(grid), dim3(rgf::Cfg<3>::NTH), 0, s, n, nbatch, a, o, w, info); break;
    case 128: hipLaunchKernelGGL((rgf::k_inv_batched<4, 128>), dim3(grid), dim3(rgf::Cfg<4>::NTH), 0, s, n, nbatch, a, o, w, info); break;
    case 192: hipLaunchKernelGGL((rgf::k_inv_batched<3, 192>), dim3(grid), dim3(rgf::Cfg<3>::NTH), 0, s, n, nbatch, a, o, w, info); break;
    case 256: hipLaunchKernelGGL((rgf::k_inv_batched<4, 256>), dim3(grid), dim3(rgf::Cfg<4>::NTH), 0, s, n, nbatch, a, o, w, info); break;
    default:
      if (n <= 256)
        hipLaunchKernelGGL((rgf::k_inv_batched<1, 256>), dim3(grid), dim3(rgf::Cfg<1>::NTH), 0, s, n, nbatch, a, o, w, info);
      else if (v.T == 4)
        hipLaunchKernelGGL((rgf::k_inv_batched<4, kNmaxBig>), dim3(grid), dim3(rgf::Cfg<4>::NTH), 0, s, n, nbatch, a, o, w, info);
      else
        hipLaunchKernelGGL((rgf::k_inv_batched<2, kNmaxBig>), dim3(grid), dim3(rgf::Cfg<2>::NTH), 0, s, n, nbatch, a, o, w, info);
      break;
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int rgf_w_solve_batched(const rgf_w_batch* b, void* stream) {
  g_err.clear();
  if (!b) return fail(RGF_EINVAL, "rgf_w_solve_batched: batch is required");
  if (b->nm < 1 || b->nx < 2 || b->nE < 0) return fail(RGF_EINVAL, "rgf_w_solve_batched: need nm >= 1, nx >= 2, nE >= 0");
  if (b->nbc != 0)
    return fail(RGF_EUNSUP, "rgf_w_solve_batched: open-boundary blocks (NBC > 0) are not supported (every driver passes 0)");
  if (b->nm > kNmaxBig) return fail(RGF_EUNSUP, "rgf_w_solve_batched: nm > 512");
  if (!b->Vii || !b->V1i || !b->PL || !b->PG || !b->PR || !b->WL || !b->WG || !b->WR)
    return fail(RGF_EINVAL, "rgf_w_solve_batched: Vii, V1i, PL, PG, PR, WL, WG, WR are required");
  if (b->nE == 0) return 0;
  const int n = b->nm, np = w_block(n), nx = b->nx;
  WVariant v;
  if (!pick_w(np, &v)) return fail(RGF_EUNSUP, "rgf_w_solve_batched: unsupported block size");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDev) return fail(RGF_EINVAL, "device index out of range");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  int per_cu = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, v.fn, v.nth, 0));
  if (per_cu < 1) return fail(RGF_EHIP, "W sweep kernel does not fit on a CU");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const size_t nn = (size_t)np * np;
  const long long nblk = (long long)nx * b->nE;
  // device layout: [conj(V1i) | padded V (2 nx) | padded P, W (6 or 7 nx nE) | energy slots]
  const bool pad = np != n;
  const size_t vc_el = nn * nx;
  const size_t pad_el = pad ? nn * (2 * (size_t)nx + (b->WRi1 ? 7 : 6) * (size_t)nblk) : 0;
  const long long slot_el = (long long)rgf::kWTmpBlocks * (long long)nn;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  Buf& ws = g_ws[dev];
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  const size_t fixed = (vc_el + pad_el) * sizeof(double2);
  const size_t avail = (size_t)((double)(fr + ws.bytes) * 0.9);
  if (avail <= fixed) return fail(RGF_ENOMEM, "rgf_w_solve_batched: not enough device memory for the padded blocks");
  long long slots = std::min<long long>(b->nE, (long long)prop.multiProcessorCount * per_cu);
  slots = std::min<long long>(slots, (long long)((avail - fixed) / (slot_el * sizeof(double2))));
  if (slots < 1) return fail(RGF_ENOMEM, "rgf_w_solve_batched: not enough device memory for one energy slot");
  if (int rc = ws.ensure(fixed + (size_t)slots * slot_el * sizeof(double2))) return rc;
  double2* base = reinterpret_cast<double2*>(ws.p);
  rgf::WArgs a{};
  a.n = np;
  a.nx = nx;
  a.nE = b->nE;
  a.info = b->info;
  double2* Vc = base;
  double2* q = base + vc_el;
  const unsigned g = 2048;
  const double2 zero = make_double2(0, 0);
  auto padded = [&](const double* src, long long nb) -> const double2* {
    double2* d = q;
    q += nn * nb;
    hipLaunchKernelGGL(pad_blocks, dim3(g), dim3(256), 0, s, reinterpret_cast<const double2*>(src), n, d, np, nb, zero);
    return d;
  };
  if (pad) {
    a.Vii = padded(b->Vii, nx);
    a.V1i = padded(b->V1i, nx);
    a.PL = padded(b->PL, nblk);
    a.PG = padded(b->PG, nblk);
    a.PR = padded(b->PR, nblk);
    a.WL = q; q += nn * nblk;
    a.WG = q; q += nn * nblk;
    a.WR = q; q += nn * nblk;
    a.WRi1 = nullptr;
    if (b->WRi1) { a.WRi1 = q; q += nn * nblk; }
  } else {
    a.Vii = reinterpret_cast<const double2*>(b->Vii);
    a.V1i = reinterpret_cast<const double2*>(b->V1i);
    a.PL = reinterpret_cast<const double2*>(b->PL);
    a.PG = reinterpret_cast<const double2*>(b->PG);
    a.PR = reinterpret_cast<const double2*>(b->PR);
    a.WL = reinterpret_cast<double2*>(b->WL);
    a.WG = reinterpret_cast<double2*>(b->WG);
    a.WR = reinterpret_cast<double2*>(b->WR);
    a.WRi1 = reinterpret_cast<double2*>(b->WRi1);
  }
  hipLaunchKernelGGL(conj_blocks, dim3(256), dim3(256), 0, s, a.V1i, Vc, vc_el);
  a.Vc = Vc;
  a.ws = base + vc_el + pad_el;
  a.ws_slot = slot_el;
  HIPCHK(hipGetLastError());
  v.launch(a, (int)slots, s);
  HIPCHK(hipGetLastError());
  if (pad) {
    const std::pair<const double2*, double*> outs[4] = {
        {a.WL, b->WL}, {a.WG, b->WG}, {a.WR, b->WR}, {a.WRi1, b->WRi1}};
    for (const auto& o : outs) {
      if (!o.second) continue;
      hipLaunchKernelGGL(unpad_blocks<double2>, dim3(g), dim3(256), 0, s, o.first, np, reinterpret_cast<double2*>(o.second),
                         n, nblk);
    }
    HIPCHK(hipGetLastError());
  }
  return 0;
}

int rgf_integrate(int nk, int nE, int nx, const double* E, const double* midgap, const double* dens_n,
                  const double* dens_p, const double* tr, double scale, double* nelec, double* pelec,
                  double* current, void* stream) {
  g_err.clear();
  if (nk < 1 || nx < 1) return fail(RGF_EINVAL, "rgf_integrate: need nk >= 1, nx >= 1");
  if (nE < 2) return fail(RGF_EINVAL, "rgf_integrate: need nE >= 2 (dE = E[1] - E[0])");
  if (!E) return fail(RGF_EINVAL, "rgf_integrate: E is required");
  const bool dens = dens_n || dens_p;
  if (dens && (!midgap || !nelec || !pelec))
    return fail(RGF_EINVAL, "rgf_integrate: densities need midgap, nelec and pelec");
  if (tr && !current) return fail(RGF_EINVAL, "rgf_integrate: tr needs current");
  if (!dens && !tr) return 0;
  // pi as the reference's module constant (src/green_rgf.f90:23): dE = (E2 - E1) / (2 pi), dK = 2 pi / nk
  const double pi_ref = 3.14159265359;
  rgf::IntegrateArgs a;
  a.nE = nE; a.nk = nk; a.nx = nx;
  a.E = E; a.midgap = midgap; a.dens_n = dens_n; a.dens_p = dens_p; a.tr = tr;
  a.two_pi = 2.0 * pi_ref;
  a.dK = 2.0 * pi_ref / nk;
  a.scale = scale;
  a.nelec = dens ? nelec : nullptr;
  a.pelec = dens ? pelec : nullptr;
  a.current = tr ? current : nullptr;
  hipLaunchKernelGGL(rgf::k_integrate, dim3(nx + 1), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), a);
  HIPCHK(hipGetLastError());
  return 0;
}

void rgf_release(void) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) return;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  g_ws[dev].release();
  g_stage[dev].release();
  g_adj_dh[dev].release();
}

}  // extern "C"
