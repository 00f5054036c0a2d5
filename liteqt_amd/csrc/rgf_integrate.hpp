// rgf_integrate.hpp -- energy (and k-point) integrals of the batched solve's
// per-energy outputs: the carrier densities of calc_n_electron
// (src/green_rgf.f90:3920-3945) and the terminal current of the SCBA drivers
// (src/green_rgf.f90:3189, 788).
//
// Inputs are the per-energy reductions the sweep kernels already write
// (dens_n = Im Tr G^<_ll, dens_p = -Im Tr G^>_ll, tr), laid out
// [(ik * nE + ie) * nx + l], so the integral reads nk*nE*(2*nx + 1) doubles
// instead of the full (nm, nm, nx, nE, nk) Green's functions.  One workgroup
// per slab (and one for the current) sums in a fixed order (strided partial
// sums per thread, then a wave tree and the four wave totals in order): the
// result does not depend on the launch or on timing.
#pragma once
#include <hip/hip_runtime.h>

namespace rgf {

struct IntegrateArgs {
  int nE, nk, nx;
  const double* E;       // nE energies (eV), uniform grid: dE = E[1] - E[0]
  const double* midgap;  // nx midgap energies (eV)
  const double* dens_n;  // (nk, nE, nx) Im Tr G^<_ll, may be null
  const double* dens_p;  // (nk, nE, nx) -Im Tr G^>_ll, may be null
  const double* tr;      // (nk, nE), may be null
  double two_pi;         // 2 pi as the caller's code defines it
  double dK;             // k-point weight (the reference: 2 pi / nk)
  double scale;          // current per unit of dE * sum tr
  double* nelec;         // nx
  double* pelec;         // nx
  double* current;       // 1
};

__device__ __forceinline__ double wave_sum_fixed(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// 256 threads.  Blocks 0..nx-1: nelec[l], pelec[l]; block nx: the current.
__global__ __launch_bounds__(256) void k_integrate(IntegrateArgs a) {
  __shared__ double red[2][4];
  const int l = blockIdx.x, tid = threadIdx.x;
  const int total = a.nk * a.nE;
  double sn = 0.0, sp = 0.0;
  if (l < a.nx && a.nelec) {
    const double mg = a.midgap[l];
    for (int j = tid; j < total; j += 256) {
      const double e = a.E[j % a.nE];
      const size_t o = (size_t)j * a.nx + l;
      if (e > mg) {
        if (a.dens_n) sn += a.dens_n[o];
      } else {
        if (a.dens_p) sp += a.dens_p[o];
      }
    }
  } else if (a.tr) {
    for (int j = tid; j < total; j += 256) sn += a.tr[j];
  }
  sn = wave_sum_fixed(sn);
  sp = wave_sum_fixed(sp);
  const int lane = tid & 63, wave = tid >> 6;
  if (lane == 0) {
    red[0][wave] = sn;
    red[1][wave] = sp;
  }
  __syncthreads();
  if (tid == 0) {
    const double dE = a.E[1] - a.E[0];
    const double n = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
    const double p = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
    if (l < a.nx) {
      if (!a.nelec) return;
      const double w = dE / a.two_pi * a.dK;
      a.nelec[l] = n * w;
      a.pelec[l] = p * w;
    } else if (a.current) {
      a.current[0] = n * dE * a.scale;
    }
  }
}

}  // namespace rgf
