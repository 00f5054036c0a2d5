// rgf_multi.hpp -- many independent devices (each its own Hii/H1i, block
// size nm, slab count nx and energy grid) in one persistent launch.
//
// The host flattens all (device, energy) pairs into one work list sorted by
// predicted cost, longest first.  Every workgroup of rgf_sweep_multi takes the
// next item from a global counter (one atomic per item, issued by thread 0 and
// broadcast through LDS) and runs that energy's full solve with the tile
// variant chosen for its block size, so devices of different nm share the
// machine dynamically and the small items fill in at the end (longest-
// processing-time-first list scheduling).  Every workgroup leaves the loop
// once the counter passes the end of the list.
#pragma once
#include "rgf_sweep.hpp"

namespace rgf {

// Tile variants compiled into the multi-device kernel: (T, NMAX).
enum MultiVariant : int {
  MV_1_32 = 0, MV_2_64, MV_1_64, MV_3_96, MV_1_96, MV_4_128, MV_2_128, MV_3_192, MV_1_192, MV_4_256, MV_2_256,
  MV_1_256, MV_COUNT
};

struct MultiArgs {
  const SweepArgs* jobs;  // one per device (ws, ws_slot unused)
  const int* variant;     // MultiVariant per device
  const int2* items;      // (device, energy), longest first
  int nitems;
  int* counter;           // zero before the launch
  double2* ws;
  long long ws_slot;      // complex elements per workgroup (fits the largest device)
};

template <int T, int NMAX>
__device__ __forceinline__ void run_variant(const SweepArgs& a, int e, double2* ws, ldsd* lds, ldsd* red) {
  solve_energy<T, NMAX>(a, e, ws, lds, red);
}

struct MultiLds {
  static constexpr int A = SweepLds<4, 256>::DOUBLES;
  static constexpr int B = SweepLds<3, 192>::DOUBLES;
  static constexpr int C = SweepLds<2, 256>::DOUBLES;
  static constexpr int AB = A > B ? A : B;
  static constexpr int DOUBLES = AB > C ? AB : C;
};

__global__ __launch_bounds__(256, 1) void rgf_sweep_multi(MultiArgs m) {
  __shared__ double lds[MultiLds::DOUBLES];
  __shared__ int next;
  static_assert(Cfg<1>::NTH == 256 && Cfg<2>::NTH == 256 && Cfg<3>::NTH == 256 && Cfg<4>::NTH == 256,
                "all variants run 256 threads");
  double2* ws = m.ws + (size_t)blockIdx.x * (size_t)m.ws_slot;
  ldsd* L = (ldsd*)lds;
  ldsd* R = L;  // block reductions run between block operations, when the LDS is free
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(m.counter, 1);
    __syncthreads();
    const int it = uni(next);
    __syncthreads();
    if (it >= m.nitems) break;
    const int2 wv = m.items[it];
    const int2 w = make_int2(uni(wv.x), uni(wv.y));
    const SweepArgs& a = m.jobs[w.x];
    switch (uni(m.variant[w.x])) {
      case MV_1_32: run_variant<1, 32>(a, w.y, ws, L, R); break;
      case MV_2_64: run_variant<2, 64>(a, w.y, ws, L, R); break;
      case MV_1_64: run_variant<1, 64>(a, w.y, ws, L, R); break;
      case MV_3_96: run_variant<3, 96>(a, w.y, ws, L, R); break;
      case MV_1_96: run_variant<1, 96>(a, w.y, ws, L, R); break;
      case MV_4_128: run_variant<4, 128>(a, w.y, ws, L, R); break;
      case MV_2_128: run_variant<2, 128>(a, w.y, ws, L, R); break;
      case MV_3_192: run_variant<3, 192>(a, w.y, ws, L, R); break;
      case MV_1_192: run_variant<1, 192>(a, w.y, ws, L, R); break;
      case MV_4_256: run_variant<4, 256>(a, w.y, ws, L, R); break;
      case MV_2_256: run_variant<2, 256>(a, w.y, ws, L, R); break;
      default: run_variant<1, 256>(a, w.y, ws, L, R); break;
    }
  }
}

__global__ void k_fill_int(int* p, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace rgf
