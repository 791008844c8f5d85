// Work-item / workgroup ids and sizes read straight from the hardware registers and the dispatch
// packet.  The library is compiled without packed FP32 (-packed-fp32-ops), so the device library's
// id functions, built with that feature, cannot be inlined into these kernels: threadIdx, blockIdx,
// blockDim and gridDim would become out-of-line calls whose results the compiler treats as
// lane-varying (vector registers, divergent loop exits).  These builtins are inlined and uniform.
#pragma once
#include <stdint.h>

__device__ __forceinline__ uint32_t tid_x() { return __builtin_amdgcn_workitem_id_x(); }
__device__ __forceinline__ uint32_t bid_x() { return __builtin_amdgcn_workgroup_id_x(); }
__device__ __forceinline__ uint32_t bid_y() { return __builtin_amdgcn_workgroup_id_y(); }
__device__ __forceinline__ uint32_t bdim_x() { return __builtin_amdgcn_workgroup_size_x(); }
__device__ __forceinline__ uint32_t gdim_x() { return __builtin_amdgcn_grid_size_x() / __builtin_amdgcn_workgroup_size_x(); }
__device__ __forceinline__ uint32_t gdim_y() { return __builtin_amdgcn_grid_size_y() / __builtin_amdgcn_workgroup_size_y(); }

// Workgroup b's place among nb when the hardware deals workgroups to the eight XCDs round-robin
// (workgroup b to XCD b mod 8): each XCD gets one contiguous run of places, so that neighbouring
// places, whose rows share candidates, share that XCD's L2.  A bijection of [0, nb) for any nb,
// so a kernel that maps its workgroups through it still covers each place once.
__device__ __forceinline__ uint32_t xcd_place(uint32_t b, uint32_t nb) {
  constexpr uint32_t X = 8;
  const uint32_t q = nb / X, r = nb % X, x = b % X;
  return x * q + (x < r ? x : r) + b / X;
}
