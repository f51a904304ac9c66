// Host-side launch of the hop kernel family for one storage precision.
#pragma once

#include "hop_impl.hpp"

namespace qphix_amd {

template <int P, bool C12, int ISIGN>
static hipError_t launch_hop3(int kind, const HopArgs &a, int block, hipStream_t st) {
  const int nr = a.nrhs > 1 ? a.nrhs : 1;
  const dim3 grid((unsigned)(((a.vh + block - 1) / block) * nr));
  switch (kind) {
    case KD_WILSON: hop_kernel<P, C12, ISIGN, KD_WILSON><<<grid, block, a.lds_cap, st>>>(a); break;
    case KA_WILSON: hop_kernel<P, C12, ISIGN, KA_WILSON><<<grid, block, a.lds_cap, st>>>(a); break;
    case KD_CLOVER: hop_kernel<P, C12, ISIGN, KD_CLOVER><<<grid, block, a.lds_cap, st>>>(a); break;
    case KA_CLOVER: hop_kernel<P, C12, ISIGN, KA_CLOVER><<<grid, block, a.lds_cap, st>>>(a); break;
    case KD_TM: hop_kernel<P, C12, ISIGN, KD_TM><<<grid, block, a.lds_cap, st>>>(a); break;
    case KA_TM: hop_kernel<P, C12, ISIGN, KA_TM><<<grid, block, a.lds_cap, st>>>(a); break;
    case KD_FULL: hop_kernel<P, C12, ISIGN, KD_FULL><<<grid, block, a.lds_cap, st>>>(a); break;
    case KD_DIR:  // isign = +1 only (the reference hard-wires it)
      if constexpr (ISIGN == 1) {
        if (a.nrhs > 1) return hipErrorInvalidValue;
        hop_kernel<P, C12, ISIGN, KD_DIR><<<grid, block, 0, st>>>(a);
        break;
      }
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int P>
hipError_t launch_hop_prec(int kind, bool c12, int isign, const HopArgs &a, int block, hipStream_t st) {
  if (block <= 0 || block > 256 || (block & 63)) return hipErrorInvalidValue;
  if (c12) return isign > 0 ? launch_hop3<P, true, 1>(kind, a, block, st) : launch_hop3<P, true, -1>(kind, a, block, st);
  if constexpr (!PT<P>::scaled) {
    return isign > 0 ? launch_hop3<P, false, 1>(kind, a, block, st) : launch_hop3<P, false, -1>(kind, a, block, st);
  }
  return hipErrorInvalidValue;
}

// register-tiled multi-RHS launch: R right-hand sides per work-item, a.nrhs
// a multiple of R
template <int P, bool C12, int ISIGN, int R>
static hipError_t launch_tile3(int kind, const HopArgs &a, int block, hipStream_t st) {
  const dim3 grid((unsigned)(((a.vh + block - 1) / block) * (a.nrhs / R)));
  switch (kind) {
#define TK(K) \
  case K: hop_tile_kernel<P, C12, ISIGN, K, R><<<grid, block, a.lds_cap, st>>>(a); break;
    TK(KD_WILSON) TK(KA_WILSON) TK(KD_CLOVER) TK(KA_CLOVER) TK(KD_TM) TK(KA_TM) TK(KD_FULL)
#undef TK
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int P, bool C12>
static hipError_t launch_tile2(int kind, int isign, int rt, const HopArgs &a, int block, hipStream_t st) {
  if (rt == 2)
    return isign > 0 ? launch_tile3<P, C12, 1, 2>(kind, a, block, st) : launch_tile3<P, C12, -1, 2>(kind, a, block, st);
  if (rt == 4)
    return isign > 0 ? launch_tile3<P, C12, 1, 4>(kind, a, block, st) : launch_tile3<P, C12, -1, 4>(kind, a, block, st);
  return hipErrorInvalidValue;
}

template <int P>
hipError_t launch_hop_tile_prec(int kind, bool c12, int isign, int rt, const HopArgs &a, int block, hipStream_t st) {
  if (block <= 0 || block > 256 || (block & 63) || a.nrhs % rt) return hipErrorInvalidValue;
  if (c12) return launch_tile2<P, true>(kind, isign, rt, a, block, st);
  if constexpr (!PT<P>::scaled) return launch_tile2<P, false>(kind, isign, rt, a, block, st);
  return hipErrorInvalidValue;
}

constexpr int kBatchBlock = 256;

template <int P, bool C12, int ISIGN>
static hipError_t launch_batch3(int kind, const HopSeg *segs, const int *off, int nseg, const int *active, int grid,
                                hipStream_t st) {
  switch (kind) {
#define BK(K) \
  case K: hop_batch_kernel<P, C12, ISIGN, K><<<grid, kBatchBlock, 0, st>>>(segs, off, nseg, active); break;
    BK(KD_WILSON) BK(KA_WILSON) BK(KD_CLOVER) BK(KA_CLOVER) BK(KD_TM) BK(KA_TM) BK(KD_FULL)
#undef BK
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// tiled batch: every segment carries kBatchTile systems sharing one operator
template <int P, bool C12, int ISIGN>
static hipError_t launch_batch_tile3(int kind, const HopSeg *segs, const int *off, int nseg, const int *active,
                                     int grid, hipStream_t st) {
  switch (kind) {
#define BT(K) \
  case K: hop_batch_tile_kernel<P, C12, ISIGN, K, kBatchTile><<<grid, kBatchBlock, 0, st>>>(segs, off, nseg, active); break;
    BT(KD_WILSON) BT(KA_WILSON) BT(KD_CLOVER) BT(KA_CLOVER) BT(KD_TM) BT(KA_TM) BT(KD_FULL)
#undef BT
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int P>
hipError_t launch_hop_batch_tile_prec(int kind, bool c12, int isign, const HopSeg *segs, const int *off, int nseg,
                                      const int *active, int grid, hipStream_t st) {
  if (c12)
    return isign > 0 ? launch_batch_tile3<P, true, 1>(kind, segs, off, nseg, active, grid, st)
                     : launch_batch_tile3<P, true, -1>(kind, segs, off, nseg, active, grid, st);
  if constexpr (!PT<P>::scaled) {
    return isign > 0 ? launch_batch_tile3<P, false, 1>(kind, segs, off, nseg, active, grid, st)
                     : launch_batch_tile3<P, false, -1>(kind, segs, off, nseg, active, grid, st);
  }
  return hipErrorInvalidValue;
}

// segs / off / active are device arrays; off has nseg + 1 entries
template <int P>
hipError_t launch_hop_batch_prec(int kind, bool c12, int isign, const HopSeg *segs, const int *off, int nseg,
                                 const int *active, int grid, hipStream_t st) {
  if (c12)
    return isign > 0 ? launch_batch3<P, true, 1>(kind, segs, off, nseg, active, grid, st)
                     : launch_batch3<P, true, -1>(kind, segs, off, nseg, active, grid, st);
  if constexpr (!PT<P>::scaled) {
    return isign > 0 ? launch_batch3<P, false, 1>(kind, segs, off, nseg, active, grid, st)
                     : launch_batch3<P, false, -1>(kind, segs, off, nseg, active, grid, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace qphix_amd
