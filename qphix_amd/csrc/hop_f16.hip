// Hop kernel instantiations for P16 storage.
#include "hop_launch.hpp"
#include "hop_mfma.hpp"

namespace qphix_amd {
template hipError_t launch_hop_prec<P16>(int, bool, int, const HopArgs &, int, hipStream_t);
template hipError_t launch_hop_tile_prec<P16>(int, bool, int, int, const HopArgs &, int, hipStream_t);
hipError_t launch_hop_mfma_f16(int kind, bool c12, int isign, const HopArgs &a, hipStream_t st) {
  return launch_hop_mfma16(kind, c12, isign, a, st);
}
template hipError_t launch_hop_batch_prec<P16>(int, bool, int, const HopSeg *, const int *, int, const int *,
                                                  int, hipStream_t);
template hipError_t launch_hop_batch_tile_prec<P16>(int, bool, int, const HopSeg *, const int *, int, const int *,
                                                       int, hipStream_t);
}
