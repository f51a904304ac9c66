// Hop kernel instantiations for bfloat16 (PB16) storage.
#include "hop_launch.hpp"

namespace qphix_amd {
template hipError_t launch_hop_prec<PB16>(int, bool, int, const HopArgs &, int, hipStream_t);
template hipError_t launch_hop_tile_prec<PB16>(int, bool, int, int, const HopArgs &, int, hipStream_t);
template hipError_t launch_hop_batch_prec<PB16>(int, bool, int, const HopSeg *, const int *, int, const int *,
                                                  int, hipStream_t);
template hipError_t launch_hop_batch_tile_prec<PB16>(int, bool, int, const HopSeg *, const int *, int, const int *,
                                                       int, hipStream_t);
}
