// Hop kernel instantiations for P64 storage.
#include "hop_launch.hpp"

namespace qphix_amd {
template hipError_t launch_hop_prec<P64>(int, bool, int, const HopArgs &, int, hipStream_t);
template hipError_t launch_hop_tile_prec<P64>(int, bool, int, int, const HopArgs &, int, hipStream_t);
template hipError_t launch_hop_batch_prec<P64>(int, bool, int, const HopSeg *, const int *, int, const int *,
                                                  int, hipStream_t);
template hipError_t launch_hop_batch_tile_prec<P64>(int, bool, int, const HopSeg *, const int *, int, const int *,
                                                       int, hipStream_t);
}
