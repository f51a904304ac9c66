// Block BLAS over sets of spinor fields (block Krylov solvers).
#pragma once

#include "fields.hpp"

namespace qphix_amd {

struct BlockArgs {
  const void *a[kMaxRhs];  // first set (na fields)
  const void *b[kMaxRhs];  // Gram: second set (nb fields)
  void *o[kMaxRhs];        // combine: outputs (nb fields)
  const void *y[kMaxRhs];  // combine: added to the outputs (null entries: nothing)
  int na, nb;
};

constexpr int kMaxGramBlocks = 512;
// device partial-sum buffer of the Gram kernel, in doubles
constexpr size_t kGramPartials = (size_t)kMaxGramBlocks * kMaxRhs * kMaxRhs * 2;

// out[(i * nb + j) * 2 + {0, 1}] = sum over sites of conj(a_i) . b_j (FP64),
// for FP64 and FP32 storage
hipError_t block_gram_launch(const GeomDesc &g, const BlockArgs &ba, double *partials, double *out, hipStream_t st);
// o_j = y_j + sum_i a_i mat_ij, j < nb; mat is a device array of na x nb
// complex doubles, row-major; o_j may alias any a_i or y_j
hipError_t block_combine_launch(const GeomDesc &g, const BlockArgs &ba, const double *mat, hipStream_t st);

}  // namespace qphix_amd
