// Drop-in FullSpinor (full_spinor.h:15-60 of JeffersonLab/qphix): both
// checkerboards of a spinor field, allocated from and freed to the geometry.
#pragma once

#include "geometry.h"

namespace QPhiX {

template <typename T, int V, int S, bool compress12>
class FullSpinor {
 public:
  using CBSpinor = typename Geometry<T, V, S, compress12>::FourSpinorBlock;
  explicit FullSpinor(Geometry<T, V, S, compress12> &geom) : geom_(geom)
  {
    data_[0] = geom.allocCBFourSpinor();
    data_[1] = geom.allocCBFourSpinor();
  }
  ~FullSpinor()
  {
    geom_.free(data_[0]);
    geom_.free(data_[1]);
  }
  FullSpinor(const FullSpinor &) = delete;
  FullSpinor &operator=(const FullSpinor &) = delete;
  CBSpinor *getCBData(int cb) const { return data_[cb]; }
  Geometry<T, V, S, compress12> &getGeometry() { return geom_; }

 private:
  Geometry<T, V, S, compress12> &geom_;
  CBSpinor *data_[2];
};

}  // namespace QPhiX
