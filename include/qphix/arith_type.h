// ArithType (arith_type.h:8-17 of JeffersonLab/qphix): the type arithmetic is
// done in for a storage type; half-precision fields compute in float.
#pragma once

#include "geometry.h"

namespace QPhiX {

template <typename FT> struct ArithType {
  typedef FT Type;
};
template <> struct ArithType<half> {
  typedef float Type;
};

}  // namespace QPhiX
