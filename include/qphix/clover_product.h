// clover_product (clover_product.h:250-339 of JeffersonLab/qphix): out += A in
// for one checkerboard, A the packed Hermitian clover blocks (CloverBlock) or
// the full complex blocks (FullCloverBlock). As in the reference the product
// is accumulated into out, which the caller zeroes first when it wants A in.
#pragma once

#include <type_traits>

#include "blas_new_c.h"

namespace QPhiX {

template <typename FT, int veclen, int soalen, bool compress12, typename Clover>
void clover_product(typename Geometry<FT, veclen, soalen, compress12>::FourSpinorBlock *const out,
                    typename Geometry<FT, veclen, soalen, compress12>::FourSpinorBlock const *const in,
                    Clover *clover, Geometry<FT, veclen, soalen, compress12> &geom)
{
  typedef Geometry<FT, veclen, soalen, compress12> G;
  static_assert(std::is_same<typename std::remove_const<Clover>::type, typename G::CloverBlock>::value ||
                    std::is_same<typename std::remove_const<Clover>::type, typename G::FullCloverBlock>::value,
                "clover_product takes the geometry's CloverBlock or FullCloverBlock");
  const int full = std::is_same<typename std::remove_const<Clover>::type, typename G::FullCloverBlock>::value;
  typename G::FourSpinorBlock *tmp = geom.allocCBFourSpinor();
  int rc = qphix_clover_apply(geom.handle(), tmp, in, clover, full);
  if (rc == QPHIX_OK) rc = qphix_ypeqx(geom.handle(), tmp, out);
  geom.free(tmp);
  detail::check(rc, "clover_product");
}

}  // namespace QPhiX
