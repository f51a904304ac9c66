// Drop-in QPhiX::Geometry for the gfx950 library (replaces
// include/qphix/geometry.h:115-441 of JeffersonLab/qphix).
//
// Template arguments keep the reference's meaning: FT is the storage and
// arithmetic precision (double, float or QPhiX::half), compress12 selects
// two-row gauge storage. VECLEN and SOALEN describe CPU SIMD blocking and have
// no effect on the GPU layout; the blocking / threading constructor arguments
// are accepted and ignored. Block pointers returned by the alloc functions are
// device addresses owned by the caller until free().
#pragma once

#include <stdexcept>
#include <string>
#include <type_traits>

#include "../qphix_amd.h"

namespace QPhiX {

// 16-bit storage tag; arithmetic on half fields happens in float
// (include/qphix/arith_type.h:8-17 of the reference)
struct half {
  unsigned short bits;
};

namespace detail {
template <typename FT> struct PrecOf;
template <> struct PrecOf<double> { static constexpr int value = QPHIX_F64; };
template <> struct PrecOf<float> { static constexpr int value = QPHIX_F32; };
template <> struct PrecOf<half> { static constexpr int value = QPHIX_F16; };

inline void check(int status, const char *what) {
  if (status == QPHIX_OK) return;
  std::string msg = std::string(what) + ": " + qphix_strerror(status) + ": " + qphix_last_error();
  if (status == QPHIX_EDOMAIN) throw std::domain_error(msg);
  throw std::runtime_error(msg);
}
}  // namespace detail

template <typename FT, int veclen, int soalen, bool compress12>
class Geometry {
 public:
  // opaque device blocks: only pointers to them are ever formed
  struct FourSpinorBlock;
  struct TwoSpinorBlock;
  struct SU3MatrixBlock;
  struct CloverBlock;
  struct FullCloverBlock;

  Geometry(const int latt_size[], int By = 1, int Bz = 1, int NCores = 1, int Sy = 1, int Sz = 1,
           int PadXY = 0, int PadXYZ = 0, int MinCt = 1, bool verbose = false)
  {
    (void)By; (void)Bz; (void)NCores; (void)Sy; (void)Sz; (void)PadXY; (void)PadXYZ; (void)MinCt;
    (void)verbose;
    for (int k = 0; k < 4; k++) latt_[k] = latt_size[k];
    detail::check(qphix_geom_create(latt_, detail::PrecOf<FT>::value, compress12 ? 1 : 0, &g_),
                  "Geometry");
  }
  ~Geometry() { qphix_geom_destroy(g_); }
  Geometry(const Geometry &) = delete;
  Geometry &operator=(const Geometry &) = delete;

  int Nx() const { return latt_[0]; }
  int Nxh() const { return latt_[0] / 2; }
  int Ny() const { return latt_[1]; }
  int Nz() const { return latt_[2]; }
  int Nt() const { return latt_[3]; }
  int getVolCB() const { return latt_[0] * latt_[1] * latt_[2] * latt_[3] / 2; }
  int getNSIMT() const { return 1; }
  int getNumCores() const { return 1; }

  FourSpinorBlock *allocCBFourSpinor() { return static_cast<FourSpinorBlock *>(alloc(qphix_alloc_spinor)); }
  SU3MatrixBlock *allocCBGauge() { return static_cast<SU3MatrixBlock *>(alloc(qphix_alloc_gauge)); }
  CloverBlock *allocCBClov() { return static_cast<CloverBlock *>(alloc(qphix_alloc_clover)); }
  FullCloverBlock *allocCBFullClov() { return static_cast<FullCloverBlock *>(alloc(qphix_alloc_full_clover)); }
  void free(FourSpinorBlock *p) { qphix_free(p); }
  void free(SU3MatrixBlock *p) { qphix_free(p); }
  void free(CloverBlock *p) { qphix_free(p); }
  void free(FullCloverBlock *p) { qphix_free(p); }

  qphix_geom *handle() const { return g_; }

 private:
  void *alloc(int (*fn)(const qphix_geom *, void **))
  {
    void *p = nullptr;
    detail::check(fn(g_, &p), "Geometry alloc");
    return p;
  }
  int latt_[4];
  qphix_geom *g_ = nullptr;
};

// Canonical-array packers (the QDP++-free counterparts of
// include/qphix/qdp_packer_parscalar.h); layouts are documented in qphix_amd.h.
template <typename FT, int V, int S, bool C>
void pack_spinor(const Geometry<FT, V, S, C> &g, typename Geometry<FT, V, S, C>::FourSpinorBlock *dst,
                 const double *src, int cb)
{
  detail::check(qphix_pack_spinor(g.handle(), dst, src, cb), "pack_spinor");
}
template <typename FT, int V, int S, bool C>
void unpack_spinor(const Geometry<FT, V, S, C> &g, double *dst,
                   const typename Geometry<FT, V, S, C>::FourSpinorBlock *src, int cb)
{
  detail::check(qphix_unpack_spinor(g.handle(), dst, src, cb), "unpack_spinor");
}
template <typename FT, int V, int S, bool C>
void pack_gauge(const Geometry<FT, V, S, C> &g, typename Geometry<FT, V, S, C>::SU3MatrixBlock *u_cb0,
                typename Geometry<FT, V, S, C>::SU3MatrixBlock *u_cb1, const double *src)
{
  detail::check(qphix_pack_gauge(g.handle(), u_cb0, u_cb1, src), "pack_gauge");
}
template <typename FT, int V, int S, bool C>
void pack_clover(const Geometry<FT, V, S, C> &g, typename Geometry<FT, V, S, C>::CloverBlock *dst,
                 const double *src, int cb)
{
  detail::check(qphix_pack_clover(g.handle(), dst, src, cb), "pack_clover");
}
template <typename FT, int V, int S, bool C>
void pack_full_clover(const Geometry<FT, V, S, C> &g, typename Geometry<FT, V, S, C>::FullCloverBlock *dst,
                      const double *src, int cb)
{
  detail::check(qphix_pack_full_clover(g.handle(), dst, src, cb), "pack_full_clover");
}

// Owning handles of one checkerboard field (geometry.h:443-548): allocated
// from the geometry on construction, freed on destruction, movable.
namespace detail {
template <class G, class Block, Block *(G::*Alloc)()> class FieldHandle {
 public:
  typedef G Geom;
  typedef Block ValueType;
  explicit FieldHandle(Geom &geom) : value_((geom.*Alloc)()), geom_(geom) {}
  FieldHandle(FieldHandle &&other) noexcept : value_(other.value_), geom_(other.geom_) { other.value_ = nullptr; }
  FieldHandle(const FieldHandle &) = delete;
  FieldHandle &operator=(const FieldHandle &) = delete;
  ~FieldHandle()
  {
    if (value_) geom_.free(value_);
  }
  ValueType *get() const { return value_; }

 private:
  ValueType *value_;
  Geom &geom_;
};
}  // namespace detail

#define QPHIX_HANDLE(Name, Block, Alloc)                                                                          \
  template <typename FT, int V, int S, bool C>                                                                    \
  class Name : public detail::FieldHandle<Geometry<FT, V, S, C>, typename Geometry<FT, V, S, C>::Block,           \
                                          &Geometry<FT, V, S, C>::Alloc> {                                       \
    typedef detail::FieldHandle<Geometry<FT, V, S, C>, typename Geometry<FT, V, S, C>::Block,                     \
                                &Geometry<FT, V, S, C>::Alloc>                                                    \
        Base;                                                                                                     \
                                                                                                                  \
   public:                                                                                                        \
    using Base::Base;                                                                                             \
  };                                                                                                              \
  template <typename FT, int V, int S, bool C> Name<FT, V, S, C> make##Name(Geometry<FT, V, S, C> &geom)          \
  {                                                                                                               \
    return Name<FT, V, S, C>(geom);                                                                               \
  }
QPHIX_HANDLE(FourSpinorHandle, FourSpinorBlock, allocCBFourSpinor)
QPHIX_HANDLE(GaugeHandle, SU3MatrixBlock, allocCBGauge)
QPHIX_HANDLE(CloverHandle, CloverBlock, allocCBClov)
QPHIX_HANDLE(FullCloverHandle, FullCloverBlock, allocCBFullClov)
#undef QPHIX_HANDLE

}  // namespace QPhiX
