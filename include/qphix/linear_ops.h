// Drop-in even-odd operators for the gfx950 library: EvenOddLinearOperator /
// TwoFlavEvenOddLinearOperator (linearOp.h:8-147), EvenOddWilsonOperator
// (wilson.h:11-100), EvenOddCloverOperator (clover.h:15-267),
// EvenOddTMWilsonOperator (twisted_mass.h:12-113), EvenOddTMCloverOperator
// (twisted_mass_clover.h:13-137), EvenOddNDTMWilsonReuseOperator
// (ndtm_reuse_operator.h:14-150) and EvenOddNDTMCloverReuseOperator
// (ndtm_reuse_operator_clover.h:25-169) of JeffersonLab/qphix.
//
// The concrete operators hold a library operator (qphix_op) whose
// operator() is one C-ABI call; solvers recognise them through native().
// Field pointers are not owned, as in the reference.
#pragma once

#include "dslash.h"
#include "full_spinor.h"

namespace QPhiX {

template <typename FT, int veclen, int soalen, bool compress12>
class EvenOddLinearOperator {
 public:
  typedef typename Geometry<FT, veclen, soalen, compress12>::FourSpinorBlock FourSpinorBlock;
  static constexpr int num_flav = 1;
  virtual ~EvenOddLinearOperator() {}
  virtual void operator()(FourSpinorBlock *res, const FourSpinorBlock *in, int isign, int target_cb = 1) const = 0;
  // multigrid-facing pieces (linearOp.h:26-35): the library operators provide
  // them for Wilson and clover, the others report "not yet implemented" by a
  // std::runtime_error as the reference prints it
  virtual void M_offdiag(FourSpinorBlock *res, FourSpinorBlock const *in, int isign, int target_cb) const
  {
    void *r[1] = {res};
    const void *i[1] = {in};
    detail::check(qphix_op_m_offdiag(need_native(), r, i, isign, target_cb), "M_offdiag");
  }
  virtual void M_diag_inv(FourSpinorBlock *res, FourSpinorBlock const *in, int isign, int cb = 0) const
  {
    void *r[1] = {res};
    const void *i[1] = {in};
    detail::check(qphix_op_m_diag_inv(need_native(), r, i, isign, cb), "M_diag_inv");
  }
  virtual Geometry<FT, veclen, soalen, compress12> &getGeometry() = 0;
  // library operator backing this object, or nullptr for user-defined operators
  virtual qphix_op *native() const { return nullptr; }

 private:
  qphix_op *need_native() const
  {
    if (!native()) throw std::runtime_error("operator has no library implementation of this piece");
    return native();
  }
};

template <typename FT, int veclen, int soalen, bool compress12>
class TwoFlavEvenOddLinearOperator {
 public:
  typedef typename Geometry<FT, veclen, soalen, compress12>::FourSpinorBlock FourSpinorBlock;
  static constexpr int num_flav = 2;
  virtual ~TwoFlavEvenOddLinearOperator() {}
  virtual void operator()(FourSpinorBlock *const res[2], const FourSpinorBlock *const in[2], int isign,
                          int target_cb = 1) = 0;
  virtual Geometry<FT, veclen, soalen, compress12> &getGeometry() = 0;
  virtual qphix_op *native() const { return nullptr; }
};

namespace detail {
template <typename FT, int V, int S, bool C> class NativeOp {
 public:
  NativeOp(Geometry<FT, V, S, C> *geom, int kind, double t_boundary, double aniso_s, double aniso_t,
           const bool use_tbc[4], const double tbc_phases[4][2])
      : geom_(geom)
  {
    p_ = qphix_op_params{};
    p_.kind = kind;
    p_.hop = hop_params(t_boundary, aniso_s, aniso_t, use_tbc, tbc_phases);
  }
  ~NativeOp()
  {
    if (op_) qphix_op_destroy(op_);
  }
  void create()
  {
    if (op_) qphix_op_destroy(op_);
    op_ = nullptr;
    check(qphix_op_create(geom_->handle(), &p_, &op_), "EvenOdd operator");
  }
  void apply(void *const *res, const void *const *in, int isign, int cb) const
  {
    if (!op_) throw std::runtime_error("operator fields not set");
    check(qphix_op_apply(op_, res, in, isign, cb), "EvenOdd operator()");
  }
  Geometry<FT, V, S, C> *geom_;
  qphix_op_params p_;
  qphix_op *op_ = nullptr;
};
}  // namespace detail

#define QPHIX_NATIVE_OP_COMMON                                                           \
 public:                                                                               \
  Geometry<FT, veclen, soalen, compress12> &getGeometry() override { return *n_.geom_; } \
  qphix_op *native() const override { return n_.op_; }                                 \
                                                                                       \
 private:                                                                              \
  detail::NativeOp<FT, veclen, soalen, compress12> n_;                                 \
                                                                                       \
 public:

template <typename FT, int veclen, int soalen, bool compress12>
class EvenOddWilsonOperator : public EvenOddLinearOperator<FT, veclen, soalen, compress12> {
  QPHIX_NATIVE_OP_COMMON
  typedef typename Geometry<FT, veclen, soalen, compress12>::FourSpinorBlock FourSpinorBlock;
  typedef typename Geometry<FT, veclen, soalen, compress12>::SU3MatrixBlock SU3MatrixBlock;
  EvenOddWilsonOperator(const double Mass, SU3MatrixBlock *u[2], Geometry<FT, veclen, soalen, compress12> *geom,
                        double t_boundary, double aniso_fac_s, double aniso_fac_t, bool use_tbc[4] = nullptr,
                        double tbc_phases[4][2] = nullptr)
      : n_(geom, QPHIX_OP_WILSON, t_boundary, aniso_fac_s, aniso_fac_t, use_tbc, tbc_phases)
  {
    n_.p_.mass = Mass;
    n_.p_.u[0] = u[0];
    n_.p_.u[1] = u[1];
    n_.create();
  }
  void operator()(FourSpinorBlock *res, const FourSpinorBlock *in, int isign, int target_cb = 1) const override
  {
    void *r[1] = {res};
    const void *i[1] = {in};
    n_.apply(r, i, isign, target_cb);
  }
};

template <typename FT, int veclen, int soalen, bool compress12>
class EvenOddCloverOperator : public EvenOddLinearOperator<FT, veclen, soalen, compress12> {
  QPHIX_NATIVE_OP_COMMON
  typedef typename Geometry<FT, veclen, soalen, compress12>::FourSpinorBlock FourSpinorBlock;
  typedef typename Geometry<FT, veclen, soalen, compress12>::SU3MatrixBlock SU3MatrixBlock;
  typedef typename Geometry<FT, veclen, soalen, compress12>::CloverBlock CloverBlock;
  // single clover parity (the odd target) and the inverse on the even sites
  EvenOddCloverOperator(SU3MatrixBlock *u[2], CloverBlock *clov, CloverBlock *invclov_even,
                        Geometry<FT, veclen, soalen, compress12> *geom, double t_boundary, double aniso_coeff_s,
                        double aniso_coeff_t, bool use_tbc[4] = nullptr, double tbc_phases[4][2] = nullptr,
                        double const prec_mass_rho = 0.0, CloverBlock *invclov_odd = nullptr)
      : n_(geom, QPHIX_OP_CLOVER, t_boundary, aniso_coeff_s, aniso_coeff_t, use_tbc, tbc_phases)
  {
    n_.p_.prec_mass_rho = prec_mass_rho;
    set(u, nullptr, clov, invclov_even, invclov_odd);
  }
  // both clover parities
  EvenOddCloverOperator(SU3MatrixBlock *u[2], CloverBlock *clov[2], CloverBlock *invclov_even,
                        Geometry<FT, veclen, soalen, compress12> *geom, double t_boundary, double aniso_coeff_s,
                        double aniso_coeff_t, bool use_tbc[4] = nullptr, double tbc_phases[4][2] = nullptr,
                        double const prec_mass_rho = 0.0, CloverBlock *invclov_odd = nullptr)
      : n_(geom, QPHIX_OP_CLOVER, t_boundary, aniso_coeff_s, aniso_coeff_t, use_tbc, tbc_phases)
  {
    n_.p_.prec_mass_rho = prec_mass_rho;
    set(u, clov[0], clov[1], invclov_even, invclov_odd);
  }
  EvenOddCloverOperator(Geometry<FT, veclen, soalen, compress12> *geom, double t_boundary, double aniso_coeff_s,
                        double aniso_coeff_t, bool use_tbc[4] = nullptr, double tbc_phases[4][2] = nullptr,
                        double const prec_mass_rho = 0.0)
      : n_(geom, QPHIX_OP_CLOVER, t_boundary, aniso_coeff_s, aniso_coeff_t, use_tbc, tbc_phases)
  {
    n_.p_.prec_mass_rho = prec_mass_rho;
  }
  void setFields(SU3MatrixBlock *u[2], CloverBlock *clov, CloverBlock *invclov) { set(u, nullptr, clov, invclov, nullptr); }
  void setFields(SU3MatrixBlock *u[2], CloverBlock *clov[2], CloverBlock *invclov)
  {
    set(u, clov[0], clov[1], invclov, nullptr);
  }
  void operator()(FourSpinorBlock *res, const FourSpinorBlock *in, int isign, int target_cb = 1) const override
  {
    void *r[1] = {res};
    const void *i[1] = {in};
    n_.apply(r, i, isign, target_cb);
  }

  // clover.h:196-251: A on both (or one) checkerboards, the unpreconditioned
  // operator A - 1/2 D, and one direction of the hop (isign +1)
  using SpinorFull = FullSpinor<FT, veclen, soalen, compress12>;
  void M_diag(SpinorFull &res, const SpinorFull &in, int isign) const { diag(res, in, isign, -1); }
  void M_diag(SpinorFull &res, const SpinorFull &in, int isign, int cb) const { diag(res, in, isign, cb); }
  void M_unprec(SpinorFull &res, const SpinorFull &in, int isign) const
  {
    void *r[2] = {res.getCBData(0), res.getCBData(1)};
    const void *i[2] = {in.getCBData(0), in.getCBData(1)};
    detail::check(qphix_op_m_unprec(n_.op_, r, i, isign), "M_unprec");
  }
  void DslashDir(SpinorFull &res, const SpinorFull &in, int dir) const
  {
    void *r[2] = {res.getCBData(0), res.getCBData(1)};
    const void *i[2] = {in.getCBData(0), in.getCBData(1)};
    detail::check(qphix_op_dslash_dir(n_.op_, r, i, dir), "DslashDir");
  }

 private:
  void diag(SpinorFull &res, const SpinorFull &in, int isign, int cb) const
  {
    void *r[2] = {res.getCBData(0), res.getCBData(1)};
    const void *i[2] = {in.getCBData(0), in.getCBData(1)};
    detail::check(qphix_op_m_diag(n_.op_, r, i, isign, cb), "M_diag");
  }
  void set(SU3MatrixBlock *u[2], CloverBlock *c0, CloverBlock *c1, CloverBlock *i0, CloverBlock *i1)
  {
    n_.p_.u[0] = u[0];
    n_.p_.u[1] = u[1];
    n_.p_.clov[0] = c0;
    n_.p_.clov[1] = c1;
    n_.p_.invclov[0] = i0;
    n_.p_.invclov[1] = i1;
    n_.create();
  }
};

template <typename FT, int veclen, int soalen, bool compress12>
class EvenOddTMWilsonOperator : public EvenOddLinearOperator<FT, veclen, soalen, compress12> {
  QPHIX_NATIVE_OP_COMMON
  typedef typename Geometry<FT, veclen, soalen, compress12>::FourSpinorBlock FourSpinorBlock;
  typedef typename Geometry<FT, veclen, soalen, compress12>::SU3MatrixBlock SU3MatrixBlock;
  EvenOddTMWilsonOperator(const double Mass, const double TwistedMass, SU3MatrixBlock *u[2],
                          Geometry<FT, veclen, soalen, compress12> *geom, double t_boundary, double aniso_fac_s,
                          double aniso_fac_t, bool use_tbc[4] = nullptr, double tbc_phases[4][2] = nullptr)
      : n_(geom, QPHIX_OP_TMWILSON, t_boundary, aniso_fac_s, aniso_fac_t, use_tbc, tbc_phases)
  {
    n_.p_.mass = Mass;
    n_.p_.twisted_mass = TwistedMass;
    n_.p_.u[0] = u[0];
    n_.p_.u[1] = u[1];
    n_.create();
  }
  void operator()(FourSpinorBlock *res, const FourSpinorBlock *in, int isign, int target_cb = 1) const override
  {
    void *r[1] = {res};
    const void *i[1] = {in};
    n_.apply(r, i, isign, target_cb);
  }
};

template <typename FT, int veclen, int soalen, bool compress12>
class EvenOddTMCloverOperator : public EvenOddLinearOperator<FT, veclen, soalen, compress12> {
  QPHIX_NATIVE_OP_COMMON
  typedef typename Geometry<FT, veclen, soalen, compress12>::FourSpinorBlock FourSpinorBlock;
  typedef typename Geometry<FT, veclen, soalen, compress12>::SU3MatrixBlock SU3MatrixBlock;
  typedef typename Geometry<FT, veclen, soalen, compress12>::CloverBlock CloverBlock;
  typedef typename Geometry<FT, veclen, soalen, compress12>::FullCloverBlock FullCloverBlock;
  // clov: A on the odd target; invclov[0/1]: (A +/- i mu g5)^-1 on the even sites
  EvenOddTMCloverOperator(SU3MatrixBlock *const u[2], CloverBlock *clov, FullCloverBlock *const invclov[2],
                          Geometry<FT, veclen, soalen, compress12> *geom, double t_boundary, double aniso_coeff_s,
                          double aniso_coeff_t, bool use_tbc[4] = nullptr, double tbc_phases[4][2] = nullptr,
                          double const prec_mass_rho = 0.0)
      : n_(geom, QPHIX_OP_TMCLOVER, t_boundary, aniso_coeff_s, aniso_coeff_t, use_tbc, tbc_phases)
  {
    n_.p_.prec_mass_rho = prec_mass_rho;
    n_.p_.twisted_mass = prec_mass_rho;
    setFields(u, clov, invclov);
  }
  void setFields(SU3MatrixBlock *const u[2], CloverBlock *clov, FullCloverBlock *const invclov[2])
  {
    n_.p_.u[0] = u[0];
    n_.p_.u[1] = u[1];
    n_.p_.clov[1] = clov;
    n_.p_.fullclov[0][0] = invclov[0];
    n_.p_.fullclov[0][1] = invclov[1];
    n_.create();
  }
  void operator()(FourSpinorBlock *res, const FourSpinorBlock *in, int isign, int target_cb = 1) const override
  {
    void *r[1] = {res};
    const void *i[1] = {in};
    n_.apply(r, i, isign, target_cb);
  }
};

template <typename FT, int veclen, int soalen, bool compress12>
class EvenOddNDTMWilsonReuseOperator : public TwoFlavEvenOddLinearOperator<FT, veclen, soalen, compress12> {
  QPHIX_NATIVE_OP_COMMON
  typedef typename Geometry<FT, veclen, soalen, compress12>::FourSpinorBlock FourSpinorBlock;
  typedef typename Geometry<FT, veclen, soalen, compress12>::SU3MatrixBlock SU3MatrixBlock;
  EvenOddNDTMWilsonReuseOperator(const double mass, const double mu, const double epsilon, SU3MatrixBlock *u[2],
                                 Geometry<FT, veclen, soalen, compress12> *geom, double t_boundary,
                                 double aniso_coeff_s, double aniso_coeff_t, bool use_tbc[4] = nullptr,
                                 double tbc_phases[4][2] = nullptr)
      : n_(geom, QPHIX_OP_NDTMWILSON, t_boundary, aniso_coeff_s, aniso_coeff_t, use_tbc, tbc_phases)
  {
    n_.p_.mass = mass;
    n_.p_.twisted_mass = mu;
    n_.p_.epsilon = epsilon;
    n_.p_.u[0] = u[0];
    n_.p_.u[1] = u[1];
    n_.create();
  }
  void operator()(FourSpinorBlock *const res[2], const FourSpinorBlock *const in[2], int isign,
                  int target_cb = 1) override
  {
    void *r[2] = {res[0], res[1]};
    const void *i[2] = {in[0], in[1]};
    n_.apply(r, i, isign, target_cb);
  }
};

template <typename FT, int veclen, int soalen, bool compress12>
class EvenOddNDTMCloverReuseOperator : public TwoFlavEvenOddLinearOperator<FT, veclen, soalen, compress12> {
  QPHIX_NATIVE_OP_COMMON
  typedef typename Geometry<FT, veclen, soalen, compress12>::FourSpinorBlock FourSpinorBlock;
  typedef typename Geometry<FT, veclen, soalen, compress12>::SU3MatrixBlock SU3MatrixBlock;
  typedef typename Geometry<FT, veclen, soalen, compress12>::CloverBlock CloverBlock;
  typedef typename Geometry<FT, veclen, soalen, compress12>::FullCloverBlock FullCloverBlock;
  EvenOddNDTMCloverReuseOperator(double TwistedMass, double Epsilon, SU3MatrixBlock *u[2], CloverBlock *clov,
                                 CloverBlock *invclov_odiag, FullCloverBlock *invclov[2],
                                 Geometry<FT, veclen, soalen, compress12> *geom, double t_boundary,
                                 double aniso_coeff_s, double aniso_coeff_t, bool use_tbc[4] = nullptr,
                                 double tbc_phases[4][2] = nullptr)
      : n_(geom, QPHIX_OP_NDTMCLOVER, t_boundary, aniso_coeff_s, aniso_coeff_t, use_tbc, tbc_phases)
  {
    n_.p_.twisted_mass = TwistedMass;
    n_.p_.epsilon = Epsilon;
    setFields(u, clov, invclov_odiag, invclov);
  }
  void setFields(SU3MatrixBlock *u[2], CloverBlock *clov, CloverBlock *invclov_odiag, FullCloverBlock *invclov[2])
  {
    n_.p_.u[0] = u[0];
    n_.p_.u[1] = u[1];
    n_.p_.clov[1] = clov;
    n_.p_.clov_odiag[0] = invclov_odiag;
    n_.p_.fullclov[0][0] = invclov[0];
    n_.p_.fullclov[0][1] = invclov[1];
    n_.create();
  }
  void operator()(FourSpinorBlock *const res[2], const FourSpinorBlock *const in[2], int isign,
                  int target_cb = 1) override
  {
    void *r[2] = {res[0], res[1]};
    const void *i[2] = {in[0], in[1]};
    n_.apply(r, i, isign, target_cb);
  }
};

#undef QPHIX_NATIVE_OP_COMMON

}  // namespace QPhiX
