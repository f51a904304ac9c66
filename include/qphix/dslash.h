// Drop-in Dslash classes for the gfx950 library: Dslash (dslash_def.h:16-172),
// ClovDslash (clover_dslash_def.h:13-178), TMDslash (tm_dslash_def.h:11-235)
// and TMClovDslash (tm_clov_dslash_def.h:10-317) of JeffersonLab/qphix. Each
// method launches one stream-ordered HIP kernel through the C ABI.
#pragma once

#include "geometry.h"

namespace QPhiX {

namespace detail {
inline qphix_dslash_params hop_params(double t_boundary, double aniso_s, double aniso_t, const bool use_tbc[4],
                                      const double tbc_phases[4][2])
{
  qphix_dslash_params p{};
  p.t_boundary = t_boundary;
  p.aniso_s = aniso_s;
  p.aniso_t = aniso_t;
  if (use_tbc != nullptr && tbc_phases != nullptr)
    for (int d = 0; d < 4; d++) {
      p.use_tbc[d] = use_tbc[d] ? 1 : 0;
      p.tbc_phases[d][0] = tbc_phases[d][0];
      p.tbc_phases[d][1] = tbc_phases[d][1];
    }
  return p;
}

template <typename FT, int V, int S, bool C> class DslashBase {
 public:
  typedef typename Geometry<FT, V, S, C>::FourSpinorBlock FourSpinorBlock;
  typedef typename Geometry<FT, V, S, C>::SU3MatrixBlock SU3MatrixBlock;
  typedef typename Geometry<FT, V, S, C>::CloverBlock CloverBlock;
  typedef typename Geometry<FT, V, S, C>::FullCloverBlock FullCloverBlock;

  DslashBase(Geometry<FT, V, S, C> *geom, const qphix_dslash_params &p) : geom_(geom)
  {
    check(qphix_dslash_create(geom->handle(), &p, &d_), "Dslash");
  }
  ~DslashBase() { qphix_dslash_destroy(d_); }
  DslashBase(const DslashBase &) = delete;
  DslashBase &operator=(const DslashBase &) = delete;
  Geometry<FT, V, S, C> &getGeometry() { return *geom_; }
  qphix_dslash *handle() const { return d_; }

 protected:
  Geometry<FT, V, S, C> *geom_;
  qphix_dslash *d_ = nullptr;
};
}  // namespace detail

template <typename FT, int veclen, int soalen, bool compress12>
class Dslash : public detail::DslashBase<FT, veclen, soalen, compress12> {
  typedef detail::DslashBase<FT, veclen, soalen, compress12> B;

 public:
  using typename B::FourSpinorBlock;
  using typename B::SU3MatrixBlock;
  Dslash(Geometry<FT, veclen, soalen, compress12> *geom, double t_boundary, double aniso_coeff_S,
         double aniso_coeff_T, bool use_tbc[4] = nullptr, double tbc_phases[4][2] = nullptr)
      : B(geom, detail::hop_params(t_boundary, aniso_coeff_S, aniso_coeff_T, use_tbc, tbc_phases)) {}

  void dslash(FourSpinorBlock *res, const FourSpinorBlock *psi, const SU3MatrixBlock *u, int isign, int cb)
  {
    detail::check(qphix_wilson_dslash(this->d_, res, psi, u, isign, cb), "Dslash::dslash");
  }
  void dslashAChiMinusBDPsi(FourSpinorBlock *res, const FourSpinorBlock *psi, const FourSpinorBlock *chi,
                            const SU3MatrixBlock *u, double alpha, double beta, int isign, int cb)
  {
    detail::check(qphix_wilson_achimbdpsi(this->d_, res, psi, chi, u, alpha, beta, isign, cb),
                  "Dslash::dslashAChiMinusBDPsi");
  }
  // dslash_def.h:54-58: one term of the hop, isign +1
  void dslashDir(FourSpinorBlock *res, const FourSpinorBlock *psi, const SU3MatrixBlock *u, int cb, int dir)
  {
    detail::check(qphix_wilson_dslash_dir(this->d_, res, psi, u, cb, dir), "Dslash::dslashDir");
  }
};

template <typename FT, int veclen, int soalen, bool compress12>
class ClovDslash : public detail::DslashBase<FT, veclen, soalen, compress12> {
  typedef detail::DslashBase<FT, veclen, soalen, compress12> B;
  static qphix_dslash_params params(double tb, double as, double at, bool *ut, double (*tp)[2], double rho)
  {
    qphix_dslash_params p = detail::hop_params(tb, as, at, ut, tp);
    p.prec_mass_rho = rho;
    return p;
  }

 public:
  using typename B::CloverBlock;
  using typename B::FourSpinorBlock;
  using typename B::SU3MatrixBlock;
  ClovDslash(Geometry<FT, veclen, soalen, compress12> *geom, double t_boundary, double dslash_aniso_s,
             double dslash_aniso_t, bool use_tbc[4] = nullptr, double tbc_phases[4][2] = nullptr,
             double const prec_mass_rho = 0.0)
      : B(geom, params(t_boundary, dslash_aniso_s, dslash_aniso_t, use_tbc, tbc_phases, prec_mass_rho)) {}

  void dslash(FourSpinorBlock *res, const FourSpinorBlock *psi, const SU3MatrixBlock *u,
              const CloverBlock *invclov, int isign, int cb)
  {
    detail::check(qphix_clov_dslash(this->d_, res, psi, u, invclov, isign, cb), "ClovDslash::dslash");
  }
  void dslashAChiMinusBDPsi(FourSpinorBlock *res, const FourSpinorBlock *psi, const FourSpinorBlock *chi,
                            const SU3MatrixBlock *u, const CloverBlock *clov, double beta, int isign, int cb)
  {
    detail::check(qphix_clov_achimbdpsi(this->d_, res, psi, chi, u, clov, beta, isign, cb),
                  "ClovDslash::dslashAChiMinusBDPsi");
  }
  void clovMult(FourSpinorBlock *res, const FourSpinorBlock *chi, const CloverBlock *clov)
  {
    detail::check(qphix_clov_mult(this->d_, res, chi, clov), "ClovDslash::clovMult");
  }
};

template <typename FT, int veclen, int soalen, bool compress12>
class TMDslash : public detail::DslashBase<FT, veclen, soalen, compress12> {
  typedef detail::DslashBase<FT, veclen, soalen, compress12> B;
  static qphix_dslash_params params(double tb, double as, double at, double m, double mu, bool *ut,
                                    double (*tp)[2])
  {
    qphix_dslash_params p = detail::hop_params(tb, as, at, ut, tp);
    p.mass = m;
    p.twisted_mass = mu;
    return p;
  }

 public:
  using typename B::FourSpinorBlock;
  using typename B::SU3MatrixBlock;
  TMDslash(Geometry<FT, veclen, soalen, compress12> *geom, double t_boundary, double aniso_coeff_S,
           double aniso_coeff_T, double Mass, double TwistedMass, bool use_tbc[4] = nullptr,
           double tbc_phases[4][2] = nullptr)
      : B(geom, params(t_boundary, aniso_coeff_S, aniso_coeff_T, Mass, TwistedMass, use_tbc, tbc_phases)) {}

  void dslash(FourSpinorBlock *res, const FourSpinorBlock *psi, const SU3MatrixBlock *u, int isign, int cb)
  {
    detail::check(qphix_tm_dslash(this->d_, res, psi, u, isign, cb), "TMDslash::dslash");
  }
  void dslashAChiMinusBDPsi(FourSpinorBlock *res, const FourSpinorBlock *psi, const FourSpinorBlock *chi,
                            const SU3MatrixBlock *u, double alpha, double beta, int isign, int cb,
                            const double musign = 1.0)
  {
    detail::check(qphix_tm_achimbdpsi(this->d_, res, psi, chi, u, alpha, beta, isign, cb, musign),
                  "TMDslash::dslashAChiMinusBDPsi");
  }
  void two_flav_AChiMinusBDPsi(FourSpinorBlock *const res[2], const FourSpinorBlock *const psi[2],
                               const FourSpinorBlock *const chi[2], const SU3MatrixBlock *u, double alpha,
                               double beta, double epsilon, int isign, int cb)
  {
    void *r[2] = {res[0], res[1]};
    const void *p[2] = {psi[0], psi[1]};
    const void *c[2] = {chi[0], chi[1]};
    detail::check(qphix_tm_two_flav_achimbdpsi(this->d_, r, p, c, u, alpha, beta, epsilon, isign, cb),
                  "TMDslash::two_flav_AChiMinusBDPsi");
  }
};

template <typename FT, int veclen, int soalen, bool compress12>
class TMClovDslash : public detail::DslashBase<FT, veclen, soalen, compress12> {
  typedef detail::DslashBase<FT, veclen, soalen, compress12> B;
  static qphix_dslash_params params(double tb, double as, double at, bool *ut, double (*tp)[2], double rho)
  {
    qphix_dslash_params p = detail::hop_params(tb, as, at, ut, tp);
    p.prec_mass_rho = rho;
    return p;
  }

 public:
  using typename B::CloverBlock;
  using typename B::FourSpinorBlock;
  using typename B::FullCloverBlock;
  using typename B::SU3MatrixBlock;
  TMClovDslash(Geometry<FT, veclen, soalen, compress12> *geom, double t_boundary, double dslash_aniso_s,
               double dslash_aniso_t, bool use_tbc[4] = nullptr, double tbc_phases[4][2] = nullptr,
               double const prec_mass_rho = 0.0)
      : B(geom, params(t_boundary, dslash_aniso_s, dslash_aniso_t, use_tbc, tbc_phases, prec_mass_rho)) {}

  void dslash(FourSpinorBlock *res, const FourSpinorBlock *psi, const SU3MatrixBlock *u,
              const FullCloverBlock *const invclov[2], int isign, int cb, int fl = 0)
  {
    const void *ic[2] = {invclov[0], invclov[1]};
    detail::check(qphix_tmclov_dslash(this->d_, res, psi, u, ic, isign, cb, fl), "TMClovDslash::dslash");
  }
  void dslashAChiMinusBDPsi(FourSpinorBlock *res, const FourSpinorBlock *psi, const FourSpinorBlock *chi,
                            const SU3MatrixBlock *u, const CloverBlock *clov, double beta, int isign, int cb,
                            int fl = 0)
  {
    detail::check(qphix_tmclov_achimbdpsi(this->d_, res, psi, chi, u, clov, beta, isign, cb, fl),
                  "TMClovDslash::dslashAChiMinusBDPsi");
  }
  void two_flav_AChiMinusBDPsi(FourSpinorBlock *const res[2], const FourSpinorBlock *const psi[2],
                               const FourSpinorBlock *const chi[2], const SU3MatrixBlock *u,
                               const CloverBlock *clov, double beta, double epsilon, int isign, int cb)
  {
    void *r[2] = {res[0], res[1]};
    const void *p[2] = {psi[0], psi[1]};
    const void *c[2] = {chi[0], chi[1]};
    detail::check(qphix_tmclov_two_flav_achimbdpsi(this->d_, r, p, c, u, clov, beta, epsilon, isign, cb),
                  "TMClovDslash::two_flav_AChiMinusBDPsi");
  }
  void two_flav_inverse_clover_term(FourSpinorBlock *const res[2], const FourSpinorBlock *const psi[2],
                                    const FullCloverBlock *const fcl[2], const CloverBlock *const clOffDiag,
                                    int isign)
  {
    void *r[2] = {res[0], res[1]};
    const void *p[2] = {psi[0], psi[1]};
    const void *f[2] = {fcl[0], fcl[1]};
    detail::check(qphix_tmclov_two_flav_inverse_clover_term(this->d_, r, p, f, clOffDiag, isign),
                  "TMClovDslash::two_flav_inverse_clover_term");
  }
};

}  // namespace QPhiX
