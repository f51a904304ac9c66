// Chroma LinOpSysSolver surface (README.markdown:163-256): a reader for the
// <InvertParam> XML group of QPHIX_CLOVER_INVERTER and
// QPHIX_CLOVER_ITER_REFINE_BICGSTAB_INVERTER, and the odd-checkerboard clover
// system solver those parameters describe, built on the library's C ABI.
#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <string>
#include <utility>
#include <vector>

#include "../../include/qphix_amd.h"
#include "fields.hpp"

using qphix_amd::api_fail;

namespace {

int failf(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return api_fail(code, buf);
}
#define TRY(x)         \
  do {                 \
    int r_ = (x);      \
    if (r_) return r_; \
  } while (0)
#define CHKPTR(p) \
  do {            \
    if (!(p)) return failf(QPHIX_EINVAL, "%s: null argument '%s'", __func__, #p); \
  } while (0)

// ---------------------------------------------------------------------------
// XML: elements, attributes (read and ignored), character data with the five
// predefined and numeric character references, comments, processing
// instructions, CDATA sections and a DOCTYPE without an internal subset
// ---------------------------------------------------------------------------
struct XNode {
  std::string name, text;
  int line = 0;
  std::vector<XNode> kids;
};

class XReader {
 public:
  explicit XReader(const char *s) : s_(s), n_(strlen(s)) {}

  int document(XNode &root) {
    TRY(prolog());
    if (!at("<")) return err("expected the root element");
    TRY(element(root, 0));
    TRY(prolog());
    if (i_ < n_) return err("content after the root element");
    return QPHIX_OK;
  }

 private:
  static constexpr int kMaxDepth = 64;
  const char *s_;
  size_t n_, i_ = 0;
  int line_ = 1;

  bool at(const char *t) const { return strncmp(s_ + i_, t, strlen(t)) == 0; }
  void adv(size_t k) {
    for (size_t j = 0; j < k && i_ < n_; j++, i_++)
      if (s_[i_] == '\n') line_++;
  }
  void ws() {
    while (i_ < n_ && isspace((unsigned char)s_[i_])) adv(1);
  }
  int err(const char *what) { return failf(QPHIX_EINVAL, "InvertParam XML, line %d: %s", line_, what); }
  int skip_past(const char *end, const char *what) {
    const char *p = strstr(s_ + i_, end);
    if (!p) return err(what);
    adv((size_t)(p - (s_ + i_)) + strlen(end));
    return QPHIX_OK;
  }
  // one comment, processing instruction or DOCTYPE at the cursor, if any
  int misc(bool *any) {
    *any = true;
    if (at("<!--")) return skip_past("-->", "unterminated comment");
    if (at("<?")) return skip_past("?>", "unterminated processing instruction");
    if (at("<!DOCTYPE")) {
      const char *gt = strchr(s_ + i_, '>'), *br = strchr(s_ + i_, '[');
      if (br && (!gt || br < gt)) return err("DOCTYPE internal subsets are not supported");
      return skip_past(">", "unterminated DOCTYPE");
    }
    *any = false;
    return QPHIX_OK;
  }
  int prolog() {
    for (bool any = true; any;) {
      ws();
      TRY(misc(&any));
    }
    return QPHIX_OK;
  }
  static bool name_start(char c) { return isalpha((unsigned char)c) || c == '_' || c == ':'; }
  static bool name_char(char c) { return name_start(c) || isdigit((unsigned char)c) || c == '-' || c == '.'; }
  bool name(std::string &out) {
    if (i_ >= n_ || !name_start(s_[i_])) return false;
    size_t b = i_;
    while (i_ < n_ && name_char(s_[i_])) i_++;
    out.assign(s_ + b, i_ - b);
    return true;
  }
  int reference(std::string &out) {
    const char *semi = (const char *)memchr(s_ + i_, ';', n_ - i_ < 12 ? n_ - i_ : 12);
    if (!semi) return err("unterminated character reference");
    std::string ref(s_ + i_ + 1, semi);
    static const char *const names[5][2] = {{"lt", "<"}, {"gt", ">"}, {"amp", "&"}, {"quot", "\""}, {"apos", "'"}};
    bool ok = false;
    for (auto &nm : names)
      if (ref == nm[0]) {
        out += nm[1];
        ok = true;
      }
    if (!ok && ref.size() > 1 && ref[0] == '#') {
      char *end = nullptr;
      const bool hex = ref[1] == 'x';
      long v = strtol(ref.c_str() + (hex ? 2 : 1), &end, hex ? 16 : 10);
      if (*end || v <= 0 || v > 127) return err("character reference outside ASCII");
      out += (char)v;
      ok = true;
    }
    if (!ok) return err("unknown entity reference");
    adv((size_t)(semi - (s_ + i_)) + 1);
    return QPHIX_OK;
  }
  int element(XNode &e, int depth) {
    if (depth > kMaxDepth) return err("elements nested too deeply");
    adv(1);  // '<'
    e.line = line_;
    if (!name(e.name)) return err("'<' not followed by an element name");
    for (;;) {  // attributes
      ws();
      if (at("/>")) {
        adv(2);
        return QPHIX_OK;
      }
      if (at(">")) {
        adv(1);
        break;
      }
      std::string an;
      if (!name(an)) return err("malformed attribute");
      ws();
      if (!at("=")) return err("attribute without a value");
      adv(1);
      ws();
      if (!at("\"") && !at("'")) return err("attribute value not quoted");
      const char q[2] = {s_[i_], 0};
      adv(1);
      TRY(skip_past(q, "unterminated attribute value"));
    }
    for (;;) {  // content
      if (i_ >= n_) return failf(QPHIX_EINVAL, "InvertParam XML: element <%s> (line %d) is not closed",
                                 e.name.c_str(), e.line);
      if (at("</")) {
        adv(2);
        std::string nm;
        if (!name(nm) || nm != e.name)
          return err(("closing tag </" + nm + "> does not match <" + e.name + ">").c_str());
        ws();
        if (!at(">")) return err("malformed closing tag");
        adv(1);
        return QPHIX_OK;
      }
      if (at("<![CDATA[")) {
        adv(9);
        const char *p = strstr(s_ + i_, "]]>");
        if (!p) return err("unterminated CDATA section");
        e.text.append(s_ + i_, p);
        adv((size_t)(p - (s_ + i_)) + 3);
        continue;
      }
      bool any = false;
      TRY(misc(&any));
      if (any) continue;
      if (at("<")) {
        XNode k;
        TRY(element(k, depth + 1));
        e.kids.push_back(std::move(k));
        continue;
      }
      if (at("&")) {
        TRY(reference(e.text));
        continue;
      }
      e.text += s_[i_];
      adv(1);
    }
  }
};

const XNode *find_node(const XNode &n, const char *name) {
  if (n.name == name) return &n;
  for (const XNode &k : n.kids)
    if (const XNode *f = find_node(k, name)) return f;
  return nullptr;
}

// the unique child element `tag` of e, or null when absent
int child(const XNode &e, const char *tag, const XNode **out) {
  *out = nullptr;
  for (const XNode &k : e.kids)
    if (k.name == tag) {
      if (*out) return failf(QPHIX_EINVAL, "InvertParam XML, line %d: <%s> given twice in <%s>", k.line, tag,
                             e.name.c_str());
      *out = &k;
    }
  return QPHIX_OK;
}

int value(const XNode &e, const char *tag, std::string *v, bool *present, bool required) {
  const XNode *k = nullptr;
  TRY(child(e, tag, &k));
  *present = k != nullptr;
  if (!k) {
    if (required) return failf(QPHIX_EINVAL, "InvertParam XML: <%s> (line %d) has no <%s>", e.name.c_str(), e.line, tag);
    return QPHIX_OK;
  }
  if (!k->kids.empty()) return failf(QPHIX_EINVAL, "InvertParam XML, line %d: <%s> holds elements, not a value", k->line, tag);
  size_t b = 0, en = k->text.size();
  while (b < en && isspace((unsigned char)k->text[b])) b++;
  while (en > b && isspace((unsigned char)k->text[en - 1])) en--;
  *v = k->text.substr(b, en - b);
  if (v->empty()) return failf(QPHIX_EINVAL, "InvertParam XML, line %d: <%s> is empty", k->line, tag);
  return QPHIX_OK;
}

int read_double(const XNode &e, const char *tag, double *out, bool required, bool *present = nullptr) {
  std::string v;
  bool pr = false;
  TRY(value(e, tag, &v, &pr, required));
  if (present) *present = pr;
  if (!pr) return QPHIX_OK;
  char *end = nullptr;
  double d = strtod(v.c_str(), &end);
  if (*end || !isfinite(d)) return failf(QPHIX_EINVAL, "InvertParam XML: <%s> = '%s' is not a number", tag, v.c_str());
  *out = d;
  return QPHIX_OK;
}

int read_int(const XNode &e, const char *tag, int *out, bool required) {
  std::string v;
  bool pr = false;
  TRY(value(e, tag, &v, &pr, required));
  if (!pr) return QPHIX_OK;
  char *end = nullptr;
  long d = strtol(v.c_str(), &end, 10);
  if (*end || d < -2147483647L || d > 2147483647L)
    return failf(QPHIX_EINVAL, "InvertParam XML: <%s> = '%s' is not an integer", tag, v.c_str());
  *out = (int)d;
  return QPHIX_OK;
}

int read_bool(const XNode &e, const char *tag, int *out, bool required) {
  std::string v;
  bool pr = false;
  TRY(value(e, tag, &v, &pr, required));
  if (!pr) return QPHIX_OK;
  if (v == "true" || v == "1") *out = 1;
  else if (v == "false" || v == "0") *out = 0;
  else return failf(QPHIX_EINVAL, "InvertParam XML: <%s> = '%s' is not a boolean (true / false)", tag, v.c_str());
  return QPHIX_OK;
}

int read_prec(const XNode &e, const char *tag, int *out, const char *allowed) {
  std::string v;
  bool pr = false;
  TRY(value(e, tag, &v, &pr, false));
  if (!pr) return QPHIX_OK;
  if (v.size() != 1 || !strchr(allowed, v[0]))
    return failf(QPHIX_EDOMAIN, "InvertParam XML: <%s> = '%s', expected one of %s", tag, v.c_str(), allowed);
  *out = v[0] == 'd' ? QPHIX_F64 : v[0] == 'f' ? QPHIX_F32 : QPHIX_F16;
  return QPHIX_OK;
}

int check_param(const qphix_invert_param &p) {
  if (p.inv_type != QPHIX_INV_CLOVER && p.inv_type != QPHIX_INV_CLOVER_ITER_REFINE)
    return failf(QPHIX_EDOMAIN, "invert param: unknown inv_type %d", p.inv_type);
  if (p.solver_type != QPHIX_SOLVER_CG && p.solver_type != QPHIX_SOLVER_BICGSTAB)
    return failf(QPHIX_EDOMAIN, "invert param: unknown solver_type %d", p.solver_type);
  if (p.max_iter < 1) return failf(QPHIX_EDOMAIN, "invert param: MaxIter %d < 1", p.max_iter);
  if (!(p.rsd_target > 0)) return failf(QPHIX_EDOMAIN, "invert param: RsdTarget must be positive");
  if (p.inv_type == QPHIX_INV_CLOVER_ITER_REFINE && !(p.delta > 0 && p.delta < 1))
    return failf(QPHIX_EDOMAIN, "invert param: Delta %g outside (0, 1)", p.delta);
  if (p.t_dir != 3) return failf(QPHIX_EDOMAIN, "invert param: t_dir %d, only 3 is supported", p.t_dir);
  if (p.aniso_p && !(p.xi_0 > 0 && p.nu > 0)) return failf(QPHIX_EDOMAIN, "invert param: xi_0 and nu must be positive");
  if (!(p.rsd_tolerance_factor > 0)) return failf(QPHIX_EDOMAIN, "invert param: RsdToleranceFactor must be positive");
  if (p.ncores < 1 || p.threads_per_core < 1 || p.by < 1 || p.bz < 1 || p.sy < 1 || p.sz < 1 || p.pad_xy < 0 ||
      p.pad_xyz < 0 || p.min_ct < 1)
    return failf(QPHIX_EDOMAIN, "invert param: blocking parameters must be positive (pads non-negative)");
  if (p.prec != QPHIX_F64 && p.prec != QPHIX_F32)
    return failf(QPHIX_EDOMAIN, "invert param: the outer precision must be F64 or F32");
  if (p.inner_prec != QPHIX_F64 && p.inner_prec != QPHIX_F32 && p.inner_prec != QPHIX_F16)
    return failf(QPHIX_EDOMAIN, "invert param: unknown inner precision %d", p.inner_prec);
  if (p.inv_type == QPHIX_INV_CLOVER_ITER_REFINE && p.inner_prec == QPHIX_F16 && !p.inner_compress12)
    return failf(QPHIX_EDOMAIN, "invert param: a half-precision inner solver needs two-row compression");
  return QPHIX_OK;
}

void coeffs(const qphix_invert_param &p, double c[6]) {
  // tests/clover_term_qdp_w.h:156-180: the spatial clover coefficient and the
  // spatial hops carry 1/xi_0 and nu/xi_0 on an anisotropic lattice, and the
  // diagonal mass the hop terms' share 1 + (Nd-1) nu/xi_0
  const double ff = p.aniso_p ? p.nu / p.xi_0 : 1.0;
  c[0] = ff;
  c[1] = 1.0;
  c[2] = p.antiperiodic_t ? -1.0 : 1.0;
  c[3] = p.aniso_p ? p.clov_coeff_r / p.xi_0 : p.clov_coeff_r;
  c[4] = p.clov_coeff_t;
  c[5] = 1.0 + 3.0 * ff + p.mass;
}

}  // namespace

extern "C" int qphix_invert_param_default(qphix_invert_param *p) {
  CHKPTR(p);
  memset(p, 0, sizeof(*p));
  p->inv_type = QPHIX_INV_CLOVER;
  p->solver_type = QPHIX_SOLVER_BICGSTAB;
  p->max_iter = 1000;
  p->rsd_target = 1e-7;
  p->delta = 0.1;
  p->clov_coeff_r = p->clov_coeff_t = 1.0;
  p->t_dir = 3;
  p->xi_0 = p->nu = 1.0;
  p->rsd_tolerance_factor = 5.0;
  p->ncores = p->threads_per_core = p->by = p->bz = p->sy = p->sz = p->min_ct = 1;
  p->prec = QPHIX_F64;
  p->inner_prec = QPHIX_F32;
  return QPHIX_OK;
}

extern "C" int qphix_invert_param_read_xml(const char *xml, qphix_invert_param *out) {
  CHKPTR(xml);
  CHKPTR(out);
  XNode doc;
  TRY(XReader(xml).document(doc));
  const XNode *ip = find_node(doc, "InvertParam");
  if (!ip) return failf(QPHIX_EINVAL, "InvertParam XML: no <InvertParam> element");
  qphix_invert_param p;
  qphix_invert_param_default(&p);
  bool pr = false;
  std::string v;
  TRY(value(*ip, "invType", &v, &pr, true));
  if (v == "QPHIX_CLOVER_INVERTER") p.inv_type = QPHIX_INV_CLOVER;
  else if (v == "QPHIX_CLOVER_ITER_REFINE_BICGSTAB_INVERTER") p.inv_type = QPHIX_INV_CLOVER_ITER_REFINE;
  else
    return failf(QPHIX_EDOMAIN, "InvertParam XML: invType '%s' is not QPHIX_CLOVER_INVERTER or "
                 "QPHIX_CLOVER_ITER_REFINE_BICGSTAB_INVERTER", v.c_str());
  TRY(value(*ip, "SolverType", &v, &pr, false));
  if (pr) {
    if (v == "CG") p.solver_type = QPHIX_SOLVER_CG;
    else if (v == "BICGSTAB") p.solver_type = QPHIX_SOLVER_BICGSTAB;
    else return failf(QPHIX_EDOMAIN, "InvertParam XML: SolverType '%s' is not CG or BICGSTAB", v.c_str());
  }
  TRY(read_int(*ip, "MaxIter", &p.max_iter, true));
  TRY(read_double(*ip, "RsdTarget", &p.rsd_target, true));
  TRY(read_double(*ip, "Delta", &p.delta, false));
  TRY(read_double(*ip, "RsdToleranceFactor", &p.rsd_tolerance_factor, false));
  TRY(read_bool(*ip, "AntiPeriodicT", &p.antiperiodic_t, false));
  TRY(read_bool(*ip, "Verbose", &p.verbose, false));
  TRY(read_bool(*ip, "Tune", &p.tune, false));
  TRY(read_int(*ip, "NCores", &p.ncores, false));
  TRY(read_int(*ip, "ThreadsPerCore", &p.threads_per_core, false));
  TRY(read_int(*ip, "By", &p.by, false));
  TRY(read_int(*ip, "Bz", &p.bz, false));
  TRY(read_int(*ip, "Sy", &p.sy, false));
  TRY(read_int(*ip, "Sz", &p.sz, false));
  TRY(read_int(*ip, "PadXY", &p.pad_xy, false));
  TRY(read_int(*ip, "PadXYZ", &p.pad_xyz, false));
  TRY(read_int(*ip, "MinCt", &p.min_ct, false));
  TRY(read_prec(*ip, "OuterPrecision", &p.prec, "df"));
  TRY(read_prec(*ip, "InnerPrecision", &p.inner_prec, "hfd"));
  int c12 = -1;
  TRY(read_bool(*ip, "Compress12", &c12, false));
  if (c12 >= 0) p.compress12 = p.inner_compress12 = c12;
  TRY(read_bool(*ip, "InnerCompress12", &p.inner_compress12, false));

  const XNode *cp = nullptr;
  TRY(child(*ip, "CloverParams", &cp));
  if (!cp) return failf(QPHIX_EINVAL, "InvertParam XML: no <CloverParams>");
  bool has_mass = false, has_kappa = false;
  double kappa = 0;
  TRY(read_double(*cp, "Mass", &p.mass, false, &has_mass));
  TRY(read_double(*cp, "Kappa", &kappa, false, &has_kappa));
  if (has_mass && has_kappa) return failf(QPHIX_EDOMAIN, "InvertParam XML: <CloverParams> gives both Mass and Kappa");
  if (!has_mass && !has_kappa) return failf(QPHIX_EINVAL, "InvertParam XML: <CloverParams> gives neither Mass nor Kappa");
  if (has_kappa) {
    if (!(kappa > 0)) return failf(QPHIX_EDOMAIN, "InvertParam XML: Kappa must be positive");
    p.mass = 1.0 / (2.0 * kappa) - 4.0;
  }
  // <clovCoeff> sets both planes; <clovCoeffR> / <clovCoeffT> set their own
  bool has_c = false, has_r = false, has_t = false;
  double c = 0;
  TRY(read_double(*cp, "clovCoeff", &c, false, &has_c));
  if (has_c) p.clov_coeff_r = p.clov_coeff_t = c;
  TRY(read_double(*cp, "clovCoeffR", &p.clov_coeff_r, false, &has_r));
  TRY(read_double(*cp, "clovCoeffT", &p.clov_coeff_t, false, &has_t));
  if (!has_c && !(has_r && has_t))
    return failf(QPHIX_EINVAL, "InvertParam XML: <CloverParams> needs clovCoeff or both clovCoeffR and clovCoeffT");
  const XNode *ap = nullptr;
  TRY(child(*cp, "AnisoParam", &ap));
  if (ap) {
    TRY(read_bool(*ap, "anisoP", &p.aniso_p, true));
    TRY(read_int(*ap, "t_dir", &p.t_dir, false));
    TRY(read_double(*ap, "xi_0", &p.xi_0, p.aniso_p != 0));
    TRY(read_double(*ap, "nu", &p.nu, p.aniso_p != 0));
  }
  TRY(check_param(p));
  *out = p;
  return QPHIX_OK;
}

extern "C" int qphix_invert_param_coeffs(const qphix_invert_param *p, double out[6]) {
  CHKPTR(p);
  CHKPTR(out);
  TRY(check_param(*p));
  coeffs(*p, out);
  return QPHIX_OK;
}

// ---------------------------------------------------------------------------
// the system solver
// ---------------------------------------------------------------------------
struct qphix_syssolver {
  qphix_invert_param p;
  double c[6];
  long vh = 0;
  qphix_geom *g = nullptr, *gi = nullptr;
  void *u[2] = {nullptr, nullptr}, *A = nullptr, *Ainv = nullptr;
  void *ui[2] = {nullptr, nullptr}, *Ai = nullptr, *Ainvi = nullptr;
  qphix_op *M = nullptr, *Mi = nullptr;
  qphix_solver *inner = nullptr, *solver = nullptr;
  void *x = nullptr, *b = nullptr, *t = nullptr, *r = nullptr;
};

extern "C" int qphix_syssolver_destroy(qphix_syssolver *s) {
  if (!s) return QPHIX_OK;
  qphix_solver_destroy(s->solver);
  qphix_solver_destroy(s->inner);
  qphix_op_destroy(s->M);
  qphix_op_destroy(s->Mi);
  void *f[] = {s->u[0], s->u[1], s->A, s->Ainv, s->ui[0], s->ui[1], s->Ai, s->Ainvi, s->x, s->b, s->t, s->r};
  for (void *q : f) qphix_free(q);
  qphix_geom_destroy(s->g);
  qphix_geom_destroy(s->gi);
  delete s;
  return QPHIX_OK;
}

namespace {

// gauge field and clover blocks of one geometry
struct CloverFields {
  qphix_geom *g = nullptr;
  void *u[2] = {nullptr, nullptr}, *A = nullptr, *Ainv = nullptr;
  void release() {
    void *f[] = {u[0], u[1], A, Ainv};
    for (void *q : f) qphix_free(q);
    qphix_geom_destroy(g);
    *this = CloverFields();
  }
  int alloc(const int latt[4], int prec, int c12) {
    TRY(qphix_geom_create(latt, prec, c12, &g));
    TRY(qphix_alloc_gauge(g, &u[0]));
    TRY(qphix_alloc_gauge(g, &u[1]));
    TRY(qphix_alloc_clover(g, &A));
    TRY(qphix_alloc_clover(g, &Ainv));
    return QPHIX_OK;
  }
  int convert_from(const CloverFields &src) {
    for (int cb = 0; cb < 2; cb++) TRY(qphix_convert_gauge(g, u[cb], src.g, src.u[cb]));
    TRY(qphix_convert_clover(g, A, src.g, src.A, 0));
    TRY(qphix_convert_clover(g, Ainv, src.g, src.Ainv, 0));
    return QPHIX_OK;
  }
};

int clover_op(const CloverFields &f, const double c[6], qphix_op **op) {
  qphix_op_params p;
  memset(&p, 0, sizeof(p));
  p.kind = QPHIX_OP_CLOVER;
  p.hop.aniso_s = c[0];
  p.hop.aniso_t = c[1];
  p.hop.t_boundary = c[2];
  p.u[0] = f.u[0];
  p.u[1] = f.u[1];
  p.clov[1] = f.A;      // A_oo on the odd target checkerboard
  p.invclov[0] = f.Ainv;  // A_ee^-1 on the even one
  return qphix_op_create(f.g, &p, op);
}

// one application of each sign so the autotuner runs before the first solve
int warm_up(qphix_op *M, const qphix_geom *g) {
  void *a = nullptr, *b = nullptr;
  int rc = qphix_alloc_spinor(g, &a);
  if (!rc) rc = qphix_alloc_spinor(g, &b);
  for (int isign = 1; !rc && isign >= -1; isign -= 2) {
    void *o[1] = {b};
    const void *in[1] = {a};
    rc = qphix_op_apply(M, o, in, isign, 1);
  }
  if (!rc) rc = qphix_sync();
  qphix_free(a);
  qphix_free(b);
  return rc;
}

int build(qphix_syssolver *s, const int latt[4], const double *u) {
  const qphix_invert_param &p = s->p;
  // the clover term is built in FP64 from the uncompressed gauge field and
  // then converted to the solver precisions
  CloverFields f64;
  struct Guard {
    CloverFields *f;
    ~Guard() { f->release(); }
  } guard{&f64};
  TRY(f64.alloc(latt, QPHIX_F64, 0));
  TRY(qphix_pack_gauge(f64.g, f64.u[0], f64.u[1], u));
  TRY(qphix_clover_build(f64.g, f64.u[0], f64.u[1], s->c[3], s->c[4], s->c[5], 1, QPHIX_CLOV_A, 0, 0, f64.A));
  TRY(qphix_clover_build(f64.g, f64.u[0], f64.u[1], s->c[3], s->c[4], s->c[5], 0, QPHIX_CLOV_AINV, 0, 0, f64.Ainv));
  TRY(qphix_geom_info(f64.g, nullptr, nullptr, nullptr, &s->vh, nullptr));

  CloverFields fo, fi;
  const bool refine = p.inv_type == QPHIX_INV_CLOVER_ITER_REFINE;
  if (refine) {
    int rc = fi.alloc(latt, p.inner_prec, p.inner_compress12);
    if (!rc) rc = fi.convert_from(f64);
    s->gi = fi.g;
    s->ui[0] = fi.u[0];
    s->ui[1] = fi.u[1];
    s->Ai = fi.A;
    s->Ainvi = fi.Ainv;
    TRY(rc);
  }
  if (p.prec == QPHIX_F64 && !p.compress12) {
    fo = f64;
    f64 = CloverFields();
  } else {
    int rc = fo.alloc(latt, p.prec, p.compress12);
    if (!rc) rc = fo.convert_from(f64);
    if (rc) fo.release();
    TRY(rc);
  }
  s->g = fo.g;
  s->u[0] = fo.u[0];
  s->u[1] = fo.u[1];
  s->A = fo.A;
  s->Ainv = fo.Ainv;
  TRY(qphix_sync());
  f64.release();

  TRY(clover_op(fo, s->c, &s->M));
  if (refine) {
    TRY(clover_op(fi, s->c, &s->Mi));
    if (p.solver_type == QPHIX_SOLVER_CG) TRY(qphix_cg_create(s->Mi, p.max_iter, &s->inner));
    else TRY(qphix_bicgstab_create(s->Mi, p.max_iter, &s->inner));
    TRY(qphix_richardson_create(s->M, s->inner, p.delta, p.max_iter, p.solver_type == QPHIX_SOLVER_CG, &s->solver));
  } else if (p.solver_type == QPHIX_SOLVER_CG) {
    TRY(qphix_cg_create(s->M, p.max_iter, &s->solver));
  } else {
    TRY(qphix_bicgstab_create(s->M, p.max_iter, &s->solver));
  }
  void **w[4] = {&s->x, &s->b, &s->t, &s->r};
  for (void **q : w) TRY(qphix_alloc_spinor(s->g, q));
  if (p.tune) {
    TRY(warm_up(s->M, s->g));
    if (refine) TRY(warm_up(s->Mi, s->gi));
  }
  return QPHIX_OK;
}

}  // namespace

extern "C" int qphix_syssolver_create(const int latt[4], const qphix_invert_param *p, const double *u,
                                      qphix_syssolver **out) {
  CHKPTR(latt);
  CHKPTR(p);
  CHKPTR(u);
  CHKPTR(out);
  *out = nullptr;
  TRY(check_param(*p));
  qphix_syssolver *s = new qphix_syssolver();
  s->p = *p;
  coeffs(*p, s->c);
  int rc = build(s, latt, u);
  if (rc) {
    std::string msg = qphix_last_error();
    qphix_syssolver_destroy(s);
    return api_fail(rc, msg.c_str());
  }
  *out = s;
  return QPHIX_OK;
}

extern "C" int qphix_syssolver_geom(const qphix_syssolver *s, const qphix_geom **g) {
  CHKPTR(s);
  CHKPTR(g);
  *g = s->g;
  return QPHIX_OK;
}

extern "C" int qphix_syssolver_solve_dev(qphix_syssolver *s, void *psi, const void *chi,
                                         qphix_syssolver_results *res) {
  CHKPTR(s);
  CHKPTR(psi);
  CHKPTR(chi);
  CHKPTR(res);
  memset(res, 0, sizeof(*res));
  const qphix_invert_param &p = s->p;
  TRY(qphix_sync());
  const auto t0 = std::chrono::steady_clock::now();
  unsigned long extra = 0;
  const void *rhs = chi;
  if (p.solver_type == QPHIX_SOLVER_CG) {
    // normal equations: M^+M psi = M^+ chi
    void *o[1] = {s->t};
    const void *in[1] = {chi};
    TRY(qphix_op_apply(s->M, o, in, -1, 1));
    rhs = s->t;
    extra++;
  }
  qphix_solve_stats st;
  void *xs[1] = {psi};
  const void *bs[1] = {rhs};
  TRY(qphix_solve(s->solver, xs, bs, p.rsd_target, 1, 0, 1, QPHIX_RESID_RELATIVE, &st));
  // true residual of M psi = chi in the outer precision, FP64 reductions
  void *o[1] = {s->t};
  const void *in[1] = {psi};
  TRY(qphix_op_apply(s->M, o, in, 1, 1));
  extra++;
  double rn2 = 0, bn2 = 0;
  TRY(qphix_xmy_norm2(s->g, s->r, chi, s->t, &rn2));
  TRY(qphix_norm2_spinor(s->g, chi, &bn2));
  const auto t1 = std::chrono::steady_clock::now();
  res->n_count = st.n_iters;
  res->resid = sqrt(rn2);
  res->rel_resid = bn2 > 0 ? sqrt(rn2 / bn2) : sqrt(rn2);
  res->mv_apps = st.mv_apps + extra;
  res->seconds = std::chrono::duration<double>(t1 - t0).count();
  res->status = st.status;
  if (res->status == QPHIX_OK && !(res->rel_resid <= p.rsd_target * p.rsd_tolerance_factor))
    res->status = QPHIX_ENOCONV;
  if (p.verbose)
    fprintf(stderr, "QPHIX_CLOVER_%s%s_SOLVER: %d iterations, |r| / |b| = %.6e, %.4f s%s\n",
            p.inv_type == QPHIX_INV_CLOVER_ITER_REFINE ? "ITER_REFINE_" : "",
            p.solver_type == QPHIX_SOLVER_CG ? "CG" : "BICGSTAB", res->n_count, res->rel_resid, res->seconds,
            res->status == QPHIX_OK ? "" : " -- SOLVE FAILED");
  return QPHIX_OK;
}

extern "C" int qphix_syssolver_solve(qphix_syssolver *s, double *psi, const double *chi,
                                     qphix_syssolver_results *res) {
  CHKPTR(s);
  CHKPTR(psi);
  CHKPTR(chi);
  CHKPTR(res);
  const long off = s->vh * 24;  // the odd checkerboard's half of the canonical arrays
  TRY(qphix_pack_spinor(s->g, s->b, chi + off, 1));
  TRY(qphix_pack_spinor(s->g, s->x, psi + off, 1));
  TRY(qphix_syssolver_solve_dev(s, s->x, s->b, res));
  TRY(qphix_unpack_spinor(s->g, psi + off, s->x, 1));
  return QPHIX_OK;
}
