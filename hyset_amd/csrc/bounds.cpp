// Similarity bounds as integer tables.
//
// Jaccard in HS_BOUNDS_REFERENCE mode evaluates the reference's IEEE-double
// expressions (include/hyset/similarity.hpp:40-50) exactly as written:
//   minoverlap = (unsigned) ceil((l1 + l2) * t / (1 + t))
//   minsize    = (unsigned) ceil(t * l)
//   maxsize    = (unsigned) (l / t)
// then the generic clamps (similarity.hpp:13-23):
//   minoverlap = min(l1, l2, minoverlap)
//   maxprefix  = min(l, l - minsize + 1)
//   midprefix  = min(l, l - minoverlap(l, l) + 1)
// These differ from exact arithmetic at t = 0.8 and 0.9 for some l1 + l2
// (the double quotient rounds above an integer), and the join reproduces
// that on purpose.  Every other combination uses t = p/q as an exact
// rational (shortest decimal that round-trips to t) and integer arithmetic:
//   Jaccard  minoverlap = ceil(p(l1+l2)/(p+q))   minsize = ceil(pl/q)          maxsize = floor(ql/p)
//   Dice     minoverlap = ceil(p(l1+l2)/(2q))    minsize = ceil(pl/(2q-p))     maxsize = floor((2q-p)l/p)
//   cosine   minoverlap = min o: o^2 q^2 >= p^2 l1 l2
//                                                 minsize = ceil(p^2 l/q^2)     maxsize = floor(q^2 l/p^2)
// The file compiles with -ffp-contract=off so the double expressions are
// never fused.
#include <cmath>
#include <cstdint>

#include "hs_internal.h"

namespace hs {

using u128 = unsigned __int128;

static int imin(int a, int b) { return a > b ? b : a; }

bool threshold_rational(double t, uint64_t* p, uint64_t* q) {
  if (!(t > 0.0) || t > 1.0) return false;
  uint64_t den = 1;
  for (int k = 0; k <= 9; ++k, den *= 10) {
    uint64_t num = (uint64_t)std::llround(t * (double)den);
    if ((double)num / (double)den == t) {
      uint64_t a = num, b = den;
      while (b) {
        uint64_t r = a % b;
        a = b;
        b = r;
      }
      *p = num / a;
      *q = den / a;
      return true;
    }
  }
  return false;
}

static uint64_t cdiv(u128 a, u128 b) { return (uint64_t)((a + b - 1) / b); }

static bool use_double(int kind, int mode) { return kind == HS_JACCARD && mode == HS_BOUNDS_REFERENCE; }

uint32_t bound_minsize(int kind, int mode, double t, uint64_t p, uint64_t q, uint32_t l) {
  if (use_double(kind, mode)) return (unsigned int)(std::ceil(t * l));
  switch (kind) {
    case HS_JACCARD: return (uint32_t)cdiv((u128)p * l, q);
    case HS_DICE: return (uint32_t)cdiv((u128)p * l, 2 * q - p);
    default: return (uint32_t)cdiv((u128)p * p * l, (u128)q * q);
  }
}

uint32_t bound_maxsize(int kind, int mode, double t, uint64_t p, uint64_t q, uint32_t l) {
  if (use_double(kind, mode)) return (unsigned int)((l / t));
  u128 v;
  switch (kind) {
    case HS_JACCARD: v = (u128)q * l / p; break;
    case HS_DICE: v = (u128)(2 * q - p) * l / p; break;
    default: v = (u128)q * q * l / ((u128)p * p); break;
  }
  return v > 0xFFFFFFFFu ? 0xFFFFFFFFu : (uint32_t)v;
}

uint32_t bound_minoverlap_raw(int kind, int mode, double t, uint64_t p, uint64_t q, uint32_t l1, uint32_t l2) {
  if (use_double(kind, mode)) return (unsigned int)(std::ceil((l1 + l2) * t / (1 + t)));
  switch (kind) {
    case HS_JACCARD: return (uint32_t)cdiv((u128)p * ((uint64_t)l1 + l2), p + q);
    case HS_DICE: return (uint32_t)cdiv((u128)p * ((uint64_t)l1 + l2), 2 * q);
    default: {
      u128 need = (u128)p * p * l1 * l2;
      u128 qq = (u128)q * q;
      double est = (double)p / (double)q * std::sqrt((double)l1 * (double)l2);
      uint64_t o = est > 2.0 ? (uint64_t)est - 2 : 0;
      while ((u128)o * o * qq < need) ++o;
      while (o > 0 && (u128)(o - 1) * (o - 1) * qq >= need) --o;
      return (uint32_t)o;
    }
  }
}

void BoundTables::build(int kind_, int mode_, double t_, uint32_t L_) {
  uint64_t p = 1, q = 1;
  bool rational = threshold_rational(t_, &p, &q);
  HS_REQUIRE(t_ > 0.0 && t_ <= 1.0, HS_E_INVALID, "threshold must be in (0, 1]");
  HS_REQUIRE(kind_ == HS_JACCARD || kind_ == HS_COSINE || kind_ == HS_DICE, HS_E_INVALID, "unknown similarity");
  HS_REQUIRE(mode_ == HS_BOUNDS_REFERENCE || mode_ == HS_BOUNDS_EXACT, HS_E_INVALID, "unknown bounds mode");
  HS_REQUIRE(rational || use_double(kind_, mode_), HS_E_INVALID,
             "threshold has no short decimal form; exact bounds need one");
  HS_REQUIRE(kind_ != HS_COSINE || L_ <= 4096, HS_E_UNSUPPORTED, "cosine tables cover set sizes up to 4096");
  kind = kind_;
  mode = mode_;
  t = t_;
  L = L_;
  minsize.assign(L + 2, 0);
  maxsize.assign(L + 2, 0);
  maxprefix.assign(L + 2, 0);
  midprefix.assign(L + 2, 0);
  for (uint32_t l = 0; l <= L + 1; ++l) {
    minsize[l] = bound_minsize(kind, mode, t, p, q, l);
    maxsize[l] = bound_maxsize(kind, mode, t, p, q, l);
  }
  if (kind == HS_COSINE) {
    mo.assign((size_t)(L + 1) * (L + 1), 0);
    for (uint32_t a = 0; a <= L; ++a)
      for (uint32_t b = 0; b <= L; ++b) {
        uint32_t raw = bound_minoverlap_raw(kind, mode, t, p, q, a, b);
        mo[(size_t)a * (L + 1) + b] = (uint32_t)imin((int)b, imin((int)a, (int)raw));
      }
  } else {
    mo.assign(2 * (size_t)L + 1, 0);
    // The raw value depends on l1 + l2 only; evaluate it as (k, 0) so the
    // double expression sees the same sum the reference computes.
    for (uint32_t k = 0; k <= 2 * L; ++k) mo[k] = bound_minoverlap_raw(kind, mode, t, p, q, k, 0);
  }
  for (uint32_t l = 0; l <= L + 1; ++l) {
    maxprefix[l] = (uint32_t)imin((int)l, (int)(l - minsize[l] + 1));
    uint32_t mol = l <= L ? minoverlap(l, l) : 0;
    midprefix[l] = (uint32_t)imin((int)l, (int)(l - mol + 1));
  }
}

uint32_t BoundTables::minoverlap(uint32_t l1, uint32_t l2) const {
  uint32_t m = l1 < l2 ? l1 : l2;
  uint32_t v = (kind == HS_COSINE) ? mo[(size_t)l1 * (L + 1) + l2] : mo[l1 + l2];
  return v < m ? v : m;
}

}  // namespace hs
