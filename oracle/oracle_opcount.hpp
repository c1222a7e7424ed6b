// Counting double for the op-counting build of the oracle (TEST
// INFRASTRUCTURE ONLY): oracle_engine.c and oracle_tables.c are compiled as
// C++ with this header force-included, so every `double` becomes `ocd`,
// which performs the same IEEE operation and adds one to orc_ops[phase] per
// add / sub / mul / div and per libm call (an FMA-free build, so a*b+c
// counts 2).  fabs and floor count 0; comparisons and conversions count 0.
// The values are those of the plain oracle bit for bit.
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <limits.h>
#include <float.h>
#include <type_traits>
extern "C" long long orc_ops[16];
extern "C" int orc_phase;
struct ocd {
  double v;
  ocd() = default;
  constexpr ocd(double x) : v(x) {}
  operator double() const { return v; }
  ocd &operator+=(ocd o) { orc_ops[orc_phase]++; v += o.v; return *this; }
  ocd &operator-=(ocd o) { orc_ops[orc_phase]++; v -= o.v; return *this; }
  ocd &operator*=(ocd o) { orc_ops[orc_phase]++; v *= o.v; return *this; }
  ocd &operator/=(ocd o) { orc_ops[orc_phase]++; v /= o.v; return *this; }
  ocd operator-() const { return ocd(-v); }
  ocd operator+() const { return *this; }
};
template <class T> using arith = std::enable_if_t<std::is_arithmetic<T>::value, int>;
#define OC_BIN(op) \
  inline ocd operator op(ocd a, ocd b) { orc_ops[orc_phase]++; return ocd(a.v op b.v); } \
  template <class T, arith<T> = 0> inline ocd operator op(ocd a, T b) { orc_ops[orc_phase]++; return ocd(a.v op (double)b); } \
  template <class T, arith<T> = 0> inline ocd operator op(T a, ocd b) { orc_ops[orc_phase]++; return ocd((double)a op b.v); }
OC_BIN(+) OC_BIN(-) OC_BIN(*) OC_BIN(/)
#define OC_CMP(op) \
  inline bool operator op(ocd a, ocd b) { return a.v op b.v; } \
  template <class T, arith<T> = 0> inline bool operator op(ocd a, T b) { return a.v op (double)b; } \
  template <class T, arith<T> = 0> inline bool operator op(T a, ocd b) { return (double)a op b.v; }
OC_CMP(<) OC_CMP(>) OC_CMP(<=) OC_CMP(>=) OC_CMP(==) OC_CMP(!=)
#define OC_F1(f, c) inline ocd f(ocd a) { orc_ops[orc_phase] += c; return ocd(::f(a.v)); }
OC_F1(exp, 1) OC_F1(sqrt, 1) OC_F1(log, 1) OC_F1(acos, 1) OC_F1(atan, 1) OC_F1(sin, 1) OC_F1(cos, 1)
OC_F1(fabs, 0) OC_F1(floor, 0)
inline ocd pow(ocd a, ocd b) { orc_ops[orc_phase]++; return ocd(::pow(a.v, b.v)); }
template <class T, arith<T> = 0> inline ocd pow(ocd a, T b) { orc_ops[orc_phase]++; return ocd(::pow(a.v, (double)b)); }
template <class T, arith<T> = 0> inline ocd pow(T a, ocd b) { orc_ops[orc_phase]++; return ocd(::pow((double)a, b.v)); }
inline ocd ldexp(ocd a, int e) { return ocd(::ldexp(a.v, e)); }
inline int isinf(ocd a) { return std::isinf(a.v); }
inline int isnan(ocd a) { return std::isnan(a.v); }
#define double ocd
