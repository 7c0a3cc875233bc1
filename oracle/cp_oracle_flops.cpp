// Flop-counting build of the CPU oracle (test infrastructure only, like cp_oracle.cpp).
//
// cp_oracle.cpp is compiled a second time with every `double` replaced by Counted, a scalar that
// holds one double and counts each floating-point operation applied to it.  Nothing else changes:
// the same source, the same control flow, the same results (Counted does its arithmetic in
// double), so the counts are those of the oracle's algorithm on the given inputs.  The library
// exports the oracle's functions unchanged (Counted has the layout and the calling convention of a
// double: one double member, trivially copyable) plus the readout below.
//
// Counting convention (algorithmic flops): + - * / each 1, sqrt 1, exp/log/pow/tgamma/atan/cos 1 each
// (kept apart as "functions"); comparisons, fabs, negation, copies and conversions 0.  Written
// a*b + c counts 2.  Single-threaded (built without OpenMP).
//
// Readout: or_flop_counts(step, pass, out[8]) -> {add/sub, mul, div, sqrt, functions, qp calls,
// flops up to the first numerically singular local system, calls that met one} of the qp_update
// calls in Newton step `step` (increment * max_newton + iteration) of or_simulate; pass 1 = the
// assembly pass (flag 1: return map + consistent tangent), pass 0 = the post-increment pass
// (flag 0: stress only; attributed to the increment's last step).  or_flop_total(out[5]) reads
// the running totals of every counted operation since or_flop_reset.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <type_traits>
#include <vector>

namespace flops {
long long g_count[5];  // add/sub, mul, div, sqrt, functions
enum { ADD, MUL, DIV, SQRT, FUN };
}  // namespace flops

struct Counted {
  double v;
  Counted() = default;
  constexpr Counted(double x) : v(x) {}
  template <class T, class = typename std::enable_if<std::is_integral<T>::value>::type>
  constexpr Counted(T x) : v((double)x) {}
  explicit operator double() const { return v; }
  explicit operator int() const { return (int)v; }
  explicit operator unsigned() const { return (unsigned)v; }
  explicit operator long long() const { return (long long)v; }
  explicit operator bool() const { return v != 0.0; }
  Counted &operator+=(Counted o) { ++flops::g_count[flops::ADD]; v += o.v; return *this; }
  Counted &operator-=(Counted o) { ++flops::g_count[flops::ADD]; v -= o.v; return *this; }
  Counted &operator*=(Counted o) { ++flops::g_count[flops::MUL]; v *= o.v; return *this; }
  Counted &operator/=(Counted o) { ++flops::g_count[flops::DIV]; v /= o.v; return *this; }
};
static_assert(sizeof(Counted) == sizeof(double) && std::is_trivially_copyable<Counted>::value, "layout");

inline Counted operator+(Counted a, Counted b) { return a += b; }
inline Counted operator-(Counted a, Counted b) { return a -= b; }
inline Counted operator*(Counted a, Counted b) { return a *= b; }
inline Counted operator/(Counted a, Counted b) { return a /= b; }
inline Counted operator-(Counted a) { return Counted(-a.v); }
inline Counted operator+(Counted a) { return a; }
inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
inline bool operator!=(Counted a, Counted b) { return a.v != b.v; }

namespace std {
inline Counted fabs(Counted a) { return Counted(std::fabs(a.v)); }
inline Counted abs(Counted a) { return Counted(std::fabs(a.v)); }
inline Counted sqrt(Counted a) { ++flops::g_count[flops::SQRT]; return Counted(std::sqrt(a.v)); }
inline Counted exp(Counted a) { ++flops::g_count[flops::FUN]; return Counted(std::exp(a.v)); }
inline Counted log(Counted a) { ++flops::g_count[flops::FUN]; return Counted(std::log(a.v)); }
inline Counted atan(Counted a) { ++flops::g_count[flops::FUN]; return Counted(std::atan(a.v)); }
inline Counted cos(Counted a) { ++flops::g_count[flops::FUN]; return Counted(std::cos(a.v)); }
inline Counted log10(Counted a) { ++flops::g_count[flops::FUN]; return Counted(std::log10(a.v)); }
inline Counted tgamma(Counted a) { ++flops::g_count[flops::FUN]; return Counted(std::tgamma(a.v)); }
inline Counted pow(Counted a, Counted b) { ++flops::g_count[flops::FUN]; return Counted(std::pow(a.v, b.v)); }
inline Counted pow(Counted a, int b) { ++flops::g_count[flops::FUN]; return Counted(std::pow(a.v, b)); }
inline Counted floor(Counted a) { return Counted(std::floor(a.v)); }
inline Counted fmod(Counted a, Counted b) { ++flops::g_count[flops::DIV]; return Counted(std::fmod(a.v, b.v)); }
inline Counted copysign(Counted a, Counted b) { return Counted(std::copysign(a.v, b.v)); }
inline bool isfinite(Counted a) { return std::isfinite(a.v); }
inline bool isnan(Counted a) { return std::isnan(a.v); }
}  // namespace std

namespace flops {
constexpr int kMaxSteps = 4096;
// per step and pass: add/sub, mul, div, sqrt, functions, qp calls, flops of the calls up to their
// first numerically singular local system (all of a call's flops when it meets none), calls that
// met one
long long g_step[kMaxSteps][2][8];
int g_cur_step = 0;
long long g_enter[5], g_sing[5];
bool g_sing_seen = false;
inline long long sum5(const long long *c) { return c[0] + c[1] + c[2] + c[3] + c[4]; }
}  // namespace flops

#define OR_FLOP_STEP(inc, it, max_newton) (flops::g_cur_step = (inc) * (max_newton) + (it))
#define OR_FLOP_QP_BEGIN(flag) \
  (std::memcpy(flops::g_enter, flops::g_count, sizeof(flops::g_enter)), flops::g_sing_seen = false)
#define OR_FLOP_SINGULAR()                                                                        \
  do {                                                                                            \
    if (!flops::g_sing_seen) {                                                                    \
      std::memcpy(flops::g_sing, flops::g_count, sizeof(flops::g_sing));                          \
      flops::g_sing_seen = true;                                                                  \
    }                                                                                             \
  } while (0)
#define OR_FLOP_QP_END(flag)                                                                        \
  do {                                                                                            \
    if (flops::g_cur_step >= 0 && flops::g_cur_step < flops::kMaxSteps) {                         \
      long long *o = flops::g_step[flops::g_cur_step][(flag) ? 1 : 0];                            \
      for (int k_ = 0; k_ < 5; ++k_) o[k_] += flops::g_count[k_] - flops::g_enter[k_];            \
      o[5] += 1;                                                                                  \
      o[6] += flops::sum5(flops::g_sing_seen ? flops::g_sing : flops::g_count) - flops::sum5(flops::g_enter); \
      o[7] += flops::g_sing_seen ? 1 : 0;                                                         \
    }                                                                                             \
  } while (0)

// cp_oracle.h is first seen here, so its structs and declarations take Counted too (same layout)
#define double Counted
#include "cp_oracle.cpp"
#undef double

extern "C" {
void or_flop_reset(void) {
  std::memset(flops::g_count, 0, sizeof(flops::g_count));
  std::memset(flops::g_step, 0, sizeof(flops::g_step));
  flops::g_cur_step = 0;
}
int or_flop_counts(int step, int pass, long long out[8]) {
  if (step < 0 || step >= flops::kMaxSteps || pass < 0 || pass > 1) return -1;
  for (int k = 0; k < 8; ++k) out[k] = flops::g_step[step][pass][k];
  return 0;
}
void or_flop_total(long long out[5]) {
  for (int k = 0; k < 5; ++k) out[k] = flops::g_count[k];
}
}
