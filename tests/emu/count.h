// count.h -- TEST INFRASTRUCTURE ONLY: a double that counts the FP64
// operations made on it, by class, for the algorithmic operation count of
// the column step (the weights of the measurement: +, -, *, min, max and
// compare 1; divide and sqrt 10; exp, log, pow and tanh 20).  Each
// workgroup lane of the host build runs in its own thread, so the counters
// are thread-local and summed by the caller.
#pragma once

#include <cmath>

namespace cnt {

struct Counts {
  long long add = 0, mul = 0, div = 0, sqrt = 0, trans = 0, cmp = 0;
};
inline thread_local Counts tally;

struct fd {
  double v;
  fd() : v(0.0) {}
  fd(double x) : v(x) {}  // constants are free
};

inline fd operator+(fd a, fd b) { tally.add++; return fd(a.v + b.v); }
inline fd operator+(fd a, double b) { tally.add++; return fd(a.v + b); }
inline fd operator+(double a, fd b) { tally.add++; return fd(a + b.v); }
inline fd operator-(fd a, fd b) { tally.add++; return fd(a.v - b.v); }
inline fd operator-(fd a, double b) { tally.add++; return fd(a.v - b); }
inline fd operator-(double a, fd b) { tally.add++; return fd(a - b.v); }
inline fd operator-(fd a) { return fd(-a.v); }  // a sign modifier, free
inline fd operator*(fd a, fd b) { tally.mul++; return fd(a.v * b.v); }
inline fd operator*(fd a, double b) { tally.mul++; return fd(a.v * b); }
inline fd operator*(double a, fd b) { tally.mul++; return fd(a * b.v); }
inline fd operator/(fd a, fd b) { tally.div++; return fd(a.v / b.v); }
inline fd operator/(fd a, double b) { tally.div++; return fd(a.v / b); }
inline fd operator/(double a, fd b) { tally.div++; return fd(a / b.v); }
inline fd &operator+=(fd &a, fd b) { a = a + b; return a; }
inline fd &operator*=(fd &a, fd b) { a = a * b; return a; }

#define CNT_CMP(OP)                                                             \
  inline bool operator OP(fd a, fd b) { tally.cmp++; return a.v OP b.v; }       \
  inline bool operator OP(fd a, double b) { tally.cmp++; return a.v OP b; }     \
  inline bool operator OP(double a, fd b) { tally.cmp++; return a OP b.v; }
CNT_CMP(<)
CNT_CMP(>)
CNT_CMP(<=)
CNT_CMP(>=)
CNT_CMP(==)
CNT_CMP(!=)
#undef CNT_CMP

inline fd exp(fd a) { tally.trans++; return fd(::exp(a.v)); }
inline fd log(fd a) { tally.trans++; return fd(::log(a.v)); }
inline fd tanh(fd a) { tally.trans++; return fd(::tanh(a.v)); }
inline fd pow(fd a, double p) { tally.trans++; return fd(::pow(a.v, p)); }
inline fd sqrt(fd a) { tally.sqrt++; return fd(::sqrt(a.v)); }
inline fd fabs(fd a) { return fd(::fabs(a.v)); }  // a sign modifier, free

inline double weighted(const Counts &c) {
  return (double)(c.add + c.mul + c.cmp) + 10.0 * (double)(c.div + c.sqrt) +
         20.0 * (double)c.trans;
}

}  // namespace cnt
