// CPU restatement of the reference hot path -- TEST INFRASTRUCTURE ONLY.
//
// This directory is the oracle: a plain C++ restatement of the reference's
// fEuler/face_flux/stability (src/utilities.cpp), ghost fill and stencil packing
// (src/euler3D.hpp), the primordial chemistry network (src/raja_primordial_network.cpp),
// the Dengo scaling helpers (src/primordial_blast.cpp) and the slow/fast RHS
// wrappers (src/multirate_chem_hydro_main.cpp). It is compiled with
// -O2 -ffp-contract=off (the reference targets baseline x86-64, where no FMA is
// formed) and loaded only by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py, always as the checker. The shipped library
// never links or calls anything here.
//
// Every algorithmic step keeps the reference's operation order so that the
// oracle reproduces the reference arithmetic; the comments cite the lines.
#pragma once
#include <cmath>
#include <cstdint>
#include <algorithm>

namespace orc {

constexpr int NVMAX = 5 + 16;   // fluid fields + up to 16 advected tracers

// Boundary condition codes, euler3D.hpp:90-93.
enum { BC_PERIODIC = 0, BC_NEUMANN = 1, BC_DIRICHLET = 2, BC_REFLECTING = 3 };

// Op-counting scalar: instantiating the templated kernels with this type gives
// the exact per-face / per-cell operation counts used for the roofline.
struct OpCounter {
  static uint64_t add, mul, div, sqrt_, cmp, trans;
  static void reset() { add = mul = div = sqrt_ = cmp = trans = 0; }
};

struct Cnt {
  double v;
  Cnt() : v(0) {}
  Cnt(double x) : v(x) {}
};
inline Cnt operator+(Cnt a, Cnt b) { OpCounter::add++; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { OpCounter::add++; return Cnt(a.v - b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline Cnt operator*(Cnt a, Cnt b) { OpCounter::mul++; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { OpCounter::div++; return Cnt(a.v / b.v); }
inline Cnt& operator+=(Cnt& a, Cnt b) { a = a + b; return a; }
inline Cnt& operator-=(Cnt& a, Cnt b) { a = a - b; return a; }
inline Cnt& operator*=(Cnt& a, Cnt b) { a = a * b; return a; }
inline bool operator<(Cnt a, Cnt b) { OpCounter::cmp++; return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { OpCounter::cmp++; return a.v > b.v; }
inline bool operator<=(Cnt a, Cnt b) { OpCounter::cmp++; return a.v <= b.v; }
inline Cnt sqrt(Cnt a) { OpCounter::sqrt_++; return Cnt(std::sqrt(a.v)); }
inline Cnt fabs(Cnt a) { OpCounter::cmp++; return Cnt(std::fabs(a.v)); }
inline Cnt log(Cnt a) { OpCounter::trans++; return Cnt(std::log(a.v)); }
inline Cnt exp(Cnt a) { OpCounter::trans++; return Cnt(std::exp(a.v)); }
inline Cnt pow(Cnt a, double e) { OpCounter::trans++; return Cnt(std::pow(a.v, e)); }
inline Cnt fmax(Cnt a, Cnt b) { OpCounter::cmp++; return Cnt(std::fmax(a.v, b.v)); }
inline Cnt fmin(Cnt a, Cnt b) { OpCounter::cmp++; return Cnt(std::fmin(a.v, b.v)); }
inline double value(Cnt a) { return a.v; }
inline double value(double a) { return a; }

// double overloads so that unqualified calls inside namespace orc resolve to libm
inline double sqrt(double a) { return std::sqrt(a); }
inline double fabs(double a) { return std::fabs(a); }
inline double log(double a) { return std::log(a); }
inline double exp(double a) { return std::exp(a); }
inline double pow(double a, double e) { return std::pow(a, e); }
inline double fmax(double a, double b) { return std::fmax(a, b); }
inline double fmin(double a, double b) { return std::fmin(a, b); }

// SUNRsqrt semantics (sundials_math.h): non-positive arguments give 0.
template <class R> inline R sun_sqrt(R x) { return (x <= R(0.0)) ? R(0.0) : sqrt(x); }

}  // namespace orc
