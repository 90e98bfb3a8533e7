// Operation-counting scalar for the oracle -- TEST INFRASTRUCTURE.
//
// Instantiating the oracle templates with orc::Cnt runs the same algorithm in
// double precision while tallying every floating-point add/subtract,
// multiply, divide and square root. The tallies give the algorithmic FP32
// operation count per segment-substep that the roofline figures use, derived
// from the algorithm itself rather than estimated by hand.
#pragma once

#include <cmath>

namespace orc {

struct OpTally {
    long long add = 0, mul = 0, div = 0, sqrt = 0, cmp = 0;
};
inline OpTally& tally() {
    static thread_local OpTally t;
    return t;
}

struct Cnt {
    double v = 0;
    constexpr Cnt() = default;
    constexpr Cnt(double x) : v(x) {}
    explicit operator int() const { return int(v); }
    explicit operator double() const { return v; }
};

inline Cnt operator+(Cnt a, Cnt b) { ++tally().add; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { ++tally().add; return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { ++tally().mul; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { ++tally().div; return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline Cnt& operator+=(Cnt& a, Cnt b) { a = a + b; return a; }
inline Cnt& operator-=(Cnt& a, Cnt b) { a = a - b; return a; }
inline bool operator<(Cnt a, Cnt b) { ++tally().cmp; return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { ++tally().cmp; return a.v > b.v; }
inline bool operator<=(Cnt a, Cnt b) { ++tally().cmp; return a.v <= b.v; }
inline bool operator>=(Cnt a, Cnt b) { ++tally().cmp; return a.v >= b.v; }
inline bool operator==(Cnt a, Cnt b) { ++tally().cmp; return a.v == b.v; }
inline bool operator!=(Cnt a, Cnt b) { ++tally().cmp; return a.v != b.v; }
inline Cnt oSqrt(Cnt a) { ++tally().sqrt; return Cnt(std::sqrt(a.v)); }
inline Cnt oFloor(Cnt a) { return Cnt(std::floor(a.v)); }
inline Cnt oExp(Cnt a) { return Cnt(std::exp(a.v)); }
inline Cnt oCos(Cnt a) { return Cnt(std::cos(a.v)); }
inline Cnt oSin(Cnt a) { return Cnt(std::sin(a.v)); }

}  // namespace orc
