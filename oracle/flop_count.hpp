// flop_count.hpp — a counting scalar for the oracle.
//
// TEST INFRASTRUCTURE ONLY.  `Flop` wraps a double and counts every
// arithmetic operation applied to it (add/sub, mul, div, sqrt; comparisons,
// negation, abs and min/max are not counted), so instantiating the oracle's
// World<Flop> yields the algorithmic floating-point operation count of
// newt's step on a given state: the FLOP figure the step's compute roofline
// is priced against.
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>

namespace oracle {

struct FlopCounts {
    uint64_t add = 0, mul = 0, div = 0, sqrt = 0;
};
inline FlopCounts& flop_counts() {
    static thread_local FlopCounts c;
    return c;
}

struct Flop {
    double v = 0.0;
    Flop() = default;
    constexpr Flop(double x) : v(x) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    explicit operator int() const { return (int)v; }
    Flop& operator+=(Flop o) { ++flop_counts().add; v += o.v; return *this; }
    Flop& operator-=(Flop o) { ++flop_counts().add; v -= o.v; return *this; }
    Flop& operator*=(Flop o) { ++flop_counts().mul; v *= o.v; return *this; }
    Flop& operator/=(Flop o) { ++flop_counts().div; v /= o.v; return *this; }
};
inline Flop operator+(Flop a, Flop b) { ++flop_counts().add; return Flop(a.v + b.v); }
inline Flop operator-(Flop a, Flop b) { ++flop_counts().add; return Flop(a.v - b.v); }
inline Flop operator*(Flop a, Flop b) { ++flop_counts().mul; return Flop(a.v * b.v); }
inline Flop operator/(Flop a, Flop b) { ++flop_counts().div; return Flop(a.v / b.v); }
inline Flop operator-(Flop a) { return Flop(-a.v); }
inline bool operator<(Flop a, Flop b) { return a.v < b.v; }
inline bool operator>(Flop a, Flop b) { return a.v > b.v; }
inline bool operator<=(Flop a, Flop b) { return a.v <= b.v; }
inline bool operator>=(Flop a, Flop b) { return a.v >= b.v; }
inline bool operator==(Flop a, Flop b) { return a.v == b.v; }
inline bool operator!=(Flop a, Flop b) { return a.v != b.v; }

}  // namespace oracle

namespace std {
template <> class numeric_limits<oracle::Flop> : public numeric_limits<double> {
  public:
    static constexpr oracle::Flop infinity() noexcept { return oracle::Flop(numeric_limits<double>::infinity()); }
    static constexpr oracle::Flop max() noexcept { return oracle::Flop(numeric_limits<double>::max()); }
    static constexpr oracle::Flop lowest() noexcept { return oracle::Flop(numeric_limits<double>::lowest()); }
    static constexpr oracle::Flop epsilon() noexcept { return oracle::Flop(numeric_limits<double>::epsilon()); }
};
inline oracle::Flop sqrt(oracle::Flop a) { ++oracle::flop_counts().sqrt; return oracle::Flop(std::sqrt(a.v)); }
inline oracle::Flop fabs(oracle::Flop a) { return oracle::Flop(std::fabs(a.v)); }
inline oracle::Flop abs(oracle::Flop a) { return oracle::Flop(std::fabs(a.v)); }
inline bool isfinite(oracle::Flop a) { return std::isfinite(a.v); }
inline bool isnan(oracle::Flop a) { return std::isnan(a.v); }
inline oracle::Flop floor(oracle::Flop a) { return oracle::Flop(std::floor(a.v)); }
inline oracle::Flop fmax(oracle::Flop a, oracle::Flop b) { return oracle::Flop(std::fmax(a.v, b.v)); }
inline oracle::Flop fmin(oracle::Flop a, oracle::Flop b) { return oracle::Flop(std::fmin(a.v, b.v)); }
inline oracle::Flop copysign(oracle::Flop a, oracle::Flop b) { return oracle::Flop(std::copysign(a.v, b.v)); }
inline bool signbit(oracle::Flop a) { return std::signbit(a.v); }
}  // namespace std
