// ORACLE (test infrastructure only). A double that counts the floating-point operations done on it, so that the
// restatement instantiated on it reports the FP work of each stage (the FLOP figure of the step's VALU roofline).
//
// Counting rules: each binary +, -, * and / is one operation (the device library is built without FMA contraction,
// so every product and every sum is its own instruction there too); sqrt is one operation; negation, fabs,
// comparisons, fmin/fmax, floor/ceil and conversions are free. Library calls (sin, cos, atan2, acos, pow, fmod) are
// one operation each and are also counted apart in flops_transcendental. The counters are plain globals: count on
// one thread.
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

struct FlopCounter {
    static inline uint64_t ops = 0;
    static inline uint64_t transcendental = 0;
};

struct CF {
    double v;
    CF() = default;
    constexpr CF(double x) : v(x) {}
    explicit constexpr operator double() const { return v; }
    explicit constexpr operator float() const { return (float)v; }
    explicit constexpr operator int() const { return (int)v; }
    explicit constexpr operator long() const { return (long)v; }
    explicit constexpr operator unsigned long() const { return (unsigned long)v; }
    CF& operator+=(CF b);
    CF& operator-=(CF b);
    CF& operator*=(CF b);
    CF& operator/=(CF b);
};

inline CF operator+(CF a, CF b) { FlopCounter::ops++; return CF(a.v + b.v); }
inline CF operator-(CF a, CF b) { FlopCounter::ops++; return CF(a.v - b.v); }
inline CF operator*(CF a, CF b) { FlopCounter::ops++; return CF(a.v * b.v); }
inline CF operator/(CF a, CF b) { FlopCounter::ops++; return CF(a.v / b.v); }
inline CF operator-(CF a) { return CF(-a.v); }
inline CF operator+(CF a) { return a; }
inline CF& CF::operator+=(CF b) { return *this = *this + b; }
inline CF& CF::operator-=(CF b) { return *this = *this - b; }
inline CF& CF::operator*=(CF b) { return *this = *this * b; }
inline CF& CF::operator/=(CF b) { return *this = *this / b; }
inline bool operator<(CF a, CF b) { return a.v < b.v; }
inline bool operator<=(CF a, CF b) { return a.v <= b.v; }
inline bool operator>(CF a, CF b) { return a.v > b.v; }
inline bool operator>=(CF a, CF b) { return a.v >= b.v; }
inline bool operator==(CF a, CF b) { return a.v == b.v; }
inline bool operator!=(CF a, CF b) { return a.v != b.v; }

inline CF cf_call(double r) {
    FlopCounter::ops++;
    FlopCounter::transcendental++;
    return CF(r);
}

}  // namespace orc

// the restatement calls the std:: math functions by their qualified names
namespace std {
inline orc::CF fabs(orc::CF a) { return orc::CF(std::fabs(a.v)); }
inline orc::CF abs(orc::CF a) { return orc::CF(std::fabs(a.v)); }
inline orc::CF fmin(orc::CF a, orc::CF b) { return orc::CF(std::fmin(a.v, b.v)); }
inline orc::CF fmax(orc::CF a, orc::CF b) { return orc::CF(std::fmax(a.v, b.v)); }
inline orc::CF floor(orc::CF a) { return orc::CF(std::floor(a.v)); }
inline orc::CF ceil(orc::CF a) { return orc::CF(std::ceil(a.v)); }
inline orc::CF sqrt(orc::CF a) { orc::FlopCounter::ops++; return orc::CF(std::sqrt(a.v)); }
inline orc::CF sin(orc::CF a) { return orc::cf_call(std::sin(a.v)); }
inline orc::CF cos(orc::CF a) { return orc::cf_call(std::cos(a.v)); }
inline orc::CF acos(orc::CF a) { return orc::cf_call(std::acos(a.v)); }
inline orc::CF atan2(orc::CF a, orc::CF b) { return orc::cf_call(std::atan2(a.v, b.v)); }
inline orc::CF pow(orc::CF a, orc::CF b) { return orc::cf_call(std::pow(a.v, b.v)); }
inline orc::CF fmod(orc::CF a, orc::CF b) { return orc::cf_call(std::fmod(a.v, b.v)); }
inline bool isnan(orc::CF a) { return std::isnan(a.v); }
inline bool isinf(orc::CF a) { return std::isinf(a.v); }
inline bool isfinite(orc::CF a) { return std::isfinite(a.v); }
}  // namespace std
