// gi_host_internal.h -- helpers shared by the host-side sources of
// libgiant_impact.so (not part of the C ABI).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

namespace gi_host {



// Physical constants, Code/CommonCode/commonSetup.h:29-35.
constexpr double kPi = 3.141592654;
constexpr double kUniversalGravity = 6.67408e-20;  // km^3 kg^-1 s^-2
constexpr double kMassOfEarth = 5.97219e24;        // kg

// Value printed with "%f" and read back with operator>> (the parameter file
// round trip of generateTargetImpactorBasicFunctions.h:693-709 ->
// impactBasicFunctions.h:39-75).
inline double round_trip_f(double v) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "%f", v);
    return std::strtod(buf, nullptr);
}

// Reads "the value after each '='" in order, as the reference's
// getline(data, name, '=') ; data >> value pairs do.
class PositionalReader {
   public:
    explicit PositionalReader(const char* path) : in_(path) {}
    bool ok() const { return (bool)in_; }
    template <typename T>
    bool next(T& v) {
        std::string skip;
        if (!std::getline(in_, skip, '=')) return false;
        return (bool)(in_ >> v);
    }

   private:
    std::ifstream in_;
};

// SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014).
struct SplitMix64 {
    unsigned long long s;
    unsigned long long next() {
        unsigned long long z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct LatticePoint {
    int x, y, z;
    long long d2;
};

// The n simple-cubic lattice sites (spacing 1) nearest the origin, ordered by
// distance, ties broken by (x, y, z).
inline std::vector<LatticePoint> nearest_sites(int n) {
    int h = (int)std::ceil(std::cbrt(3.0 * n / (4.0 * kPi))) + 2;
    std::vector<LatticePoint> pts;
    pts.reserve((size_t)(2 * h + 1) * (2 * h + 1) * (2 * h + 1));
    for (int x = -h; x <= h; ++x)
        for (int y = -h; y <= h; ++y)
            for (int z = -h; z <= h; ++z) pts.push_back({x, y, z, (long long)x * x + (long long)y * y + (long long)z * z});
    auto less = [](const LatticePoint& a, const LatticePoint& b) {
        if (a.d2 != b.d2) return a.d2 < b.d2;
        if (a.x != b.x) return a.x < b.x;
        if (a.y != b.y) return a.y < b.y;
        return a.z < b.z;
    };
    std::partial_sort(pts.begin(), pts.begin() + n, pts.end(), less);
    pts.resize(n);
    return pts;
}

struct Body {
    std::vector<double> x, y, z;  // jittered positions, radius order
    int nfe = 0;
};

inline Body make_body(int n, double fe_fraction, double jitter, SplitMix64& rng) {
    Body b;
    std::vector<LatticePoint> sites = nearest_sites(n);
    b.x.resize(n);
    b.y.resize(n);
    b.z.resize(n);
    for (int k = 0; k < n; ++k) {
        b.x[k] = sites[k].x + (2.0 * rng.uniform() - 1.0) * jitter;
        b.y[k] = sites[k].y + (2.0 * rng.uniform() - 1.0) * jitter;
        b.z[k] = sites[k].z + (2.0 * rng.uniform() - 1.0) * jitter;
    }
    b.nfe = (int)std::floor(fe_fraction * n + 0.5);
    if (b.nfe > n) b.nfe = n;
    return b;
}


}  // namespace gi_host

namespace gi_host {

// A body of n lattice sites whose innermost nfe sites are iron.
inline Body make_body_count(int n, int nfe, double jitter, SplitMix64& rng) {
    Body b = make_body(n, 0.0, jitter, rng);
    b.nfe = std::min(std::max(nfe, 0), n);
    return b;
}

}  // namespace gi_host
