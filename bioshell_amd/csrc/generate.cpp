// generate.cpp -- deterministic synthetic ensembles.
//
// Chain: a random-walk CA trace placed atom by atom with the NeRF construction
// of bioshell-builder/src/nerf.rs: the first three atoms form a stub
// (create_stub, nerf.rs:6-11: b = a + (r,0,0), c = b + r(cos(PI-theta),
// sin(PI-theta), 0)), every later atom is restored from its three predecessors
// (restore_atom, nerf.rs:13-30: local frame [bc, n x bc, n] with
// n = (b-a) x bc, local offset r(cos(PI-theta), sin(PI-theta)cos(phi),
// sin(PI-theta)sin(phi))).  Bond 3.8 A, theta ~ U(85, 150) deg,
// phi ~ U(-180, 180] deg.
//
// Conformation i: base + N(0, s_i^2) per coordinate, s_i ~ U(sigma_lo, sigma_hi)
// (or s_i = near_sigma with probability near_fraction), then a uniform random
// rotation (Shoemake's unit quaternion from three uniforms) and a translation
// U(-50, 50)^3.  Conformation i draws only from its own stream, seeded from
// (seed, i), so generation is parallel and any subset can be regenerated.
//
// Random numbers: SplitMix64 seeding xoshiro256**; uniforms from the top 53
// bits; normals by Box-Muller.
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "common.hpp"

namespace bs {
namespace {

inline uint64_t splitmix64(uint64_t& x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Xoshiro256ss {
    uint64_t s[4];
    explicit Xoshiro256ss(uint64_t seed)
    {
        uint64_t x = seed;
        for (auto& v : s) v = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
    double normal()
    {
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

struct V3 {
    double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 unit(V3 a)
{
    const double n = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return {a.x / n, a.y / n, a.z / n};
}

void base_chain(uint64_t seed, size_t L, std::vector<V3>& out)
{
    Xoshiro256ss rng(seed ^ 0x5DEECE66DB1A5E5Dull);
    const double r = 3.8;
    const double deg = M_PI / 180.0;
    out.assign(L, V3{0, 0, 0});
    if (L >= 2) out[1] = {r, 0.0, 0.0};
    if (L >= 3) {
        const double th = rng.uniform(85.0, 150.0) * deg;
        const double ang = M_PI - th;
        out[2] = {out[1].x + r * std::cos(ang), out[1].y + r * std::sin(ang), out[1].z};
    }
    for (size_t i = 3; i < L; ++i) {
        const double th = rng.uniform(85.0, 150.0) * deg;
        // U(-180, 180]: 180 - U[0, 360)
        const double ph = (180.0 - rng.uniform(0.0, 360.0)) * deg;
        const V3 a = out[i - 3], b = out[i - 2], c = out[i - 1];
        const V3 bc = unit(sub(c, b));
        const V3 n = unit(cross(unit(sub(b, a)), bc));
        const V3 m = cross(n, bc);
        const double ang = M_PI - th;
        const double lx = r * std::cos(ang), ly = r * std::sin(ang) * std::cos(ph),
                     lz = r * std::sin(ang) * std::sin(ph);
        out[i] = {c.x + bc.x * lx + m.x * ly + n.x * lz, c.y + bc.y * lx + m.y * ly + n.y * lz,
                  c.z + bc.z * lx + m.z * ly + n.z * lz};
    }
}

void one_conformation(uint64_t seed, size_t i, const std::vector<V3>& base, double sigma_lo,
                      double sigma_hi, double near_fraction, double near_sigma, double* out)
{
    uint64_t sx = seed * 0x9E3779B97F4A7C15ull + (uint64_t)i * 0xD1B54A32D192ED03ull + 1;
    Xoshiro256ss rng(splitmix64(sx));
    const double u = rng.uniform();
    const double s = u < near_fraction ? near_sigma : rng.uniform(sigma_lo, sigma_hi);
    // uniform rotation (Shoemake)
    const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
    const double q1 = std::sqrt(1.0 - u1) * std::sin(2.0 * M_PI * u2);
    const double q2 = std::sqrt(1.0 - u1) * std::cos(2.0 * M_PI * u2);
    const double q3 = std::sqrt(u1) * std::sin(2.0 * M_PI * u3);
    const double q0 = std::sqrt(u1) * std::cos(2.0 * M_PI * u3);
    const double R[9] = {q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3, 2 * (q1 * q2 - q0 * q3),
                         2 * (q1 * q3 + q0 * q2),               2 * (q1 * q2 + q0 * q3),
                         q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3, 2 * (q2 * q3 - q0 * q1),
                         2 * (q1 * q3 - q0 * q2),               2 * (q2 * q3 + q0 * q1),
                         q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3};
    const double t[3] = {rng.uniform(-50.0, 50.0), rng.uniform(-50.0, 50.0),
                         rng.uniform(-50.0, 50.0)};
    for (size_t k = 0; k < base.size(); ++k) {
        const double p[3] = {base[k].x + s * rng.normal(), base[k].y + s * rng.normal(),
                             base[k].z + s * rng.normal()};
        for (int c = 0; c < 3; ++c)
            out[3 * k + c] = R[3 * c] * p[0] + R[3 * c + 1] * p[1] + R[3 * c + 2] * p[2] + t[c];
    }
}

}  // namespace
}  // namespace bs

extern "C" bs_status bs_generate_ensemble(uint64_t seed, size_t n_conf, size_t n_atoms,
                                          double sigma_lo, double sigma_hi, double near_fraction,
                                          double near_sigma, double* out_xyz)
{
    return bs::guard([&] {
        BS_REQUIRE(out_xyz != nullptr || n_conf * n_atoms == 0, "out_xyz is null");
        BS_REQUIRE(sigma_lo >= 0.0 && sigma_hi >= sigma_lo, "need 0 <= sigma_lo <= sigma_hi");
        BS_REQUIRE(near_fraction >= 0.0 && near_fraction <= 1.0, "near_fraction must be in [0,1]");
        if (n_conf == 0 || n_atoms == 0) return;
        std::vector<bs::V3> base;
        bs::base_chain(seed, n_atoms, base);
        unsigned hw = std::thread::hardware_concurrency();
        const size_t nth = std::max<size_t>(1, std::min<size_t>(hw ? hw : 1, 16));
        std::vector<std::thread> th;
        for (size_t w = 0; w < nth; ++w)
            th.emplace_back([&, w] {
                for (size_t i = w; i < n_conf; i += nth)
                    bs::one_conformation(seed, i, base, sigma_lo, sigma_hi, near_fraction, near_sigma,
                                         out_xyz + 3 * n_atoms * i);
            });
        for (auto& t : th) t.join();
    });
}
