// Deterministic synthetic RGBA8 volumes for tests and benchmarks.
//
// Every random number comes from an in-repo SplitMix64 stream (no
// std::*_distribution, whose output is implementation-defined), and the
// arithmetic is plain IEEE double, so a seed gives the same bytes on every
// x86-64 host with this toolchain.  xds_hash (FNV-1a 64) fingerprints a
// buffer so two machines can check they hold the same volume.
//
// Conventions of the reference's data-prep scripts: RGBA8, x fastest, alpha
// 255 on every non-empty voxel (tools/make-bunny-volume.py:48).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double range(double a, double b) { return a + (b - a) * uniform(); }
};

template <typename F>
void parallel_z(uint64_t Z, F f) {
    unsigned n = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < n; ++t)
        th.emplace_back([=, &f] {
            for (uint64_t z = t; z < Z; z += n) f(z);
        });
    for (auto& t : th) t.join();
}

inline uint32_t pack(uint32_t r, uint32_t g, uint32_t b, uint32_t a) { return r | (g << 8) | (b << 16) | (a << 24); }

}  // namespace

extern "C" {

__attribute__((visibility("default"))) uint64_t xds_hash(const void* data, uint64_t n) {
    const uint8_t* p = static_cast<const uint8_t*>(data);
    uint64_t h = 1469598103934665603ull;
    for (uint64_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

// Ellipsoidal emission shells plus sparse uniform noise: n_shells shells with
// semi-axes U[rmin, rmax] voxels and wall thickness U[tmin, tmax], each with a
// uniform random RGB; then noise_frac of all voxels get a random colour.
__attribute__((visibility("default"))) void xds_shells(uint8_t* out, uint64_t X, uint64_t Y, uint64_t Z,
                                                       uint64_t seed, uint32_t n_shells, double rmin, double rmax,
                                                       double tmin, double tmax, double noise_frac) {
    struct Shell {
        double c[3], r[3], t;
        uint32_t col;
    };
    Rng rng(seed);
    std::vector<Shell> shells(n_shells);
    for (auto& s : shells) {
        s.c[0] = rng.range(0, (double)X);
        s.c[1] = rng.range(0, (double)Y);
        s.c[2] = rng.range(0, (double)Z);
        for (int i = 0; i < 3; ++i) s.r[i] = rng.range(rmin, rmax);
        s.t = rng.range(tmin, tmax);
        uint32_t r = (uint32_t)(rng.next() & 255), g = (uint32_t)(rng.next() & 255), b = (uint32_t)(rng.next() & 255);
        if ((r | g | b) == 0) r = 1;
        s.col = pack(r, g, b, 255);
    }
    uint32_t* o = reinterpret_cast<uint32_t*>(out);
    parallel_z(Z, [&](uint64_t z) {
        for (uint64_t y = 0; y < Y; ++y)
            for (uint64_t x = 0; x < X; ++x) {
                uint32_t v = 0;
                const double p[3] = {(double)x + 0.5, (double)y + 0.5, (double)z + 0.5};
                for (const auto& s : shells) {
                    // distance to the ellipsoid surface ~ |rn - 1| / |grad rn|
                    double q = 0, g2 = 0;
                    for (int i = 0; i < 3; ++i) {
                        double d = (p[i] - s.c[i]) / s.r[i];
                        q += d * d;
                        g2 += (d / s.r[i]) * (d / s.r[i]);
                    }
                    double rn = std::sqrt(q);
                    if (rn <= 0.0) continue;
                    double grad = std::sqrt(g2) / rn;
                    if (std::fabs(rn - 1.0) <= 0.5 * s.t * grad) v = s.col;
                }
                o[(z * Y + y) * X + x] = v;
            }
    });
    // sparse noise: a separate stream so the shells do not depend on it
    Rng nr(seed ^ 0xA5A5A5A55A5A5A5Aull);
    const uint64_t total = X * Y * Z;
    const uint64_t count = (uint64_t)(noise_frac * (double)total);
    for (uint64_t k = 0; k < count; ++k) {
        uint64_t i = nr.next() % total;
        uint32_t r = (uint32_t)(nr.next() & 255), g = (uint32_t)(nr.next() & 255), b = (uint32_t)(nr.next() & 255);
        o[i] = pack(r | 1, g, b, 255);
    }
}

// CT-scan-like body: a sum of random Gaussian blobs smoothed with sigma
// (analytically: each blob of radius r becomes a Gaussian of variance
// r^2 + sigma^2), thresholded into a closed wall of density 1 around an
// interior whose density ramps from 0.15 at the wall to 0.55 at the core;
// density maps to colour through a linear ramp.  empty_frac of the voxels
// (by the field's quantile) are left empty; wall_frac of the body is wall.
// The field is evaluated on a grid 4x coarser per axis and upsampled
// trilinearly (the smoothed field varies over >= sigma voxels).
__attribute__((visibility("default"))) void xds_ct_body(uint8_t* out, uint64_t X, uint64_t Y, uint64_t Z,
                                                        uint64_t seed, uint32_t n_blobs, double sigma,
                                                        double empty_frac, double wall_frac) {
    Rng rng(seed);
    struct Blob {
        double c[3], inv2s2, w;
    };
    std::vector<Blob> blobs(n_blobs);
    const double dims[3] = {(double)X, (double)Y, (double)Z};
    for (auto& b : blobs) {
        for (int i = 0; i < 3; ++i) b.c[i] = dims[i] * rng.range(0.3, 0.7);
        double r = rng.range(0.05, 0.14) * std::min({dims[0], dims[1], dims[2]}) + 4.0;
        b.inv2s2 = 1.0 / (2.0 * (r * r + sigma * sigma));
        b.w = rng.range(0.6, 1.0);
    }
    const uint64_t cs = 4;
    const uint64_t CX = X / cs + 2, CY = Y / cs + 2, CZ = Z / cs + 2;
    std::vector<float> coarse(CX * CY * CZ);
    parallel_z(CZ, [&](uint64_t cz) {
        for (uint64_t cy = 0; cy < CY; ++cy)
            for (uint64_t cx = 0; cx < CX; ++cx) {
                const double p[3] = {(double)(cx * cs), (double)(cy * cs), (double)(cz * cs)};
                double f = 0;
                for (const auto& b : blobs) {
                    double d2 = 0;
                    for (int i = 0; i < 3; ++i) d2 += (p[i] - b.c[i]) * (p[i] - b.c[i]);
                    f += b.w * std::exp(-d2 * b.inv2s2);
                }
                coarse[(cz * CY + cy) * CX + cx] = (float)f;
            }
    });
    auto field = [&](uint64_t x, uint64_t y, uint64_t z) {
        const uint64_t x0 = x / cs, y0 = y / cs, z0 = z / cs;
        const float fx = (float)(x % cs) / (float)cs, fy = (float)(y % cs) / (float)cs,
                    fz = (float)(z % cs) / (float)cs;
        auto at = [&](uint64_t a, uint64_t b, uint64_t c) { return coarse[(c * CY + b) * CX + a]; };
        float c00 = at(x0, y0, z0) * (1 - fx) + at(x0 + 1, y0, z0) * fx;
        float c10 = at(x0, y0 + 1, z0) * (1 - fx) + at(x0 + 1, y0 + 1, z0) * fx;
        float c01 = at(x0, y0, z0 + 1) * (1 - fx) + at(x0 + 1, y0, z0 + 1) * fx;
        float c11 = at(x0, y0 + 1, z0 + 1) * (1 - fx) + at(x0 + 1, y0 + 1, z0 + 1) * fx;
        float c0 = c00 * (1 - fy) + c10 * fy;
        float c1 = c01 * (1 - fy) + c11 * fy;
        return c0 * (1 - fz) + c1 * fz;
    };
    // thresholds from the coarse field's quantiles (one coarse sample stands
    // for cs^3 voxels, so its quantiles match the fine field's closely)
    std::vector<float> sorted;
    sorted.reserve((X / cs) * (Y / cs) * (Z / cs));
    for (uint64_t cz = 0; cz < Z / cs; ++cz)
        for (uint64_t cy = 0; cy < Y / cs; ++cy)
            for (uint64_t cx = 0; cx < X / cs; ++cx) sorted.push_back(coarse[(cz * CY + cy) * CX + cx]);
    std::sort(sorted.begin(), sorted.end());
    auto q = [&](double frac) {
        size_t i = (size_t)std::min<double>((double)(sorted.size() - 1), frac * (double)sorted.size());
        return sorted[i];
    };
    const float t_body = q(empty_frac);
    const float t_core = q(empty_frac + wall_frac * (1.0 - empty_frac));
    const float f_max = sorted.back();
    uint32_t* o = reinterpret_cast<uint32_t*>(out);
    parallel_z(Z, [&](uint64_t z) {
        for (uint64_t y = 0; y < Y; ++y)
            for (uint64_t x = 0; x < X; ++x) {
                float f = field(x, y, z);
                uint32_t v = 0;
                if (f >= t_body) {
                    double d;
                    if (f < t_core)
                        d = 1.0;  // wall
                    else
                        d = 0.15 + 0.40 * std::min(1.0, (double)(f - t_core) / (double)(f_max - t_core + 1e-12));
                    // linear colour ramp: dark blue (low density) -> warm white (high)
                    uint32_t r = (uint32_t)std::lround(30 + d * (255 - 30));
                    uint32_t g = (uint32_t)std::lround(50 + d * (220 - 50));
                    uint32_t b = (uint32_t)std::lround(160 + d * (150 - 160));
                    v = pack(r, g, b, 255);
                }
                o[(z * Y + y) * X + x] = v;
            }
    });
}

}  // extern "C"
