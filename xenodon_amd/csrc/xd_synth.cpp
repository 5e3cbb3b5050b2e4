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

// In-place radix-2 complex FFT of n = 2^k points spaced `stride` apart
// (inverse when inv; unnormalised).  Twiddles come from a table computed
// once per size with std::cos/std::sin.
void fft1d(double* re, double* im, uint64_t n, uint64_t stride, bool inv, const std::vector<double>& cs,
           const std::vector<double>& sn) {
    for (uint64_t i = 1, j = 0; i < n; ++i) {
        uint64_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            std::swap(re[i * stride], re[j * stride]);
            std::swap(im[i * stride], im[j * stride]);
        }
    }
    for (uint64_t len = 2; len <= n; len <<= 1) {
        const uint64_t step = n / len;
        for (uint64_t i = 0; i < n; i += len)
            for (uint64_t k = 0; k < len / 2; ++k) {
                const double wr = cs[k * step], wi = inv ? sn[k * step] : -sn[k * step];
                double* ar = re + (i + k) * stride;
                double* ai = im + (i + k) * stride;
                double* br = re + (i + k + len / 2) * stride;
                double* bi = im + (i + k + len / 2) * stride;
                const double tr = *br * wr - *bi * wi, ti = *br * wi + *bi * wr;
                *br = *ar - tr;
                *bi = *ai - ti;
                *ar += tr;
                *ai += ti;
            }
    }
}

// 3-D FFT of an n^3 cube (x fastest) along all three axes.
void fft3d(std::vector<double>& re, std::vector<double>& im, uint64_t n, bool inv) {
    std::vector<double> cs(n / 2), sn(n / 2);
    for (uint64_t k = 0; k < n / 2; ++k) {
        cs[k] = std::cos(2.0 * M_PI * (double)k / (double)n);
        sn[k] = std::sin(2.0 * M_PI * (double)k / (double)n);
    }
    const uint64_t strides[3] = {1, n, n * n};
    for (int axis = 0; axis < 3; ++axis) {
        const uint64_t st = strides[axis];
        // the n^2 lines along this axis, split over threads by their outer index
        parallel_z(n, [&](uint64_t a) {
            for (uint64_t b = 0; b < n; ++b) {
                uint64_t base;
                if (axis == 0) base = (a * n + b) * n;
                else if (axis == 1) base = a * n * n + b;
                else base = a * n + b;
                fft1d(re.data() + base, im.data() + base, n, st, inv, cs, sn);
            }
        });
    }
}

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

// Cosmological-gas-like volume: a Gaussian random field with power spectrum
// P(k) ~ k^slope on a base^3 periodic grid (white noise filtered in Fourier
// space by sqrt(P)), exponentiated into a log-normal density, upsampled
// trilinearly to X x Y x Z, and thresholded at the `keep_above` quantile of
// the density (voxels below it are empty, RGBA 0).  Emission is the
// log-density through a colour ramp (deep red -> orange -> pale yellow),
// alpha 255.
__attribute__((visibility("default"))) void xds_lognormal(void* out, uint64_t X, uint64_t Y, uint64_t Z,
                                                          uint64_t seed, uint32_t base, double slope,
                                                          double sigma_ln, double keep_above) {
    const uint64_t n = base, n3 = n * n * n;
    std::vector<double> re(n3), im(n3, 0.0);
    Rng rng(seed);
    for (uint64_t i = 0; i < n3; i += 2) {  // Box-Muller pairs
        const double u1 = 1.0 - rng.uniform(), u2 = rng.uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        re[i] = r * std::cos(2.0 * M_PI * u2);
        if (i + 1 < n3) re[i + 1] = r * std::sin(2.0 * M_PI * u2);
    }
    fft3d(re, im, n, false);
    parallel_z(n, [&](uint64_t z) {
        const double kz = (double)(z <= n / 2 ? z : n - z);
        for (uint64_t y = 0; y < n; ++y) {
            const double ky = (double)(y <= n / 2 ? y : n - y);
            for (uint64_t x = 0; x < n; ++x) {
                const double kx = (double)(x <= n / 2 ? x : n - x);
                const double k2 = kx * kx + ky * ky + kz * kz;
                const double amp = k2 > 0 ? std::pow(k2, slope / 4.0) : 0.0;  // sqrt(P) = k^(slope/2)
                const uint64_t i = (z * n + y) * n + x;
                re[i] *= amp;
                im[i] *= amp;
            }
        }
    });
    fft3d(re, im, n, true);
    // unit variance, then log-normal density
    double s1 = 0, s2 = 0;
    for (uint64_t i = 0; i < n3; ++i) {
        s1 += re[i];
        s2 += re[i] * re[i];
    }
    const double mean = s1 / (double)n3, sd = std::sqrt(std::max(s2 / (double)n3 - mean * mean, 1e-300));
    std::vector<float> lnrho(n3);
    for (uint64_t i = 0; i < n3; ++i) lnrho[i] = (float)(sigma_ln * (re[i] - mean) / sd);
    std::vector<double>().swap(re);
    std::vector<double>().swap(im);
    // trilinear upsampling of the log-density (cell-centred, clamped at the
    // faces); the density is its exponential, so thresholds and the colour
    // ramp work on the interpolated value directly.  Per-axis cells and
    // weights are tabulated once.
    struct Axis {
        std::vector<uint64_t> i0, i1;
        std::vector<double> a;
    };
    auto axis = [&](uint64_t m) {
        Axis A;
        A.i0.resize(m);
        A.i1.resize(m);
        A.a.resize(m);
        for (uint64_t v = 0; v < m; ++v) {
            const double f = std::min(std::max(((double)v + 0.5) * (double)n / (double)m - 0.5, 0.0), (double)(n - 1));
            A.i0[v] = (uint64_t)f;
            A.i1[v] = std::min(A.i0[v] + 1, n - 1);
            A.a[v] = f - (double)A.i0[v];
        }
        return A;
    };
    const Axis ax_x = axis(X), ax_y = axis(Y), ax_z = axis(Z);
    auto log_density = [&](uint64_t x, uint64_t y, uint64_t z) {
        const uint64_t x0 = ax_x.i0[x], x1 = ax_x.i1[x], y0 = ax_y.i0[y], y1 = ax_y.i1[y];
        const uint64_t z0 = ax_z.i0[z], z1 = ax_z.i1[z];
        const double ax = ax_x.a[x], ay = ax_y.a[y], az = ax_z.a[z];
        auto at = [&](uint64_t i, uint64_t j, uint64_t k) { return (double)lnrho[(k * n + j) * n + i]; };
        const double c00 = at(x0, y0, z0) * (1 - ax) + at(x1, y0, z0) * ax;
        const double c10 = at(x0, y1, z0) * (1 - ax) + at(x1, y1, z0) * ax;
        const double c01 = at(x0, y0, z1) * (1 - ax) + at(x1, y0, z1) * ax;
        const double c11 = at(x0, y1, z1) * (1 - ax) + at(x1, y1, z1) * ax;
        const double c0 = c00 * (1 - ay) + c10 * ay, c1 = c01 * (1 - ay) + c11 * ay;
        return c0 * (1 - az) + c1 * az;
    };
    // the threshold: quantile of the upsampled log-density from a histogram
    // of every voxel (exact up to one bin of 1/65536 of the log range)
    float lo = lnrho[0], hi = lnrho[0];
    for (float v : lnrho) {
        lo = std::min(lo, v);
        hi = std::max(hi, v);
    }
    const int bins = 1 << 16;
    const double blo = (double)lo, bw = ((double)hi - (double)lo) / bins + 1e-12;
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::vector<uint64_t>> hist(nt, std::vector<uint64_t>(bins, 0));
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t)
            th.emplace_back([&, t] {
                for (uint64_t z = t; z < Z; z += nt)
                    for (uint64_t y = 0; y < Y; ++y)
                        for (uint64_t x = 0; x < X; ++x) {
                            const int b = (int)std::min<double>(bins - 1, std::max(0.0, (log_density(x, y, z) - blo) / bw));
                            ++hist[t][b];
                        }
            });
        for (auto& t : th) t.join();
    }
    const uint64_t total = X * Y * Z, want = (uint64_t)(keep_above * (double)total);
    uint64_t acc = 0;
    int cut = 0;
    for (; cut < bins; ++cut) {
        uint64_t c = 0;
        for (unsigned t = 0; t < nt; ++t) c += hist[t][cut];
        if (acc + c > want) break;
        acc += c;
    }
    const double ln_thr = blo + cut * bw, ln_max = (double)hi;
    uint32_t* o = reinterpret_cast<uint32_t*>(out);
    parallel_z(Z, [&](uint64_t z) {
        for (uint64_t y = 0; y < Y; ++y)
            for (uint64_t x = 0; x < X; ++x) {
                const double l = log_density(x, y, z);
                uint32_t v = 0;
                if (l >= ln_thr) {
                    const double t = std::min(1.0, (l - ln_thr) / (ln_max - ln_thr + 1e-12));
                    const uint32_t r = (uint32_t)std::lround(90 + t * (255 - 90));
                    const uint32_t g = (uint32_t)std::lround(10 + t * (240 - 10));
                    const uint32_t b = (uint32_t)std::lround(20 + t * t * (170 - 20));
                    v = pack(r, g, b, 255);
                }
                o[(z * Y + y) * X + x] = v;
            }
    });
}

}  // extern "C"
