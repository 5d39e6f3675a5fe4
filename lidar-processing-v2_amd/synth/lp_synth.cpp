/*
 * lp_synth.cpp -- seeded synthetic LiDAR scans for tests and benchmarks.
 *
 * Casts one ray per (beam, azimuth bin) of a spinning multi-beam sensor at
 * the origin into a scene of a tilted noisy ground plane, cars (boxes),
 * poles (vertical cylinders), wall slabs and optional building facades, then
 * applies i.i.d. dropout, injects spurious below-ground returns and optional
 * rain/snow points.  Output is SoA x,y,z,intensity (f32) + ring (u16),
 * ordered ring-major (ring 0 = lowest beam) with ascending azimuth, which is
 * the layout of an organized cloud (src/dataloader/src/dataloader.cpp:68-137
 * assigns the same ring convention to scan-ordered KITTI clouds, counting
 * down from the top beam).
 *
 * The PRNG is splitmix64 (Steele, Lea, Flood 2014), so a seed fixes the
 * scene independent of any library's random distributions.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {
typedef struct lp_synth_spec {
    int32_t height;          /* beams = range-image rows */
    int32_t width;           /* azimuth bins */
    float fov_up_deg;
    float fov_down_deg;
    int32_t beam_layout;     /* 0: two-block HDL-64E-like spacing, 1: uniform */
    float dropout;           /* i.i.d. probability that a ray returns nothing */
    int32_t n_cars;          /* 4.5 x 1.8 x 1.5 m boxes */
    int32_t n_poles;         /* r 0.15 m, h 3 m cylinders */
    int32_t n_walls;         /* 8-20 m x 0.3 m x 3 m slabs */
    int32_t facades;         /* 1: two continuous facades at 8-25 m lateral offset */
    float obj_range_min;
    float obj_range_max;
    int32_t n_below;         /* spurious returns below the ground within 10 m */
    int32_t n_rain;          /* rain/snow points uniform in a 1-30 m shell */
    float max_range;         /* rays that hit nothing closer return nothing */
    float enclosure_radius;  /* >0: distant 25 m high 16-sided terrain ring (returns for upward beams) */
    uint64_t seed;
} lp_synth_spec;
}

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        double u1 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

struct Box {  // oriented box standing on the ground
    double cx, cy, z0, z1, c, s, hl, hw;
};
struct Cyl {
    double cx, cy, r, z0, z1;
};

constexpr double kInf = 1e300;

double hitBox(const Box& b, double dx, double dy, double dz) {
    // Ray from the origin; box frame: translate by (cx,cy), rotate by -yaw.
    const double ox = -b.cx, oy = -b.cy;
    const double lox = b.c * ox + b.s * oy, loy = -b.s * ox + b.c * oy;
    const double ldx = b.c * dx + b.s * dy, ldy = -b.s * dx + b.c * dy;
    double t0 = 0.0, t1 = kInf;
    auto slab = [&](double o, double d, double lo, double hi) {
        if (std::fabs(d) < 1e-15) return o >= lo && o <= hi;
        double a = (lo - o) / d, c = (hi - o) / d;
        if (a > c) std::swap(a, c);
        t0 = std::max(t0, a);
        t1 = std::min(t1, c);
        return t0 <= t1;
    };
    if (!slab(lox, ldx, -b.hl, b.hl)) return kInf;
    if (!slab(loy, ldy, -b.hw, b.hw)) return kInf;
    if (!slab(0.0, dz, b.z0, b.z1)) return kInf;
    return t0 > 1e-6 ? t0 : kInf;
}

double hitCyl(const Cyl& c, double dx, double dy, double dz) {
    const double a = dx * dx + dy * dy;
    if (a < 1e-15) return kInf;
    const double bq = -2.0 * (dx * c.cx + dy * c.cy);
    const double cq = c.cx * c.cx + c.cy * c.cy - c.r * c.r;
    const double disc = bq * bq - 4.0 * a * cq;
    if (disc < 0.0) return kInf;
    const double t = (-bq - std::sqrt(disc)) / (2.0 * a);
    if (t <= 1e-6) return kInf;
    const double z = t * dz;
    return (z >= c.z0 && z <= c.z1) ? t : kInf;
}

}  // namespace

extern "C" {

/* Fills up to `capacity` points; returns the number written, or -1 if the
 * capacity is too small. */
int32_t lp_synth_frame(const lp_synth_spec* sp, int32_t capacity, float* x, float* y, float* z,
                       float* intensity, uint16_t* ring) {
    Rng rng{sp->seed};
    const int32_t H = sp->height, W = sp->width;
    // Beam elevations, ring 0 = lowest.
    std::vector<double> elev(H);
    const double up = sp->fov_up_deg, down = sp->fov_down_deg;
    if (sp->beam_layout == 0 && H >= 4) {
        // Upper half of the beams share the top 39% of the field of view
        // (HDL-64E: +2.0 .. -8.33 deg), the lower half the rest (-8.83 .. -24.8).
        const int32_t hu = H / 2, hl = H - hu;
        const double split_hi = up + (down - up) * 0.3866;
        const double split_lo = up + (down - up) * 0.4052;
        for (int32_t k = 0; k < hl; ++k) elev[k] = down + (split_lo - down) * k / (hl - 1);
        for (int32_t k = 0; k < hu; ++k) elev[hl + k] = split_hi + (up - split_hi) * k / (hu - 1);
    } else {
        for (int32_t k = 0; k < H; ++k) elev[k] = down + (up - down) * k / (H > 1 ? H - 1 : 1);
    }
    // Ground: z = c0 + ga*x + gb*y (sensor pitch/roll within +-2 deg plus a 3% grade).
    const double pitch = rng.uniform(-2.0, 2.0) * M_PI / 180.0;
    const double roll = rng.uniform(-2.0, 2.0) * M_PI / 180.0;
    const double grade_dir = rng.uniform(0.0, 2.0 * M_PI);
    const double ga = std::tan(pitch) + 0.03 * std::cos(grade_dir);
    const double gb = std::tan(roll) + 0.03 * std::sin(grade_dir);
    const double c0 = -1.73;
    auto ground_z = [&](double px, double py) { return c0 + ga * px + gb * py; };

    std::vector<Box> boxes;
    std::vector<Cyl> cyls;
    double encl = 0.0;
    auto place = [&](double& px, double& py) {
        const double r = rng.uniform(sp->obj_range_min, sp->obj_range_max);
        const double a = rng.uniform(0.0, 2.0 * M_PI);
        px = r * std::cos(a);
        py = r * std::sin(a);
    };
    for (int32_t i = 0; i < sp->n_cars; ++i) {
        double px, py;
        place(px, py);
        const double yaw = rng.uniform(0.0, M_PI);
        const double g = ground_z(px, py);
        boxes.push_back({px, py, g, g + 1.5, std::cos(yaw), std::sin(yaw), 2.25, 0.9});
    }
    for (int32_t i = 0; i < sp->n_walls; ++i) {
        double px, py;
        place(px, py);
        const double yaw = rng.uniform(0.0, M_PI);
        const double len = rng.uniform(8.0, 20.0);
        const double g = ground_z(px, py);
        boxes.push_back({px, py, g, g + 3.0, std::cos(yaw), std::sin(yaw), len * 0.5, 0.15});
    }
    if (sp->facades) {
        for (int side = -1; side <= 1; side += 2) {
            const double off = rng.uniform(8.0, 25.0);
            const double cy = side * (off + 2.5);
            boxes.push_back({0.0, cy, c0 - 1.0, c0 + rng.uniform(10.0, 20.0), 1.0, 0.0, 150.0, 2.5});
        }
    }
    for (int32_t i = 0; i < sp->n_poles; ++i) {
        double px, py;
        place(px, py);
        const double g = ground_z(px, py);
        cyls.push_back({px, py, 0.15, g, g + 3.0});
    }

    if (sp->enclosure_radius > 0.0F) {
        // Distant terrain ring: a vertical cylinder wall seen from the inside.
        encl = sp->enclosure_radius;
    }

    struct Ret { double x, y, z; double t; bool ground; };
    std::vector<Ret> rets(static_cast<size_t>(H) * W);
    std::vector<uint8_t> valid(static_cast<size_t>(H) * W, 0);
    for (int32_t k = 0; k < H; ++k) {
        const double el = elev[k] * M_PI / 180.0;
        const double ce = std::cos(el), se = std::sin(el);
        for (int32_t j = 0; j < W; ++j) {
            const double az = (j + 0.5) * 2.0 * M_PI / W;
            const double dx = ce * std::cos(az), dy = ce * std::sin(az), dz = se;
            double best = sp->max_range;
            bool ground = false;
            const double den = dz - ga * dx - gb * dy;
            if (den < -1e-12) {
                const double t = c0 / den;
                if (t > 0.0 && t < best) {
                    best = t;
                    ground = true;
                }
            }
            for (const auto& b : boxes) {
                const double t = hitBox(b, dx, dy, dz);
                if (t < best) {
                    best = t;
                    ground = false;
                }
            }
            for (const auto& c : cyls) {
                const double t = hitCyl(c, dx, dy, dz);
                if (t < best) {
                    best = t;
                    ground = false;
                }
            }
            if (encl > 0.0) {
                // Terrain ring: 16 straight wall segments whose vertices lie
                // on a circle of radius encl, seen from the inside.
                const double seg = 2.0 * M_PI / 16.0;
                const int k = static_cast<int>(az / seg) % 16;
                const double nang = (k + 0.5) * seg;
                const double den_w = std::cos(nang) * dx + std::sin(nang) * dy;
                if (den_w > 1e-9) {
                    const double t = encl * std::cos(0.5 * seg) / den_w;
                    const double zt = t * dz;
                    if (t < best && zt <= c0 + 25.0) {
                        best = t;
                        ground = false;
                    }
                }
            }
            const size_t idx = static_cast<size_t>(k) * W + j;
            if (best >= sp->max_range) continue;
            Ret r{best * dx, best * dy, best * dz, best, ground};
            if (ground) r.z += 0.02 * rng.normal();
            rets[idx] = r;
            valid[idx] = 1;
        }
    }
    // Spurious below-ground returns: multipath that lengthens a near ground ray.
    if (sp->n_below > 0) {
        std::vector<size_t> near;
        for (size_t i = 0; i < rets.size(); ++i)
            if (valid[i] && rets[i].ground && std::hypot(rets[i].x, rets[i].y) < 10.0) near.push_back(i);
        for (int32_t b = 0; b < sp->n_below && !near.empty(); ++b) {
            const size_t pick = near[rng.next() % near.size()];
            const double f = rng.uniform(1.15, 1.45);
            rets[pick].x *= f;
            rets[pick].y *= f;
            rets[pick].z = rets[pick].z * f;
            rets[pick].ground = false;
        }
    }
    int32_t n = 0;
    for (int32_t k = 0; k < H; ++k)
        for (int32_t j = 0; j < W; ++j) {
            const size_t idx = static_cast<size_t>(k) * W + j;
            const double u = rng.uniform();
            const double inten = rng.uniform();
            if (!valid[idx] || u < sp->dropout) continue;
            if (n >= capacity) return -1;
            x[n] = static_cast<float>(rets[idx].x);
            y[n] = static_cast<float>(rets[idx].y);
            z[n] = static_cast<float>(rets[idx].z);
            intensity[n] = static_cast<float>(inten);
            ring[n] = static_cast<uint16_t>(k);
            ++n;
        }
    for (int32_t i = 0; i < sp->n_rain; ++i) {
        const double r = rng.uniform(1.0, 30.0);
        const double a = rng.uniform(0.0, 2.0 * M_PI);
        const double pz = rng.uniform(-1.2, 3.0);
        const double ph = std::sqrt(std::max(r * r - pz * pz, 0.25));
        if (n >= capacity) return -1;
        x[n] = static_cast<float>(ph * std::cos(a));
        y[n] = static_cast<float>(ph * std::sin(a));
        z[n] = static_cast<float>(pz);
        intensity[n] = static_cast<float>(rng.uniform());
        const double el = std::atan2(pz, ph) * 180.0 / M_PI;
        int32_t best = 0;
        for (int32_t k = 1; k < H; ++k)
            if (std::fabs(elev[k] - el) < std::fabs(elev[best] - el)) best = k;
        ring[n] = static_cast<uint16_t>(best);
        ++n;
    }
    return n;
}

}  // extern "C"
