// synth.cpp -- seeded synthetic binary inputs for the thinning workloads.
//
// Every generator is deterministic: a SplitMix64-seeded xoshiro256** stream
// (both published generators, Steele et al. 2014 / Blackman & Vigna 2018),
// uniforms from the top 53 bits, no std::*_distribution, and only +,-,*,/,
// sqrt, sin, cos on doubles (built with -ffp-contract=off).
//
//   polylines : W x H uint8, n_lines polylines of 3..8 vertices uniform in the
//               frame, each a tube of width U[3,15) px (radius w/2), plus salt
//               noise with probability p_salt per pixel.
//   tree2d    : W x H uint8 binary branching trees, each rooted on a random
//               border point heading inward +-U(45 deg); depth 6 (127
//               branches), branch length U[40,200), radius 8 - 7*depth/6,
//               child heading parent +- U(45 deg); trees are added until the
//               foreground fraction reaches a target; salt noise p_salt.
//   tree3d    : W x H x D uint8 vessel trees, each rooted on a random face,
//               depth 7 (255 branches), length U[10,60), radius 8 - depth,
//               child heading the parent's rotated by U[0,45) deg about a
//               random perpendicular axis; plus n_ellipsoids solid
//               axis-aligned ellipsoids with semi-axes U[3,12) and uniform
//               centres; trees are added until a target foreground fraction.
// A pixel/voxel with integer coordinates p is inside a tube iff its Euclidean
// distance to the branch segment is <= radius.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t t = seed;
        for (auto &v : s) v = splitmix64(t);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uni(double a, double b) { return a + (b - a) * uni(); }
    // integer uniform in [lo, hi]
    int uint_in(int lo, int hi) { return lo + (int)(uni() * (double)(hi - lo + 1)); }
};

const double kPi = 3.14159265358979323846;

// ---- 2D -------------------------------------------------------------------

inline double seg_dist2_2d(double px, double py, double x0, double y0, double x1, double y1)
{
    double dx = x1 - x0, dy = y1 - y0;
    double l2 = dx * dx + dy * dy;
    double t = 0.0;
    if (l2 > 0.0) {
        t = ((px - x0) * dx + (py - y0) * dy) / l2;
        if (t < 0.0) t = 0.0;
        if (t > 1.0) t = 1.0;
    }
    double qx = x0 + t * dx - px, qy = y0 + t * dy - py;
    return qx * qx + qy * qy;
}

// Row interval of the capsule (convex), widened; every candidate is then
// decided by the exact distance test.
void capsule2d(uint8_t *img, int w, int h, double x0, double y0, double x1, double y1, double r)
{
    double r2 = r * r;
    int ylo = (int)std::floor(std::fmin(y0, y1) - r) - 1;
    int yhi = (int)std::ceil(std::fmax(y0, y1) + r) + 1;
    if (ylo < 0) ylo = 0;
    if (yhi > h - 1) yhi = h - 1;
    double dx = x1 - x0, dy = y1 - y0, L = std::sqrt(dx * dx + dy * dy);
    for (int y = ylo; y <= yhi; y++) {
        double lo = 1e300, hi = -1e300;
        auto disc = [&](double cx, double cy) {
            double e = r2 - (y - cy) * (y - cy);
            if (e >= 0.0) {
                double s = std::sqrt(e);
                lo = std::fmin(lo, cx - s);
                hi = std::fmax(hi, cx + s);
            }
        };
        disc(x0, y0);
        disc(x1, y1);
        if (L > 0.0) {
            double ux = dx / L, uy = dy / L;
            // t(x) = (x-x0)ux + (y-y0)uy in [0,L];  |(x-x0)(-uy) + (y-y0)ux| <= r
            double a = -1e300, b = 1e300;
            double cy_ = (y - y0);
            if (std::fabs(ux) > 1e-12) {
                double ta = (0.0 - cy_ * uy) / ux + x0, tb = (L - cy_ * uy) / ux + x0;
                a = std::fmax(a, std::fmin(ta, tb));
                b = std::fmin(b, std::fmax(ta, tb));
            } else if (!(cy_ * uy >= 0.0 && cy_ * uy <= L)) {
                b = -1e300;
            }
            if (std::fabs(uy) > 1e-12) {
                double pa = (cy_ * ux - r) / uy + x0, pb = (cy_ * ux + r) / uy + x0;
                a = std::fmax(a, std::fmin(pa, pb));
                b = std::fmin(b, std::fmax(pa, pb));
            } else if (!(std::fabs(cy_ * ux) <= r)) {
                b = -1e300;
            }
            if (a <= b) {
                lo = std::fmin(lo, a);
                hi = std::fmax(hi, b);
            }
        }
        if (lo > hi) continue;
        int xa = (int)std::floor(lo) - 1, xb = (int)std::ceil(hi) + 1;
        if (xa < 0) xa = 0;
        if (xb > w - 1) xb = w - 1;
        uint8_t *row = img + (size_t)y * w;
        for (int x = xa; x <= xb; x++)
            if (!row[x] && seg_dist2_2d(x, y, x0, y0, x1, y1) <= r2) row[x] = 1;
    }
}

void salt(uint8_t *img, size_t n, double p, Rng &rng)
{
    if (p <= 0.0) return;
    for (size_t i = 0; i < n; i++)
        if (rng.uni() < p) img[i] = 1;
}

// ---- 3D -------------------------------------------------------------------

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 unit(V3 a)
{
    double l = std::sqrt(dot(a, a));
    return l > 0 ? (1.0 / l) * a : V3{1, 0, 0};
}

inline double seg_dist2_3d(V3 p, V3 a, V3 b)
{
    V3 d = b - a;
    double l2 = dot(d, d), t = 0.0;
    if (l2 > 0.0) {
        t = dot(p - a, d) / l2;
        if (t < 0.0) t = 0.0;
        if (t > 1.0) t = 1.0;
    }
    V3 q = a + t * d - p;
    return dot(q, q);
}

void capsule3d(uint8_t *vol, int w, int h, int dd, V3 a, V3 b, double r)
{
    double r2 = r * r;
    auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
    int zlo = clampi((int)std::floor(std::fmin(a.z, b.z) - r) - 1, 0, dd - 1);
    int zhi = clampi((int)std::ceil(std::fmax(a.z, b.z) + r) + 1, 0, dd - 1);
    int ylo = clampi((int)std::floor(std::fmin(a.y, b.y) - r) - 1, 0, h - 1);
    int yhi = clampi((int)std::ceil(std::fmax(a.y, b.y) + r) + 1, 0, h - 1);
    int xlo = clampi((int)std::floor(std::fmin(a.x, b.x) - r) - 1, 0, w - 1);
    int xhi = clampi((int)std::ceil(std::fmax(a.x, b.x) + r) + 1, 0, w - 1);
    if (a.z - r > dd + 1 && b.z - r > dd + 1) return;
    V3 d = b - a;
    double L = std::sqrt(dot(d, d));
    V3 u = L > 0 ? (1.0 / L) * d : V3{1, 0, 0};
    for (int z = zlo; z <= zhi; z++)
        for (int y = ylo; y <= yhi; y++) {
            double lo = 1e300, hi = -1e300;
            auto sphere = [&](V3 c) {
                double e = r2 - (y - c.y) * (y - c.y) - (z - c.z) * (z - c.z);
                if (e >= 0.0) {
                    double s = std::sqrt(e);
                    lo = std::fmin(lo, c.x - s);
                    hi = std::fmax(hi, c.x + s);
                }
            };
            sphere(a);
            sphere(b);
            if (L > 0.0) {
                // p(x) = (x, y, z); w = p - a = (x - a.x, wy, wz)
                double wy = y - a.y, wz = z - a.z;
                double ct = wy * u.y + wz * u.z;  // t(x) = u.x (x - a.x) + ct
                // perp^2 = |w|^2 - t^2 = s^2 + wy^2 + wz^2 - (u.x s + ct)^2, s = x - a.x
                double A = 1.0 - u.x * u.x, B = -2.0 * u.x * ct, C = wy * wy + wz * wz - ct * ct - r2;
                double s0 = -1e300, s1 = 1e300;
                bool ok = true;
                if (A > 1e-12) {
                    double disc = B * B - 4 * A * C;
                    if (disc < 0) ok = false;
                    else {
                        double q = std::sqrt(disc);
                        s0 = (-B - q) / (2 * A);
                        s1 = (-B + q) / (2 * A);
                    }
                } else if (C > 0) {
                    ok = false;
                }
                if (ok) {
                    if (std::fabs(u.x) > 1e-12) {
                        double ta = (0.0 - ct) / u.x, tb = (L - ct) / u.x;
                        s0 = std::fmax(s0, std::fmin(ta, tb));
                        s1 = std::fmin(s1, std::fmax(ta, tb));
                    } else if (!(ct >= 0.0 && ct <= L)) {
                        ok = false;
                    }
                }
                if (ok && s0 <= s1) {
                    lo = std::fmin(lo, s0 + a.x);
                    hi = std::fmax(hi, s1 + a.x);
                }
            }
            if (lo > hi) continue;
            int xa = clampi((int)std::floor(lo) - 1, xlo, xhi);
            int xb = clampi((int)std::ceil(hi) + 1, xlo, xhi);
            uint8_t *row = vol + ((size_t)z * h + y) * w;
            for (int x = xa; x <= xb; x++)
                if (!row[x] && seg_dist2_3d(V3{(double)x, (double)y, (double)z}, a, b) <= r2) row[x] = 1;
        }
}

void ellipsoid(uint8_t *vol, int w, int h, int dd, V3 c, V3 s)
{
    auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
    int z0 = clampi((int)std::floor(c.z - s.z), 0, dd - 1), z1 = clampi((int)std::ceil(c.z + s.z), 0, dd - 1);
    int y0 = clampi((int)std::floor(c.y - s.y), 0, h - 1), y1 = clampi((int)std::ceil(c.y + s.y), 0, h - 1);
    int x0 = clampi((int)std::floor(c.x - s.x), 0, w - 1), x1 = clampi((int)std::ceil(c.x + s.x), 0, w - 1);
    for (int z = z0; z <= z1; z++)
        for (int y = y0; y <= y1; y++)
            for (int x = x0; x <= x1; x++) {
                double ex = (x - c.x) / s.x, ey = (y - c.y) / s.y, ez = (z - c.z) / s.z;
                if (ex * ex + ey * ey + ez * ez <= 1.0) vol[((size_t)z * h + y) * w + x] = 1;
            }
}

// rotate v by angle t about unit axis k (Rodrigues)
V3 rotate(V3 v, V3 k, double t)
{
    double c = std::cos(t), s = std::sin(t);
    return c * v + s * cross(k, v) + ((1 - c) * dot(k, v)) * k;
}

V3 random_perpendicular(V3 d, Rng &rng)
{
    V3 a = std::fabs(d.x) < 0.9 ? V3{1, 0, 0} : V3{0, 1, 0};
    V3 p = unit(cross(d, a));
    V3 q = cross(d, p);
    double phi = rng.uni(0.0, 2 * kPi);
    return unit(std::cos(phi) * p + std::sin(phi) * q);
}

}  // namespace

extern "C" {

uint64_t bgt_synth_item_seed(int config, int item)
{
    uint64_t s = 0xBE6E0000ull + (uint64_t)config + ((uint64_t)item << 32);
    return splitmix64(s);
}

int bgt_synth_polylines(uint8_t *img, int w, int h, uint64_t seed, int n_lines, double p_salt)
{
    if (!img || w <= 0 || h <= 0) return -1;
    std::memset(img, 0, (size_t)w * h);
    Rng rng(seed);
    for (int l = 0; l < n_lines; l++) {
        int k = rng.uint_in(3, 8);
        double width = rng.uni(3.0, 15.0);
        double r = width * 0.5;
        double px = rng.uni(0.0, (double)w), py = rng.uni(0.0, (double)h);
        for (int v = 1; v < k; v++) {
            double qx = rng.uni(0.0, (double)w), qy = rng.uni(0.0, (double)h);
            capsule2d(img, w, h, px, py, qx, qy, r);
            px = qx; py = qy;
        }
    }
    salt(img, (size_t)w * h, p_salt, rng);
    return 0;
}

static void one_tree2d(uint8_t *img, int w, int h, int depth, Rng &rng);

// Trees are added (each drawn from the same stream) until the foreground
// fraction reaches target_fraction (at least one tree, at most 64).
int bgt_synth_tree2d(uint8_t *img, int w, int h, uint64_t seed, int depth, double p_salt,
                     double target_fraction)
{
    if (!img || w <= 0 || h <= 0) return -1;
    std::memset(img, 0, (size_t)w * h);
    Rng rng(seed);
    const size_t n = (size_t)w * h;
    for (int t = 0; t < 64; ++t) {
        one_tree2d(img, w, h, depth, rng);
        size_t fg = 0;
        for (size_t i = 0; i < n; ++i) fg += img[i];
        if ((double)fg >= target_fraction * (double)n) break;
    }
    salt(img, n, p_salt, rng);
    return 0;
}

static void one_tree2d(uint8_t *img, int w, int h, int depth, Rng &rng)
{
    // root on a random border point, heading inward +- 45 deg
    int side = rng.uint_in(0, 3);
    double t = rng.uni();
    double x0, y0, head;
    switch (side) {
    case 0: x0 = t * (w - 1); y0 = 0; head = kPi / 2; break;          // top, heading +y
    case 1: x0 = w - 1; y0 = t * (h - 1); head = kPi; break;          // right, heading -x
    case 2: x0 = t * (w - 1); y0 = h - 1; head = -kPi / 2; break;     // bottom, heading -y
    default: x0 = 0; y0 = t * (h - 1); head = 0; break;               // left, heading +x
    }
    head += rng.uni(-kPi / 4, kPi / 4);
    struct B { double x, y, a; int d; };
    std::vector<B> stack{{x0, y0, head, 0}};
    // depth-first, left child before right child
    while (!stack.empty()) {
        B b = stack.back();
        stack.pop_back();
        double len = rng.uni(40.0, 200.0);
        double r = 8.0 - 7.0 * b.d / (double)(depth > 0 ? depth : 1);
        double x1 = b.x + len * std::cos(b.a), y1 = b.y + len * std::sin(b.a);
        capsule2d(img, w, h, b.x, b.y, x1, y1, r);
        if (b.d < depth) {
            double a1 = b.a + rng.uni(-kPi / 4, kPi / 4);
            double a2 = b.a + rng.uni(-kPi / 4, kPi / 4);
            stack.push_back({x1, y1, a2, b.d + 1});
            stack.push_back({x1, y1, a1, b.d + 1});
        }
    }
}

static void one_tree3d(uint8_t *vol, int w, int h, int d, int depth, Rng &rng);

// One vessel tree, the ellipsoids, then further trees (same stream) until the
// foreground fraction reaches target_fraction (at most 64 trees).
int bgt_synth_tree3d(uint8_t *vol, int w, int h, int d, uint64_t seed, int depth, int n_ellipsoids,
                     double target_fraction)
{
    if (!vol || w <= 0 || h <= 0 || d <= 0) return -1;
    const size_t n = (size_t)w * h * d;
    std::memset(vol, 0, n);
    Rng rng(seed);
    one_tree3d(vol, w, h, d, depth, rng);
    for (int i = 0; i < n_ellipsoids; i++) {
        V3 c = {rng.uni(0.0, (double)w), rng.uni(0.0, (double)h), rng.uni(0.0, (double)d)};
        V3 sx = {rng.uni(3.0, 12.0), rng.uni(3.0, 12.0), rng.uni(3.0, 12.0)};
        ellipsoid(vol, w, h, d, c, sx);
    }
    for (int t = 1; t < 64; ++t) {
        size_t fg = 0;
        for (size_t i = 0; i < n; ++i) fg += vol[i];
        if ((double)fg >= target_fraction * (double)n) break;
        one_tree3d(vol, w, h, d, depth, rng);
    }
    return 0;
}

static void one_tree3d(uint8_t *vol, int w, int h, int d, int depth, Rng &rng)
{
    int face = rng.uint_in(0, 5);
    double s = rng.uni(), t = rng.uni();
    V3 p, dir;
    switch (face) {
    case 0: p = {0, s * (h - 1), t * (d - 1)}; dir = {1, 0, 0}; break;
    case 1: p = {(double)(w - 1), s * (h - 1), t * (d - 1)}; dir = {-1, 0, 0}; break;
    case 2: p = {s * (w - 1), 0, t * (d - 1)}; dir = {0, 1, 0}; break;
    case 3: p = {s * (w - 1), (double)(h - 1), t * (d - 1)}; dir = {0, -1, 0}; break;
    case 4: p = {s * (w - 1), t * (h - 1), 0}; dir = {0, 0, 1}; break;
    default: p = {s * (w - 1), t * (h - 1), (double)(d - 1)}; dir = {0, 0, -1}; break;
    }
    {
        V3 ax = random_perpendicular(dir, rng);
        dir = unit(rotate(dir, ax, rng.uni(0.0, kPi / 4)));
    }
    struct B { V3 p, dir; int dep; };
    std::vector<B> stack{{p, dir, 0}};
    while (!stack.empty()) {
        B b = stack.back();
        stack.pop_back();
        double len = rng.uni(10.0, 60.0);
        double r = 8.0 - 7.0 * b.dep / (double)(depth > 0 ? depth : 1);
        V3 e = b.p + len * b.dir;
        capsule3d(vol, w, h, d, b.p, e, r);
        if (b.dep < depth) {
            V3 ax1 = random_perpendicular(b.dir, rng);
            V3 d1 = unit(rotate(b.dir, ax1, rng.uni(0.0, kPi / 4)));
            V3 ax2 = random_perpendicular(b.dir, rng);
            V3 d2 = unit(rotate(b.dir, ax2, rng.uni(0.0, kPi / 4)));
            stack.push_back({e, d2, b.dep + 1});
            stack.push_back({e, d1, b.dep + 1});
        }
    }
}

// Mixed-size batch description: n2 images with w,h ~ U{lo2..hi2} and n3
// volumes with w,h,d ~ U{lo3..hi3}, drawn from one stream seeded by `seed`.
// dims2: n2*2 ints (w,h); dims3: n3*3 ints (w,h,d).
int bgt_synth_batch_dims(uint64_t seed, int n2, int lo2, int hi2, int *dims2, int n3, int lo3,
                         int hi3, int *dims3)
{
    Rng rng(seed);
    for (int i = 0; i < n2; i++) {
        dims2[2 * i] = rng.uint_in(lo2, hi2);
        dims2[2 * i + 1] = rng.uint_in(lo2, hi2);
    }
    for (int i = 0; i < n3; i++) {
        dims3[3 * i] = rng.uint_in(lo3, hi3);
        dims3[3 * i + 1] = rng.uint_in(lo3, hi3);
        dims3[3 * i + 2] = rng.uint_in(lo3, hi3);
    }
    return 0;
}

// Exponential inter-arrival gaps (seconds) at rate qps, and a per-request
// uniform in [0,1) for the mix; n entries each.
int bgt_synth_arrivals(uint64_t seed, int n, double qps, double *gaps, double *mix)
{
    Rng rng(seed);
    for (int i = 0; i < n; i++) {
        double u = rng.uni();
        gaps[i] = -std::log(1.0 - u) / qps;
        mix[i] = rng.uni();
    }
    return 0;
}

}  // extern "C"
