// Deterministic procedural scenes (no reference counterpart: the reference
// loads OBJ files only).  Generation uses a splitmix64 stream and double
// arithmetic rounded to float at the end, so the triangle bits depend only on
// the arguments.
#include <cmath>
#include <vector>

#include "scene.hpp"

namespace pray {

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 unit(V3 a) {
    const double l = std::sqrt(dot(a, a));
    return a * (1.0 / l);
}
inline Vec3 f3(V3 a) { return Vec3((float)a.x, (float)a.y, (float)a.z); }

void addTri(Scene &s, V3 a, V3 b, V3 c, Vec3 col, bool emit = false) {
    s.triangles.emplace_back(f3(a), f3(b), f3(c), col, emit);
}

// ------------------------------------------------------------- terrain --

// Value noise: lattice values hashed from (seed, octave, ix, iz), smoothstep
// interpolation.
double lattice(uint64_t seed, int oct, int64_t ix, int64_t iz) {
    const uint64_t h = mix64(seed ^ mix64(0x51ED27ull * (uint64_t)(oct + 1) ^ mix64((uint64_t)ix * 0x9E3779B1ull) ^
                                          mix64((uint64_t)iz * 0x85EBCA77ull + 0xC2B2AE3Dull)));
    return (h >> 11) * (1.0 / 9007199254740992.0);
}

double valueNoise(uint64_t seed, int oct, double x, double z) {
    const double fx = std::floor(x), fz = std::floor(z);
    const int64_t ix = (int64_t)fx, iz = (int64_t)fz;
    double tx = x - fx, tz = z - fz;
    tx = tx * tx * (3 - 2 * tx);
    tz = tz * tz * (3 - 2 * tz);
    const double v00 = lattice(seed, oct, ix, iz), v10 = lattice(seed, oct, ix + 1, iz);
    const double v01 = lattice(seed, oct, ix, iz + 1), v11 = lattice(seed, oct, ix + 1, iz + 1);
    const double a = v00 + (v10 - v00) * tx, b = v01 + (v11 - v01) * tx;
    return a + (b - a) * tz;
}

// ---------------------------------------------------------- icospheres --

void icosphere(Scene &s, V3 c, double r, int level, Vec3 col) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    const V3 v[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                      {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const int f[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8},  {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
                          {3, 8, 9},  {4, 9, 5},  {2, 4, 11}, {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
    std::vector<V3> tris;
    for (auto &fa : f) {
        V3 a = unit(v[fa[0]]), b = unit(v[fa[1]]), d = unit(v[fa[2]]);
        if (dot(cross(b - a, d - a), a + b + d) < 0) std::swap(b, d);  // outward winding
        tris.push_back(a), tris.push_back(b), tris.push_back(d);
    }
    for (int l = 0; l < level; l++) {
        std::vector<V3> nt;
        nt.reserve(tris.size() * 4);
        for (size_t i = 0; i < tris.size(); i += 3) {
            const V3 a = tris[i], b = tris[i + 1], d = tris[i + 2];
            const V3 ab = unit((a + b) * 0.5), bd = unit((b + d) * 0.5), da = unit((d + a) * 0.5);
            const V3 q[12] = {a, ab, da, b, bd, ab, d, da, bd, ab, bd, da};
            nt.insert(nt.end(), q, q + 12);
        }
        tris.swap(nt);
    }
    for (size_t i = 0; i < tris.size(); i += 3)
        addTri(s, c + tris[i] * r, c + tris[i + 1] * r, c + tris[i + 2] * r, col);
}

// Quad p0 + u*U + v*V (u,v in [0,1]) tessellated into mu x mv cells of two
// triangles; the face normal is U x V.
void quad(Scene &s, V3 p0, V3 U, V3 V, int mu, int mv, Vec3 col, bool emit = false) {
    for (int j = 0; j < mv; j++) {
        for (int i = 0; i < mu; i++) {
            const V3 a = p0 + U * (double(i) / mu) + V * (double(j) / mv);
            const V3 b = p0 + U * (double(i + 1) / mu) + V * (double(j) / mv);
            const V3 c = p0 + U * (double(i + 1) / mu) + V * (double(j + 1) / mv);
            const V3 d = p0 + U * (double(i) / mu) + V * (double(j + 1) / mv);
            addTri(s, a, b, c, col, emit);
            addTri(s, a, c, d, col, emit);
        }
    }
}

// Box of half extents h centred at c on the floor, rotated by `deg` about y,
// 6 faces with outward normals, each tessellated with cell size `cell`.
void box(Scene &s, V3 c, V3 h, double deg, double cell, Vec3 col) {
    const double a = deg * 3.14159265358979323846 / 180.0;
    const V3 ex = {std::cos(a), 0, std::sin(a)}, ey = {0, 1, 0}, ez = {-std::sin(a), 0, std::cos(a)};
    const V3 X = ex * (2 * h.x), Y = ey * (2 * h.y), Z = ez * (2 * h.z);
    const V3 o = c - ex * h.x - ey * h.y - ez * h.z;
    auto cells = [&](double len) { return std::max(1, (int)std::lround(len / cell)); };
    const int nx = cells(2 * h.x), ny = cells(2 * h.y), nz = cells(2 * h.z);
    // quad() faces U x V; (ex, ey, ez) is right-handed.
    quad(s, o, X, Z, nx, nz, col);                // bottom (-y)
    quad(s, o + Y, Z, X, nz, nx, col);            // top (+y)
    quad(s, o, Y, X, ny, nx, col);                // front (-z)
    quad(s, o + Z, X, Y, nx, ny, col);            // back (+z)
    quad(s, o, Z, Y, nz, ny, col);                // left (-x)
    quad(s, o + X, Y, Z, ny, nz, col);            // right (+x)
}

}  // namespace

void makeTerrainScene(Scene &s, uint32_t grid, uint64_t seed) {
    s = Scene();
    const uint32_t nv = grid + 1;
    std::vector<float> h((size_t)nv * nv);
    double norm = 0;
    for (int o = 0; o < 6; o++) norm += std::pow(0.5, o);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)nv; j++) {
        for (uint32_t i = 0; i < nv; i++) {
            const double x = double(i) / grid, z = double(j) / grid;
            double acc = 0, amp = 1, freq = 4;
            for (int o = 0; o < 6; o++) {
                acc += amp * valueNoise(seed, o, x * freq, z * freq);
                amp *= 0.5;
                freq *= 2;
            }
            h[(size_t)j * nv + i] = (float)(0.1 * acc / norm);
        }
    }
    s.triangles.resize((size_t)2 * grid * grid);
    const Vec3 col(0.5f, 0.5f, 0.5f);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)grid; j++) {
        for (uint32_t i = 0; i < grid; i++) {
            auto P = [&](uint32_t ii, uint32_t jj) {
                return Vec3((float)(double(ii) / grid), h[(size_t)jj * nv + ii], (float)(double(jj) / grid));
            };
            const Vec3 v00 = P(i, j), v10 = P(i + 1, j), v01 = P(i, j + 1), v11 = P(i + 1, j + 1);
            const size_t k = 2 * ((size_t)j * grid + i);
            s.triangles[k] = Triangle(v00, v01, v11, col);
            s.triangles[k + 1] = Triangle(v00, v11, v10, col);
        }
    }
    s.width = s.height = 512;
    s.field_of_view = 60;
    s.camera_position = Vec3(0.5f, 0.6f, 1.6f);
    s.camera_look = Vec3(0.f, -0.45f, -1.f);
    s.lights.emplace_back(Vec3(0.5f, 2.f, 0.5f), Vec3(4.f, 4.f, 4.f));
    s.backgroundColor = Vec3(0.4f, 0.5f, 0.7f);
    s.seed = seed;
}

void makeSpheresScene(Scene &s, uint64_t seed) {
    s = Scene();
    SplitMix rng{seed};
    auto albedo = [&]() {
        return Vec3((float)rng.uniform(0.2, 0.9), (float)rng.uniform(0.2, 0.9), (float)rng.uniform(0.2, 0.9));
    };
    icosphere(s, {0, 1, 0}, 1.0, 4, albedo());
    for (int k = 0; k < 16; k++) {
        const double r = rng.uniform(0.1, 0.4);
        const double gx = -3.0 + 2.0 * (k % 4), gz = -3.0 + 2.0 * (k / 4);
        const double jx = rng.uniform(-0.4, 0.4), jz = rng.uniform(-0.4, 0.4);
        icosphere(s, {gx + jx, r, gz + jz}, r, 3, albedo());
    }
    const Vec3 g = albedo();
    addTri(s, {-6, 0, -6}, {-6, 0, 6}, {6, 0, 6}, g);
    addTri(s, {-6, 0, -6}, {6, 0, 6}, {6, 0, -6}, g);
    s.mode = RAYTRACING;
    s.width = s.height = 512;
    s.field_of_view = 60;
    s.camera_position = Vec3(0.f, 3.f, 8.f);
    s.camera_look = Vec3(0.f, -0.35f, -1.f);
    s.lights.emplace_back(Vec3(4.f, 8.f, 6.f), Vec3(80.f, 80.f, 80.f));
    s.backgroundColor = Vec3(0.1f, 0.1f, 0.15f);
    s.treeKind = 0;
    s.seed = seed;
}

void makeCornellScene(Scene &s, uint64_t seed, uint32_t target_tris) {
    SplitMix rng{seed};
    auto albedo = [&]() {
        return Vec3((float)rng.uniform(0.1, 0.8), (float)rng.uniform(0.1, 0.8), (float)rng.uniform(0.1, 0.8));
    };
    Vec3 cols[7];
    for (auto &c : cols) c = albedo();
    // Cell size chosen so that the total triangle count is closest to the
    // target (monotone in the cell size; bisection).
    auto build = [&](double cell, Scene &out) {
        out = Scene();
        auto n = [&](double len) { return std::max(1, (int)std::lround(len / cell)); };
        quad(out, {-1, 0, -1}, {0, 0, 2}, {2, 0, 0}, n(2), n(2), cols[0]);   // floor, normal +y
        quad(out, {-1, 2, -1}, {2, 0, 0}, {0, 0, 2}, n(2), n(2), cols[1]);   // ceiling, normal -y
        quad(out, {-1, 0, -1}, {2, 0, 0}, {0, 2, 0}, n(2), n(2), cols[2]);   // back, normal +z
        quad(out, {-1, 0, -1}, {0, 2, 0}, {0, 0, 2}, n(2), n(2), cols[3]);   // left, normal +x
        quad(out, {1, 0, -1}, {0, 0, 2}, {0, 2, 0}, n(2), n(2), cols[4]);    // right, normal -x
        const Vec3 le(15.f, 15.f, 15.f);
        quad(out, {-0.25, 1.98, -0.25}, {0.5, 0, 0}, {0, 0, 0.5}, n(0.5), n(0.5), le, true);  // light, normal -y
        box(out, {0.35, 0.3, 0.3}, {0.3, 0.3, 0.3}, -18.0, cell, cols[5]);
        box(out, {-0.35, 0.6, -0.35}, {0.3, 0.6, 0.3}, 18.0, cell, cols[6]);
    };
    double lo = 0.001, hi = 1.0;
    Scene tmp;
    for (int it = 0; it < 40; it++) {
        const double mid = std::sqrt(lo * hi);
        build(mid, tmp);
        if (tmp.triangles.size() > target_tris)
            lo = mid;
        else
            hi = mid;
    }
    Scene a, b;
    build(lo, a);
    build(hi, b);
    const long da = std::labs((long)a.triangles.size() - (long)target_tris);
    const long db = std::labs((long)b.triangles.size() - (long)target_tris);
    s = std::move(da <= db ? a : b);
    s.mode = PATHTRACING;
    s.maxDepth = 8;
    s.numSamples = 64;
    s.rrDepth = 3;
    s.width = s.height = 1024;
    s.field_of_view = 40;
    s.camera_position = Vec3(0.f, 1.f, 3.8f);
    s.camera_look = Vec3(0.f, 0.f, -1.f);
    s.backgroundColor = Vec3(0.f, 0.f, 0.f);
    s.detEps = 1e-9f;
    s.treeKind = 1;
    s.seed = seed;
}

void makeRandomScene(Scene &s, uint64_t seed, uint32_t n_tris) {
    s = Scene();
    SplitMix rng{seed};
    auto albedo = [&]() {
        return Vec3((float)rng.uniform(0.1, 0.8), (float)rng.uniform(0.1, 0.8), (float)rng.uniform(0.1, 0.8));
    };
    // Ground terrain patch takes about a third of the budget.
    const uint32_t budget_terrain = n_tris / 3;
    const uint32_t g = std::max(2u, (uint32_t)std::sqrt(budget_terrain / 2.0));
    {
        Scene t;
        makeTerrainScene(t, g, seed ^ 0xA5A5ull);
        const Vec3 col = albedo();
        for (auto &tr : t.triangles) {
            auto m = [](Vec3 p) { return Vec3(p.x * 8.f - 4.f, p.y * 4.f, p.z * 8.f - 4.f); };
            s.triangles.emplace_back(m(tr.a), m(tr.b), m(tr.c), col);
        }
    }
    // Objects until the budget is spent: icospheres (levels 1..6) and boxes.
    int guard = 0;
    while (s.triangles.size() < n_tris && guard++ < 10000) {
        const size_t left = n_tris - s.triangles.size();
        const double x = rng.uniform(-3.5, 3.5), z = rng.uniform(-3.5, 3.5);
        if (rng.uniform() < 0.6) {
            int level = 0;
            while (level < 6 && 20u * (1u << (2 * (level + 1))) <= left) level++;
            const double r = rng.uniform(0.15, 0.8);
            icosphere(s, {x, 0.4 + r, z}, r, level, albedo());
        } else {
            const double cell = std::max(0.02, std::sqrt(24.0 / std::max<size_t>(left, 12)));
            box(s, {x, 0.8, z}, {rng.uniform(0.1, 0.5), rng.uniform(0.1, 0.8), rng.uniform(0.1, 0.5)},
                rng.uniform(-45, 45), std::min(cell, 0.3), albedo());
        }
    }
    quad(s, {-1, 6, -1}, {2, 0, 0}, {0, 0, 2}, 1, 1, Vec3(12.f, 12.f, 12.f), true);
    s.mode = PATHTRACING;
    s.maxDepth = 6;
    s.numSamples = 16;
    s.rrDepth = 3;
    s.field_of_view = 60;
    s.camera_position = Vec3(0.f, 4.f, 9.f);
    s.camera_look = Vec3(0.f, -0.45f, -1.f);
    s.backgroundColor = Vec3(0.3f, 0.35f, 0.45f);
    s.detEps = 1e-9f;
    s.treeKind = 1;
    s.seed = seed;
}

}  // namespace pray
