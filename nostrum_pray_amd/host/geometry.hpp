// Host-side geometry for the Nostrum Pray drop-in host.
//
// Semantics follow the reference's include/geometry.h: Vec3 dot/cross/normalize
// (geometry.h:28-38), Ray (geometry.h:76-88), BoundingBox expand/clip/split/
// surfaceArea (geometry.h:112-216), Triangle precomputed fields
// (geometry.h:219-239).  The floating-point operation ORDER of dot, cross and
// normalize is kept identical because triangle edges (ab, ac), normals and ray
// directions computed here are consumed bit-for-bit by the GPU kernels and the
// CPU oracle.  Everything in this library is compiled with -ffp-contract=off.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>

namespace pray {

// Absolute determinant cull of the reference (geometry.h:12).
constexpr float EPSILON = 1.e-4f;
constexpr float INFTY = std::numeric_limits<float>::max();

inline float pi_f() { return static_cast<float>(std::acos(-1.0)); }

inline float clampf(float lo, float hi, float v) { return std::max(lo, std::min(hi, v)); }

struct Vec3 {
    float x, y, z;

    Vec3() = default;
    constexpr Vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}

    // (x*x' + y*y') + z*z'  -- left to right, as geometry.h:36
    float dot(const Vec3 &v) const { return x * v.x + y * v.y + z * v.z; }
    // geometry.h:38
    Vec3 cross(const Vec3 &v) const { return Vec3(y * v.z - z * v.y, z * v.x - x * v.z, x * v.y - y * v.x); }
    float norm() const { return x * x + y * y + z * z; }
    // geometry.h:28-34: one IEEE sqrt, one IEEE division, three multiplies.
    void normalize() {
        float n2 = norm();
        if (n2 > 0) {
            float inv = 1 / std::sqrt(n2);
            x *= inv, y *= inv, z *= inv;
        }
    }
    Vec3 inverse() const { return Vec3(1 / x, 1 / y, 1 / z); }

    Vec3 operator+(const Vec3 &v) const { return Vec3(x + v.x, y + v.y, z + v.z); }
    Vec3 operator+(float r) const { return Vec3(x + r, y + r, z + r); }
    Vec3 operator-(const Vec3 &v) const { return Vec3(x - v.x, y - v.y, z - v.z); }
    Vec3 operator-() const { return Vec3(-x, -y, -z); }
    Vec3 operator*(float r) const { return Vec3(x * r, y * r, z * r); }
    Vec3 operator*(const Vec3 &v) const { return Vec3(x * v.x, y * v.y, z * v.z); }
    Vec3 operator/(float r) const { return Vec3(x / r, y / r, z / r); }
    Vec3 &operator+=(const Vec3 &v) { x += v.x, y += v.y, z += v.z; return *this; }
    bool operator==(const Vec3 &v) const { return x == v.x && y == v.y && z == v.z; }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float &at(int i) { return i == 0 ? x : (i == 1 ? y : z); }
    Vec3 clamp(float lo, float hi) const { return Vec3(clampf(lo, hi, x), clampf(lo, hi, y), clampf(lo, hi, z)); }
};

// A ray; the constructor normalises the direction and stores its reciprocal
// (geometry.h:84-87).
struct Ray {
    Vec3 position, direction, inverseDirection;
    Ray() = default;
    Ray(Vec3 p, Vec3 d) : position(p), direction(d) {
        direction.normalize();
        inverseDirection = direction.inverse();
    }
};

// Split plane of a KD node (geometry.h:91-100).
struct SplitPlane {
    uint8_t axis = 255;
    float distance = 0.f;
    SplitPlane() = default;
    SplitPlane(uint8_t a, float d) : axis(a), distance(d) {}
    bool operator==(const SplitPlane &o) const { return axis == o.axis && distance == o.distance; }
};

struct BoundingBox {
    Vec3 min, max;
    BoundingBox() = default;
    BoundingBox(Vec3 lo, Vec3 hi) : min(lo), max(hi) {}

    // Union (geometry.h:112-119); std::min/std::max argument order kept.
    BoundingBox expand(const BoundingBox &b) const {
        return BoundingBox(Vec3(std::min(min.x, b.min.x), std::min(min.y, b.min.y), std::min(min.z, b.min.z)),
                           Vec3(std::max(max.x, b.max.x), std::max(max.y, b.max.y), std::max(max.z, b.max.z)));
    }
    // Intersection (geometry.h:121-129); no containment assert here.
    BoundingBox clip(const BoundingBox &b) const {
        return BoundingBox(Vec3(std::max(min.x, b.min.x), std::max(min.y, b.min.y), std::max(min.z, b.min.z)),
                           Vec3(std::min(max.x, b.max.x), std::min(max.y, b.max.y), std::min(max.z, b.max.z)));
    }
    bool isPlanar(int k) const { return min[k] - max[k] == 0; }
    // Longest axis, ties resolved x before y before z (geometry.h:135-149).
    uint8_t getLongestAxis() const {
        float dx = max.x - min.x, dy = max.y - min.y, dz = max.z - min.z;
        if (dx >= dy && dx >= dz) return 0;
        if (dy >= dx && dy >= dz) return 1;
        return 2;
    }
    // 2*((dx*dy + dx*dz) + dy*dz)  (geometry.h:185-191)
    float surfaceArea() const {
        float dx = max.x - min.x, dy = max.y - min.y, dz = max.z - min.z;
        return 2.f * (dx * dy + dx * dz + dy * dz);
    }
    // Halves at a split plane; axis values other than 0 and 1 split z
    // (geometry.h:194-211).
    void split(const SplitPlane &p, BoundingBox &bl, BoundingBox &br) const {
        bl = *this;
        br = *this;
        if (p.axis == 0) {
            bl.max.x = p.distance;
            br.min.x = p.distance;
        } else if (p.axis == 1) {
            bl.max.y = p.distance;
            br.min.y = p.distance;
        } else {
            bl.max.z = p.distance;
            br.min.z = p.distance;
        }
    }
};

// Triangle with the per-triangle precompute of initialiseTriangles
// (scene_preprocessor.cpp:74-96).
struct Triangle {
    Vec3 a, b, c;
    Vec3 color;
    bool isEmitting = false;
    Vec3 normal, ab, ac, midPoint;
    BoundingBox boundingBox;

    Triangle() = default;
    Triangle(Vec3 a_, Vec3 b_, Vec3 c_, Vec3 col, bool emit = false) : a(a_), b(b_), c(c_), color(col), isEmitting(emit) {}

    void precompute() {
        boundingBox.min = Vec3(std::min(a.x, std::min(b.x, c.x)), std::min(a.y, std::min(b.y, c.y)),
                               std::min(a.z, std::min(b.z, c.z)));
        boundingBox.max = Vec3(std::max(a.x, std::max(b.x, c.x)), std::max(a.y, std::max(b.y, c.y)),
                               std::max(a.z, std::max(b.z, c.z)));
        midPoint = (a + b + c) / 3;
        ab = b - a;
        ac = c - a;
        normal = ab.cross(ac);
        normal.normalize();
    }
};

// Row-major 4x4 matrix with the reference's Gauss-Jordan inverse and
// row-vector direction transform (geometry.h:278-402).
struct Matrix44 {
    float x[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    Matrix44() = default;
    Matrix44(float a, float b, float c, float d, float e, float f, float g, float h, float i, float j, float k,
             float l, float m, float n, float o, float p) {
        const float v[16] = {a, b, c, d, e, f, g, h, i, j, k, l, m, n, o, p};
        for (int r = 0; r < 4; ++r)
            for (int q = 0; q < 4; ++q) x[r][q] = v[4 * r + q];
    }
    Vec3 multDirMatrix(const Vec3 &s) const {
        float a = s.x * x[0][0] + s.y * x[1][0] + s.z * x[2][0];
        float b = s.x * x[0][1] + s.y * x[1][1] + s.z * x[2][1];
        float c = s.x * x[0][2] + s.y * x[1][2] + s.z * x[2][2];
        return Vec3(a, b, c);
    }
    // Gauss-Jordan with partial pivoting over the first three columns, then
    // back substitution; singular input yields the identity.
    Matrix44 inverse() const {
        Matrix44 s;
        Matrix44 t(*this);
        for (int i = 0; i < 3; i++) {
            int pivot = i;
            float pivotsize = t.x[i][i];
            if (pivotsize < 0) pivotsize = -pivotsize;
            for (int j = i + 1; j < 4; j++) {
                float tmp = t.x[j][i];
                if (tmp < 0) tmp = -tmp;
                if (tmp > pivotsize) {
                    pivot = j;
                    pivotsize = tmp;
                }
            }
            if (pivotsize == 0) return Matrix44();
            if (pivot != i) {
                for (int j = 0; j < 4; j++) {
                    std::swap(t.x[i][j], t.x[pivot][j]);
                    std::swap(s.x[i][j], s.x[pivot][j]);
                }
            }
            for (int j = i + 1; j < 4; j++) {
                float f = t.x[j][i] / t.x[i][i];
                for (int k = 0; k < 4; k++) {
                    t.x[j][k] -= f * t.x[i][k];
                    s.x[j][k] -= f * s.x[i][k];
                }
            }
        }
        for (int i = 3; i >= 0; --i) {
            float f = t.x[i][i];
            if (f == 0) return Matrix44();
            for (int j = 0; j < 4; j++) {
                t.x[i][j] /= f;
                s.x[i][j] /= f;
            }
            for (int j = 0; j < i; j++) {
                f = t.x[j][i];
                for (int k = 0; k < 4; k++) {
                    t.x[j][k] -= f * t.x[i][k];
                    s.x[j][k] -= f * s.x[i][k];
                }
            }
        }
        return s;
    }
};

}  // namespace pray
