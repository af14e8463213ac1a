// Scene I/O of the drop-in host (reference src/io.cpp:27-137, main.cpp:93-98).
// The reference links nlohmann json 2.0.7, tinyobjloader 1.0.x and
// stb_image_write 1.02; this file implements the parts it uses.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "scene.hpp"

namespace pray {

// ------------------------------------------------------------------ JSON --
namespace {

struct JVal {
    enum Kind { NUL, BOOL, INT, FLOAT, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    int64_t i = 0;
    double f = 0;
    std::string s;
    std::vector<JVal> a;
    std::map<std::string, JVal> o;

    bool isNum() const { return kind == INT || kind == FLOAT; }
    // JSON numbers convert like nlohmann's get<float>(): integers by
    // static_cast from int64, decimals through strtod.
    float asFloat() const { return kind == INT ? static_cast<float>(i) : static_cast<float>(f); }
    const JVal *get(const std::string &k) const {
        auto it = o.find(k);
        return it == o.end() ? nullptr : &it->second;
    }
};

struct JParser {
    const char *p, *e;
    std::string err;

    void ws() {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
    }
    bool lit(const char *w) {
        size_t n = std::strlen(w);
        if ((size_t)(e - p) >= n && std::strncmp(p, w, n) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    bool value(JVal &v) {
        ws();
        if (p >= e) return fail("unexpected end");
        if (*p == '{') {
            v.kind = JVal::OBJ;
            p++;
            ws();
            if (p < e && *p == '}') return p++, true;
            while (true) {
                JVal key;
                ws();
                if (p >= e || *p != '"' || !str(key.s)) return fail("expected key");
                ws();
                if (p >= e || *p != ':') return fail("expected ':'");
                p++;
                JVal val;
                if (!value(val)) return false;
                v.o[key.s] = std::move(val);
                ws();
                if (p < e && *p == ',') {
                    p++;
                    continue;
                }
                if (p < e && *p == '}') return p++, true;
                return fail("expected ',' or '}'");
            }
        }
        if (*p == '[') {
            v.kind = JVal::ARR;
            p++;
            ws();
            if (p < e && *p == ']') return p++, true;
            while (true) {
                JVal val;
                if (!value(val)) return false;
                v.a.push_back(std::move(val));
                ws();
                if (p < e && *p == ',') {
                    p++;
                    continue;
                }
                if (p < e && *p == ']') return p++, true;
                return fail("expected ',' or ']'");
            }
        }
        if (*p == '"') {
            v.kind = JVal::STR;
            return str(v.s);
        }
        if (lit("true")) return v.kind = JVal::BOOL, v.b = true, true;
        if (lit("false")) return v.kind = JVal::BOOL, v.b = false, true;
        if (lit("null")) return v.kind = JVal::NUL, true;
        return number(v);
    }
    bool str(std::string &out) {
        p++;  // opening quote
        while (p < e && *p != '"') {
            if (*p == '\\' && p + 1 < e) {
                p++;
                const char c = *p;
                out.push_back(c == 'n' ? '\n' : c == 't' ? '\t' : c == 'r' ? '\r' : c);
                p++;
            } else {
                out.push_back(*p++);
            }
        }
        if (p >= e) return fail("unterminated string");
        p++;
        return true;
    }
    bool number(JVal &v) {
        const char *s = p;
        bool isFloat = false;
        if (p < e && *p == '-') p++;
        while (p < e && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-')) {
            if (*p == '.' || *p == 'e' || *p == 'E') isFloat = true;
            p++;
        }
        if (p == s) return fail("bad value");
        std::string tok(s, p);
        if (!isFloat) {
            errno = 0;
            char *end = nullptr;
            long long iv = std::strtoll(tok.c_str(), &end, 10);
            if (errno == 0 && end && *end == '\0') {
                v.kind = JVal::INT;
                v.i = iv;
                return true;
            }
        }
        v.kind = JVal::FLOAT;
        v.f = std::strtod(tok.c_str(), nullptr);
        return true;
    }
    bool fail(const char *m) {
        if (err.empty()) err = m;
        return false;
    }
};

// ------------------------------------------------------------------- OBJ --

// A decimal number of an OBJ token [s, e), computed in the order that
// decides the reference's vertex bits (its loader, tinyobjloader 1.0.x,
// ext/tiny_obj_loader.h:459-572): integer digits as m = m*10 + d; the k-th
// fraction digit adds d * 10^-k, with 10^-k taken from the table below for
// k < 8 and from pow(10, -k) beyond; an exponent x gives
// ldexp(m * 5^x, x), i.e. m * 10^x with the power of two applied exactly.
// Accepted: [+-]digits[.digits][(e|E)[+-]digits]; a token that does not start
// that way, or whose exponent has no digit, is rejected (false, *result
// untouched); anything after the number is ignored.
constexpr double kNegPow10[8] = {1.0, 1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7};

bool objParseDouble(const char *s, const char *e, double *result) {
    const char *p = s;
    const auto at_digit = [&p, e] { return p < e && *p >= '0' && *p <= '9'; };
    double sign = 1.0;
    if (p < e && (*p == '+' || *p == '-')) sign = *p++ == '-' ? -1.0 : 1.0;
    if (!at_digit()) return false;
    double m = 0.0;
    for (; at_digit(); ++p) {
        m *= 10;
        m += *p - '0';
    }
    if (p < e && *p == '.') {
        ++p;
        for (int k = 1; at_digit(); ++p, ++k) m += (*p - '0') * (k < 8 ? kNegPow10[k] : std::pow(10.0, -k));
    }
    int x = 0;
    if (p < e && (*p == 'e' || *p == 'E')) {
        ++p;
        bool neg = false;
        if (p < e && (*p == '+' || *p == '-')) neg = *p++ == '-';
        if (!at_digit()) return false;
        for (; at_digit(); ++p) x = std::min(10 * x + (*p - '0'), 99999);  // beyond +-99999: 0 or inf either way
        if (neg) x = -x;
    }
    *result = sign * (x ? std::ldexp(m * std::pow(5.0, x), x) : m);
    return true;
}

float objFloat(const char *&tok, double def = 0.0) {
    tok += std::strspn(tok, " \t");
    const char *end = tok + std::strcspn(tok, " \t\r");
    double val = def;
    objParseDouble(tok, end, &val);
    tok = end;
    return static_cast<float>(val);
}

struct Material {
    float diffuse[3] = {0, 0, 0}, emission[3] = {0, 0, 0};
};

bool loadMtl(const std::string &path, std::vector<Material> &mats, std::map<std::string, int> &ids, std::string &err) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot open material file " + path;
        return false;
    }
    std::string line;
    bool have = false;
    Material cur;
    std::string curName;
    auto flush = [&]() {
        if (have) {
            ids[curName] = (int)mats.size();
            mats.push_back(cur);
        }
    };
    while (std::getline(in, line)) {
        const char *t = line.c_str();
        t += std::strspn(t, " \t");
        if (std::strncmp(t, "newmtl", 6) == 0 && (t[6] == ' ' || t[6] == '\t')) {
            flush();
            cur = Material();
            char name[4096] = {0};
            std::sscanf(t + 7, "%4095s", name);
            curName = name;
            have = true;
        } else if (t[0] == 'K' && t[1] == 'd' && (t[2] == ' ' || t[2] == '\t')) {
            t += 2;
            for (float &c : cur.diffuse) c = objFloat(t);
        } else if (t[0] == 'K' && t[1] == 'e' && (t[2] == ' ' || t[2] == '\t')) {
            t += 2;
            for (float &c : cur.emission) c = objFloat(t);
        }
    }
    flush();
    return true;
}

int objIndex(int idx, int n) {
    if (idx > 0) return idx - 1;
    if (idx == 0) return 0;
    return n + idx;
}

}  // namespace

bool readSceneDescription(const std::string &path, Scene &scene, std::string &err) {
    std::ifstream fin(path);
    if (!fin) {
        err = "Error opening file: " + path;
        return false;
    }
    std::stringstream ss;
    ss << fin.rdbuf();
    const std::string text = ss.str();
    JVal root;
    JParser jp{text.data(), text.data() + text.size(), {}};
    if (!jp.value(root) || root.kind != JVal::OBJ) {
        err = "bad scene JSON " + path + ": " + jp.err;
        return false;
    }
    std::string base = path;
    const size_t slash = base.find_last_of('/');
    base = (slash == std::string::npos) ? std::string() : base.substr(0, slash + 1);

    auto num = [&](const JVal *v, const char *what, float &out) -> bool {
        if (!v || !v->isNum()) {
            err = std::string("missing number: ") + what;
            return false;
        }
        out = v->asFloat();
        return true;
    };
    auto vec = [&](const char *key, Vec3 &out) -> bool {
        const JVal *v = root.get(key);
        if (!v || v->kind != JVal::ARR || v->a.size() < 3) {
            err = std::string("missing vector: ") + key;
            return false;
        }
        return num(&v->a[0], key, out.x) && num(&v->a[1], key, out.y) && num(&v->a[2], key, out.z);
    };
    auto uint = [&](const char *key, uint64_t &out) -> bool {
        const JVal *v = root.get(key);
        if (!v || !v->isNum()) {
            err = std::string("missing number: ") + key;
            return false;
        }
        out = v->kind == JVal::INT ? (uint64_t)v->i : (uint64_t)v->f;
        return true;
    };

    const JVal *m = root.get("method");
    if (m && m->kind == JVal::STR && m->s == "path_tracing") {
        scene.mode = PATHTRACING;
        uint64_t md = 0, ns = 0;
        if (!uint("max_depth", md) || !uint("num_samples", ns)) return false;
        scene.maxDepth = md;
        scene.numSamples = ns;
    } else {
        scene.mode = RAYTRACING;
        scene.maxDepth = 0;
        scene.numSamples = 0;
    }
    if (!uint("resolution_x", scene.width) || !uint("resolution_y", scene.height) || !uint("fov", scene.field_of_view))
        return false;
    if (!vec("camera_position", scene.camera_position) || !vec("camera_look", scene.camera_look)) return false;
    // Lights are read in ray-tracing mode only (io.cpp:91 evaluates
    // `!scene.mode == Mode::PATHTRACING`, true exactly for RAYTRACING).
    if (scene.mode == RAYTRACING) {
        const JVal *ls = root.get("lights");
        if (ls && ls->kind == JVal::ARR) {
            for (const JVal &l : ls->a) {
                const JVal *pp = l.get("position"), *cc = l.get("color");
                if (!pp || !cc || pp->a.size() < 3 || cc->a.size() < 3) {
                    err = "bad light";
                    return false;
                }
                Vec3 p, c;
                if (!num(&pp->a[0], "light", p.x) || !num(&pp->a[1], "light", p.y) || !num(&pp->a[2], "light", p.z) ||
                    !num(&cc->a[0], "light", c.x) || !num(&cc->a[1], "light", c.y) || !num(&cc->a[2], "light", c.z))
                    return false;
                scene.lights.emplace_back(p, c);
            }
        }
    }
    if (!vec("background", scene.backgroundColor)) return false;
    // Optional extensions of this build.
    if (const JVal *v = root.get("det_eps")) scene.detEps = v->asFloat();
    if (const JVal *v = root.get("rr_depth")) scene.rrDepth = v->kind == JVal::INT ? (size_t)v->i : (size_t)v->f;
    if (const JVal *v = root.get("seed")) scene.seed = v->kind == JVal::INT ? (uint64_t)v->i : (uint64_t)v->f;
    if (const JVal *v = root.get("tree")) {
        if (v->kind == JVal::STR) scene.treeKind = v->s == "sah" ? 1 : (v->s == "median" ? 0 : -1);
    }

    const JVal *of = root.get("obj_file");
    if (!of || of->kind != JVal::STR) {
        err = "missing obj_file";
        return false;
    }
    const std::string objPath = base + of->s;
    std::ifstream obj(objPath);
    if (!obj) {
        err = "Error loading file: " + of->s;
        return false;
    }
    std::vector<float> v;
    std::vector<Material> mats;
    std::map<std::string, int> ids;
    int material = -1;
    std::string line;
    while (std::getline(obj, line)) {
        const char *t = line.c_str();
        t += std::strspn(t, " \t");
        if (t[0] == '\0' || t[0] == '#') continue;
        if (t[0] == 'v' && (t[1] == ' ' || t[1] == '\t')) {
            t += 2;
            for (int k = 0; k < 3; k++) v.push_back(objFloat(t));
        } else if (t[0] == 'f' && (t[1] == ' ' || t[1] == '\t')) {
            t += 2;
            std::vector<int> face;
            const int nv = (int)(v.size() / 3);
            while (true) {
                t += std::strspn(t, " \t\r");
                if (*t == '\0') break;
                face.push_back(objIndex(std::atoi(t), nv));
                t += std::strcspn(t, " \t\r");
            }
            if (face.size() < 3) continue;
            if (material < 0 || material >= (int)mats.size()) {
                err = "face without a known material in " + of->s;
                return false;
            }
            const Material &mt = mats[material];
            const Vec3 em(mt.emission[0], mt.emission[1], mt.emission[2]);
            const Vec3 df(mt.diffuse[0], mt.diffuse[1], mt.diffuse[2]);
            auto V = [&](int i) { return Vec3(v[3 * i], v[3 * i + 1], v[3 * i + 2]); };
            for (size_t k = 2; k < face.size(); k++) {
                const int i0 = face[0], i1 = face[k - 1], i2 = face[k];
                if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= nv || i1 >= nv || i2 >= nv) {
                    err = "face index out of range in " + of->s;
                    return false;
                }
                if (em == Vec3(0.f, 0.f, 0.f))
                    scene.triangles.emplace_back(V(i0), V(i1), V(i2), df, false);
                else
                    scene.triangles.emplace_back(V(i0), V(i1), V(i2), em, true);
            }
        } else if (std::strncmp(t, "usemtl", 6) == 0 && (t[6] == ' ' || t[6] == '\t')) {
            char name[4096] = {0};
            std::sscanf(t + 7, "%4095s", name);
            auto it = ids.find(name);
            material = it == ids.end() ? -1 : it->second;
        } else if (std::strncmp(t, "mtllib", 6) == 0 && (t[6] == ' ' || t[6] == '\t')) {
            char name[4096] = {0};
            std::sscanf(t + 7, "%4095s", name);
            if (!loadMtl(base + name, mats, ids, err)) return false;
        }
    }
    return true;
}

bool writeSceneFiles(const Scene &scene, const std::string &json_path, std::string &err) {
    std::string stem = json_path;
    if (stem.size() > 5 && stem.substr(stem.size() - 5) == ".json") stem = stem.substr(0, stem.size() - 5);
    const size_t slash = stem.find_last_of('/');
    const std::string name = slash == std::string::npos ? stem : stem.substr(slash + 1);
    FILE *mtl = std::fopen((stem + ".mtl").c_str(), "w");
    FILE *obj = std::fopen((stem + ".obj").c_str(), "w");
    FILE *js = std::fopen(json_path.c_str(), "w");
    if (!mtl || !obj || !js) {
        if (mtl) std::fclose(mtl);
        if (obj) std::fclose(obj);
        if (js) std::fclose(js);
        err = "cannot write scene files for " + json_path;
        return false;
    }
    // One material per distinct (colour, emissive) pair in first-use order.
    std::map<std::tuple<float, float, float, bool>, int> matIds;
    std::fprintf(obj, "mtllib %s.mtl\n", name.c_str());
    int cur = -1;
    for (size_t i = 0; i < scene.triangles.size(); i++) {
        const Triangle &t = scene.triangles[i];
        std::fprintf(obj, "v %.9g %.9g %.9g\nv %.9g %.9g %.9g\nv %.9g %.9g %.9g\n", t.a.x, t.a.y, t.a.z, t.b.x, t.b.y,
                     t.b.z, t.c.x, t.c.y, t.c.z);
    }
    for (size_t i = 0; i < scene.triangles.size(); i++) {
        const Triangle &t = scene.triangles[i];
        auto key = std::make_tuple(t.color.x, t.color.y, t.color.z, t.isEmitting);
        auto it = matIds.find(key);
        int id;
        if (it == matIds.end()) {
            id = (int)matIds.size();
            matIds[key] = id;
            std::fprintf(mtl, "newmtl m%d\n", id);
            if (t.isEmitting)
                std::fprintf(mtl, "Kd 0 0 0\nKe %.9g %.9g %.9g\n\n", t.color.x, t.color.y, t.color.z);
            else
                std::fprintf(mtl, "Kd %.9g %.9g %.9g\nKe 0 0 0\n\n", t.color.x, t.color.y, t.color.z);
        } else {
            id = it->second;
        }
        if (id != cur) {
            std::fprintf(obj, "usemtl m%d\n", id);
            cur = id;
        }
        std::fprintf(obj, "f %zu %zu %zu\n", 3 * i + 1, 3 * i + 2, 3 * i + 3);
    }
    std::fprintf(js, "{\n\t\"obj_file\": \"%s.obj\",\n", name.c_str());
    if (scene.mode == PATHTRACING)
        std::fprintf(js, "\t\"method\": \"path_tracing\",\n\t\"max_depth\": %zu,\n\t\"num_samples\": %zu,\n",
                     scene.maxDepth, scene.numSamples);
    std::fprintf(js, "\t\"resolution_x\": %llu,\n\t\"resolution_y\": %llu,\n", (unsigned long long)scene.width,
                 (unsigned long long)scene.height);
    std::fprintf(js, "\t\"camera_position\": [%.9g, %.9g, %.9g],\n", scene.camera_position.x, scene.camera_position.y,
                 scene.camera_position.z);
    std::fprintf(js, "\t\"camera_look\": [%.9g, %.9g, %.9g],\n", scene.camera_look.x, scene.camera_look.y,
                 scene.camera_look.z);
    std::fprintf(js, "\t\"fov\": %llu,\n\t\"lights\": [", (unsigned long long)scene.field_of_view);
    for (size_t i = 0; i < scene.lights.size(); i++) {
        const LightSource &l = scene.lights[i];
        std::fprintf(js, "%s\n\t\t{\"position\": [%.9g, %.9g, %.9g], \"color\": [%.9g, %.9g, %.9g]}", i ? "," : "",
                     l.position.x, l.position.y, l.position.z, l.color.x, l.color.y, l.color.z);
    }
    std::fprintf(js, "\n\t],\n\t\"background\": [%.9g, %.9g, %.9g]", scene.backgroundColor.x, scene.backgroundColor.y,
                 scene.backgroundColor.z);
    std::fprintf(js, ",\n\t\"det_eps\": %.9g", scene.detEps);
    if (scene.rrDepth != SIZE_MAX) std::fprintf(js, ",\n\t\"rr_depth\": %zu", scene.rrDepth);
    std::fprintf(js, ",\n\t\"seed\": %llu", (unsigned long long)scene.seed);
    if (scene.treeKind >= 0) std::fprintf(js, ",\n\t\"tree\": \"%s\"", scene.treeKind == 1 ? "sah" : "median");
    std::fprintf(js, "\n}\n");
    std::fclose(mtl);
    std::fclose(obj);
    std::fclose(js);
    return true;
}

// 24-bit BMP in the byte layout stb_image_write 1.02 produces
// (stbi_write_bmp): 14+40 byte header, rows bottom-up, BGR, rows padded to 4.
bool writeBMP(const std::string &path, uint32_t w, uint32_t h, const uint8_t *rgb) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const uint32_t pad = (uint32_t)(-(int64_t)w * 3) & 3u;
    const uint32_t fsize = 14 + 40 + (w * 3 + pad) * h;
    auto u16 = [&](uint32_t x) { std::fputc(x & 0xFF, f), std::fputc((x >> 8) & 0xFF, f); };
    auto u32 = [&](uint32_t x) { u16(x & 0xFFFF), u16(x >> 16); };
    std::fputc('B', f), std::fputc('M', f);
    u32(fsize), u16(0), u16(0), u32(14 + 40);
    u32(40), u32(w), u32(h), u16(1), u16(24), u32(0), u32(0), u32(0), u32(0), u32(0), u32(0);
    std::vector<uint8_t> row(w * 3 + pad, 0);
    for (int64_t y = (int64_t)h - 1; y >= 0; y--) {
        const uint8_t *src = rgb + (size_t)y * w * 3;
        for (uint32_t x = 0; x < w; x++) {
            row[3 * x] = src[3 * x + 2];
            row[3 * x + 1] = src[3 * x + 1];
            row[3 * x + 2] = src[3 * x];
        }
        std::fwrite(row.data(), 1, row.size(), f);
    }
    return std::fclose(f) == 0;
}

}  // namespace pray
