// Float image output and the state of a progressive render (no reference
// counterpart: the reference writes only the clamped 8-bit BMP,
// main.cpp:93-98, and renders every sample in one call).
//
// PFM: "PF\n<w> <h>\n-1.0\n" then little-endian float RGB rows, bottom row
// first (the Portable Float Map convention; a negative scale means
// little-endian).
//
// Accumulation file: a 32-byte header, the running sum of per-sample
// radiance (w*h*3 floats, top row first) and an FNV-1a 64-bit checksum of
// everything before it.
//   char     magic[8]      "PRAYACC1"
//   uint32   width, height, samples, reserved
//   uint64   fingerprint   renderFingerprint of the scene it belongs to
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "scene.hpp"

namespace pray {

namespace {

constexpr char ACC_MAGIC[8] = {'P', 'R', 'A', 'Y', 'A', 'C', 'C', '1'};

struct AccHeader {
    char magic[8];
    uint32_t width, height, samples, reserved;
    uint64_t fingerprint;
};
static_assert(sizeof(AccHeader) == 32, "header layout");

struct Fnv {
    uint64_t h = 0xCBF29CE484222325ull;
    void add(const void *p, size_t n) {
        const uint8_t *b = static_cast<const uint8_t *>(p);
        for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 0x100000001B3ull;
    }
    template <class T> void val(const T &v) { add(&v, sizeof v); }
};

}  // namespace

bool writePFM(const std::string &path, uint32_t w, uint32_t h, const float *rgb) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = std::fprintf(f, "PF\n%u %u\n-1.0\n", w, h) > 0;
    for (int64_t y = (int64_t)h - 1; ok && y >= 0; y--)
        ok = std::fwrite(rgb + (size_t)y * w * 3, sizeof(float), (size_t)w * 3, f) == (size_t)w * 3;
    return (std::fclose(f) == 0) && ok;
}

bool readPFM(const std::string &path, uint32_t &w, uint32_t &h, std::vector<float> &rgb, std::string &err) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    char tag[3] = {0, 0, 0};
    unsigned ww = 0, hh = 0;
    double scale = 0;
    const bool head = std::fscanf(f, "%2s %u %u %lf", tag, &ww, &hh, &scale) == 4 && std::strcmp(tag, "PF") == 0 &&
                      std::fgetc(f) == '\n' && ww > 0 && hh > 0 && scale != 0 && (uint64_t)ww * hh <= (1ull << 28);
    if (!head) {
        std::fclose(f);
        err = path + " is not an RGB PFM file";
        return false;
    }
    w = ww, h = hh;
    rgb.assign((size_t)w * h * 3, 0.f);
    bool ok = true;
    for (int64_t y = (int64_t)h - 1; ok && y >= 0; y--)
        ok = std::fread(rgb.data() + (size_t)y * w * 3, sizeof(float), (size_t)w * 3, f) == (size_t)w * 3;
    std::fclose(f);
    if (!ok) {
        err = path + ": truncated";
        return false;
    }
    if (scale > 0) {  // big-endian payload
        for (float &v : rgb) {
            uint32_t u;
            std::memcpy(&u, &v, 4);
            u = __builtin_bswap32(u);
            std::memcpy(&v, &u, 4);
        }
    }
    return true;
}

// Everything that decides a path-traced pixel except the sample count:
// geometry and materials, camera, image size, depth limits, seed,
// background and determinant epsilon.
uint64_t renderFingerprint(const Scene &s) {
    Fnv f;
    f.val((uint64_t)s.triangles.size());
    for (const Triangle &t : s.triangles) {
        const float v[12] = {t.a.x, t.a.y, t.a.z, t.b.x, t.b.y, t.b.z, t.c.x, t.c.y, t.c.z,
                             t.color.x, t.color.y, t.color.z};
        f.add(v, sizeof v);
        f.val((uint8_t)t.isEmitting);
    }
    float cam[16];
    cameraParams(s, cam);
    f.add(cam, sizeof cam);
    f.val((uint64_t)s.width), f.val((uint64_t)s.height), f.val((uint64_t)s.maxDepth);
    // the estimator as preprocessing chooses it (from the mode and depth),
    // so that a scene fingerprints the same before and after preprocessing
    const PtMode pm = (s.mode == PATHTRACING && s.maxDepth > MAX_DEPTH_DIFFUSE) ? NORMAL : DIFFUSE;
    f.val((uint64_t)s.rrDepth), f.val(s.seed), f.val((uint32_t)pm);
    const float bg[4] = {s.backgroundColor.x, s.backgroundColor.y, s.backgroundColor.z, s.detEps};
    f.add(bg, sizeof bg);
    return f.h;
}

bool saveAccum(const std::string &path, uint32_t w, uint32_t h, uint32_t samples, uint64_t fingerprint,
               const float *sum, std::string &err) {
    AccHeader hd{};
    std::memcpy(hd.magic, ACC_MAGIC, 8);
    hd.width = w, hd.height = h, hd.samples = samples, hd.fingerprint = fingerprint;
    // Written next to the target and renamed over it, so an interrupted
    // save leaves the previous state intact.
    const std::string tmp = path + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) {
        err = "cannot write " + tmp;
        return false;
    }
    Fnv sum_h;
    sum_h.add(&hd, sizeof hd);
    sum_h.add(sum, (size_t)w * h * 12);
    const uint64_t digest = sum_h.h;
    bool ok = std::fwrite(&hd, sizeof hd, 1, f) == 1;
    ok = ok && std::fwrite(sum, 12, (size_t)w * h, f) == (size_t)w * h;
    ok = ok && std::fwrite(&digest, 8, 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    ok = ok && std::rename(tmp.c_str(), path.c_str()) == 0;
    if (!ok) {
        std::remove(tmp.c_str());
        err = "write error on " + path;
    }
    return ok;
}

bool loadAccum(const std::string &path, uint32_t &w, uint32_t &h, uint32_t &samples, uint64_t &fingerprint,
               std::vector<float> &sum, std::string &err) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    AccHeader hd{};
    if (std::fread(&hd, sizeof hd, 1, f) != 1 || std::memcmp(hd.magic, ACC_MAGIC, 8) != 0 ||
        (uint64_t)hd.width * hd.height > (1ull << 28)) {
        std::fclose(f);
        err = path + " is not an accumulation file";
        return false;
    }
    sum.assign((size_t)hd.width * hd.height * 3, 0.f);
    uint64_t digest = 0;
    bool ok = sum.empty() || std::fread(sum.data(), 12, (size_t)hd.width * hd.height, f) == (size_t)hd.width * hd.height;
    ok = ok && std::fread(&digest, 8, 1, f) == 1;
    const bool extra = ok && std::fgetc(f) != EOF;
    std::fclose(f);
    if (!ok || extra) {
        err = path + (ok ? ": trailing bytes" : ": truncated");
        return false;
    }
    Fnv sum_h;
    sum_h.add(&hd, sizeof hd);
    sum_h.add(sum.data(), sum.size() * 4);
    if (sum_h.h != digest) {
        err = path + ": checksum mismatch";
        return false;
    }
    w = hd.width, h = hd.height, samples = hd.samples, fingerprint = hd.fingerprint;
    return true;
}

}  // namespace pray
