// On-disk form of a flattened tree (the device layout of gpu_layout.hpp), so
// that large scenes are built once and later uploaded without a rebuild.
//
// File: a 96-byte header, the arrays in upload order, and an FNV-1a 64-bit
// checksum of everything before it.
//   char     magic[8]      "PRAYFLT1"
//   uint32   version       1
//   uint32   kind          0 median, 1 SAH
//   uint32   n_nodes, n_refs, n_tris, max_depth, n_leaves, n_empty_leaves
//   uint32   flags         bit 0: shading arrays present
//   uint32   reserved
//   float    root_box[6]
//   uint8    pad[24]
// then nodes (8 B each), leaf_refs (4 B), a, ab, ac (12 B per triangle each),
// and with shading data normal, color (12 B) and emit (1 B) per triangle.
#include <sys/types.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gpu_layout.hpp"

namespace pray {

namespace {

constexpr char MAGIC[8] = {'P', 'R', 'A', 'Y', 'F', 'L', 'T', '1'};

struct FlatHeader {
    char magic[8];
    uint32_t version, kind;
    uint32_t n_nodes, n_refs, n_tris, max_depth, n_leaves, n_empty;
    uint32_t flags, reserved;
    float root_box[6];
    uint8_t pad[24];
};
static_assert(sizeof(FlatHeader) == 96, "header layout");

struct Fnv {
    uint64_t h = 0xCBF29CE484222325ull;
    void add(const void *p, size_t n) {
        const uint8_t *b = static_cast<const uint8_t *>(p);
        for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 0x100000001B3ull;
    }
};

}  // namespace

bool saveFlat(const KDTreeGPU &g, int kind, const std::string &path, std::string &err) {
    FlatHeader h{};
    std::memcpy(h.magic, MAGIC, 8);
    h.version = 1;
    h.kind = (uint32_t)kind;
    h.n_nodes = g.numNodes();
    h.n_refs = (uint32_t)g.leafRefs.size();
    h.n_tris = g.numTris();
    h.max_depth = g.maxDepth;
    h.n_leaves = g.numLeaves;
    h.n_empty = g.numEmptyLeaves;
    const bool shading = g.triNormal.size() == 3 * (size_t)h.n_tris && g.triColor.size() == 3 * (size_t)h.n_tris;
    h.flags = shading ? 1u : 0u;
    std::memcpy(h.root_box, g.rootBox, sizeof h.root_box);
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        err = "cannot write " + path;
        return false;
    }
    Fnv sum;
    bool ok = true;
    auto put = [&](const void *p, size_t n) {
        if (!ok || n == 0) return;
        sum.add(p, n);
        ok = std::fwrite(p, 1, n, f) == n;
    };
    put(&h, sizeof h);
    put(g.nodes.data(), g.nodes.size() * 4);
    put(g.leafRefs.data(), g.leafRefs.size() * 4);
    put(g.triA.data(), g.triA.size() * 4);
    put(g.triAB.data(), g.triAB.size() * 4);
    put(g.triAC.data(), g.triAC.size() * 4);
    if (shading) {
        put(g.triNormal.data(), g.triNormal.size() * 4);
        put(g.triColor.data(), g.triColor.size() * 4);
        put(g.triEmit.data(), g.triEmit.size());
    }
    const uint64_t digest = sum.h;
    ok = ok && std::fwrite(&digest, 1, 8, f) == 8;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) err = "write error on " + path;
    return ok;
}

bool loadFlat(const std::string &path, KDTreeGPU &g, int &kind, std::string &err) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    Fnv sum;
    bool ok = true;
    auto get = [&](void *p, size_t n) {
        if (!ok || n == 0) return;
        ok = std::fread(p, 1, n, f) == n;
        if (ok) sum.add(p, n);
    };
    FlatHeader h{};
    get(&h, sizeof h);
    if (!ok || std::memcmp(h.magic, MAGIC, 8) != 0 || h.version != 1) {
        std::fclose(f);
        err = path + " is not a flat tree file";
        return false;
    }
    // Sizes must match the file length before anything is allocated.
    const bool shading = (h.flags & 1u) != 0;
    const uint64_t body = (uint64_t)h.n_nodes * 8 + (uint64_t)h.n_refs * 4 + (uint64_t)h.n_tris * 36 +
                          (shading ? (uint64_t)h.n_tris * 25 : 0);
    fseeko(f, 0, SEEK_END);
    const long long len = (long long)ftello(f);
    fseeko(f, (off_t)sizeof h, SEEK_SET);
    if (len < 0 || (uint64_t)len != sizeof h + body + 8) {
        std::fclose(f);
        err = path + ": length does not match its header";
        return false;
    }
    g = KDTreeGPU();
    g.nodes.resize(2 * (size_t)h.n_nodes);
    g.leafRefs.resize(h.n_refs);
    g.triA.resize(3 * (size_t)h.n_tris);
    g.triAB.resize(3 * (size_t)h.n_tris);
    g.triAC.resize(3 * (size_t)h.n_tris);
    get(g.nodes.data(), g.nodes.size() * 4);
    get(g.leafRefs.data(), g.leafRefs.size() * 4);
    get(g.triA.data(), g.triA.size() * 4);
    get(g.triAB.data(), g.triAB.size() * 4);
    get(g.triAC.data(), g.triAC.size() * 4);
    if (shading) {
        g.triNormal.resize(3 * (size_t)h.n_tris);
        g.triColor.resize(3 * (size_t)h.n_tris);
        g.triEmit.resize(h.n_tris);
        get(g.triNormal.data(), g.triNormal.size() * 4);
        get(g.triColor.data(), g.triColor.size() * 4);
        get(g.triEmit.data(), g.triEmit.size());
    } else {
        g.triEmit.resize(h.n_tris);  // numTris() counts emit flags
    }
    uint64_t digest = 0;
    const uint64_t want = sum.h;
    ok = ok && std::fread(&digest, 1, 8, f) == 8;
    std::fclose(f);
    if (!ok) {
        err = path + ": truncated";
        return false;
    }
    if (digest != want) {
        err = path + ": checksum mismatch";
        return false;
    }
    std::memcpy(g.rootBox, h.root_box, sizeof g.rootBox);
    g.maxDepth = h.max_depth;
    g.numLeaves = h.n_leaves;
    g.numEmptyLeaves = h.n_empty;
    kind = (int)h.kind;
    return true;
}

}  // namespace pray
