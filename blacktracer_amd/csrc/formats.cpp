// formats.cpp -- the reference's on-disk caches and its star catalogue
// reader, host side.
//
// The reference keeps grids and star sets in cereal binary archives
// (BlackTracer.cpp:20-47 writes them, :226-261 and :300-336 read them back):
//   .grid  Grid::serialize (Grid.h:33-39): MAXLEVEL, N, M, then the
//          PSHOffsetTable (PSHOffsetTable.h:44-48): hashTable (float2),
//          offsetTable (int2), hashPosTag (int2), hashTableWidth,
//          offsetTableWidth.
//   .star  StarProcessor::serialize (StarProcessor.h:19-23): binaryStarTree
//          (int), starPos (float), starMag (float), starSize, treeSize,
//          treeLevel.
// cereal's binary archive writes an int as its 4 bytes, a vector as a
// 64-bit element count followed by the elements (an arithmetic element type
// as one block of raw bytes), all in the host's byte order (little endian
// here and on the reference's x64 build).  float2 and int2 elements are
// written as their two members, x then y.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "bt_grid.h"

using namespace bt;

namespace {

struct Out {
    FILE* f;
    bool ok = true;
    void bytes(const void* p, size_t n) {
        if (ok && n && std::fwrite(p, 1, n, f) != n) ok = false;
    }
    void i32(int32_t v) { bytes(&v, 4); }
    void count(uint64_t n) { bytes(&n, 8); }
};

struct In {
    FILE* f;
    bool ok = true;
    void bytes(void* p, size_t n) {
        if (ok && n && std::fread(p, 1, n, f) != n) ok = false;
    }
    int32_t i32() {
        int32_t v = 0;
        bytes(&v, 4);
        return v;
    }
    uint64_t count() {
        uint64_t n = 0;
        bytes(&n, 8);
        return n;
    }
    // a vector of `elem`-byte elements: into dst (nullable), at most `cap` elements
    uint64_t vec(void* dst, size_t elem, uint64_t cap) {
        const uint64_t n = count();
        if (!ok || n > cap) {
            ok = false;
            return 0;
        }
        if (dst) {
            bytes(dst, (size_t)n * elem);
        } else if (n && std::fseek(f, (long)(n * elem), SEEK_CUR) != 0) {
            ok = false;
        }
        return n;
    }
};

// metric::wrapToPi(float&, float&) (Metric.h:187-196): each statement in
// double (PI2 and PI are double constants), stored back to float.
void wrap_to_pi_f(float& thetaW, float& phiW) {
    const double PI = 3.14159265358979323846, PI2 = 2 * PI;
    thetaW = (float)std::fmod((double)thetaW, PI2);
    while (thetaW < 0) thetaW = (float)(thetaW + PI2);
    if (thetaW > PI) {
        thetaW = (float)(thetaW - 2 * (thetaW - PI));
        phiW = (float)(phiW + PI);
    }
    while (phiW < 0) phiW = (float)(phiW + PI2);
    phiW = (float)std::fmod((double)phiW, PI2);
}

}  // namespace

extern "C" {

// readStars2 (StarProcessor.h:71-96) and the magnitude cut of the
// StarProcessor constructor (StarProcessor.h:34-46).  Records of five
// whitespace-separated numbers: magnitude, B-V colour, right ascension,
// declination, one ignored value, angles in radians; theta = pi/2 - dec,
// wrapped with (theta, ra).  Stars brighter than or at the cut are kept, in
// file order.  The reference's loop tests eof before reading and so appends
// one more record, made of a failed read, after the last one; here a
// trailing incomplete record is dropped instead.
int bt_star_catalog_parse(const char* text, size_t len, float magnitude_cut, float* stars4, int capacity,
                          int* count) {
    if ((!text && len) || !count || capacity < 0 || (capacity > 0 && !stars4))
        return fail(BT_ERR_INVALID, "bt_star_catalog_parse: bad arguments");
    std::string buf(text ? text : "", len);
    const char* p = buf.c_str();
    const char* end = p + buf.size();
    int n = 0;
    while (true) {
        float v[5];
        int got = 0;
        for (; got < 5; got++) {
            while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\f' || *p == '\v'))
                p++;
            if (p >= end) break;
            char* q = nullptr;
            v[got] = std::strtof(p, &q);
            if (q == p) return fail(BT_ERR_INVALID, "bt_star_catalog_parse: not a number in the catalogue");
            p = q;
        }
        if (got < 5) break;
        float mag = v[0], col = v[1], ra = v[2], dec = v[3];
        dec = (float)(3.14159265358979323846 / 2. - dec);
        wrap_to_pi_f(dec, ra);
        if (mag > magnitude_cut) continue;
        if (n < capacity) {
            stars4[4 * (size_t)n] = dec;
            stars4[4 * (size_t)n + 1] = ra;
            stars4[4 * (size_t)n + 2] = mag;
            stars4[4 * (size_t)n + 3] = col;
        }
        n++;
    }
    *count = n;
    if (n > capacity && capacity > 0) return fail(BT_ERR_INVALID, "bt_star_catalog_parse: capacity too small");
    return BT_OK;
}

int bt_star_file_write(const char* path, const int32_t* tree, int tree_level, const float* pos, const float* mag,
                       int star_count) {
    if (!path || !tree || tree_level < 1 || tree_level > 20 || star_count < 0 ||
        (star_count > 0 && (!pos || !mag)))
        return fail(BT_ERR_INVALID, "bt_star_file_write: bad arguments");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(BT_ERR_INVALID, std::string("bt_star_file_write: cannot open ") + path);
    Out o{f};
    const int tree_size = (1 << (tree_level + 1)) - 1;
    o.count((uint64_t)tree_size);
    o.bytes(tree, sizeof(int32_t) * (size_t)tree_size);
    o.count(2 * (uint64_t)star_count);
    o.bytes(pos, sizeof(float) * 2 * (size_t)star_count);
    o.count(2 * (uint64_t)star_count);
    o.bytes(mag, sizeof(float) * 2 * (size_t)star_count);
    o.i32(star_count);
    o.i32(tree_size);
    o.i32(tree_level);
    const bool ok = o.ok && std::fclose(f) == 0;
    return ok ? BT_OK : fail(BT_ERR_INVALID, "bt_star_file_write: write failed");
}

// Null arrays: the sizes only (star_count, tree_level).
int bt_star_file_read(const char* path, int32_t* tree, float* pos, float* mag, int* star_count, int* tree_level) {
    if (!path || !star_count || !tree_level) return fail(BT_ERR_INVALID, "bt_star_file_read: bad arguments");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(BT_ERR_NOTFOUND, std::string("bt_star_file_read: cannot open ") + path);
    In in{f};
    const uint64_t kMax = uint64_t(1) << 31;
    // sizes first (the trailing ints), then the arrays
    const uint64_t nt = in.vec(nullptr, 4, kMax);
    const uint64_t np = in.vec(nullptr, 4, kMax);
    const uint64_t nm = in.vec(nullptr, 4, kMax);
    const int32_t S = in.i32(), tsize = in.i32(), level = in.i32();
    bool ok = in.ok && S >= 0 && np == 2 * (uint64_t)S && nm == np && level >= 1 && level <= 20 &&
              tsize == (1 << (level + 1)) - 1 && nt == (uint64_t)tsize;
    if (ok && (tree || pos || mag)) {
        std::rewind(f);
        if (in.vec(tree, 4, nt) != nt || in.vec(pos, 4, np) != np || in.vec(mag, 4, nm) != nm || !in.ok) ok = false;
    }
    std::fclose(f);
    if (!ok) return fail(BT_ERR_INVALID, std::string("bt_star_file_read: not a star file: ") + path);
    *star_count = S;
    *tree_level = level;
    return BT_OK;
}

int bt_grid_file_write(const char* path, const bt_grid* g) {
    if (!path || !g) return fail(BT_ERR_INVALID, "bt_grid_file_write: bad arguments");
    int hw = 0, ow = 0;
    int rc;
    // Grid.h:623: uniform grids (start level == max level) build no table
    const bool table = g->start_level != g->max_level;
    std::vector<float> hash;
    std::vector<int32_t> offs, tags;
    if (table) {
        if ((rc = bt_grid_export_psh(g, &hw, &ow, nullptr, nullptr, nullptr))) return rc;
        hash.resize(2 * (size_t)hw * hw);
        tags.resize(2 * (size_t)hw * hw);
        offs.resize(2 * (size_t)ow * ow);
        if ((rc = bt_grid_export_psh(g, &hw, &ow, hash.data(), offs.data(), tags.data()))) return rc;
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(BT_ERR_INVALID, std::string("bt_grid_file_write: cannot open ") + path);
    Out o{f};
    o.i32(g->max_level);
    o.i32(g->N);
    o.i32(g->M);
    o.count(hash.size() / 2);
    o.bytes(hash.data(), sizeof(float) * hash.size());
    o.count(offs.size() / 2);
    o.bytes(offs.data(), sizeof(int32_t) * offs.size());
    o.count(tags.size() / 2);
    o.bytes(tags.data(), sizeof(int32_t) * tags.size());
    o.i32(hw);
    o.i32(ow);
    const bool ok = o.ok && std::fclose(f) == 0;
    return ok ? BT_OK : fail(BT_ERR_INVALID, "bt_grid_file_write: write failed");
}

// Null arrays: the header only (levels, sizes, table widths).  hash: 2*hw*hw
// floats, offsets: 2*ow*ow ints, tags: 2*hw*hw ints.
int bt_grid_file_read(const char* path, int* max_level, int* N, int* M, int* hash_width, int* offset_width,
                      float* hash, int32_t* offsets, int32_t* tags) {
    if (!path || !max_level || !N || !M || !hash_width || !offset_width)
        return fail(BT_ERR_INVALID, "bt_grid_file_read: bad arguments");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(BT_ERR_NOTFOUND, std::string("bt_grid_file_read: cannot open ") + path);
    In in{f};
    const uint64_t kMax = uint64_t(1) << 31;
    const int32_t ml = in.i32(), n = in.i32(), m = in.i32();
    const long tables = std::ftell(f);
    const uint64_t nh = in.vec(nullptr, 8, kMax), no = in.vec(nullptr, 8, kMax), ntag = in.vec(nullptr, 8, kMax);
    const int32_t hw = in.i32(), ow = in.i32();
    bool ok = in.ok && ml >= 0 && n > 0 && m > 0 && hw >= 0 && ow >= 0 && nh == (uint64_t)hw * hw &&
              ntag == nh && no == (uint64_t)ow * ow;
    if (ok && (hash || offsets || tags)) {
        std::fseek(f, tables, SEEK_SET);
        if (in.vec(hash, 8, nh) != nh || in.vec(offsets, 8, no) != no || in.vec(tags, 8, ntag) != ntag || !in.ok)
            ok = false;
    }
    std::fclose(f);
    if (!ok) return fail(BT_ERR_INVALID, std::string("bt_grid_file_read: not a grid file: ") + path);
    *max_level = ml;
    *N = n;
    *M = m;
    *hash_width = hw;
    *offset_width = ow;
    return BT_OK;
}

}  // extern "C"
