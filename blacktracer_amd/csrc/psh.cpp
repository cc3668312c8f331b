// psh.cpp -- deterministic perfect-spatial-hash builder (see psh.h).
#include "psh.h"

#include <algorithm>
#include <cmath>
#include <numeric>

#include "bt_internal.h"

namespace bt {

namespace {

int gcd(int a, int b) { return b == 0 ? a : gcd(b, a % b); }

inline int wrap(int v, int w) { return ((v % w) + w) % w; }

// PSHOffsetTable.h:59-62 / 70-78
int hash_width_for(int n) {
    float d = (float)std::pow(n * 1.1f, 1.f / 2.f);
    return (int)(d + 1.1f);
}
int offset_width_for(int n, int hw) {
    float d = (float)std::pow(n / 4.f, 1.f / 2.f);
    int w = (int)(d + 1.1f);
    while (gcd(w, hw) > 1) w++;
    return w;
}

struct Rng {
    uint64_t s;
    uint32_t next() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (uint32_t)(s >> 33);
    }
};

}  // namespace

bool psh_lookup(const int32_t* ot, const float* ht, const int32_t* tag, int hw, int ow, int x,
                int y, float* out2) {
    int ix = (x + ow) % ow, iy = (y + ow) % ow;
    int ox = ot[2 * (ix * ow + iy)], oy = ot[2 * (ix * ow + iy) + 1];
    int ax = (x + hw) % hw + ox, ay = (y + hw) % hw + oy;
    int hx = (ax + hw) % hw, hy = (ay + hw) % hw;
    int s = hx * hw + hy;
    if (tag[2 * s] != x || tag[2 * s + 1] != y) {
        out2[0] = -2.f;
        out2[1] = -2.f;
        return false;
    }
    out2[0] = ht[2 * s];
    out2[1] = ht[2 * s + 1];
    return true;
}

int psh_build(const std::vector<int32_t>& keys, const std::vector<float>& vals, PshTable* out) {
    const int n = (int)(keys.size() / 2);
    if (n <= 0) return fail(BT_ERR_INVALID, "psh_build: no elements");
    const int hw = hash_width_for(n);
    int ow = offset_width_for(n, hw);
    Rng rng{0x9e3779b97f4a7c15ull};
    for (int attempt = 0; attempt < 64; attempt++) {
        const int nb = ow * ow;
        // buckets by hash1(key) (PSHOffsetTable.h:269-279)
        std::vector<int> count(nb, 0), start(nb + 1, 0), members(n);
        for (int e = 0; e < n; e++) count[wrap(keys[2 * e], ow) * ow + wrap(keys[2 * e + 1], ow)]++;
        for (int b = 0; b < nb; b++) start[b + 1] = start[b] + count[b];
        std::vector<int> fillp(start.begin(), start.end() - 1);
        for (int e = 0; e < n; e++) {
            int b = wrap(keys[2 * e], ow) * ow + wrap(keys[2 * e + 1], ow);
            members[fillp[b]++] = e;
        }
        // largest buckets first (PSHOffsetTable.h:281-294)
        std::vector<int> order(nb);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return count[a] > count[b]; });
        std::vector<uint8_t> filled((size_t)hw * hw, 0);
        std::vector<int32_t> offsets(2 * (size_t)nb, 0);
        std::vector<int> empty;   // lazily built free list for one-element buckets
        bool ok = true;
        for (int bi = 0; bi < nb && ok; bi++) {
            int b = order[bi];
            int cnt = count[b];
            if (cnt == 0) break;
            const int* mem = &members[start[b]];
            int best_x = 0, best_y = 0;
            bool found = false;
            if (cnt == 1) {
                if (empty.empty()) {
                    for (int s = (int)filled.size() - 1; s >= 0; s--)
                        if (!filled[s]) empty.push_back(s);
                }
                while (!empty.empty() && filled[empty.back()]) empty.pop_back();
                if (!empty.empty()) {
                    int s = empty.back();
                    int e = mem[0];
                    best_x = s / hw - wrap(keys[2 * e], hw);
                    best_y = s % hw - wrap(keys[2 * e + 1], hw);
                    found = true;
                }
            } else {
                // scan offsets from a random seed, as findOffsetRandom does
                int sx = (int)(rng.next() % (uint32_t)hw), sy = (int)(rng.next() % (uint32_t)hw);
                for (int dx = 0; dx < hw && !found; dx++) {
                    for (int dy = 0; dy < hw && !found; dy++) {
                        int ox = (sx + dx) % hw, oy = (sy + dy) % hw;
                        bool fits = true;
                        for (int q = 0; q < cnt && fits; q++) {
                            int e = mem[q];
                            int hx = (wrap(keys[2 * e], hw) + ox) % hw;
                            int hy = (wrap(keys[2 * e + 1], hw) + oy) % hw;
                            if (filled[(size_t)hx * hw + hy]) fits = false;
                            for (int r = 0; r < q && fits; r++) {
                                int e2 = mem[r];
                                if ((wrap(keys[2 * e2], hw) + ox) % hw == hx &&
                                    (wrap(keys[2 * e2 + 1], hw) + oy) % hw == hy)
                                    fits = false;
                            }
                        }
                        if (fits) {
                            best_x = ox;
                            best_y = oy;
                            found = true;
                        }
                    }
                }
            }
            if (!found) {
                ok = false;
                break;
            }
            offsets[2 * b] = best_x;
            offsets[2 * b + 1] = best_y;
            for (int q = 0; q < cnt; q++) {
                int e = mem[q];
                int hx = wrap(wrap(keys[2 * e], hw) + best_x, hw);
                int hy = wrap(wrap(keys[2 * e + 1], hw) + best_y, hw);
                filled[(size_t)hx * hw + hy] = 1;
            }
        }
        if (!ok) {
            // PSHOffsetTable.h:236-244: grow the offset table and retry
            ow += 5;
            while (gcd(ow, hw % ow) > 1) ow++;
            continue;
        }
        out->hash_width = hw;
        out->offset_width = ow;
        out->offset_table = offsets;
        out->hash_table.assign(2 * (size_t)hw * hw, 0.f);
        out->hash_pos_tag.assign(2 * (size_t)hw * hw, 0);
        for (int e = 0; e < n; e++) {
            int b = wrap(keys[2 * e], ow) * ow + wrap(keys[2 * e + 1], ow);
            int hx = wrap(wrap(keys[2 * e], hw) + offsets[2 * b], hw);
            int hy = wrap(wrap(keys[2 * e + 1], hw) + offsets[2 * b + 1], hw);
            size_t s = (size_t)hx * hw + hy;
            out->hash_table[2 * s] = vals[2 * e];
            out->hash_table[2 * s + 1] = vals[2 * e + 1];
            out->hash_pos_tag[2 * s] = keys[2 * e];
            out->hash_pos_tag[2 * s + 1] = keys[2 * e + 1];
        }
        return BT_OK;
    }
    return fail(BT_ERR_STATE, "psh_build: no perfect hash found");
}

}  // namespace bt
