// psh.h -- perfect spatial hash (Lefebvre & Hoppe 2006) of the traced grid
// points, in the table layout of PSHOffsetTable (PSHOffsetTable.h:327-333):
// hash_table 2*hw*hw floats, offset_table 2*ow*ow ints, hash_pos_tag
// 2*hw*hw ints.  Lookup (PSHOffsetTable.h:221-234,335-340, kernel.cuh:705-733):
//   o = offset[hash1(key)], slot = hash0(hash0(key) + o),
//   value = hash_pos_tag[slot] == key ? hash_table[slot] : (-2, -2).
// The reference seeds its offset search with srand(time(NULL)); this builder
// uses a fixed-seed generator, so a table is a deterministic function of the
// key set.  Any valid table answers every lookup identically.
#pragma once

#include <cstdint>
#include <vector>

namespace bt {

struct PshTable {
    int hash_width = 0;
    int offset_width = 0;
    std::vector<float> hash_table;
    std::vector<int32_t> offset_table;
    std::vector<int32_t> hash_pos_tag;
};

// keys: pairs (i, j) (non-negative); vals: pairs (theta, phi).
int psh_build(const std::vector<int32_t>& keys, const std::vector<float>& vals, PshTable* out);

// Host lookup with the reference's semantics; returns false for (-2,-2).
bool psh_lookup(const int32_t* offset_table, const float* hash_table, const int32_t* hash_pos_tag,
                int hw, int ow, int x, int y, float* out2);

}  // namespace bt
