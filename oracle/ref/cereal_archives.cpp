// cereal_archives.cpp -- writes .star and .grid archives through the cereal
// 1.2.2 binary archive vendored with the reference (libraries/cereal-1.2.2),
// into oracle/_ref/libbtcereal.so.
//
// TEST INFRASTRUCTURE ONLY.  It lets the tests compare the library's own
// writers (formats.cpp) with what cereal itself writes, byte for byte.
// StarProcessor.h and Grid.h cannot be compiled here (they include
// cuda_runtime.h, which this image lacks), so the two archive shapes are
// declared below with the members their serialize() functions name, in the
// order they name them:
//   StarProcessor::serialize  (StarProcessor.h:19-23):
//       binaryStarTree (vector<int>), starPos, starMag (vector<float>),
//       starSize, treeSize, treeLevel (int)
//   Grid::serialize (Grid.h:34-39): MAXLEVEL, N, M (int), hasher
//   PSHOffsetTable::serialize (PSHOffsetTable.h:44-48):
//       hashTable (vector<float2>), offsetTable, hashPosTag (vector<int2>),
//       hashTableWidth, offsetTableWidth (int)
// The reference's tree has no serializer for CUDA's float2/int2; the one
// below writes x then y, as any member-wise serializer does.
#include <cstdint>
#include <fstream>
#include <vector>

#include <cereal/archives/binary.hpp>
#include <cereal/types/vector.hpp>

namespace {

struct F2 {
    float x, y;
    template <class A>
    void serialize(A& ar) { ar(x, y); }
};
struct I2 {
    int x, y;
    template <class A>
    void serialize(A& ar) { ar(x, y); }
};

struct StarArchive {
    std::vector<int> binaryStarTree;
    std::vector<float> starPos, starMag;
    int starSize = 0, treeSize = 0, treeLevel = 0;
    template <class A>
    void serialize(A& ar) { ar(binaryStarTree, starPos, starMag, starSize, treeSize, treeLevel); }
};

struct PshArchive {
    std::vector<F2> hashTable;
    std::vector<I2> offsetTable, hashPosTag;
    int hashTableWidth = 0, offsetTableWidth = 0;
    template <class A>
    void serialize(A& ar) { ar(hashTable, offsetTable, hashPosTag, hashTableWidth, offsetTableWidth); }
};

struct GridArchive {
    int MAXLEVEL = 0, N = 0, M = 0;
    PshArchive hasher;
    template <class A>
    void serialize(A& ar) { ar(MAXLEVEL, N, M, hasher); }
};

}  // namespace

extern "C" {

// tree: tree_size ints; pos, mag: 2 * star_size floats each.
int btcereal_star_write(const char* path, const int* tree, int tree_size, const float* pos, const float* mag,
                        int star_size, int tree_level) {
    StarArchive s;
    s.binaryStarTree.assign(tree, tree + tree_size);
    s.starPos.assign(pos, pos + 2 * (size_t)star_size);
    s.starMag.assign(mag, mag + 2 * (size_t)star_size);
    s.starSize = star_size;
    s.treeSize = tree_size;
    s.treeLevel = tree_level;
    std::ofstream os(path, std::ios::binary);
    if (!os) return -1;
    {
        cereal::BinaryOutputArchive ar(os);
        ar(s);
    }
    return os.good() ? 0 : -1;
}

// hash: 2 * hw * hw floats; offs: 2 * ow * ow ints; tags: 2 * hw * hw ints
// (hw = ow = 0 and null arrays for a grid without a table).
int btcereal_grid_write(const char* path, int max_level, int N, int M, const float* hash, const int* offs,
                        const int* tags, int hw, int ow) {
    GridArchive g;
    g.MAXLEVEL = max_level;
    g.N = N;
    g.M = M;
    const size_t nh = (size_t)hw * hw, no = (size_t)ow * ow;
    g.hasher.hashTable.resize(nh);
    g.hasher.hashPosTag.resize(nh);
    g.hasher.offsetTable.resize(no);
    for (size_t i = 0; i < nh; i++) {
        g.hasher.hashTable[i] = F2{hash[2 * i], hash[2 * i + 1]};
        g.hasher.hashPosTag[i] = I2{tags[2 * i], tags[2 * i + 1]};
    }
    for (size_t i = 0; i < no; i++) g.hasher.offsetTable[i] = I2{offs[2 * i], offs[2 * i + 1]};
    g.hasher.hashTableWidth = hw;
    g.hasher.offsetTableWidth = ow;
    std::ofstream os(path, std::ios::binary);
    if (!os) return -1;
    {
        cereal::BinaryOutputArchive ar(os);
        ar(g);
    }
    return os.good() ? 0 : -1;
}

}  // extern "C"
