// world.hpp -- host-side sparse voxel tree store (C++ mirror of src/voxel/*).
//
// Layout and semantics follow the reference so the pools it produces are the
// bytes the reference uploads:
//   - 64^3 chunk -> 4^3 L1 cells of 16^3 -> 4^3 leaves of 4^3 voxels; a leaf is a
//     u64 occupancy mask; cell bit = x | y<<2 | z<<4 (src/voxel/tree.rs:8-35).
//   - Inner = {mask u64, run_ptr u32 (bit 31 FULL), leaf_prefix u32 (low 28 bits
//     prefix, high 4 bits water line)} (src/voxel/geometry.rs:12-52).
//   - Runs are hash-consed and refcounted; freed runs go to a per-length LIFO free
//     list and are reused before the pool grows (src/voxel/geometry.rs:133-211).
//     Pool offsets therefore depend only on the intern/release order.
//   - Attribute words live in a u32 pool with power-of-two size classes, minimum 4
//     (src/voxel/attributes.rs:6-88).
#pragma once

#include <cstddef>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace vxh {

using BlockId = uint16_t;

constexpr BlockId AIR = 0, STONE = 1, DIRT = 2, GRASS = 3, SAND = 4, GRAVEL = 5, LOG = 6, LEAVES = 7,
                  PLANKS = 8, COBBLE = 9, GLASS = 10, GLOWSTONE = 11, BEDROCK = 12, WATER = 13, SNOW = 14,
                  PODZOL = 15, PINE_LEAVES = 16, LICHEN = 17, TALL_GRASS = 18, MEADOW = 19, AMBER_LAMP = 20,
                  AZURE_LAMP = 21, VERDANT_LAMP = 22, MEADOW_SIDE = 23, GRASS_TALL = 24, REEDS = 25;
constexpr int BLOCK_COUNT = 26;

// Block predicates from the BLOCKS table (src/block.rs:133-185, :270-287).
bool is_foliage(BlockId id);
bool is_cutout(BlockId id);
bool is_glass(BlockId id);
bool is_emitter(BlockId id);

constexpr uint32_t FULL_FLAG = 1u << 31;
constexpr uint32_t PREFIX_MASK = 0x0FFFFFFFu;
constexpr uint32_t UNIFORM_BIT = 1u << 31;
constexpr int DIM = 64;
constexpr int VOL = DIM * DIM * DIM;

inline size_t dense_index(uint32_t x, uint32_t y, uint32_t z) { return x | (z << 6) | (y << 12); }
inline uint32_t cell_bit(uint32_t x, uint32_t y, uint32_t z) { return x | (y << 2) | (z << 4); }
inline uint32_t l1_bit(uint32_t x, uint32_t y, uint32_t z) { return cell_bit(x >> 4, y >> 4, z >> 4); }
inline uint32_t leaf_bit(uint32_t x, uint32_t y, uint32_t z) {
    return cell_bit((x >> 2) & 3, (y >> 2) & 3, (z >> 2) & 3);
}
inline uint32_t voxel_bit(uint32_t x, uint32_t y, uint32_t z) { return cell_bit(x & 3, y & 3, z & 3); }
inline uint32_t below(uint64_t mask, uint32_t bit) {
    return (uint32_t)__builtin_popcountll(bit >= 64 ? mask : (mask & ((1ull << bit) - 1)));
}
inline bool has(uint64_t mask, uint32_t bit) { return (mask >> bit) & 1; }

struct Inner {
    uint64_t mask = 0;
    uint32_t run_ptr = 0;
    uint32_t leaf_prefix = 0;
    static Inner full(uint32_t leaf_prefix) { return Inner{~0ull, FULL_FLAG, leaf_prefix}; }
    bool is_full() const { return (run_ptr & FULL_FLAG) != 0; }
    bool is_empty() const { return mask == 0; }
    uint32_t run() const { return run_ptr & ~FULL_FLAG; }
    uint32_t count() const { return (uint32_t)__builtin_popcountll(mask); }
    bool operator==(const Inner &o) const {
        return mask == o.mask && run_ptr == o.run_ptr && leaf_prefix == o.leaf_prefix;
    }
};
static_assert(sizeof(Inner) == 16, "Inner is one vec4<u32>");

struct LocalL1 {
    uint64_t mask = 0;
    uint32_t leaf_start = 0;
    uint32_t leaf_prefix = 0;
    bool full = false;
    uint8_t water_line = 0;
};

struct LocalTree {
    uint64_t root_mask = 0, dry_mask = 0;
    std::vector<LocalL1> l1;
    std::vector<uint64_t> leaves;
    std::vector<uint32_t> attrs, palette;
    bool has_uniform = false;
    BlockId uniform = 0;
    bool has_water = false, has_foliage = false, has_cutout = false, has_glass = false, has_emitter = false;
    uint32_t emitters = 0, solid_count = 0, leaf_count = 0;
    uint8_t aabb_min[3] = {0, 0, 0}, aabb_max[3] = {0, 0, 0};
    bool is_empty() const { return root_mask == 0; }
    bool is_full_solid() const;
};

LocalTree build_local(const BlockId *dense);
BlockId local_block(const LocalTree &t, uint32_t x, uint32_t y, uint32_t z);
uint32_t encode_leaf(const BlockId ids[64], uint64_t present, std::vector<uint32_t> &out);
BlockId decode_leaf(const uint32_t *words, uint32_t entry, uint32_t vb);

template <class T> class RunPool {
public:
    RunPool() : free_(65) {}
    const std::vector<T> &data() const { return data_; }
    const T *get(uint32_t off) const { return data_.data() + off; }
    std::pair<uint32_t, bool> intern(const T *run, uint32_t len);
    void add_ref(uint32_t off);
    bool release(uint32_t off);
    uint32_t ref_count(uint32_t off) const { return refs_[off]; }
    std::vector<std::pair<uint32_t, uint32_t>> take_dirty() {
        std::vector<std::pair<uint32_t, uint32_t>> d;
        d.swap(dirty_);
        return d;
    }
    const std::vector<std::pair<uint32_t, uint32_t>> &dirty() const { return dirty_; }
    uint64_t live_elems = 0, referenced_elems = 0;
    size_t live_runs() const { return live_runs_; }

private:
    static uint64_t hash_run(const T *run, uint32_t len);
    uint32_t alloc(uint32_t len);
    std::vector<T> data_;
    std::vector<uint32_t> refs_;
    std::vector<uint8_t> lens_;
    std::unordered_multimap<uint64_t, uint32_t> table_;
    std::vector<std::vector<uint32_t>> free_;
    std::vector<std::pair<uint32_t, uint32_t>> dirty_;
    size_t live_runs_ = 0;
};

class WordPool {
public:
    WordPool() : free_(32) {}
    const std::vector<uint32_t> &data() const { return data_; }
    uint32_t alloc(size_t len);
    void free(uint32_t off, size_t len);
    void write(uint32_t off, const uint32_t *words, size_t n);
    uint32_t alloc_write(const uint32_t *words, size_t n) {
        uint32_t off = alloc(n);
        write(off, words, n);
        return off;
    }
    std::vector<std::pair<uint32_t, uint32_t>> take_dirty() {
        std::vector<std::pair<uint32_t, uint32_t>> d;
        d.swap(dirty_);
        return d;
    }
    const std::vector<std::pair<uint32_t, uint32_t>> &dirty() const { return dirty_; }
    uint64_t used_words = 0;

private:
    std::vector<uint32_t> data_;
    std::vector<std::vector<uint32_t>> free_;
    std::vector<std::pair<uint32_t, uint32_t>> dirty_;
};

// ChunkKey (src/voxel/chunk.rs:6-60): a chunk of 64^3 voxels of size 2^lod at
// position pos, covering [pos * (64 << lod), (pos + 1) * (64 << lod)) in blocks.
struct Key {
    int32_t x, y, z;
    int32_t lod = 0;
    bool operator==(const Key &o) const { return x == o.x && y == o.y && z == o.z && lod == o.lod; }
    int32_t size() const { return 64 << lod; }
    int32_t voxel_size() const { return 1 << lod; }
    Key parent() const { return Key{x >> 1, y >> 1, z >> 1, lod + 1}; }
    Key child(int32_t dx, int32_t dy, int32_t dz) const { return Key{x * 2 + dx, y * 2 + dy, z * 2 + dz, lod - 1}; }
};
struct KeyHash {
    size_t operator()(const Key &k) const {
        uint64_t h = (uint64_t)(uint32_t)k.x * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)(uint32_t)k.y * 0xC2B2AE3D27D4EB4Full + (h << 6);
        h ^= (uint64_t)(uint32_t)k.z * 0x165667B19E3779F9ull + (h >> 2);
        h ^= (uint64_t)(uint32_t)k.lod * 0xD6E8FEB86659FD93ull + (h << 3);
        return (size_t)h;
    }
};

struct ChunkRecord {
    Key key{};
    Inner root{};
    bool attr_uniform = true;
    uint32_t attr_base = 0;  // uniform id or table base
    uint32_t attr_len = 0;
    std::vector<std::pair<uint32_t, size_t>> regions;
    int32_t aabb_min[3] = {0, 0, 0}, aabb_max[3] = {0, 0, 0};
    uint32_t solid_count = 0, leaf_count = 0;
    bool has_water = false, has_foliage = false, has_cutout = false, has_glass = false, has_emitter = false;
    uint32_t emitters = 0;
    uint64_t dry_mask = 0;
    uint32_t version = 0;
};

class World {
public:
    RunPool<uint64_t> leaves;
    RunPool<Inner> inners;
    WordPool bricks;
    std::unordered_map<Key, ChunkRecord, KeyHash> chunks;
    std::vector<Key> order;  // insertion order of resident chunks (render set order)
    uint64_t version = 0;
    uint32_t emitter_chunks = 0;

    void insert_local(Key key, const LocalTree &t);
    bool remove_chunk(Key key);
    BlockId get_block(int32_t x, int32_t y, int32_t z) const;
    bool set_block(int32_t x, int32_t y, int32_t z, BlockId id);
    bool extract_dense(Key key, BlockId *out) const;

private:
    Inner make_root(uint64_t root_mask, const std::vector<Inner> &run);
    void release_root(Inner root);
    bool l1_node(Inner root, uint32_t l1b, Inner &out) const;
    uint64_t leaf_mask(Inner l1, uint32_t lfb) const;
    BlockId leaf_block(const ChunkRecord &rec, uint32_t leaf_index, uint32_t vb) const;
};

constexpr int32_t WORLD_HEIGHT = 512;

}  // namespace vxh

// Opaque handle behind the vxw_* C ABI.
struct vxw_world {
    vxh::World w;
};
