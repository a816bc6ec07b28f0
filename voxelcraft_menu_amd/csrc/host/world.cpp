// world.cpp -- host-side sparse voxel tree store (C++ mirror of src/voxel/*).
#include "world.hpp"

#include <algorithm>
#include <cstring>

namespace vxh {

namespace {
struct BlockFlags {
    bool foliage, cutout, emitter;
};
// foliage / cutout / light != 0 columns of the BLOCKS table (src/block.rs:133-185).
constexpr BlockFlags BLOCKS[BLOCK_COUNT] = {
    {false, false, false},  // air
    {false, false, false},  // stone
    {false, false, false},  // dirt
    {false, false, false},  // grass
    {false, false, false},  // sand
    {false, false, false},  // gravel
    {false, false, false},  // log
    {false, true, false},   // leaves
    {false, false, false},  // planks
    {false, false, false},  // cobblestone
    {false, false, false},  // glass
    {false, false, true},   // glowstone
    {false, false, false},  // bedrock
    {false, false, false},  // water
    {false, false, false},  // snow
    {false, false, false},  // podzol
    {false, true, false},   // pine leaves
    {false, false, false},  // lichen
    {true, false, false},   // tall grass
    {false, false, false},  // meadow
    {false, false, true},   // amber lamp
    {false, false, true},   // azure lamp
    {false, false, true},   // verdant lamp
    {false, false, false},  // meadow-side
    {true, false, false},   // grass, tall
    {true, false, false},   // reeds
};
const BlockFlags &def(BlockId id) { return BLOCKS[std::min<int>(id, BLOCK_COUNT - 1)]; }
}  // namespace

bool is_foliage(BlockId id) { return def(id).foliage; }
bool is_cutout(BlockId id) { return def(id).cutout; }
bool is_glass(BlockId id) { return id == GLASS; }
bool is_emitter(BlockId id) { return def(id).emitter; }

bool LocalTree::is_full_solid() const {
    if (l1.size() != 64) return false;
    for (const auto &n : l1)
        if (!n.full) return false;
    return true;
}

// ---- palette encoding (src/voxel/attributes.rs:90-166) ----
static uint32_t bits_for(size_t n) {
    if (n <= 2) return 1;
    if (n <= 4) return 2;
    if (n <= 16) return 4;
    return 8;
}

uint32_t encode_leaf(const BlockId ids[64], uint64_t present, std::vector<uint32_t> &out) {
    BlockId pal[64] = {0};
    size_t n = 0;
    for (int i = 0; i < 64; i++) {
        if (!((present >> i) & 1)) continue;
        BlockId id = ids[i];
        bool seen = false;
        for (size_t k = 0; k < n; k++)
            if (pal[k] == id) { seen = true; break; }
        if (!seen) pal[n++] = id;
    }
    if (n <= 1) return UNIFORM_BIT | (uint32_t)pal[0];
    uint32_t bits = bits_for(n);
    uint32_t off = (uint32_t)out.size();
    out.push_back((uint32_t)n | (bits << 8));
    for (size_t w = 0; w < (n + 1) / 2; w++) {
        uint32_t lo = pal[2 * w];
        uint32_t hi = (2 * w + 1 < n) ? pal[2 * w + 1] : 0;
        out.push_back(lo | (hi << 16));
    }
    size_t idx_words = (64 * bits) / 32;
    size_t idx_base = out.size();
    out.resize(idx_base + idx_words, 0);
    for (int i = 0; i < 64; i++) {
        if (!((present >> i) & 1)) continue;
        uint32_t idx = 0;
        while (pal[idx] != ids[i]) idx++;
        uint32_t bitpos = (uint32_t)i * bits;
        out[idx_base + (bitpos >> 5)] |= idx << (bitpos & 31);
    }
    return off;
}

BlockId decode_leaf(const uint32_t *words, uint32_t entry, uint32_t vb) {
    if (entry & UNIFORM_BIT) return (BlockId)entry;
    size_t base = entry;
    uint32_t hdr = words[base];
    size_t n = hdr & 0xFF;
    uint32_t bits = (hdr >> 8) & 0xFF;
    size_t idx_base = base + 1 + (n + 1) / 2;
    uint32_t bitpos = vb * bits;
    uint32_t w = words[idx_base + (bitpos >> 5)];
    size_t idx = (w >> (bitpos & 31)) & ((1u << bits) - 1);
    uint32_t pw = words[base + 1 + idx / 2];
    return (BlockId)(pw >> ((idx & 1) * 16));
}

// ---- build_local (src/voxel/tree.rs:89-204) ----
LocalTree build_local(const BlockId *dense) {
    LocalTree t;
    BlockId ids[64];
    std::fill(ids, ids + 64, AIR);
    bool have_first = false;
    BlockId first = 0;
    bool all_same = true;
    uint8_t amin[3] = {64, 64, 64}, amax[3] = {0, 0, 0};
    for (uint32_t l1z = 0; l1z < 4; l1z++)
        for (uint32_t l1y = 0; l1y < 4; l1y++)
            for (uint32_t l1x = 0; l1x < 4; l1x++) {
                uint64_t l1mask = 0;
                uint32_t leaf_start = (uint32_t)t.leaves.size();
                bool all_full = true, l1_dry = false;
                uint32_t max_nonwater_y = 0;
                for (uint32_t lz = 0; lz < 4; lz++)
                    for (uint32_t ly = 0; ly < 4; ly++)
                        for (uint32_t lx = 0; lx < 4; lx++) {
                            uint64_t m = 0;
                            uint32_t bx = l1x * 16 + lx * 4, by = l1y * 16 + ly * 4, bz = l1z * 16 + lz * 4;
                            for (uint32_t vz = 0; vz < 4; vz++)
                                for (uint32_t vy = 0; vy < 4; vy++) {
                                    size_t row = dense_index(bx, by + vy, bz + vz);
                                    for (uint32_t vx = 0; vx < 4; vx++) {
                                        BlockId id = dense[row + vx];
                                        if (id == AIR) continue;
                                        uint32_t vb = cell_bit(vx, vy, vz);
                                        m |= 1ull << vb;
                                        ids[vb] = id;
                                        t.has_water |= id == WATER;
                                        l1_dry |= id != WATER;
                                        if (id != WATER) max_nonwater_y = std::max(max_nonwater_y, ly * 4 + vy);
                                        t.has_foliage |= is_foliage(id);
                                        t.has_cutout |= is_cutout(id);
                                        t.has_glass |= is_glass(id);
                                        bool e = is_emitter(id);
                                        t.has_emitter |= e;
                                        t.emitters += e ? 1u : 0u;
                                        if (!have_first) { have_first = true; first = id; }
                                        else if (first != id) all_same = false;
                                        uint8_t x = (uint8_t)(bx + vx), y = (uint8_t)(by + vy), z = (uint8_t)(bz + vz);
                                        amin[0] = std::min(amin[0], x); amin[1] = std::min(amin[1], y);
                                        amin[2] = std::min(amin[2], z);
                                        amax[0] = std::max<uint8_t>(amax[0], x + 1);
                                        amax[1] = std::max<uint8_t>(amax[1], y + 1);
                                        amax[2] = std::max<uint8_t>(amax[2], z + 1);
                                    }
                                }
                            if (m != 0) {
                                l1mask |= 1ull << cell_bit(lx, ly, lz);
                                t.leaves.push_back(m);
                                uint32_t e = encode_leaf(ids, m, t.palette);
                                t.attrs.push_back(e);
                                if (m != ~0ull) all_full = false;
                            } else {
                                all_full = false;
                            }
                        }
                if (l1mask != 0) {
                    if (l1_dry) t.dry_mask |= 1ull << cell_bit(l1x, l1y, l1z);
                    bool full = all_full && l1mask == ~0ull;
                    if (full) t.leaves.resize(leaf_start);
                    LocalL1 n;
                    n.mask = l1mask;
                    n.leaf_start = leaf_start;
                    n.leaf_prefix = t.leaf_count;
                    n.full = full;
                    n.water_line = l1_dry ? (uint8_t)max_nonwater_y : 15;
                    t.l1.push_back(n);
                    t.leaf_count += (uint32_t)__builtin_popcountll(l1mask);
                    t.root_mask |= 1ull << cell_bit(l1x, l1y, l1z);
                }
            }
    uint32_t sc = 0;
    for (uint64_t m : t.leaves) sc += (uint32_t)__builtin_popcountll(m);
    for (const auto &n : t.l1)
        if (n.full) sc += 4096;
    t.solid_count = sc;
    if (have_first && all_same) {
        t.has_uniform = true;
        t.uniform = first;
        t.attrs.clear();
        t.palette.clear();
    }
    if (t.solid_count > 0) {
        std::memcpy(t.aabb_min, amin, 3);
        std::memcpy(t.aabb_max, amax, 3);
    }
    return t;
}

// local_block (src/voxel/tree.rs:206-236)
BlockId local_block(const LocalTree &t, uint32_t x, uint32_t y, uint32_t z) {
    uint32_t l1b = l1_bit(x, y, z);
    if (!has(t.root_mask, l1b)) return AIR;
    const LocalL1 &n = t.l1[below(t.root_mask, l1b)];
    uint32_t lfb = leaf_bit(x, y, z);
    if (!has(n.mask, lfb)) return AIR;
    uint32_t li = below(n.mask, lfb);
    uint32_t vb = voxel_bit(x, y, z);
    uint64_t m = n.full ? ~0ull : t.leaves[n.leaf_start + li];
    if (!has(m, vb)) return AIR;
    if (t.has_uniform) return t.uniform;
    uint32_t entry = t.attrs[n.leaf_prefix + li];
    if (entry & UNIFORM_BIT) return (BlockId)entry;
    return decode_leaf(t.palette.data(), entry, vb);
}

// ---- RunPool (src/voxel/geometry.rs:74-221) ----
template <class T> uint64_t RunPool<T>::hash_run(const T *run, uint32_t len) {
    // A multiply-rotate hash over the run's 64-bit words (the element types are 8 and 16
    // bytes); only lookups depend on it, never offsets.
    static_assert(sizeof(T) % 8 == 0, "runs hash as 64-bit words");
    const uint64_t *w = reinterpret_cast<const uint64_t *>(run);
    const size_t n = sizeof(T) / 8 * len;
    uint64_t h = 0x9E3779B97F4A7C15ull ^ len;
    for (size_t i = 0; i < n; i++) h = (((h << 5) | (h >> 59)) ^ w[i]) * 0x517CC1B727220A95ull;
    return h ^ (h >> 29);
}

template <class T> uint32_t RunPool<T>::alloc(uint32_t len) {
    if (!free_[len].empty()) {
        uint32_t off = free_[len].back();
        free_[len].pop_back();
        return off;
    }
    uint32_t off = (uint32_t)data_.size();
    data_.resize(off + len, T{});
    refs_.resize(off + len, 0);
    lens_.resize(off + len, 0);
    return off;
}

template <class T> std::pair<uint32_t, bool> RunPool<T>::intern(const T *run, uint32_t len) {
    uint64_t h = hash_run(run, len);
    auto range = table_.equal_range(h);
    for (auto it = range.first; it != range.second; ++it) {
        uint32_t o = it->second;
        if (lens_[o] == len && std::equal(run, run + len, data_.begin() + o)) {
            refs_[o] += 1;
            referenced_elems += len;
            return {o, false};
        }
    }
    uint32_t off = alloc(len);
    std::copy(run, run + len, data_.begin() + off);
    refs_[off] = 1;
    lens_[off] = (uint8_t)len;
    table_.emplace(h, off);
    dirty_.push_back({off, len});
    live_elems += len;
    referenced_elems += len;
    live_runs_++;
    return {off, true};
}

template <class T> void RunPool<T>::add_ref(uint32_t off) {
    refs_[off] += 1;
    referenced_elems += lens_[off];
}

template <class T> bool RunPool<T>::release(uint32_t off) {
    uint32_t len = lens_[off];
    refs_[off] -= 1;
    referenced_elems -= len;
    if (refs_[off] != 0) return false;
    uint64_t h = hash_run(data_.data() + off, len);
    auto range = table_.equal_range(h);
    for (auto it = range.first; it != range.second; ++it)
        if (it->second == off) {
            table_.erase(it);
            break;
        }
    free_[len].push_back(off);
    live_elems -= len;
    live_runs_--;
    return true;
}

template class RunPool<uint64_t>;
template class RunPool<Inner>;

// ---- WordPool (src/voxel/attributes.rs:6-88) ----
static void class_of(size_t len, size_t &cls, size_t &size) {
    size_t s = std::max<size_t>(len, 4);
    size_t p = 1;
    while (p < s) p <<= 1;
    size = p;
    cls = (size_t)__builtin_ctzll(p);
}

uint32_t WordPool::alloc(size_t len) {
    size_t cls, size;
    class_of(len, cls, size);
    used_words += size;
    if (!free_[cls].empty()) {
        uint32_t off = free_[cls].back();
        free_[cls].pop_back();
        return off;
    }
    uint32_t off = (uint32_t)data_.size();
    data_.resize(off + size, 0);
    return off;
}

void WordPool::free(uint32_t off, size_t len) {
    size_t cls, size;
    class_of(len, cls, size);
    used_words -= size;
    free_[cls].push_back(off);
}

void WordPool::write(uint32_t off, const uint32_t *words, size_t n) {
    std::copy(words, words + n, data_.begin() + off);
    dirty_.push_back({off, (uint32_t)n});
}

// ---- World (src/voxel/world.rs:63-629) ----
Inner World::make_root(uint64_t root_mask, const std::vector<Inner> &run) {
    if (run.empty()) return Inner{};
    if (run.size() == 64) {
        bool all_full = true;
        for (const auto &n : run)
            if (!n.is_full()) { all_full = false; break; }
        if (all_full) return Inner::full(0);
    }
    auto r = inners.intern(run.data(), (uint32_t)run.size());
    if (r.second)
        for (const auto &n : run)
            if (!n.is_full()) leaves.add_ref(n.run());
    return Inner{root_mask, r.first, 0};
}

void World::release_root(Inner root) {
    if (root.is_full() || root.is_empty()) return;
    std::vector<Inner> nodes(inners.get(root.run()), inners.get(root.run()) + root.count());
    if (inners.release(root.run()))
        for (const auto &n : nodes)
            if (!n.is_full()) leaves.release(n.run());
}

void World::insert_local(Key key, const LocalTree &t) {
    if (chunks.count(key)) remove_chunk(key);
    std::vector<Inner> run;
    run.reserve(t.l1.size());
    std::vector<uint32_t> temp;
    for (const auto &n : t.l1) {
        if (n.full) {
            run.push_back(Inner::full(n.leaf_prefix | ((uint32_t)n.water_line << 28)));
        } else {
            uint32_t cnt = (uint32_t)__builtin_popcountll(n.mask);
            auto r = leaves.intern(t.leaves.data() + n.leaf_start, cnt);
            temp.push_back(r.first);
            run.push_back(Inner{n.mask, r.first, n.leaf_prefix | ((uint32_t)n.water_line << 28)});
        }
    }
    Inner root = make_root(t.root_mask, run);
    for (uint32_t off : temp) leaves.release(off);

    ChunkRecord rec;
    rec.key = key;
    rec.root = root;
    if (t.has_uniform) {
        rec.attr_uniform = true;
        rec.attr_base = t.uniform;
    } else if (t.attrs.empty()) {
        rec.attr_uniform = true;
        rec.attr_base = AIR;
    } else {
        uint32_t pal_base = 0;
        if (!t.palette.empty()) {
            pal_base = bricks.alloc_write(t.palette.data(), t.palette.size());
            rec.regions.push_back({pal_base, t.palette.size()});
        }
        std::vector<uint32_t> table(t.attrs.size());
        for (size_t i = 0; i < t.attrs.size(); i++)
            table[i] = (t.attrs[i] & UNIFORM_BIT) ? t.attrs[i] : t.attrs[i] + pal_base;
        rec.attr_uniform = false;
        rec.attr_base = bricks.alloc_write(table.data(), table.size());
        rec.attr_len = (uint32_t)table.size();
    }
    for (int i = 0; i < 3; i++) {
        rec.aabb_min[i] = t.aabb_min[i];
        rec.aabb_max[i] = t.aabb_max[i];
    }
    rec.solid_count = t.solid_count;
    rec.leaf_count = t.leaf_count;
    rec.has_water = t.has_water;
    rec.dry_mask = t.dry_mask;
    rec.has_foliage = t.has_foliage;
    rec.has_cutout = t.has_cutout;
    rec.has_glass = t.has_glass;
    rec.has_emitter = t.has_emitter;
    rec.emitters = t.emitters;
    chunks[key] = std::move(rec);
    order.push_back(key);
    if (chunks[key].emitters != 0) emitter_chunks++;
    version++;
}

bool World::remove_chunk(Key key) {
    auto it = chunks.find(key);
    if (it == chunks.end()) return false;
    ChunkRecord rec = std::move(it->second);
    chunks.erase(it);
    order.erase(std::find(order.begin(), order.end(), key));
    if (rec.has_emitter) emitter_chunks--;
    release_root(rec.root);
    if (!rec.attr_uniform) bricks.free(rec.attr_base, rec.attr_len);
    for (const auto &r : rec.regions) bricks.free(r.first, r.second);
    version++;
    return true;
}

bool World::l1_node(Inner root, uint32_t l1b, Inner &out) const {
    if (!has(root.mask, l1b)) return false;
    if (root.is_full()) {
        out = Inner::full(l1b * 64);
        return true;
    }
    out = inners.get(root.run())[below(root.mask, l1b)];
    return true;
}

uint64_t World::leaf_mask(Inner l1, uint32_t lfb) const {
    if (!has(l1.mask, lfb)) return 0;
    if (l1.is_full()) return ~0ull;
    return leaves.get(l1.run())[below(l1.mask, lfb)];
}

BlockId World::leaf_block(const ChunkRecord &rec, uint32_t leaf_index, uint32_t vb) const {
    if (rec.attr_uniform) return (BlockId)rec.attr_base;
    const uint32_t *words = bricks.data().data();
    return decode_leaf(words, words[rec.attr_base + leaf_index], vb);
}

static Key key_of_block(int32_t x, int32_t y, int32_t z) { return Key{x >> 6, y >> 6, z >> 6}; }

BlockId World::get_block(int32_t x, int32_t y, int32_t z) const {
    if (y < 0 || y >= WORLD_HEIGHT) return AIR;
    auto it = chunks.find(key_of_block(x, y, z));
    if (it == chunks.end()) return AIR;
    const ChunkRecord &rec = it->second;
    uint32_t lx = (uint32_t)(x & 63), ly = (uint32_t)(y & 63), lz = (uint32_t)(z & 63);
    Inner l1;
    if (!l1_node(rec.root, l1_bit(lx, ly, lz), l1)) return AIR;
    uint32_t lfb = leaf_bit(lx, ly, lz);
    uint64_t m = leaf_mask(l1, lfb);
    uint32_t vb = voxel_bit(lx, ly, lz);
    if (!has(m, vb)) return AIR;
    return leaf_block(rec, (l1.leaf_prefix & PREFIX_MASK) + below(l1.mask, lfb), vb);
}

// set_block (src/voxel/world.rs:288-529)
bool World::set_block(int32_t x, int32_t y, int32_t z, BlockId id) {
    if (y < 0 || y >= WORLD_HEIGHT) return false;
    Key key = key_of_block(x, y, z);
    auto it = chunks.find(key);
    if (it == chunks.end()) return false;
    Inner old_root = it->second.root;
    bool attr_uniform = it->second.attr_uniform;
    uint32_t attr_base = it->second.attr_base, attr_len = it->second.attr_len;
    uint32_t leaf_count = it->second.leaf_count;
    uint32_t lx = (uint32_t)(x & 63), ly = (uint32_t)(y & 63), lz = (uint32_t)(z & 63);
    uint32_t l1b = l1_bit(lx, ly, lz), lfb = leaf_bit(lx, ly, lz), vb = voxel_bit(lx, ly, lz);

    Inner l1_nodes[64];
    if (old_root.is_full()) {
        for (uint32_t b = 0; b < 64; b++) l1_nodes[b] = Inner::full((b * 64) | 0xF0000000u);
    } else if (!old_root.is_empty()) {
        const Inner *run = inners.get(old_root.run());
        uint32_t k = 0;
        for (uint32_t b = 0; b < 64; b++)
            if (has(old_root.mask, b)) l1_nodes[b] = run[k++];
    }
    Inner old_l1_nodes[64];
    std::copy(l1_nodes, l1_nodes + 64, old_l1_nodes);
    Inner old_l1 = l1_nodes[l1b];

    uint64_t leafs[64] = {0};
    if (old_l1.is_full()) {
        for (auto &m : leafs) m = ~0ull;
    } else if (!old_l1.is_empty()) {
        const uint64_t *run = leaves.get(old_l1.run());
        uint32_t k = 0;
        for (uint32_t b = 0; b < 64; b++)
            if (has(old_l1.mask, b)) leafs[b] = run[k++];
    }
    uint64_t old_leaf = leafs[lfb];
    bool was_solid = has(old_leaf, vb);
    bool now_solid = id != AIR;
    if (!was_solid && !now_solid) return true;

    uint32_t base_prefix = 0;
    for (uint32_t b = 0; b < l1b; b++) base_prefix += l1_nodes[b].count();
    size_t leaf_pos = base_prefix + below(old_l1.mask, lfb);

    const ChunkRecord &rec0 = it->second;
    BlockId ids[64];
    std::fill(ids, ids + 64, AIR);
    if (old_leaf != 0)
        for (uint32_t b = 0; b < 64; b++)
            if (has(old_leaf, b)) ids[b] = leaf_block(rec0, (uint32_t)leaf_pos, b);
    if (was_solid && now_solid && ids[vb] == id) return true;

    uint64_t new_leaf = now_solid ? (old_leaf | (1ull << vb)) : (old_leaf & ~(1ull << vb));
    BlockId old_id = ids[vb];
    ids[vb] = id;

    std::vector<uint32_t> table;
    if (attr_uniform) table.assign(leaf_count, UNIFORM_BIT | attr_base);
    else table.assign(bricks.data().begin() + attr_base, bricks.data().begin() + attr_base + attr_len);
    std::vector<std::pair<uint32_t, size_t>> new_regions;
    if (new_leaf != 0) {
        std::vector<uint32_t> pw;
        uint32_t e = encode_leaf(ids, new_leaf, pw);
        uint32_t entry;
        if (e & UNIFORM_BIT) {
            entry = e;
        } else {
            uint32_t off = bricks.alloc_write(pw.data(), pw.size());
            new_regions.push_back({off, pw.size()});
            entry = off;
        }
        if (old_leaf != 0) table[leaf_pos] = entry;
        else table.insert(table.begin() + (ptrdiff_t)leaf_pos, entry);
    } else {
        table.erase(table.begin() + (ptrdiff_t)leaf_pos);
    }

    leafs[lfb] = new_leaf;
    uint64_t new_mask = 0;
    bool all_max = true;
    for (uint32_t b = 0; b < 64; b++) {
        if (leafs[b] != 0) new_mask |= 1ull << b;
        if (leafs[b] != ~0ull) all_max = false;
    }
    Inner new_l1;
    if (new_mask == 0) {
        new_l1 = Inner{};
    } else if (new_mask == ~0ull && all_max) {
        new_l1 = Inner::full(0xF0000000u);
    } else {
        std::vector<uint64_t> run;
        for (uint32_t b = 0; b < 64; b++)
            if (leafs[b] != 0) run.push_back(leafs[b]);
        auto r = leaves.intern(run.data(), (uint32_t)run.size());
        new_l1 = Inner{new_mask, r.first, 0};
    }
    bool has_temp = !new_l1.is_full() && !new_l1.is_empty();
    uint32_t temp_leaf = has_temp ? new_l1.run() : 0;
    l1_nodes[l1b] = new_l1;

    std::vector<Inner> run;
    run.reserve(64);
    uint32_t prefix = 0;
    uint64_t root_mask = 0;
    for (uint32_t b = 0; b < 64; b++) {
        if (l1_nodes[b].is_empty()) continue;
        Inner n = l1_nodes[b];
        n.leaf_prefix = prefix | (n.leaf_prefix & ~PREFIX_MASK);
        prefix += n.count();
        run.push_back(n);
        root_mask |= 1ull << b;
    }
    Inner new_root = make_root(root_mask, run);
    if (has_temp) leaves.release(temp_leaf);
    if (!old_root.is_full() && !old_root.is_empty() && inners.release(old_root.run()))
        for (uint32_t b = 0; b < 64; b++) {
            const Inner &n = old_l1_nodes[b];
            if (!n.is_empty() && !n.is_full()) leaves.release(n.run());
        }

    bool uniform_all;
    if (table.empty()) {
        uniform_all = true;
    } else {
        uint32_t f = table[0];
        uniform_all = (f & UNIFORM_BIT) != 0;
        for (size_t i = 1; uniform_all && i < table.size(); i++) uniform_all = table[i] == f;
    }
    bool new_uniform;
    uint32_t new_base, new_len = 0;
    if (uniform_all) {
        if (!attr_uniform) bricks.free(attr_base, attr_len);
        new_uniform = true;
        new_base = table.empty() ? AIR : (uint32_t)(BlockId)table[0];
    } else if (!attr_uniform && attr_len == table.size()) {
        bricks.write(attr_base, table.data(), table.size());
        new_uniform = false;
        new_base = attr_base;
        new_len = attr_len;
    } else if (!attr_uniform) {
        bricks.free(attr_base, attr_len);
        new_uniform = false;
        new_base = bricks.alloc_write(table.data(), table.size());
        new_len = (uint32_t)table.size();
    } else {
        new_uniform = false;
        new_base = bricks.alloc_write(table.data(), table.size());
        new_len = (uint32_t)table.size();
    }

    ChunkRecord &rec = chunks[key];
    rec.root = new_root;
    rec.attr_uniform = new_uniform;
    rec.attr_base = new_base;
    rec.attr_len = new_len;
    for (auto &r : new_regions) rec.regions.push_back(r);
    rec.leaf_count = prefix;
    if (now_solid && !was_solid) {
        rec.solid_count += 1;
        int32_t v[3] = {(int32_t)lx, (int32_t)ly, (int32_t)lz};
        if (rec.solid_count == 1) {
            for (int i = 0; i < 3; i++) { rec.aabb_min[i] = v[i]; rec.aabb_max[i] = v[i] + 1; }
        } else {
            for (int i = 0; i < 3; i++) {
                rec.aabb_min[i] = std::min(rec.aabb_min[i], v[i]);
                rec.aabb_max[i] = std::max(rec.aabb_max[i], v[i] + 1);
            }
        }
    } else if (!now_solid && was_solid) {
        rec.solid_count -= 1;
    }
    rec.has_water |= id == WATER;
    if (id != WATER && id != AIR) rec.dry_mask |= 1ull << l1b;
    rec.has_foliage |= is_foliage(id);
    rec.has_cutout |= is_cutout(id);
    rec.has_glass |= is_glass(id);
    bool now_e = is_emitter(id), was_e = is_emitter(old_id);
    if (!was_e && now_e) {
        rec.emitters += 1;
        if (rec.emitters == 1) {
            rec.has_emitter = true;
            emitter_chunks++;
        }
    } else if (was_e && !now_e) {
        rec.emitters -= 1;
        if (rec.emitters == 0) {
            rec.has_emitter = false;
            emitter_chunks--;
        }
    }
    rec.version++;
    version++;
    return true;
}

// extract_dense (src/voxel/world.rs:594-629)
bool World::extract_dense(Key key, BlockId *out) const {
    auto it = chunks.find(key);
    if (it == chunks.end()) return false;
    std::fill(out, out + VOL, AIR);
    const ChunkRecord &rec = it->second;
    if (rec.root.is_empty()) return true;
    for (uint32_t l1b = 0; l1b < 64; l1b++) {
        Inner l1;
        if (!l1_node(rec.root, l1b, l1)) continue;
        uint32_t l1x = l1b & 3, l1y = (l1b >> 2) & 3, l1z = (l1b >> 4) & 3;
        for (uint32_t lfb = 0; lfb < 64; lfb++) {
            uint64_t mask = leaf_mask(l1, lfb);
            if (mask == 0) continue;
            uint32_t leaf_index = (l1.leaf_prefix & PREFIX_MASK) + below(l1.mask, lfb);
            uint32_t lfx = lfb & 3, lfy = (lfb >> 2) & 3, lfz = (lfb >> 4) & 3;
            for (uint32_t vb = 0; vb < 64; vb++) {
                if (!has(mask, vb)) continue;
                uint32_t vx = vb & 3, vy = (vb >> 2) & 3, vz = (vb >> 4) & 3;
                out[dense_index(l1x * 16 + lfx * 4 + vx, l1y * 16 + lfy * 4 + vy, l1z * 16 + lfz * 4 + vz)] =
                    leaf_block(rec, leaf_index, vb);
            }
        }
    }
    return true;
}

}  // namespace vxh
