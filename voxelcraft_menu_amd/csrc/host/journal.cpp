// journal.cpp -- edit journal files (VXJ1): the byte format of src/journal.rs, read into
// edit records for vxw_apply_edits and written back as EditJournal::save writes it.
//
// File layout (all little-endian, src/journal.rs:13-139):
//   Header   magic "VXJ1", version, seed, sea_level, count, flags      24 bytes
//   Player   pos[3], yaw, pitch, hotbar, fly, pad     (flags & 1)      32 bytes
//   Bar      10 block ids, 2 pad u16                  (flags & 4)      24 bytes
//   Records  x, y, z, id, pad                         count times      16 bytes each
// flags & 2 (APPEND) marks a file an append log grew: records past `count` are whole
// edits the header had not counted yet, and a torn trailing record is dropped.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "vxhost.h"

namespace {

constexpr uint32_t MAGIC = 0x314A5856u;  // "VXJ1" as a little-endian u32
constexpr uint32_t VERSION = 3;
constexpr uint32_t FLAG_PLAYER = 1, FLAG_APPEND = 2, FLAG_BAR = 4;
constexpr size_t HEAD = 24, PLAYER = 32, BAR = 24, REC = 16;
constexpr uint32_t SLOTS = 10;  // HOTBAR.len() (src/block.rs:222-226)
// block ids of bar_slot_refusal (src/block.rs:3-35, :230-240)
constexpr uint16_t AIR = 0, WATER = 13, TALL_GRASS = 18, GRASS_TALL = 24, REEDS = 25, BLOCK_COUNT = 26;

int fail(char *err, size_t cap, const char *fmt, ...) {
    if (err && cap) {
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(err, cap, fmt, ap);
        va_end(ap);
    }
    return VXW_E_JOURNAL;
}

uint32_t rd32(const uint8_t *p) {
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}

const char *slot_refusal(uint16_t id) {
    if (id == AIR) return "air is not a block you can place";
    if (id == WATER) return "water would let a run with no sea in it acquire a water pixel";
    if (id == TALL_GRASS || id == GRASS_TALL || id == REEDS)
        return "a placed tuft is invisible in a build compiled without foliage";
    if (id >= BLOCK_COUNT) return "no block has that id";
    return nullptr;
}

size_t extras_bytes(uint32_t flags) {
    return ((flags & FLAG_PLAYER) ? PLAYER : 0) + ((flags & FLAG_BAR) ? BAR : 0);
}

}  // namespace

extern "C" {

// EditJournal::load (src/journal.rs:391-475), with the same checks in the same order.
int vxw_journal_parse(const void *bytes, size_t n, int32_t seed, int32_t sea_level, vxw_journal_info *info,
                      vx_edit *edits, size_t cap, char *err, size_t err_cap) {
    if ((!bytes && n) || !info) return VXW_E_ARG;
    std::memset(info, 0, sizeof *info);
    const uint8_t *b = (const uint8_t *)bytes;
    if (n < HEAD) return fail(err, err_cap, "too short to be an edit journal");
    const uint32_t magic = rd32(b), version = rd32(b + 4), count = rd32(b + 16), flags = rd32(b + 20);
    int32_t h_seed, h_sea;
    std::memcpy(&h_seed, b + 8, 4);
    std::memcpy(&h_sea, b + 12, 4);
    if (magic != MAGIC) return fail(err, err_cap, "not an edit journal");
    if (version == 0 || version > VERSION)
        return fail(err, err_cap, "journal version %u, this build reads 1..=%u", version, VERSION);
    if (h_seed != seed || h_sea != sea_level)
        return fail(err, err_cap,
                    "journal is for seed %d sea-level %d, this world is seed %d sea-level %d -- the edits "
                    "would land on terrain that is not there",
                    h_seed, h_sea, seed, sea_level);
    const size_t extras = extras_bytes(flags);
    if (n - HEAD < extras)
        return fail(err, err_cap, "the header claims a player or bar block the file is too short to hold");
    info->version = version;
    info->flags = flags;
    info->seed = h_seed;
    info->sea_level = h_sea;
    const uint8_t *x = b + HEAD;
    if (flags & FLAG_PLAYER) {
        std::memcpy(&info->player, x, PLAYER);
        info->has_player = 1;
        x += PLAYER;
    }
    if (flags & FLAG_BAR) {
        std::memcpy(info->bar, x, SLOTS * 2);
        info->has_bar = 1;
    }
    if (info->has_player) {  // PlayerState::check (src/journal.rs:60-80)
        const vxw_player &p = info->player;
        if (!std::isfinite(p.pos[0]) || !std::isfinite(p.pos[1]) || !std::isfinite(p.pos[2]) ||
            !std::isfinite(p.yaw) || !std::isfinite(p.pitch))
            return fail(err, err_cap, "the journal's player position or look angle is not a finite number");
        if (p.hotbar >= SLOTS)
            return fail(err, err_cap, "journal selects hotbar slot %u, this build has %u", p.hotbar, SLOTS);
    }
    if (info->has_bar) {  // BarState::check (src/journal.rs:92-104)
        for (uint32_t i = 0; i < SLOTS; i++)
            if (const char *why = slot_refusal(info->bar[i]))
                return fail(err, err_cap, "journal puts block %u in bar slot %u, and %s", info->bar[i], i, why);
    }
    const size_t rest = n - HEAD - extras;
    const size_t want = (size_t)count * REC;
    const bool appended = (flags & FLAG_APPEND) != 0;
    if (!(rest == want || (appended && rest > want)))
        return fail(err, err_cap, "header says %u edits (%zu bytes), file carries %zu", count, want, rest);
    const size_t whole = rest / REC;
    info->count = (uint32_t)whole;
    info->recovered = (uint32_t)(whole - count);
    info->torn_bytes = (uint32_t)(rest - whole * REC);
    if (edits) {
        if (cap < whole) return VXW_E_ARG;
        std::memcpy(edits, b + HEAD + extras, whole * REC);
    }
    return VXW_OK;
}

// EditJournal::save (src/journal.rs:477-508): the last edit of each voxel, sorted by
// (x, z, y), after the header (count, PLAYER / BAR flags) and the player and bar blocks.
int vxw_journal_write(const vx_edit *edits, size_t n, int32_t seed, int32_t sea_level, const vxw_player *player,
                      const uint16_t *bar, void *out, size_t cap, size_t *written) {
    if ((!edits && n) || !written) return VXW_E_ARG;
    std::map<std::tuple<int32_t, int32_t, int32_t>, uint16_t> last;
    for (size_t i = 0; i < n; i++) last[{edits[i].x, edits[i].z, edits[i].y}] = edits[i].id;
    const uint32_t flags = (player ? FLAG_PLAYER : 0u) | (bar ? FLAG_BAR : 0u);
    const size_t total = HEAD + extras_bytes(flags) + last.size() * REC;
    *written = total;
    if (!out) return VXW_OK;
    if (cap < total) return VXW_E_ARG;
    uint8_t *o = (uint8_t *)out;
    const uint32_t head[6] = {MAGIC, VERSION, (uint32_t)seed, (uint32_t)sea_level, (uint32_t)last.size(), flags};
    std::memcpy(o, head, HEAD);
    o += HEAD;
    if (player) {
        vxw_player p = *player;
        p.pad = 0;
        std::memcpy(o, &p, PLAYER);
        o += PLAYER;
    }
    if (bar) {
        uint16_t b[SLOTS + 2] = {};
        std::memcpy(b, bar, SLOTS * 2);
        std::memcpy(o, b, BAR);
        o += BAR;
    }
    for (const auto &kv : last) {  // map order is (x, z, y)
        vx_edit r{std::get<0>(kv.first), std::get<2>(kv.first), std::get<1>(kv.first), kv.second, 0};
        std::memcpy(o, &r, REC);
        o += REC;
    }
    return VXW_OK;
}

}  // extern "C"
