// tile_order.hpp -- the order in which the blocks of an ensemble launch take
// the tiles of the flat tile space (ensemble_kernel.hpp).
//
// Plain: block b takes tiles b, b + G, b + 2G, ... (G blocks).
// XCD-aware: the dispatcher places block b on XCD b % 8.  The tiles are dealt
// to the XCDs in chunks of kXcdChunk neighbouring tiles, chunk c to XCD
// c % 8, and the blocks of XCD x take that XCD's tiles in order: its local
// indices kl = b / 8, b / 8 + g_x, ... (g_x blocks on XCD x) map to tiles
// (kl / kXcdChunk * 8 + x) * kXcdChunk + kl % kXcdChunk.  Neighbouring tiles,
// which share the cache lines straddled by shifted loads, then run on one
// XCD's L2 (all but one pair in kXcdChunk), and every XCD gets a share of
// every part of the tile space, so cavities of different cost spread evenly
// over the XCDs.  A grid of fewer than 8 blocks uses the plain order (every
// XCD index needs a block).  Either way every tile is taken exactly once, and
// a block's tile numbers grow, so its loop ends at the first tile beyond the
// table (checked on the CPU by tests/test_tile_order.py).
//
// Plain constexpr C++: host and device.
#pragma once

#include <cstdint>

namespace lbm {

constexpr uint32_t kXcds = 8;
constexpr uint32_t kXcdChunk = 16;

struct TileOrder {
    bool xcd;          // XCD-aware order in use
    uint32_t x;        // the block's XCD index (b % 8)
    uint32_t first;    // the block's first local index
    uint32_t step;     // distance between the block's local indices
};

constexpr TileOrder tile_order(bool xcd_flag, uint32_t block, uint32_t grid) {
    TileOrder o{};
    o.xcd = xcd_flag && grid >= kXcds;
    o.x = block % kXcds;
    const uint32_t gx = grid / kXcds + (o.x < grid % kXcds ? 1u : 0u);  // blocks on XCD x
    o.first = o.xcd ? block / kXcds : block;
    o.step = o.xcd ? gx : grid;
    return o;
}

// Tile number of local index kl.
constexpr uint64_t tile_of(const TileOrder &o, uint32_t kl) {
    return o.xcd ? (static_cast<uint64_t>(kl / kXcdChunk) * kXcds + o.x) * kXcdChunk + kl % kXcdChunk
                 : static_cast<uint64_t>(kl);
}

}  // namespace lbm
