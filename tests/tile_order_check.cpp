// Host check of the ensemble launch's tile order (lbm_ddf_amd/csrc/tile_order.hpp):
// for every tile count and grid size given, in both orders, every tile is
// taken by exactly one block and each block's tile numbers grow.
// usage: tile_order_check [plant]  (plant 1: one block too many counted on each XCD, which must fail)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tile_order.hpp"

int main(int argc, char **argv) {
    const int plant = argc > 1 ? atoi(argv[1]) : 0;
    std::vector<uint32_t> tiles, grids;
    for (uint32_t n = 1; n <= 400; ++n) tiles.push_back(n);
    for (uint32_t n : {1000u, 4097u, 155648u, 155649u}) tiles.push_back(n);
    for (uint32_t g : {1u, 3u, 7u, 8u, 9u, 15u, 16u, 17u, 255u, 256u, 512u, 2048u, 4096u}) grids.push_back(g);
    long cases = 0;
    for (uint32_t n : tiles) {
        std::vector<uint32_t> gs = grids;
        gs.push_back(n);  // one block per tile
        for (uint32_t g : gs) {
            for (int xcd = 0; xcd < 2; ++xcd) {
                std::vector<int> seen(n, 0);
                for (uint32_t b = 0; b < g; ++b) {
                    lbm::TileOrder o = lbm::tile_order(xcd != 0, b, g);
                    if (plant == 1 && o.xcd) o.step += 1;
                    uint64_t prev = 0;
                    bool first = true;
                    for (uint32_t kl = o.first;; kl += o.step) {
                        const uint64_t t = lbm::tile_of(o, kl);
                        if (!first && t <= prev) {
                            printf("FAIL tiles %u grid %u xcd %d block %u: tile numbers do not grow\n", n, g, xcd, b);
                            return 1;
                        }
                        if (t >= n) break;
                        seen[t] += 1;
                        prev = t;
                        first = false;
                    }
                }
                for (uint32_t t = 0; t < n; ++t)
                    if (seen[t] != 1) {
                        printf("FAIL tiles %u grid %u xcd %d: tile %u taken %d times\n", n, g, xcd, t, seen[t]);
                        return 1;
                    }
                ++cases;
            }
        }
    }
    printf("ok %ld\n", cases);
    return 0;
}
