// Host check of the 16-bit g storage's row reference temperatures
// (lbm_ddf_amd/csrc/tref.hpp), built with g++ by tests/test_tref_rows.py.
//
// For each height LY and wall pair given on the command line, every row
// and y direction: the temperature a decoding cell uses for a pulled value
// (its source row, A/B pattern) and the one an encoding cell uses for a slot
// it writes (its destination row, in-place pattern) must equal, bit for bit,
// the one the owning row encodes or decodes with.  Prints the number of
// mismatches, and how many rows the binary32 recomputation from the own
// row's value (T_ref(iy) + c_y * dtref) would get wrong, to show that the
// check can fail.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "tref.hpp"

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    const double t_hot = std::atof(argv[1]), t_cold = std::atof(argv[2]);
    long bad = 0, naive_bad = 0, pairs = 0;
    for (int k = 3; k < argc; ++k) {
        const int ly = std::atoi(argv[k]);
        const double dtref = (t_hot - t_cold) / ly;
        const double tref0 = t_hot - 0.5 * dtref;
        for (int iy = 0; iy < ly; ++iy) {
            const float own = lbm::tref_row<float>(tref0, dtref, iy);
            for (int cy = -1; cy <= 1; cy += 2) {
                const int src = lbm::source_row(iy, cy), dst = lbm::dest_row(iy, cy);
                if (src >= 0 && src < ly) {
                    ++pairs;
                    const float dec = lbm::tref_row<float>(tref0, dtref, src);
                    const float enc = lbm::tref_row<float>(tref0, dtref, iy - cy);
                    if (std::memcmp(&dec, &enc, sizeof(float)) != 0) ++bad;
                    const float naive = own + static_cast<float>(cy) * static_cast<float>(dtref);
                    if (std::memcmp(&naive, &enc, sizeof(float)) != 0) ++naive_bad;
                }
                if (dst >= 0 && dst < ly) {
                    const float enc = lbm::tref_row<float>(tref0, dtref, dst);
                    const float dec = lbm::tref_row<float>(tref0, dtref, iy + cy);
                    if (std::memcmp(&dec, &enc, sizeof(float)) != 0) ++bad;
                }
            }
        }
    }
    std::printf("pairs %ld mismatches %ld naive_mismatches %ld\n", pairs, bad, naive_bad);
    return bad == 0 ? 0 : 1;
}
