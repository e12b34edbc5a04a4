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

// The in-place odd pass writes f*_j(x) into slot (j, x + c_j) when x + c_j is
// inside the domain, else the wall value into slot (opp(j), x) (aa_kernel.hpp).
// Replays those writes on a small grid, records the row of each slot's writer
// and compares it with writer_row(); returns the number of mismatches.
static long check_writer_rows(int lx, int ly, int lz) {
    static const int CXv[27] = {0, 1, 0, -1, 0, 0, 0, 1, -1, -1, 1, 1, 0, -1, 0, 1, 0, -1, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    static const int CYv[27] = {0, 0, 1, 0, -1, 0, 0, 1, 1, -1, -1, 0, 1, 0, -1, 0, 1, 0, -1, 1, 1, -1, -1, 1, 1, -1, -1};
    static const int CZv[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
    int opp[27];
    for (int i = 0; i < 27; ++i)
        for (int k = 0; k < 27; ++k)
            if (CXv[k] == -CXv[i] && CYv[k] == -CYv[i] && CZv[k] == -CZv[i]) opp[i] = k;
    const long n = static_cast<long>(lx) * ly * lz;
    long *writer = new long[27 * n];
    for (long k = 0; k < 27 * n; ++k) writer[k] = -1;
    long bad = 0;
    for (int z = 0; z < lz; ++z)
        for (int y = 0; y < ly; ++y)
            for (int x = 0; x < lx; ++x)
                for (int j = 0; j < 27; ++j) {
                    const int dx = x + CXv[j], dy = y + CYv[j], dz = z + CZv[j];
                    const bool inside = dx >= 0 && dx < lx && dy >= 0 && dy < ly && dz >= 0 && dz < lz;
                    const long slot = inside ? j * n + (static_cast<long>(dz) * ly + dy) * lx + dx
                                             : opp[j] * n + (static_cast<long>(z) * ly + y) * lx + x;
                    if (writer[slot] != -1) ++bad;  // written twice
                    writer[slot] = y;
                }
    for (int i = 0; i < 27; ++i)
        for (int z = 0; z < lz; ++z)
            for (int y = 0; y < ly; ++y)
                for (int x = 0; x < lx; ++x) {
                    const long slot = i * n + (static_cast<long>(z) * ly + y) * lx + x;
                    if (writer[slot] != lbm::writer_row(x, y, z, lx, ly, lz, CXv[i], CYv[i], CZv[i])) ++bad;
                }
    delete[] writer;
    return bad;
}

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    const long wr_bad = check_writer_rows(5, 4, 3) + check_writer_rows(1, 3, 2) + check_writer_rows(4, 5, 1);
    std::printf("writer_rows mismatches %ld\n", wr_bad);
    if (wr_bad) return 1;
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
