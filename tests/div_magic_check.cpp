// Host check of the multiply-shift division (lbm_ddf_amd/csrc/div_magic.hpp)
// against the integer division: every n below 2^16 and the top 2^16 values
// below 2^31 for every divisor up to 2048; multiples of d and their
// neighbours up to 2^31; pseudo-random n for divisors up to 2^31 - 1.
// usage: div_magic_check [plant]  (plant 1: m rounded down, which must fail)
#include <cstdio>
#include <cstdlib>
#include <initializer_list>

#include "div_magic.hpp"

static int plant = 0;
static long cases = 0;

static bool check(uint32_t n, uint32_t d) {
    lbm::DivMagic k = lbm::div_magic(d);
    if (plant == 1) {
        const uint64_t p = 1ull << k.sh;
        k.m = static_cast<uint32_t>(p / d);
    }
    ++cases;
    const uint32_t q = lbm::div_by(n, k.m, k.sh);
    if (q != n / d) {
        printf("FAIL n %u d %u: %u instead of %u (m %u sh %u)\n", n, d, q, n / d, k.m, k.sh);
        return false;
    }
    return true;
}

int main(int argc, char **argv) {
    plant = argc > 1 ? atoi(argv[1]) : 0;
    const uint32_t top = 1u << 31;
    for (uint32_t d = 1; d <= 2048; ++d) {
        for (uint32_t n = 0; n < (1u << 16); ++n)
            if (!check(n, d) || !check(top - 1 - n, d)) return 1;
    }
    uint64_t x = 0x9e3779b97f4a7c15ull;
    auto next = [&x]() {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return x;
    };
    for (uint32_t d : {2049u, 3000u, 4095u, 4096u, 4097u, 10007u, 32767u, 32768u, 32769u, 65535u, 65536u,
                       1000003u, (1u << 24) + 1u, (1u << 30) - 1u, 1u << 30, top - 1u}) {
        for (uint64_t k = top / d; k > 0 && k + 2000 > top / d; --k) {
            const uint64_t n = k * d;
            for (int dn = -1; dn <= 1; ++dn)
                if (n + dn < top && !check(static_cast<uint32_t>(n + dn), d)) return 1;
        }
        for (int i = 0; i < 200000; ++i)
            if (!check(static_cast<uint32_t>(next() % top), d)) return 1;
    }
    for (int i = 0; i < 2000000; ++i) {
        const uint32_t d = 1 + static_cast<uint32_t>(next() % (top - 1));
        if (!check(static_cast<uint32_t>(next() % top), d)) return 1;
    }
    printf("ok %ld cases\n", cases);
    return 0;
}
