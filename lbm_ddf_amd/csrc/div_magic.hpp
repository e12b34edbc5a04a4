// div_magic.hpp -- unsigned division of a cell index by a row length or a
// column height as one multiply and one shift (plain constexpr C++: the
// ensemble kernel and the host share it, and it is checked on the CPU).
//
// For 0 <= n < 2^31 and d >= 1, with l = ceil(log2 d),
// m = ceil(2^(31+l) / d) and sh = 31 + l:  floor(n / d) = floor(m * n / 2^sh).
// m < 2^32 because d > 2^(l-1) (or d = 1: m = 2^31).  This is Theorem 4.2 of
// T. Granlund and P. L. Montgomery, "Division by invariant integers using
// multiplication" (PLDI 1994) with N = 31: m * d - 2^(N+l) < d <= 2^l.
#pragma once

#include <cstdint>

namespace lbm {

// div_by is exact for n below this (N = 31 above).
constexpr uint64_t kDivLimit = 1ull << 31;

struct DivMagic {
    uint32_t m, sh;
};

constexpr DivMagic div_magic(uint32_t d) {
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    const uint64_t p = 1ull << (31 + l);
    return {static_cast<uint32_t>((p + d - 1) / d), 31 + l};
}

constexpr uint32_t div_by(uint32_t n, uint32_t m, uint32_t sh) {
    return static_cast<uint32_t>((static_cast<uint64_t>(m) * n) >> sh);
}

}  // namespace lbm
