// lattice.hpp -- D3Q27 velocity set, weights and compile-time helpers.
//
// The velocity ordering and weights are the reference's
// (src/initial.cu:71-83); the opposite-direction table is derived from them
// and checked at compile time below.  Everything here is constexpr so that
// fully unrolled kernels fold every table access into immediates.
#pragma once

#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <utility>

namespace lbm {

constexpr int Q = 27;

// src/initial.cu:71-73
constexpr int CX[Q] = {0, 1, 0, -1, 0, 0, 0, 1, -1, -1, 1, 1, 0, -1, 0, 1, 0, -1, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int CY[Q] = {0, 0, 1, 0, -1, 0, 0, 1, 1, -1, -1, 0, 1, 0, -1, 0, 1, 0, -1, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1};

// Opposite direction: c_opp(i) = -c_i.
constexpr int OPP[Q] = {0, 3, 4, 1, 2, 6, 5, 9, 10, 7, 8, 17, 18, 15, 16, 13, 14, 11, 12, 25, 26, 23, 24, 21, 22, 19, 20};

// src/initial.cu:77-83, evaluated in binary64 exactly as the reference does.
constexpr double weight(int i) {
    return i == 0 ? 8.0 / 27.0 : (i <= 6 ? 2.0 / 27.0 : (i <= 18 ? 1.0 / 54.0 : 1.0 / 216.0));
}

constexpr bool opp_table_ok() {
    for (int i = 0; i < Q; ++i) {
        int o = OPP[i];
        if (CX[o] != -CX[i] || CY[o] != -CY[i] || CZ[o] != -CZ[i]) return false;
        if (OPP[o] != i) return false;
    }
    return true;
}
static_assert(opp_table_ok(), "OPP must map every direction to its negation");

// A direction is "canonical" when it is the lower index of its opposite pair;
// the equilibrium projection c_i.u of the other member is the exact negation.
constexpr bool is_canonical(int i) { return i != 0 && i < OPP[i]; }

// Compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1.
template <typename F, int... Is>
__host__ __device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__host__ __device__ __forceinline__ void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// c_i . u summed over the non-zero components in the order x, y, z.
// This equals the reference's (cix*u + ciy*v + ciz*w) (src/collision.cu:25)
// bit for bit: multiplying by +-1 is exact and adding a signed zero does not
// change a sum.
template <int I, typename T>
__host__ __device__ __forceinline__ T cdotu(T u, T v, T w) {
    constexpr int cx = CX[I], cy = CY[I], cz = CZ[I];
    T s = T(0);
    bool started = false;
    if constexpr (cx != 0) {
        s = cx > 0 ? u : -u;
        started = true;
    }
    if constexpr (cy != 0) {
        if (started) s = cy > 0 ? s + v : s - v;
        else s = cy > 0 ? v : -v;
        started = true;
    }
    if constexpr (cz != 0) {
        if (started) s = cz > 0 ? s + w : s - w;
        else s = cz > 0 ? w : -w;
        started = true;
    }
    (void)started;
    return s;
}

}  // namespace lbm
