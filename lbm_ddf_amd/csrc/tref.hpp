// tref.hpp -- reference temperature of a row for the 16-bit g storage.
//
// 16-bit storage keeps g_i - w_i * T_ref(row) instead of g_i, with
// T_ref(row) = tref0 - row * dtref the conductive profile between isothermal
// walls half a cell outside the end rows.  A stored value is often decoded
// by a cell of another row than the one that encoded it: in the pull scheme
// a cell decodes what its source row stored, and in the in-place pattern a
// cell writes into its neighbours' slots.  Both sides must use
// bit-identical temperatures for a row, so every encode and decode of every
// kernel, the initial state and the host-copy codec take T_ref from
// tref_row() with the row index of the cell that wrote the stored value
// (source_row and writer_row below).  Recomputing a neighbour row's value from
// the own row's binary32 value (T_ref(iy) + c_y * dtref in binary32) would
// disagree by an ulp for about a quarter of the rows unless LY is a power of
// two, and inject a fixed per-row heat bias.
//
// Plain constexpr C++: HIP treats constexpr functions as host and device
// functions, and a host test compiles this header with g++.
#pragma once

namespace lbm {

// T_ref(row) evaluated in binary64 and rounded once to T.  The product is a
// statement of its own so that no build contracts it into a fused
// multiply-add (the 16-bit kernels allow contraction within an expression).
template <class T>
constexpr T tref_row(double tref0, double dtref, int row) {
    const double drop = static_cast<double>(row) * dtref;
    return static_cast<T>(tref0 - drop);
}

// Row of the cell that stored the value a cell of row iy pulls along a
// direction with y component cy (the source cell x - c).
constexpr int source_row(int iy, int cy) { return iy - cy; }

// Row of the cell that will read a slot written along a direction with y
// component cy (the destination cell x + c).
constexpr int dest_row(int iy, int cy) { return iy + cy; }

// In-place pattern, layout N (the reference's post-streaming layout): the row
// of the cell that wrote slot (i, x) of cell x = (ix, iy, iz), i of velocity
// c = (cx, cy, cz).  The slot holds f*_i(x - c) streamed from the source cell
// when that lies inside the domain, else the wall value that x itself wrote.
// 16-bit in-place storage encodes every value against its writer's row, so
// the stored deviations are as small as the A/B scheme's.
constexpr int writer_row(int ix, int iy, int iz, int lx, int ly, int lz, int cx, int cy, int cz) {
    const bool inside = ix - cx >= 0 && ix - cx < lx && iy - cy >= 0 && iy - cy < ly && iz - cz >= 0 && iz - cz < lz;
    return inside ? iy - cy : iy;
}

}  // namespace lbm
