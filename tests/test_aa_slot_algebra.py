"""The slot algebra of the in-place (AA) pattern, checked on the CPU.

aa_kernel.hpp updates one buffer in place.  An even pass reads each cell's
own slots (i, x) and writes slot (opp(j), x); an odd pass reads slot
(opp(i), x - c_i) (or (i, x) when x - c_i is outside) and writes slot
(j, x + c_j) (or (opp(j), x) when x + c_j is outside).  Wall rules are
applied by the writer.  This test restates those index maps on small odd
shapes and checks, with a toy per-cell collision standing in for BGK:

  * race freedom: in both passes each cell reads and writes the same slot
    set, and the sets of different cells partition all 27*N slots;
  * equivalence: after any number of passes (and the pairwise exchange that
    completes a pending streaming) the populations equal those of the pull
    A/B scheme with the reference's wall rules -- half-way bounce-back for f,
    anti-bounce-back with (1 + rho') at the y walls for g unless the source
    is also beyond a z wall (src/streaming.cu:3-251).
"""
import math

import numpy as np
import pytest

import oracle

CX, CY, CZ, W = (list(map(float, a)) if k == 3 else list(map(int, a)) for k, a in enumerate(oracle.tables()))
Q = 27
OPP = [next(j for j in range(Q) if (CX[j], CY[j], CZ[j]) == (-CX[i], -CY[i], -CZ[i])) for i in range(Q)]
T_HOT, T_COLD = 0.5, -0.5


class Grid:
    def __init__(self, lx, ly, lz):
        self.lx, self.ly, self.lz, self.n = lx, ly, lz, lx * ly * lz

    def coords(self, idx):
        r = idx // self.lx
        return idx - r * self.lx, r % self.ly, r // self.ly

    def inside(self, x, y, z):
        return 0 <= x < self.lx and 0 <= y < self.ly and 0 <= z < self.lz

    def shift(self, idx, i, sign):
        return idx + sign * (CX[i] + CY[i] * self.lx + CZ[i] * self.lx * self.ly)


def collide(f, g):
    """Toy per-cell map standing in for BGK: nonlinear, mixes directions."""
    rho, phi = sum(f), sum(g)
    fs = [0.9 * f[i] + 0.1 * W[i] * rho + 1e-3 * math.sin(f[(i + 1) % Q]) for i in range(Q)]
    gs = [0.8 * g[i] + 0.2 * W[i] * phi + 1e-3 * math.cos(g[(i + 5) % Q]) for i in range(Q)]
    return fs, gs, rho


def wall_g(G, k, v, iy, iz, rw):
    """Value of g_k at a cell whose source x - c_k is outside, from the
    cell's own post-collision v = g*_opp(k)."""
    sy, sz = iy - CY[k], iz - CZ[k]
    if CY[k] != 0 and not (0 <= sy < G.ly) and 0 <= sz < G.lz:
        tw = T_HOT if iy == 0 else T_COLD
        return -v + 2.0 * W[k] * tw * (1.0 + rw)
    return v


def pull_ab(G, f0, g0, rho0, passes):
    fs = [[0.0] * G.n for _ in range(Q)]
    gs = [[0.0] * G.n for _ in range(Q)]
    rw = list(rho0)
    for x in range(G.n):
        a, b, _ = collide([f0[i][x] for i in range(Q)], [g0[i][x] for i in range(Q)])
        for i in range(Q):
            fs[i][x], gs[i][x] = a[i], b[i]

    def gather(fs, gs, rw):
        F = [[0.0] * G.n for _ in range(Q)]
        Gg = [[0.0] * G.n for _ in range(Q)]
        for x in range(G.n):
            ix, iy, iz = G.coords(x)
            for i in range(Q):
                if G.inside(ix - CX[i], iy - CY[i], iz - CZ[i]):
                    s = G.shift(x, i, -1)
                    F[i][x], Gg[i][x] = fs[i][s], gs[i][s]
                else:
                    F[i][x] = fs[OPP[i]][x]
                    Gg[i][x] = wall_g(G, i, gs[OPP[i]][x], iy, iz, rw[x])
        return F, Gg

    for _ in range(passes):
        F, Gg = gather(fs, gs, rw)
        nrw = [0.0] * G.n
        for x in range(G.n):
            a, b, r = collide([F[i][x] for i in range(Q)], [Gg[i][x] for i in range(Q)])
            for i in range(Q):
                fs[i][x], gs[i][x] = a[i], b[i]
            nrw[x] = r
        rw = nrw
    return gather(fs, gs, rw)


def aa_slots(G, x, odd):
    """(read slots, write slots) of cell x in a pass, as (population, cell)."""
    ix, iy, iz = G.coords(x)
    rd, wr = [], []
    for i in range(Q):
        if odd and i != 0 and G.inside(ix - CX[i], iy - CY[i], iz - CZ[i]):
            rd.append((OPP[i], G.shift(x, i, -1)))
        elif odd and i != 0:
            rd.append((i, x))
        else:
            rd.append((i, x))
        if odd and i != 0 and G.inside(ix + CX[i], iy + CY[i], iz + CZ[i]):
            wr.append((i, G.shift(x, i, +1)))
        else:
            wr.append((OPP[i], x))
    return rd, wr


def in_place(G, f0, g0, rho0, passes):
    A = {(i, x): f0[i][x] for i in range(Q) for x in range(G.n)}
    B = {(i, x): g0[i][x] for i in range(Q) for x in range(G.n)}

    def write(x, fs, gs, rw, odd):
        ix, iy, iz = G.coords(x)
        _, wr = aa_slots(G, x, odd)
        for j, slot in enumerate(wr):
            inside = G.inside(ix + CX[j], iy + CY[j], iz + CZ[j])
            A[slot] = fs[j]
            B[slot] = gs[j] if inside else wall_g(G, OPP[j], gs[j], iy, iz, rw)

    # prologue: collide the natural layout, write the exchanged layout
    for x in range(G.n):
        rd, _ = aa_slots(G, x, False)
        fs, gs, _ = collide([A[s] for s in rd], [B[s] for s in rd])
        write(x, fs, gs, rho0[x], False)
    odd = True
    for _ in range(passes):
        res = []
        for x in range(G.n):  # every read of a pass precedes its writes (disjoint slot sets)
            rd, _ = aa_slots(G, x, odd)
            res.append(collide([A[s] for s in rd], [B[s] for s in rd]))
        for x in range(G.n):
            fs, gs, rho = res[x]
            write(x, fs, gs, rho, odd)
        odd = not odd
    if odd:  # layout E pending: the pairwise exchange completes the streaming
        for x in range(G.n):
            ix, iy, iz = G.coords(x)
            for i in range(Q):
                if i < OPP[i] and G.inside(ix + CX[i], iy + CY[i], iz + CZ[i]):
                    a, b = (OPP[i], x), (i, G.shift(x, i, +1))
                    A[a], A[b] = A[b], A[a]
                    B[a], B[b] = B[b], B[a]
    return ([[A[(i, x)] for x in range(G.n)] for i in range(Q)],
            [[B[(i, x)] for x in range(G.n)] for i in range(Q)])


SHAPES = [(4, 5, 3), (1, 4, 2), (3, 3, 1), (5, 3, 4), (2, 6, 2)]


@pytest.mark.parametrize("shape", SHAPES)
def test_passes_are_race_free_in_place(shape):
    G = Grid(*shape)
    for odd in (False, True):
        owner = {}
        for x in range(G.n):
            rd, wr = aa_slots(G, x, odd)
            assert sorted(rd) == sorted(wr), (odd, x)
            for s in wr:
                assert 0 <= s[1] < G.n
                assert s not in owner, (odd, s, x, owner.get(s))
                owner[s] = x
        assert len(owner) == Q * G.n


@pytest.mark.parametrize("passes", [0, 1, 2, 3])
@pytest.mark.parametrize("shape", SHAPES)
def test_in_place_equals_pull_streaming(shape, passes):
    G = Grid(*shape)
    rng = np.random.default_rng(7)
    f0 = (rng.standard_normal((Q, G.n)) * 0.01).tolist()
    g0 = (rng.standard_normal((Q, G.n)) * 0.1 + 0.2).tolist()
    rho0 = [sum(f0[i][x] for i in range(Q)) for x in range(G.n)]
    fa, ga = in_place(G, f0, g0, rho0, passes)
    fb, gb = pull_ab(G, f0, g0, rho0, passes)
    assert fa == fb
    assert ga == gb


def test_a_misplaced_wall_value_is_caught():
    """The equivalence check is sensitive: dropping the anti-bounce-back of
    the writer (plain bounce-back everywhere) changes the result."""
    G = Grid(4, 5, 3)
    rng = np.random.default_rng(7)
    f0 = (rng.standard_normal((Q, G.n)) * 0.01).tolist()
    g0 = (rng.standard_normal((Q, G.n)) * 0.1 + 0.2).tolist()
    rho0 = [sum(f0[i][x] for i in range(Q)) for x in range(G.n)]
    global wall_g
    keep = wall_g
    try:
        fb, gb = pull_ab(G, f0, g0, rho0, 1)
        wall_g = lambda G_, k, v, iy, iz, rw: v  # noqa: E731
        fa, ga = in_place(G, f0, g0, rho0, 1)
    finally:
        wall_g = keep
    assert fa == fb and ga != gb
