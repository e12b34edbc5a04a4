"""Pins of the FP64 oracle (CPU only).

The reference ships no tests, fixtures or golden vectors, so the oracle is
pinned by closed forms, invariants, an independent implementation and
textbook solutions, each chosen so that a plausible transcription mistake
(wrong weight, sign, index, wall set or override order) fails one of them:

  lattice         weights/velocities: isotropy up to 4th order, opp table
  derive          src/globals.cu:11-23 against values computed by hand
  collision_BGK   mass and momentum change in closed form, incl. the force
                  term's (1 - 1/(2 tau)) F (src/collision.cu:22-32)
  collision g     conserves phi (src/collision.cu:74-81)
  streaming       permutation of values; equals an independent push-form
                  implementation derived from the velocity vectors, with the
                  reference's x, y, z wall order for g
  g subsystem     conduction between anti-bounce-back walls: the steady
                  profile puts the walls half a cell outside the end nodes
  f subsystem     hydrostatic balance under a uniform force: u -> 0 and
                  d rho'/dy = 3 F
  full step       rest state is a fixed point; mass is conserved; x and z
                  mirror symmetry; the x <-> z swap symmetry is broken at the
                  wall edges by the wall order; conduction below onset gives
                  Nu_bulk -> 1
  output_nu       Nu formulas on fields with known answers
"""
import itertools
import math

import numpy as np
import pytest

import oracle

CX, CY, CZ, W = oracle.tables()
C = np.stack([CX, CY, CZ], axis=1)
OPP = [int(np.where((C == -C[i]).all(axis=1))[0][0]) for i in range(27)]


# ----------------------------------------------------------------- lattice
def test_lattice_isotropy():
    """Exact integer moments of 216*w: sum 216, first and third moments 0,
    second 72*delta (w: 1/3), fourth 24*(dd+dd+dd) (w: 1/9)."""
    Wi = np.round(W * 216).astype(np.int64)
    assert np.array_equal(W, Wi / 216.0)  # 8/27, 2/27, 1/54, 1/216 correctly rounded
    assert sorted(set(Wi)) == [1, 4, 16, 64] and Wi.sum() == 216
    assert abs(W.sum() - 1.0) <= 2 * np.finfo(float).eps  # not exactly 1.0 in binary64
    delta = lambda p, q: int(p == q)
    for a in range(3):
        assert (Wi * C[:, a]).sum() == 0
        for b in range(3):
            assert (Wi * C[:, a] * C[:, b]).sum() == 72 * delta(a, b)
            for c in range(3):
                assert (Wi * C[:, a] * C[:, b] * C[:, c]).sum() == 0
                for d in range(3):
                    want = 24 * (delta(a, b) * delta(c, d) + delta(a, c) * delta(b, d) + delta(a, d) * delta(b, c))
                    assert (Wi * C[:, a] * C[:, b] * C[:, c] * C[:, d]).sum() == want


def test_opposites_match_survey_table():
    assert OPP == [0, 3, 4, 1, 2, 6, 5, 9, 10, 7, 8, 17, 18, 15, 16, 13, 14, 11, 12, 25, 26, 23, 24, 21, 22, 19, 20]


# ----------------------------------------------------------------- derive
@pytest.mark.parametrize("L,ra,tau,tauc", [(64, 1e5, 0.57235, 0.60190), (512, 1e7, 0.55788, 0.58152),
                                           (512, 1e8, 0.51830, 0.52578)])
def test_derived_parameters(L, ra, tau, tauc):
    p = oracle.make_params(L, L, L, rayleigh=ra, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
    # g*beta*dT*LY = 0.02 with beta = 0.1/(5 LY): nu = LY sqrt(0.02 Pr / Ra)
    nu = L * math.sqrt(0.02 * 0.71 / ra)
    assert p.visc == pytest.approx(nu, rel=1e-14)
    assert p.tau == pytest.approx(3 * nu + 0.5, rel=1e-14)
    assert p.tauc == pytest.approx(3 * nu / 0.71 + 0.5, rel=1e-14)
    assert round(p.tau, 5) == tau and round(p.tauc, 5) == tauc
    assert p.u0 == pytest.approx(math.sqrt(0.02), rel=1e-14)
    assert p.t0 == 0.0


# ----------------------------------------------------------------- collision
def _single_cells(n=64, seed=0, **kw):
    """A 4x4x4 grid used as 64 independent cells for local operators."""
    s = oracle.OracleState(4, 4, 4, rayleigh=1e6, prandtl=0.71, t_hot=0.5, t_cold=-0.5, **kw)
    rng = np.random.default_rng(seed)
    s["f"] = rng.uniform(-1e-2, 1e-2, s["f"].shape)
    s["g"] = rng.uniform(0, 1e-1, s["g"].shape)
    s["ux"] = rng.uniform(-0.05, 0.05, s.shape)
    s["uy"] = rng.uniform(-0.05, 0.05, s.shape)
    s["uz"] = rng.uniform(-0.05, 0.05, s.shape)
    s["phi"] = rng.uniform(-0.5, 0.5, s.shape)
    s["rho"] = rng.uniform(-0.01, 0.01, s.shape)
    s["fy"] = rng.uniform(-1e-3, 1e-3, s.shape)
    return s


def test_collision_f_mass_and_momentum_closed_form():
    s = _single_cells()
    tau = s.p.tau
    f0 = s["f"].copy()
    s.collision_BGK()
    df = s["f"] - f0
    rho, u, v, w, F = s["rho"], s["ux"], s["uy"], s["uz"], s["fy"]
    # mass: -(sum f - rho')/tau ; momentum: -(j - u)/tau + (1 - 1/(2 tau)) F e_y
    m0 = f0.sum(axis=0)
    np.testing.assert_allclose(df.sum(axis=0), -(m0 - rho) / tau, atol=1e-15)
    j = np.einsum("i,i...->...", CX.astype(float), f0), np.einsum("i,i...->...", CY.astype(float), f0), \
        np.einsum("i,i...->...", CZ.astype(float), f0)
    dj = [np.einsum("i,i...->...", C[:, a].astype(float), df) for a in range(3)]
    np.testing.assert_allclose(dj[0], -(j[0] - u) / tau, atol=1e-15)
    np.testing.assert_allclose(dj[1], -(j[1] - v) / tau + (1 - 0.5 / tau) * F, atol=1e-15)
    np.testing.assert_allclose(dj[2], -(j[2] - w) / tau, atol=1e-15)


def test_collision_f_equilibrium_is_fixed_point_without_force():
    s = _single_cells(seed=1)
    s["fy"] = 0.0
    rho, u, v, w = s["rho"], s["ux"], s["uy"], s["uz"]
    eu = np.einsum("ia,a...->i...", C.astype(float), np.stack([u, v, w])) * 3
    uv = (u * u + v * v + w * w) * 3
    feq = W[:, None, None, None] * (rho + eu + 0.5 * (eu * eu - uv))
    s["f"] = feq
    s.collision_BGK()
    np.testing.assert_allclose(s["f"], feq, atol=1e-16)


def test_collision_g_conserves_phi():
    s = _single_cells(seed=2)
    tauc = s.p.tauc
    g0 = s["g"].copy()
    s.collision_BGK_scalar()
    dphi = (s["g"] - g0).sum(axis=0)
    np.testing.assert_allclose(dphi, -(g0.sum(axis=0) - s["phi"]) / tauc, atol=1e-15)


# ----------------------------------------------------------------- streaming
def push_stream(src, rho, t_hot, t_cold, scalar):
    """Independent push-form streaming with the reference's wall rules.

    Every post-collision value src_i(x) moves to x + c_i if that node exists.
    Otherwise it returns to x as direction opp(i): bounce-back, or for g at a
    y wall the anti-bounce-back -src_i + 2 w T_w (1 + rho').  The reference
    applies wall blocks in the order x, y, z (last writer wins), so a value
    leaving through z bounces back even when it also leaves through y.
    """
    Q, LZ, LY, LX = src.shape
    out = np.full_like(src, np.nan)
    for iz, iy, ix in itertools.product(range(LZ), range(LY), range(LX)):
        for i in range(27):
            d = (ix + CX[i], iy + CY[i], iz + CZ[i])
            out_x = not 0 <= d[0] < LX
            out_y = not 0 <= d[1] < LY
            out_z = not 0 <= d[2] < LZ
            if not (out_x or out_y or out_z):
                out[i, d[2], d[1], d[0]] = src[i, iz, iy, ix]
                continue
            o = OPP[i]
            if scalar and out_y and not out_z:
                tw = t_hot if CY[i] < 0 else t_cold  # leaving through y=0 or y=LY-1
                out[o, iz, iy, ix] = -src[i, iz, iy, ix] + 2.0 * W[o] * tw * (1.0 + rho[iz, iy, ix])
            else:
                out[o, iz, iy, ix] = src[i, iz, iy, ix]
    assert not np.isnan(out).any()
    return out


@pytest.mark.parametrize("shape", [(3, 4, 5), (5, 3, 4), (1, 3, 2), (4, 5, 1)])
def test_streaming_equals_push_form(shape):
    lx, ly, lz = shape
    s = oracle.OracleState(lx, ly, lz, rayleigh=1e6, prandtl=0.71, t_hot=0.7, t_cold=-0.3)
    rng = np.random.default_rng(3)
    s["f"] = rng.standard_normal(s["f"].shape)
    s["g"] = rng.standard_normal(s["g"].shape)
    s["rho"] = rng.uniform(-0.1, 0.1, s.shape)
    f0, g0, rho = s["f"].copy(), s["g"].copy(), s["rho"].copy()
    s.streaming()
    s.streaming_scalar()
    np.testing.assert_array_equal(s["f"], push_stream(f0, rho, 0.7, -0.3, False))
    np.testing.assert_array_equal(s["g"], push_stream(g0, rho, 0.7, -0.3, True))


def test_streaming_f_is_a_permutation():
    s = oracle.OracleState(6, 5, 7)
    s["f"] = np.random.default_rng(4).standard_normal(s["f"].shape)
    before = np.sort(s["f"].ravel())
    s.streaming()
    np.testing.assert_array_equal(np.sort(s["f"].ravel()), before)


# ----------------------------------------------------------------- g subsystem
def _diffuse(s, nsteps):
    for _ in range(nsteps):
        s.collision_BGK_scalar()
        s.streaming_scalar()
        s.macrovar()


def test_thermal_steady_profile_and_wall_position():
    """Anti-bounce-back places the isothermal walls half a cell outside the
    end nodes: the conductive steady state in a column away from the x and z
    walls is
    T(iy) = T_hot - dT (iy + 1/2)/LY.  (A wall on the end node would give
    T_hot - dT iy/(LY-1), 0.0625 away at iy = 0 for LY = 8.)  Cells next to
    the x and z walls deviate slightly: full bounce-back also reverses the
    tangential component, and at y/z edges the adiabatic z rule overrides the
    isothermal one (reference order x, y, z)."""
    LX, LY, LZ = 24, 8, 24
    s = oracle.OracleState(LX, LY, LZ, rayleigh=200.0, prandtl=1.0, t_hot=0.5, t_cold=-0.5)
    assert s.p.diff == pytest.approx((s.p.tauc - 0.5) / 3, rel=1e-12)
    s["phi"] = 0.0
    s["g"] = 0.0
    _diffuse(s, 12000)
    y = np.arange(LY) + 0.5
    halfway = 0.5 - y / LY
    centre = s["phi"][LZ // 2, :, LX // 2]
    np.testing.assert_allclose(centre, halfway, atol=2e-6)
    assert abs(centre[0] - 0.5) > 0.05
    # antisymmetric about the mid-plane, to rounding
    np.testing.assert_allclose(centre, -centre[::-1], atol=1e-13)


# ----------------------------------------------------------------- f subsystem
def test_hydrostatic_balance_under_uniform_force():
    LY, F = 16, 1.0e-5
    s = oracle.OracleState(2, LY, 2, rayleigh=1e4, prandtl=1.0)
    s["fy"] = F
    for _ in range(6000):
        s.collision_BGK()
        s.streaming()
        s.macrovar()
    for k in ("ux", "uy", "uz"):
        assert np.max(np.abs(s[k])) < 1e-9, k  # spurious corner currents ~5e-11
    d = np.diff(s["rho"][0, :, 0])
    np.testing.assert_allclose(d[2:-2], 3.0 * F, rtol=1e-6)
    assert abs(s["rho"].sum()) < 1e-14


# ----------------------------------------------------------------- full step
def test_rest_state_is_fixed_point():
    s = oracle.OracleState(5, 6, 4, rayleigh=1e5, prandtl=0.71, t_hot=0.25, t_cold=0.25)
    s.init_conductive(1, 0.0)  # phi = t0 everywhere, F = 0
    f0, g0 = s["f"].copy(), s["g"].copy()
    s.run(20)
    # sum(w) != 1 in binary64, so phi drifts by an ulp and F ~ 1e-19 appears
    np.testing.assert_allclose(s["f"], f0, rtol=0, atol=1e-17)
    np.testing.assert_allclose(s["g"], g0, rtol=0, atol=1e-15)


def test_mass_conservation():
    s = oracle.OracleState(12, 10, 8, rayleigh=1e6, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
    s.init_conductive(5, 1e-2)
    s.run(100)
    assert abs(s["rho"].sum()) < 1e-12
    assert abs(s["f"].sum()) < 1e-12


def test_x_mirror_symmetry():
    LX = 8
    s = oracle.OracleState(LX, 10, 6, rayleigh=1e6, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
    s.init_conductive(9, 1e-2)
    phi = s["phi"]
    sym = 0.5 * (phi + phi[:, :, ::-1])
    s["phi"] = sym
    s["g"] = W[:, None, None, None] * sym
    s.forcing()
    s.run(40)
    np.testing.assert_allclose(s["phi"], s["phi"][:, :, ::-1], rtol=0, atol=1e-13)
    np.testing.assert_allclose(s["ux"], -s["ux"][:, :, ::-1], rtol=0, atol=1e-15)


def test_z_mirror_symmetry():
    LZ = 8
    s = oracle.OracleState(6, 10, LZ, rayleigh=1e6, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
    s.init_conductive(4, 1e-2)
    phi = s["phi"]
    sym = 0.5 * (phi + phi[::-1, :, :])
    s["phi"] = sym
    s["g"] = W[:, None, None, None] * sym
    s.forcing()
    s.run(40)
    np.testing.assert_allclose(s["phi"], s["phi"][::-1, :, :], rtol=0, atol=1e-13)
    np.testing.assert_allclose(s["uz"], -s["uz"][::-1, :, :], rtol=0, atol=1e-15)


def test_x_z_swap_asymmetry_at_the_wall_edges():
    """The reference applies the wall rules x, then y, then z
    (src/streaming.cu:127-251): a g population whose source lies beyond a y
    wall and a z wall gets the adiabatic bounce-back, one beyond a y wall and
    an x wall the isothermal anti-bounce-back.  A state symmetric under the
    x <-> z swap therefore loses that symmetry, first on the edges where the
    y walls meet the z walls; the interior away from the edges keeps it."""
    L = 8
    s = oracle.OracleState(L, 6, L, rayleigh=1e4, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
    s.init_conductive(2, 0.0)  # laterally uniform: symmetric under x <-> z
    s.run(3)
    phi = s["phi"]  # [z][y][x]
    swapped = np.transpose(phi, (2, 1, 0))
    diff = np.abs(phi - swapped)
    assert diff.max() > 1e-6
    # after 3 steps the asymmetry has spread at most 3 cells from a wall edge
    interior = diff[3:L - 3, :, 3:L - 3]
    assert interior.max() < 1e-14
    # the edges where the y walls meet the z walls differ from their x
    # counterparts
    assert np.abs(phi[0, 0, 3:L - 3] - phi[3:L - 3, 0, 0]).max() > 1e-6


def test_conduction_below_onset_gives_unit_bulk_nusselt():
    s = oracle.OracleState(6, 8, 6, rayleigh=500.0, prandtl=1.0, t_hot=0.5, t_cold=-0.5)
    s.init_conductive(3, 1e-3)
    s.run(4000)
    nu_bulk, nu_top, nu_bottom = s.output_nu()
    # edge cells make the field weakly 3-D, which drives a faint circulation
    assert abs(nu_bulk - 1.0) < 2e-3
    assert np.max(np.abs(s["uy"])) < 1e-4
    assert nu_top == pytest.approx(nu_bottom, rel=1e-6)


# ----------------------------------------------------------------- output_nu
def test_output_nu_known_fields():
    LX, LY, LZ = 3, 7, 4
    s = oracle.OracleState(LX, LY, LZ, rayleigh=1e5, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
    b = 0.37
    s["phi"] = (0.2 - b * np.arange(LY))[None, :, None] * np.ones(s.shape)
    s["uy"] = 0.0
    nu_bulk, nu_top, nu_bottom = s.output_nu()
    assert nu_bulk == 1.0
    assert nu_top == pytest.approx(b * LY / 1.0, rel=1e-14)
    assert nu_bottom == pytest.approx(b * LY / 1.0, rel=1e-14)
    s["uy"] = 0.01
    s["phi"] = 2.0
    nu_bulk, _, _ = s.output_nu()
    assert nu_bulk == pytest.approx(1.0 + 0.02 * LY / s.p.diff, rel=1e-14)
    means = s.diag_flow()
    assert means[1] == pytest.approx(0.01) and means[3] == pytest.approx(2.0)


def test_perturbation_bounds_and_determinism():
    v = np.array([oracle.perturbation(1, i, 1e-3) for i in range(20000)])
    assert np.all(np.abs(v) < 1e-3)
    assert abs(v.mean()) < 2e-5
    assert v.std() == pytest.approx(1e-3 / math.sqrt(3), rel=0.02)
    assert oracle.perturbation(1, 0, 1e-3) == v[0]
    assert oracle.perturbation(2, 0, 1e-3) != v[0]
