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


# ----------------------------------------------------------------- moments of the collisions
def _second_moment(pops):
    """Pi_ab = sum_i c_ia c_ib pops_i, shape (3, 3) + cell shape."""
    Cf = C.astype(float)
    return np.einsum("ia,ib,i...->ab...", Cf, Cf, pops)


def test_collision_g_first_and_second_moments_closed_form():
    """g_eq = w phi (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2) has sum c_i g_eq = phi u
    and sum c_a c_b g_eq = phi (delta_ab / 3 + u_a u_b) (lattice isotropy up to
    4th order), so the BGK change of src/collision.cu:74-81 obeys
    sum c_i dg_i = -(sum c_i g_i - phi u) / tau_c and the same for the second
    moment.  A sign slip in c.u, a dropped 1/2 or a wrong 1/RT fails here."""
    s = _single_cells(seed=5)
    tauc = s.p.tauc
    g0 = s["g"].copy()
    s.collision_BGK_scalar()
    dg = s["g"] - g0
    phi = s["phi"]
    u = np.stack([s["ux"], s["uy"], s["uz"]])
    Cf = C.astype(float)
    j0 = np.einsum("ia,i...->a...", Cf, g0)
    dj = np.einsum("ia,i...->a...", Cf, dg)
    np.testing.assert_allclose(dj, -(j0 - phi * u) / tauc, rtol=0, atol=1e-15)
    pi0, dpi = _second_moment(g0), _second_moment(dg)
    eq = phi * (np.eye(3)[:, :, None, None, None] / 3.0 + u[:, None] * u[None, :])
    np.testing.assert_allclose(dpi, -(pi0 - eq) / tauc, rtol=0, atol=1e-15)


def test_collision_f_second_moment_without_force():
    """sum c_a c_b f_eq = rho' delta_ab / 3 + u_a u_b (f is stored as a
    deviation, the 1 of the density sits outside the quadratic terms,
    src/collision.cu:26): without force the BGK change of the momentum flux is
    -(Pi - Pi_eq) / tau."""
    s = _single_cells(seed=6)
    s["fy"] = 0.0
    tau = s.p.tau
    f0 = s["f"].copy()
    s.collision_BGK()
    df = s["f"] - f0
    u = np.stack([s["ux"], s["uy"], s["uz"]])
    eq = s["rho"] * np.eye(3)[:, :, None, None, None] / 3.0 + u[:, None] * u[None, :]
    np.testing.assert_allclose(_second_moment(df), -(_second_moment(f0) - eq) / tau, rtol=0, atol=1e-15)


# ----------------------------------------------------------------- output_profile
def test_output_profile_known_planes():
    """Per-plane means and phi_rms = sqrt(<phi^2> - <phi>^2) of
    src/output.cu:336-358 on fields built from dyadic values, so that every
    sum is exact: plane iy holds phi = a_y + b_y s(x, z) with s = +-1 balanced
    over the plane, giving mean a_y and rms |b_y| exactly; plane 0 is constant
    (rms exactly 0).  u_x, u_y, u_z are plane constants plus balanced
    patterns."""
    LX, LY, LZ = 4, 5, 6
    s = oracle.OracleState(LX, LY, LZ, rayleigh=1e5, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
    iz, iy, ix = np.meshgrid(np.arange(LZ), np.arange(LY), np.arange(LX), indexing="ij")
    sgn = np.where((ix + iz) % 2 == 0, 1.0, -1.0)       # balanced in every plane
    sgn2 = np.where(ix < LX // 2, 1.0, -1.0)            # balanced too
    a = 0.25 * iy - 0.5
    b = 0.125 * iy
    s["phi"] = a + b * sgn
    s["ux"] = 0.0625 * iy + 0.5 * sgn2
    s["uy"] = -0.03125 * iy
    s["uz"] = 0.75 * sgn
    prof = s.profile()
    y = np.arange(LY)
    np.testing.assert_array_equal(prof[:, 0], 0.0625 * y)
    np.testing.assert_array_equal(prof[:, 1], -0.03125 * y)
    np.testing.assert_array_equal(prof[:, 2], np.zeros(LY))
    np.testing.assert_array_equal(prof[:, 3], 0.25 * y - 0.5)
    np.testing.assert_array_equal(prof[:, 4], 0.125 * y)
    assert prof[0, 4] == 0.0


def test_output_profile_population_statistics():
    """On a random field the profile equals numpy's plane mean and population
    (ddof=0) standard deviation: the rms is about the plane mean, not zero and
    not with the n-1 normalisation."""
    LX, LY, LZ = 7, 6, 5
    s = oracle.OracleState(LX, LY, LZ, rayleigh=1e5, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
    rng = np.random.default_rng(11)
    for k in ("ux", "uy", "uz"):
        s[k] = rng.uniform(-0.1, 0.1, s.shape)
    s["phi"] = 0.3 + rng.uniform(-0.2, 0.2, s.shape)
    prof = s.profile()
    for col, k in enumerate(("ux", "uy", "uz", "phi")):
        np.testing.assert_allclose(prof[:, col], s[k].mean(axis=(0, 2)), rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(prof[:, 4], s["phi"].std(axis=(0, 2), ddof=0), rtol=1e-9)


# ----------------------------------------------------------------- energy and mass
def test_energy_mass_known_fields():
    """kinetic energy = 0.5 <|u|^2> over cells, mass = sum rho' (no
    normalisation): checked on dyadic fields (exact) and against numpy on a
    random field."""
    LX, LY, LZ = 4, 3, 2
    s = oracle.OracleState(LX, LY, LZ, rayleigh=1e5, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
    s["ux"], s["uy"], s["uz"] = 0.5, -0.25, 0.125
    s["rho"] = 0.0
    s["rho"][0, 0, 0] = 0.375
    s["rho"][1, 2, 3] = -0.125
    ke, mass = s.energy_mass()
    assert ke == 0.5 * (0.25 + 0.0625 + 0.015625)
    assert mass == 0.25
    rng = np.random.default_rng(12)
    for k in ("ux", "uy", "uz", "rho"):
        s[k] = rng.uniform(-0.1, 0.1, s.shape)
    ke, mass = s.energy_mass()
    usq = s["ux"] ** 2 + s["uy"] ** 2 + s["uz"] ** 2
    assert ke == pytest.approx(0.5 * usq.mean(), rel=1e-14)
    assert mass == pytest.approx(s["rho"].sum(), rel=1e-12, abs=1e-16)


# ----------------------------------------------------------------- initial states
@pytest.mark.parametrize("ly", [5, 8, 33])
def test_init_conductive_profile(ly):
    """Conductive start without perturbation: phi falls linearly from T_hot at
    iy = 0 to T_cold at iy = LY-1 (Parallel/initial.cu:89), the same in every
    x-z position; rho' = u = 0; F_y = g beta (phi - t0); f = 0 (f_eq of rho' = 0,
    src/initial.cu:237-238) and g = w phi (:240-241)."""
    t_hot, t_cold = 0.7, -0.3
    s = oracle.OracleState(3, ly, 4, rayleigh=1e5, prandtl=0.71, t_hot=t_hot, t_cold=t_cold)
    s.init_conductive(7, 0.0)
    phi = s["phi"]
    assert np.all(phi == phi[0:1, :, 0:1])  # laterally uniform
    col = phi[0, :, 0]
    assert col[0] == t_hot and col[-1] == pytest.approx(t_cold, abs=1e-15)
    d = np.diff(col)
    np.testing.assert_allclose(d, -(t_hot - t_cold) / (ly - 1), rtol=1e-12)  # constant slope
    if ly % 2 == 1:
        assert col[ly // 2] == pytest.approx(0.5 * (t_hot + t_cold), abs=1e-15)
    for k in ("rho", "ux", "uy", "uz", "fx", "fz"):
        assert np.all(s[k] == 0.0), k
    np.testing.assert_array_equal(s["fy"], s.p.grav0 * s.p.beta * (phi - s.p.t0))
    assert np.all(s["f"] == 0.0)
    np.testing.assert_array_equal(s["g"], W[:, None, None, None] * phi[None])


def test_init_arrays_reference_state():
    """init_arrays (src/initial.cu:208-242): phi = 0.5 everywhere, rho' = u = 0,
    f = 0, g = w / 2, F_y = g beta (0.5 - t0)."""
    s = oracle.OracleState(3, 4, 5, rayleigh=1e5, prandtl=0.71, t_hot=1.0, t_cold=0.0)
    s.init_arrays()
    assert np.all(s["phi"] == 0.5)
    assert np.all(s["f"] == 0.0)
    np.testing.assert_array_equal(s["g"], np.broadcast_to(W[:, None, None, None] * 0.5, s["g"].shape))
    assert np.all(s["fy"] == s.p.grav0 * s.p.beta * (0.5 - 0.5))


# ----------------------------------------------------------------- perturbation
_M64 = (1 << 64) - 1
_GAMMA = 0x9E3779B97F4A7C15


class _SplitMix64:
    """The published SplitMix64 generator (Steele, Lea and Flood, OOPSLA 2014;
    Vigna's reference splitmix64.c), restated with Python integers: the state
    advances by the golden gamma, the output is the state passed through the
    variant-13 mixer."""

    def __init__(self, seed):
        self.x = seed & _M64

    def next(self):
        self.x = (self.x + _GAMMA) & _M64
        z = self.x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)


def test_splitmix64_restatement_matches_published_vectors():
    """The restatement reproduces the published known answers (Rosetta Code,
    'Pseudo-random numbers/Splitmix64'): the first five outputs for seed
    1234567, and the bins of floor(5 u) over 100000 draws for seed 987654321."""
    r = _SplitMix64(1234567)
    assert [r.next() for _ in range(5)] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                            4593380528125082431, 16408922859458223821]
    r = _SplitMix64(987654321)
    counts = [0] * 5
    for _ in range(100000):
        counts[int(((r.next() >> 11) * 2.0 ** -53) * 5)] += 1
    assert counts == [20027, 19892, 20073, 19978, 20030]


def test_perturbation_is_splitmix64_of_seed_and_cell():
    """The draw for cell idx is the (idx+1)-th SplitMix64 output of the seed,
    its top 53 bits as u in [0, 1), returned as amp (2u - 1).  With amp = 1/2
    every step is exact, so u is recovered bit for bit."""
    for seed in (1234567, 1, 2 ** 63 + 5):
        r = _SplitMix64(seed)
        for idx in range(64):
            u = (r.next() >> 11) * 2.0 ** -53
            assert oracle.perturbation(seed, idx, 0.5) + 0.5 == u
    # the first outputs for seed 1234567 in the form the initial state uses
    want = [(v >> 11) * 2.0 ** -53 for v in (6457827717110365317, 3203168211198807973, 9817491932198370423)]
    got = [oracle.perturbation(1234567, i, 1.0e-3) for i in range(3)]
    np.testing.assert_array_equal(got, [1.0e-3 * (2.0 * u - 1.0) for u in want])


def test_init_conductive_perturbation_per_cell():
    """phi - (conductive profile) is the draw of the cell's own index
    (idx = (iz LY + iy) LX + ix), not of another cell or seed."""
    LX, LY, LZ = 3, 4, 2
    s0 = oracle.OracleState(LX, LY, LZ, rayleigh=1e5, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
    s0.init_conductive(9, 0.0)
    s = oracle.OracleState(LX, LY, LZ, rayleigh=1e5, prandtl=0.71, t_hot=0.5, t_cold=-0.5)
    s.init_conductive(9, 1.0e-3)
    r = _SplitMix64(9)
    draws = np.array([1.0e-3 * (2.0 * ((r.next() >> 11) * 2.0 ** -53) - 1.0) for _ in range(LX * LY * LZ)])
    np.testing.assert_array_equal(s["phi"].ravel(), s0["phi"].ravel() + draws)


# ----------------------------------------------------------------- macrovar, forcing
def test_macrovar_moments_with_half_force():
    """src/macrovar.cu:3-34: rho' = sum f, u = sum c f + F/2 (F as stored,
    i.e. from the previous phi), phi = sum g."""
    s = _single_cells(seed=13)
    rng = np.random.default_rng(14)
    s["fx"] = rng.uniform(-1e-3, 1e-3, s.shape)
    s["fz"] = rng.uniform(-1e-3, 1e-3, s.shape)
    f, g = s["f"].copy(), s["g"].copy()
    F = [s["fx"].copy(), s["fy"].copy(), s["fz"].copy()]
    s.macrovar()
    np.testing.assert_allclose(s["rho"], f.sum(axis=0), rtol=0, atol=1e-16)
    np.testing.assert_allclose(s["phi"], g.sum(axis=0), rtol=0, atol=1e-15)
    for a, k in enumerate(("ux", "uy", "uz")):
        j = np.einsum("i,i...->...", C[:, a].astype(float), f)
        np.testing.assert_allclose(s[k], j + 0.5 * F[a], rtol=0, atol=1e-16)


def test_forcing_boussinesq():
    """src/macrovar.cu:36-47: F = (0, g beta (phi - t0), 0)."""
    s = _single_cells(seed=15)
    s["fx"] = 1.0
    s["fz"] = 1.0
    s.forcing()
    assert np.all(s["fx"] == 0.0) and np.all(s["fz"] == 0.0)
    beta = 0.1 / (5 * 4)  # the reference default 0.1 / (5 LY), LY = 4
    assert s.p.beta == beta
    np.testing.assert_allclose(s["fy"], 1.0 * beta * (s["phi"] - 0.0), rtol=1e-15)
