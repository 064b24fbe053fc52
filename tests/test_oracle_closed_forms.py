"""Pins of the CPU oracle's per-node arithmetic against what the mathematics fixes.

The reference ships no tests or golden vectors, so each oracle building block is checked against
closed forms, invariants and special cases chosen so that a dropped term, a wrong sign or a
transposed index fails: Hermite round trips, the textbook equilibrium, equilibrium-in /
equilibrium-out of every closed-form boundary condition, mirror symmetry of the BCFLUID tags,
rotation invariance of the curved-wall moments and the literal extrapolation coefficients.
"""
import numpy as np
import pytest

from mrlbm_3d_cuda_amd import _abi, configs

# D3Q27 velocity set in the reference ordering (core/D3Q27/latticeProperties.cuh:17-21)
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1])
C = np.stack([CX, CY, CZ], axis=1)
T = _abi.NODE
RNG = np.random.default_rng(20261203)


def moment_vec(rho, u, m):
    """(rho, ux, uy, uz, mxx, myy, mzz, mxy, mxz, myz)"""
    return np.array([rho, *u, m[0, 0], m[1, 1], m[2, 2], m[0, 1], m[0, 2], m[1, 2]], dtype=np.float64)


def textbook_equilibrium(rho, u):
    cu = C @ u
    w = weights()
    return w * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u @ u))


def weights():
    w1 = np.where(C == 0, 2 / 3, 1 / 6)
    return w1.prod(axis=1)


def random_moments(scale=0.02):
    u = RNG.uniform(-scale, scale, 3)
    mneq = RNG.uniform(-scale, scale, (3, 3))
    mneq = 0.5 * (mneq + mneq.T) * 0.1
    return moment_vec(1 + RNG.uniform(-0.01, 0.01), u, np.outer(u, u) + mneq)


def test_lattice_weights_and_isotropy(oracle_lib):
    w = oracle_lib.pop_reconstruction(moment_vec(1.0, np.zeros(3), np.zeros((3, 3))))
    assert np.allclose(w, weights(), rtol=0, atol=1e-16)
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose(w @ C, 0, atol=1e-16)
    assert np.allclose((w[:, None, None] * C[:, :, None] * C[:, None, :]).sum(0), np.eye(3) / 3, atol=1e-16)
    # opposite directions are odd/even neighbours
    for q in range(1, 27, 2):
        assert (C[q] == -C[q + 1]).all()


@pytest.mark.parametrize("trial", range(5))
def test_reconstruction_moment_round_trip(oracle_lib, trial):
    m = random_moments()
    f = oracle_lib.pop_reconstruction(m)
    back = oracle_lib.eval_moments(f)
    assert np.allclose(back, m, rtol=0, atol=2e-16 * 10), back - m


def test_reconstruction_equals_textbook_equilibrium(oracle_lib):
    for _ in range(5):
        rho = 1 + RNG.uniform(-0.05, 0.05)
        u = RNG.uniform(-0.05, 0.05, 3)
        f = oracle_lib.pop_reconstruction(moment_vec(rho, u, np.outer(u, u)))
        assert np.allclose(f, textbook_equilibrium(rho, u), rtol=0, atol=1e-16)


def test_moments_match_direct_sums(oracle_lib):
    f = RNG.uniform(0.001, 0.05, 27)
    m = oracle_lib.eval_moments(f)
    rho = f.sum()
    j = f @ C
    P = (f[:, None, None] * C[:, :, None] * C[:, None, :]).sum(0)
    assert np.isclose(m[0], rho, rtol=1e-15)
    assert np.allclose(m[1:4], j / rho, rtol=1e-14, atol=1e-18)
    ref = P / rho - np.eye(3) / 3
    assert np.allclose(m[4:], [ref[0, 0], ref[1, 1], ref[2, 2], ref[0, 1], ref[0, 2], ref[1, 2]], atol=1e-15)


def test_collision_fixed_point_and_conservation(oracle_lib):
    m = random_moments()
    u = m[1:4]
    uu = [u[0] ** 2, u[1] ** 2, u[2] ** 2, u[0] * u[1], u[0] * u[2], u[1] * u[2]]
    out = oracle_lib.collide(1.0, m)  # omega = 1: relaxes fully to u u
    assert np.allclose(out[4:], uu, atol=1e-18)
    assert np.array_equal(out[:4], m[:4])
    eq = m.copy()
    eq[4:] = uu
    assert np.allclose(oracle_lib.collide(1.7, eq)[4:], uu, atol=1e-18)
    omega = 1.3
    out = oracle_lib.collide(omega, m)
    assert np.allclose(out[4:], (1 - omega) * m[4:] + omega * np.array(uu), atol=1e-18)


# ---- domain-face closed forms: equilibrium in -> equilibrium out

def eq_pops(oracle_lib, rho, u):
    return oracle_lib.pop_reconstruction(moment_vec(rho, u, np.outer(u, u)))


@pytest.mark.parametrize("name", ["NORTH", "SOUTH", "NORTH_WEST", "SOUTH_WEST", "FRONT", "BACK", "NORTH_FRONT",
                                  "NORTH_BACK", "SOUTH_FRONT", "SOUTH_BACK", "WEST_FRONT", "WEST_BACK",
                                  "NORTH_WEST_FRONT", "NORTH_WEST_BACK", "SOUTH_WEST_FRONT", "SOUTH_WEST_BACK"])
@pytest.mark.parametrize("omega", [0.8, 1.25, 1.9])
def test_no_slip_faces_at_rest(oracle_lib, name, omega):
    f = eq_pops(oracle_lib, 1.0, np.zeros(3))
    old = moment_vec(1.0, np.zeros(3), np.zeros((3, 3)))
    out, hit = oracle_lib.node_bc(T[name], f, old, omega=omega)
    assert hit
    assert np.allclose(out, old, atol=1e-15), out


@pytest.mark.parametrize("U", [0.0, 0.03, 0.08])
def test_inlet_equilibrium(oracle_lib, U):
    u = np.array([U, 0, 0])
    f = eq_pops(oracle_lib, 1.0, u)
    old = moment_vec(1.0, u, np.outer(u, u))
    out, hit = oracle_lib.node_bc(T["WEST"], f, old, u_in=U)
    assert hit
    assert np.allclose(out, old, atol=1e-15), out - old
    # the density closure uses the node's previous u_x (boundaries.cuh:216), not the prescribed one
    old2 = old.copy()
    old2[1] = 0.0
    out2, _ = oracle_lib.node_bc(T["WEST"], f, old2, u_in=U)
    if U:
        assert not np.isclose(out2[0], 1.0)


@pytest.mark.parametrize("name", ["EAST", "NORTH_EAST", "SOUTH_EAST", "EAST_FRONT", "EAST_BACK", "NORTH_EAST_FRONT",
                                  "NORTH_EAST_BACK", "SOUTH_EAST_FRONT", "SOUTH_EAST_BACK"])
def test_outlet_blend_and_equilibrium(oracle_lib, name):
    U = 0.04 if name == "EAST" else 0.0
    u = np.array([U, 0, 0])
    f = eq_pops(oracle_lib, 1.0, u)
    old = moment_vec(1.0, u, np.outer(u, u))
    out, hit = oracle_lib.node_bc(T[name], f, old, old, omega=1.2)
    assert hit
    assert np.allclose(out, old, atol=1e-15), out - old
    # convective blend: phi = (1 - Uc) phi_b + Uc phi_{x-1}, Uc = |u_x(x-1)|
    ob = moment_vec(1.01, [0.02, 0.001, -0.002], np.zeros((3, 3)))
    oi = moment_vec(0.99, [-0.03, 0.004, 0.001], np.zeros((3, 3)))
    out, _ = oracle_lib.node_bc(T[name], f, ob, oi, omega=1.2)
    Uc = 0.03
    assert np.allclose(out[:4], (1 - Uc) * ob[:4] + Uc * oi[:4], atol=1e-16)


def bcfluid_masks():
    from oracle.binding import Oracle
    o = Oracle(configs.c2(precision=_abi.FP64))
    t = o.bcfluid_table(0)
    return {int(tag): int(t["in_mask"][tag]) for tag in (17, 34, 68, 136)}


@pytest.mark.parametrize("tag", [17, 34, 68, 136])
def test_bcfluid_tags_at_rest(oracle_lib, tag):
    masks = bcfluid_masks()
    f = eq_pops(oracle_lib, 1.0, np.zeros(3))
    old = moment_vec(1.0, np.zeros(3), np.zeros((3, 3)))
    out, hit = oracle_lib.node_bc(T["BCFLUID_INNER"], f, old, tag=tag, in_mask=masks[tag], omega=1.0)
    assert hit
    assert np.allclose(out, old, atol=1e-13), out - old


def mirror_perm(axis):
    """q -> index of the direction mirrored along `axis`."""
    Cm = C.copy()
    Cm[:, axis] *= -1
    return np.array([int(np.where((C == Cm[q]).all(1))[0][0]) for q in range(27)])


def test_bcfluid_mirror_symmetry(oracle_lib):
    """Tag 34 is the x-mirror of tag 17, 68 the y-mirror, 136 both: applying a tag to mirrored
    populations equals mirroring tag 17's result."""
    masks = bcfluid_masks()
    px, py = mirror_perm(0), mirror_perm(1)
    f = oracle_lib.pop_reconstruction(random_moments(0.03))
    base, _ = oracle_lib.node_bc(T["BCFLUID_INNER"], f, np.ones(10), tag=17, in_mask=masks[17])

    def mirrored_mask(mask, perm):
        return sum(1 << int(perm[q]) for q in range(27) if mask >> q & 1)

    # sign flips of (rho, ux, uy, uz, mxx, myy, mzz, mxy, mxz, myz) under x and y mirrors
    sx = np.array([1, -1, 1, 1, 1, 1, 1, -1, -1, 1])
    sy = np.array([1, 1, -1, 1, 1, 1, 1, -1, 1, -1])
    for tag, perm, sign in [(34, px, sx), (68, py, sy), (136, px[py], sx * sy)]:
        assert masks[tag] == mirrored_mask(masks[17], perm)
        fm = np.empty(27)
        fm[perm] = f
        out, hit = oracle_lib.node_bc(T["BCFLUID_INNER"], fm, np.ones(10), tag=tag, in_mask=masks[tag])
        assert hit
        assert np.allclose(out, sign * base, rtol=1e-13, atol=1e-16), (tag, out - sign * base)


def test_bcfluid_other_tags_leave_node_unchanged(oracle_lib):
    f = oracle_lib.pop_reconstruction(random_moments())
    old = random_moments()
    out, hit = oracle_lib.node_bc(T["BCFLUID_INNER"], f, old, tag=51, in_mask=(1 << 27) - 1)
    assert not hit and np.array_equal(out, old)


# ---- BCSOLID end-wall nodes (bcsolid_boundary_condition, boundaries.cuh:1039-1214)

BCS_TAGS = [7, 11, 13, 14, 112, 176, 208, 224]


@pytest.mark.parametrize("tag", BCS_TAGS)
@pytest.mark.parametrize("omega", [0.8, 1.25, 1.9])
def test_bcsolid_at_rest(oracle_lib, tag, omega):
    f = eq_pops(oracle_lib, 1.0, np.zeros(3))
    old = moment_vec(1.0, np.zeros(3), np.zeros((3, 3)))
    out, hit = oracle_lib.node_bc(T["BCSOLID_INNER"], f, np.ones(10), tag=tag, omega=omega)
    assert hit
    assert np.allclose(out, old, atol=1e-14), out - old


def test_bcsolid_tags_are_mirror_images(oracle_lib):
    """The eight tags are one geometry seen through the lattice symmetries: inverting every
    velocity maps tag 7 to 224, 11 to 208, 13 to 176, 14 to 112 (m_xz, m_yz even), and reflecting
    x and y maps 7 to 14 and 11 to 13 (m_xz, m_yz odd). A sign slip in any closed form breaks one
    of these."""
    inv = np.array([int(np.where((C == -C[q]).all(1))[0][0]) for q in range(27)])
    pxy = mirror_perm(0)[mirror_perm(1)]
    even = np.ones(10)
    odd_xy = np.array([1, -1, -1, 1, 1, 1, 1, 1, -1, -1])
    f = oracle_lib.pop_reconstruction(random_moments(0.03))
    for a, b, perm, sign in [(7, 224, inv, even), (11, 208, inv, even), (13, 176, inv, even), (14, 112, inv, even),
                             (7, 14, pxy, odd_xy), (11, 13, pxy, odd_xy)]:
        fa, _ = oracle_lib.node_bc(T["BCSOLID_INNER"], f, np.ones(10), tag=a, omega=1.3)
        fm = np.empty(27)
        fm[perm] = f
        fb, _ = oracle_lib.node_bc(T["BCSOLID_INNER"], fm, np.ones(10), tag=b, omega=1.3)
        assert np.allclose(fb, sign * fa, rtol=1e-12, atol=1e-15), (a, b, fb - sign * fa)


def test_bcsolid_other_tags_leave_node_unchanged(oracle_lib):
    f = oracle_lib.pop_reconstruction(random_moments())
    old = random_moments()
    out, hit = oracle_lib.node_bc(T["BCSOLID_INNER"], f, old, tag=3)
    assert not hit and np.array_equal(out, old)


# ---- curved wall

def test_incoming_rotated_full_mask_is_moments(oracle_lib):
    f = oracle_lib.pop_reconstruction(random_moments())
    m = oracle_lib.eval_moments(f)
    rho, m6 = oracle_lib.incoming_rotated(1.0, 0.0, (1 << 27) - 1, f)
    assert np.isclose(rho, m[0], rtol=1e-15)
    assert np.allclose(m6, m[4:], atol=1e-16)


def test_incoming_rotated_is_a_rotation(oracle_lib):
    f = oracle_lib.pop_reconstruction(random_moments())
    m = oracle_lib.eval_moments(f)
    th = 0.7
    c, s = np.cos(th), np.sin(th)
    R = np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]])
    M = np.array([[m[4], m[7], m[8]], [m[7], m[5], m[9]], [m[8], m[9], m[6]]])
    Mp = R @ M @ R.T
    _, m6 = oracle_lib.incoming_rotated(c, s, (1 << 27) - 1, f)
    assert np.allclose(m6, [Mp[0, 0], Mp[1, 1], Mp[2, 2], Mp[0, 1], Mp[0, 2], Mp[1, 2]], atol=1e-15)


@pytest.mark.parametrize("delta", [0.0, 0.3, 0.71, 1.0])
def test_extrapolation_coefficients_literal(oracle_lib, delta):
    c = oracle_lib.extrapolation_coeffs(delta)
    # literal coefficients of extrapolation_utils.cuh:52-58
    assert np.allclose(c, [1 - 0.5 * delta ** 2 + 1.5 * delta, delta ** 2 - 2 * delta, -0.5 * (delta ** 2 - delta)],
                       atol=1e-16)
    assert abs(c.sum() - 1) < 1e-15  # constants reproduced
    assert abs(c[1] + 2 * c[2] - (-delta)) < 1e-15  # linear data extrapolated to s = -delta
    # quadratic data give -delta^2 (the reference's sign), not the Lagrange +delta^2
    assert abs(c[1] + 4 * c[2] - (-delta ** 2)) < 1e-15


@pytest.mark.parametrize("theta", [0.0, 0.4, 2.2, -1.1])
def test_curved_closure_reproduces_equilibrium(oracle_lib, theta):
    """Incoming populations of an equilibrium at the wall velocity, with extrapolated wall-frame
    moments of that equilibrium, close to rho = 1 and m = u u."""
    nx, ny = np.cos(theta), np.sin(theta)
    # an incoming mask of a wall whose fluid side is +n: directions with c.n <= 0.3 are unknown
    mask = 1
    for q in range(1, 27):
        if -(C[q, 0] * nx + C[q, 1] * ny) <= 0.3:  # population arriving from x - c is from the fluid
            mask |= 1 << q
    up = np.array([0.0, 0.04, 0.01])  # wall-frame velocity
    Rt = np.array([[nx, -ny, 0], [ny, nx, 0], [0, 0, 1]])  # wall frame -> lattice
    u = Rt @ up
    f = eq_pops(oracle_lib, 1.0, u)
    rho_i, m6 = oracle_lib.incoming_rotated(nx, ny, mask, f)
    prime7 = [up[0], up[1], up[2], up[0] ** 2, up[1] ** 2, up[2] ** 2, up[1] * up[2]]
    out = oracle_lib.curved_solution(nx, ny, mask, prime7, rho_i, m6[3], m6[4])
    uu = np.outer(u, u)
    assert abs(out[0] - 1) < 1e-13
    assert np.allclose(out[1:4], u, atol=1e-15)
    assert np.allclose(out[4:], [uu[0, 0], uu[1, 1], uu[2, 2], uu[0, 1], uu[0, 2], uu[1, 2]], atol=1e-13)


def opposite_mask(mask):
    """Directions q whose opposite is in mask (find_incomings_outgoings' outgoing set) plus rest."""
    opp = [int(np.where((C == -C[q]).all(1))[0][0]) for q in range(27)]
    return 1 | sum(1 << q for q in range(27) if mask >> opp[q] & 1)


def wall_mask(nx, ny):
    """An incoming mask of a wall whose fluid side is +n (rest direction included)."""
    mask = 1
    for q in range(1, 27):
        if -(C[q, 0] * nx + C[q, 1] * ny) <= 0.3:
            mask |= 1 << q
    return mask


@pytest.mark.parametrize("theta,omega", [(0.0, 1.0), (0.4, 1.6), (2.2, 0.8), (-1.1, 1.95)])
def test_curved_closure_strong_conserves_mass(oracle_lib, theta, omega):
    """MassBC::Strong closure (numerical_solution.cu:124-279): after the closure and the
    collision, the populations leaving the wall node along its outgoing directions carry exactly
    the mass rho_I that arrived along the incoming ones, for any state; at rest it gives rho = 1,
    m = 0."""
    nx, ny = np.cos(theta), np.sin(theta)
    mask = wall_mask(nx, ny)
    out_mask = opposite_mask(mask)
    rng = np.random.default_rng(int(1000 * omega))
    f = oracle_lib.pop_reconstruction(random_moments(0.04))
    rho_i, m6 = oracle_lib.incoming_rotated(nx, ny, mask, f)
    prime7 = rng.uniform(-0.03, 0.03, 7)
    out = oracle_lib.curved_solution_strong(nx, ny, mask, out_mask, omega, prime7, rho_i, m6[3], m6[4])
    fpost = oracle_lib.pop_reconstruction(oracle_lib.collide(omega, out))
    sent = sum(fpost[q] for q in range(27) if out_mask >> q & 1)
    assert abs(sent - rho_i) < 1e-13 * rho_i, (sent, rho_i)
    # the Equilibrium closure does not balance this (it sets rho = rho_I / E')
    eq = oracle_lib.curved_solution(nx, ny, mask, prime7, rho_i, m6[3], m6[4])
    sent_eq = sum(oracle_lib.pop_reconstruction(oracle_lib.collide(omega, eq))[q] for q in range(27)
                  if out_mask >> q & 1)
    assert abs(sent_eq - rho_i) > 1e-6
    # at rest
    f0 = eq_pops(oracle_lib, 1.0, np.zeros(3))
    rho0_i, m60 = oracle_lib.incoming_rotated(nx, ny, mask, f0)
    rest = oracle_lib.curved_solution_strong(nx, ny, mask, out_mask, omega, np.zeros(7), rho0_i, m60[3], m60[4])
    assert abs(rest[0] - 1) < 1e-13 and np.allclose(rest[1:], 0, atol=1e-14)


@pytest.mark.parametrize("tag", [17, 34, 68, 136])
@pytest.mark.parametrize("omega", [0.7, 1.0, 1.9])
def test_bcfluid_strong_at_rest(oracle_lib, tag, omega):
    """The Strong BCFLUID quadratic (boundaries.cuh:894-904 and its mirrors) has rho = 1 as the
    root selected for populations at rest, for any omega; u and m stay zero."""
    masks = bcfluid_masks()
    f = eq_pops(oracle_lib, 1.0, np.zeros(3))
    out = oracle_lib.bcfluid(tag, masks[tag], omega, True, f)
    assert out is not None
    assert abs(out[0] - 1) < 1e-13
    assert np.allclose(out[1:], 0, atol=1e-14)


def test_bcfluid_strong_mirror_symmetry(oracle_lib):
    masks = bcfluid_masks()
    px, py = mirror_perm(0), mirror_perm(1)
    f = oracle_lib.pop_reconstruction(random_moments(0.03))
    base = oracle_lib.bcfluid(17, masks[17], 1.3, True, f)
    sx = np.array([1, -1, 1, 1, 1, 1, 1, -1, -1, 1])
    sy = np.array([1, 1, -1, 1, 1, 1, 1, -1, 1, -1])
    for tag, perm, sign in [(34, px, sx), (68, py, sy), (136, px[py], sx * sy)]:
        fm = np.empty(27)
        fm[perm] = f
        out = oracle_lib.bcfluid(tag, masks[tag], 1.3, True, fm)
        assert np.allclose(out, sign * base, rtol=1e-12, atol=1e-15), (tag, out - sign * base)
    # the Strong density differs from the Equilibrium one away from rest
    eq = oracle_lib.bcfluid(17, masks[17], 1.3, False, f)
    assert abs(eq[0] - base[0]) > 1e-9


# ---- wall-pressure statistics (surface_pressure_extrapolation, extrapolation_utils.cuh:109-154)

@pytest.mark.parametrize("theta,R", [(0.0, 10.0), (0.9, 7.3), (2.6, 12.5), (-2.0, 5.0)])
def test_pressure_extrapolation_reproduces_quadratics(theta, R):
    """Pressures from an exact quadratic p(r) = a + b r + c r^2 at the three fluid points give
    that quadratic's value at the wall radius."""
    from oracle.binding import surface_pressure_extrapolation
    xc, yc = 31.5, 20.25
    nx, ny = np.cos(theta), np.sin(theta)
    a, b, c = 0.3312, 0.0021, -0.00037
    pts = [(xc + (R + k) * nx, yc + (R + k) * ny) for k in range(4)]  # wall point, then 1, 2, 3 out
    rho = [(a + b * (R + k) + c * (R + k) ** 2) * 3.0 for k in (1, 2, 3)]
    xy8 = [v for p in pts for v in p]
    ps = surface_pressure_extrapolation(xc, yc, xy8, rho)
    assert abs(ps - (a + b * R + c * R * R)) < 1e-13


def test_wall_pressure_of_uniform_and_linear_density():
    from oracle.binding import Oracle
    d = configs.cylinder(nx=64, ny=40, nz=3, diameter=12.0, xc=20.0, yc=19.5, precision=_abi.FP64, noise=0.0)
    o = Oracle(d).init()
    assert np.allclose(o.surface_pressure(), 1.0 / 3.0, atol=1e-15)
    # bilinear interpolation is exact on a linear field, and p(r) along a normal is then linear
    z, y, x = np.meshgrid(np.arange(d.nz), np.arange(d.ny), np.arange(d.nx), indexing="ij")
    alpha, beta = 2e-4, -3e-4
    o.set_field(0, alpha * x + beta * y)
    t = o.wall_table(0)
    R = d.r_inner
    expect = (1.0 + alpha * (d.xc + R * t["nx"]) + beta * (d.yc + R * t["ny"])) / 3.0
    assert np.allclose(o.surface_pressure(), expect, atol=1e-14)
