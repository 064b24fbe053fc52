"""Term-by-term pins of the oracle's boundary closed forms: each closure is the exact solution of the
moment system it closes.

The reference derives its node boundary conditions (cases/cylinder/boundaries.cuh:201-843 faces,
edges and corners; :845-1037 BCFLUID; :1039-1214 BCSOLID) as closed forms of a regularised
reconstruction f^R(rho, u, m)_q = w_q rho [1 + 3 c.u + 4.5 sum_ab H_ab(c) m_ab] whose moments over
the known (incoming) populations I must equal the incoming moments, with some moments prescribed
and the density fixed by a mass condition. This file restates each system independently of the
reference's algebra -- the known set from the node's geometry, the prescribed moments, the matched
Hermite moments and the mass condition -- and checks, for random non-equilibrium populations and
several omega:

* every residual of the system vanishes (to 1e-13) at the oracle's output, and
* the system's Jacobian there is non-singular, so the solution is locally unique.

A dropped term, a wrong coefficient, sign or index in any closed form moves the output away from
the unique solution and fails the residual check (test_residuals_catch_a_mutation shows that a
1e-7 change of any single output component is caught). The mass conditions found, per family:

* faces, edges, corners, BCSOLID and the Strong BCFLUID density: the populations leaving towards
  the fluid after collision carry the incoming mass, sum_{q: -c_q in I} f^R(rho, u, m*)_q = rho_I
  with m* = (1 - omega) m + omega u u (for faces and corners this coincides with sum_I f^R = rho_I);
* inlet: sum_I f^R = rho_I evaluated with the node's previous u_x (boundaries.cuh:216);
* Equilibrium BCFLUID density: sum_I f^eq(rho, u) = rho_I.

Outlet nodes take rho and u from the convective blend of the previous moments
(boundaries.cuh:163-199), which is checked directly; their second moments solve the matching
system with that rho and u (faces) or with u = 0 in the reconstruction (edges, as the reference's
forms are written).
"""
import zlib

import numpy as np
import pytest

from mrlbm_3d_cuda_amd import _abi

# D3Q27 in the reference ordering (core/D3Q27/latticeProperties.cuh:17-21)
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1])
C = np.stack([CX, CY, CZ], axis=1).astype(float)
W = np.where(C == 0, 2 / 3, 1 / 6).prod(axis=1)
OPP = np.array([int(np.nonzero((C == -C[q]).all(1))[0][0]) for q in range(27)])
PAIRS = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
NODE = _abi.NODE
OMEGAS = [0.7, 1.3, 1.93]


def hermite(a, b):
    return C[:, a] * C[:, b] - (1.0 / 3.0 if a == b else 0.0)


PHI = {"x": C[:, 0], "y": C[:, 1], "z": C[:, 2], "xx": hermite(0, 0), "yy": hermite(1, 1), "zz": hermite(2, 2),
       "xy": hermite(0, 1), "xz": hermite(0, 2), "yz": hermite(1, 2), "yy-zz": hermite(1, 1) - hermite(2, 2)}
FIELD = {"rho": 0, "ux": 1, "uy": 2, "uz": 3, "mxx": 4, "myy": 5, "mzz": 6, "mxy": 7, "mxz": 8, "myz": 9}


def recon(x):
    """Regularised second-order populations of x = (rho, ux, uy, uz, mxx, myy, mzz, mxy, mxz, myz)."""
    s = 1.0 + 3.0 * C @ x[1:4]
    for k, (a, b) in enumerate(PAIRS):
        s = s + 4.5 * (1.0 if a == b else 2.0) * hermite(a, b) * x[4 + k]
    return W * x[0] * s


def uu(u):
    return np.array([u[0] * u[0], u[1] * u[1], u[2] * u[2], u[0] * u[1], u[0] * u[2], u[1] * u[2]])


def collided(x, omega):
    y = x.copy()
    y[4:] = (1.0 - omega) * x[4:] + omega * uu(x[1:4])
    return y


def equilibrium(x):
    y = x.copy()
    y[4:] = uu(x[1:4])
    return y


# ---- known sets from the geometry: population q is known when its source node x - c_q is fluid

SIDE = {"W": (0, -1), "E": (0, 1), "S": (1, -1), "N": (1, 1), "B": (2, -1), "F": (2, 1)}


def known_walls(sides):
    """Domain sides: the source lies outside the domain when c_q points inward through the side."""
    ok = np.ones(27, bool)
    for s in sides:
        a, sign = SIDE[s]
        ok &= ~(C[:, a] * sign < 0)
    return ok


def octants_of(d):
    """Octants (bit k set: +k direction) that contain the neighbour displacement d."""
    return [b for b in range(8) if all(d[k] == 0 or d[k] == (1 if (b >> k) & 1 else -1) for k in range(3))]


def known_bcfluid(tag):
    """BCFLUID: the tag holds the octants with SOLID neighbours; a neighbour is SOLID when every octant
    that contains it is in the tag (z-extruded walls)."""
    return np.array([q == 0 or not all((tag >> b) & 1 for b in octants_of(-C[q])) for q in range(27)])


def known_bcsolid(tag):
    """BCSOLID: the tag holds the octants with BULK neighbours; a neighbour is fluid when some octant
    that contains it is in the tag."""
    return np.array([q == 0 or any((tag >> b) & 1 for b in octants_of(-C[q])) for q in range(27)])


def mask_of(I):
    return int(sum(1 << q for q in range(27) if I[q]))


# ---- the systems

def residuals(spec, x, f, I, omega, old=None, oxm=None):
    """The 10 equations of a closure at x; each vanishes at the solution."""
    r = []
    u_mode = spec["u"]
    if u_mode == "wall":
        r += list(x[1:4])
    elif u_mode == "inlet":
        r += [x[1] - spec["U"], x[2], x[3]]
    elif u_mode == "blend":
        r += list(x[0:4] - outlet_blend(old, oxm))
    for name, value in spec.get("fixed", {}).items():
        r.append(x[FIELD[name]] - value(x))
    for extra in spec.get("extra", []):
        r.append(extra(x))
    y = x.copy()
    if spec.get("match_u0"):
        y[1:4] = 0.0
    fr = recon(y)
    for phi in spec.get("match", []):
        r.append((fr[I] * PHI[phi][I]).sum() - (f[I] * PHI[phi][I]).sum())
    rho_i = f[I].sum()
    mass = spec.get("mass")
    if mass == "out_post":
        r.append(recon(collided(x, omega))[I[OPP]].sum() - rho_i)
    elif mass == "in":
        r.append(recon(x)[I].sum() - rho_i)
    elif mass == "in_eq":
        r.append(recon(equilibrium(x))[I].sum() - rho_i)
    elif mass == "in_old_u":
        z = x.copy()
        z[1], z[4] = old[1], old[1] ** 2
        r.append(recon(z)[I].sum() - rho_i)
    assert len(r) == 10, (spec, len(r))
    return np.array(r)


def outlet_blend(old, oxm):
    """phi = (1 - Uc) phi_b + Uc phi_{x-1}, Uc = |u_x(x-1)|, on rho and u (boundaries.cuh:163-199)."""
    uc = abs(oxm[1])
    return (1.0 - uc) * old[0:4] + uc * oxm[0:4]


def jacobian(fun, x, h=1e-7):
    J = np.empty((10, 10))
    for k in range(10):
        d = np.zeros(10)
        d[k] = h * max(1.0, abs(x[k]))
        J[:, k] = (fun(x + d) - fun(x - d)) / (2 * d[k])
    return J


def zero(_):
    return 0.0


def m_fixed(names):
    return {n: zero for n in names}


WALL_FACE = {"u": "wall", "mass": "out_post"}
SPECS = {
    # no-slip faces: shears across the wall matched, the rest of m zero
    "NORTH": dict(WALL_FACE, sides="N", fixed=m_fixed(["mxx", "myy", "mzz", "mxz"]), match=["xy", "yz"]),
    "SOUTH": dict(WALL_FACE, sides="S", fixed=m_fixed(["mxx", "myy", "mzz", "mxz"]), match=["xy", "yz"]),
    "FRONT": dict(WALL_FACE, sides="F", fixed=m_fixed(["mxx", "myy", "mzz", "mxy"]), match=["xz", "yz"]),
    "BACK": dict(WALL_FACE, sides="B", fixed=m_fixed(["mxx", "myy", "mzz", "mxy"]), match=["xz", "yz"]),
    # inlet: prescribed velocity, m_xx = U^2, shears across the inlet matched
    "WEST": dict(sides="W", u="inlet", fixed={"mxx": lambda x: x[1] ** 2, "myy": zero, "mzz": zero, "myz": zero},
                 match=["xy", "xz"], mass="in_old_u"),
    # outlet: rho, u from the blend, m_xx = u_x^2, tr_yz m = u_y^2 + u_z^2, the rest matched
    "EAST": dict(sides="E", u="blend", fixed={"mxx": lambda x: x[1] ** 2},
                 extra=[lambda x: x[5] + x[6] - x[2] ** 2 - x[3] ** 2], match=["yy-zz", "xy", "xz", "yz"]),
    # wall edges: the edge shear and the shear along the free direction matched
    "NORTH_WEST": dict(WALL_FACE, sides="NW", fixed=m_fixed(["mxx", "myy", "mzz", "mxz"]), match=["xy", "yz"]),
    "SOUTH_WEST": dict(WALL_FACE, sides="SW", fixed=m_fixed(["mxx", "myy", "mzz", "mxz"]), match=["xy", "yz"]),
    "WEST_FRONT": dict(WALL_FACE, sides="WF", fixed=m_fixed(["mxx", "myy", "mzz", "mxy"]), match=["xz", "yz"]),
    "WEST_BACK": dict(WALL_FACE, sides="WB", fixed=m_fixed(["mxx", "myy", "mzz", "mxy"]), match=["xz", "yz"]),
    "NORTH_FRONT": dict(WALL_FACE, sides="NF", fixed=m_fixed(["mxx", "myy", "mzz", "mxy", "mxz"]), match=["yz"]),
    "NORTH_BACK": dict(WALL_FACE, sides="NB", fixed=m_fixed(["mxx", "myy", "mzz", "mxy", "mxz"]), match=["yz"]),
    "SOUTH_FRONT": dict(WALL_FACE, sides="SF", fixed=m_fixed(["mxx", "myy", "mzz", "mxy", "mxz"]), match=["yz"]),
    "SOUTH_BACK": dict(WALL_FACE, sides="SB", fixed=m_fixed(["mxx", "myy", "mzz", "mxy", "mxz"]), match=["yz"]),
    # outlet edges: blend, shears matched with u = 0 in the reconstruction
    "NORTH_EAST": dict(sides="NE", u="blend", fixed=m_fixed(["mxx", "myy", "mzz", "mxz"]), match=["xy", "yz"],
                       match_u0=True),
    "SOUTH_EAST": dict(sides="SE", u="blend", fixed=m_fixed(["mxx", "myy", "mzz", "mxz"]), match=["xy", "yz"],
                       match_u0=True),
    "EAST_FRONT": dict(sides="EF", u="blend", fixed=m_fixed(["mxx", "myy", "mzz", "mxy"]), match=["xz", "yz"],
                       match_u0=True),
    "EAST_BACK": dict(sides="EB", u="blend", fixed=m_fixed(["mxx", "myy", "mzz", "mxy"]), match=["xz", "yz"],
                      match_u0=True),
}
for _c in ("NORTH_WEST_FRONT", "NORTH_WEST_BACK", "SOUTH_WEST_FRONT", "SOUTH_WEST_BACK"):
    SPECS[_c] = dict(WALL_FACE, sides=_c[0] + "W" + _c.split("_")[-1][0],
                     fixed=m_fixed(["mxx", "myy", "mzz", "mxy", "mxz", "myz"]))
for _c in ("NORTH_EAST_FRONT", "NORTH_EAST_BACK", "SOUTH_EAST_FRONT", "SOUTH_EAST_BACK"):
    SPECS[_c] = dict(sides=_c[0] + "E" + _c.split("_")[-1][0], u="blend",
                     fixed=m_fixed(["mxx", "myy", "mzz", "mxy", "mxz", "myz"]))

BCSOLID_SPEC = dict(WALL_FACE, fixed=m_fixed(["mxx", "myy", "mzz", "mxy"]), match=["xz", "yz"])
BCSOLID_TAGS = (7, 11, 13, 14, 112, 176, 208, 224)
BCFLUID_TAGS = (17, 34, 68, 136)


def random_state(rng):
    """Random non-equilibrium populations around a random regularised state."""
    u = rng.uniform(-0.06, 0.06, 3)
    m = uu(u) + rng.uniform(-0.01, 0.01, 6)
    f = recon(np.r_[1.0 + rng.uniform(-0.03, 0.03), u, m])
    return f * (1.0 + rng.uniform(-0.05, 0.05, 27))


def random_old(rng):
    u = np.array([rng.uniform(0.0, 0.08), rng.uniform(-0.02, 0.02), rng.uniform(-0.02, 0.02)])
    return np.r_[1.0 + rng.uniform(-0.02, 0.02), u, uu(u) + rng.uniform(-0.005, 0.005, 6)]


def known_set(name):
    return known_walls(SPECS[name]["sides"])


def check_system(spec, out, f, I, omega, old=None, oxm=None):
    r = residuals(spec, out, f, I, omega, old, oxm)
    assert np.max(np.abs(r)) < 1e-13, r
    J = jacobian(lambda x: residuals(spec, x, f, I, omega, old, oxm), out)
    assert np.linalg.matrix_rank(J, tol=1e-9) == 10


@pytest.mark.parametrize("name", sorted(SPECS))
def test_face_edge_corner_closures_solve_their_systems(oracle_lib, name):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    spec, I = SPECS[name], known_set(name)
    for omega in OMEGAS:
        for _ in range(8):
            f, old, oxm = random_state(rng), random_old(rng), random_old(rng)
            out, hit = oracle_lib.node_bc(NODE[name], f, old, oxm, omega=omega, u_in=0.047)
            assert hit
            if spec.get("u") == "inlet":
                spec = dict(spec, U=0.047)
            check_system(spec, out, f, I, omega, old, oxm)


@pytest.mark.parametrize("name", sorted(SPECS))
def test_closures_read_only_known_populations(oracle_lib, name):
    """The output does not depend on populations from outside the domain (the known set is the one
    the geometry fixes)."""
    rng = np.random.default_rng(7)
    I = known_set(name)
    f, old = random_state(rng), random_old(rng)
    base, _ = oracle_lib.node_bc(NODE[name], f, old, old, omega=1.2, u_in=0.04)
    for q in np.nonzero(~I)[0]:
        g = f.copy()
        g[q] *= 1.7
        out, _ = oracle_lib.node_bc(NODE[name], g, old, old, omega=1.2, u_in=0.04)
        assert np.array_equal(out, base), q


@pytest.mark.parametrize("strong", [False, True])
@pytest.mark.parametrize("tag", BCFLUID_TAGS)
def test_bcfluid_closures_solve_their_systems(oracle_lib, tag, strong):
    """BCFLUID (fluid node beside a solid corner): u and all six m match the incoming moments; the
    density is the root of the Equilibrium or the Strong mass condition."""
    rng = np.random.default_rng(1000 + tag + int(strong))
    I = known_bcfluid(tag)
    assert (~I).sum() == 3  # the three populations from the solid column
    spec = dict(u="free", match=["x", "y", "z", "xx", "yy", "zz", "xy", "xz", "yz"],
                mass="out_post" if strong else "in_eq")
    for omega in OMEGAS:
        for _ in range(8):
            f = random_state(rng)
            out = oracle_lib.bcfluid(tag, mask_of(I), omega, strong, f)
            check_system(spec, out, f, I, omega)


@pytest.mark.parametrize("tag", BCSOLID_TAGS)
def test_bcsolid_closures_solve_their_systems(oracle_lib, tag):
    """BCSOLID (end-wall node beside the curved wall): u = 0, the two shears across the end wall
    matched, the outgoing post-collision mass equal to the incoming mass."""
    rng = np.random.default_rng(2000 + tag)
    I = known_bcsolid(tag)
    assert I.sum() == 16
    for omega in OMEGAS:
        for _ in range(8):
            f, old = random_state(rng), random_old(rng)
            out, hit = oracle_lib.node_bc(NODE["BCSOLID_INNER"], f, old, old, tag=tag, omega=omega)
            assert hit
            check_system(BCSOLID_SPEC, out, f, I, omega)


@pytest.mark.parametrize("name", ["NORTH_WEST", "EAST", "WEST", "SOUTH_BACK", "NORTH_EAST"])
def test_residuals_catch_a_mutation(oracle_lib, name):
    """Changing any single output component by 1e-7 (relative to its scale) violates the system:
    the check is sensitive to every term of the closed form."""
    rng = np.random.default_rng(5)
    spec, I = dict(SPECS[name], U=0.047), known_set(name)
    f, old, oxm = random_state(rng), random_old(rng), random_old(rng)
    out, _ = oracle_lib.node_bc(NODE[name], f, old, oxm, omega=1.3, u_in=0.047)
    for k in range(10):
        bad = out.copy()
        bad[k] += 1e-7 * max(1.0, abs(out[k]))
        assert np.max(np.abs(residuals(spec, bad, f, I, 1.3, old, oxm))) > 1e-10, k


def test_bcfluid_mutation_caught(oracle_lib):
    rng = np.random.default_rng(6)
    I = known_bcfluid(17)
    spec = dict(u="free", match=["x", "y", "z", "xx", "yy", "zz", "xy", "xz", "yz"], mass="in_eq")
    f = random_state(rng)
    out = oracle_lib.bcfluid(17, mask_of(I), 1.3, False, f)
    for k in range(10):
        bad = out.copy()
        bad[k] += 1e-7
        assert np.max(np.abs(residuals(spec, bad, f, I, 1.3))) > 1e-10, k
