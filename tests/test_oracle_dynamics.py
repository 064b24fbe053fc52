"""Invariants of the oracle's full time step (CPU): conservation, translation invariance,
shear-wave viscosity, geometry counts, and the annular Couette profile."""
import numpy as np
import pytest

from mrlbm_3d_cuda_amd import _abi, configs
from oracle.binding import Oracle


def test_periodic_box_conserves_mass_and_momentum():
    d = configs.periodic_box(24, 16, 8, tau=0.7, noise=1e-3)
    o = Oracle(d).init()
    F0 = o.fields()
    o.step(30)
    F = o.fields()
    rho0, rho1 = 1 + F0[0], 1 + F[0]
    assert abs(rho1.sum() - rho0.sum()) < 1e-11 * rho0.size
    for a in (1, 2, 3):
        assert abs((rho1 * F[a]).sum() - (rho0 * F0[a]).sum()) < 1e-13 * rho0.size


def test_translation_invariance_bitwise():
    """Shifting the initial field by whole lattice vectors shifts the result (periodic box)."""
    d = configs.periodic_box(20, 12, 8, tau=0.9, noise=1e-3)
    a = Oracle(d).init()
    F0 = a.fields()
    b = Oracle(d).init()
    shift = (3, -5, 7)  # (z, y, x)
    for f in range(10):
        b.set_field(f, np.roll(F0[f], shift, axis=(0, 1, 2)))
    a.step(12)
    b.step(12)
    assert np.array_equal(np.roll(a.fields(), shift, axis=(1, 2, 3)), b.fields())


@pytest.mark.parametrize("tau", [0.6, 0.9])
def test_shear_wave_decay_rate(tau):
    """u_x = U sin(k y) decays as exp(-nu k^2 t) with nu = (tau - 1/2) / 3."""
    nx, ny, nz = 4, 32, 4
    d = configs.periodic_box(nx, ny, nz, tau=tau, noise=0.0)
    o = Oracle(d).init()
    k = 2 * np.pi / ny
    y = np.arange(ny)
    U = 1e-3
    ux = np.broadcast_to(U * np.sin(k * y)[None, :, None], (nz, ny, nx))
    o.set_field(1, ux)
    o.set_field(4, ux ** 2)  # m = u u
    steps = 200
    o.step(steps)
    amp = np.sum(o.field(1)[0, :, 0] * np.sin(k * y)) * 2 / ny
    nu = (tau - 0.5) / 3
    expected = U * np.exp(-nu * k * k * steps)
    assert abs(amp - expected) / expected < 5e-3, (amp, expected)


def test_survey_wall_counts():
    """Wall and BCFLUID counts per z plane of the two workload geometries."""
    o = Oracle(configs.c1())
    i = o.info()
    assert (i.nb_inner, i.nb_outer) == (68 * 16, 160 * 16)
    assert i.nb_fluid_inner + i.nb_fluid_outer == 88 * 16
    assert i.active_tiles == 49 * 4  # 49 of 64 8x8 columns, 4 tiles deep
    o = Oracle(configs.c2(precision=_abi.FP64))
    i = o.info()
    assert i.nb_inner == 60 * 64 and i.nb_fluid_inner == 16 * 64
    tags = set(o.bcfluid_table(0)["tag"].tolist())
    assert tags == {17, 34, 68, 136}


def test_wall_tables_geometry():
    """Unit normals point away from the centre; delta is the distance to the wall circle."""
    d = configs.c1()
    o = Oracle(d)
    for g, radius in ((0, d.r_inner), (1, d.r_outer)):
        t = o.wall_table(g)
        x = t["node"] % d.nx
        y = (t["node"] // d.nx) % d.ny
        r = np.hypot(x - d.xc, y - d.yc)
        assert np.allclose(np.hypot(t["nx"], t["ny"]), 1, atol=1e-15)
        assert np.allclose(t["nx"] * r, x - d.xc, atol=1e-12)
        assert np.allclose(t["delta"], np.abs(radius - r), atol=1e-12)
        # the rest direction is always incoming and outgoing; masks are opposite-symmetric
        assert (t["in_mask"] & 1).all() and (t["out_mask"] & 1).all()


def test_annular_couette_profile_c1():
    """C1 after 2000 steps: u_theta(r) follows U_i R_i (R_o^2/r - r) / (R_o^2 - R_i^2) within a few %
    of U_i in the gap interior."""
    d = configs.c1()
    o = Oracle(d).init()
    o.step(2000)
    ux, uy = o.field(1)[0], o.field(2)[0]
    types = o.node_types()[0] >> _abi.TYPE_SHIFT
    yy, xx = np.mgrid[0:d.ny, 0:d.nx]
    r = np.maximum(np.hypot(xx - d.xc, yy - d.yc), 1e-9)
    ut = (-(yy - d.yc) * ux + (xx - d.xc) * uy) / r
    Ri, Ro, Ui = d.r_inner, d.r_outer, d.u_wall_inner[1]
    analytic = Ui * Ri * (Ro ** 2 / r - r) / (Ro ** 2 - Ri ** 2)
    gap = (types == _abi.NODE["BULK"]) & (r > Ri + 1.5) & (r < Ro - 1.5)
    err = np.max(np.abs(ut[gap] - analytic[gap])) / Ui
    assert err < 0.05, err


def test_plane_means_follow_the_state_after_the_step():
    """Inlet / outlet plane means (inlet_avg_density, outlet_avg_ux): the mean of rho0 + rho over
    x = 0 and of u_x over x = nx - 2 of the state after each step queued with the statistics flag;
    steps without the flag record zeros."""
    from oracle.binding import Oracle
    d = configs.cylinder(nx=48, ny=24, nz=4, diameter=6.0, xc=14.0, yc=12.0, u_max=0.05, re=40.0,
                         precision=_abi.FP64)
    o = Oracle(d).init()
    o.step(3, _abi.STEP_STATS)
    F = o.fields()
    o.step(1, 0)
    P = o.plane_means(0, 4)
    assert np.array_equal(P[3], [0.0, 0.0])
    rho_in = (d.rho0 + F[0][:, :, 0]).mean()
    ux_out = F[1][:, :, d.nx - 2].mean()
    assert abs(P[2, 0] - rho_in) < 1e-15 and abs(P[2, 1] - ux_out) < 1e-16
    assert abs(P[2, 0] - 1.0) < 0.05 and P[0, 1] != P[1, 1]
