"""Workload descriptors: the five configurations the solver is measured and checked on.

C1  annular Couette flow, 64x64x16, FP64          (parity and analytic profile)
C2  confined cylinder, 256x128x64, FP32           (per-step Cd/Cl; the benchmark workload)
C3  cylinder wake, 512x256x256, BF16/FP16 storage (error vs FP64)
C4  64 batched annuli, FP32
C5  256 batched cylinders, BF16 storage

Conventions that the reference fixes at compile time (cases/*/constants.cuh) are runtime fields
here; the choices made for the configurations are:
  * Re -> tau with the reference's convention nu = u_max * D / Re, tau = 0.5 + 3 nu
    (cases/cylinder/constants.cuh:52-54);
  * wall centres on lattice nodes;
  * inlet/initial noise from the counter hash in the library (seed per configuration).
"""
from ._abi import (CaseDesc, CASE_ANNULUS, CASE_CYLINDER, CASE_PERIODIC_BOX, FP32, FP32_BF16,
                   FP32_FP16, FP64, INLET_PARABOLIC, INLET_UNIFORM)

SEED_BASE = 0x4D524C4200000000  # "MRLB"


def tau_from_re(u_max, diameter, re):
    nu = u_max * diameter / re
    return 0.5 + 3.0 * nu


def periodic_box(nx=64, ny=64, nz=16, tau=0.8, precision=FP64, noise=1e-4, seed=SEED_BASE | 0xB0):
    d = CaseDesc()
    d.kind = CASE_PERIODIC_BOX
    d.precision = precision
    d.nx, d.ny, d.nz = nx, ny, nz
    d.periodic_x = d.periodic_y = d.periodic_z = 1
    d.tau = tau
    d.rho0 = 1.0
    d.u_max = 0.05
    d.init_noise = noise
    d.seed = seed
    return d


def annulus(nx=64, ny=64, nz=16, xc=32.0, yc=32.0, r_in=12.0, r_out=28.0, u_inner=0.05,
            u_outer=0.0, tau=0.8, precision=FP64, noise=1e-4, seed=SEED_BASE | 0x01):
    """C1: annular Couette flow; inner wall tangential speed u_inner, outer at u_outer."""
    d = CaseDesc()
    d.kind = CASE_ANNULUS
    d.precision = precision
    d.nx, d.ny, d.nz = nx, ny, nz
    d.periodic_x = d.periodic_y = 0
    d.periodic_z = 1
    d.tau = tau
    d.rho0 = 1.0
    d.u_max = max(abs(u_inner), abs(u_outer))
    d.xc, d.yc = xc, yc
    d.r_inner, d.r_outer = r_in, r_out
    d.u_wall_inner[:] = [0.0, u_inner, 0.0]
    d.u_wall_outer[:] = [0.0, u_outer, 0.0]
    d.init_noise = noise
    d.seed = seed
    return d


def cylinder(nx=256, ny=128, nz=64, diameter=20.0, xc=64.0, yc=64.0, u_max=0.05, re=100.0,
             profile=INLET_PARABOLIC, periodic_y=0, precision=FP32, noise=1e-4,
             inlet_noise=0.0, seed=SEED_BASE | 0x02):
    """C2: confined flow past a circular cylinder (walls at y=0, ny-1 unless periodic_y)."""
    d = CaseDesc()
    d.kind = CASE_CYLINDER
    d.precision = precision
    d.nx, d.ny, d.nz = nx, ny, nz
    d.periodic_x = 0
    d.periodic_y = periodic_y
    d.periodic_z = 1
    d.tau = tau_from_re(u_max, diameter, re)
    d.rho0 = 1.0
    d.u_max = u_max
    d.inlet_profile = profile
    d.xc, d.yc = xc, yc
    d.r_inner = 0.5 * diameter
    d.r_outer = 0.0
    d.init_noise = noise
    d.inlet_noise = inlet_noise
    d.seed = seed
    return d


def c1(precision=FP64):
    return annulus(precision=precision)


def c2(precision=FP32):
    return cylinder(precision=precision)


def c3(precision=FP32_BF16):
    """C3: 512x256x256 wake, D=32, uniform 0.04 inflow with 1e-3 inlet perturbation, Re=1000."""
    return cylinder(nx=512, ny=256, nz=256, diameter=32.0, xc=128.0, yc=128.0, u_max=0.04,
                    re=1000.0, profile=INLET_UNIFORM, periodic_y=1, precision=precision,
                    noise=1e-4, inlet_noise=1e-3, seed=SEED_BASE | 0x03)


def c4_cases(n=64, precision=FP32, seed=SEED_BASE | 0x04):
    """C4: seeded annulus sweep. R_in ~ U[8, 20], R_out = R_in + U[10, 30], nz ~ U{16..64},
    nx = ny = ceil(2 R_out + 8) rounded up to a multiple of 8, centre at the grid middle, inner
    wall speed ~ U[0.02, 0.08] (outer at rest), tau ~ U[0.55, 1.2]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        r_in = float(rng.uniform(8.0, 20.0))
        r_out = r_in + float(rng.uniform(10.0, 30.0))
        nz = int(rng.integers(16, 65))
        nxy = int(np.ceil(2.0 * r_out + 8.0))
        nxy = (nxy + 7) // 8 * 8
        u = float(rng.uniform(0.02, 0.08))
        tau = float(rng.uniform(0.55, 1.2))
        c = 0.5 * (nxy - 1)
        out.append(annulus(nx=nxy, ny=nxy, nz=nz, xc=c, yc=c, r_in=r_in, r_out=r_out, u_inner=u, tau=tau,
                           precision=precision, seed=seed + 0x100 * (i + 1)))
    return out


def c5_cases(n=256, precision=FP32_BF16, seed=SEED_BASE | 0x05):
    """C5: seeded cylinder sweep in confined channels as C2: nx ~ U{64..192}, ny ~ U{64..96},
    nz ~ U{32..96} (multiples of 8), D ~ U[8, 24], Re ~ U[20, 300], u_max ~ U[0.03, 0.06]; the
    cylinder sits at x = max(nx / 4, D + 8), y = ny / 2."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nx = int(rng.integers(8, 25)) * 8
        ny = int(rng.integers(8, 13)) * 8
        nz = int(rng.integers(4, 13)) * 8
        dia = float(rng.uniform(8.0, 24.0))
        re = float(rng.uniform(20.0, 300.0))
        u = float(rng.uniform(0.03, 0.06))
        xc = float(max(nx // 4, int(dia) + 8))
        out.append(cylinder(nx=nx, ny=ny, nz=nz, diameter=dia, xc=xc, yc=float(ny // 2), u_max=u, re=re,
                            precision=precision, seed=seed + 0x100 * (i + 1)))
    return out


CONFIGS = {
    "c1_annulus_64x64x16_f64": c1,
    "c2_cylinder_256x128x64_f32": c2,
    "c3_wake_512x256x256": c3,
}

__all__ = ["periodic_box", "annulus", "cylinder", "c1", "c2", "c3", "c4_cases", "c5_cases", "tau_from_re",
           "CONFIGS", "FP64", "FP32", "FP32_BF16", "FP32_FP16"]
