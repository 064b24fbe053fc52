"""Python binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module. It
builds oracle/liboracle.so with the committed Makefile if the library is missing.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from mrlbm_3d_cuda_amd import _abi
from mrlbm_3d_cuda_amd._abi import CaseDesc, Info, as_ptr

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_P = C.POINTER
_dp = _P(C.c_double)
_u32p = _P(C.c_uint32)
_i64p = _P(C.c_int64)

_PROTOS = {
    "oracle_create": (C.c_int, [_P(CaseDesc), C.c_int32, _P(C.c_void_p)]),
    "oracle_init": (C.c_int, [C.c_void_p]),
    "oracle_step": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32]),
    "oracle_get_moments": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "oracle_set_moments": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "oracle_get_forces": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, _dp]),
    "oracle_get_node_types": (C.c_int, [C.c_void_p, _u32p]),
    "oracle_get_plane_means": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, _dp]),
    "oracle_get_wall_table": (C.c_int, [C.c_void_p, C.c_int32, _i64p, _dp, _dp, _dp, _u32p, _u32p]),
    "oracle_get_bcfluid_table": (C.c_int, [C.c_void_p, C.c_int32, _i64p, _u32p, _u32p, _u32p]),
    "oracle_get_info": (C.c_int, [C.c_void_p, _P(Info)]),
    "oracle_destroy": (C.c_int, [C.c_void_p]),
    "oracle_num_threads": (C.c_int32, []),
    "oracle_pop_reconstruction": (None, [_dp, _dp]),
    "oracle_eval_moments": (None, [_dp, _dp]),
    "oracle_collide": (None, [C.c_double, _dp]),
    "oracle_node_bc": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double, _dp, _dp, _dp, _dp]),
    "oracle_incoming_rotated": (None, [C.c_double, C.c_double, C.c_uint32, _dp, _dp, _dp]),
    "oracle_extrapolation_coeffs": (None, [C.c_double, _dp]),
    "oracle_surface_pressure": (C.c_int, [C.c_void_p, _dp]),
    "oracle_surface_pressure_extrapolation": (C.c_double, [C.c_double, C.c_double, _dp, _dp]),
    "oracle_vti_points": (None, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                 C.c_double, _dp, _dp, _dp, _dp, _P(C.c_uint8), _P(C.c_float)]),
    "oracle_write_vti": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P(C.c_float)]),
    "oracle_write_master_pvd": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32]),
    "oracle_curved_solution_strong": (None, [C.c_double, C.c_double, C.c_uint32, C.c_uint32, C.c_double, _dp,
                                             C.c_double, C.c_double, C.c_double, _dp]),
    "oracle_bcfluid": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_int32, _dp, _dp]),
    "oracle_curved_solution": (None, [C.c_double, C.c_double, C.c_uint32, _dp, C.c_double, C.c_double, C.c_double, _dp]),
}

_lib = None


def build(force=False):
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = _abi.bind(C.CDLL(LIB_PATH), _PROTOS)
    return _lib


class OracleError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        raise OracleError(f"{what} failed with status {rc}")


class Oracle:
    """CPU restatement of the reference step on one case."""

    def __init__(self, desc: CaseDesc, real_bytes=8):
        self.desc = desc.copy()
        self.lib = lib()
        h = C.c_void_p()
        _check(self.lib.oracle_create(C.byref(self.desc), real_bytes, C.byref(h)), "oracle_create")
        self.h = h
        self.shape = (desc.nz, desc.ny, desc.nx)
        self.n = desc.num_nodes

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    __del__ = close

    def init(self):
        _check(self.lib.oracle_init(self.h), "oracle_init")
        return self

    def step(self, n=1, flags=0):
        _check(self.lib.oracle_step(self.h, n, flags), "oracle_step")
        return self

    def field(self, f):
        out = np.empty(self.n, dtype=np.float64)
        _check(self.lib.oracle_get_moments(self.h, f, as_ptr(out, C.c_double)), "oracle_get_moments")
        return out.reshape(self.shape)

    def fields(self):
        return np.stack([self.field(f) for f in range(_abi.NUM_FIELDS)])

    def set_field(self, f, values):
        arr = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        _check(self.lib.oracle_set_moments(self.h, f, as_ptr(arr, C.c_double)), "oracle_set_moments")

    def forces(self, first, count):
        out = np.empty((count, 4), dtype=np.float64)
        _check(self.lib.oracle_get_forces(self.h, first, count, as_ptr(out, C.c_double)), "oracle_get_forces")
        return out

    def node_types(self):
        out = np.empty(self.n, dtype=np.uint32)
        _check(self.lib.oracle_get_node_types(self.h, as_ptr(out, C.c_uint32)), "oracle_get_node_types")
        return out.reshape(self.shape)

    def info(self):
        i = Info()
        _check(self.lib.oracle_get_info(self.h, C.byref(i)), "oracle_get_info")
        return i

    def wall_table(self, group):
        info = self.info()
        n = info.nb_inner if group == 0 else info.nb_outer
        node = np.empty(n, np.int64)
        nx, ny, delta = (np.empty(n, np.float64) for _ in range(3))
        fin, fout = np.empty(n, np.uint32), np.empty(n, np.uint32)
        _check(self.lib.oracle_get_wall_table(self.h, group, as_ptr(node, C.c_int64), as_ptr(nx, C.c_double),
                                              as_ptr(ny, C.c_double), as_ptr(delta, C.c_double),
                                              as_ptr(fin, C.c_uint32), as_ptr(fout, C.c_uint32)),
               "oracle_get_wall_table")
        return dict(node=node, nx=nx, ny=ny, delta=delta, in_mask=fin, out_mask=fout)

    def surface_pressure(self):
        """Instantaneous wall pressure of every inner-wall node (compute_surface_pressure)."""
        out = np.empty(self.info().nb_inner, np.float64)
        _check(self.lib.oracle_surface_pressure(self.h, as_ptr(out, C.c_double)), "surface_pressure")
        return out

    def plane_means(self, first, count):
        out = np.empty((count, 2), np.float64)
        _check(self.lib.oracle_get_plane_means(self.h, first, count, as_ptr(out, C.c_double)), "oracle_get_plane_means")
        return out

    def bcfluid_table(self, group):
        info = self.info()
        n = info.nb_fluid_inner if group == 0 else info.nb_fluid_outer
        node, tag = np.empty(n, np.int64), np.empty(n, np.uint32)
        fin, fout = np.empty(256, np.uint32), np.empty(256, np.uint32)
        _check(self.lib.oracle_get_bcfluid_table(self.h, group, as_ptr(node, C.c_int64), as_ptr(tag, C.c_uint32),
                                                 as_ptr(fin, C.c_uint32), as_ptr(fout, C.c_uint32)),
               "oracle_get_bcfluid_table")
        return dict(node=node, tag=tag, in_mask=fin, out_mask=fout)


def num_threads():
    return int(lib().oracle_num_threads())


# ---- per-node building blocks (double precision)

def pop_reconstruction(m10):
    m = np.ascontiguousarray(m10, np.float64)
    f = np.empty(27, np.float64)
    lib().oracle_pop_reconstruction(as_ptr(m, C.c_double), as_ptr(f, C.c_double))
    return f


def eval_moments(f27):
    f = np.ascontiguousarray(f27, np.float64)
    m = np.empty(10, np.float64)
    lib().oracle_eval_moments(as_ptr(f, C.c_double), as_ptr(m, C.c_double))
    return m


def collide(omega, m10):
    m = np.array(m10, np.float64)
    lib().oracle_collide(omega, as_ptr(m, C.c_double))
    return m


def node_bc(node_type, f27, old10, old10_xm1=None, tag=0, in_mask=0, omega=1.0, u_in=0.0):
    f = np.ascontiguousarray(f27, np.float64)
    o = np.ascontiguousarray(old10, np.float64)
    ox = np.ascontiguousarray(old10 if old10_xm1 is None else old10_xm1, np.float64)
    out = np.empty(10, np.float64)
    hit = lib().oracle_node_bc(node_type, tag, in_mask, omega, u_in, as_ptr(f, C.c_double), as_ptr(o, C.c_double),
                               as_ptr(ox, C.c_double), as_ptr(out, C.c_double))
    return out, bool(hit)


def incoming_rotated(nx, ny, in_mask, f27):
    f = np.ascontiguousarray(f27, np.float64)
    rho = C.c_double()
    m6 = np.empty(6, np.float64)
    lib().oracle_incoming_rotated(nx, ny, in_mask, as_ptr(f, C.c_double), C.byref(rho), as_ptr(m6, C.c_double))
    return rho.value, m6


def surface_pressure_extrapolation(xc, yc, xy8, rho3):
    a = np.ascontiguousarray(xy8, np.float64)
    r = np.ascontiguousarray(rho3, np.float64)
    return lib().oracle_surface_pressure_extrapolation(xc, yc, as_ptr(a, C.c_double), as_ptr(r, C.c_double))


def extrapolation_coeffs(delta):
    c = np.empty(3, np.float64)
    lib().oracle_extrapolation_coeffs(delta, as_ptr(c, C.c_double))
    return c


def curved_solution_strong(nx, ny, in_mask, out_mask, omega, prime7, rho_i, mxy_i, mxz_i):
    p = np.ascontiguousarray(prime7, np.float64)
    out = np.empty(10, np.float64)
    lib().oracle_curved_solution_strong(nx, ny, in_mask, out_mask, omega, as_ptr(p, C.c_double), rho_i, mxy_i, mxz_i,
                                        as_ptr(out, C.c_double))
    return out


def bcfluid(tag, in_mask, omega, strong, f27):
    f = np.ascontiguousarray(f27, np.float64)
    out = np.empty(10, np.float64)
    ok = lib().oracle_bcfluid(tag, in_mask, omega, int(strong), as_ptr(f, C.c_double), as_ptr(out, C.c_double))
    return out if ok else None


def curved_solution(nx, ny, in_mask, prime7, rho_i, mxy_i, mxz_i):
    p = np.ascontiguousarray(prime7, np.float64)
    out = np.empty(10, np.float64)
    lib().oracle_curved_solution(nx, ny, in_mask, as_ptr(p, C.c_double), rho_i, mxy_i, mxz_i, as_ptr(out, C.c_double))
    return out


# ---- VTK output

def vti_points(layout, shape, rho0, xc, yc, vel_norm, rho_dev, ux, uy, uz, solid):
    """Point arrays (4, nz, ny, nx) float32 of the reference writer from FP64 moments."""
    nz, ny, nx = shape
    arrs = [np.ascontiguousarray(a, np.float64).reshape(-1) for a in (rho_dev, ux, uy, uz)]
    sol = np.ascontiguousarray(solid, np.uint8).reshape(-1)
    out = np.empty(4 * nx * ny * nz, np.float32)
    lib().oracle_vti_points(layout, nx, ny, nz, rho0, xc, yc, vel_norm, *(as_ptr(a, C.c_double) for a in arrs),
                            as_ptr(sol, C.c_uint8), as_ptr(out, C.c_float))
    return out.reshape((4,) + tuple(shape))


def write_vti(path, layout, points):
    """Write points (4, nz, ny, nx) as the reference writer's .vti file."""
    p = np.ascontiguousarray(points, np.float32)
    _, nz, ny, nx = p.shape
    _check(lib().oracle_write_vti(os.fsencode(path), layout, nx, ny, nz, as_ptr(p, C.c_float)), "oracle_write_vti")


def write_master_pvd(path, max_iter, save_every):
    _check(lib().oracle_write_master_pvd(os.fsencode(path), max_iter, save_every), "oracle_write_master_pvd")
