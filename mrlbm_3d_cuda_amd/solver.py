"""Python side of the C ABI: the reference driver's Simulation, case setup and step.

Mirrors the reference's host flow (src/main.cu:7-40, src/simulation.cu:13-172):

    sim = Simulation(configs.c1())     # setup_case + initialize_simulation + case initialize
    sim.init()                         # gpu_initialize_Moments_GhostInterface
    sim.step(100, forces=True)         # streaming, boundary_treatment, post_streaming,
                                       # collision, post_collision (x100)
    rho = sim.field("rho")             # copyMomentsDeviceToHost
    F = sim.forces(0, 100)             # d_TotalFx/Fy/Fz/m per step

All compute runs in the HIP library libmrlbm.so built from csrc/; there is no CPU fallback. If
the library is missing this module raises at import of the first simulation.
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import CaseDesc, Info, TuneResult, as_ptr

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmrlbm.so")

_lib = None


class MrlbmError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"mrlbm status {status}: {message}")
        self.status = status


def lib():
    """Load libmrlbm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MrlbmError(_abi.ERR_DEVICE, f"{LIB_PATH} is missing: run python -m mrlbm_3d_cuda_amd.build")
        _lib = _abi.bind(C.CDLL(LIB_PATH), _abi.PRODUCT_PROTOTYPES)
        _abi.bind(_lib, _abi.BATCH_PROTOTYPES)
        _abi.bind(_lib, {"mrlbm_create_geometry_only": (C.c_int, [C.POINTER(CaseDesc), C.POINTER(C.c_void_p)])})
    return _lib


def _check(rc):
    if rc != 0:
        raise MrlbmError(rc, lib().mrlbm_last_error().decode(errors="replace"))


def default_desc(kind):
    d = CaseDesc()
    _check(lib().mrlbm_case_desc_default(kind, C.byref(d)))
    return d


class Simulation:
    """One case on the GPU, owned by the HIP library."""

    def __init__(self, desc: CaseDesc, geometry_only=False, _handle=None):
        self.lib = lib()
        self.desc = desc.copy()
        self.owned = _handle is None
        if _handle is None:
            h = C.c_void_p()
            fn = self.lib.mrlbm_create_geometry_only if geometry_only else self.lib.mrlbm_create
            _check(fn(C.byref(self.desc), C.byref(h)))
        else:
            h = _handle
        self.h = h
        self.shape = (desc.nz, desc.ny, desc.nx)
        self.n = desc.num_nodes

    def close(self):
        if getattr(self, "h", None):
            if self.owned:
                self.lib.mrlbm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- time stepping
    def init(self):
        _check(self.lib.mrlbm_init(self.h))
        return self

    def step(self, n=1, forces=False, stats=False):
        flags = (_abi.STEP_FORCES if forces else 0) | (_abi.STEP_STATS if stats else 0)
        _check(self.lib.mrlbm_step(self.h, n, flags))
        return self

    def stats_reset(self):
        _check(self.lib.mrlbm_stats_reset(self.h))

    def surface_pressure(self):
        """(mean wall pressure per inner-wall node in wall-table order, steps averaged)."""
        out = np.empty(self.info().nb_inner, np.float64)
        n = C.c_int64()
        _check(self.lib.mrlbm_get_surface_pressure(self.h, as_ptr(out, C.c_double), C.byref(n)))
        return out, n.value

    def sync(self):
        _check(self.lib.mrlbm_sync(self.h))

    def time_steps(self, n, forces=False, per_kernel=True):
        """Device time of n steps (ms). With per_kernel, steps are queued one by one with HIP
        events around each launch and the mean per-launch ms of the fused and wall kernels is
        returned too; without, the steps run as captured graph replays (as mrlbm_step does)."""
        total = C.c_double()
        per = (C.c_double * 2)()
        _check(self.lib.mrlbm_time_steps(self.h, n, _abi.STEP_FORCES if forces else 0, C.byref(total),
                                         per if per_kernel else None))
        return total.value, ((per[0], per[1]) if per_kernel else None)

    def time_replays(self, n, forces=False):
        """Device time of n steps run as graph replays captured beforehand: (ms for all n steps,
        median ms per step over the replays, number of replays)."""
        total, med, reps = C.c_double(), C.c_double(), C.c_int32()
        _check(self.lib.mrlbm_time_replays(self.h, n, _abi.STEP_FORCES if forces else 0, C.byref(total),
                                           C.byref(med), C.byref(reps)))
        return total.value, med.value, reps.value

    def time_memory_twin(self, launches=20):
        """Mean ms of one launch of the fused pass's memory-only twin (the access pattern's floor)."""
        ms = C.c_double()
        _check(self.lib.mrlbm_time_memory_twin(self.h, launches, C.byref(ms)))
        return ms.value

    # -- state
    def field(self, f):
        fi = _abi.FIELDS.index(f) if isinstance(f, str) else int(f)
        out = np.empty(self.n, dtype=np.float64)
        _check(self.lib.mrlbm_get_moments(self.h, fi, as_ptr(out, C.c_double)))
        return out.reshape(self.shape)

    def fields(self):
        return np.stack([self.field(f) for f in range(_abi.NUM_FIELDS)])

    def field_range(self, f, first, count):
        """Field f of the nodes first .. first + count - 1 (x-fastest index) through the library's
        fixed staging array (no device allocation; count <= _abi.RANGE_MAX_VALUES)."""
        fi = _abi.FIELDS.index(f) if isinstance(f, str) else int(f)
        out = np.empty(count, dtype=np.float64)
        _check(self.lib.mrlbm_get_moments_range(self.h, fi, first, count, as_ptr(out, C.c_double)))
        return out

    def set_field(self, f, values):
        fi = _abi.FIELDS.index(f) if isinstance(f, str) else int(f)
        arr = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        assert arr.size == self.n
        _check(self.lib.mrlbm_set_moments(self.h, fi, as_ptr(arr, C.c_double)))

    def forces(self, first, count):
        out = np.empty((count, 4), dtype=np.float64)
        _check(self.lib.mrlbm_get_forces(self.h, first, count, as_ptr(out, C.c_double)))
        return out

    def checkpoint_write(self, path):
        _check(self.lib.mrlbm_checkpoint_write(self.h, os.fsencode(path)))

    def checkpoint_read(self, path):
        _check(self.lib.mrlbm_checkpoint_read(self.h, os.fsencode(path)))

    def checkpoint_import_reference(self, path, real_bytes=4):
        """Load a restart.bin written by the reference (block-tiled moments, `real` of real_bytes)."""
        _check(self.lib.mrlbm_checkpoint_import_reference(self.h, os.fsencode(path), real_bytes))

    # -- output
    def default_vti_layout(self):
        return _abi.VTI_ANNULUS if self.desc.kind == _abi.CASE_ANNULUS else _abi.VTI_CARTESIAN

    def vti_points(self, layout=None, vel_norm=0.0):
        """The four float32 point arrays of a VTK output file, shape (4, nz, ny, nx); names in
        _abi.VTI_FIELDS[layout]. vel_norm 0 selects the case's velocity norm."""
        layout = self.default_vti_layout() if layout is None else layout
        out = np.empty(4 * self.n, dtype=np.float32)
        _check(self.lib.mrlbm_vti_points(self.h, layout, vel_norm, as_ptr(out, C.c_float)))
        return out.reshape((4,) + self.shape)

    def write_vti(self, directory, timestep=None, layout=None, vel_norm=0.0):
        """Write directory/data<10000000 + timestep>.vti (timestep: steps done) and return its path."""
        layout = self.default_vti_layout() if layout is None else layout
        timestep = self.info().steps_done if timestep is None else timestep
        _check(self.lib.mrlbm_write_vti(self.h, os.fsencode(directory), timestep, layout, vel_norm))
        return os.path.join(directory, f"data{10000000 + timestep}.vti")

    def write_forces(self, directory, first, count):
        """forces.dat and mass_flux.dat of the records first .. first + count - 1."""
        _check(self.lib.mrlbm_write_forces(self.h, os.fsencode(directory), first, count))

    def plane_means(self, first, count):
        """(count, 2) float64: mean inlet density (x = 0) and mean u_x at x = nx - 2 per step."""
        out = np.empty((count, 2), dtype=np.float64)
        _check(self.lib.mrlbm_get_plane_means(self.h, first, count, as_ptr(out, C.c_double)))
        return out

    def write_inlet_density(self, directory, first, count):
        _check(self.lib.mrlbm_write_inlet_density(self.h, os.fsencode(directory), first, count))

    def write_pressure(self, directory):
        """pressure.dat: angle, z and mean wall pressure per inner-wall node."""
        _check(self.lib.mrlbm_write_pressure(self.h, os.fsencode(directory)))

    def info_text(self):
        n = C.c_int64()
        _check(self.lib.mrlbm_sim_info(self.h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        _check(self.lib.mrlbm_sim_info(self.h, buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    # -- geometry
    def info(self):
        i = Info()
        _check(self.lib.mrlbm_get_info(self.h, C.byref(i)))
        return i

    def set_geometry(self, node_types, walls=(), bcf_in=None, bcf_out=None):
        """Replace the geometry (mrlbm_set_geometry): node types in the reference encoding
        (type << 20 | index, x fastest), one dict per wall group with the keys of wall_table()
        plus xc, yc, radius, fluid_side (+1 fluid outside the circle, -1 inside) and u_wall (3 values),
        and the 256-entry BCFLUID mask tables. A device simulation restarts from init()."""
        types = np.ascontiguousarray(node_types, dtype=np.uint32).reshape(-1)
        assert types.size == self.n
        keep = [types]
        arr = (_abi.WallDesc * max(len(walls), 1))()
        for g, w in enumerate(walls):
            cols = {}
            for key, ct, dt in (("node", C.c_int64, np.int64), ("nx", C.c_double, np.float64),
                                ("ny", C.c_double, np.float64), ("delta", C.c_double, np.float64),
                                ("in_mask", C.c_uint32, np.uint32), ("out_mask", C.c_uint32, np.uint32)):
                a = np.ascontiguousarray(w[key], dtype=dt)
                keep.append(a)
                cols[key] = as_ptr(a, ct)
            arr[g] = _abi.WallDesc(len(w["node"]), 0, cols["node"], cols["nx"], cols["ny"], cols["delta"],
                                   cols["in_mask"], cols["out_mask"], float(w["xc"]), float(w["yc"]),
                                   float(w["radius"]), float(w["fluid_side"]),
                                   (C.c_double * 3)(*w["u_wall"]))
        masks = []
        for m in (bcf_in, bcf_out):
            if m is None:
                masks.append(None)
            else:
                a = np.ascontiguousarray(m, dtype=np.uint32)
                assert a.size == 256
                keep.append(a)
                masks.append(as_ptr(a, C.c_uint32))
        _check(self.lib.mrlbm_set_geometry(self.h, as_ptr(types, C.c_uint32), len(walls), arr, masks[0], masks[1]))
        return self

    def node_types(self):
        out = np.empty(self.n, dtype=np.uint32)
        _check(self.lib.mrlbm_get_node_types(self.h, as_ptr(out, C.c_uint32)))
        return out.reshape(self.shape)

    def wall_table(self, group):
        info = self.info()
        n = info.nb_inner if group == 0 else info.nb_outer
        node = np.empty(n, np.int64)
        nx, ny, delta = (np.empty(n, np.float64) for _ in range(3))
        fin, fout = np.empty(n, np.uint32), np.empty(n, np.uint32)
        _check(self.lib.mrlbm_get_wall_table(self.h, group, as_ptr(node, C.c_int64), as_ptr(nx, C.c_double),
                                             as_ptr(ny, C.c_double), as_ptr(delta, C.c_double),
                                             as_ptr(fin, C.c_uint32), as_ptr(fout, C.c_uint32)))
        return dict(node=node, nx=nx, ny=ny, delta=delta, in_mask=fin, out_mask=fout)

    def bcfluid_table(self, group):
        info = self.info()
        n = info.nb_fluid_inner if group == 0 else info.nb_fluid_outer
        node, tag = np.empty(n, np.int64), np.empty(n, np.uint32)
        fin, fout = np.empty(256, np.uint32), np.empty(256, np.uint32)
        _check(self.lib.mrlbm_get_bcfluid_table(self.h, group, as_ptr(node, C.c_int64), as_ptr(tag, C.c_uint32),
                                                as_ptr(fin, C.c_uint32), as_ptr(fout, C.c_uint32)))
        return dict(node=node, tag=tag, in_mask=fin, out_mask=fout)


def write_master_pvd(directory, max_iter, save_every):
    """Write directory/master.pvd listing data<10000000 + it>.vti for it = 0, save_every, ... <= max_iter."""
    _check(lib().mrlbm_write_master_pvd(os.fsencode(directory), max_iter, save_every))
    return os.path.join(directory, "master.pvd")


class Batch:
    """Many cases of one precision stepped together (one fused and one wall launch per step).

    `cases` holds a Simulation view of every admitted case: fields, forces and tables read as for
    a standalone case; the batch steps them."""

    def __init__(self, descs, hbm_budget_bytes=0):
        self.lib = lib()
        arr = (CaseDesc * len(descs))(*[d.copy() for d in descs])
        h, adm = C.c_void_p(), C.c_int32()
        _check(self.lib.mrlbm_batch_create(arr, len(descs), hbm_budget_bytes, C.byref(h), C.byref(adm)))
        self.h = h
        self.admitted = adm.value
        self.cases = []
        for i in range(self.admitted):
            ch = C.c_void_p()
            _check(self.lib.mrlbm_batch_case(self.h, i, C.byref(ch)))
            self.cases.append(Simulation(descs[i], _handle=ch))
        self.num_nodes = sum(d.num_nodes for d in descs[:self.admitted])

    @staticmethod
    def plan(descs, hbm_budget_bytes):
        """Host-only admission: (admitted cases, device bytes) within the budget."""
        arr = (CaseDesc * len(descs))(*[d.copy() for d in descs])
        adm, nbytes = C.c_int32(), C.c_uint64()
        _check(lib().mrlbm_batch_plan(arr, len(descs), hbm_budget_bytes, C.byref(adm), C.byref(nbytes)))
        return adm.value, nbytes.value

    def close(self):
        if getattr(self, "h", None):
            for c in self.cases:
                c.h = None
            self.lib.mrlbm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def init(self):
        _check(self.lib.mrlbm_batch_init(self.h))
        return self

    def step(self, n=1, forces=False, stats=False):
        flags = (_abi.STEP_FORCES if forces else 0) | (_abi.STEP_STATS if stats else 0)
        _check(self.lib.mrlbm_batch_step(self.h, n, flags))
        return self

    def sync(self):
        _check(self.lib.mrlbm_batch_sync(self.h))

    def time_steps(self, n, forces=False):
        ms = C.c_double()
        _check(self.lib.mrlbm_batch_time_steps(self.h, n, _abi.STEP_FORCES if forces else 0, C.byref(ms)))
        return ms.value

    def autotune(self, trial_steps=10):
        r = TuneResult()
        _check(self.lib.mrlbm_batch_autotune(self.h, trial_steps, C.byref(r)))
        return {k: getattr(r, k) for k, _ in r._fields_}

    def launch(self):
        r = TuneResult()
        _check(self.lib.mrlbm_batch_get_launch(self.h, C.byref(r)))
        return {"bx": r.bx, "by": r.by, "variant": r.variant, "lz": r.lz, "blocks_per_cu": r.blocks_per_cu,
                "packing": r.packing, "slab_axis": r.slab_axis, "lz_special": r.lz_special}


def algorithmic_bytes_per_update(precision):
    """Read 10 moments + write 10 moments (B_alg)."""
    return 20 * _abi.STORAGE_BYTES[precision]
