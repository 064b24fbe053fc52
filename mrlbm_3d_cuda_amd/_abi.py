"""ctypes mirror of include/mrlbm.h (structures, enums and prototypes).

The structures here must match the C header field for field; tests/test_abi.py checks the
sizes and offsets against the compiled library's own sizeof reports.
"""
import ctypes as C

ABI_VERSION = 2

# status codes
OK = 0
ERR_INVALID = 1
ERR_TAU = 2
ERR_MACH = 3
ERR_GEOMETRY = 4
ERR_DEVICE = 5
ERR_NOMEM = 6
ERR_IO = 7
ERR_CORRUPT = 8
ERR_NONFINITE = 9
ERR_UNSUPPORTED = 10

# case kinds
CASE_PERIODIC_BOX = 0
CASE_CYLINDER = 1
CASE_ANNULUS = 2

# precisions
FP64 = 0
FP32 = 1
FP32_BF16 = 2
FP32_FP16 = 3

PRECISION_NAMES = {FP64: "f64", FP32: "f32", FP32_BF16: "f32-bf16", FP32_FP16: "f32-fp16"}
STORAGE_BYTES = {FP64: 8, FP32: 4, FP32_BF16: 2, FP32_FP16: 2}

INLET_UNIFORM = 0
INLET_PARABOLIC = 1

FIELDS = ("rho", "ux", "uy", "uz", "mxx", "myy", "mzz", "mxy", "mxz", "myz")
NUM_FIELDS = 10

COMPAT_INIT_M_ZERO = 0x1
COMPAT_MASS_STRONG = 0x2
COMPAT_BCF_MASS_STRONG = 0x4
STEP_FORCES = 0x1
STEP_STATS = 0x2

VTI_CARTESIAN = 0  # rho, ux, uy, uz
VTI_ANNULUS = 1    # rho, uz, u_theta, u_radial
VTI_FIELDS = {VTI_CARTESIAN: ("rho", "ux", "uy", "uz"), VTI_ANNULUS: ("rho", "uz", "u_theta", "u_radial")}

# reference node types (common/nodeTypeMap.cuh:38-80)
NODE_TYPES = (
    "BULK", "SOLID", "INNER", "OUTER", "BCFLUID_INNER", "BCFLUID_OUTER", "BCSOLID_INNER",
    "BCSOLID_OUTER", "NORTH", "SOUTH", "EAST", "WEST", "FRONT", "BACK", "NORTH_EAST",
    "NORTH_WEST", "NORTH_FRONT", "NORTH_BACK", "SOUTH_EAST", "SOUTH_WEST", "SOUTH_FRONT",
    "SOUTH_BACK", "EAST_FRONT", "EAST_BACK", "WEST_FRONT", "WEST_BACK", "NORTH_WEST_FRONT",
    "NORTH_WEST_BACK", "SOUTH_WEST_FRONT", "SOUTH_WEST_BACK", "NORTH_EAST_FRONT",
    "NORTH_EAST_BACK", "SOUTH_EAST_FRONT", "SOUTH_EAST_BACK",
)
NODE = {name: i for i, name in enumerate(NODE_TYPES)}
TYPE_SHIFT = 20


class CaseDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("precision", C.c_int32),
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("periodic_x", C.c_int32),
        ("periodic_y", C.c_int32),
        ("periodic_z", C.c_int32),
        ("tau", C.c_double),
        ("rho0", C.c_double),
        ("u_max", C.c_double),
        ("inlet_profile", C.c_int32),
        ("compat_flags", C.c_uint32),
        ("inlet_noise", C.c_double),
        ("xc", C.c_double),
        ("yc", C.c_double),
        ("r_inner", C.c_double),
        ("r_outer", C.c_double),
        ("u_wall_inner", C.c_double * 3),
        ("u_wall_outer", C.c_double * 3),
        ("init_noise", C.c_double),
        ("seed", C.c_uint64),
    ]

    def copy(self):
        other = CaseDesc()
        C.memmove(C.byref(other), C.byref(self), C.sizeof(CaseDesc))
        return other

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out

    @property
    def num_nodes(self):
        return int(self.nx) * int(self.ny) * int(self.nz)


class Info(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int64),
        ("nb_inner", C.c_int32),
        ("nb_outer", C.c_int32),
        ("nb_fluid_inner", C.c_int32),
        ("nb_fluid_outer", C.c_int32),
        ("num_donors", C.c_int32),
        ("active_tiles", C.c_int32),
        ("total_tiles", C.c_int32),
        ("steps_done", C.c_int64),
        ("precision", C.c_int32),
        ("bytes_per_value", C.c_int32),
    ]


class TuneResult(C.Structure):
    _fields_ = [
        ("bx", C.c_int32),
        ("by", C.c_int32),
        ("variant", C.c_int32),
        ("lz", C.c_int32),
        ("candidates", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("ms_per_step", C.c_double),
        ("ms_per_step_default", C.c_double),
        ("packing", C.c_int32),
        ("slab_axis", C.c_int32),
        ("lz_special", C.c_int32),
    ]


class WallDesc(C.Structure):
    _fields_ = [
        ("nb", C.c_int32),
        ("reserved", C.c_int32),
        ("node", C.POINTER(C.c_int64)),
        ("nx", C.POINTER(C.c_double)),
        ("ny", C.POINTER(C.c_double)),
        ("delta", C.POINTER(C.c_double)),
        ("in_mask", C.POINTER(C.c_uint32)),
        ("out_mask", C.POINTER(C.c_uint32)),
        ("xc", C.c_double),
        ("yc", C.c_double),
        ("radius", C.c_double),
        ("fluid_side", C.c_double),
        ("u_wall", C.c_double * 3),
    ]


P = C.POINTER
_dp = P(C.c_double)
_u32p = P(C.c_uint32)
_i64p = P(C.c_int64)

# name -> (restype, argtypes); shared by the product library and the oracle ("mrlbm_" / "oracle_")
PRODUCT_PROTOTYPES = {
    "mrlbm_case_desc_default": (C.c_int, [C.c_int32, P(CaseDesc)]),
    "mrlbm_create": (C.c_int, [P(CaseDesc), P(C.c_void_p)]),
    "mrlbm_init": (C.c_int, [C.c_void_p]),
    "mrlbm_step": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32]),
    "mrlbm_sync": (C.c_int, [C.c_void_p]),
    "mrlbm_get_moments": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "mrlbm_set_moments": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "mrlbm_get_forces": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, _dp]),
    "mrlbm_get_node_types": (C.c_int, [C.c_void_p, _u32p]),
    "mrlbm_get_wall_table": (C.c_int, [C.c_void_p, C.c_int32, _i64p, _dp, _dp, _dp, _u32p, _u32p]),
    "mrlbm_get_bcfluid_table": (C.c_int, [C.c_void_p, C.c_int32, _i64p, _u32p, _u32p, _u32p]),
    "mrlbm_get_info": (C.c_int, [C.c_void_p, P(Info)]),
    "mrlbm_set_geometry": (C.c_int, [C.c_void_p, _u32p, C.c_int32, P(WallDesc), _u32p, _u32p]),
    "mrlbm_time_steps": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, _dp, _dp]),
    "mrlbm_time_replays": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, _dp, _dp, P(C.c_int32)]),
    "mrlbm_time_memory_twin": (C.c_int, [C.c_void_p, C.c_int32, _dp]),
    "mrlbm_checkpoint_write": (C.c_int, [C.c_void_p, C.c_char_p]),
    "mrlbm_checkpoint_read": (C.c_int, [C.c_void_p, C.c_char_p]),
    "mrlbm_destroy": (C.c_int, [C.c_void_p]),
    "mrlbm_stats_reset": (C.c_int, [C.c_void_p]),
    "mrlbm_get_surface_pressure": (C.c_int, [C.c_void_p, _dp, P(C.c_int64)]),
    "mrlbm_vti_points": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, P(C.c_float)]),
    "mrlbm_write_vti": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.c_int32, C.c_double]),
    "mrlbm_write_master_pvd": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32]),
    "mrlbm_sim_info": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64, P(C.c_int64)]),
    "mrlbm_write_forces": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64, C.c_int32]),
    "mrlbm_write_pressure": (C.c_int, [C.c_void_p, C.c_char_p]),
    "mrlbm_get_plane_means": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, _dp]),
    "mrlbm_checkpoint_import_reference": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32]),
    "mrlbm_write_inlet_density": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64, C.c_int32]),
    "mrlbm_last_error": (C.c_char_p, []),
    "mrlbm_abi_version": (C.c_int32, []),
    "mrlbm_get_moments_range": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_int64, _dp]),
}

RANGE_MAX_VALUES = 1 << 21

BATCH_PROTOTYPES = {
    "mrlbm_batch_plan": (C.c_int, [P(CaseDesc), C.c_int32, C.c_uint64, P(C.c_int32), P(C.c_uint64)]),
    "mrlbm_batch_create": (C.c_int, [P(CaseDesc), C.c_int32, C.c_uint64, P(C.c_void_p), P(C.c_int32)]),
    "mrlbm_batch_init": (C.c_int, [C.c_void_p]),
    "mrlbm_batch_step": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32]),
    "mrlbm_batch_sync": (C.c_int, [C.c_void_p]),
    "mrlbm_batch_num_cases": (C.c_int, [C.c_void_p, P(C.c_int32)]),
    "mrlbm_batch_case": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_void_p)]),
    "mrlbm_batch_time_steps": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, _dp]),
    "mrlbm_batch_autotune": (C.c_int, [C.c_void_p, C.c_int32, P(TuneResult)]),
    "mrlbm_batch_get_launch": (C.c_int, [C.c_void_p, P(TuneResult)]),
    "mrlbm_batch_destroy": (C.c_int, [C.c_void_p]),
}


def bind(lib, prototypes):
    for name, (res, args) in prototypes.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def as_ptr(arr, ctype):
    """Pointer to a C-contiguous numpy array's data."""
    assert arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(C.POINTER(ctype))
