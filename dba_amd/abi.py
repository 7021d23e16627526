"""ctypes mirror of include/dba/dba.h (the C ABI of libdba_hip.so).

The library is the product path; this module only binds it. Loading fails
loudly (ImportError/OSError) when the HIP library has not been built: there is
no CPU fallback behind this interface.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdba_hip.so")

# status codes
OK = 0
ERR_INVALID = -1
ERR_HIP = -2
ERR_LAUNCH = -3
ERR_NOMEM = -4
ERR_NOTFOUND = -5
ERR_NUMERIC = -6

# models
MODEL_DENDRITE_K94 = 1
MODEL_DENDRITE_K93 = 2
MODEL_MODELB = 3
MODEL_CH = 4
MODEL_KH = 5

# modes
MODE_REGULAR = 0
MODE_DBA = 1
MODE_CONSERVED = 2

# boundary conditions
BC_PERIODIC = 0
BC_NOFLUX = 1
BC_WALL = 2

DT_FIXED = 0
DT_CFL = 1
KH_METRIC_ALLGRAD = 0
KH_METRIC_DENSITY = 1

# field ids (dba_upload / dba_download)
F_PHI, F_T = 0, 1                      # dendrite
F_C = 0                                # CH
F_X, F_Y, F_Z, F_MBPHI = 0, 1, 2, 3    # Model B
F_MASS, F_MOMX, F_MOMY, F_ENERGY = 0, 1, 2, 3  # KH

SET_MAIN = 0
SET_T = 1


class Params(C.Structure):
    """dba_params (include/dba/dba.h)."""

    _fields_ = [
        ("model", C.c_int32),
        ("mode", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("tile_x", C.c_int32), ("tile_y", C.c_int32), ("tile_z", C.c_int32),
        ("bc_x", C.c_int32), ("bc_y", C.c_int32), ("bc_z", C.c_int32),
        ("stencil", C.c_int32),
        ("aniso_j", C.c_int32),
        ("dt_mode", C.c_int32),
        ("metric_kind", C.c_int32),
        ("threads", C.c_int32),
        ("history_cap", C.c_int32),
        ("reserved_i", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("thresh_a", C.c_double),
        ("thresh_b", C.c_double),
        ("tau", C.c_double), ("eps_bar", C.c_double), ("delta", C.c_double), ("alpha", C.c_double),
        ("gamma", C.c_double), ("theta0", C.c_double), ("latent_k", C.c_double), ("t_eq", C.c_double),
        ("mobility", C.c_double), ("kappa", C.c_double), ("react_gamma", C.c_double), ("cbar", C.c_double),
        ("chi_xy", C.c_double), ("chi_xz", C.c_double), ("chi_yz", C.c_double),
        ("chi_xb", C.c_double), ("chi_yb", C.c_double), ("chi_zb", C.c_double),
        ("v_x", C.c_double), ("v_y", C.c_double), ("v_z", C.c_double),
        ("r_x", C.c_double), ("r_y", C.c_double), ("r_z", C.c_double),
        ("n_exp", C.c_double), ("m_exp", C.c_double),
        ("epsx_sq", C.c_double), ("epsy_sq", C.c_double), ("epsz_sq", C.c_double), ("epsphi_sq", C.c_double),
        ("mob_x", C.c_double), ("mob_y", C.c_double), ("mob_z", C.c_double), ("mob_phi", C.c_double),
        ("k_eq", C.c_double), ("z_crit", C.c_double), ("p_gel", C.c_double), ("k_0", C.c_double),
        ("gas_gamma", C.c_double), ("cfl", C.c_double),
    ]

    def copy(self):
        q = Params()
        C.pointer(q)[0] = self
        return q

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Info(C.Structure):
    _fields_ = [
        ("steps_done", C.c_int64),
        ("ntiles", C.c_int32),
        ("ntiles_x", C.c_int32), ("ntiles_y", C.c_int32), ("ntiles_z", C.c_int32),
        ("sim_time", C.c_double),
        ("last_dt", C.c_double),
        ("cons_offset", C.c_double),
        ("footprint_bytes", C.c_int64),
        ("cell_updates", C.c_int64),
        ("single_launch", C.c_int32),
        ("reserved", C.c_int32),
        ("field_updates", C.c_int64),
    ]


# (name, restype, argtypes) of every exported entry point, in header order.
SIGNATURES = [
    ("dba_init", C.c_int32, [C.c_int32, C.POINTER(C.c_void_p)]),
    ("dba_finalize", C.c_int32, [C.c_void_p]),
    ("dba_last_error", C.c_char_p, []),
    ("dba_footprint", C.c_int32, [C.POINTER(Params), C.POINTER(C.c_int64)]),
    ("dba_device_memory", C.c_int32, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("dba_instance_create", C.c_int32, [C.c_void_p, C.POINTER(Params), C.POINTER(C.c_int32)]),
    ("dba_instance_destroy", C.c_int32, [C.c_void_p, C.c_int32]),
    ("dba_instance_info", C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(Info)]),
    ("dba_upload", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_size_t]),
    ("dba_download", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_size_t]),
    ("dba_step", C.c_int32, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.c_int32]),
    ("dba_step_graph", C.c_int32, [C.c_void_p, C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_int32]),
    ("dba_sync", C.c_int32, [C.c_void_p]),
    ("dba_active_tiles", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.c_void_p]),
    ("dba_worklist", C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    ("dba_active_history", C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    ("dba_timing_enable", C.c_int32, [C.c_void_p, C.c_int32]),
    ("dba_timing_query", C.c_int32, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                     C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    ("dba_marker", C.c_int32, [C.c_void_p, C.c_int32]),
    ("dba_marker_elapsed", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    ("dba_check_finite", C.c_int32, [C.c_void_p, C.c_int32]),
    ("dba_field_sum", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    ("dba_growth_front", C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p]),
    ("dba_kh_vy_mode", C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    ("dba_set_single_launch", C.c_int32, [C.c_void_p, C.c_int32]),
    ("dba_droplets", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int64), C.c_void_p, C.c_int64]),
    ("dba_droplets_noalloc", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_double),
                                         C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_void_p, C.c_int64]),
]

_lib = None


def load(path=None):
    """Load libdba_hip.so and bind every entry point (raises if absent)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(
            f"{p} is missing: the HIP library must be built first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(p)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class DbaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dba error {code}: {msg}")
        self.code = code


def check(code):
    if code != OK:
        raise DbaError(code, load().dba_last_error().decode(errors="replace"))
    return code
