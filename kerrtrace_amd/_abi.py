"""ctypes mirror of the structs in include/kerrtrace.h.

The field order and types must match the C declarations exactly; the
layout test in tests/test_abi.py compares ``ctypes.sizeof`` and field
offsets against the values the shared library reports.
"""
import ctypes as C

KT_OK = 0
KT_ERR_INVALID_CONFIG = -1
KT_ERR_DEVICE = -2
KT_ERR_LAUNCH = -3
KT_ERR_LUT_OVERFLOW = -4
KT_ERR_QUEUE_ORDER = -5
KT_ERR_INVALID_ARGUMENT = -6
KT_ERR_CAPACITY = -7
KT_ADJOINT_FP64 = 0
KT_ADJOINT_FP32 = 1

KT_PRECISION_FP64 = 0
KT_PRECISION_MIXED = 1
KT_DISK_ROOT_SECANT_REF = 0
KT_DISK_ROOT_TRUE = 1
KT_RHS_REFERENCE = 0
KT_RHS_HAMILTONIAN = 1

KT_END_OTHER = 0
KT_END_HORIZON = 1
KT_END_ESCAPE = 2
KT_FLAG_DISK = 4
KT_FLAG_SKY_GRID = 8
KT_FLAG_LUT_OVERFLOW = 16
KT_DISK_COUNT_SHIFT = 5


class KtConfig(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("spp", C.c_uint32),
        ("sky_grid_divisions", C.c_uint32),
        ("exposure_scale", C.c_double),
        ("sky_line_thickness", C.c_double),
        ("sky_intensity", C.c_double),
        ("spin", C.c_double),
        ("mass", C.c_double),
        ("disk_outer_radius", C.c_double),
        ("disk_temperature_scale", C.c_double),
        ("initial_step", C.c_double),
        ("tolerance", C.c_double),
        ("max_steps", C.c_uint32),
        ("max_attempts", C.c_uint32),
        ("transmittance_cutoff", C.c_double),
        ("horizon_epsilon", C.c_double),
        ("escape_radius", C.c_double),
        ("lut_size", C.c_uint32),
        ("reserved0", C.c_uint32),
        ("lut_max_temp", C.c_double),
        ("wavelength_start", C.c_double),
        ("wavelength_end", C.c_double),
        ("wavelength_step", C.c_double),
        ("noise_enabled", C.c_int32),
        ("noise_detail", C.c_uint32),
        ("noise_scale", C.c_double),
        ("noise_strength", C.c_double),
        ("noise_winding", C.c_double),
        ("noise_seed", C.c_uint64),
        ("bloom_enabled", C.c_int32),
        ("reserved1", C.c_int32),
        ("bloom_intensity", C.c_double),
        ("bloom_radius", C.c_double),
        ("bloom_threshold", C.c_double),
        ("tol_floor", C.c_double),
        ("disk_peak_temperature", C.c_double),
        ("mdot_scale", C.c_double),
        ("disk_root_mode", C.c_int32),
        ("precision", C.c_int32),
        ("rhs_mode", C.c_int32),
        ("reserved2", C.c_int32),
        ("mixed_min_tolerance", C.c_double),
    ]

    def copy(self):
        out = KtConfig()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(KtConfig))
        return out

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class KtCamera(C.Structure):
    _fields_ = [
        ("pos", C.c_double * 3),
        ("fwd", C.c_double * 3),
        ("rgt", C.c_double * 3),
        ("up", C.c_double * 3),
        ("fov_scale", C.c_double),
    ]

    def copy(self):
        out = KtCamera()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(KtCamera))
        return out


class KtDerived(C.Structure):
    _fields_ = [
        ("rh", C.c_double),
        ("disk_inner", C.c_double),
        ("disk_outer", C.c_double),
        ("temperature_scale", C.c_double),
        ("disk_lut_max", C.c_double),
        ("bloom_sigma", C.c_double),
        ("bloom_radius_int", C.c_int32),
        ("bloom_active", C.c_int32),
    ]


class KtTraceStats(C.Structure):
    _fields_ = [
        ("rays", C.c_uint64),
        ("rhs_evals", C.c_uint64),
        ("attempts", C.c_uint64),
        ("accepted", C.c_uint64),
        ("crossings", C.c_uint64),
        ("max_attempts_ray", C.c_uint64),
        ("lut_overflow", C.c_uint32),
        ("overflow_temperature", C.c_float),
        ("kernel_us", C.c_double),
        ("tail_us", C.c_double),
        ("kernel_event_ms", C.c_double),
    ]


class KtLaunchOpts(C.Structure):
    _fields_ = [
        ("waves_per_cu", C.c_int32),
        ("refill_threshold", C.c_int32),
        ("queue_chunk", C.c_int32),
        ("cost_order", C.c_int32),
        ("long_ray_attempts", C.c_int32),
        ("isolate_level", C.c_int32),
    ]


class KtFrameView(C.Structure):
    _fields_ = [
        ("index", C.c_int32),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("reserved", C.c_uint32),
        ("data", C.POINTER(C.c_uint32)),
    ]


class KtGradOpts(C.Structure):
    _fields_ = [
        ("fov_deg", C.c_double),
        ("checkpoint_interval", C.c_int32),
        ("log_steps_per_ray", C.c_int32),
        ("adjoint_precision", C.c_int32),
        ("reserved", C.c_int32),
    ]


class KtGradStats(C.Structure):
    _fields_ = [
        ("forward_ms", C.c_double),
        ("backward_ms", C.c_double),
        ("rhs_evals", C.c_uint64),
        ("attempts", C.c_uint64),
        ("chunks", C.c_uint32),
        ("log_overflow", C.c_uint32),
        ("checkpoint_interval", C.c_uint32),
        ("reserved", C.c_uint32),
    ]
