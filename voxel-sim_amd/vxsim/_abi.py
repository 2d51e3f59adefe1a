"""ctypes declarations of include/vx_solver.h."""
from __future__ import annotations

import ctypes

c_fp = ctypes.POINTER(ctypes.c_float)
c_dp = ctypes.POINTER(ctypes.c_double)
c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u8p = ctypes.POINTER(ctypes.c_uint8)

VX_OK = 0
VX_ERR_INIT = 1
VX_ERR_TOO_LARGE = 2
VX_ERR_NONFINITE_RESIDUAL = 3
VX_ERR_NOT_CONVERGED = 4
VX_ERR_NONFINITE_CORRECTION = 5
VX_ERR_INVALID_ARGUMENT = 6
VX_ERR_DEVICE = 7
VX_EMPTY = 0xFFFFFFFF
VX_LAUNCH_GRAPHS = 0
VX_LAUNCH_DIRECT = 1
# Launch shape XCD deals (vx_scene_set_shape)
VX_DEAL_AUTO = 0
VX_DEAL_CONTIGUOUS = 1
VX_DEAL_SPLIT = 2
VX_DEAL_ROUND_ROBIN = 3

VX_SOLVE_MIXED = 0
VX_SOLVE_F64 = 1
VX_PRECOND_BLOCK_JACOBI = 0
VX_PRECOND_RED_BLACK = 1
VX_PRECOND_MULTIGRID = 2

VX_SCENE_FREE = 0
VX_SCENE_RUNNING = 1
VX_SCENE_DONE = 2
VX_SCENE_FAILED = 3


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_uint32), ("ny", ctypes.c_uint32), ("nz", ctypes.c_uint32),
        ("positions", c_fp), ("rotations", c_fp), ("velocities", c_fp), ("mass", c_fp),
        ("kinematic", c_u8p), ("young", c_fp), ("young_perp", c_fp), ("shear", c_fp), ("solid", c_fp),
        ("fibers", c_fp), ("anisotropy", c_fp), ("delta_temp", c_fp), ("alpha_par", c_fp),
        ("alpha_perp", c_fp), ("neighbor", c_u32p), ("damage", c_fp), ("kscale", c_fp),
        ("formation_offset", c_fp), ("plastic_strain", c_fp), ("targets", c_fp), ("loads", c_fp),
        ("strength", c_fp), ("bond_history", c_fp),
    ]


class FractureParams(ctypes.Structure):
    _fields_ = [
        ("enabled", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("damage_rate", ctypes.c_double),
        ("energy_release", ctypes.c_double), ("fatigue_cycles", ctypes.c_double),
        ("fatigue_exponent", ctypes.c_double), ("fatigue_threshold", ctypes.c_double),
        ("plastic_heat_fraction", ctypes.c_double),
    ]


class FractureReport(ctypes.Structure):
    _fields_ = [
        ("bonds_broken", ctypes.c_uint64), ("plastic_heat", ctypes.c_double), ("stored_plastic", ctypes.c_double),
        ("surface_work", ctypes.c_double), ("damage_heat", ctypes.c_double), ("released", ctypes.c_double),
        ("seconds", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class SolveParams(ctypes.Structure):
    _fields_ = [
        ("tolerance", ctypes.c_double), ("inner_tolerance", ctypes.c_double),
        ("max_inner", ctypes.c_uint32), ("max_outer", ctypes.c_uint32),
        ("warm_start", ctypes.c_uint32), ("mode", ctypes.c_uint32),
        ("preconditioner", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
    ]


class StepParams(ctypes.Structure):
    pass


class ToolDesc(ctypes.Structure):
    _fields_ = [("n_cells", ctypes.c_uint32), ("has_mount", ctypes.c_uint32), ("cells", ctypes.POINTER(ctypes.c_float)),
                ("mass", ctypes.c_double), ("inverse_inertia", ctypes.c_double * 3), ("position", ctypes.c_double * 3),
                ("rotation", ctypes.c_double * 4), ("velocity", ctypes.c_double * 3),
                ("angular_velocity", ctypes.c_double * 3), ("external_force", ctypes.c_double * 3),
                ("external_torque", ctypes.c_double * 3), ("friction", ctypes.c_double),
                ("mount_local_base", ctypes.c_double * 3), ("mount_world_base", ctypes.c_double * 3),
                ("mount_direction", ctypes.c_double * 3), ("mount_rotation", ctypes.c_double * 4)]


class ToolState(ctypes.Structure):
    _fields_ = [("position", ctypes.c_double * 3), ("rotation", ctypes.c_double * 4), ("velocity", ctypes.c_double * 3),
                ("angular_velocity", ctypes.c_double * 3), ("friction_heat", ctypes.c_double),
                ("microsteps", ctypes.c_uint32), ("contacts", ctypes.c_uint32), ("candidates", ctypes.c_uint32),
                ("status", ctypes.c_uint32)]


class KernelStats(ctypes.Structure):
    _fields_ = [("dispatches", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("mean_us", ctypes.c_double),
                ("median_us", ctypes.c_double), ("min_us", ctypes.c_double), ("max_us", ctypes.c_double),
                ("total_us", ctypes.c_double)]


class KernelTimes(ctypes.Structure):
    _fields_ = [("spmv", KernelStats), ("update", KernelStats)]


class SolveReport(ctypes.Structure):
    _fields_ = [
        ("inner_iterations", ctypes.c_uint32), ("outer_sweeps", ctypes.c_uint32),
        ("active_rows", ctypes.c_uint32), ("status", ctypes.c_uint32),
        ("relative_residual", ctypes.c_double), ("rhs_norm", ctypes.c_double),
    ]


StepParams._fields_ = [
    ("fixed_dt", ctypes.c_double), ("gravity", ctypes.c_double * 3), ("substeps", ctypes.c_uint32),
    ("corotation_iterations", ctypes.c_uint32), ("corotation_tolerance", ctypes.c_double),
    ("solve", SolveParams), ("fracture", FractureParams),
]


class StepReport(ctypes.Structure):
    _fields_ = [
        ("solves", ctypes.c_uint32), ("inner_iterations", ctypes.c_uint32), ("outer_sweeps", ctypes.c_uint32),
        ("active_rows", ctypes.c_uint32), ("last_relative_residual", ctypes.c_double),
        ("last_angle", ctypes.c_double), ("seconds", ctypes.c_double), ("fracture", FractureReport),
    ]


class BatchSceneReport(ctypes.Structure):
    _fields_ = [
        ("state", ctypes.c_uint32), ("status", ctypes.c_uint32), ("steps_done", ctypes.c_uint32),
        ("solves", ctypes.c_uint32), ("inner_iterations", ctypes.c_uint64), ("outer_sweeps", ctypes.c_uint64),
        ("rows", ctypes.c_uint32), ("active_rows", ctypes.c_uint32), ("last_relative_residual", ctypes.c_double),
        ("last_angle", ctypes.c_double), ("fracture", FractureReport),
    ]


class Contact(ctypes.Structure):
    _fields_ = [("i", ctypes.c_uint32), ("j", ctypes.c_uint32), ("normal", ctypes.c_float * 3),
                ("new_impact", ctypes.c_uint32)]


class ContactReport(ctypes.Structure):
    _fields_ = [("candidate_pairs", ctypes.c_uint64), ("pairs", ctypes.c_uint64), ("contacts", ctypes.c_uint64),
                ("passes", ctypes.c_uint32), ("rounds", ctypes.c_uint32), ("device_us", ctypes.c_double),
                ("projection_us", ctypes.c_double), ("evaluated", ctypes.c_uint64), ("redone", ctypes.c_uint32),
                ("relaxed_passes", ctypes.c_uint32), ("relaxation_rounds", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


# (name, restype, argtypes) for every exported function of include/vx_solver.h.
SIGNATURES = [
    ("vx_last_error", ctypes.c_char_p, []),
    ("vx_init", ctypes.c_int32, []),
    ("vx_copy_bandwidth", ctypes.c_int32, [ctypes.c_uint64, ctypes.c_uint32, c_dp]),
    ("vx_set_launch_mode", ctypes.c_int32, [ctypes.c_int32]),
    ("vx_solve_pcg", ctypes.c_int32,
     [c_fp, c_u32p, c_fp, c_fp, c_fp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, c_fp, c_u32p, c_fp]),
    ("vx_solve_pcg_f64", ctypes.c_int32,
     [c_fp, c_u32p, c_fp, c_fp, c_fp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, c_fp, c_u32p, c_fp]),
    ("vx_solve_pcg_noalloc", ctypes.c_int32,
     [c_fp, c_u32p, c_fp, c_fp, c_fp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, c_fp, c_u32p, c_fp]),
    ("vx_scene_create", ctypes.c_int32, [ctypes.POINTER(SceneDesc), ctypes.POINTER(ctypes.c_void_p)]),
    ("vx_scene_destroy", ctypes.c_int32, [ctypes.c_void_p]),
    ("vx_scene_predict", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_double, c_dp]),
    ("vx_scene_fit_rotations", ctypes.c_int32, [ctypes.c_void_p, c_dp]),
    ("vx_scene_assemble", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_double]),
    ("vx_scene_solve", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.POINTER(SolveParams), ctypes.POINTER(SolveReport)]),
    ("vx_scene_apply", ctypes.c_int32, [ctypes.c_void_p]),
    ("vx_scene_finish_substep", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_double]),
    ("vx_scene_download", ctypes.c_int32, [ctypes.c_void_p, c_dp, c_dp, c_dp, c_dp]),
    ("vx_scene_export_rows", ctypes.c_int32, [ctypes.c_void_p, c_u32p, c_u32p, c_dp, c_dp, c_dp, c_u32p]),
    ("vx_scene_multigrid_level", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32, c_u32p, c_fp, c_fp, c_u8p]),
    ("vx_scene_time_spmv", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32, c_dp]),
    ("vx_scene_solve_timed", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.POINTER(SolveParams), ctypes.POINTER(SolveReport), ctypes.POINTER(KernelTimes)]),
    ("vx_scene_step", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.POINTER(StepParams), ctypes.POINTER(StepReport)]),
    ("vx_scene_set_schedule", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32]),
    ("vx_scene_set_shape", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    ("vx_batch_set_schedule", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    ("vx_batch_bytes_per_row", ctypes.c_uint64, []),
    ("vx_batch_create", ctypes.c_int32,
     [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]),
    ("vx_batch_destroy", ctypes.c_int32, [ctypes.c_void_p]),
    ("vx_batch_free_rows", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    ("vx_batch_admit", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.POINTER(SceneDesc), ctypes.POINTER(StepParams), ctypes.c_uint32, c_u32p]),
    ("vx_batch_tick", ctypes.c_int32, [ctypes.c_void_p, c_u32p]),
    ("vx_batch_report", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(BatchSceneReport)]),
    ("vx_batch_states", ctypes.c_int32, [ctypes.c_void_p, c_u32p]),
    ("vx_batch_download", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32, c_dp, c_dp, c_dp]),
    ("vx_batch_release", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32]),
    ("vx_batch_hold", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    ("vx_scene_download_bonds", ctypes.c_int32, [ctypes.c_void_p, c_u32p, c_fp, c_fp, c_fp]),
    ("vx_scene_set_targets_and_loads", ctypes.c_int32, [ctypes.c_void_p, c_fp, c_fp]),
    ("vx_batch_download_bonds", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32, c_u32p, c_fp, c_fp, c_fp]),
    ("vx_scene_download_bond_forces", ctypes.c_int32, [ctypes.c_void_p, c_dp, c_dp]),
    ("vx_scene_set_tool", ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(ToolDesc)]),
    ("vx_scene_tool_state", ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(ToolState)]),
    ("vx_batch_download_bond_forces", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_uint32, c_dp, c_dp]),
    ("vx_contact_set_method", ctypes.c_int32, [ctypes.c_int32, ctypes.c_uint32]),
    ("vx_contact_project", ctypes.c_int32,
     [ctypes.c_uint32, c_fp, c_fp, c_fp, c_fp, c_u8p, c_u8p, c_u32p, ctypes.POINTER(Contact), ctypes.c_uint64,
      ctypes.POINTER(ContactReport)]),
]


def declare(L: ctypes.CDLL) -> None:
    for name, restype, argtypes in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = restype
        fn.argtypes = argtypes
