"""ctypes wrapper over the CPU oracle (oracle/libvxoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py. The product path never imports this module.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libvxoracle.so")

F32 = 0
F64 = 1
IC0 = 0
BJ = 1
EMPTY = 0xFFFFFFFF

_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)


class _Scene(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint32),
        ("positions", _dp), ("rotations", _dp), ("velocities", _dp),
        ("mass", _dp), ("kinematic", _u8p),
        ("young", _dp), ("young_perp", _dp), ("shear", _dp), ("solid", _dp),
        ("fibers", _dp), ("anisotropy", _dp),
        ("delta_temp", _dp), ("alpha_par", _dp), ("alpha_perp", _dp),
        ("neighbor", _u32p), ("damage", _dp), ("kscale", _dp),
        ("formation_offset", _dp), ("plastic_strain", _dp),
    ]


class _StepConfig(ctypes.Structure):
    _fields_ = [
        ("fixed_dt", ctypes.c_double), ("substeps", ctypes.c_uint32), ("iterations", ctypes.c_uint32),
        ("corotation_iterations", ctypes.c_uint32), ("solver_relative_tolerance", ctypes.c_double),
        ("corotation_tolerance", ctypes.c_double), ("gravity", ctypes.c_double * 3),
        ("precond", ctypes.c_int), ("exact", ctypes.c_int), ("exact_tolerance", ctypes.c_double),
    ]


class _StepReport(ctypes.Structure):
    _fields_ = [("pcg_iterations", ctypes.c_uint32), ("passes", ctypes.c_uint32),
                ("solver_residual", ctypes.c_double), ("last_angle", ctypes.c_double)]


class _FractureConfig(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("damage_rate", "energy_release", "fatigue_cycles",
                                               "fatigue_exponent", "fatigue_threshold", "plastic_heat_fraction")]


class _Bond(ctypes.Structure):
    _fields_ = [("neighbor", ctypes.c_uint32), ("plastic_strain", ctypes.c_double * 3)] + [
        (k, ctypes.c_double) for k in ("damage", "equivalent_plastic_strain", "plastic_work", "fatigue_damage",
                                       "previous_stress_ratio", "reversal_stress_ratio", "loading_increment")]


class _Dissipation(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("plastic_heat", "stored_plastic", "surface_work", "damage_heat",
                                               "released")]


class _FractureState(ctypes.Structure):
    _fields_ = [("neighbor", _u32p), ("damage", _dp), ("plastic_strain", _dp),
                ("equivalent_plastic_strain", _dp), ("plastic_work", _dp), ("fatigue_damage", _dp),
                ("previous_stress_ratio", _dp), ("reversal_stress_ratio", _dp), ("loading_increment", _dp),
                ("strength", _dp), ("config", _FractureConfig), ("dissipated", _Dissipation),
                ("broken", ctypes.c_uint32)]


class _Tool(ctypes.Structure):
    _fields_ = [("n_cells", ctypes.c_uint32), ("has_mount", ctypes.c_uint32), ("cells", _dp), ("mass", ctypes.c_double),
                ("inverse_inertia", ctypes.c_double * 3), ("position", ctypes.c_double * 3),
                ("rotation", ctypes.c_double * 4), ("velocity", ctypes.c_double * 3),
                ("angular_velocity", ctypes.c_double * 3), ("external_force", ctypes.c_double * 3),
                ("external_torque", ctypes.c_double * 3), ("friction", ctypes.c_double),
                ("mount_local_base", ctypes.c_double * 3), ("mount_world_base", ctypes.c_double * 3),
                ("mount_direction", ctypes.c_double * 3), ("mount_rotation", ctypes.c_double * 4),
                ("friction_heat", ctypes.c_double), ("microsteps", ctypes.c_uint32), ("contacts", ctypes.c_uint32)]


HISTORY_FIELDS = ("equivalent_plastic_strain", "plastic_work", "fatigue_damage", "previous_stress_ratio",
                  "reversal_stress_ratio", "loading_increment")
BOND_FIELDS = ("damage",) + HISTORY_FIELDS

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library missing: build it with `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        L.vxo_build_topology.restype = ctypes.c_uint32
        L.vxo_build_topology.argtypes = [ctypes.c_uint32, _i32p, _u8p, _u32p]
        L.vxo_bond_frame.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_uint32, _dp]
        L.vxo_stiffness_matrix.argtypes = [ctypes.c_int, ctypes.POINTER(_Scene), ctypes.c_uint32, ctypes.c_uint32, _dp, _dp]
        L.vxo_bond_thermal_rest.argtypes = [ctypes.c_int, ctypes.POINTER(_Scene), ctypes.c_uint32, ctypes.c_uint32, _dp, _dp]
        L.vxo_fit_rotations.restype = ctypes.c_double
        L.vxo_fit_rotations.argtypes = [ctypes.c_int, ctypes.POINTER(_Scene)]
        L.vxo_assemble.restype = ctypes.c_void_p
        L.vxo_assemble.argtypes = [ctypes.c_int, ctypes.POINTER(_Scene), ctypes.c_double, _dp, _u8p, _dp]
        L.vxo_system_free.argtypes = [ctypes.c_void_p]
        L.vxo_system_active_count.restype = ctypes.c_uint32
        L.vxo_system_active_count.argtypes = [ctypes.c_void_p]
        L.vxo_system_constraint_count.restype = ctypes.c_uint32
        L.vxo_system_constraint_count.argtypes = [ctypes.c_void_p]
        L.vxo_system_export_rows.argtypes = [ctypes.c_void_p, _u32p, _u32p, _dp, _dp, _dp, _dp]
        L.vxo_system_export_full.argtypes = [ctypes.c_void_p, _dp, _dp, _u32p, _u8p]
        L.vxo_system_export_constraints.argtypes = [ctypes.c_void_p, _u32p, _dp, _dp, _dp, _dp]
        L.vxo_solve_cpu.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_uint32, _dp, _u32p, _dp, _dp]
        L.vxo_solve_device_semantics.restype = ctypes.c_int
        L.vxo_solve_device_semantics.argtypes = [ctypes.c_int, ctypes.c_uint32, _dp, _u32p, _dp, _dp, _dp,
                                                 ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, _dp, _u32p, _dp]
        L.vxo_solve_exact.restype = ctypes.c_double
        L.vxo_solve_exact.argtypes = [ctypes.c_uint32, _dp, _u32p, _dp, _dp, ctypes.c_double, ctypes.c_uint32, _dp, _u32p]
        L.vxo_solve_red_black.restype = ctypes.c_double
        L.vxo_solve_red_black.argtypes = [ctypes.c_uint32, _dp, _u32p, _dp, _dp, _u8p, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_uint32, _dp, _u32p, _dp]
        L.vxo_matvec.argtypes = [ctypes.c_int, ctypes.c_uint32, _u32p, _dp, _dp, _dp, _dp]
        L.vxo_step.argtypes = [ctypes.c_int, ctypes.POINTER(_Scene), ctypes.POINTER(_StepConfig), _dp, _dp,
                               ctypes.POINTER(_StepReport), _dp, _dp]
        L.vxo_pair_limits.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.vxo_fatigue.argtypes = [ctypes.c_int, ctypes.POINTER(_Bond), ctypes.c_double, ctypes.POINTER(_FractureConfig)]
        L.vxo_fracture_update.argtypes = [ctypes.c_int, ctypes.POINTER(_Bond), _dp, _dp, _dp, ctypes.c_double,
                                          ctypes.POINTER(_FractureConfig), ctypes.POINTER(_Dissipation)]
        L.vxo_tool_advance.argtypes = [ctypes.c_int, ctypes.POINTER(_Tool), ctypes.c_uint32, _dp, _dp, _dp, _u8p, _dp,
                                       ctypes.c_double, _dp, _dp]
        L.vxo_step_tool.argtypes = [ctypes.c_int, ctypes.POINTER(_Scene), ctypes.POINTER(_StepConfig), _dp, _dp,
                                    ctypes.POINTER(_Tool), ctypes.c_void_p, ctypes.POINTER(_StepReport)]
        L.vxo_step_fracture.argtypes = [ctypes.c_int, ctypes.POINTER(_Scene), ctypes.POINTER(_StepConfig), _dp, _dp,
                                        ctypes.POINTER(_FractureState), ctypes.POINTER(_StepReport), _dp, _dp]
        L.vxo_dot_rows.restype = ctypes.c_double
        L.vxo_dot_rows.argtypes = [ctypes.c_uint32, _dp, _dp]
        L.vxo_guarded_ratio.restype = ctypes.c_double
        L.vxo_guarded_ratio.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        _fp = ctypes.POINTER(ctypes.c_float)
        _u64p = ctypes.POINTER(ctypes.c_uint64)
        L.vxo_contact_project.restype = ctypes.c_uint64
        L.vxo_contact_project.argtypes = [ctypes.c_uint32, _fp, _fp, _fp, _fp, _u8p, _u8p, _u32p, ctypes.c_uint64,
                                          _u32p, _u32p, _fp, _u8p, _u64p]
        L.vxo_contact_candidates.restype = ctypes.c_uint64
        L.vxo_contact_candidates.argtypes = [ctypes.c_uint32, _fp, _fp, _u8p, _u32p, ctypes.c_uint64]
        L.vxo_local_continuum_pair.restype = ctypes.c_int
        L.vxo_local_continuum_pair.argtypes = [ctypes.c_uint32, _fp, _u32p, ctypes.c_uint32, ctypes.c_uint32]
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleScene:
    """float64 copies of a vxsim.scenes.Scene, kept alive with the struct."""

    def __init__(self, scene):
        self.src = scene
        self.positions = _f64(scene.positions).copy()
        self.rotations = _f64(scene.rotations).copy()
        self.velocities = _f64(scene.velocities).copy()
        self.mass = _f64(scene.mass)
        self.kinematic = np.ascontiguousarray(scene.kinematic, dtype=np.uint8)
        self.young = _f64(scene.young)
        self.young_perp = _f64(scene.young_perp)
        self.shear = _f64(scene.shear)
        self.solid = _f64(scene.solid)
        self.fibers = _f64(scene.fibers)
        self.anisotropy = _f64(scene.anisotropy)
        self.delta_temp = _f64(scene.delta_temp)
        self.alpha_par = _f64(scene.alpha_par)
        self.alpha_perp = _f64(scene.alpha_perp)
        # Topology, damage and plastic strain are copies: the fracture pass
        # writes them.
        self.neighbor = np.array(scene.neighbor, dtype=np.uint32, order="C")
        self.damage = np.array(scene.damage, dtype=np.float64, order="C")
        self.kscale = _f64(scene.kscale)
        self.formation_offset = _f64(scene.formation_offset)
        self.plastic_strain = np.array(scene.plastic_strain, dtype=np.float64, order="C")
        self.struct = _Scene(
            scene.n,
            _p(self.positions, _dp), _p(self.rotations, _dp), _p(self.velocities, _dp),
            _p(self.mass, _dp), _p(self.kinematic, _u8p),
            _p(self.young, _dp), _p(self.young_perp, _dp), _p(self.shear, _dp), _p(self.solid, _dp),
            _p(self.fibers, _dp), _p(self.anisotropy, _dp),
            _p(self.delta_temp, _dp), _p(self.alpha_par, _dp), _p(self.alpha_perp, _dp),
            _p(self.neighbor, _u32p), _p(self.damage, _dp), _p(self.kscale, _dp),
            _p(self.formation_offset, _dp), _p(self.plastic_strain, _dp),
        )

    @property
    def n(self):
        return self.src.n


def build_topology(keys, kinematic):
    keys = np.ascontiguousarray(keys, dtype=np.int32)
    kin = np.ascontiguousarray(kinematic, dtype=np.uint8)
    n = keys.shape[0]
    nb = np.empty((3, n), np.uint32)
    count = lib().vxo_build_topology(n, _p(keys, _i32p), _p(kin, _u8p), _p(nb, _u32p))
    return nb, int(count)


def bond_frame(qi, qj, axis, mode=F64):
    out = np.empty(9)
    lib().vxo_bond_frame(mode, _p(_f64(qi), _dp), _p(_f64(qj), _dp), axis, _p(out, _dp))
    return out.reshape(3, 3)


def stiffness_matrix(oscene: OracleScene, i, j, frame, mode=F64):
    out = np.empty(9)
    lib().vxo_stiffness_matrix(mode, ctypes.byref(oscene.struct), i, j, _p(_f64(frame).ravel(), _dp), _p(out, _dp))
    return out.reshape(3, 3)


def bond_thermal_rest(oscene: OracleScene, i, j, frame, mode=F64):
    out = np.empty(3)
    lib().vxo_bond_thermal_rest(mode, ctypes.byref(oscene.struct), i, j, _p(_f64(frame).ravel(), _dp), _p(out, _dp))
    return out


def fit_rotations(oscene: OracleScene, mode=F64) -> float:
    """Updates oscene.rotations in place; returns the largest angle change."""
    return float(lib().vxo_fit_rotations(mode, ctypes.byref(oscene.struct)))


class System:
    def __init__(self, oscene: OracleScene, mode, h, predicted, free_mask, loads):
        self.oscene = oscene
        self.mode = mode
        self._pred = _f64(predicted).ravel()
        self._free = np.ascontiguousarray(free_mask, dtype=np.uint8)
        self._loads = None if loads is None else _f64(loads).ravel()
        self.h = lib().vxo_assemble(mode, ctypes.byref(oscene.struct), h, _p(self._pred, _dp),
                                    _p(self._free, _u8p), _p(self._loads, _dp))
        self.na = int(lib().vxo_system_active_count(self.h))
        self.nc = int(lib().vxo_system_constraint_count(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().vxo_system_free(self.h)
            self.h = None

    def rows(self):
        na = self.na
        active = np.empty(na, np.uint32)
        nbr = np.empty((na, 6), np.uint32)
        blocks = np.empty((na, 54))
        diag = np.empty((na, 9))
        inv = np.empty((na, 9))
        rhs = np.empty((na, 3))
        lib().vxo_system_export_rows(self.h, _p(active, _u32p), _p(nbr, _u32p), _p(blocks, _dp), _p(diag, _dp),
                                     _p(inv, _dp), _p(rhs, _dp))
        return dict(active=active, neighbors=nbr, blocks=blocks, diagonal=diag, inverse=inv, rhs=rhs)

    def full(self):
        n = self.oscene.n
        diag = np.empty((n, 9))
        rhs = np.empty((n, 3))
        inc = np.empty((n, 6), np.uint32)
        inc_len = np.empty(n, np.uint8)
        lib().vxo_system_export_full(self.h, _p(diag, _dp), _p(rhs, _dp), _p(inc, _u32p), _p(inc_len, _u8p))
        return dict(diagonal=diag, rhs=rhs, incidence=inc, incidence_len=inc_len)

    def constraints(self):
        nc = self.nc
        ija = np.empty((nc, 3), np.uint32)
        frame = np.empty((nc, 9))
        mk = np.empty((nc, 9))
        k = np.empty((nc, 9))
        rest = np.empty((nc, 3))
        lib().vxo_system_export_constraints(self.h, _p(ija, _u32p), _p(frame, _dp), _p(mk, _dp), _p(k, _dp), _p(rest, _dp))
        return dict(ija=ija, frame=frame, material_k=mk, k=k, rest=rest)

    def solve_cpu(self, precond=BJ, max_iterations=80, tolerance=1e-5, threads=1, parallel_cutoff=32768):
        dx = np.empty((self.na, 3))
        it = ctypes.c_uint32()
        res = ctypes.c_double()
        sec = ctypes.c_double()
        lib().vxo_solve_cpu(self.h, precond, max_iterations, tolerance, threads, parallel_cutoff, _p(dx, _dp),
                            ctypes.byref(it), ctypes.byref(res), ctypes.byref(sec))
        return dx, int(it.value), float(res.value), float(sec.value)


def assemble(oscene: OracleScene, mode, h, predicted, free_mask, loads=None) -> System:
    return System(oscene, mode, h, predicted, free_mask, loads)


def solve_device_semantics(rows, max_iterations, tolerance, mode=F32, check_every=16):
    n = rows["rhs"].shape[0]
    dx = np.zeros((n, 3))
    it = ctypes.c_uint32()
    rel = ctypes.c_double()
    nbr = np.ascontiguousarray(rows["neighbors"], np.uint32)
    status = lib().vxo_solve_device_semantics(
        mode, n, _p(_f64(rows["rhs"]), _dp), _p(nbr, _u32p), _p(_f64(rows["diagonal"]), _dp),
        _p(_f64(rows["blocks"]), _dp), _p(_f64(rows["inverse"]), _dp), max_iterations, tolerance, check_every,
        _p(dx, _dp), ctypes.byref(it), ctypes.byref(rel))
    return int(status), dx, int(it.value), float(rel.value)


def solve_exact(rows, tolerance=1e-13, max_iterations=200000):
    n = rows["rhs"].shape[0]
    x = np.zeros((n, 3))
    it = ctypes.c_uint32()
    nbr = np.ascontiguousarray(rows["neighbors"], np.uint32)
    rel = lib().vxo_solve_exact(n, _p(_f64(rows["rhs"]), _dp), _p(nbr, _u32p), _p(_f64(rows["diagonal"]), _dp),
                                _p(_f64(rows["blocks"]), _dp), tolerance, max_iterations, _p(x, _dp),
                                ctypes.byref(it))
    return x, int(it.value), float(rel)


def black_rows(active, dims):
    """1 for the rows whose REV has odd x + y + z (the black colour of the
    red-black ordering), over the active REV indices of a box."""
    nx, ny, _ = dims
    a = np.asarray(active, np.int64)
    return (((a % nx) + (a // nx) % ny + a // (nx * ny)) & 1).astype(np.uint8)


def solve_red_black(rows, black, tolerance, inner_tolerance=0.0, max_iterations=100000, history=False):
    """FP64 CG on the red-black reduced system (vxo_solve_red_black):
    (x, iterations, full true relative residual, ||r|| history or None)."""
    n = rows["rhs"].shape[0]
    x = np.zeros((n, 3))
    it = ctypes.c_uint32()
    nbr = np.ascontiguousarray(rows["neighbors"], np.uint32)
    blk = np.ascontiguousarray(black, np.uint8)
    hist = np.zeros(max_iterations + 1) if history else None
    rel = lib().vxo_solve_red_black(n, _p(_f64(rows["rhs"]), _dp), _p(nbr, _u32p), _p(_f64(rows["diagonal"]), _dp),
                                    _p(_f64(rows["blocks"]), _dp), _p(blk, _u8p), tolerance, inner_tolerance,
                                    max_iterations, _p(x, _dp), ctypes.byref(it), _p(hist, _dp))
    k = int(it.value)
    return x, k, float(rel), (hist[:k + 1] if history else None)


def matvec(rows, x, mode=F64):
    n = rows["diagonal"].shape[0]
    y = np.empty((n, 3))
    nbr = np.ascontiguousarray(rows["neighbors"], np.uint32)
    lib().vxo_matvec(mode, n, _p(nbr, _u32p), _p(_f64(rows["diagonal"]), _dp), _p(_f64(rows["blocks"]), _dp),
                     _p(_f64(x), _dp), _p(y, _dp))
    return y


def dot_rows(a, b):
    a = _f64(a).ravel()
    b = _f64(b).ravel()
    return float(lib().vxo_dot_rows(a.size // 3, _p(a, _dp), _p(b, _dp)))


def guarded_ratio(num, den, mode=F32):
    return float(lib().vxo_guarded_ratio(mode, num, den))


def step(oscene: OracleScene, cfg, mode=F64, precond=BJ, exact=False, exact_tolerance=1e-12,
         targets=None, loads=None, want_forces=False):
    """One fixed step (solve part). cfg: vxsim.scenes.StepConfig."""
    c = _StepConfig(cfg.fixed_dt, cfg.substeps, cfg.iterations, cfg.corotation_iterations,
                    cfg.solver_relative_tolerance, cfg.corotation_tolerance,
                    (ctypes.c_double * 3)(*cfg.gravity), precond, 1 if exact else 0, exact_tolerance)
    rep = _StepReport()
    tg = None if targets is None else _f64(targets).ravel()
    ld = None if loads is None else _f64(loads).ravel()
    forces = np.zeros((3, oscene.n, 3)) if want_forces else None
    lam = np.zeros((3, oscene.n, 3)) if want_forces else None
    lib().vxo_step(mode, ctypes.byref(oscene.struct), ctypes.byref(c), _p(tg, _dp), _p(ld, _dp),
                   ctypes.byref(rep), _p(forces, _dp), _p(lam, _dp))
    out = dict(pcg_iterations=rep.pcg_iterations, passes=rep.passes, solver_residual=rep.solver_residual,
               last_angle=rep.last_angle)
    if want_forces:
        out["last_force"] = forces
        out["lambda"] = lam
    return out


# ---- fracture --------------------------------------------------------------
def _fcfg(c):
    return _FractureConfig(c.damage_rate, c.energy_release, c.fatigue_cycles, c.fatigue_exponent,
                           c.fatigue_threshold, c.plastic_heat_fraction)


def pair_limits(a, b, alignment, anisotropy, mode=F64):
    """pair_limits (src/fracture.rs:76-133) for two strength rows."""
    out = np.empty(8)
    lib().vxo_pair_limits(mode, _p(_f64(a), _dp), _p(_f64(b), _dp), _p(_f64(alignment), _dp),
                          _p(_f64(anisotropy), _dp), _p(out, _dp))
    return out


def bond(neighbor=1, **kw):
    b = _Bond()
    b.neighbor = neighbor
    for k, v in kw.items():
        if k == "plastic_strain":
            b.plastic_strain[:] = list(v)
        else:
            setattr(b, k, v)
    return b


def bond_dict(b):
    d = {k: getattr(b, k) for k, _ in _Bond._fields_ if k != "plastic_strain"}
    d["plastic_strain"] = np.array(b.plastic_strain[:])
    return d


def fatigue(b, signed_ratio, cfg, mode=F64):
    lib().vxo_fatigue(mode, ctypes.byref(b), float(signed_ratio), ctypes.byref(_fcfg(cfg)))
    return b


def fracture_update(b, u_total, k, limits, dt, cfg, mode=F64):
    """fracture::update (src/fracture.rs:210-292) on a _Bond in place; returns
    the dissipation dict."""
    d = _Dissipation()
    lib().vxo_fracture_update(mode, ctypes.byref(b), _p(_f64(u_total), _dp), _p(_f64(np.asarray(k)).ravel(), _dp),
                              _p(_f64(limits), _dp), float(dt), ctypes.byref(_fcfg(cfg)), ctypes.byref(d))
    return {k: getattr(d, k) for k, _ in _Dissipation._fields_}


class FractureHistory:
    """Bond history beyond the scene (3 x n each), zero for a new scene."""

    def __init__(self, n):
        for k in HISTORY_FIELDS:
            setattr(self, k, np.zeros((3, n)))


def step_fracture(oscene: OracleScene, cfg, fcfg, history: FractureHistory, strength, mode=F64, precond=BJ,
                  exact=False, exact_tolerance=1e-12, targets=None, loads=None, want_forces=False):
    """One fixed step (solve part) with the fracture pass of every substep
    (src/solver.rs:627-665). Updates oscene.neighbor / damage / plastic_strain
    and `history`; returns the step report with the broken-bond count and
    the dissipated energies of the step."""
    c = _StepConfig(cfg.fixed_dt, cfg.substeps, cfg.iterations, cfg.corotation_iterations,
                    cfg.solver_relative_tolerance, cfg.corotation_tolerance,
                    (ctypes.c_double * 3)(*cfg.gravity), precond, 1 if exact else 0, exact_tolerance)
    st = np.ascontiguousarray(strength, dtype=np.float64)
    fs = _FractureState(_p(oscene.neighbor, _u32p), _p(oscene.damage, _dp), _p(oscene.plastic_strain, _dp),
                        *[_p(getattr(history, k), _dp) for k in HISTORY_FIELDS], _p(st, _dp), _fcfg(fcfg),
                        _Dissipation(), 0)
    rep = _StepReport()
    tg = None if targets is None else _f64(targets).ravel()
    ld = None if loads is None else _f64(loads).ravel()
    forces = np.zeros((3, oscene.n, 3)) if want_forces else None
    lam = np.zeros((3, oscene.n, 3)) if want_forces else None
    lib().vxo_step_fracture(mode, ctypes.byref(oscene.struct), ctypes.byref(c), _p(tg, _dp), _p(ld, _dp),
                            ctypes.byref(fs), ctypes.byref(rep), _p(forces, _dp), _p(lam, _dp))
    out = dict(pcg_iterations=rep.pcg_iterations, passes=rep.passes, solver_residual=rep.solver_residual,
               last_angle=rep.last_angle, broken=fs.broken,
               dissipated={k: getattr(fs.dissipated, k) for k, _ in _Dissipation._fields_})
    if want_forces:
        out["last_force"] = forces
        out["lambda"] = lam
    return out



# ---- contact projection (src/collision.rs:33-123) ------------------------------

def _f32c(a):
    return np.ascontiguousarray(a, dtype=np.float32).reshape(-1).copy()


def _targets_bits(targets, n):
    if targets is None:
        return None
    t = np.asarray(targets, dtype=bool).reshape(n, 3)
    return np.ascontiguousarray(t[:, 0] * 1 + t[:, 1] * 2 + t[:, 2] * 4, dtype=np.uint8)


def contact_project(positions, previous, rest_positions, mass, kinematic, bond_neighbor, targets=None):
    """collision::project on the CPU: (positions, previous, contacts, stats), as vxsim.contact.project."""
    L = lib()
    _fp = ctypes.POINTER(ctypes.c_float)
    pos, prev, rest = _f32c(positions), _f32c(previous), _f32c(rest_positions)
    n = pos.size // 3
    m = _f32c(mass)
    kin = np.ascontiguousarray(np.asarray(kinematic).reshape(-1) != 0, dtype=np.uint8)
    bonds = np.ascontiguousarray(bond_neighbor, dtype=np.uint32).reshape(-1)
    tgt = _targets_bits(targets, n)
    cap = max(1, 6 * n)
    while True:
        p, q = pos.copy(), prev.copy()
        ci = np.zeros(cap, np.uint32)
        cj = np.zeros(cap, np.uint32)
        cn = np.zeros(3 * cap, np.float32)
        cimp = np.zeros(cap, np.uint8)
        stats = np.zeros(4, np.uint64)
        k = L.vxo_contact_project(n, p.ctypes.data_as(_fp), q.ctypes.data_as(_fp), rest.ctypes.data_as(_fp),
                                  m.ctypes.data_as(_fp), _p(kin, _u8p), _p(tgt, _u8p) if tgt is not None else None,
                                  _p(bonds, _u32p), cap, _p(ci, _u32p), _p(cj, _u32p), cn.ctypes.data_as(_fp),
                                  _p(cimp, _u8p), stats.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        if k <= cap:
            break
        cap = int(k)
    pos, prev = p, q
    contacts = {"i": ci[:k].copy(), "j": cj[:k].copy(), "normal": cn[:3 * k].reshape(k, 3).copy(),
                "new_impact": cimp[:k] != 0}
    st = {"candidate_pairs": int(stats[0]), "pairs": int(stats[1]), "passes": int(stats[2]),
          "contacts": int(stats[3])}
    return pos.reshape(n, 3), prev.reshape(n, 3), contacts, st


def contact_candidates(positions, previous, kinematic):
    """The broad phase alone: sorted (i, j) candidate pairs, shape (k, 2)."""
    L = lib()
    _fp = ctypes.POINTER(ctypes.c_float)
    pos, prev = _f32c(positions), _f32c(previous)
    n = pos.size // 3
    kin = np.ascontiguousarray(np.asarray(kinematic).reshape(-1) != 0, dtype=np.uint8)
    k = L.vxo_contact_candidates(n, pos.ctypes.data_as(_fp), prev.ctypes.data_as(_fp), _p(kin, _u8p), None, 0)
    out = np.zeros(2 * max(int(k), 1), np.uint32)
    L.vxo_contact_candidates(n, pos.ctypes.data_as(_fp), prev.ctypes.data_as(_fp), _p(kin, _u8p), _p(out, _u32p), k)
    return out[:2 * k].reshape(int(k), 2)


def local_continuum_pair(rest_positions, bond_neighbor, i, j):
    L = lib()
    rest = _f32c(rest_positions)
    bonds = np.ascontiguousarray(bond_neighbor, dtype=np.uint32).reshape(-1)
    return bool(L.vxo_local_continuum_pair(rest.size // 3, rest.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                           _p(bonds, _u32p), i, j))


# ---- rigid tool (src/tool.rs:192-431) -------------------------------------------

class OracleTool:
    """The oracle's copy of a tool's state (vxsim.tool.Tool fields)."""

    def __init__(self, tool):
        self.cells = np.ascontiguousarray(tool.cells, dtype=np.float64)
        t = _Tool()
        t.n_cells = len(self.cells)
        t.has_mount = 1 if tool.mount is not None else 0
        t.cells = _p(self.cells, _dp)
        t.mass = float(tool.mass)
        for name in ("inverse_inertia", "position", "rotation", "velocity", "angular_velocity", "external_force",
                     "external_torque"):
            arr = getattr(t, name)
            for k, x in enumerate(getattr(tool, name)):
                arr[k] = float(x)
        t.friction = float(tool.friction)
        if tool.mount is not None:
            lb, wb, d, rot = tool.mount
            for k in range(3):
                t.mount_local_base[k], t.mount_world_base[k], t.mount_direction[k] = lb[k], wb[k], d[k]
            for k in range(4):
                t.mount_rotation[k] = rot[k]
        self.struct = t

    def state(self) -> dict:
        t = self.struct
        return dict(position=np.array(t.position[:]), rotation=np.array(t.rotation[:]),
                    velocity=np.array(t.velocity[:]), angular_velocity=np.array(t.angular_velocity[:]),
                    friction_heat=t.friction_heat, microsteps=t.microsteps, contacts=t.contacts)


def tool_advance(tool: OracleTool, positions, velocities, mass, kinematic, targets, dt, gravity, mode=F64):
    """Tool::advance over the given scene state; velocities (n, 3) float64 are
    updated in place; returns the contact velocity (n, 3)."""
    pos = np.ascontiguousarray(positions, dtype=np.float64)
    n = pos.shape[0]
    assert velocities.dtype == np.float64 and velocities.flags.c_contiguous
    m = np.ascontiguousarray(mass, dtype=np.float64)
    kin = np.ascontiguousarray(kinematic, dtype=np.uint8)
    tg = None if targets is None else _f64(targets).ravel()
    g = np.ascontiguousarray(gravity, dtype=np.float64)
    cv = np.zeros((n, 3))
    lib().vxo_tool_advance(mode, ctypes.byref(tool.struct), n, _p(pos, _dp), _p(velocities, _dp), _p(m, _dp),
                           _p(kin, _u8p), _p(tg, _dp), float(dt), _p(g, _dp), _p(cv, _dp))
    return cv


def step_tool(oscene: OracleScene, cfg, tool: OracleTool, mode=F64, precond=BJ, exact=False, exact_tolerance=1e-12,
              targets=None, loads=None, fracture=None):
    """One fixed step (solve part) with the tool advancing each substep after
    the prediction; `fracture` is an optional (fcfg, history, strength)."""
    c = _StepConfig(cfg.fixed_dt, cfg.substeps, cfg.iterations, cfg.corotation_iterations,
                    cfg.solver_relative_tolerance, cfg.corotation_tolerance,
                    (ctypes.c_double * 3)(*cfg.gravity), precond, 1 if exact else 0, exact_tolerance)
    rep = _StepReport()
    tg = None if targets is None else _f64(targets).ravel()
    ld = None if loads is None else _f64(loads).ravel()
    fs = None
    keep = None
    if fracture is not None:
        fcfg, history, strength = fracture
        keep = np.ascontiguousarray(strength, dtype=np.float64)
        fs = _FractureState(_p(oscene.neighbor, _u32p), _p(oscene.damage, _dp), _p(oscene.plastic_strain, _dp),
                            *[_p(getattr(history, k), _dp) for k in HISTORY_FIELDS], _p(keep, _dp), _fcfg(fcfg),
                            _Dissipation(), 0)
    lib().vxo_step_tool(mode, ctypes.byref(oscene.struct), ctypes.byref(c), _p(tg, _dp), _p(ld, _dp),
                        ctypes.byref(tool.struct), ctypes.byref(fs) if fs is not None else None, ctypes.byref(rep))
    out = dict(pcg_iterations=rep.pcg_iterations, passes=rep.passes, solver_residual=rep.solver_residual,
               last_angle=rep.last_angle)
    if fs is not None:
        out["broken"] = fs.broken
    return out
