"""Synthetic REV-lattice scenes for the corotational PCG solve.

The state follows the reference's scene data model (src/scene.rs:309-351,
src/voxel.rs:46-60, src/bond.rs:4-16) flattened into structure-of-arrays
planes of the reference's f32 values:

* voxel index v = x + nx*(y + ny*z) (the z, y, x loop order of Scene::new,
  src/scene.rs:1607-1619), centre = key*CELL + CELL/2 with key =
  (x - nx//2, y - ny//2, z) (src/scene.rs:1589-1611);
* bonds are stored per +axis (`neighbor[axis][v]`, EMPTY = 0xFFFFFFFF) and
  follow the lattice rule "both cells occupied and equal kinematic flag"
  (src/scene.rs:563, 1674). Voxels fixed after construction keep their bonds
  (Scene::fix_voxels, src/scene.rs:901-930); voxels spawned fixed get none.

The reference has no scene generator; the draws below (SplitMix64 streams)
define the synthetic workloads.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

CELL = np.float32(0.01)  # src/scene.rs:10
EMPTY = np.uint32(0xFFFFFFFF)  # src/bond.rs:2

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Counter form of SplitMix64 (Steele, Lea, Flood 2014): the k-th output is
    mix(seed + (k + 1) * golden), so blocks of draws vectorise."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed)
        self.count = 0

    def next_u64(self, size: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(self.count + 1, self.count + 1 + size, dtype=np.uint64)
            z = self.state + k * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        self.count += size
        return z

    def uniform(self, size: int) -> np.ndarray:
        """Doubles in [0, 1) from the top 53 bits."""
        return (self.next_u64(size) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, size: int) -> np.ndarray:
        """Box-Muller pairs from two uniform streams."""
        m = (size + 1) // 2
        u1 = 1.0 - self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(u1))
        out = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return out[:size]


BRITTLE, DUCTILE, MIXED = 0, 1, 2  # FractureMode (src/material.rs:4-8)
# Order of the per-REV strength row (Material fields, src/material.rs:20-63).
STRENGTH_FIELDS = ("tension", "tension_perp", "compression", "compression_perp", "shear", "shear_perp",
                   "yield", "hardening", "plastic_failure", "mode", "gc", "gc_perp")
# Per-bond history beyond damage and plastic strain (Bond, src/bond.rs:4-16).
HISTORY_FIELDS = ("equivalent_plastic_strain", "plastic_work", "fatigue_damage", "previous_stress_ratio",
                  "reversal_stress_ratio", "loading_increment")


@dataclasses.dataclass(frozen=True)
class Material:
    """Elastic, thermal and strength constants used by the solve and the
    fracture law (src/material.rs:20-63)."""

    name: str
    density: float
    young: float
    young_perp: float
    shear: float
    alpha_par: float = 0.0
    alpha_perp: float = 0.0
    strength: tuple = (0.0,) * 12  # STRENGTH_FIELDS order, Pa / - / J m^-2


# Built-in values of src/material.rs:105-150, 277-295 for iron and stone.
# Wood uses the values of the synthetic workloads (600 kg/m^3, 12 / 0.8 GPa),
# with the built-in shear modulus, expansion and strength (src/material.rs:
# 96-116, 277-286): limits 100 / 50 / 12 MPa along the grain and 2.5 / 2 / 3
# MPa across (tension / compression / shear), yield 25 MPa, mixed mode.
_WOOD_STRENGTH = (100e6, 2.5e6, 50e6, 2e6, 12e6, 3e6, 25e6, 2e9, 0.04, MIXED, 350.0, 80.0)
MATERIALS = {
    "wood": Material("wood", 600.0, 12e9, 0.8e9, 0.65e9, 3.0e-6, 30.0e-6, _WOOD_STRENGTH),
    "wood_builtin": Material("wood_builtin", 650.0, 10e9, 0.6e9, 0.65e9, 3.0e-6, 30.0e-6, _WOOD_STRENGTH),
    "iron": Material("iron", 7870.0, 200e9, 200e9, 77e9, 12.0e-6, 12.0e-6,
                     (350e6, 350e6, 350e6, 350e6, 200e6, 200e6, 200e6, 1e9, 0.15, DUCTILE, 30000.0, 30000.0)),
    "stone": Material("stone", 2700.0, 50e9, 50e9, 20e9, 8.0e-6, 8.0e-6,
                      (7e6, 7e6, 120e6, 120e6, 12e6, 12e6, 100e6, 0.0, 0.0, BRITTLE, 80.0, 80.0)),
}


@dataclasses.dataclass
class FractureConfig:
    """FractureConfig of scenes/wood_chop.ron (src/fracture.rs:11-25)."""

    damage_rate: float = 35.0
    energy_release: float = 0.005
    fatigue_cycles: float = 1000.0
    fatigue_exponent: float = 3.0
    fatigue_threshold: float = 0.5
    plastic_heat_fraction: float = 0.9

AXES = np.eye(3, dtype=np.float32)


@dataclasses.dataclass
class StepConfig:
    """Solver fields of SceneConfig (src/scene.rs:215-266)."""

    fixed_dt: float = 0.001
    substeps: int = 1
    iterations: int = 80
    cuda_iterations: int = 160
    corotation_iterations: int = 2
    solver_relative_tolerance: float = 1e-5
    corotation_tolerance: float = 1e-5
    gravity: tuple = (0.0, 0.0, -9.81)


@dataclasses.dataclass
class Scene:
    """SoA scene state; every float plane is float32 like the reference."""

    dims: tuple
    positions: np.ndarray      # (n, 3)
    rotations: np.ndarray      # (n, 4) quaternion x, y, z, w
    velocities: np.ndarray     # (n, 3)
    mass: np.ndarray           # (n,)
    kinematic: np.ndarray      # (n,) uint8
    young: np.ndarray
    young_perp: np.ndarray
    shear: np.ndarray
    solid: np.ndarray
    fibers: np.ndarray         # (n, 3) body-frame fibre direction
    anisotropy: np.ndarray
    delta_temp: np.ndarray
    alpha_par: np.ndarray
    alpha_perp: np.ndarray
    neighbor: np.ndarray       # (3, n) uint32
    damage: np.ndarray         # (3, n)
    kscale: np.ndarray         # (3, n)
    formation_offset: np.ndarray  # (3, n, 3)
    plastic_strain: np.ndarray    # (3, n, 3)
    targets: np.ndarray        # (n, 3), NaN where the axis is free
    loads: np.ndarray          # (n, 3) applied loads (zero on kinematic voxels)
    keys: np.ndarray           # (n, 3) int32 lattice keys
    config: StepConfig = dataclasses.field(default_factory=StepConfig)
    name: str = ""
    strength: np.ndarray = None  # (n, 12) STRENGTH_FIELDS per REV
    bond_history: np.ndarray = None  # (3, n, 6) HISTORY_FIELDS per +axis bond; None = zero

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])

    def free_mask(self) -> np.ndarray:
        """Bit a set when axis a is solved for (src/solver.rs:1554-1565)."""
        free = np.isnan(self.targets) & (self.kinematic == 0)[:, None]
        return (free[:, 0] * 1 + free[:, 1] * 2 + free[:, 2] * 4).astype(np.uint8)

    def copy(self) -> "Scene":
        kw = {}
        for f in dataclasses.fields(self):
            v = getattr(self, f.name)
            kw[f.name] = v.copy() if isinstance(v, np.ndarray) else (dataclasses.replace(v) if isinstance(v, StepConfig) else v)
        return Scene(**kw)

    def bond_count(self) -> int:
        return int(np.count_nonzero(self.neighbor != EMPTY))

    def active_count(self) -> int:
        return int(np.count_nonzero(self.free_mask()))


def lattice_topology(dims, kinematic: np.ndarray) -> np.ndarray:
    """+axis bonds of a full nx*ny*nz box: i -> i + stride_a when the
    neighbour cell exists and both kinematic flags are equal
    (src/scene.rs:558-570)."""
    nx, ny, nz = dims
    n = nx * ny * nz
    idx = np.arange(n, dtype=np.int64)
    x = idx % nx
    y = (idx // nx) % ny
    z = idx // (nx * ny)
    nb = np.full((3, n), EMPTY, dtype=np.uint32)
    for axis, (coord, extent, stride) in enumerate(((x, nx, 1), (y, ny, nx), (z, nz, nx * ny))):
        has = coord + 1 < extent
        j = idx + stride
        ok = has.copy()
        ok[has] = kinematic[idx[has]] == kinematic[j[has]]
        nb[axis, ok] = j[ok].astype(np.uint32)
    return nb


def _box(dims, material_ids, grain_axis, anisotropy, kinematic, materials) -> Scene:
    nx, ny, nz = dims
    n = nx * ny * nz
    idx = np.arange(n)
    x = idx % nx
    y = (idx // nx) % ny
    z = idx // (nx * ny)
    keys = np.stack([x - nx // 2, y - ny // 2, z], axis=1).astype(np.int32)
    positions = keys.astype(np.float32) * CELL + np.float32(CELL * np.float32(0.5))
    table = [MATERIALS[m] for m in materials]
    mid = np.asarray(material_ids)
    f32 = lambda field: np.array([getattr(m, field) for m in table], dtype=np.float32)[mid]  # noqa: E731
    mass = (f32("density") * np.float32(CELL) ** 3).astype(np.float32)
    kin = np.asarray(kinematic, dtype=np.uint8)
    return Scene(
        dims=(nx, ny, nz),
        positions=positions.astype(np.float32),
        rotations=np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1)),
        velocities=np.zeros((n, 3), np.float32),
        mass=mass,
        kinematic=kin,
        young=f32("young"),
        young_perp=f32("young_perp"),
        shear=f32("shear"),
        solid=np.ones(n, np.float32),
        fibers=AXES[np.asarray(grain_axis)].astype(np.float32),
        anisotropy=np.asarray(anisotropy, np.float32) * np.ones(n, np.float32),
        delta_temp=np.zeros(n, np.float32),
        alpha_par=f32("alpha_par"),
        alpha_perp=f32("alpha_perp"),
        neighbor=lattice_topology(dims, np.zeros(n, np.uint8)),
        damage=np.zeros((3, n), np.float32),
        kscale=np.ones((3, n), np.float32),
        formation_offset=np.zeros((3, n, 3), np.float32),
        plastic_strain=np.zeros((3, n, 3), np.float32),
        targets=np.full((n, 3), np.nan, np.float32),
        loads=np.zeros((n, 3), np.float32),
        keys=keys,
        strength=np.array([m.strength for m in table], np.float32)[mid],
    )


def uniform_box(dims, material="wood", grain=2, anisotropy=None) -> Scene:
    """Homogeneous box: one material, one grain axis, no loads or supports."""
    n = dims[0] * dims[1] * dims[2]
    if anisotropy is None:
        anisotropy = 1.0 if material.startswith("wood") else 0.0
    return _box(dims, np.zeros(n, np.int64), np.full(n, grain), anisotropy, np.zeros(n, np.uint8), [material])


def tension_10cube() -> Scene:
    """10x10x10 wood REVs, grain Z, anisotropy 1 (scenes/wood_chop.ron:9).
    Bottom and top faces are axis-targeted on all axes (the fixture pattern of
    src/experiments.rs:73-87); the top target is rest + 0.1 mm in Z. One 1 ms
    step with one substep, no gravity, at most 4 corotation passes."""
    s = uniform_box((10, 10, 10), "wood", grain=2)
    z = s.keys[:, 2]
    bottom = z == z.min()
    top = z == z.max()
    s.targets[bottom] = s.positions[bottom]
    s.targets[top] = s.positions[top] + np.array([0, 0, 1e-4], np.float32)
    s.config = StepConfig(fixed_dt=0.001, substeps=1, iterations=1000, corotation_iterations=4,
                          solver_relative_tolerance=1e-10, gravity=(0.0, 0.0, 0.0))
    s.name = "tension_10x10x10"
    return s


def pulled_column(material: str, pull: float, dims=(3, 3, 6), grain: int = 2) -> Scene:
    """A box of one material with its bottom layer held and its top layer
    axis-targeted `pull` metres up in Z (all axes targeted), no gravity: the
    fracture fixture of the reference's tension tests at REV scale."""
    s = uniform_box(dims, material, grain=grain)
    z = s.keys[:, 2]
    s.targets[z == z.min()] = s.positions[z == z.min()]
    s.targets[z == z.max()] = s.positions[z == z.max()] + np.array([0, 0, pull], np.float32)
    s.config = StepConfig(fixed_dt=0.001, substeps=1, iterations=1000, corotation_iterations=2,
                          solver_relative_tolerance=1e-10, gravity=(0.0, 0.0, 0.0))
    s.name = "pulled_%s_%dx%dx%d" % ((material,) + tuple(dims))
    return s


def pulled_mixed(dims, seed: int, pull: float) -> Scene:
    """A regional wood/iron/stone box (`mixed_box`, bottom face fixed) whose
    top layer is axis-targeted `pull` metres up in Z, with gravity: brittle
    stone and cross-grain wood bonds break, iron and along-grain wood yield."""
    s = mixed_box(dims, seed, "z-")
    z = s.keys[:, 2]
    top = z == z.max()
    s.targets[top] = s.positions[top] + np.array([0, 0, pull], np.float32)
    s.config = StepConfig(fixed_dt=0.001, substeps=1, iterations=1000, corotation_iterations=2,
                          solver_relative_tolerance=1e-10, gravity=(0.0, 0.0, -9.81))
    s.name = "pulled_mixed_%dx%dx%d" % tuple(dims)
    return s


def _regional_materials(dims, region, rng: SplitMix64):
    """Per region of region^3 cells: material wood / iron / stone with
    probabilities 0.6 / 0.2 / 0.2, then the wood grain axis uniform over
    X, Y, Z. Regions are drawn in z, y, x order."""
    nx, ny, nz = dims
    rx, ry, rz = (math.ceil(nx / region), math.ceil(ny / region), math.ceil(nz / region))
    nreg = rx * ry * rz
    u = rng.uniform(nreg)
    mat = np.where(u < 0.6, 0, np.where(u < 0.8, 1, 2))
    grain = np.minimum((rng.uniform(nreg) * 3).astype(np.int64), 2)
    idx = np.arange(nx * ny * nz)
    x = idx % nx
    y = (idx // nx) % ny
    z = idx // (nx * ny)
    reg = (z // region) * (ry * rx) + (y // region) * rx + (x // region)
    return mat[reg], grain[reg]


def _finish_mixed(s: Scene, mat, rng: SplitMix64):
    s.anisotropy = np.where(mat == 0, 1.0, 0.0).astype(np.float32)
    s.kscale = (0.95 + 0.1 * rng.uniform(3 * s.n)).astype(np.float32).reshape(3, s.n)


def mixed_box(dims, seed: int, fixed_face: str = "z-") -> Scene:
    """A box with regional wood/iron/stone materials, per-bond stiffness jitter
    U[0.95, 1.05] and one face fixed with Scene::fix_voxels semantics (the
    voxels become kinematic but keep their bonds)."""
    rng = SplitMix64(seed)
    mat, grain = _regional_materials(dims, 8, rng)
    n = dims[0] * dims[1] * dims[2]
    s = _box(dims, mat, grain, 0.0, np.zeros(n, np.uint8), ["wood", "iron", "stone"])
    _finish_mixed(s, mat, rng)
    axis = "xyz".index(fixed_face[0])
    coord = s.keys[:, axis]
    face = coord == (coord.min() if fixed_face[1] == "-" else coord.max())
    s.kinematic = face.astype(np.uint8)
    return s


def large_64cube(seed: int = 2, dims=(64, 64, 64)) -> Scene:
    """Single large mixed scene: 64^3 REVs, materials per 8^3 region, bottom
    layer fixed, gravity -9.81 Z, persistent +0.2 N Z force on a random 5% of
    the top layer (drawn without replacement), 1 ms steps of one substep."""
    s = mixed_box(dims, seed, "z-")
    rng = SplitMix64(seed ^ 0x5EED)
    z = s.keys[:, 2]
    top = np.nonzero(z == z.max())[0]
    count = int(round(0.05 * top.size))
    order = np.argsort(rng.uniform(top.size), kind="stable")
    chosen = top[order[:count]]
    s.loads[chosen, 2] = np.float32(0.2)
    s.config = StepConfig(fixed_dt=0.001, substeps=1, iterations=1000, corotation_iterations=2,
                          solver_relative_tolerance=1e-6, gravity=(0.0, 0.0, -9.81))
    s.name = "large_%dx%dx%d_mixed" % tuple(dims)
    return s


def wood_block_blade(seed: int = 5, dims=(192, 192, 96), blade=(40, 5, 20), speed=5.0) -> Scene:
    """The solve of the large cutting scene at its real size: a wood block
    (grain Z, per-bond stiffness jitter U[0.95, 1.05]) with the bottom layer
    fixed and gravity. The rigid iron blade (realised 40x5x20 REVs) driven at
    `speed` m/s in -Z is represented by its footprint on the top face: the
    top-layer REVs under the blade are axis-targeted in Z one step's travel
    down (contact projection and fracture are not part of this solve)."""
    rng = SplitMix64(seed)
    n = dims[0] * dims[1] * dims[2]
    s = _box(dims, np.zeros(n, np.int64), np.full(n, 2), 1.0, np.zeros(n, np.uint8), ["wood"])
    s.kscale = (0.95 + 0.1 * rng.uniform(3 * n)).astype(np.float32).reshape(3, n)
    z = s.keys[:, 2]
    s.kinematic = (z == z.min()).astype(np.uint8)
    dt = 0.001
    s.targets = blade_targets(s, blade, speed, dt, 0)
    s.config = StepConfig(fixed_dt=dt, substeps=1, iterations=1000, corotation_iterations=6,
                          solver_relative_tolerance=1e-6, gravity=(0.0, 0.0, -9.81))
    s.name = "wood_block_%dx%dx%d_blade" % tuple(dims)
    return s


def wood_block(seed: int = 5, dims=(192, 192, 96)) -> Scene:
    """The cutting scene's wood block (grain Z, per-bond stiffness jitter,
    bottom layer fixed, gravity, up to six corotation passes) without axis
    targets: the blade acts on it as a rigid tool (`blade_tool`)."""
    s = wood_block_blade(seed=seed, dims=dims)
    s.targets = np.full((s.n, 3), np.nan, np.float32)
    s.name = "wood_block_%dx%dx%d" % tuple(dims)
    return s


def blade_tool(s: Scene, tool_dimensions=(40, 2, 20), speed: float = 5.0, gap: float = 2e-4,
               offset=(0.3137, 0.1713), friction: float = 0.35, mounted: bool = False):
    """The blade of the cutting scene as a rigid tool (src/tool.rs:45-110):
    Tool::new's tapered blade (the 2-REV thickness realised as a 5-REV spine),
    iron, driven along -Z at `speed`; its edge cell `gap` metres above the
    block's top face, centred over the face (shifted by `offset` cells in X
    and Y so no cell face lies on a voxel face). `mounted` slides it on a
    guide along its direction of travel from its rear face, as the
    reference's application mounts its driven tools (src/main.rs:21-29):
    rotation frozen, motion only along -Z."""
    from .tool import BLADE, Tool
    top = float(s.keys[:, 2].max() + 1) * float(CELL)
    cx = float(s.positions[:, 0].mean()) + offset[0] * float(CELL)
    cy = float(s.positions[:, 1].mean()) + offset[1] * float(CELL)
    t = Tool.new(BLADE, tool_dimensions, [cx, cy, 0.0], [0.0, 0.0, -1.0], speed, friction=friction)
    lowest = t.world_cells()[:, 2].min()
    t.position = t.position + np.array([0.0, 0.0, top + 0.5 * float(CELL) + gap - lowest])
    if mounted:
        t.mount_base([float(t.cells[:, 0].min()), 0.0, 0.0], [0.0, 0.0, -1.0])
    return t


def blade_footprint(s: Scene, blade=(40, 5, 20)) -> np.ndarray:
    """Top-layer REVs under the blade's blade[0] x blade[1] edge, centred on
    the box's top face."""
    nx, ny, _ = s.dims
    z = s.keys[:, 2]
    x = s.keys[:, 0] - s.keys[:, 0].min()
    y = s.keys[:, 1] - s.keys[:, 1].min()
    bx0, by0 = (nx - blade[0]) // 2, (ny - blade[1]) // 2
    return (z == z.max()) & (x >= bx0) & (x < bx0 + blade[0]) & (y >= by0) & (y < by0 + blade[1])


def blade_targets(s: Scene, blade=(40, 5, 20), speed=5.0, dt=0.001, step: int = 0) -> np.ndarray:
    """Axis targets of step `step` (0-based) of a blade driven at `speed` m/s
    in -Z: the footprint REVs are targeted in Z (step + 1) steps of travel
    below their rest height; every other axis is free."""
    t = np.full((s.n, 3), np.nan, np.float32)
    under = blade_footprint(s, blade)
    rest_z = s.keys[:, 2].astype(np.float32) * CELL + np.float32(CELL * np.float32(0.5))  # as _box places REVs
    t[under, 2] = rest_z[under] - np.float32(speed * dt * (step + 1))
    return t


def batch_boxes(count: int = 256, seed: int = 3, lo: int = 8, hi: int = 48):
    """Independent boxes with integer edges U[lo, hi] per axis, regional
    materials, one random face fixed and initial velocities N(0, 0.05 m/s)
    per component on the free voxels."""
    rng = SplitMix64(seed)
    edges = lo + np.minimum((rng.uniform(3 * count) * (hi - lo + 1)).astype(np.int64), hi - lo)
    faces = np.minimum((rng.uniform(count) * 6).astype(np.int64), 5)
    scenes = []
    for b in range(count):
        dims = tuple(int(e) for e in edges[3 * b:3 * b + 3])
        face = "xyz"[faces[b] // 2] + "-+"[faces[b] % 2]
        s = mixed_box(dims, seed * 1_000_003 + b, face)
        vrng = SplitMix64(seed * 7_000_001 + b)
        v = (0.05 * vrng.normal(3 * s.n)).astype(np.float32).reshape(s.n, 3)
        v[s.kinematic != 0] = 0
        s.velocities = v
        s.config = StepConfig(fixed_dt=0.001, substeps=1, iterations=1000, corotation_iterations=2,
                              solver_relative_tolerance=1e-6, gravity=(0.0, 0.0, -9.81))
        s.name = "batch%03d_%dx%dx%d" % ((b,) + dims)
        scenes.append(s)
    return scenes


def rotate_scene(s: Scene, quat_xyzw) -> Scene:
    """Rigidly rotate all positions about the origin (as the reference's
    rest_and_rigid_rotation_have_no_strain test does, src/solver.rs:1008-1021)."""
    x, y, z, w = [float(c) for c in quat_xyzw]
    r = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    out = s.copy()
    out.positions = (s.positions.astype(np.float64) @ r.T).astype(np.float32)
    return out


def contact_box(dims, seed: int = 7, broken: float = 0.02, jitter: float = 2e-4, sag: float = 3e-6,
                movers: float = 0.002, mover_reach: float = 1.5, targeted: float = 0.01,
                fixed_floor: bool = True) -> dict:
    """Input of one contact projection (collision::project, src/collision.rs:33-123)
    on a lattice box part way through a step: the scene fields it reads, f32.

    Rest positions are the lattice centres; bonds follow the lattice rule, and
    a fraction `broken` of them is removed (cracks). With `fixed_floor` the
    z = 0 layer is kinematic (fixed, hence unbonded to the layer above). Each
    movable voxel starts (previous) within jitter/2 of its rest position and
    ends (positions) within a further `jitter`, sunk by `sag` (the load of a
    settled block pressing its bottom layer into the floor); a fraction
    `movers` starts up to `mover_reach` cells away instead, so its swept box
    spans several bins. A fraction `targeted` of the movable voxels has one
    axis held by a target. Masses are cell volumes at densities U[500, 8000].
    """
    nx, ny, nz = dims
    n = nx * ny * nz
    rng = SplitMix64(seed)
    idx = np.arange(n, dtype=np.int64)
    key = np.stack([idx % nx - nx // 2, (idx // nx) % ny - ny // 2, idx // (nx * ny)], axis=1)
    rest = (key.astype(np.float32) * CELL + CELL / np.float32(2)).astype(np.float32)
    kin = (key[:, 2] == 0) if fixed_floor else np.zeros(n, dtype=bool)
    nb = lattice_topology(dims, kin)
    cut = rng.uniform(3 * n).reshape(3, n) < broken
    nb[cut & (nb != EMPTY)] = EMPTY
    mass = (float(CELL) ** 3 * (500.0 + 7500.0 * rng.uniform(n))).astype(np.float32)
    move = ~kin
    prev = rest.astype(np.float64) + (0.5 * jitter) * (2.0 * rng.uniform(3 * n).reshape(n, 3) - 1.0) * move[:, None]
    pos = prev + jitter * (2.0 * rng.uniform(3 * n).reshape(n, 3) - 1.0) * move[:, None]
    pos[move, 2] -= sag
    far = move & (rng.uniform(n) < movers)
    direction = rng.normal(3 * n).reshape(n, 3)
    direction /= np.maximum(np.linalg.norm(direction, axis=1, keepdims=True), 1e-12)
    reach = mover_reach * float(CELL) * rng.uniform(n)
    prev[far] = pos[far] - direction[far] * reach[far, None]
    targets = np.zeros((n, 3), dtype=bool)
    held = move & (rng.uniform(n) < targeted)
    axis = (rng.uniform(n) * 3).astype(np.int64).clip(0, 2)
    targets[held, axis[held]] = True
    return {
        "positions": pos.astype(np.float32), "previous": prev.astype(np.float32), "rest_positions": rest,
        "mass": mass, "kinematic": kin, "bond_neighbor": nb, "targets": targets,
    }
