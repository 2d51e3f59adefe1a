"""Rigid tool: a voxel group moving as one body (the reference's Tool,
src/tool.rs:26-43), built on the host and advanced on the device.

`Tool.new` restates Tool::new (src/tool.rs:45-110) in f32: the shape's cells
on a unit grid (x outermost), the blade and wedge taper, the material
density of the shape's component (iron for blades and wedges, wood for rods,
stone for rounded tools; src/material.rs:73-160), cells re-centred on their
centroid, the diagonal inertia of cubic cells about it, and the orientation
that turns the body X axis onto the direction of travel (glam
Quat::from_rotation_arc). The device advance (vx_scene_set_tool, used by
vx_scene_step) restates Tool::advance (src/tool.rs:192-431).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np

from ._abi import ToolDesc, ToolState, c_fp

CELL = np.float32(0.01)
f32 = np.float32

BLADE, WEDGE, ROD, ROUNDED = "blade", "wedge", "rod", "rounded"
_DENSITY = {BLADE: 7870.0, WEDGE: 7870.0, ROD: 650.0, ROUNDED: 2700.0}


def _v(x):
    return np.asarray(x, dtype=np.float32)


def _dot(a, b):
    return f32(f32(f32(a[0] * b[0]) + f32(a[1] * b[1])) + f32(a[2] * b[2]))


def _normalize(a):
    return _v(a) * f32(f32(1.0) / f32(math.sqrt(float(_dot(a, a)))))


def _qmul_vec(q, v):
    """glam Quat::mul_vec3: v (w^2 - b.b) + b (2 v.b) + (b x v) 2w, f32."""
    b = _v(q[:3])
    w = f32(q[3])
    b2 = _dot(b, b)
    cr = _v([b[1] * v[2] - v[1] * b[2], b[2] * v[0] - v[2] * b[0], b[0] * v[1] - v[0] * b[1]])
    return (_v(v) * f32(w * w - b2) + b * f32(_dot(v, b) * f32(2.0))) + cr * f32(w * f32(2.0))


def rotation_arc(frm, to):
    """glam Quat::from_rotation_arc(from, to) for unit vectors, (x, y, z, w)."""
    frm, to = _v(frm), _v(to)
    d = _dot(frm, to)
    eps = f32(1.0) - f32(2.0) * f32(np.finfo(np.float32).eps)
    if d > eps:
        return _v([0, 0, 0, 1])
    if d < -eps:
        # any_orthonormal_vector of `from`, then half a turn about it
        sign = f32(1.0) if not math.copysign(1.0, float(frm[2])) < 0 else f32(-1.0)
        a = f32(f32(-1.0) / f32(sign + frm[2]))
        b = f32(f32(frm[0] * frm[1]) * a)
        axis = _v([b, f32(sign + f32(f32(frm[1] * frm[1]) * a)), -frm[1]])
        half = float(f32(f32(np.pi) * f32(0.5)))  # glam: sin_cos(PI * 0.5) in f32
        s, c = f32(math.sin(half)), f32(math.cos(half))
        return _v([axis[0] * s, axis[1] * s, axis[2] * s, c])
    c = _v([frm[1] * to[2] - to[1] * frm[2], frm[2] * to[0] - to[2] * frm[0], frm[0] * to[1] - to[0] * frm[1]])
    q = _v([c[0], c[1], c[2], f32(1.0) + d])
    sq = f32(f32(f32(f32(q[0] * q[0]) + f32(q[1] * q[1])) + f32(q[2] * q[2])) + f32(q[3] * q[3]))
    n = f32(math.sqrt(float(sq)))
    return q * f32(f32(1.0) / n)


@dataclasses.dataclass
class Tool:
    cells: np.ndarray          # (n, 3) float32, body frame about the centre of mass
    mass: float
    inverse_inertia: np.ndarray
    position: np.ndarray       # centre of mass, world
    rotation: np.ndarray       # x, y, z, w
    velocity: np.ndarray
    angular_velocity: np.ndarray
    external_force: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    external_torque: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))
    friction: float = 0.35     # SceneConfig fracture.friction of scenes/wood_chop.ron
    mount: tuple = None        # (local_base, world_base, direction, rotation)

    @classmethod
    def new(cls, shape: str, tool_dimensions, position, direction, speed: float, friction: float = 0.35) -> "Tool":
        d = [int(x) for x in tool_dimensions]
        if shape == BLADE:
            dims = [d[0], max(d[1], 5), d[2]]  # a 5-REV spine tapered to one REV at the edge
        elif shape == WEDGE:
            dims = d
        elif shape == ROD:
            dims = [d[0], 4, 4]
        elif shape == ROUNDED:
            dims = [14, 14, 14]
        else:
            raise ValueError("unknown tool shape %r" % shape)
        half = _v(dims) * f32(0.5)
        cells = []
        for x in range(dims[0]):
            for y in range(dims[1]):
                for z in range(dims[2]):
                    p = _v([f32(x) + f32(0.5), f32(y) + f32(0.5), f32(z) + f32(0.5)]) - half
                    if shape in (BLADE, WEDGE):
                        t = f32(f32(0.5) + f32(f32(f32(dims[0] - 1 - x) / f32(dims[0] - 1)) *
                                               f32(f32(dims[1]) * f32(0.5) - f32(0.5))))
                        blunt = f32(0.5) if shape == WEDGE else f32(0.0)
                        inc = abs(p[1]) <= f32(t + blunt)
                    elif shape == ROD:
                        inc = f32(f32(p[1] * p[1]) + f32(p[2] * p[2])) <= f32(4.0)
                    else:
                        inc = _dot(p, p) <= f32(49.0)
                    if inc:
                        cells.append(p * CELL)
        cells = np.array(cells, np.float32)
        cell_mass = f32(f32(_DENSITY[shape]) * f32(CELL * CELL * CELL))
        mass = f32(f32(len(cells)) * cell_mass)
        direction = _normalize(direction)
        rotation = rotation_arc([1, 0, 0], direction)
        total = _v([0, 0, 0])
        for p in cells:
            total = total + p
        centroid = total / f32(len(cells))
        cells = (cells - centroid).astype(np.float32)
        inertia = _v([0, 0, 0])
        six = f32(f32(CELL * CELL) / f32(6.0))
        for p in cells:
            inertia = inertia + (np.full(3, f32(_dot(p, p) + six), np.float32) - p * p) * cell_mass
        return cls(cells=cells, mass=float(mass), inverse_inertia=(f32(1.0) / inertia).astype(np.float64),
                   position=(_v(position) + _qmul_vec(rotation, centroid)).astype(np.float64),
                   rotation=rotation.astype(np.float64), velocity=(direction * f32(speed)).astype(np.float64),
                   angular_velocity=np.zeros(3), friction=float(friction))

    def mount_base(self, local_base, direction):
        """Tool::mount_base (src/tool.rs:127-142): the base slides along the
        unit direction (zero direction: fixed); the rotation is frozen."""
        d = np.asarray(direction, np.float64)
        n = np.linalg.norm(d)
        d = d / n if n > 0 and np.isfinite(1.0 / n) else np.zeros(3)
        world = self.position + _qrot64(self.rotation, np.asarray(local_base, np.float64))
        self.mount = (np.asarray(local_base, np.float64), world, d, self.rotation.copy())
        self.velocity = d * np.dot(d, self.velocity)
        self.angular_velocity = np.zeros(3)

    def radius(self) -> float:
        return float(np.sqrt((self.cells.astype(np.float64) ** 2).sum(axis=1)).max())

    def world_cells(self) -> np.ndarray:
        return self.position + np.array([_qrot64(self.rotation, c) for c in self.cells.astype(np.float64)])

    # ---- the C ABI view ---------------------------------------------------
    def desc(self):
        """(vx_tool_desc, kept arrays)."""
        cells = np.ascontiguousarray(self.cells, np.float32)
        d = ToolDesc()
        d.n_cells = len(cells)
        d.has_mount = 1 if self.mount is not None else 0
        d.cells = cells.ctypes.data_as(c_fp)
        d.mass = self.mass
        for name in ("inverse_inertia", "position", "rotation", "velocity", "angular_velocity", "external_force",
                     "external_torque"):
            arr = getattr(d, name)
            for k, x in enumerate(getattr(self, name)):
                arr[k] = float(x)
        d.friction = self.friction
        if self.mount is not None:
            lb, wb, dr, rot = self.mount
            for k in range(3):
                d.mount_local_base[k], d.mount_world_base[k], d.mount_direction[k] = lb[k], wb[k], dr[k]
            for k in range(4):
                d.mount_rotation[k] = rot[k]
        return d, cells

    def update_from(self, st: ToolState):
        self.position = np.array(st.position[:])
        self.rotation = np.array(st.rotation[:])
        self.velocity = np.array(st.velocity[:])
        self.angular_velocity = np.array(st.angular_velocity[:])


def _qrot64(q, v):
    b = np.asarray(q[:3], np.float64)
    w = float(q[3])
    return v * (w * w - b @ b) + b * (2.0 * (v @ b)) + np.cross(b, v) * (2.0 * w)
