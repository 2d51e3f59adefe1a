"""Interop with the reference's file formats.

* `SceneConfig` files (RON, src/scene.rs:215-303; the fracture block
  src/fracture.rs:10-61): `load_scene_config` reads the subset of RON those
  files use (named or anonymous structs, tuples, numbers, identifiers,
  comments), rejects unknown and duplicate fields like serde's
  `deny_unknown_fields`, fills the two defaulted fields, rounds every f32
  field to f32, and applies `SceneConfig::validate` (src/scene.rs:274-301)
  and `FractureConfig::validate` (src/fracture.rs:45-61).
* The H/F/V capture CSV written by src/bin/solver_scale_capture.rs:113-149
  (and read by validation/live_viewer.py): `CaptureWriter` writes it with the
  same record layout and Rust's shortest round-trip f32 formatting,
  `read_capture` reads it back (other record kinds such as the thermal
  S/Q/L/B lines are kept as raw rows), `surface_samples` and `surface_state`
  restate the capture's sampling of the outer surface and its per-REV damage
  and lost-bond count (src/bin/solver_scale_capture.rs:20-45, 84-97), and
  `cluster_count` the connected-component count of `rebuild_clusters`
  (src/solver.rs:828-858).
"""
from __future__ import annotations

import dataclasses
import math
import re
from typing import Dict, List, Optional, Tuple

import numpy as np

EMPTY = 0xFFFFFFFF


class ConfigError(ValueError):
    pass


# ---- RON subset ---------------------------------------------------------------
_TOKEN = re.compile(r"""
    (?P<ws>\s+) |
    (?P<line>//[^\n]*) |
    (?P<block>/\*.*?\*/) |
    (?P<num>[+-]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?) |
    (?P<ident>[A-Za-z_][A-Za-z0-9_]*) |
    (?P<str>"(?:[^"\\]|\\.)*") |
    (?P<punct>[()\[\]{}:,])
""", re.VERBOSE | re.DOTALL)


def _tokens(text: str):
    pos = 0
    out = []
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise ConfigError("unexpected character %r at offset %d" % (text[pos], pos))
        pos = m.end()
        kind = m.lastgroup
        if kind in ("ws", "line", "block"):
            continue
        out.append((kind, m.group(kind)))
    return out


class _Parser:
    def __init__(self, text: str):
        self.t = _tokens(text)
        self.i = 0

    def peek(self, k=0):
        return self.t[self.i + k] if self.i + k < len(self.t) else (None, None)

    def take(self, value=None):
        tok = self.peek()
        if tok[0] is None:
            raise ConfigError("unexpected end of input")
        if value is not None and tok[1] != value:
            raise ConfigError("expected %r, found %r" % (value, tok[1]))
        self.i += 1
        return tok

    def value(self):
        kind, v = self.peek()
        if kind == "num":
            self.take()
            return float(v) if any(c in v for c in ".eE") else int(v)
        if kind == "str":
            self.take()
            return bytes(v[1:-1], "utf-8").decode("unicode_escape")
        if kind == "ident":
            self.take()
            if v in ("true", "false"):
                return v == "true"
            if self.peek()[1] == "(":  # a named struct or tuple
                return self.group()
            return ("enum", v)
        if v == "(":
            return self.group()
        if v == "[":
            self.take("[")
            items = []
            while self.peek()[1] != "]":
                items.append(self.value())
                if self.peek()[1] == ",":
                    self.take(",")
            self.take("]")
            return items
        raise ConfigError("unexpected token %r" % (v,))

    def group(self):
        self.take("(")
        if self.peek()[0] == "ident" and self.peek(1)[1] == ":":
            fields: Dict[str, object] = {}
            while self.peek()[1] != ")":
                _, name = self.take()
                self.take(":")
                if name in fields:
                    raise ConfigError("duplicate field `%s`" % name)
                fields[name] = self.value()
                if self.peek()[1] == ",":
                    self.take(",")
            self.take(")")
            return fields
        items = []
        while self.peek()[1] != ")":
            items.append(self.value())
            if self.peek()[1] == ",":
                self.take(",")
        self.take(")")
        return tuple(items)


def parse_ron(text: str):
    p = _Parser(text)
    v = p.value()
    if p.peek()[0] is not None:
        raise ConfigError("trailing characters after the value")
    return v


# ---- SceneConfig ------------------------------------------------------------------
def _f32(x) -> float:
    if isinstance(x, bool) or not isinstance(x, (int, float)):
        raise ConfigError("expected a number, found %r" % (x,))
    return float(np.float32(x))


def _usize(x) -> int:
    if isinstance(x, bool) or not isinstance(x, int) or x < 0:
        raise ConfigError("expected a non-negative integer, found %r" % (x,))
    return int(x)


def _triple(x, conv):
    if not isinstance(x, (tuple, list)) or len(x) != 3:
        raise ConfigError("expected a 3-tuple, found %r" % (x,))
    return tuple(conv(v) for v in x)


@dataclasses.dataclass
class FractureSettings:
    """FractureConfig (src/fracture.rs:10-25), all fields."""

    damage_rate: float
    energy_release: float
    propagation_delay: int
    fatigue_cycles: float
    fatigue_exponent: float
    fatigue_threshold: float
    plastic_heat_fraction: float
    fragment_min_size: int
    cluster_interval: int
    friction: float
    restitution: float

    def valid(self) -> bool:
        """FractureConfig::validate (src/fracture.rs:45-61)."""
        f = math.isfinite
        return (f(self.damage_rate) and self.damage_rate > 0.0 and 0.0 <= self.energy_release <= 1.0
                and f(self.fatigue_cycles) and self.fatigue_cycles > 0.0 and f(self.fatigue_exponent)
                and self.fatigue_exponent >= 1.0 and 0.0 <= self.fatigue_threshold < 1.0
                and 0.0 <= self.plastic_heat_fraction <= 1.0 and self.fragment_min_size > 0
                and self.cluster_interval > 0 and f(self.friction) and self.friction >= 0.0
                and 0.0 <= self.restitution <= 1.0)

    def fracture_config(self):
        """The fields the fracture pass uses, as a `vxsim.scenes.FractureConfig`."""
        from .scenes import FractureConfig
        return FractureConfig(self.damage_rate, self.energy_release, self.fatigue_cycles, self.fatigue_exponent,
                              self.fatigue_threshold, self.plastic_heat_fraction)


_FRACTURE_FIELDS = {"damage_rate": _f32, "energy_release": _f32, "propagation_delay": _usize,
                    "fatigue_cycles": _f32, "fatigue_exponent": _f32, "fatigue_threshold": _f32,
                    "plastic_heat_fraction": _f32, "fragment_min_size": _usize, "cluster_interval": _usize,
                    "friction": _f32, "restitution": _f32}

CPU_PRECONDITIONERS = ("IncompleteCholesky0", "BlockJacobi")  # src/scene.rs:14-24


@dataclasses.dataclass
class SceneConfig:
    """SceneConfig (src/scene.rs:215-242); f32 fields hold f32 values."""

    dimensions: Tuple[int, int, int]
    branch_dimensions: Tuple[int, int, int]
    ground_dimensions: Tuple[int, int, int]
    tool_dimensions: Tuple[int, int, int]
    fibers: Tuple[float, float, float]
    anisotropy: float
    gravity: Tuple[float, float, float]
    fixed_dt: float
    substeps: int
    iterations: int
    solver_relative_tolerance: float
    corotation_tolerance: float
    corotation_iterations: int
    fracture: FractureSettings
    initial_speed: float
    tool_start: Tuple[float, float, float]
    cuda_iterations: int = 160  # default_cuda_iterations (src/scene.rs:303-305)
    cpu_preconditioner: str = "IncompleteCholesky0"

    def valid(self) -> bool:
        """SceneConfig::validate (src/scene.rs:274-301)."""
        f = math.isfinite
        fib = self.fibers
        return not (0 in self.dimensions or 0 in self.ground_dimensions or any(n < 2 for n in self.tool_dimensions)
                    or not 0.0 <= self.anisotropy <= 1.0 or not all(map(f, self.gravity))
                    or not all(map(f, self.tool_start)) or not f(self.initial_speed) or self.initial_speed < 0.0
                    or not all(map(f, fib)) or not f(self.fixed_dt) or self.fixed_dt <= 0.0 or self.substeps == 0
                    or self.iterations == 0 or self.cuda_iterations == 0 or self.corotation_iterations == 0
                    or not f(self.solver_relative_tolerance) or self.solver_relative_tolerance <= 0.0
                    or not f(self.corotation_tolerance) or self.corotation_tolerance <= 0.0
                    or not self.fracture.valid()
                    or float(np.float32(fib[0] * fib[0] + fib[1] * fib[1] + fib[2] * fib[2])) < 1e-8)

    def step_config(self):
        """The solver fields as a `vxsim.scenes.StepConfig`."""
        from .scenes import StepConfig
        return StepConfig(fixed_dt=self.fixed_dt, substeps=self.substeps, iterations=self.iterations,
                          cuda_iterations=self.cuda_iterations, corotation_iterations=self.corotation_iterations,
                          solver_relative_tolerance=self.solver_relative_tolerance,
                          corotation_tolerance=self.corotation_tolerance, gravity=self.gravity)

    def step_params(self, solve=None, fracture: bool = True):
        """vx_step_params for `vx_scene_step` / `vx_batch_admit`: the config's
        step, substeps and corotation settings, its fracture law when
        `fracture`, and `solve` (default: mixed precision to the config's
        relative tolerance)."""
        from . import grid
        if solve is None:
            solve = grid.solve_params(tolerance=self.solver_relative_tolerance)
        return grid.step_params(self.fixed_dt, self.gravity, self.substeps, self.corotation_iterations,
                                self.corotation_tolerance, solve,
                                self.fracture.fracture_config() if fracture else None)


_SCENE_FIELDS = {"dimensions": lambda v: _triple(v, _usize), "branch_dimensions": lambda v: _triple(v, _usize),
                 "ground_dimensions": lambda v: _triple(v, _usize), "tool_dimensions": lambda v: _triple(v, _usize),
                 "fibers": lambda v: _triple(v, _f32), "anisotropy": _f32, "gravity": lambda v: _triple(v, _f32),
                 "fixed_dt": _f32, "substeps": _usize, "iterations": _usize, "cuda_iterations": _usize,
                 "solver_relative_tolerance": _f32, "corotation_tolerance": _f32, "corotation_iterations": _usize,
                 "initial_speed": _f32, "tool_start": lambda v: _triple(v, _f32)}
_SCENE_DEFAULTED = ("cuda_iterations", "cpu_preconditioner")


def _struct(value, fields, defaulted, what):
    if isinstance(value, tuple) and len(value) == 0:
        value = {}
    if not isinstance(value, dict):
        raise ConfigError("expected a %s struct" % what)
    unknown = sorted(set(value) - set(fields))
    if unknown:
        raise ConfigError("unknown field `%s` in %s" % (unknown[0], what))
    missing = [k for k in fields if k not in value and k not in defaulted]
    if missing:
        raise ConfigError("missing field `%s` in %s" % (missing[0], what))
    return value


def scene_config_from_text(text: str) -> SceneConfig:
    raw = parse_ron(text)
    fields = dict(_SCENE_FIELDS)
    fields["fracture"] = None
    fields["cpu_preconditioner"] = None
    raw = _struct(raw, fields, _SCENE_DEFAULTED, "SceneConfig")
    fr = _struct(raw["fracture"], _FRACTURE_FIELDS, (), "FractureConfig")
    fracture = FractureSettings(**{k: conv(fr[k]) for k, conv in _FRACTURE_FIELDS.items()})
    kw = {k: conv(raw[k]) for k, conv in _SCENE_FIELDS.items() if k in raw}
    if "cpu_preconditioner" in raw:
        pc = raw["cpu_preconditioner"]
        if not (isinstance(pc, tuple) and len(pc) == 2 and pc[0] == "enum" and pc[1] in CPU_PRECONDITIONERS):
            raise ConfigError("unknown cpu_preconditioner %r" % (pc,))
        kw["cpu_preconditioner"] = pc[1]
    cfg = SceneConfig(fracture=fracture, **kw)
    if not cfg.valid():
        raise ConfigError("Invalid scene parameters")
    return cfg


def load_scene_config(path: str) -> SceneConfig:
    """SceneConfig::load (src/scene.rs:269-273)."""
    with open(path) as f:
        return scene_config_from_text(f.read())


# ---- H/F/V capture ------------------------------------------------------------------
def fmt_f32(x) -> str:
    """Rust's `{}` for an f32: the shortest decimal that reads back as the same
    f32, positional, no exponent."""
    return np.format_float_positional(np.float32(x), unique=True, trim="-")


def surface_samples(dims) -> np.ndarray:
    """(sample_id, index, x, y, z) rows of the capture's surface sample: REVs on
    the box boundary whose index is not a multiple of 3, in z, y, x order
    (src/bin/solver_scale_capture.rs:84-97)."""
    nx, ny, nz = dims
    idx = np.arange(nx * ny * nz, dtype=np.int64)
    x, y, z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    boundary = (x == 0) | (x + 1 == nx) | (y == 0) | (y + 1 == ny) | (z == 0) | (z + 1 == nz)
    keep = boundary & (idx % 3 != 0)
    sel = idx[keep]
    return np.stack([np.arange(sel.size), sel, x[keep], y[keep], z[keep]], axis=1)


def surface_state(neighbor: np.ndarray, damage: np.ndarray, dims, index: np.ndarray):
    """Per sampled REV: the largest damage of its intact incident bonds and
    the number of its incident lattice bonds that are gone
    (src/bin/solver_scale_capture.rs:20-45). neighbor, damage: (3, n)."""
    nx, ny, nz = dims
    index = np.asarray(index, dtype=np.int64)
    coords = (index % nx, (index // nx) % ny, index // (nx * ny))
    extents, strides = (nx, ny, nz), (1, nx, nx * ny)
    max_damage = np.zeros(index.size, np.float32)
    lost = np.zeros(index.size, np.int64)
    for a in range(3):
        for has, owner in ((coords[a] + 1 < extents[a], index), (coords[a] > 0, index - strides[a])):
            o = np.where(has, owner, 0)
            gone = has & (neighbor[a][o] == EMPTY)
            lost += gone
            intact = has & ~gone
            max_damage = np.where(intact, np.maximum(max_damage, damage[a][o].astype(np.float32)), max_damage)
    return max_damage, lost


def cluster_count(neighbor: np.ndarray, kinematic: np.ndarray) -> int:
    """Connected components over the intact bonds that hold at least one
    non-kinematic REV (rebuild_clusters, src/solver.rs:828-858)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    n = neighbor.shape[1]
    src, dst = [], []
    for a in range(3):
        ok = neighbor[a] != EMPTY
        src.append(np.nonzero(ok)[0])
        dst.append(neighbor[a][ok].astype(np.int64))
    s, d = np.concatenate(src), np.concatenate(dst)
    g = coo_matrix((np.ones(s.size, np.int8), (s, d)), shape=(n, n))
    _, labels = connected_components(g, directed=False)
    return int(np.unique(labels[np.asarray(kinematic) == 0]).size)


class CaptureWriter:
    """The capture CSV of src/bin/solver_scale_capture.rs:113-149: an H line
    (dt, steps, saved frames), then per saved frame an F line (frame, the
    tool pose fields as the capture writes them for a scene without a tool,
    broken bonds, clusters) and a V line per surface sample (frame, sample
    id, position, damage, lost bonds, REV index)."""

    def __init__(self, path: str, dt: float, steps: int, saved_frames: int):
        if steps <= 0 or saved_frames < 2 or saved_frames > steps + 1:
            raise ValueError("steps must be positive and saved_frames must be in 2..=steps+1")
        self.f = open(path, "w", newline="\n")
        self.steps, self.saved_frames = int(steps), int(saved_frames)
        self.f.write("H,%s,%d,%d\n" % (fmt_f32(dt), steps, saved_frames))

    def saves(self, frame: int) -> bool:
        """Whether `frame` is one of the evenly spaced saved frames
        (src/bin/solver_scale_capture.rs:129-133)."""
        if frame == 0 or frame == self.steps:
            return True
        iv = self.saved_frames - 1
        return (frame * iv) // self.steps > ((frame - 1) * iv) // self.steps

    def frame(self, frame: int, broken: int, clusters: int, samples: np.ndarray, positions: np.ndarray,
              damage: np.ndarray, lost: np.ndarray):
        w = self.f.write
        w("F,%d,0,0,0,0,0,0,1,0,0,%d,%d,0\n" % (frame, broken, clusters))
        for (sid, index, _x, _y, _z), d, l_ in zip(samples, damage, lost):
            p = positions[index]
            w("V,%d,%d,%s,%s,%s,%s,%d,%d\n" % (frame, sid, fmt_f32(p[0]), fmt_f32(p[1]), fmt_f32(p[2]), fmt_f32(d),
                                               l_, index))
        self.f.flush()

    def close(self):
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


@dataclasses.dataclass
class CaptureFrame:
    frame: int
    fields: List[str]                 # the F line's fields after the frame number
    broken: Optional[int] = None
    clusters: Optional[int] = None
    samples: Optional[np.ndarray] = None   # V rows: sample id, x, y, z, damage, lost[, index]


def read_capture(path: str) -> dict:
    """Header, frames (F and V records by frame) and every other record kept
    raw, of a capture CSV (the record kinds of validation/live_viewer.py)."""
    header = None
    frames: Dict[int, CaptureFrame] = {}
    vrows: Dict[int, list] = {}
    other: List[List[str]] = []
    with open(path) as f:
        for line in f:
            row = line.rstrip("\n").split(",")
            if not row or row == [""]:
                continue
            kind = row[0]
            if kind == "H":
                header = {"dt": float(row[1]), "steps": int(row[2]), "saved_frames": int(row[3]),
                          "mode": row[4] if len(row) > 4 else "mechanical"}
            elif kind == "F":
                fr = int(row[1])
                rest = row[2:]
                cf = CaptureFrame(fr, rest)
                if len(rest) >= 12:  # the scale capture's layout: ..., broken, clusters, 0
                    cf.broken, cf.clusters = int(rest[9]), int(rest[10])
                frames[fr] = cf
            elif kind == "V":
                vrows.setdefault(int(row[1]), []).append([float(v) for v in row[2:]])
            else:
                other.append(row)
    for fr, rows in vrows.items():
        frames.setdefault(fr, CaptureFrame(fr, [])).samples = np.array(rows)
    return {"header": header, "frames": frames, "other": other}


__all__ = ["ConfigError", "SceneConfig", "FractureSettings", "parse_ron", "scene_config_from_text",
           "load_scene_config", "fmt_f32", "surface_samples", "surface_state", "cluster_count", "CaptureWriter",
           "read_capture", "CaptureFrame"]
