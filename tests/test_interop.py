"""Interop formats: SceneConfig RON files and the H/F/V capture CSV.

Fixtures are the reference's own data files: scenes/wood_chop.ron (loaded
by the reference's tests, src/fracture.rs:325-329, src/solver.rs:886),
scenes/mounted_axe.ron, and the first lines of validation/melting.csv (a
thermal capture). The capture's sampling, surface state and cluster count
are checked against plain-loop restatements of
src/bin/solver_scale_capture.rs:20-45, 84-97, 129-133 and
src/solver.rs:828-858, and against the reference test
damage_splits_connectivity (src/solver.rs:1206-1217)."""
import os

import numpy as np
import pytest

from vxsim import interop, scenes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
EMPTY = interop.EMPTY


def f32(x):
    return float(np.float32(x))


def test_wood_chop_config():
    c = interop.load_scene_config(os.path.join(GOLDEN, "wood_chop.ron"))
    assert c.dimensions == (30, 30, 300) and c.branch_dimensions == (40, 4, 4)
    assert c.ground_dimensions == (120, 120, 2) and c.tool_dimensions == (50, 5, 20)
    assert c.fibers == (0.0, 0.0, 1.0) and c.anisotropy == 1.0
    assert c.gravity == (0.0, 0.0, f32(-9.81))
    assert c.fixed_dt == f32(0.016666668) and c.substeps == 4 and c.iterations == 40
    assert c.solver_relative_tolerance == f32(1e-5) and c.corotation_tolerance == f32(1e-5)
    assert c.corotation_iterations == 4 and c.initial_speed == f32(8.5) and c.tool_start == (f32(-0.65), 0.0, 1.0)
    # serde defaults of the two fields the file leaves out
    assert c.cuda_iterations == 160 and c.cpu_preconditioner == "IncompleteCholesky0"
    fr = c.fracture
    assert (fr.propagation_delay, fr.fragment_min_size, fr.cluster_interval) == (1, 10, 10)
    assert (fr.friction, fr.restitution) == (f32(0.35), f32(0.1))
    # the fracture law's settings are the ones the fracture tests load
    assert c.fracture.fracture_config() == scenes.FractureConfig(
        f32(35.0), f32(0.005), f32(1000.0), f32(3.0), f32(0.5), f32(0.9))
    sc = c.step_config()
    assert (sc.fixed_dt, sc.substeps, sc.corotation_iterations) == (c.fixed_dt, 4, 4)


def test_mounted_axe_config():
    c = interop.load_scene_config(os.path.join(GOLDEN, "mounted_axe.ron"))
    assert c.dimensions == (12, 12, 30) and c.branch_dimensions == (0, 0, 0)
    assert c.ground_dimensions == (80, 80, 1) and c.tool_dimensions == (20, 5, 14)
    assert c.fixed_dt == f32(0.001) and c.iterations == 80 and c.corotation_iterations == 2
    assert c.initial_speed == 1.0 and c.tool_start == (f32(-0.25), 0.0, f32(0.15))


def _text():
    with open(os.path.join(GOLDEN, "mounted_axe.ron")) as f:
        return f.read()


@pytest.mark.parametrize("edit", [
    ("anisotropy: 1.0,", "anisotropy: 1.0, colour: 3,"),        # unknown field
    ("anisotropy: 1.0,", "anisotropy: 1.0, anisotropy: 0.5,"),  # duplicate field
    ("    substeps: 4,\n", ""),                                   # missing field
    ("dimensions: (12, 12, 30)", "dimensions: (12, 0, 30)"),
    ("tool_dimensions: (20, 5, 14)", "tool_dimensions: (20, 1, 14)"),
    ("anisotropy: 1.0", "anisotropy: 1.5"),
    ("initial_speed: 1.0", "initial_speed: -1.0"),
    ("fixed_dt: 0.001", "fixed_dt: 0.0"),
    ("fibers: (0.0, 0.0, 1.0)", "fibers: (0.0, 0.0, 0.0)"),
    ("substeps: 4", "substeps: 4.0"),                             # integer field given a float
    ("restitution: 0.1", "restitution: 2.0"),
    ("fatigue_threshold: 0.5", "fatigue_threshold: 1.0"),
    ("fatigue_exponent: 3.0", "fatigue_exponent: 0.5"),
    ("cluster_interval: 10", "cluster_interval: 0"),
    ("anisotropy: 1.0,", "anisotropy: 1.0, cpu_preconditioner: Multigrid,"),
])
def test_invalid_configs_are_rejected(edit):
    text = _text()
    assert edit[0] in text
    with pytest.raises(interop.ConfigError):
        interop.scene_config_from_text(text.replace(edit[0], edit[1], 1))


def test_optional_fields_and_named_struct():
    text = _text().replace("anisotropy: 1.0,", "anisotropy: 1.0, cuda_iterations: 300, cpu_preconditioner: BlockJacobi,")
    c = interop.scene_config_from_text("SceneConfig" + text.split("\n", 1)[1])
    assert c.cuda_iterations == 300 and c.cpu_preconditioner == "BlockJacobi"


def test_ron_subset():
    v = interop.parse_ron("/* c */ Foo(a: 1, b: [1.5, -2e3,], c: (true, x), d: \"s\",) // end")
    assert v == {"a": 1, "b": [1.5, -2000.0], "c": (True, ("enum", "x")), "d": "s"}
    with pytest.raises(interop.ConfigError):
        interop.parse_ron("(a: 1) (b: 2)")


def test_f32_formatting_is_shortest_round_trip():
    assert [interop.fmt_f32(x) for x in (0.001, -9.81, 1.0, 0.016666668, 1e-7, 0.0, 1234.5)] == \
        ["0.001", "-9.81", "1", "0.016666668", "0.0000001", "0", "1234.5"]
    rng = np.random.default_rng(1)
    for x in rng.standard_normal(200).astype(np.float32) * np.float32(1e-2):
        assert np.float32(float(interop.fmt_f32(x))) == x


def test_surface_samples_and_state_match_loops():
    dims = (5, 4, 3)
    nx, ny, nz = dims
    want = []
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                index = (z * ny + y) * nx + x
                boundary = x == 0 or x + 1 == nx or y == 0 or y + 1 == ny or z == 0 or z + 1 == nz
                if boundary and index % 3 != 0:
                    want.append((len(want), index, x, y, z))
    got = interop.surface_samples(dims)
    assert got.tolist() == [list(w) for w in want]
    # the capture's own box: 36 is a multiple of 3, so index % 3 == x % 3
    big = interop.surface_samples((36, 32, 32))
    count = sum(1 for z in range(32) for y in range(32) for x in range(36)
                if (x in (0, 35) or y in (0, 31) or z in (0, 31)) and x % 3 != 0)
    assert big.shape[0] == count

    s = scenes.uniform_box(dims, "stone")
    rng = np.random.default_rng(3)
    nb = s.neighbor.copy()
    broken = (rng.random(nb.shape) < 0.3) & (nb != EMPTY)
    nb[broken] = EMPTY
    dmg = (rng.random(nb.shape) * 0.9).astype(np.float32)
    md, lost = interop.surface_state(nb, dmg, dims, got[:, 1])
    strides = (1, nx, nx * ny)
    for (sid, index, x, y, z), m, l_ in zip(want, md, lost):
        c, e = (x, y, z), dims
        m_ref, l_ref = 0.0, 0
        for a in range(3):
            for has, owner in ((c[a] + 1 < e[a], index), (c[a] > 0, index - strides[a])):
                if not has:
                    continue
                if nb[a][owner] == EMPTY:
                    l_ref += 1
                else:
                    m_ref = max(m_ref, float(dmg[a][owner]))
        assert (float(m), int(l_)) == (np.float32(m_ref), l_ref)


def test_cluster_count():
    s = scenes.uniform_box((2, 2, 2), "iron")
    assert interop.cluster_count(s.neighbor, s.kinematic) == 1
    # damage_splits_connectivity: every bond gone -> 8 clusters
    assert interop.cluster_count(np.full_like(s.neighbor, EMPTY), s.kinematic) == 8
    # kinematic REVs connect but are not counted
    b = scenes.uniform_box((3, 1, 1), "iron")
    b.kinematic[1] = 1
    assert interop.cluster_count(b.neighbor, b.kinematic) == 1
    nb = b.neighbor.copy()
    nb[0][0] = EMPTY
    assert interop.cluster_count(nb, b.kinematic) == 2


@pytest.mark.parametrize("steps,saved", [(176, 177), (176, 12), (10, 2), (7, 5), (1, 2)])
def test_saved_frames(tmp_path, steps, saved):
    w = interop.CaptureWriter(str(tmp_path / "c.csv"), 0.001, steps, saved)
    frames = [f for f in range(steps + 1) if w.saves(f)]
    w.close()
    iv = saved - 1
    ref = [f for f in range(steps + 1)
           if f == 0 or f == steps or (f * iv) // steps > ((f - 1) * iv) // steps]
    assert frames == ref and len(frames) == saved


def test_capture_round_trip(tmp_path):
    path = str(tmp_path / "c.csv")
    dims = (4, 3, 3)
    s = scenes.uniform_box(dims, "wood")
    samples = interop.surface_samples(dims)
    with interop.CaptureWriter(path, 0.001, 5, 2) as w:
        for frame in (0, 5):
            md, lost = interop.surface_state(s.neighbor, s.damage, dims, samples[:, 1])
            w.frame(frame, 3 * frame, 1, samples, s.positions, md, lost)
    with open(path) as f:
        lines = f.read().splitlines()
    assert lines[0] == "H,0.001,5,2"
    assert lines[1] == "F,0,0,0,0,0,0,0,1,0,0,0,1,0"
    cap = interop.read_capture(path)
    assert cap["header"] == {"dt": 0.001, "steps": 5, "saved_frames": 2, "mode": "mechanical"}
    assert sorted(cap["frames"]) == [0, 5] and cap["frames"][5].broken == 15 and cap["frames"][5].clusters == 1
    v = cap["frames"][5].samples
    assert v.shape == (samples.shape[0], 7)
    assert np.array_equal(v[:, 0], samples[:, 0]) and np.array_equal(v[:, 6], samples[:, 1])
    assert np.array_equal(v[:, 1:4].astype(np.float32), s.positions[samples[:, 1]])


def test_reads_the_reference_thermal_capture():
    cap = interop.read_capture(os.path.join(GOLDEN, "melting_head.csv"))
    assert cap["header"] == {"dt": 0.001, "steps": 1680, "saved_frames": 176, "mode": "thermal"}
    kinds = {r[0] for r in cap["other"]}
    assert "S" in kinds and "Q" in kinds and not cap["frames"]
