"""Scene generator (.boid -> .grid; generator/generator.cpp) against the
reference generator's own outputs.

Golden vectors: tests/golden/boid_golden.json, written by
tests/golden/make_boid_golden.py from the reference generator built from its
sources (oracle/_ref/GeneratorForMph); sysu_2mm is the reference's
results/SYSU/2mm/DEM.boid, whose output equals the DEM.grid shipped beside it.
The fixtures cover every shape kind, in 2D and 3D, in a file order that differs
from the generation order, and the generator's error paths.  Particle counts
are also checked against closed forms computed here from the shape sizes.
"""
import hashlib
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")
BOID = os.path.join(GOLDEN_DIR, "boid")
GOLDEN = json.load(open(os.path.join(GOLDEN_DIR, "boid_golden.json")))
CLI = os.path.join(ROOT, "dem_mph_amd", "GeneratorForMph")
REF_GEN = os.path.join(ROOT, "oracle", "_ref", "GeneratorForMph")


def _run(exe, name, tmp_path):
    shutil.copy(os.path.join(BOID, name + ".boid"), tmp_path / "scene.boid")
    r = subprocess.run([exe, "scene"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return (tmp_path / "scene.grid").read_bytes()


@pytest.mark.skipif(not os.path.exists(CLI), reason="GeneratorForMph not built")
@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_cli_writes_the_generators_bytes(name, tmp_path):
    data = _run(CLI, name, tmp_path)
    g = GOLDEN[name]
    assert len(data) == g["bytes"]
    assert hashlib.sha256(data).hexdigest() == g["sha256"]
    assert int(data.split(b"\n")[1].split()[0]) == g["n"]


@pytest.mark.skipif(not (os.path.exists(CLI) and os.path.exists(REF_GEN)), reason="reference generator not built")
@pytest.mark.parametrize("name", ["kinds3d", "kinds2d", "err_unknown_key", "err_missing_key"])
def test_cli_equals_reference_generator(name, tmp_path):
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    assert _run(CLI, name, tmp_path / "a") == _run(REF_GEN, name, tmp_path / "b")


@pytest.mark.parametrize("name", ["kinds3d", "kinds2d", "sysu_2mm"])
def test_library_generates_and_writes_the_same_file(name, tmp_path):
    import dem_mph_amd as D
    out = tmp_path / "x.grid"
    sc = D.generate_boid(os.path.join(BOID, name + ".boid"), str(out))
    data = out.read_bytes()
    assert hashlib.sha256(data).hexdigest() == GOLDEN[name]["sha256"]
    assert sc.n == GOLDEN[name]["n"] and sc.time == 0.0
    # the written file reads back through readGridFile's restatement
    back = D.read_grid(str(out))
    assert back.n == sc.n
    assert np.array_equal(back.type, sc.type)
    assert np.allclose(back.pos, sc.pos, rtol=1e-6, atol=1e-9)
    assert np.allclose(back.vel, sc.vel, rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("name", ["err_unknown_key", "err_missing_key"])
def test_parse_errors_raise(name):
    import dem_mph_amd as D
    with pytest.raises(D.DmphError):
        D.generate_boid(os.path.join(BOID, name + ".boid"))


def test_end_of_file_inside_a_block_is_an_error(tmp_path):
    """The generator loops forever here; the library reports it."""
    import dem_mph_amd as D
    p = tmp_path / "cut.boid"
    p.write_text("ParticleDistance 0.01\nStartCuboid\n Spacing 0.01\n Type 1\n Lower 0 0 0\n")
    with pytest.raises(D.DmphError, match="end of file"):
        D.generate_boid(str(p))
    with pytest.raises(D.DmphError):
        D.generate_boid(str(tmp_path / "missing.boid"))


def _types(name):
    import dem_mph_amd as D
    return D.generate_boid(os.path.join(BOID, name + ".boid"))


def test_counts_follow_the_shape_sizes():
    """kinds3d: one type per kind; counts from the shapes' sizes."""
    sc = _types("kinds3d")
    cnt = {t: int((sc.type == t).sum()) for t in range(6)}
    rnd = lambda w, s: int(math.floor(w / s + 0.5))
    # Cuboids (type 1): round(w / s) points per axis
    assert cnt[1] == rnd(0.10, 0.01) * rnd(0.07, 0.01) * rnd(0.05, 0.01) + rnd(0.05, 0.007) ** 2 * rnd(0.03, 0.007)
    # Cuboid2 (type 2): the same lattice, rotated
    assert cnt[2] == 6 * 6 * 4
    # Recboid2 (type 3): the full lattice, rotated about the origin
    assert cnt[3] == 8 * 6 * 3
    # Cyboid (type 5): floor(w / (min w / sqrt 3)) per axis
    side = 0.04 / math.sqrt(3.0)
    assert cnt[5] == math.floor(0.06 / side) * math.floor(0.05 / side) * math.floor(0.04 / side)
    # Cyboid2 (type 4): rings of max(6, int(2 pi r / s)) at z = 0.005, 0.015, ... plus top and bottom
    r1, r2, h, s = 0.04, 0.02, 0.06, 0.01
    zs = [0.005 + k * s for k in range(6)]
    ring = lambda r: max(6, int(2.0 * math.pi * r / s))
    assert cnt[4] == sum(ring(r2 + (r1 - r2) * z / h) for z in zs) + ring(r1) + ring(r2)
    # Recboid (type 0): lattice points below the line y = tan(30 deg) x
    xs = 0.01 + 0.0001 + 0.01 * np.arange(8)
    assert cnt[0] == 2 * int(sum((np.tan(30 * 3.1415 / 180) > y / x) for x in xs for y in xs))


def test_rotations_keep_distances():
    """Cuboid2 rotates about the box centre, Recboid2 about the origin: the
    rotated sets keep the unrotated lattice's pairwise distances and (Cuboid2)
    centroid."""
    sc = _types("kinds3d")
    c2 = sc.pos[sc.type == 2]
    a = 0.3
    R = np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])
    back = (c2[:, :2] - [0.15, 0.03]) @ R + [0.15, 0.03]   # inverse rotation (R^T applied to row vectors)
    xs = 0.12 + 0.0001 + 0.01 * np.arange(6)
    ys = 0.0001 + 0.01 * np.arange(6)
    grid = np.array([(x, y) for x in xs for y in ys for _ in range(4)])
    assert np.allclose(back, grid, atol=1e-12)
    r2 = sc.pos[sc.type == 3]
    b = 15 * 3.1415 / 180
    Rb = np.array([[math.cos(b), -math.sin(b)], [math.sin(b), math.cos(b)]])
    un = r2[:, :2] @ Rb   # rotate back
    xs = 0.02 + 0.0001 + 0.01 * np.arange(8)
    ys = 0.02 + 0.0001 + 0.01 * np.arange(6)
    assert np.allclose(un, np.array([(x, y) for x in xs for y in ys for _ in range(3)]), atol=1e-12)


def test_generation_order_is_kind_by_kind():
    sc = _types("kinds3d")
    order = [1, 2, 5, 4, 0, 3]   # Cuboid, Cuboid2, Cyboid, Cyboid2, Recboid, Recboid2 (types in kinds3d)
    firsts = [int(np.argmax(sc.type == t)) for t in order]
    assert firsts == sorted(firsts)
