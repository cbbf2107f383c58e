"""Host logic of the surrogate driver (CPU, no GPU): the per-scene debris-centroid
and fluid errors of one trajectory against another, which the bench reports for
the mixed-precision solver against the FP64 solver, and the position-fold
eligibility rule."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from dem_mph_amd import surrogate as SG


def _traj(pos_frames, body, cls, off):
    """A Trajectories holding given frames (CPU tensors), without running a solver."""
    t = object.__new__(SG.Trajectories)
    t.pos = torch.tensor(np.asarray(pos_frames, np.float32))
    t.body_np = np.asarray(body, np.int32)
    t.cls_np = np.asarray(cls, np.int32)
    t.off = np.asarray(off)
    t.nscene = len(off) - 1
    t.nbodies = int(t.body_np.max()) + 1
    return t


def test_frame_errors_against_hand_computed_shifts():
    # scene 0: body 0 (2 particles) and 2 fluid particles; scene 1: body 1 (2 particles) and 1 wall particle
    body = [0, 0, -1, -1, 1, 1, -1]
    cls = [1, 1, 0, 0, 1, 1, 2]
    off = [0, 4, 7]
    rng = np.random.default_rng(0)
    ref = rng.uniform(0.0, 1.0, (3, 7, 3))
    other = ref.copy()
    # frame 1: body 0 moves by (0.003, 0.004, 0) -> 5 mm; fluid particle 2 by 0.01 in x
    other[1, 0:2] += [0.003, 0.004, 0.0]
    other[1, 2, 0] += 0.01
    # frame 2: body 1 moves by 2 mm in z, one of its particles by 4 mm more (centroid: +4 mm)
    other[2, 4, 2] += 0.002
    other[2, 5, 2] += 0.006
    e = SG.frame_errors(_traj(ref, body, cls, off), _traj(other, body, cls, off), [0, 1, 2])
    # per scene maxima: scene 0 debris 5 mm, fluid RMS over 2 particles sqrt(0.01^2 / 2); scene 1 debris 4 mm
    assert e["debris_centroid_rms_m_mean"] == pytest.approx((0.005 + 0.004) / 2, rel=1e-5)
    assert e["debris_centroid_rms_m_max"] == pytest.approx(0.005, rel=1e-5)
    assert e["fluid_rms_m_mean"] == pytest.approx(np.sqrt(0.01 ** 2 / 2) / 2, rel=1e-4)
    # frames not listed are not looked at
    e0 = SG.frame_errors(_traj(ref, body, cls, off), _traj(other, body, cls, off), [0])
    assert e0["debris_centroid_rms_m_max"] == 0.0 and e0["fluid_rms_m_mean"] == 0.0


def test_position_fold_applies_only_to_the_graphed_positions_within_its_reach():
    sur = object.__new__(SG.Surrogate)
    sur.pos_fold, sur.radius = True, 0.03
    pos = torch.zeros((10, 3))
    sur._box = None
    assert sur._fold_args(pos, 10) is None                     # no graph yet
    sur._box = (np.array([0.0, 0.0, 0.0, 0.8, 0.4, 0.6]), pos.data_ptr(), 10)
    f = sur._fold_args(pos, 10)
    assert f is not None
    assert [f.origin[k] for k in range(3)] == pytest.approx([0.4, 0.2, 0.3])
    assert f.inv_r == pytest.approx(1 / 0.03)
    assert sur._fold_args(pos, 9) is None                      # another node count
    assert sur._fold_args(torch.zeros((10, 3)), 10) is None    # other positions
    sur._box = (np.array([0.0, 0.0, 0.0, 1.6, 0.4, 0.6]), pos.data_ptr(), 10)
    assert sur._fold_args(pos, 10) is None                     # 26.7 radii from the centre: beyond the reach
    sur._box = (np.array([0.0, 0.0, 0.0, 0.8, 0.4, 0.6]), pos.data_ptr(), 10)
    sur.pos_fold = False
    assert sur._fold_args(pos, 10) is None                     # switched off
