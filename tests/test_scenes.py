"""Scene generators (CPU): the workloads' stated shapes, counts and separations."""
import numpy as np
import pytest

from dem_mph_amd import scenes as S


def test_c2_counts():
    sc = S.dam_break_c2(seed=1)
    c = sc.counts()
    assert (c["fluid"], c["solid"], c["wall"], c["bodies"]) == (562500, 64000, 340908, 64)
    rho = sc.density[sc.type == S.SOLID]
    assert 500.0 <= rho.min() and rho.max() <= 900.0


def test_tank_walls_are_the_padded_lattice_without_its_interior():
    """The plane-by-plane wall builder against the definition, brute force."""
    for interior, d, closed, dim in (((0.2, 0.1, 0.15), 0.01, False, 3), ((0.2, 0.1, 0.15), 0.01, True, 3),
                                     ((0.3, 0.2, 0.01), 0.01, False, 2)):
        L = np.asarray(interior)
        lo = -3 * d * np.ones(3)
        hi = L + 3 * d
        if not closed:
            hi[2 if dim == 3 else 1] = L[2 if dim == 3 else 1]
        if dim == 2:
            lo[2], hi[2] = 0.0, d
        full = S.lattice(lo, hi, d, dim)
        inside = np.all([(full[:, a] > 0) & (full[:, a] < L[a]) for a in range(dim)], axis=0)
        np.testing.assert_array_equal(S.tank_walls(interior, d, 3, closed, dim), full[~inside])


def test_c5_counts_at_full_size():
    """8.0M fluid (100 x 400 x 200), 2048 bodies of 4-12 per side (~1.05M), ~3.9M
    walls (open top); bodies inside the channel, clear of the fluid column and of
    one another."""
    sc = S.channel_c5(seed=1)
    c = sc.counts()
    assert c["fluid"] == 8_000_000 and c["bodies"] == 2048
    assert 0.95e6 < c["solid"] < 1.15e6 and 3.8e6 < c["wall"] < 4.0e6
    sizes = np.bincount(sc.body[sc.body >= 0])
    assert sizes.min() >= 64 and sizes.max() <= 1728
    deb = sc.pos[sc.type == S.SOLID]
    assert deb[:, 0].min() > 0.3 - 0.01 * sc.d and deb[:, 0].max() < 1.6 and deb[:, 1].max() < 0.8
    assert deb[:, 2].min() > 0.0
    rho = sc.density[sc.type == S.SOLID]
    assert 500.0 <= rho.min() and rho.max() <= 900.0


def test_c5_small_scale_is_seeded_and_separated():
    from scipy.spatial import cKDTree
    a = S.channel_c5(seed=3, scale=0.25, nbodies=40)
    b = S.channel_c5(seed=3, scale=0.25, nbodies=40)
    np.testing.assert_array_equal(a.pos, b.pos)
    assert a.counts()["bodies"] == 40 and a.counts()["fluid"] == 25 * 100 * 50
    solid = a.pos[a.type == S.SOLID]
    other = a.pos[a.type != S.SOLID]
    # no debris particle within half a spacing of any other particle of another body or of the rest
    assert len(cKDTree(solid).query_ball_tree(cKDTree(other), 0.5 * a.d)[0]) == 0
    body = a.body[a.type == S.SOLID]
    pairs = cKDTree(solid).query_pairs(0.99 * a.d, output_type="ndarray")
    assert (body[pairs[:, 0]] == body[pairs[:, 1]]).all()
    with pytest.raises(ValueError):
        S.channel_c5(seed=3, scale=0.25, nbodies=10_000)


@pytest.fixture(scope="module")
def ensemble():
    return S.ensemble_c3(512, seed=0)


def test_ensemble_shares_one_domain_and_spacing(ensemble):
    lo, hi = ensemble[0].domain_min, ensemble[0].domain_max
    for sc in ensemble:
        assert sc.d == 0.01
        np.testing.assert_array_equal(sc.domain_min, lo)
        np.testing.assert_array_equal(sc.domain_max, hi)
        assert (sc.pos > lo).all() and (sc.pos < hi).all()


def test_ensemble_fluid_counts_log_uniform(ensemble):
    nf = np.array([(sc.type < 2).sum() for sc in ensemble])
    # log-uniform on [2000, 60000]: mean (60000 - 2000) / ln 30 = 17,053; column rounding shifts it slightly
    assert 1900 <= nf.min() and nf.max() <= 61000
    assert 15000 < nf.mean() < 19000
    # roughly uniform in log N: each of three log-equal bins holds about a third
    b = np.digitize(np.log(nf), np.log([2000.0, 6214.0, 19311.0, 60001.0]))
    frac = np.bincount(b, minlength=5)[1:4] / len(nf)
    assert (frac > 0.25).all()


def test_ensemble_debris_bodies(ensemble):
    nb = np.array([sc.body.max() + 1 for sc in ensemble])
    assert nb.min() >= 1 and nb.max() <= 32
    for sc in ensemble[:64]:
        sol = (sc.type >= 2) & (sc.type < 4)
        body = sc.body[sol]
        for b in np.unique(body):
            pts = sc.pos[sol][body == b]
            ext = np.round((pts.max(axis=0) - pts.min(axis=0)) / sc.d).astype(int) + 1
            assert ((3 <= ext) & (ext <= 8)).all()
            r = np.unique(sc.density[sol][body == b])
            assert len(r) == 1 and 400.0 <= r[0] <= 950.0
        # debris starts beyond RadiusP (2.5 d) of the fluid column (the drag is NaN at zero relative velocity)
        fl = sc.pos[sc.type < 2]
        assert sc.pos[sol][:, 0].min() - fl[:, 0].max() > 2.5 * sc.d


def test_ensemble_is_seeded():
    a = S.ensemble_c3(3, seed=5, d=0.02, interior=(0.6, 0.3, 0.4), fluid_range=(200.0, 3000.0))
    b = S.ensemble_c3(3, seed=5, d=0.02, interior=(0.6, 0.3, 0.4), fluid_range=(200.0, 3000.0))
    c = S.ensemble_c3(3, seed=6, d=0.02, interior=(0.6, 0.3, 0.4), fluid_range=(200.0, 3000.0))
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.pos, y.pos)
    assert any(x.n != z.n or not np.array_equal(x.pos, z.pos) for x, z in zip(a, c))


def test_surrogate_training_windows():
    """Training windows: HIST + 1 frames STRIDE steps apart; every window yields
    exactly one sample frame (the one with 4 frames before and 1 after)."""
    from dem_mph_amd import surrogate as SG
    st = SG.window_steps(50, 300, 4)
    assert len(st) == 4 * (SG.HIST + 1)
    assert st[:6] == [50, 60, 70, 80, 90, 100] and st[6] == 350
    assert SG.sample_frames(st) == [4, 10, 16, 22]
    # a continuous record: every frame with 4 before and 1 after
    assert SG.sample_frames(list(range(0, 100, 10))) == [4, 5, 6, 7, 8]
    # a gap breaks the history: only frame 7 (75) has 4 evenly spaced frames before it and one after
    assert SG.sample_frames([0, 10, 20, 35, 45, 55, 65, 75, 85]) == [7]
