"""The oracle's FoF catalogue (fof.py:510-519 group order, :650-766 reductions) pinned on the CPU:
the reference's documented cosmo-2D answer (164 groups of >= 10 particles in a catalogue of 1639
slots, docs/quickstart.md:219-251), closed forms for centres of mass, inertia radii and means, and
the group-order invariants."""
import numpy as np

from conftest import cosmo2d


def test_cosmo2d_catalogue_known_answer(oracle):
    pos = cosmo2d()
    pz, _ = oracle.zsort(pos)
    lab = oracle.fof(pz, 0.2 * 100. / 128, boxsize=100.0, key=np.arange(len(pz)))
    isort, counts = oracle.fof_order(lab)
    cat = oracle.fof_catalogue(pz[isort], counts, npart_min=10, boxsize=100.0, mass=1.0)
    assert cat["ngroups"] == 164
    assert len(cat["mass"]) == 1639  # catalogue num, shape: 164 (1639,)
    assert np.all(cat["count"][:164] >= 10) and np.all(cat["count"][164:] == 0)
    assert np.array_equal(cat["mass"][:164], cat["count"][:164].astype(np.float32))
    assert np.all(np.isnan(cat["com_pos"][164:])) and np.all(cat["offset"][164:] == len(pz))


def test_group_order_invariants(oracle):
    rng = np.random.default_rng(3)
    n = 5000
    # labels as FoF produces them: each the smallest index of its group
    parent = np.arange(n)
    for _ in range(3000):
        a, b = sorted(rng.integers(0, n, 2))
        ra, rb = parent[a], parent[b]
        lo, hi = min(ra, rb), max(ra, rb)
        parent[parent == hi] = lo
    lab = parent
    npart = n - 37
    isort, counts = oracle.fof_order(lab, npart=npart)
    assert sorted(isort.tolist()) == list(range(n))
    l_sorted = np.where(isort < npart, lab[isort], n)
    assert np.all(np.diff(l_sorted) >= 0)  # groups contiguous, ordered by their label
    starts = np.nonzero(counts)[0]
    assert np.array_equal(isort[starts], l_sorted[starts])  # a group starts at its label (first member)
    same = l_sorted[1:] == l_sorted[:-1]
    assert np.all(isort[1:][same] > isort[:-1][same])  # members in z-order
    ref = np.bincount(lab[:npart], minlength=n)
    assert np.array_equal(counts[starts], ref[isort[starts]])
    assert counts.sum() == npart


def _grouped(groups):
    """Particles of several groups in group order, with counts at each group's first member."""
    pos = np.concatenate(groups)
    counts = np.zeros(len(pos), np.int64)
    o = 0
    for g in groups:
        counts[o] = len(g)
        o += len(g)
    return pos, counts


def test_closed_forms_3d(oracle):
    h, c = 0.25, np.array([3.0, -2.0, 5.0])
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], float) * h + c
    line = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [2.0, 0.0, 0.0], [3.0, 0.0, 0.0]]) + 10.0
    tiny = np.array([[50.0, 50.0, 50.0], [50.5, 50.0, 50.0]])  # below npart_min
    pos, counts = _grouped([corners, line, tiny])
    pos = pos.astype(np.float64)
    mass = np.array([1.0] * 8 + [1.0, 1.0, 1.0, 5.0] + [1.0, 1.0])
    vel = np.zeros_like(pos)
    vel[8:12, 1] = [1.0, 2.0, 3.0, 4.0]
    cat = oracle.fof_catalogue(pos, counts, npart_min=4, mass=mass, vel=vel, v_rad=np.full(len(pos), 2.0),
                               zplus1=np.full(len(pos), 4.0))
    assert cat["ngroups"] == 2 and list(cat["count"][:2]) == [8, 4] and list(cat["offset"][:2]) == [0, 8]
    assert np.allclose(cat["com_pos"][0], c, rtol=0, atol=1e-14)
    assert np.isclose(cat["com_inertia_radius"][0], h * np.sqrt(3.0), rtol=1e-14)
    # weighted line: masses 1,1,1,5 at x = 10..13 -> x_com = (10+11+12+65)/8
    assert np.isclose(cat["com_pos"][1][0], (10 + 11 + 12 + 5 * 13) / 8.0, rtol=1e-14)
    assert np.isclose(cat["mass"][1], 8.0)
    xs = np.array([10.0, 11.0, 12.0, 13.0]) - (10 + 11 + 12 + 65) / 8.0
    assert np.isclose(cat["com_inertia_radius"][1], np.sqrt(np.sum(mass[8:12] * xs ** 2) / 8.0), rtol=1e-13)
    assert np.isclose(cat["com_vel"][1][1], (1 + 2 + 3 + 20) / 8.0, rtol=1e-14)
    assert np.isclose(cat["v_rad"][0], 2.0) and np.isclose(cat["scale_factor"][1], 0.25)


def test_periodic_group_across_the_boundary(oracle):
    """A group straddling x = 0 of a box of size 1: its centre lies on the boundary (wrapped into
    [0, 1)), and its inertia radius is that of the unwrapped points."""
    pts = np.array([[0.98, 0.5, 0.5], [0.99, 0.5, 0.5], [0.01, 0.5, 0.5], [0.02, 0.5, 0.5]], np.float64)
    pos, counts = _grouped([pts])
    cat = oracle.fof_catalogue(pos, counts, npart_min=2, boxsize=1.0, mass=1.0)
    x = cat["com_pos"][0][0]
    assert min(abs(x), abs(x - 1.0)) < 1e-12 and 0.0 <= x < 1.0
    assert np.isclose(cat["com_inertia_radius"][0], np.sqrt((0.02 ** 2 + 0.01 ** 2) * 2 / 4), rtol=1e-9)
    # the same group without the box: the centre is the plain mean
    cat0 = oracle.fof_catalogue(pos, counts, npart_min=2, mass=1.0)
    assert np.isclose(cat0["com_pos"][0][0], 0.5)


def test_2d_inertia_counts_second_component_twice(oracle):
    """fof.py:736 sums dx[...,0]^2 + dx[...,1]^2 + dx[...,2]^2; on 2D particles JAX clamps index 2
    to 1."""
    pts = np.array([[0.0, 0.0], [2.0, 0.0], [1.0, 1.0], [1.0, -1.0]])
    pos, counts = _grouped([pts])
    cat = oracle.fof_catalogue(pos, counts, npart_min=1, mass=1.0)
    assert np.allclose(cat["com_pos"][0], [1.0, 0.0])
    # dx^2 + 2 dy^2 = 1, 1, 2, 2 -> mean 1.5
    assert np.isclose(cat["com_inertia_radius"][0], np.sqrt(1.5))
