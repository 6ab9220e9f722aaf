"""The synthetic workloads: the slab-local clustered generator realises the
same particles whatever block of slabs is asked for, and the whole-box
realisation has the nominal size and a heavy-tailed occupancy."""
import numpy as np

import synth

KW = dict(grf_mesh=32)          # a coarse displacement mesh keeps the test fast


def test_clustered_slabs_do_not_depend_on_the_block():
    a = synth.clustered_slabs(96, 25, 7, range(0, 7), **KW)
    b = synth.clustered_slabs(96, 25, 7, range(4, 11), **KW)
    for x in range(4, 7):
        assert a.slab(x).n > 0
        assert np.array_equal(a.slab(x).count, b.slab(x).count), x
        assert np.array_equal(a.slab(x).pos, b.slab(x).pos), x
    assert a.slab(12).n == 0 and b.slab(0).n == 0


def test_clustered_slabs_wrap_and_full_box():
    full = synth.clustered_slabs(96, 25, 7, range(25), **KW)
    wrapped = synth.clustered_slabs(96, 25, 7, range(-2, 3), **KW)
    for x in (-2, -1, 0, 1, 2):
        assert np.array_equal(full.slab(x).pos, wrapped.slab(x).pos), x
    n = sum(s.n for s in full.slabs)
    assert abs(n - 96 ** 3) < 0.002 * 96 ** 3           # Bernoulli lattice + exact halo budget
    assert full.N == 96 ** 3
    c = full.counts()
    assert c.min() < 0.2 * c.mean() and c.max() > 20 * c.mean()     # underdense cells and halos
    # positions are cell-relative: |x| <= half a cell
    h = 0.5 / 25
    assert all(np.abs(s.pos).max() <= h * (1 + 1e-6) for s in full.slabs if s.n)


def test_box_from_global_orders_cells():
    rng = np.random.default_rng(3)
    pos = rng.random((20000, 3)) - 0.5
    box = synth.box_from_global(pos, 9)
    assert sum(s.n for s in box.slabs) == 20000
    for x in (0, 4, 8):
        s = box.slab(x)
        xyz = s.xyz.astype(np.float64)
        # rebuild global positions from cell-relative ones and check the cell of each
        cy = np.repeat(np.arange(81) // 9, s.count)
        cz = np.repeat(np.arange(81) % 9, s.count)
        g = xyz + np.stack([synth.cell_centre(np.full(s.n, x), 9), synth.cell_centre(cy, 9),
                            synth.cell_centre(cz, 9)], axis=1)
        cell = np.floor((g + 0.5) * 9).astype(int)
        assert (cell[:, 0] == x).all() and (cell[:, 1] == cy).all() and (cell[:, 2] == cz).all()


def test_reference_ewald_fixture():
    """tests/golden/ewald_pos_f32.npy: the 65,536 positions of the reference's
    Ewald test, inside the unit box, binned into a box the planner accepts;
    the planner's DirectTotal over all slabs equals the brute-force count
    (CalcDI) of the oracle."""
    import os
    pos = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ewald_pos_f32.npy"))
    assert pos.shape == (65536, 3) and pos.dtype == np.float32
    assert np.all(pos >= -0.5) and np.all(pos < 0.5)
    box = synth.box_from_global(pos.astype(np.float64), 11)
    assert box.N == 65536 and box.counts().sum() == 65536
    import nearfield as nfm
    import pyoracle as po
    from nfhelpers import R, params
    eps, b2, _ = params(box)
    for x in (0, 5, 10):
        sic = nfm.SetInteractionCollection(box, x, 0, box.cpd, R, eps, b2, "single_spline")
        assert sic.desc.DirectTotal == po.calc_di(box, R, x)



def test_clustered_slabs_parallel_matches_serial():
    """The forked-pool generator gives the serial generator's slabs, bit for
    bit, whatever the chunking (each slab depends only on the seed)."""
    want = [-1, 0, 1, 2, 5, 6]
    ref = synth.clustered_slabs(120, 31, 77, want, **KW)
    for chunk in (1, 4):
        par = synth.clustered_slabs_parallel(120, 31, 77, want, workers=3, chunk=chunk, **KW)
        assert par.N == ref.N == 120 ** 3
        for x in range(31):
            np.testing.assert_array_equal(par.slab(x).count, ref.slab(x).count)
            np.testing.assert_array_equal(par.slab(x).pos, ref.slab(x).pos)
        assert par.slab(2).n > 0 and par.slab(10).n == 0
