"""GPU checks of the tile shapes and of the snapshot batch:

  * every tile shape (8^3 / 16^3 tiles x 256 / 512 / 1024 threads) gives the
    reconstruction of the default shape (rel-L2 <= 1e-12; only the FP64
    rounding of the per-ray carries depends on where rays are cut);
  * a batch of uneven grids replayed from one hipGraph gives, snapshot by
    snapshot, bit for bit the result of the snapshot-by-snapshot schedule and
    the iteration count of a single-context reconstruction; replays after
    new fields (with eager work in between) give the new results;
  * the autotuner's choices are valid shapes and keep the results.
"""
import numpy as np
import pytest

import omni3d
from omni3d import fields as F

pytestmark = pytest.mark.gpu

SHAPES = [(20, 22, 24), (16, 16, 16), (24, 20, 18), (20, 22, 24), (16, 16, 16), (24, 20, 18)]


def single(shape, f, dirs, **kw):
    with omni3d.Context(shape, f["delta"], **kw) as ctx:
        ctx.set_directions(dirs)
        ctx.set_field(f["acc"])
        return ctx.reconstruct()


def test_tile_shapes_agree(gpu_available, ref_dirs2):
    shape = (40, 36, 33)
    print("case tile shapes", flush=True)
    f = F.taylor_green(40, noise=0.05, seed=8, shape=shape)
    ref, w_ref, rep_ref = single(shape, f, ref_dirs2, max_iterations=20)
    for bits in (3, 4):
        for threads in (256, 512, 1024):
            with omni3d.Context(shape, f["delta"], max_iterations=20) as ctx:
                ctx.set_tile_shape(bits, threads)
                ctx.set_directions(ref_dirs2)
                ctx.set_field(f["acc"])
                p, w, rep = ctx.reconstruct()
            np.testing.assert_array_equal(w, w_ref)
            assert rep.iterations == rep_ref.iterations
            assert F.rel_l2(p, ref) < 1e-12, (bits, threads)


def fields_for(seed):
    out = []
    rng = np.random.default_rng(seed)
    for shape in SHAPES:
        ph = tuple(rng.uniform(0, 2 * np.pi, 3))
        out.append(F.taylor_green(shape[2], noise=rng.uniform(0, 0.05), seed=int(rng.integers(1 << 30)),
                                  shape=shape, phase=ph))
    return out


def test_batch_graph_matches_sequential(gpu_available, ref_dirs2):
    print("case batch graph", flush=True)
    h = 2 * np.pi / 24
    kw = dict(max_iterations=60, tolerance=1e-5)
    fs = fields_for(1)
    with omni3d.Batch(SHAPES, (h, h, h), **kw) as b:
        b.set_directions(ref_dirs2)
        for i, f in enumerate(fs):
            b.set_field(i, f["acc"])
        b.run(use_graph=False)
        seq = [b.result(i) for i in range(len(SHAPES))]
        for replay in range(2):
            b.run(use_graph=True)
            for i in range(len(SHAPES)):
                p, rep = b.result(i)
                np.testing.assert_array_equal(p, seq[i][0])
                assert rep.iterations == seq[i][1].iterations
        # per-snapshot convergence: each snapshot stops where a lone context does
        for i in (0, 1, 2):
            ps, ws, rs = single(SHAPES[i], dict(fs[i], delta=(h, h, h)), ref_dirs2, **kw)
            assert rs.iterations == seq[i][1].iterations
            assert F.rel_l2(seq[i][0], ps) < 1e-12
        # new fields, eager work in between, replay: the graph reads the new fields
        fs2 = fields_for(2)
        for i, f in enumerate(fs2):
            b.set_field(i, f["acc"])
        b.run(use_graph=False)
        seq2 = [b.result(i)[0] for i in range(len(SHAPES))]
        b.run(use_graph=True)
        for i in range(len(SHAPES)):
            np.testing.assert_array_equal(b.result(i)[0], seq2[i])
            assert F.rel_l2(seq2[i], seq[i][0]) > 1e-3


def test_batch_tuner(gpu_available, ref_dirs2):
    print("case batch tuner", flush=True)
    h = 2 * np.pi / 24
    fs = fields_for(3)
    with omni3d.Batch(SHAPES, (h, h, h), max_iterations=15) as b:
        b.set_directions(ref_dirs2)
        for i, f in enumerate(fs):
            b.set_field(i, f["acc"])
        b.run(use_graph=False)
        before = [b.result(i)[0] for i in range(len(SHAPES))]
        picks = b.tune()
        assert len(picks) == 3
        for bits, threads in picks:
            assert bits in (3, 4) and threads in (256, 512, 1024)
        b.run(use_graph=True)
        for i in range(len(SHAPES)):
            assert F.rel_l2(b.result(i)[0], before[i]) < 1e-12


@pytest.mark.parametrize("method", [omni3d.PLAIN_MASK, omni3d.WEIGHTED])
def test_batch_graph_masked_methods(gpu_available, ref_dirs2, method):
    """The masked methods' tile engine (rule replay, carry chain, tile pass)
    captured in the batch graph: replays give the sequential results (bit for
    bit with unit weights; FP64 curl-weight sums within 1e-12) and match a
    lone context."""
    print(f"case batch graph masked method={method}", flush=True)
    h = 2 * np.pi / 24
    kw = dict(max_iterations=8, tolerance=0.0, method=method)
    fs = fields_for(3)
    masks = []
    for shape in SHAPES:
        Z, Y, X = shape
        k, j, i = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
        masks.append(((i - X / 2) ** 2 + (j - Y / 2) ** 2 + (k - Z / 2) ** 2 > 16).astype(float))
    with omni3d.Batch(SHAPES, (h, h, h), **kw) as b:
        b.set_directions(ref_dirs2)
        for i, f in enumerate(fs):
            b.set_field(i, f["acc"], mask=masks[i])
        b.run(use_graph=False)
        seq = [b.result(i)[0] for i in range(len(SHAPES))]
        b.run(use_graph=True)
        for i in range(len(SHAPES)):
            p = b.result(i)[0]
            if method == omni3d.PLAIN_MASK:
                np.testing.assert_array_equal(p, seq[i])
            else:
                assert F.rel_l2(p, seq[i]) < 1e-12
        for i in (0, 2):
            with omni3d.Context(SHAPES[i], (h, h, h), **kw) as ctx:
                ctx.set_directions(ref_dirs2)
                ctx.set_field(fs[i]["acc"], mask=masks[i])
                ps = ctx.reconstruct()[0]
            assert F.rel_l2(seq[i], ps) < 1e-12
