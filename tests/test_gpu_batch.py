"""GPU tests of the batched 2-D path against the per-patch drop-ins.

Bars: the batched structure-oriented median is bit-identical to somf2dc run
patch by patch (same sprays, same order statistic); the batched dip
matches dip2dc patch by patch within 1e-4 samples/trace (only the order of
the double-precision CG sums differs) with the same CG / line-search
decisions, for every smoothing tile shape; and a few patches are checked
against the oracle with the dip bars of the per-patch parity tests.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(128, 64), (192, 96), (256, 70), (130, 65), (512, 64), (128, 256), (384, 97), (200, 128)]


@pytest.fixture(scope="module")
def env():
    import torch

    from pyseistr_amd import _lib, batch
    assert torch.cuda.is_available()
    assert not _lib.missing_symbols()
    rng = np.random.default_rng(11)
    from pyseistr_amd import synth
    pats = []
    for n1, n2 in SHAPES:
        d = synth._events((n1, n2), 2, [30.0], 2.0, rng)[:, :, 0]
        d = d / max(np.abs(d).max(), 1e-12) + 0.5 * rng.standard_normal((n1, n2))
        bad = rng.random(n2) < 0.02
        d[:, bad] += rng.uniform(-1, 1, size=(n1, int(bad.sum()))) * 10
        pats.append(d.astype(np.float32))
    return batch, pats


def _bits(a):
    a = np.asarray(a, np.float32).copy()
    a[a == 0] = 0.0
    return a.view(np.uint32)


def test_somf2d_batched_bit_exact(env):
    import pyseistr_amd as ps
    batch, pats = env
    b = batch.Batch(pats)
    rng = np.random.default_rng(3)
    dips = [(rng.standard_normal(p.shape) * 0.8).astype(np.float32) for p in pats]
    db = batch.Batch(dips)
    got = b.split(batch.somf2d_batched(b, db.data, ns=6, order=2, eps=0.01))
    for p, d, g in zip(pats, dips, got):
        ref = ps.somf2dc(p, d, 6, 2, 0.01, option=1, verb=0)
        assert np.array_equal(_bits(g), _bits(ref))
    # order 1 and a smaller radius as well
    got = b.split(batch.somf2d_batched(b, db.data, ns=3, order=1, eps=0.01))
    for p, d, g in zip(pats[:3], dips[:3], got[:3]):
        ref = ps.somf2dc(p, d, 3, 1, 0.01, option=1, verb=0)
        assert np.array_equal(_bits(g), _bits(ref))
    # the space-varying median (option 2) equals the per-patch filter
    got = b.split(batch.somf2d_batched(b, db.data, ns=4, order=2, eps=0.01, option=2))
    for p, d, g in zip(pats[:4], dips[:4], got[:4]):
        ref = ps.somf2dc(p, d, 4, 2, 0.01, option=2, verb=0)
        assert np.array_equal(_bits(g), _bits(ref))


@pytest.mark.parametrize("tile", range(6))
def test_dip2d_batched_matches_per_patch(env, tile):
    from pyseistr_amd import dip
    batch, pats = env
    b = batch.Batch(pats)
    out, rec = batch.dip2d_batched(b, niter=3, liter=20, order=2, rect=(10, 10, 1), tile=tile, report=True)
    got = b.split(out)
    same = 0
    for k, (p, g) in enumerate(zip(pats, got)):
        ref, r = dip.dip2dc(p, niter=3, liter=20, order=2, rect=[10, 10, 1], report=True)
        err = np.abs(g - ref).max()
        assert err <= 1e-4, (k, err)
        same += int(np.array_equal(np.asarray(r["cg_iters"][0]), rec[k, :, 0]) and
                    np.array_equal(np.asarray(r["ls_step"][0]), rec[k, :, 1]))
    assert same == len(pats)


def test_dip2d_batched_against_oracle(env):
    from oracle import oracle as O
    batch, pats = env
    sub = [pats[0], pats[3]]
    b = batch.Batch(sub)
    got = b.split(batch.dip2d_batched(b, niter=2, liter=10, order=2, rect=(5, 5, 1)))
    for p, g in zip(sub, got):
        n1, n2 = p.shape
        ref = O.dip(p, (n1, n2, 1), niter=2, liter=10, order=2, rect=(5, 5, 1))
        ref64 = O.dip(p, (n1, n2, 1), niter=2, liter=10, order=2, rect=(5, 5, 1), dtype=np.float64)
        d = np.abs(g.ravel(order="F") - ref)
        assert d.max() <= 2e-3 and np.sqrt(np.mean(d ** 2)) <= 2e-4
        assert np.abs(g.ravel(order="F") - ref64).max() <= np.abs(ref - ref64).max() + 1e-5


def test_dip2d_batched_warm_init(env):
    """init = the batch's own result continues the iteration: 2 + 2
    iterations equal 4 within the per-patch tolerance."""
    batch, pats = env
    b = batch.Batch(pats[:4])
    two = batch.dip2d_batched(b, niter=2, liter=10, rect=(5, 5, 1))
    cont = batch.dip2d_batched(b, niter=2, liter=10, rect=(5, 5, 1), init=two.clone())
    four = batch.dip2d_batched(b, niter=4, liter=10, rect=(5, 5, 1))
    assert float((cont - four).abs().max()) <= 1e-4


def test_batched_errors(env):
    import ctypes

    from pyseistr_amd import _lib
    batch, pats = env
    b = batch.Batch(pats[:2])
    ctx = _lib.context()
    lib = _lib.load()
    offs = np.array([0, 10], np.int64)  # overlapping
    n1s = np.array([128, 128], np.int32)
    n2s = np.array([64, 64], np.int32)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = lib.sh_dip2d_batched(ctx, _lib.ptr(b.data), P(offs), P(n1s), P(n2s), 2, 1, 1, 2, 5, 5, -1, None,
                              _lib.ptr(b.data), None)
    assert rc == -1
    with pytest.raises(ValueError):
        batch.dip2d_batched(b, niter=1, tile=99)
    with pytest.raises(ValueError):
        batch.somf2d_batched(b, b.data, ns=2, option=3)


def test_batched_pipeline_at_the_mix_parameters():
    """The batched workload's exact parameters on 64 patches of its own
    uneven mix (seed 4): dip2d (niter=5, liter=20, rect 10x10) then somf2dc
    (ns=6, order=2), batched against per-patch calls.  The medians are
    bit-identical given the same dips, the dips within 1e-4 with the same
    decisions, and the whole pipeline's output within what that dip
    difference moves (the tile the tuner picks for the mix)."""
    from pyseistr_amd import _lib, batch, dip, filters, synth
    pats = synth.patches2d(count=64, seed=4)
    b = batch.Batch(pats)
    tile = batch.pick_tile(b.shapes)
    dips, rec = batch.dip2d_batched(b, niter=5, liter=20, order=2, rect=(10, 10, 1), tile=tile, report=True)
    med = b.split(batch.somf2d_batched(b, dips, ns=6, order=2, eps=0.01))
    gd = b.split(dips)
    for k, p in enumerate(pats):
        ref, r = dip.dip2dc(p, niter=5, liter=20, order=2, rect=[10, 10, 1], report=True)
        assert np.abs(gd[k] - ref).max() <= 1e-4, k
        assert np.array_equal(np.asarray(r["cg_iters"][0]), rec[k, :, 0]), k
        assert np.array_equal(np.asarray(r["ls_step"][0]), rec[k, :, 1]), k
        n1, n2 = p.shape
        # the median on the batched dips equals the per-patch filter on them
        exp = filters.somf2d_device(_lib.to_device(p), _lib.to_device(gd[k]), (n1, n2, 1), 6, order=2,
                                    eps=0.01).cpu().numpy().reshape(n1, n2, order="F")
        assert np.array_equal(_bits(med[k]), _bits(exp)), k
