"""Batched uneven k-points (the Si8 4x4x4 Monkhorst-Pack mesh: 64 k-points,
1622-1653 plane waves each, one 32^3 grid, k-dependent projector phases, 32
bands per k-point): every H psi and S psi of the 64 k-points in one
persistent kernel launch replayed from a HIP graph equals the per-k-point
device applications (themselves pinned to the oracle) within 1e-12, on
several replays with other work between them; the autotuned launch shape is
reused from the cache by a second batch of the same shape."""
import numpy as np
import pytest

from sdft_amd import native
from sdft_amd import synthetic as syn

pytestmark = pytest.mark.gpu

COLUMNS = 32


@pytest.fixture(scope="module")
def mp_batch(hip_ctx):
    ks = syn.mp_kpoints(4)
    probs = [syn.config_problem(2, kfrac=tuple(k)) for k in ks]
    ops = [native.Operator(hip_ctx, p) for p in probs]
    blocks = [syn.initial_psi(p, COLUMNS, 2, extra=i) for i, p in enumerate(probs)]
    yield probs, ops, blocks
    for op in ops:
        op.close()


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_batch_graph_replays_match_per_k_applies(hip_ctx, mp_batch):
    probs, ops, blocks = mp_batch
    assert len({p.n for p in probs}) > 1  # uneven spheres
    batch = native.Batch(hip_ctx, ops, COLUMNS, precision=0)
    try:
        plan = batch.plan()
        assert plan["threads"] in (256, 512, 1024) and plan["planes"] in (1, 2) and plan["groups"] > 0
        ref = [op.apply(b, "pair") for op, b in zip(ops, blocks)]
        for replay in range(3):
            hs, ss = batch.apply(blocks)
            for k in range(len(ops)):
                assert _rel(hs[k], ref[k][0]) < 1e-12, (replay, k)
                assert _rel(ss[k], ref[k][1]) < 1e-12, (replay, k)
            # eager work on the same stream between replays
            ops[replay].apply(blocks[replay], "pair")
        again = native.Batch(hip_ctx, ops, COLUMNS, precision=0)
        assert again.plan()["from_cache"] and again.plan()["threads"] == plan["threads"]
        again.close()
    finally:
        batch.close()


def test_batch_fp32_search_precision(hip_ctx, mp_batch):
    _, ops, blocks = mp_batch
    batch = native.Batch(hip_ctx, ops, COLUMNS, precision=1)
    try:
        hs, ss = batch.apply(blocks)
        for k in (0, 21, 63):
            rh, rs = ops[k].apply(blocks[k], "pair", precision=0)
            assert _rel(hs[k], rh) < 2e-5 and _rel(ss[k], rs) < 2e-5
    finally:
        batch.close()


def test_batch_rejects_mixed_grids(hip_ctx, si8):
    a = native.Operator(hip_ctx, si8)
    b = native.Operator(hip_ctx, si8, shape=(30, 30, 30))
    try:
        with pytest.raises(ValueError):
            native.Batch(hip_ctx, [a, b], 4)
    finally:
        a.close()
        b.close()
