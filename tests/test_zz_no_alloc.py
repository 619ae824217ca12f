"""The allocation-free Merkle commitment (azk_merkle_commit_noalloc) against the plan-based
resident commitment it is a variant of (azk_merkle_commit_resident), through the C ABI: two
calls in a row on the shared scratch, the second smaller, both byte-equal to the plan and to
the oracle.  Sizes beyond the scratch are refused."""
import numpy as np
import pytest

import pyoracle as O
from apple_amd import azk, inputs


@pytest.fixture(scope="module")
def ctx():
    c = azk.Context(0)
    yield c
    c.close()


def plan_root(ctx, buf, count, stride, length, offset=0):
    plan = azk.SHA3RawLeavesMerkleCommitPlan(ctx, count, stride, length)
    try:
        return plan.commit_resident(buf, offset)
    finally:
        plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("first,second", [((1 << 16, 64, 64), (1 << 10, 136, 136)),
                                          ((1 << 12, 72, 67), (1, 32, 32))])
def test_noalloc_matches_resident_commit(ctx, first, second):
    root = ctx.buffer(64)
    for seed, (count, stride, length) in enumerate((first, second)):
        print(f"noalloc merkle {count} x {length}/{stride}", flush=True)
        raw = inputs.random_bytes(900 + seed, count * stride + 8)
        buf = ctx.upload(raw)
        azk.merkle_commit_noalloc(ctx, buf, count, stride, length, root, leaves_offset=0, root_offset=32)
        ctx.synchronize()
        got = root.download(32, 32).tobytes()
        assert got == plan_root(ctx, buf, count, stride, length)
        assert got == O.merkle_root(raw[: count * stride], count, stride, length)
        buf.free()
    root.free()


@pytest.mark.gpu
def test_noalloc_offset_and_limits(ctx):
    count, stride = 1 << 8, 40
    raw = inputs.random_bytes(77, 12 + count * stride)
    buf = ctx.upload(raw)
    root = ctx.buffer(32)
    azk.merkle_commit_noalloc(ctx, buf, count, stride, 33, root, leaves_offset=12)
    ctx.synchronize()
    assert root.download(32).tobytes() == plan_root(ctx, buf, count, stride, 33, 12)
    for bad in [dict(leaf_count=1 << 17), dict(leaf_count=3), dict(leaf_length=137), dict(leaf_stride=8)]:
        args = dict(leaf_count=count, leaf_stride=stride, leaf_length=33)
        args.update(bad)
        with pytest.raises(azk.AZKError):
            azk.merkle_commit_noalloc(ctx, buf, args["leaf_count"], args["leaf_stride"], args["leaf_length"], root)
    with pytest.raises(azk.AZKError) as e:  # range past the buffer
        azk.merkle_commit_noalloc(ctx, buf, count, stride, 33, root, leaves_offset=13)
    assert e.value.code == 6
    buf.free()
    root.free()
