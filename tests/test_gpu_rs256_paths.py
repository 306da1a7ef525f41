"""RS256 repair kernels against each other: the three-kernel pipeline (syndromes, BM,
Chien + Forney; ecc_len <= 32) and the one-codeword-per-lane kernel give the same
repaired bytes and statuses on random codewords with errors up to four past the
correction radius; a sample is checked against the oracle."""
import numpy as np
import pytest

from libffrs_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ecc", [2, 4, 8, 16, 24, 32])
def test_pipeline_and_lane_repair_agree(ecc):
    from libffrs_amd import capi
    from oracle import oracle as O

    count, k = 32768, 255 - ecc
    msgs, counts, s = W.rs256_plain(count, ecc, ecc // 2 + 4, seed=140 + ecc)
    msgs = np.ascontiguousarray(msgs[:, :k])
    ctx = capi.Context(0)
    codec = capi.RS256Codec(ctx, ecc)
    dm, dp, dst = ctx.alloc(msgs.nbytes), ctx.alloc(count * ecc), ctx.alloc(count)
    dm.upload(msgs)
    codec.encode_device(k, dm, count, dp)
    par = np.empty((count, ecc), np.uint8)
    dp.download(par)
    cws = np.concatenate([msgs, par], 1)
    W.rs256_corrupt(cws, counts, s)
    cm, cp = np.ascontiguousarray(cws[:, :k]), np.ascontiguousarray(cws[:, k:])
    res = []
    for mode in (1, 2):
        codec.set_repair_mode(mode)
        dm.upload(cm)
        dp.upload(cp)
        codec.repair_device(k, dm, dp, count, dst)
        om, op, st = np.empty_like(cm), np.empty_like(cp), np.empty(count, np.uint8)
        dm.download(om)
        dp.download(op)
        dst.download(st)
        res.append((om, op, st))
    for b in (dm, dp, dst):
        b.free()
    codec.close()
    ctx.close()
    for a, b in zip(*res):
        assert (a == b).all()
    st = res[0][2]
    assert (st == 0).any() and (st == 1).any()
    rows = np.r_[0:48, count - 16:count]
    m, e = np.ascontiguousarray(cm[rows]), np.ascontiguousarray(cp[rows])
    ost = O.RS256(ecc).decode_batch(m, e)
    assert (ost == st[rows]).all() and (m == res[0][0][rows]).all() and (e == res[0][1][rows]).all()
