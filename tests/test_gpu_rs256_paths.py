"""RS256 kernels against each other and the oracle: the repair pipeline (syndromes, then BM +
Chien + Forney; ecc_len <= 32) and the one-codeword-per-lane kernel give the same repaired
bytes and statuses on random codewords with errors up to four past the correction radius,
a sample checked against the oracle; the replicated-table encode (ecc 17-32) matches the
oracle's parity."""
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


@pytest.mark.parametrize("ecc, k, count", [(32, 223, 1025), (32, 1, 70), (32, 61, 3000), (32, 64, 2049),
                                           (31, 224, 513), (29, 7, 100), (24, 231, 4097), (17, 238, 1),
                                           (20, 90, 2500), (32, 190, 1)])
def test_replicated_table_encode_against_oracle(ecc, k, count):
    """ecc 17-32 encode with the replicated two-slice table: every message length residue mod
    4, messages shorter than one 64-byte chunk, counts that leave partial workgroups; parity
    bit-exact against the oracle."""
    from libffrs_amd import capi
    from oracle import oracle as O

    rng = np.random.default_rng(1000 * ecc + k)
    msgs = rng.integers(0, 256, size=(count, k), dtype=np.uint8)
    ctx = capi.Context(0)
    codec = capi.RS256Codec(ctx, ecc)
    dm, dp = ctx.alloc(msgs.nbytes), ctx.alloc(count * ecc)
    dm.upload(msgs)
    codec.encode_device(k, dm, count, dp)
    par = np.empty((count, ecc), np.uint8)
    dp.download(par)
    ctx.sync()
    want = O.RS256(ecc).encode_batch(msgs)
    assert (par == want).all(), (ecc, k, count, int((par != want).any(axis=1).sum()))
