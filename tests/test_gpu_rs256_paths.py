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


@pytest.mark.parametrize("ecc, k", [(33, 222), (48, 100), (64, 191), (64, 3), (100, 155), (124, 131), (125, 130),
                                    (128, 127), (160, 95), (192, 63), (200, 40), (240, 15), (254, 1)])
def test_wave_and_lane_repair_agree(ecc, k):
    """ecc_len > 32: the wave-per-codeword repair (default) against the one-codeword-per-lane
    kernel, bit-exact, with errors up to four past the correction radius (and at most n), every
    coefficient / syndrome split of the wave kernel; a sample checked against the oracle."""
    from libffrs_amd import capi
    from oracle import oracle as O

    count = 3000
    n = k + ecc
    msgs, counts, s = W.rs256_plain(count, ecc, min(ecc // 2 + 4, n), seed=300 + ecc)
    msgs = np.ascontiguousarray(msgs[:, :k])
    # corruption positions drawn over the n bytes of these (shortened) codewords
    counts = np.minimum(counts, n)
    print(f"wave repair ecc {ecc} k {k}", flush=True)
    ctx = capi.Context(0)
    codec = capi.RS256Codec(ctx, ecc)
    dm, dp, dst = ctx.alloc(msgs.nbytes), ctx.alloc(count * ecc), ctx.alloc(count)
    try:
        dm.upload(msgs)
        codec.encode_device(k, dm, count, dp)
        par = np.empty((count, ecc), np.uint8)
        dp.download(par)
        cws = np.concatenate([msgs, par], 1)
        W.rs256_corrupt(cws, counts, s)
        cm, cp = np.ascontiguousarray(cws[:, :k]), np.ascontiguousarray(cws[:, k:])
        res = []
        for mode in (0, 2):
            codec.set_repair_mode(mode)
            dm.upload(cm)
            dp.upload(cp)
            codec.repair_device(k, dm, dp, count, dst)
            om, op, st = np.empty_like(cm), np.empty_like(cp), np.empty(count, np.uint8)
            dm.download(om)
            dp.download(op)
            dst.download(st)
            res.append((om, op, st))
    finally:
        for b in (dm, dp, dst):
            b.free()
        codec.close()
        ctx.close()
    for a, b in zip(*res):
        assert (a == b).all()
    st = res[0][2]
    assert (st == 1).any()
    rows = np.r_[0:24, count - 8:count]
    m, e = np.ascontiguousarray(cm[rows]), np.ascontiguousarray(cp[rows])
    ost = O.RS256(ecc).decode_batch(m, e)
    assert (ost == st[rows]).all() and (m == res[0][0][rows]).all() and (e == res[0][1][rows]).all()


@pytest.mark.parametrize("ecc, k, count", [(32, 223, 8193), (32, 1, 8200), (32, 61, 9000), (32, 64, 12289),
                                           (31, 224, 8197), (29, 7, 8300), (24, 231, 16385), (17, 238, 8192),
                                           (20, 90, 10000), (16, 239, 9999), (12, 100, 8250), (8, 247, 8195),
                                           (4, 251, 8209), (2, 253, 8193), (32, 190, 1), (8, 120, 700)])
def test_replicated_table_encode_against_oracle(ecc, k, count):
    """Encode over the replicated tables with cooperative message loads (launches of 8,192
    codewords and more) and over the slicing tables (smaller ones): every message length
    residue mod 4, messages shorter than one 64-byte chunk, counts that leave partial waves and
    workgroups; parity bit-exact against the oracle."""
    from libffrs_amd import capi
    from oracle import oracle as O

    rng = np.random.default_rng(1000 * ecc + k)
    msgs = rng.integers(0, 256, size=(count, k), dtype=np.uint8)
    ctx = capi.Context(0)
    codec = capi.RS256Codec(ctx, ecc)
    dm, dp = ctx.alloc(msgs.nbytes), ctx.alloc(count * ecc)
    try:
        dm.upload(msgs)
        codec.encode_device(k, dm, count, dp)
        par = np.empty((count, ecc), np.uint8)
        dp.download(par)
        ctx.sync()
    finally:
        dm.free()
        dp.free()
        codec.close()
        ctx.close()
    want = O.RS256(ecc).encode_batch(msgs)
    assert (par == want).all(), (ecc, k, count, int((par != want).any(axis=1).sum()))
