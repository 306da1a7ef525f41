"""RSi16 repair kernels against each other: the register kernel (two coefficients per
lane, the default for ecc_len <= 128) and the LDS kernel (any other lanes-per-codeword
choice; for ecc_len 256 two LDS shapes) give the same statuses and symbols on random codewords with errors up to and
beyond the correction radius, and on structured error patterns whose syndromes are
sparse (steps where the remainder degree drops by more than one); a sample is checked
against the oracle."""
import numpy as np
import pytest

from libffrs_amd import workloads as W

pytestmark = pytest.mark.gpu

CODES = [(8, 2, 8192), (16, 4, 8192), (64, 8, 16384), (256, 16, 16384), (256, 32, 32768), (512, 64, 8192),
         (1024, 128, 4096), (2048, 256, 1024)]


def _structured(n, ecc, k, s):
    """Rows of zeros / small values with offsets at regular positions."""
    rows = []
    for step in (1, 2, n // ecc, ecc, n // 2):
        for start in (0, 1, k - 1, k):
            if start >= n:
                continue
            cw = np.zeros(n, np.int64)
            cw[start::step] += 1 + (start % 3)
            rows.append(cw)
    cw = np.zeros(n, np.int64)
    cw[k:] = 7  # parity only
    rows.append(cw)
    cw = np.zeros(n, np.int64)
    cw[:ecc] = np.arange(1, ecc + 1)
    rows.append(cw)
    return np.array(rows, np.uint16)


def _run(ctx, codec, cm, cp):
    count = cm.shape[0]
    dcm, dcp, dst = ctx.alloc(cm.nbytes), ctx.alloc(cp.nbytes), ctx.alloc(count)
    dcm.upload(cm)
    dcp.upload(cp)
    codec.repair_device(dcm, dcp, count, dst)
    om, op, st = np.empty_like(cm), np.empty_like(cp), np.empty(count, np.uint8)
    dcm.download(om)
    dcp.download(op)
    dst.download(st)
    for b in (dcm, dcp, dst):
        b.free()
    return om, op, st


@pytest.mark.parametrize("n,ecc,count", CODES)
def test_register_and_lds_repair_agree(n, ecc, count):
    from libffrs_amd import capi
    from oracle import oracle as O

    k = n - ecc
    ctx = capi.Context(0)
    codec = capi.RSi16Codec(ctx, n, ecc)
    msgs, counts, s = W.rsi16_plain(count, n, ecc, ecc, seed=71 + ecc)  # 0..ecc errors
    dm, dp = ctx.alloc(msgs.nbytes), ctx.alloc(count * ecc * 2)
    dm.upload(np.ascontiguousarray(msgs))
    codec.encode_device(dm, count, dp)
    par = np.empty((count, ecc), np.uint16)
    dp.download(par)
    dm.free()
    dp.free()
    cws = np.concatenate([msgs, par], 1)
    W.rsi16_corrupt(cws, counts, s)
    cws = np.concatenate([cws, _structured(n, ecc, k, s)])
    cm, cp = np.ascontiguousarray(cws[:, :k]), np.ascontiguousarray(cws[:, k:])

    g_default = codec.groups()[1]
    assert g_default == min(64, ecc // 2)  # the register kernel up to ecc_len 128, LDS above
    reg = _run(ctx, codec, cm, cp)
    codec.set_groups(0, 2 if ecc == 2 else min(64, ecc) // 4)
    lds = _run(ctx, codec, cm, cp)
    codec.close()
    ctx.close()
    for a, b in zip(reg, lds):
        assert (a == b).all()
    st = reg[2]
    assert (st == 2).any() and (st != 2).any()

    rows = np.r_[0:64, cm.shape[0] - 24:cm.shape[0]]
    m, e = np.ascontiguousarray(cm[rows]), np.ascontiguousarray(cp[rows])
    ost = O.RSi16(n, ecc).repair_batch(m, e)
    assert (ost == st[rows]).all() and (m == reg[0][rows]).all() and (e == reg[1][rows]).all()


ENC_CODES = [(8, 2), (16, 4), (64, 8), (256, 16), (256, 32), (512, 64), (1024, 128), (2048, 256)]


@pytest.mark.parametrize("n,ecc", ENC_CODES)
@pytest.mark.parametrize("vector", [True, False], ids=["vector_words", "scalar_words"])
def test_encode_kernels_agree(n, ecc, vector):
    """Parity words of the default encode (one codeword per lane up to ecc 32, registers up to
    256) equal the LDS kernel's (another lanes-per-codeword choice), on uniform messages and on
    messages of only 0 and 65535 (zero partial sums, the lazy-residue edge); a sample matches
    the oracle."""
    from libffrs_amd import capi
    from oracle import oracle as O

    k = n - ecc
    count = max(64, min(8192, (1 << 22) // (2 * n)))
    s = W.Stream(97 + ecc)
    msgs = s.u16(count * k).reshape(count, k)
    edge = np.where(s.ints(0, 1, (count // 2) * k).reshape(count // 2, k) == 1, 65535, 0).astype(np.uint16)
    msgs[: count // 2] = edge
    ctx = capi.Context(0)
    codec = capi.RSi16Codec(ctx, n, ecc)
    dm, dp = ctx.alloc(msgs.nbytes), ctx.alloc(count * ecc * 2)
    dm.upload(np.ascontiguousarray(msgs))
    outs = []
    for g in (None, 2 if ecc <= 4 else min(64, ecc) // 4):
        if g is not None:
            codec.set_groups(g, 0)
        codec.encode_device(dm, count, dp, vector)
        par = np.empty((count, ecc), np.uint16)
        dp.download(par)
        outs.append(par)
    dm.free()
    dp.free()
    codec.close()
    ctx.close()
    assert (outs[0] == outs[1]).all()
    rows = np.r_[0:16, count - 16:count]
    ref = O.RSi16(n, ecc).encode(np.ascontiguousarray(msgs[rows]), vector_lanes=vector)
    assert (ref == outs[0][rows]).all()


MFMA_SHAPES = [(64, 16), (256, 32), (208, 16), (512, 64), (1024, 128), (4096, 256), (128, 8), (256, 128)]


@pytest.mark.parametrize("n,ecc", MFMA_SHAPES)
def test_mfma_encode_matches_butterflies(n, ecc):
    """The int8 MFMA encode (forced wherever it applies, and the automatic choice) against the
    butterfly kernels on random, constant and extreme messages, for the SIMD and the scalar
    words. (128, 8) has ecc 8 and (256, 128) a mix constant of 65536: there the contraction does
    not apply to the SIMD words (or at all for ecc 8) and the butterflies run, which must still
    agree."""
    from libffrs_amd import capi
    k = n - ecc
    msgs, _, s = W.rsi16_plain(3000, n, ecc, 0, seed=40 + n + ecc)
    edge = np.stack([np.zeros(k), np.full(k, 65535), np.full(k, 1), np.arange(k) % 2 * 65535,
                     np.arange(k) * 977 % 65536]).astype(np.uint16)
    msgs = np.concatenate([msgs, edge])
    count = msgs.shape[0]
    print(f"mfma encode ({n},{ecc}) on {count} messages", flush=True)
    ctx = capi.Context(0)
    codec = capi.RSi16Codec(ctx, n, ecc)
    dm, dp = ctx.alloc(msgs.nbytes), ctx.alloc(count * ecc * 2)
    try:
        dm.upload(msgs)
        for vec in (True, False):
            outs = []
            for kernel in (2, 1, 0):
                codec.set_encode_kernel(kernel)
                codec.encode_device(dm, count, dp, vec)
                ctx.sync()
                o = np.empty((count, ecc), np.uint16)
                dp.download(o)
                outs.append(o)
            assert (outs[0] == outs[1]).all(), (n, ecc, vec, int((outs[0] != outs[1]).any(1).sum()))
            assert (outs[2] == outs[1]).all()
        codec.set_encode_kernel(0)
    finally:
        dm.free()
        dp.free()
        codec.close()
        ctx.close()


@pytest.mark.parametrize("n,ecc", [(64, 16), (256, 16), (256, 32), (1024, 32), (512, 64), (4096, 64), (1024, 128)])
def test_mfma_syndromes_repair_matches_in_kernel(n, ecc):
    """The repair with its syndromes from the int8 MFMA contraction (ecc 16, 32, 64) against the
    repair computing them in the kernel: codewords with 0 .. ecc errors (beyond the radius too),
    clean ones and ones holding 0 / 65535 symbols; messages, parity and statuses identical.
    (1024, 128) has no contraction tables: both modes run the in-kernel syndromes."""
    from libffrs_amd import capi
    k = n - ecc
    count = 4000
    msgs, counts, s = W.rsi16_plain(count, n, ecc, ecc, seed=60 + n + ecc)
    msgs[:8] = 0
    msgs[8:16] = 65535
    print(f"mfma syndromes ({n},{ecc}) on {count} codewords", flush=True)
    ctx = capi.Context(0)
    codec = capi.RSi16Codec(ctx, n, ecc)
    dm, dp, ds = ctx.alloc(count * k * 2), ctx.alloc(count * ecc * 2), ctx.alloc(count)
    try:
        dm.upload(msgs)
        codec.encode_device(dm, count, dp, True)
        ctx.sync()
        par = np.empty((count, ecc), np.uint16)
        dp.download(par)
        cws = np.concatenate([msgs, par], 1)
        counts[:16] = 0
        W.rsi16_corrupt(cws, counts, s)
        outs = []
        for mode in (1, 2, 0):
            codec.set_repair_syndromes(mode)
            dm.upload(np.ascontiguousarray(cws[:, :k]))
            dp.upload(np.ascontiguousarray(cws[:, k:]))
            codec.repair_device(dm, dp, count, ds)
            ctx.sync()
            m, e, st = np.empty((count, k), np.uint16), np.empty((count, ecc), np.uint16), np.empty(count, np.uint8)
            dm.download(m)
            dp.download(e)
            ds.download(st)
            outs.append((m, e, st))
        for o in outs[1:]:
            assert (o[2] == outs[0][2]).all() and (o[0] == outs[0][0]).all() and (o[1] == outs[0][1]).all()
        assert (outs[0][2] == 2).sum() > count // 4  # most codewords were repaired
        codec.set_repair_syndromes(0)
    finally:
        dm.free()
        dp.free()
        ds.free()
        codec.close()
        ctx.close()
