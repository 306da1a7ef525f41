"""Allocation-free RS256 repair (ffrs_rs256_repair_device_noalloc) against the allocating
entry point (ffrs_rs256_repair_device), through the C ABI: two calls in a row on the same
codec, the second with fewer codewords, bit-exact statuses and repaired bytes; a sample is
also checked against the oracle. Covers the syndrome / Chien pipeline with BM inside the
Chien kernel (ecc 8, 32), with BM as a kernel of its own (ecc 24, whose records also live in
the workspace), and the lane-mode repair that needs no workspace (ecc 64)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(ecc, n_cw, seed):
    from libffrs_amd import workloads as W
    t = ecc // 2
    msgs, counts, s = W.rs256_plain(n_cw, ecc, t, seed=seed, uncorrectable_frac=0.02,
                                    uncorrectable_range=(t + 1, t + 3))
    return msgs, counts, s


@pytest.mark.parametrize("ecc", [8, 24, 32, 64])
def test_noalloc_repair_matches_repair_device(ecc):
    from libffrs_amd import capi
    from libffrs_amd import workloads as W
    from oracle import oracle as O

    k = 255 - ecc
    ctx = capi.Context(0)
    codec = capi.RS256Codec(ctx, ecc)
    bufs = []
    try:
        for call, n_cw in enumerate((20000, 3000)):
            print(f"ecc {ecc} call {call}: {n_cw} codewords", flush=True)
            msgs, counts, s = _case(ecc, n_cw, seed=700 + 10 * ecc + call)
            d_m, d_p = ctx.alloc(msgs.nbytes), ctx.alloc(n_cw * ecc)
            bufs += [d_m, d_p]
            d_m.upload(msgs)
            codec.encode_device(k, d_m, n_cw, d_p)
            par = np.empty((n_cw, ecc), np.uint8)
            d_p.download(par)
            cws = np.concatenate([msgs, par], 1)
            W.rs256_corrupt(cws, counts, s)
            cm, cp = np.ascontiguousarray(cws[:, :k]), np.ascontiguousarray(cws[:, k:])
            outs = []
            for noalloc in (False, True):
                dm, dp, ds = ctx.alloc(cm.nbytes), ctx.alloc(cp.nbytes), ctx.alloc(n_cw)
                bufs += [dm, dp, ds]
                dm.upload(cm)
                dp.upload(cp)
                if noalloc:
                    codec.repair_device_noalloc(k, dm, dp, n_cw, ds)
                else:
                    codec.repair_device(k, dm, dp, n_cw, ds)
                ctx.sync()
                om, op, st = np.empty_like(cm), np.empty_like(cp), np.empty(n_cw, np.uint8)
                dm.download(om)
                dp.download(op)
                ds.download(st)
                outs.append((om, op, st))
            (am, ap, ast), (bm, bp, bst) = outs
            assert (ast == bst).all(), f"status differs in {int((ast != bst).sum())} codewords"
            assert (am == bm).all() and (ap == bp).all(), "repaired bytes differ"
            assert (ast == 1).sum() > n_cw // 2 and (ast == 0).sum() > 0
            # a sample against the oracle
            rows = np.arange(0, n_cw, 37)
            om_, op_ = cm[rows].copy(), cp[rows].copy()
            ost = O.RS256(ecc).decode_batch(om_, op_)
            assert (ost == bst[rows]).all() and (om_ == bm[rows]).all() and (op_ == bp[rows]).all()
    finally:
        for b in bufs:
            b.free()
        codec.close()
        ctx.close()


def test_noalloc_rejects_calls_beyond_the_workspace():
    """ecc 24 takes 32 + 48 bytes of workspace per codeword: 500,000 codewords exceed the
    static workspace, and the call fails with the limit named before anything runs."""
    from libffrs_amd import capi
    ecc, k, n_cw = 24, 231, 500000
    ctx = capi.Context(0)
    codec = capi.RS256Codec(ctx, ecc)
    bufs = [ctx.alloc(n_cw * k), ctx.alloc(n_cw * ecc), ctx.alloc(n_cw)]
    try:
        with pytest.raises(RuntimeError, match="at most [0-9]+ codewords per call"):
            codec.repair_device_noalloc(k, bufs[0], bufs[1], n_cw, bufs[2])
    finally:
        for b in bufs:
            b.free()
        codec.close()
        ctx.close()
