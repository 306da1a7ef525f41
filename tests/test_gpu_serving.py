"""The continuous-batching server: every request's results are the same when it runs
alone and when it is batched with others, and they match the oracle (RSi16(1024, 128)
encode of message blocks; repair of corrupted blocks); at full request size (64 MiB) with
Poisson arrivals, every corrupted file comes back byte for byte."""
import numpy as np
import pytest

from libffrs_amd import workloads as W

pytestmark = pytest.mark.gpu

SIZES = [4096, 9001, 20000, 65536, 150000, 300001, 4097, 70000, 12345, 200000]


def _results(r):
    return r.msgs.copy(), r.parity.copy(), r.status.copy()


def test_alone_equals_mixed_and_oracle():
    from libffrs_amd.serving import BlockServer
    from oracle import oracle as O

    s = W.Stream(31)
    datas = [s.bytes(n) for n in SIZES]
    # small arenas and pieces so that batches cut requests and mix several of them
    with BlockServer(block_len=1024, ecc_len=128, slots=2, slot_blocks=48, piece_blocks=16) as srv:
        enc = [srv.submit("encode", d) for d in datas]
        for r in enc:
            r.done.wait()
        payload = []
        for i, r in enumerate(enc):
            if i % 2:
                m, p = r.msgs.copy(), r.parity.copy()
                W.serving_corrupt(m, p, s, max_frac=0.06)
                payload.append(("repair", m.reshape(-1).view(np.uint8), p))
            else:
                payload.append(("encode", datas[i], None))
        srv.batches.clear()
        mixed = [srv.submit(kind, d, p) for kind, d, p in payload]  # all in flight together
        for r in mixed:
            r.done.wait()
        assert any(b[1] > 1 for b in srv.batches), "no batch held more than one request"
        alone = []
        for kind, d, p in payload:
            r = srv.submit(kind, d, p)
            r.done.wait()
            alone.append(r)
    for a, b in zip(mixed, alone):
        for x, y in zip(_results(a), _results(b)):
            assert (x == y).all()

    orc = O.RSi16(1024, 128)
    for i, (r, (kind, d, p)) in enumerate(zip(alone, payload)):
        rows = slice(0, min(3, r.blocks))
        if kind == "encode":
            assert (r.parity[rows] == orc.encode(r.msgs[rows])).all()
        else:
            m = np.ascontiguousarray(np.frombuffer(d, np.uint16).reshape(-1, 896)[rows])
            e = np.ascontiguousarray(p[rows])
            st = orc.repair_batch(m, e)
            assert (st == r.status[rows]).all() and (m == r.msgs[rows]).all() and (e == r.parity[rows]).all()
            # most blocks carry fewer than ecc/2 errors and come back as they were encoded
            assert (r.msgs == enc[i].msgs).all(axis=1).mean() > 0.5


def test_full_size_requests_with_poisson_arrivals():
    """The largest requests of the serving workload (64 MiB files in RSi16(1024, 128) blocks),
    submitted with exponential gaps so that several are in flight: encode, then every block
    corrupted with 0..60 additive errors (within the correction radius of 64 even where a
    parity symbol of 65536 was stored as 0) and repaired. All but a few blocks come back as
    encoded; the others (the reference's repair leaves some within the radius unrepaired and
    reports a rare one repaired with other contents, tests/golden/rsi16_edge_radius_1024.npz)
    and a sample match the oracle bit for bit, files without such blocks come back byte for
    byte, and the first blocks' parity matches the oracle."""
    import time

    from libffrs_amd.serving import BlockServer
    from oracle import oracle as O

    sizes = [64 << 20, (64 << 20) - 1, 3 << 20, 1 << 20, 777777]
    s = W.Stream(41)
    gaps = np.random.default_rng(41).exponential(0.002, 2 * len(sizes))
    datas = [s.bytes(n).tobytes() for n in sizes]
    print("serving 64 MiB requests: encode", flush=True)
    with BlockServer(block_len=1024, ecc_len=128) as srv:
        enc = []
        for i, d in enumerate(datas):
            enc.append(srv.submit("encode", d))
            time.sleep(gaps[i])
        for r in enc:
            r.done.wait()
        print("serving 64 MiB requests: repair", flush=True)
        reps, corrupted = [], []
        for i, r in enumerate(enc):
            cws = np.concatenate([r.msgs, r.parity], 1)
            W.rsi16_corrupt(cws, s.ints(0, 60, r.blocks), s)
            cm, cp = np.ascontiguousarray(cws[:, :896]), np.ascontiguousarray(cws[:, 896:])
            assert not (cm == r.msgs).all()
            corrupted.append((cm, cp))
            reps.append(srv.submit("repair", cm.reshape(-1).view(np.uint8).tobytes(), cp))
            time.sleep(gaps[len(sizes) + i])
        for r in reps:
            r.done.wait()
    orc = O.RSi16(1024, 128)
    assert (enc[0].parity[:2] == orc.encode(enc[0].msgs[:2])).all()
    for d, e, r, (cm, cp) in zip(datas, enc, reps, corrupted):
        ok = np.isin(r.status, [0, 2])
        differs = ~((r.msgs == e.msgs).all(axis=1) & (r.parity == e.parity).all(axis=1))
        # the reference's repair leaves about 1 in 2,000 of these blocks unrepaired
        # (ErrorLocationFail) and reports about 1 in 50,000 repaired with other contents than
        # were encoded: such blocks are rare, and each is the oracle's result bit for bit
        assert (~ok).mean() < 0.002 and (ok & differs).mean() < 1e-4
        rows = np.r_[np.nonzero(~ok | differs)[0][:12], 0, r.blocks - 1]
        m, p = np.ascontiguousarray(cm[rows]), np.ascontiguousarray(cp[rows])
        st = orc.repair_batch(m, p)
        assert (st == r.status[rows]).all() and (m == r.msgs[rows]).all() and (p == r.parity[rows]).all()
        if not differs.any():
            assert r.msgs.reshape(-1).view(np.uint8)[:len(d)].tobytes() == d
