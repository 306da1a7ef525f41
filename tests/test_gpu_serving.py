"""The continuous-batching server: every request's results are the same when it runs
alone and when it is batched with others, and they match the oracle (RSi16(1024, 128)
encode of message blocks; repair of corrupted blocks)."""
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
