"""Codec backends for the golden replay (tests/golden/replay.py).

OracleBackend drives the CPU restatement (oracle/); GpuBackend drives the HIP
path through the product's public API (libffrs_amd), the same calls a user makes.
"""
import numpy as np

from oracle import oracle as O


class OracleBackend:
    def __init__(self):
        self._rs = {}
        self._ri = {}

    def rs(self, ecc):
        if ecc not in self._rs:
            self._rs[ecc] = O.RS256(ecc)
        return self._rs[ecc]

    def ri(self, n, ecc):
        if (n, ecc) not in self._ri:
            self._ri[(n, ecc)] = O.RSi16(n, ecc)
        return self._ri[(n, ecc)]

    def rs256_encode(self, block_len, ecc, buf):
        return self.rs(ecc).encode(buf, block_len - ecc)

    def rs256_encode_batch(self, ecc, msgs):
        return self.rs(ecc).encode_batch(msgs)

    def rs256_repair_batch(self, ecc, msgs, parity):
        return self.rs(ecc).decode_batch(msgs, parity)

    def rs256_synds(self, block_len, ecc, cw):
        k = block_len - ecc
        return self.rs(ecc).synds(cw[:k], cw[k:])

    def rs256_mixed_encode(self, eccs, lens, msgs):
        return [self.rs(int(e)).encode_batch(m[None])[0] for e, m in zip(eccs, msgs)]

    def rs256_mixed_repair(self, eccs, lens, cws):
        st = np.zeros(len(cws), np.uint8)
        for i, (e, l) in enumerate(zip(eccs, lens)):
            m = cws[i][:l].copy()
            p = cws[i][l:].copy()
            st[i] = self.rs(int(e)).decode(m, p)
            cws[i][:l] = m
            cws[i][l:] = p
        return st

    def rsi16_encode_batch(self, n, ecc, msgs, vector):
        return self.ri(n, ecc).encode(msgs, vector)

    def rsi16_repair_batch(self, n, ecc, msgs, parity):
        return self.ri(n, ecc).repair_batch(msgs, parity)

    def rsi16_synds(self, n, ecc, msg, parity):
        return self.ri(n, ecc).synds(msg, parity)

    def rsi16_sugiyama(self, n, ecc, synds):
        L, loc, ev = self.ri(n, ecc).sugiyama(synds)
        return None if L == 0xFFFFFFFF else (loc, ev)

    def rsi16_repair_known(self, n, ecc, msg, parity, positions):
        return self.ri(n, ecc).repair_known(msg, parity, positions)

    def rsi16_mixed_encode(self, ns, eccs, msgs):
        return [self.ri(int(n), int(e)).encode(m[None], True)[0] for n, e, m in zip(ns, eccs, msgs)]

    def rsi16_mixed_repair(self, ns, eccs, cws):
        st = np.zeros(len(cws), np.uint8)
        for i, (n, e) in enumerate(zip(ns, eccs)):
            k = int(n - e)
            m = cws[i][:k].copy()
            p = cws[i][k:].copy()
            st[i] = self.ri(int(n), int(e)).repair(m, p)
            cws[i][:k] = m
            cws[i][k:] = p
        return st


class GpuBackend:
    """The HIP path through libffrs_amd's public API."""

    def __init__(self):
        import libffrs_amd
        self.L = libffrs_amd
        self._rs = {}
        self._ri = {}

    def rs(self, ecc, block_len=255):
        key = ecc
        if key not in self._rs:
            self._rs[key] = self.L.RS256(block_len=255 if ecc < 255 else None, ecc_len=ecc)
        r = self._rs[key]
        r.block_len = block_len
        return r

    def ri(self, n, ecc, vector=True):
        key = (n, ecc, vector)
        if key not in self._ri:
            kw = {} if vector else dict(simd_x4=False, simd_x8=False, simd_x16=False)
            self._ri[key] = self.L.RSi16(n, ecc, **kw)
        return self._ri[key]

    def rs256_encode(self, block_len, ecc, buf):
        return self.rs(ecc, block_len).encode(buf)

    def rs256_encode_batch(self, ecc, msgs):
        count, k = msgs.shape
        out = self.rs(ecc, k + ecc).encode(np.ascontiguousarray(msgs).tobytes())
        return np.frombuffer(bytes(out), np.uint8).reshape(count, ecc).copy()

    def rs256_repair_batch(self, ecc, msgs, parity):
        count, k = msgs.shape
        st = self.rs(ecc, k + ecc).repair_batch(msgs, parity)
        return np.frombuffer(bytes(st), np.uint8).copy()

    def rs256_synds(self, block_len, ecc, cw):
        return np.frombuffer(bytes(self.rs(ecc, block_len)._synds(np.ascontiguousarray(cw).tobytes())), np.uint8)

    def rs256_mixed_encode(self, eccs, lens, msgs):
        out = [None] * len(msgs)
        groups = {}
        for i, (e, l) in enumerate(zip(eccs, lens)):
            groups.setdefault((int(e), int(l)), []).append(i)
        for (e, l), idx in groups.items():
            par = self.rs256_encode_batch(e, np.stack([msgs[i] for i in idx]))
            for j, i in enumerate(idx):
                out[i] = par[j]
        return out

    def rs256_mixed_repair(self, eccs, lens, cws):
        st = np.zeros(len(cws), np.uint8)
        groups = {}
        for i, (e, l) in enumerate(zip(eccs, lens)):
            groups.setdefault((int(e), int(l)), []).append(i)
        for (e, l), idx in groups.items():
            m = np.ascontiguousarray(np.stack([cws[i][:l] for i in idx]))
            p = np.ascontiguousarray(np.stack([cws[i][l:] for i in idx]))
            s = self.rs256_repair_batch(e, m, p)
            for j, i in enumerate(idx):
                st[i] = s[j]
                cws[i][:l] = m[j]
                cws[i][l:] = p[j]
        return st

    def rsi16_encode_batch(self, n, ecc, msgs, vector):
        msgs = np.ascontiguousarray(msgs, np.uint16)
        out = self.ri(n, ecc, vector).encode(msgs.tobytes())
        return np.frombuffer(bytes(out), np.uint16).reshape(-1, ecc).copy()

    def rsi16_repair_batch(self, n, ecc, msgs, parity):
        # the reference sizes buffers in bytes (util.hpp:30-55), so u16 arrays go in as byte views
        st = self.ri(n, ecc).repair_batch(msgs.view(np.uint8), parity.view(np.uint8))
        return np.frombuffer(bytes(st), np.uint8).copy()

    def rsi16_synds(self, n, ecc, msg, parity):
        return np.array(self.ri(n, ecc)._synd(np.ascontiguousarray(msg, np.uint16).tobytes(),
                                              np.ascontiguousarray(parity, np.uint16).tobytes()), np.uint32)

    def rsi16_sugiyama(self, n, ecc, synds):
        try:
            loc, ev = self.ri(n, ecc)._sugiyama([int(v) for v in synds])
        except RuntimeError:
            return None
        return np.array(loc, np.uint32), np.array(ev, np.uint32)

    def rsi16_repair_known(self, n, ecc, msg, parity, positions):
        try:
            return int(self.ri(n, ecc).repair(msg.view(np.uint8), parity.view(np.uint8), error_pos=list(positions)))
        except RuntimeError:
            return 255

    def rsi16_mixed_encode(self, ns, eccs, msgs):
        out = [None] * len(msgs)
        groups = {}
        for i, (n, e) in enumerate(zip(ns, eccs)):
            groups.setdefault((int(n), int(e)), []).append(i)
        for (n, e), idx in groups.items():
            par = self.rsi16_encode_batch(n, e, np.stack([msgs[i] for i in idx]), True)
            for j, i in enumerate(idx):
                out[i] = par[j]
        return out

    def rsi16_mixed_repair(self, ns, eccs, cws):
        st = np.zeros(len(cws), np.uint8)
        groups = {}
        for i, (n, e) in enumerate(zip(ns, eccs)):
            groups.setdefault((int(n), int(e)), []).append(i)
        for (n, e), idx in groups.items():
            k = n - e
            m = np.ascontiguousarray(np.stack([cws[i][:k] for i in idx]))
            p = np.ascontiguousarray(np.stack([cws[i][k:] for i in idx]))
            s = self.rsi16_repair_batch(n, e, m, p)
            for j, i in enumerate(idx):
                st[i] = s[j]
                cws[i][:k] = m[j]
                cws[i][k:] = p[j]
        return st
