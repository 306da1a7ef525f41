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
