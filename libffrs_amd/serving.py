"""Continuous-batching server for par-style RSi16 block coding on one GPU.

A request is a file buffer (bytes) to encode, or a file buffer plus its parity to verify
and repair. The file is split into blocks of `message_len` u16 symbols (the last one zero
padded), each coded as one RSi16(block_len, ecc_len) codeword, as the reference's `par`
front end does with `RSi16.encode` / `RSi16.repair` per block (ffrs/par/backup.py:74,
ffrs/par/repair.py:97).

Continuous batching: requests are cut into pieces of at most `piece_blocks` blocks and
queued in arrival order. Each of `slots` workers owns a HIP stream (its own context), a
device arena and pinned host staging of `slot_blocks` blocks; the arenas together take at
most `hbm_fraction` of the HBM the device reports free when the server starts (the HBM
budget: `slot_blocks` is lowered to fit it). A free worker takes queued pieces, from any
requests, until its arena is full (admission by the HBM budget), then
copies them in, runs one encode launch and one repair launch over all of them, and copies
the results out. Requests that arrive while other batches are on the GPU are admitted by
the next free worker, so copies and kernels of different batches overlap on the device.
A request completes when its last piece does; its results do not depend on how it was
batched (every codeword is independent).
"""
import collections
import threading
import time

import numpy as np

from . import capi


class Request:
    """One encode or verify+repair request; `done` is set when all its blocks are coded."""

    _ids = 0

    def __init__(self, kind, data, parity, k, ecc):
        assert kind in ("encode", "repair")
        Request._ids += 1
        self.id = Request._ids
        self.kind = kind
        self.nbytes = len(data)
        sym = np.frombuffer(data, np.uint8)
        if sym.size % 2:
            sym = np.concatenate([sym, np.zeros(1, np.uint8)])
        sym = sym.view(np.uint16)
        self.blocks = max(1, -(-sym.size // k))
        self.msgs = np.zeros((self.blocks, k), np.uint16)
        self.msgs.reshape(-1)[:sym.size] = sym
        if kind == "repair":
            self.parity = np.array(parity, np.uint16).reshape(self.blocks, ecc)
        else:
            self.parity = np.zeros((self.blocks, ecc), np.uint16)
        self.status = np.zeros(self.blocks, np.uint8)
        self.pending = 0
        self.t_submit = self.t_done = None
        self.done = threading.Event()

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_submit

    def file_bytes(self):
        """The (repaired) file contents."""
        return self.msgs.reshape(-1).view(np.uint8)[:self.nbytes].tobytes()


class _Slot:
    def __init__(self, device, n, ecc, blocks):
        self.ctx = capi.Context(device)
        self.codec = capi.RSi16Codec(self.ctx, n, ecc)
        k = n - ecc
        self.blocks = blocks
        self.d_msg = self.ctx.alloc(blocks * k * 2)
        self.d_par = self.ctx.alloc(blocks * ecc * 2)
        self.d_st = self.ctx.alloc(blocks)
        self.h_msg = self.ctx.host_alloc(blocks * k * 2).view(np.uint16).reshape(blocks, k)
        self.h_par = self.ctx.host_alloc(blocks * ecc * 2).view(np.uint16).reshape(blocks, ecc)
        self.h_st = self.ctx.host_alloc(blocks)

    def close(self):
        for a in (self.h_msg, self.h_par, self.h_st):
            self.ctx.host_free(a)
        for b in (self.d_msg, self.d_par, self.d_st):
            b.free()
        self.codec.close()
        self.ctx.close()


class BlockServer:
    """RSi16 block coding server; see the module docstring. Defaults: RSi16(1024, 128)."""

    def __init__(self, block_len=1024, ecc_len=128, slots=3, slot_blocks=8192, piece_blocks=2048, device=0,
                 hbm_fraction=0.25):
        self.n, self.ecc, self.k = block_len, ecc_len, block_len - ecc_len
        per_block = 2 * block_len + 1  # message + parity symbols and a status byte
        probe = capi.Context(device)
        self.hbm_free = int(probe.info()[2])
        probe.close()
        self.hbm_budget_cap = int(self.hbm_free * hbm_fraction)
        slot_blocks = max(1, min(slot_blocks, self.hbm_budget_cap // (slots * per_block)))
        self.piece_blocks = max(1, min(piece_blocks, slot_blocks))
        self._slots = [_Slot(device, block_len, ecc_len, slot_blocks) for _ in range(slots)]
        self.hbm_budget = slots * slot_blocks * per_block
        self._queue = collections.deque()
        self._cv = threading.Condition()
        self._stop = False
        self.batches = []  # (slot, requests in the batch, blocks) of every launch, for inspection
        self._threads = [threading.Thread(target=self._worker, args=(s,), daemon=True) for s in self._slots]
        for t in self._threads:
            t.start()

    def submit(self, kind, data, parity=None):
        req = Request(kind, data, parity, self.k, self.ecc)
        pieces = [(req, b, min(self.piece_blocks, req.blocks - b)) for b in range(0, req.blocks, self.piece_blocks)]
        req.pending = len(pieces)
        with self._cv:
            req.t_submit = time.perf_counter()
            self._queue.extend(pieces)
            self._cv.notify_all()
        return req

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        for s in self._slots:
            s.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _take(self, cap):
        with self._cv:
            while not self._queue and not self._stop:
                self._cv.wait()
            if not self._queue:
                return None
            taken, used = [], 0
            while self._queue and used + self._queue[0][2] <= cap:
                piece = self._queue.popleft()
                taken.append(piece)
                used += piece[2]
            return taken

    def _worker(self, slot):
        k, ecc = self.k, self.ecc
        while True:
            taken = self._take(slot.blocks)
            if taken is None:
                return
            enc = [p for p in taken if p[0].kind == "encode"]
            rep = [p for p in taken if p[0].kind == "repair"]
            ne = sum(p[2] for p in enc)
            nr = sum(p[2] for p in rep)
            row = 0
            for req, b, c in enc + rep:
                slot.h_msg[row:row + c] = req.msgs[b:b + c]
                if req.kind == "repair":
                    slot.h_par[row:row + c] = req.parity[b:b + c]
                row += c
            slot.d_msg.upload_async(slot.h_msg[:ne + nr])
            if nr:
                slot.d_par.upload_async(slot.h_par[ne:ne + nr], offset=ne * ecc * 2)
            if ne:
                slot.codec.encode_device(slot.d_msg.view(0, ne * k * 2), ne, slot.d_par.view(0, ne * ecc * 2))
            if nr:
                slot.codec.repair_device(slot.d_msg.view(ne * k * 2, nr * k * 2), slot.d_par.view(ne * ecc * 2,
                                         nr * ecc * 2), nr, slot.d_st.view(ne, nr))
                slot.d_msg.download_async(slot.h_msg[ne:ne + nr], offset=ne * k * 2)
                slot.d_st.download_async(slot.h_st[ne:ne + nr], offset=ne)
            slot.d_par.download_async(slot.h_par[:ne + nr])
            slot.ctx.sync()
            row = 0
            finished = []
            for req, b, c in enc + rep:
                req.parity[b:b + c] = slot.h_par[row:row + c]
                if req.kind == "repair":
                    req.msgs[b:b + c] = slot.h_msg[row:row + c]
                    req.status[b:b + c] = slot.h_st[row:row + c]
                row += c
                finished.append(req)
            now = time.perf_counter()
            with self._cv:
                self.batches.append((self._slots.index(slot), len({id(p[0]) for p in taken}), ne + nr))
                for req in finished:
                    req.pending -= 1
                    if req.pending == 0:
                        req.t_done = now
                        req.done.set()
