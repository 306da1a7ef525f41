"""ctypes wrapper around liboracle.so, the CPU restatement of libffrs' RS256/RSi16 path.

TEST INFRASTRUCTURE ONLY. This module is the checker the HIP path is compared
against. Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg
may import it; the product package ``libffrs_amd`` never does.

Reference routines restated (paths relative to the reference repository root):
  RS256 encode  libffrs/pyrs256.hpp:99-127, libffrs/reed_solomon.hpp:169-389
  RS256 repair  libffrs/reed_solomon.hpp:555-744
  RSi16 encode  libffrs/pyrsi16.hpp:128-140, libffrs/rsi16v_impl.hpp:93-119, libffrs/ntt.hpp:180-302
  RSi16 repair  libffrs/pyrsi16.hpp:588-631, libffrs/rsi16v_impl.hpp:122-512, libffrs/ntt.hpp:133-579
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_u16p = ctypes.POINTER(ctypes.c_uint16)
c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_szp = ctypes.POINTER(ctypes.c_size_t)

NO_ERRORS, NO_ERRORS_ZERO, REPAIR_OK, REPAIR_FAIL, ERROR_LOCATION_FAIL = 0, 1, 2, 3, 4
ASSERT_RAISED = 255


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_gf256_tables.argtypes = [ctypes.c_uint, ctypes.c_uint, c_u8p, c_u8p]
        L.orc_gf256_tables.restype = ctypes.c_int
        L.orc_rs256_new.argtypes = [ctypes.c_uint, ctypes.c_uint, ctypes.c_uint]
        L.orc_rs256_new.restype = ctypes.c_void_p
        L.orc_rs256_free.argtypes = [ctypes.c_void_p]
        L.orc_rs256_generator.argtypes = [ctypes.c_void_p, c_u8p]
        L.orc_rs256_encode.argtypes = [ctypes.c_void_p, ctypes.c_size_t, c_u8p, ctypes.c_size_t, c_u8p]
        L.orc_rs256_encode.restype = ctypes.c_size_t
        L.orc_rs256_synds.argtypes = [ctypes.c_void_p, c_u8p, ctypes.c_size_t, c_u8p, c_u8p]
        L.orc_rs256_decode.argtypes = [ctypes.c_void_p, c_u8p, ctypes.c_size_t, c_u8p]
        L.orc_rs256_decode.restype = ctypes.c_int
        L.orc_rs256_encode_batch.argtypes = [ctypes.c_void_p, c_u8p, ctypes.c_size_t, ctypes.c_size_t, c_u8p,
                                             ctypes.c_int]
        L.orc_rs256_decode_batch.argtypes = [ctypes.c_void_p, c_u8p, c_u8p, ctypes.c_size_t, ctypes.c_size_t,
                                             c_u8p, ctypes.c_int]
        L.orc_rsi16_new.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_uint32]
        L.orc_rsi16_new.restype = ctypes.c_void_p
        L.orc_rsi16_free.argtypes = [ctypes.c_void_p]
        L.orc_rsi16_root.argtypes = [ctypes.c_void_p]
        L.orc_rsi16_root.restype = ctypes.c_uint32
        L.orc_rsi16_ntt_len.argtypes = [ctypes.c_void_p]
        L.orc_rsi16_ntt_len.restype = ctypes.c_size_t
        L.orc_rsi16_table.argtypes = [ctypes.c_void_p, ctypes.c_int, c_u32p, ctypes.c_size_t]
        L.orc_rsi16_table.restype = ctypes.c_size_t
        L.orc_rsi16_encode.argtypes = [ctypes.c_void_p, c_u16p, ctypes.c_size_t, c_u16p, ctypes.c_int, ctypes.c_int]
        L.orc_rsi16_synds.argtypes = [ctypes.c_void_p, c_u16p, c_u16p, c_u32p]
        L.orc_rsi16_sugiyama.argtypes = [ctypes.c_void_p, c_u32p, c_u32p, c_u32p]
        L.orc_rsi16_sugiyama.restype = ctypes.c_uint32
        L.orc_rsi16_repair.argtypes = [ctypes.c_void_p, c_u16p, c_u16p]
        L.orc_rsi16_repair.restype = ctypes.c_int
        L.orc_rsi16_repair_batch.argtypes = [ctypes.c_void_p, c_u16p, c_u16p, ctypes.c_size_t, c_u8p, ctypes.c_int]
        L.orc_rsi16_repair_known.argtypes = [ctypes.c_void_p, c_u16p, c_u16p, c_szp, ctypes.c_size_t]
        L.orc_rsi16_repair_known.restype = ctypes.c_int
        L.orc_rsi16_repair_interleaved_simd.argtypes = [ctypes.c_void_p, ctypes.c_uint, c_u16p, c_u16p,
                                                        ctypes.c_size_t]
        L.orc_rsi16_repair_interleaved_simd.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(arr, ct):
    return arr.ctypes.data_as(ct)


def default_threads():
    return max(1, min(16, os.cpu_count() or 1))


def gf256_tables(primitive=2, polynomial=0x11D):
    e = np.zeros(256, np.uint8)
    l = np.zeros(256, np.uint8)
    ok = lib().orc_gf256_tables(primitive, polynomial, _p(e, c_u8p), _p(l, c_u8p))
    return ok, e, l


class RS256:
    """Oracle RS256 codec for one ecc_len (field: primitive, polynomial)."""

    def __init__(self, ecc_len, primitive=2, polynomial=0x11D):
        self.ecc_len = int(ecc_len)
        self._h = lib().orc_rs256_new(self.ecc_len, primitive, polynomial)
        if not self._h:
            raise ValueError("unsupported RS256 parameters")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_rs256_free(self._h)
            self._h = None

    def generator(self):
        g = np.zeros(self.ecc_len + 1, np.uint8)
        lib().orc_rs256_generator(self._h, _p(g, c_u8p))
        return g

    def encode(self, buf, message_len):
        """Mirror of RS256.encode(buf) for a codec with the given message_len."""
        a = np.frombuffer(bytes(buf), np.uint8).copy()
        if a.size == 0:
            return bytearray()
        nblk = -(-a.size // message_len)
        out = np.zeros(nblk * self.ecc_len, np.uint8)
        n = lib().orc_rs256_encode(self._h, message_len, _p(a, c_u8p), a.size, _p(out, c_u8p))
        return bytearray(out[:n].tobytes())

    def encode_batch(self, msgs, threads=None):
        msgs = np.ascontiguousarray(msgs, np.uint8)
        count, k = msgs.shape
        out = np.zeros((count, self.ecc_len), np.uint8)
        lib().orc_rs256_encode_batch(self._h, _p(msgs, c_u8p), count, k, _p(out, c_u8p),
                                     threads or default_threads())
        return out

    def synds(self, msg, ecc):
        msg = np.ascontiguousarray(msg, np.uint8)
        ecc = np.ascontiguousarray(ecc, np.uint8)
        s = np.zeros(self.ecc_len, np.uint8)
        lib().orc_rs256_synds(self._h, _p(msg, c_u8p), msg.size, _p(ecc, c_u8p), _p(s, c_u8p))
        return s

    def decode(self, msg, ecc):
        """In-place on numpy uint8 arrays; returns bool like RS256.repair."""
        return bool(lib().orc_rs256_decode(self._h, _p(msg, c_u8p), msg.size, _p(ecc, c_u8p)))

    def decode_batch(self, msgs, eccs, threads=None):
        """In place on contiguous [count, k] / [count, ecc] uint8 arrays; returns uint8 status."""
        assert msgs.flags.c_contiguous and eccs.flags.c_contiguous
        count, k = msgs.shape
        st = np.zeros(count, np.uint8)
        lib().orc_rs256_decode_batch(self._h, _p(msgs, c_u8p), _p(eccs, c_u8p), count, k, _p(st, c_u8p),
                                     threads or default_threads())
        return st


class RSi16:
    """Oracle RSi16 codec (interleave 1)."""

    TABLES = {"exp": 0, "log": 1, "roots_ntt": 2, "pntt_shift": 3, "pintt_shift": 4, "mix": 5,
              "roots_ecc": 6, "roots_i_ecc": 7}

    def __init__(self, block_len, ecc_len, primitive=3):
        self.block_len, self.ecc_len = int(block_len), int(ecc_len)
        self.message_len = self.block_len - self.ecc_len
        self._h = lib().orc_rsi16_new(self.block_len, self.ecc_len, primitive)
        if not self._h:
            raise RuntimeError("unsupported RSi16 parameters")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_rsi16_free(self._h)
            self._h = None

    @property
    def root(self):
        return lib().orc_rsi16_root(self._h)

    @property
    def ntt_len(self):
        return lib().orc_rsi16_ntt_len(self._h)

    def table(self, name):
        w = self.TABLES[name]
        n = lib().orc_rsi16_table(self._h, w, None, 0)
        out = np.zeros(n, np.uint32)
        lib().orc_rsi16_table(self._h, w, _p(out, c_u32p), n)
        return out

    def encode(self, msgs, vector_lanes=True, threads=None):
        msgs = np.ascontiguousarray(msgs, np.uint16).reshape(-1, self.message_len)
        out = np.zeros((msgs.shape[0], self.ecc_len), np.uint16)
        lib().orc_rsi16_encode(self._h, _p(msgs, c_u16p), msgs.shape[0], _p(out, c_u16p),
                               1 if vector_lanes else 0, threads or default_threads())
        return out

    def synds(self, msg, ecc):
        msg = np.ascontiguousarray(msg, np.uint16)
        ecc = np.ascontiguousarray(ecc, np.uint16)
        s = np.zeros(self.ecc_len, np.uint32)
        lib().orc_rsi16_synds(self._h, _p(msg, c_u16p), _p(ecc, c_u16p), _p(s, c_u32p))
        return s

    def sugiyama(self, synds):
        synds = np.ascontiguousarray(synds, np.uint32)
        loc = np.zeros(self.ecc_len, np.uint32)
        ev = np.zeros(self.ecc_len, np.uint32)
        L = lib().orc_rsi16_sugiyama(self._h, _p(synds, c_u32p), _p(loc, c_u32p), _p(ev, c_u32p))
        return L, loc, ev

    def repair(self, msg, ecc):
        """In place on uint16 arrays; returns a status code (255: reference raises)."""
        return lib().orc_rsi16_repair(self._h, _p(msg, c_u16p), _p(ecc, c_u16p))

    def repair_batch(self, msgs, eccs, threads=None):
        assert msgs.flags.c_contiguous and eccs.flags.c_contiguous
        count = msgs.shape[0]
        st = np.zeros(count, np.uint8)
        lib().orc_rsi16_repair_batch(self._h, _p(msgs, c_u16p), _p(eccs, c_u16p), count, _p(st, c_u8p),
                                     threads or default_threads())
        return st

    def repair_interleaved_simd(self, width, msg, ecc, interleave):
        """Interleaved repair with the SIMD words: columns in coupled groups of `width`
        (4, 8, 16), in place on the interleaved uint16 buffers; returns the call's status
        (255: the reference raises, with the groups before the raising one written back)."""
        assert msg.flags.c_contiguous and ecc.flags.c_contiguous
        return lib().orc_rsi16_repair_interleaved_simd(self._h, width, _p(msg, c_u16p), _p(ecc, c_u16p), interleave)

    def repair_known(self, msg, ecc, positions):
        pos = np.ascontiguousarray(positions, np.uint64).astype(np.uintp)
        return lib().orc_rsi16_repair_known(self._h, _p(msg, c_u16p), _p(ecc, c_u16p), _p(pos, c_szp), pos.size)
