"""ctypes binding of the C ABI in include/ffrs_gpu.h (libffrs_gpu.so).

The pybind11 module covers the reference's host-buffer API; this binding also
exposes the device-pointer entry points, device memory and stream timing, for
callers whose data already sits in HBM (benchmarks, pipelines, other runtimes).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libffrs_gpu.so")

FFRS_OK = 0
_lib = None

c_vp = ctypes.c_void_p
c_sz = ctypes.c_size_t
c_int = ctypes.c_int
c_uint = ctypes.c_uint
c_u32 = ctypes.c_uint32
c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_u16p = ctypes.POINTER(ctypes.c_uint16)
c_u32p = ctypes.POINTER(ctypes.c_uint32)

# (name, restype, argtypes) for every entry point of include/ffrs_gpu.h
SIGNATURES = [
    ("ffrs_abi_version", c_int, []),
    ("ffrs_last_error", ctypes.c_char_p, []),
    ("ffrs_ctx_create", c_int, [c_int, ctypes.POINTER(c_vp)]),
    ("ffrs_ctx_destroy", None, [c_vp]),
    ("ffrs_ctx_synchronize", c_int, [c_vp]),
    ("ffrs_ctx_device_info", c_int, [c_vp, ctypes.c_char_p, c_sz, ctypes.POINTER(c_int), ctypes.POINTER(c_sz),
                                     ctypes.POINTER(c_sz)]),
    ("ffrs_device_alloc", c_int, [c_vp, c_sz, ctypes.POINTER(c_vp)]),
    ("ffrs_device_free", c_int, [c_vp, c_vp]),
    ("ffrs_host_alloc", c_int, [c_vp, c_sz, ctypes.POINTER(c_vp)]),
    ("ffrs_host_free", c_int, [c_vp, c_vp]),
    ("ffrs_memcpy_h2d_async", c_int, [c_vp, c_vp, c_vp, c_sz]),
    ("ffrs_memcpy_d2h_async", c_int, [c_vp, c_vp, c_vp, c_sz]),
    ("ffrs_memcpy_h2d", c_int, [c_vp, c_vp, c_vp, c_sz]),
    ("ffrs_memcpy_d2h", c_int, [c_vp, c_vp, c_vp, c_sz]),
    ("ffrs_memcpy_d2d", c_int, [c_vp, c_vp, c_vp, c_sz]),
    ("ffrs_timer_begin", c_int, [c_vp]),
    ("ffrs_timer_end", c_int, [c_vp, ctypes.POINTER(ctypes.c_float)]),
    ("ffrs_ctx_kernel_timing", c_int, [c_vp, c_int]),
    ("ffrs_ctx_kernel_times", c_int, [c_vp, ctypes.POINTER(ctypes.c_float), c_int, ctypes.POINTER(c_int)]),
    ("ffrs_ctx_capture_begin", c_int, [c_vp]),
    ("ffrs_ctx_capture_end", c_int, [c_vp, ctypes.POINTER(c_vp)]),
    ("ffrs_graph_launch", c_int, [c_vp]),
    ("ffrs_graph_destroy", None, [c_vp]),
    ("ffrs_rs256_create", c_int, [c_vp, c_uint, c_uint, c_uint, ctypes.POINTER(c_vp)]),
    ("ffrs_rs256_destroy", None, [c_vp]),
    ("ffrs_rs256_generator", c_int, [c_vp, c_u8p, c_u8p]),
    ("ffrs_rs256_set_repair_mode", c_int, [c_vp, c_int]),
    ("ffrs_rs256_encode", c_int, [c_vp, c_sz, c_vp, c_sz, c_vp, c_sz]),
    ("ffrs_rs256_repair", c_int, [c_vp, c_sz, c_vp, c_vp, c_sz, c_vp]),
    ("ffrs_rs256_synds", c_int, [c_vp, c_sz, c_vp, c_vp, c_sz, c_vp]),
    ("ffrs_rs256_encode_device", c_int, [c_vp, c_sz, c_vp, c_sz, c_vp]),
    ("ffrs_rs256_repair_device", c_int, [c_vp, c_sz, c_vp, c_vp, c_sz, c_vp]),
    ("ffrs_rs256_repair_device_noalloc", c_int, [c_vp, c_sz, c_vp, c_vp, c_sz, c_vp]),
    ("ffrs_rsi16_create", c_int, [c_vp, c_sz, c_sz, c_u32, ctypes.POINTER(c_vp)]),
    ("ffrs_rsi16_destroy", None, [c_vp]),
    ("ffrs_rsi16_set_groups", c_int, [c_vp, c_uint, c_uint]),
    ("ffrs_rsi16_get_groups", c_int, [c_vp, ctypes.POINTER(c_uint), ctypes.POINTER(c_uint)]),
    ("ffrs_rsi16_root", c_u32, [c_vp]),
    ("ffrs_rsi16_ntt_len", c_sz, [c_vp]),
    ("ffrs_rsi16_encode", c_int, [c_vp, c_vp, c_sz, c_vp, c_int]),
    ("ffrs_rsi16_encode_interleaved", c_int, [c_vp, c_vp, c_sz, c_sz, c_vp, c_int]),
    ("ffrs_rsi16_encode_interleaved_device", c_int, [c_vp, c_vp, c_sz, c_sz, c_vp, c_int]),
    ("ffrs_rsi16_repair", c_int, [c_vp, c_vp, c_vp, c_sz, c_vp]),
    ("ffrs_rsi16_repair_known", c_int, [c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    ("ffrs_rsi16_synds", c_int, [c_vp, c_vp, c_vp, c_sz, c_vp]),
    ("ffrs_rsi16_sugiyama", c_int, [c_vp, c_vp, c_sz, c_vp, c_vp, c_vp]),
    ("ffrs_rsi16_roots", c_int, [c_vp, c_vp, c_vp]),
    ("ffrs_rsi16_encode_device", c_int, [c_vp, c_vp, c_sz, c_vp, c_int]),
    ("ffrs_rsi16_repair_device", c_int, [c_vp, c_vp, c_vp, c_sz, c_vp]),
    ("ffrs_rsi16_repair_interleaved_simd", c_int, [c_vp, c_vp, c_vp, c_sz, ctypes.c_uint, c_vp]),
    ("ffrs_rsi16_set_encode_kernel", c_int, [c_vp, c_int]),
    ("ffrs_rsi16_set_repair_syndromes", c_int, [c_vp, c_int]),
    ("ffrs_rs256_plan_create", c_int, [c_vp, c_sz, c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(c_vp)]),
    ("ffrs_rs256_plan_destroy", None, [c_vp]),
    ("ffrs_rs256_plan_use_graph", c_int, [c_vp, c_int]),
    ("ffrs_rs256_plan_set_class_mode", c_int, [c_vp, ctypes.c_uint, c_int]),
    ("ffrs_rs256_plan_encode", c_int, [c_vp, c_vp, c_sz, c_vp, c_sz]),
    ("ffrs_rs256_plan_repair", c_int, [c_vp, c_vp, c_sz, c_vp, c_sz, c_vp]),
    ("ffrs_rsi16_plan_create", c_int, [c_vp, c_sz, c_vp, c_vp, c_vp, c_vp, ctypes.c_uint32, ctypes.POINTER(c_vp)]),
    ("ffrs_rsi16_plan_destroy", None, [c_vp]),
    ("ffrs_rsi16_plan_use_graph", c_int, [c_vp, c_int]),
    ("ffrs_rsi16_ntt_op", c_int, [c_vp, c_int, ctypes.c_uint, c_vp, c_vp, ctypes.c_uint32, c_vp, c_vp, c_vp]),
    ("ffrs_rsi16_plan_set_class_groups", c_int, [c_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint, ctypes.c_uint]),
    ("ffrs_rsi16_plan_encode", c_int, [c_vp, c_vp, c_sz, c_vp, c_sz, c_int]),
    ("ffrs_circ16_create", c_int, [c_vp, c_vp, c_sz, ctypes.c_uint, ctypes.POINTER(c_vp)]),
    ("ffrs_circ16_destroy", None, [c_vp]),
    ("ffrs_circ16_message_len", c_sz, [c_vp]),
    ("ffrs_circ16_ecc_len", c_sz, [c_vp]),
    ("ffrs_circ16_encode", c_int, [c_vp, c_vp, c_sz, c_vp]),
    ("ffrs_circ16_repair", c_int, [c_vp, c_vp, c_vp, ctypes.POINTER(c_int)]),
    ("ffrs_circ16_find_outer_error_locations", c_int, [c_vp, c_vp, c_vp, c_sz, c_vp, ctypes.POINTER(c_sz)]),
    ("ffrs_rsi16_plan_repair", c_int, [c_vp, c_vp, c_sz, c_vp, c_sz, c_vp]),
]


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class FfrsError(RuntimeError):
    pass


def check(code, what):
    if code != FFRS_OK:
        raise FfrsError(f"{what} failed ({code}): {lib().ffrs_last_error().decode()}")


class Context:
    """One device and one stream (ffrs_ctx)."""

    def __init__(self, device=0):
        h = c_vp()
        check(lib().ffrs_ctx_create(device, ctypes.byref(h)), "ffrs_ctx_create")
        self.h = h

    def close(self):
        if self.h:
            lib().ffrs_ctx_destroy(self.h)
            self.h = None

    def sync(self):
        check(lib().ffrs_ctx_synchronize(self.h), "synchronize")

    def info(self):
        name = ctypes.create_string_buffer(256)
        cus, fr, tot = c_int(), c_sz(), c_sz()
        check(lib().ffrs_ctx_device_info(self.h, name, 256, ctypes.byref(cus), ctypes.byref(fr), ctypes.byref(tot)),
              "device_info")
        return name.value.decode(), cus.value, fr.value, tot.value

    def alloc(self, nbytes):
        p = c_vp()
        check(lib().ffrs_device_alloc(self.h, nbytes, ctypes.byref(p)), "alloc")
        return DeviceBuffer(self, p.value, nbytes)

    def host_alloc(self, nbytes):
        """Pinned host memory as a uint8 numpy array (freed with host_free)."""
        import numpy as np
        p = c_vp()
        check(lib().ffrs_host_alloc(self.h, nbytes, ctypes.byref(p)), "host_alloc")
        arr = np.ctypeslib.as_array((ctypes.c_uint8 * nbytes).from_address(p.value))
        return arr

    def host_free(self, arr):
        check(lib().ffrs_host_free(self.h, arr.ctypes.data), "host_free")

    def timer_begin(self):
        check(lib().ffrs_timer_begin(self.h), "timer_begin")

    def timer_end(self):
        ms = ctypes.c_float()
        check(lib().ffrs_timer_end(self.h, ctypes.byref(ms)), "timer_end")
        return ms.value

    def capture_begin(self):
        check(lib().ffrs_ctx_capture_begin(self.h), "capture_begin")

    def capture_end(self):
        g = c_vp()
        check(lib().ffrs_ctx_capture_end(self.h, ctypes.byref(g)), "capture_end")
        return Graph(g)

    def kernel_timing(self, enable=True):
        check(lib().ffrs_ctx_kernel_timing(self.h, 1 if enable else 0), "kernel_timing")

    def kernel_times(self):
        buf = (ctypes.c_float * 8)()
        n = c_int()
        check(lib().ffrs_ctx_kernel_times(self.h, buf, 8, ctypes.byref(n)), "kernel_times")
        return [buf[i] for i in range(min(n.value, 8))]


class Graph:
    """A captured sequence of device work, replayed on its context's stream."""

    def __init__(self, h):
        self.h = h

    def launch(self):
        check(lib().ffrs_graph_launch(self.h), "graph launch")

    def close(self):
        if self.h:
            lib().ffrs_graph_destroy(self.h)
            self.h = None


class DeviceBuffer:
    def __init__(self, ctx, ptr, nbytes):
        self.ctx, self.ptr, self.nbytes = ctx, ptr, nbytes

    def free(self):
        if self.ptr:
            check(lib().ffrs_device_free(self.ctx.h, self.ptr), "free")
            self.ptr = None

    def upload(self, arr):
        assert arr.nbytes <= self.nbytes and arr.flags.c_contiguous
        check(lib().ffrs_memcpy_h2d(self.ctx.h, self.ptr, arr.ctypes.data, arr.nbytes), "h2d")

    def download(self, arr):
        assert arr.nbytes <= self.nbytes and arr.flags.c_contiguous
        check(lib().ffrs_memcpy_d2h(self.ctx.h, arr.ctypes.data, self.ptr, arr.nbytes), "d2h")

    def copy_from(self, other, nbytes=None):
        n = other.nbytes if nbytes is None else nbytes
        check(lib().ffrs_memcpy_d2d(self.ctx.h, self.ptr, other.ptr, n), "d2d")

    def upload_async(self, arr, offset=0):
        """Enqueue a copy from (pinned) host memory into [offset, offset + arr.nbytes)."""
        assert offset + arr.nbytes <= self.nbytes and arr.flags.c_contiguous
        check(lib().ffrs_memcpy_h2d_async(self.ctx.h, self.ptr + offset, arr.ctypes.data, arr.nbytes), "h2d async")

    def download_async(self, arr, offset=0):
        assert offset + arr.nbytes <= self.nbytes and arr.flags.c_contiguous
        check(lib().ffrs_memcpy_d2h_async(self.ctx.h, arr.ctypes.data, self.ptr + offset, arr.nbytes), "d2h async")

    def view(self, offset, nbytes):
        """A non-owning window [offset, offset + nbytes) of this buffer."""
        assert 0 <= offset and offset + nbytes <= self.nbytes
        return BufferView(self.ptr + offset, nbytes)


class BufferView:
    def __init__(self, ptr, nbytes):
        self.ptr, self.nbytes = ptr, nbytes


class RS256Codec:
    def __init__(self, ctx, ecc_len, primitive=2, polynomial=0x11D):
        h = c_vp()
        check(lib().ffrs_rs256_create(ctx.h, ecc_len, primitive, polynomial, ctypes.byref(h)), "ffrs_rs256_create")
        self.h, self.ctx, self.ecc_len = h, ctx, ecc_len

    def close(self):
        if self.h:
            lib().ffrs_rs256_destroy(self.h)
            self.h = None

    def set_repair_mode(self, mode):
        check(lib().ffrs_rs256_set_repair_mode(self.h, int(mode)), "set_repair_mode")

    def encode_device(self, message_len, d_msg, count, d_ecc):
        check(lib().ffrs_rs256_encode_device(self.h, message_len, d_msg.ptr, count, d_ecc.ptr), "rs256 encode")

    def repair_device(self, message_len, d_msg, d_ecc, count, d_status):
        check(lib().ffrs_rs256_repair_device(self.h, message_len, d_msg.ptr, d_ecc.ptr, count, d_status.ptr),
              "rs256 repair")

    def repair_device_noalloc(self, message_len, d_msg, d_ecc, count, d_status):
        """ffrs_rs256_repair_device_noalloc: the same repair over the library's static device
        workspace (no device allocation)."""
        check(lib().ffrs_rs256_repair_device_noalloc(self.h, message_len, d_msg.ptr, d_ecc.ptr, count,
                                                     d_status.ptr), "rs256 repair (allocation-free)")


class RSi16Codec:
    def __init__(self, ctx, block_len, ecc_len, primitive=3):
        h = c_vp()
        check(lib().ffrs_rsi16_create(ctx.h, block_len, ecc_len, primitive, ctypes.byref(h)), "ffrs_rsi16_create")
        self.h, self.ctx = h, ctx
        self.block_len, self.ecc_len, self.message_len = block_len, ecc_len, block_len - ecc_len

    def set_repair_syndromes(self, mode):
        check(lib().ffrs_rsi16_set_repair_syndromes(self.h, int(mode)), "set_repair_syndromes")

    def set_encode_kernel(self, kernel):
        """0 automatic, 1 the butterfly kernels, 2 the int8 MFMA contraction wherever it applies."""
        check(lib().ffrs_rsi16_set_encode_kernel(self.h, int(kernel)), "set_encode_kernel")

    def close(self):
        if self.h:
            lib().ffrs_rsi16_destroy(self.h)
            self.h = None

    def set_groups(self, g_encode=0, g_repair=0):
        check(lib().ffrs_rsi16_set_groups(self.h, int(g_encode), int(g_repair)), "set_groups")

    def groups(self):
        a, b = c_uint(), c_uint()
        check(lib().ffrs_rsi16_get_groups(self.h, ctypes.byref(a), ctypes.byref(b)), "get_groups")
        return a.value, b.value

    def encode_device(self, d_msg, count, d_ecc, vector_lanes=True):
        check(lib().ffrs_rsi16_encode_device(self.h, d_msg.ptr, count, d_ecc.ptr, 1 if vector_lanes else 0),
              "rsi16 encode")

    def repair_device(self, d_msg, d_ecc, count, d_status):
        check(lib().ffrs_rsi16_repair_device(self.h, d_msg.ptr, d_ecc.ptr, count, d_status.ptr), "rsi16 repair")


def _arr(a, dtype):
    import numpy as np
    return np.ascontiguousarray(a, dtype=dtype)


class RS256Plan:
    """A mixed-size RS256 work list (per-codeword ecc_len, message_len and byte offsets)."""

    def __init__(self, ctx, ecc_len, message_len, msg_off, ecc_off, graph=False):
        import numpy as np
        self._keep = [_arr(ecc_len, np.uint8), _arr(message_len, np.uint8), _arr(msg_off, np.uint64),
                      _arr(ecc_off, np.uint64)]
        h = c_vp()
        check(lib().ffrs_rs256_plan_create(ctx.h, len(self._keep[0]), *[a.ctypes.data for a in self._keep],
                                           ctypes.byref(h)), "ffrs_rs256_plan_create")
        self.h, self.ctx = h, ctx
        if graph:
            check(lib().ffrs_rs256_plan_use_graph(self.h, 1), "use_graph")

    def close(self):
        if self.h:
            lib().ffrs_rs256_plan_destroy(self.h)
            self.h = None

    def encode(self, d_msgs, d_eccs):
        check(lib().ffrs_rs256_plan_encode(self.h, d_msgs.ptr, d_msgs.nbytes, d_eccs.ptr, d_eccs.nbytes),
              "rs256 plan encode")

    def repair(self, d_msgs, d_eccs, d_status):
        check(lib().ffrs_rs256_plan_repair(self.h, d_msgs.ptr, d_msgs.nbytes, d_eccs.ptr, d_eccs.nbytes,
                                           d_status.ptr), "rs256 plan repair")

    def classes(self):
        """The distinct ecc_len of the plan's codewords."""
        import numpy as np
        return sorted(int(e) for e in np.unique(self._keep[0]))

    def class_sizes(self):
        """{ecc_len: number of codewords}."""
        import numpy as np
        e, c = np.unique(self._keep[0], return_counts=True)
        return {int(a): int(b) for a, b in zip(e, c)}

    def set_class_mode(self, ecc_len, mode):
        """Repair kernel of the codewords of one ecc_len: 0 automatic, 1 one per lane, 2 one per wave."""
        check(lib().ffrs_rs256_plan_set_class_mode(self.h, int(ecc_len), int(mode)), "set_class_mode")


class RSi16Plan:
    """A mixed-size RSi16 work list (per-codeword block_len, ecc_len and u16 offsets)."""

    def __init__(self, ctx, block_len, ecc_len, msg_off, ecc_off, graph=False, primitive=3):
        import numpy as np
        self._keep = [_arr(block_len, np.uint32), _arr(ecc_len, np.uint32), _arr(msg_off, np.uint64),
                      _arr(ecc_off, np.uint64)]
        h = c_vp()
        check(lib().ffrs_rsi16_plan_create(ctx.h, len(self._keep[0]), *[a.ctypes.data for a in self._keep],
                                           primitive, ctypes.byref(h)), "ffrs_rsi16_plan_create")
        self.h, self.ctx = h, ctx
        if graph:
            check(lib().ffrs_rsi16_plan_use_graph(self.h, 1), "use_graph")

    def close(self):
        if self.h:
            lib().ffrs_rsi16_plan_destroy(self.h)
            self.h = None

    def encode(self, d_msgs, d_eccs, vector_lanes=True):
        check(lib().ffrs_rsi16_plan_encode(self.h, d_msgs.ptr, d_msgs.nbytes // 2, d_eccs.ptr, d_eccs.nbytes // 2,
                                           1 if vector_lanes else 0), "rsi16 plan encode")

    def repair(self, d_msgs, d_eccs, d_status):
        check(lib().ffrs_rsi16_plan_repair(self.h, d_msgs.ptr, d_msgs.nbytes // 2, d_eccs.ptr, d_eccs.nbytes // 2,
                                           d_status.ptr), "rsi16 plan repair")

    def classes(self):
        """The distinct (block_len, ecc_len) of the plan's codewords."""
        return sorted(self.class_sizes())

    def class_sizes(self):
        """{(block_len, ecc_len): number of codewords}."""
        import collections
        return dict(collections.Counter(zip((int(x) for x in self._keep[0]), (int(x) for x in self._keep[1]))))

    def set_class_groups(self, block_len, ecc_len, l_encode=0, g_repair=0):
        """Lanes per codeword of one class: l_encode for the encode (ecc_len <= 32; 0 keeps it),
        g_repair for the repair (0: the register code)."""
        check(lib().ffrs_rsi16_plan_set_class_groups(self.h, int(block_len), int(ecc_len), int(l_encode),
                                                     int(g_repair)), "set_class_groups")
