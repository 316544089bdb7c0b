"""libxstream-amd: MI355X-native DBCSR stack multiply behind the reference's C API.

Python bindings (ctypes) to ``libxstream_amd.so``: the libxstream runtime (streams, events,
memory), the c_dbcsr_acc_* shim and the libsmm_acc_* entry points, mirroring
samples/smm/acc_libsmm.h:31-57 and libxstream/libxstream.h:15-79 of hfp/libxstream.

The HIP library is the only compute path: if it is not built, ``lib()`` raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libxstream_amd.so")

# samples/smm/acc_libsmm.h:31-36
REAL_4 = 1
REAL_8 = 3
COMPLEX_4 = 5
COMPLEX_8 = 7
MAX_KERNEL_DIM = 80  # samples/smm/acc_bench.h:22-24

_lib = None


class LibxstreamError(RuntimeError):
    pass


class SmmConfig(ctypes.Structure):
    """libsmm_acc_smm_config_t (include/libxstream_amd.h)."""

    _fields_ = [
        ("variant", ctypes.c_int),
        ("tm", ctypes.c_int),
        ("tn", ctypes.c_int),
        ("nt", ctypes.c_int),
        ("bs", ctypes.c_int),
        ("kt", ctypes.c_int),
        ("flush", ctypes.c_int),
        ("gflops", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}

    def __repr__(self):
        return "SmmConfig(%s)" % ", ".join("%s=%s" % kv for kv in self.as_dict().items())


class BatchStack(ctypes.Structure):
    """libsmm_acc_batch_stack_t (include/libxstream_amd.h)."""

    _fields_ = [
        ("dev_param_stack", ctypes.c_void_p),
        ("stack_size", ctypes.c_int),
        ("m", ctypes.c_int),
        ("n", ctypes.c_int),
        ("k", ctypes.c_int),
        ("dev_a_data", ctypes.c_void_p),
        ("dev_b_data", ctypes.c_void_p),
        ("dev_c_data", ctypes.c_void_p),
    ]


def _declare(L):
    vp, i, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
    pvp, pi, psz = ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_size_t)
    sigs = {
        "libxstream_init": (i, []),
        "libxstream_finalize": (i, []),
        "libxstream_device_count": (i, [pi]),
        "libxstream_device_set_active": (i, [i]),
        "libxstream_device_sync": (i, []),
        "libxstream_device_info": (i, [pi, ctypes.c_char_p, sz]),
        "libxstream_stream_create": (i, [pvp, ctypes.c_char_p, i]),
        "libxstream_stream_destroy": (i, [vp]),
        "libxstream_stream_sync": (i, [vp]),
        "libxstream_stream_native": (i, [vp, pvp]),
        "libxstream_event_create": (i, [pvp]),
        "libxstream_event_destroy": (i, [vp]),
        "libxstream_event_record": (i, [vp, vp]),
        "libxstream_event_sync": (i, [vp]),
        "libxstream_event_elapsed": (i, [vp, vp, ctypes.POINTER(ctypes.c_double)]),
        "libxstream_mem_allocate": (i, [pvp, sz]),
        "libxstream_mem_deallocate": (i, [vp]),
        "libxstream_mem_copy_h2d": (i, [vp, vp, sz, vp]),
        "libxstream_mem_copy_d2h": (i, [vp, vp, sz, vp]),
        "libxstream_mem_copy_d2d": (i, [vp, vp, sz, vp]),
        "libxstream_mem_zero": (i, [vp, sz, sz, vp]),
        "libxstream_mem_info": (i, [psz, psz]),
        "libxstream_error_consume": (i, [ctypes.c_char_p, sz]),
        "libxstream_profile_enable": (i, [i]),
        "libxstream_profile_reset": (i, []),
        "libxstream_profile_get": (i, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_long)] + [ctypes.POINTER(ctypes.c_double)] * 3),
        "libsmm_acc_init": (i, []),
        "libsmm_acc_finalize": (i, []),
        "libsmm_acc_is_thread_safe": (i, []),
        "libsmm_acc_process_suitable": (i, [i, i, i, i, i, i, i]),
        "libsmm_acc_process": (i, [vp, vp, i, i, vp, vp, vp, i, i, i, i, i, vp, vp]),
        "libsmm_acc_transpose": (i, [vp, i, i, vp, i, i, i, i, vp]),
        "libsmm_acc_process_config": (i, [ctypes.POINTER(SmmConfig), vp, i, i, vp, vp, vp, i, i, i, vp]),
        "libsmm_acc_config_get": (i, [i, i, i, i, ctypes.POINTER(SmmConfig), pi]),
        "libsmm_acc_config_set": (i, [i, i, i, i, ctypes.POINTER(SmmConfig)]),
        "libsmm_acc_config_valid": (i, [i, i, i, i, ctypes.POINTER(SmmConfig)]),
        "libsmm_acc_kernel_variants": (i, [i, pi, i]),
        "libsmm_acc_params_load": (i, [ctypes.c_char_p]),
        "libsmm_acc_params_save": (i, [ctypes.c_char_p, ctypes.c_char_p]),
        "libsmm_acc_read_ref_params": (i, [ctypes.c_char_p, i, i, ctypes.c_char_p, sz, pi, pi, ctypes.POINTER(ctypes.c_double)]),
        "libsmm_acc_batch_create": (i, [i, ctypes.POINTER(BatchStack), i, pvp]),
        "libsmm_acc_batch_process": (i, [vp, vp]),
        "libsmm_acc_batch_replay": (i, [vp, vp]),
        "libsmm_acc_batch_info": (i, [vp, pi, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "libsmm_acc_batch_destroy": (i, [vp]),
        "c_dbcsr_acc_init": (i, []),
        "c_dbcsr_acc_finalize": (i, []),
        "c_dbcsr_acc_get_ndevices": (i, [pi]),
        "c_dbcsr_acc_set_active_device": (i, [i]),
        "c_calculate_norms": (i, [vp, i, vp, vp, vp, vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """The loaded HIP library; raises when it is not built (there is no other compute path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibxstreamError("%s is not built (run __graft_entry__.build())" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def check(code, what="libxstream"):
    if code != 0:
        msg = ctypes.create_string_buffer(256)
        lib().libxstream_error_consume(msg, 256)
        raise LibxstreamError("%s failed (code=%d): %s" % (what, code, msg.value.decode(errors="replace")))
    return code


def typeid(dtype):
    return REAL_8 if np.dtype(dtype) == np.float64 else REAL_4


def nptype(tid):
    return np.float64 if tid == REAL_8 else np.float32


def init():
    check(lib().libsmm_acc_init(), "libsmm_acc_init")


def finalize():
    check(lib().libsmm_acc_finalize(), "libsmm_acc_finalize")


def device_count():
    n = ctypes.c_int(0)
    check(lib().libxstream_device_count(ctypes.byref(n)))
    return n.value


def device_info():
    ncus = ctypes.c_int(0)
    name = ctypes.create_string_buffer(256)
    check(lib().libxstream_device_info(ctypes.byref(ncus), name, 256), "device_info")
    return ncus.value, name.value.decode()


class Stream:
    def __init__(self, name="stream"):
        h = ctypes.c_void_p()
        check(lib().libxstream_stream_create(ctypes.byref(h), name.encode(), 0), "stream_create")
        self.handle = h

    def sync(self):
        check(lib().libxstream_stream_sync(self.handle), "stream_sync")

    def native(self):
        h = ctypes.c_void_p()
        check(lib().libxstream_stream_native(self.handle, ctypes.byref(h)))
        return h.value

    def close(self):
        if self.handle:
            check(lib().libxstream_stream_destroy(self.handle), "stream_destroy")
            self.handle = None


class Event:
    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().libxstream_event_create(ctypes.byref(h)), "event_create")
        self.handle = h

    def record(self, stream):
        check(lib().libxstream_event_record(self.handle, stream.handle if stream else None), "event_record")

    def sync(self):
        check(lib().libxstream_event_sync(self.handle), "event_sync")

    def elapsed_ms(self, end):
        ms = ctypes.c_double(0)
        check(lib().libxstream_event_elapsed(self.handle, end.handle, ctypes.byref(ms)), "event_elapsed")
        return ms.value

    def close(self):
        if self.handle:
            check(lib().libxstream_event_destroy(self.handle))
            self.handle = None


class DeviceBuffer:
    """Device allocation (libxstream_mem_allocate) with numpy transfers."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().libxstream_mem_allocate(ctypes.byref(p), self.nbytes), "mem_allocate(%d)" % self.nbytes)
        self.ptr = p

    @classmethod
    def from_array(cls, array, stream=None):
        array = np.ascontiguousarray(array)
        buf = cls(max(array.nbytes, 1))
        buf.upload(array, stream)
        return buf

    def upload(self, array, stream=None):
        array = np.ascontiguousarray(array)
        assert array.nbytes <= self.nbytes
        check(lib().libxstream_mem_copy_h2d(array.ctypes.data, self.ptr, array.nbytes,
                                            stream.handle if stream else None), "copy_h2d")
        if stream is not None:
            stream.sync()

    def download(self, array, stream=None):
        assert array.flags.c_contiguous and array.nbytes <= self.nbytes
        check(lib().libxstream_mem_copy_d2h(self.ptr, array.ctypes.data, array.nbytes,
                                            stream.handle if stream else None), "copy_d2h")
        if stream is not None:
            stream.sync()
        return array

    def zero(self, stream=None, offset=0, nbytes=None):
        n = self.nbytes - offset if nbytes is None else nbytes
        check(lib().libxstream_mem_zero(self.ptr, offset, n, stream.handle if stream else None), "mem_zero")

    def free(self):
        if self.ptr:
            check(lib().libxstream_mem_deallocate(self.ptr), "mem_deallocate")
            self.ptr = None


def process(stack_dev, stack_size, tid, a_dev, b_dev, c_dev, m, n, k, stream=None,
            max_kernel_dim=MAX_KERNEL_DIM, def_mnk=1, host_stack=None):
    """libsmm_acc_process; returns its code (0, or -1 for host fallback); raises on other errors."""
    code = lib().libsmm_acc_process(host_stack, stack_dev.ptr if stack_dev else None, stack_size, tid,
                                    a_dev.ptr if a_dev else None, b_dev.ptr if b_dev else None,
                                    c_dev.ptr if c_dev else None, m, n, k, max_kernel_dim, def_mnk,
                                    stream.handle if stream else None, None)
    if code not in (0, -1):
        check(code, "libsmm_acc_process")
    return code


def process_config(config, stack_dev, stack_size, tid, a_dev, b_dev, c_dev, m, n, k, stream=None):
    return check(lib().libsmm_acc_process_config(ctypes.byref(config), stack_dev.ptr, stack_size, tid, a_dev.ptr,
                                                 b_dev.ptr, c_dev.ptr, m, n, k,
                                                 stream.handle if stream else None), "libsmm_acc_process_config")


def transpose(trs_dev, offset, stack_size, data_dev, tid, m, n, stream=None, max_kernel_dim=MAX_KERNEL_DIM):
    return check(lib().libsmm_acc_transpose(trs_dev.ptr if trs_dev else None, offset, stack_size,
                                            data_dev.ptr if data_dev else None, tid, m, n, max_kernel_dim,
                                            stream.handle if stream else None), "libsmm_acc_transpose")


def config_get(tid, m, n, k):
    cfg = SmmConfig()
    src = ctypes.c_int(0)
    check(lib().libsmm_acc_config_get(tid, m, n, k, ctypes.byref(cfg), ctypes.byref(src)), "config_get")
    return cfg, src.value


def config_set(tid, m, n, k, cfg):
    check(lib().libsmm_acc_config_set(tid, m, n, k, ctypes.byref(cfg)), "config_set")


def config_valid(tid, m, n, k, cfg):
    return lib().libsmm_acc_config_valid(tid, m, n, k, ctypes.byref(cfg)) == 0


def kernel_variants(tid):
    """Compiled runtime-shape stack kernels for a datatype: (variant, tm, tn, nt, kt) tuples."""
    cap = 256
    buf = (ctypes.c_int * (5 * cap))()
    count = lib().libsmm_acc_kernel_variants(tid, buf, cap)
    return [tuple(buf[5 * i + j] for j in range(5)) for i in range(min(count, cap))]


def profile_enable(on=True):
    lib().libxstream_profile_enable(1 if on else 0)


def profile_reset():
    lib().libxstream_profile_reset()


def profile_get(kernel):
    cnt = ctypes.c_long(0)
    ms, fl, by = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0)
    check(lib().libxstream_profile_get(kernel.encode(), ctypes.byref(cnt), ctypes.byref(ms), ctypes.byref(fl),
                                       ctypes.byref(by)), "profile_get")
    return {"count": cnt.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}


class Batch:
    """Several stacks of different shapes in one persistent launch (libsmm_acc_batch_*)."""

    def __init__(self, tid, stacks):
        """stacks: list of (stack DeviceBuffer, stack_size, m, n, k, a DeviceBuffer, b DeviceBuffer, c DeviceBuffer)."""
        arr = (BatchStack * len(stacks))()
        for x, (st, size, m, n, k, a, b, c) in zip(arr, stacks):
            x.dev_param_stack = st.ptr.value if st and st.ptr else None
            x.stack_size, x.m, x.n, x.k = size, m, n, k
            x.dev_a_data, x.dev_b_data, x.dev_c_data = a.ptr.value, b.ptr.value, c.ptr.value
        h = ctypes.c_void_p()
        check(lib().libsmm_acc_batch_create(tid, arr, len(stacks), ctypes.byref(h)), "libsmm_acc_batch_create")
        self.handle = h

    def process(self, stream=None):
        check(lib().libsmm_acc_batch_process(self.handle, stream.handle if stream else None), "libsmm_acc_batch_process")

    def replay(self, stream=None):
        check(lib().libsmm_acc_batch_replay(self.handle, stream.handle if stream else None), "libsmm_acc_batch_replay")

    def info(self):
        nt, fl, by = ctypes.c_int(0), ctypes.c_double(0), ctypes.c_double(0)
        check(lib().libsmm_acc_batch_info(self.handle, ctypes.byref(nt), ctypes.byref(fl), ctypes.byref(by)))
        return {"ntasks": nt.value, "flops": fl.value, "bytes": by.value}

    def close(self):
        if self.handle:
            check(lib().libsmm_acc_batch_destroy(self.handle), "libsmm_acc_batch_destroy")
            self.handle = None
