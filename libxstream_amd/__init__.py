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
        ("s", ctypes.c_int),
        ("split", ctypes.c_int),
        ("bs_short", ctypes.c_int),
        ("split_short", ctypes.c_int),
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


class BatchParams(ctypes.Structure):
    """libsmm_acc_batch_params_t (include/libxstream_amd.h): the batch's schedule."""

    _fields_ = [("chunk_elements", ctypes.c_int), ("blocks_per_cu", ctypes.c_int), ("per_stack", ctypes.c_int)]

    def as_dict(self):
        return {"chunk_elements": self.chunk_elements, "blocks_per_cu": self.blocks_per_cu,
                "per_stack": self.per_stack}


class OzakiContext(ctypes.Structure):
    """include/ozaki.h: ozaki_context_t (caller-allocated, filled by ozaki_init)."""
    _fields_ = [(name, ctypes.c_int) for name in ("use_double", "kind", "ndecomp", "nslices", "nprimes", "ozflags",
                                                  "oztrim", "verbosity", "tm", "tn", "maxk", "flush")] + [
        ("impl", ctypes.c_void_p)]


# samples/smm/smm_acc_opencl.h:115-116: (stream, alpha, ntasks, param_format, params_host, params, a, b, c)
DBM_LAUNCH_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)
PFORMAT = 8  # bits per extent of a shape packed into param_format


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
        "libxstream_stream_wait_event": (i, [vp, vp]),
        "libxstream_event_create": (i, [pvp]),
        "libxstream_event_destroy": (i, [vp]),
        "libxstream_event_record": (i, [vp, vp]),
        "libxstream_event_sync": (i, [vp]),
        "libxstream_event_elapsed": (i, [vp, vp, ctypes.POINTER(ctypes.c_double)]),
        "libxstream_mem_allocate": (i, [pvp, sz]),
        "libxstream_mem_host_allocate": (i, [pvp, sz, vp]),
        "libxstream_mem_host_deallocate": (i, [vp, vp]),
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
        "libxstream_profile_stats": (i, [ctypes.c_char_p] + [ctypes.POINTER(ctypes.c_double)] * 4 +
                                     [ctypes.POINTER(ctypes.c_long)]),
        "libxstream_profile_print": (i, [vp]),
        "libxstream_profile_mem_enable": (i, [i]),
        "libsmm_acc_init": (i, []),
        "libsmm_acc_finalize": (i, []),
        "libsmm_acc_is_thread_safe": (i, []),
        "libsmm_acc_process_suitable": (i, [i, i, i, i, i, i, i]),
        "libsmm_acc_process": (i, [vp, vp, i, i, vp, vp, vp, i, i, i, i, i, vp, vp]),
        "opencl_libsmm_acc_process": (i, [vp, vp, i, i, vp, vp, vp, i, i, i, i, i, vp, vp, i, vp]),
        "opencl_libsmm_acc_set_dbm_launch_fn": (None, [DBM_LAUNCH_FN]),
        "libsmm_acc_transpose": (i, [vp, i, i, vp, i, i, i, i, vp]),
        "libsmm_acc_process_config": (i, [ctypes.POINTER(SmmConfig), vp, i, i, vp, vp, vp, i, i, i, vp]),
        "libsmm_acc_config_get": (i, [i, i, i, i, ctypes.POINTER(SmmConfig), pi]),
        "libsmm_acc_config_heuristic": (i, [i, i, i, i, ctypes.POINTER(SmmConfig)]),
        "libsmm_acc_config_set": (i, [i, i, i, i, ctypes.POINTER(SmmConfig)]),
        "libsmm_acc_config_valid": (i, [i, i, i, i, ctypes.POINTER(SmmConfig)]),
        "libsmm_acc_kernel_variants": (i, [i, pi, i]),
        "libsmm_acc_kernel_candidates": (i, [i, i, i, i, pi, i]),
        "libsmm_acc_params_load": (i, [ctypes.c_char_p]),
        "libsmm_acc_params_load_device": (i, [ctypes.c_char_p, ctypes.c_char_p]),
        "libsmm_acc_params_save": (i, [ctypes.c_char_p, ctypes.c_char_p]),
        "libsmm_acc_read_ref_params": (i, [ctypes.c_char_p, i, i, ctypes.c_char_p, sz, pi, pi, ctypes.POINTER(ctypes.c_double)]),
        "libsmm_acc_batch_create": (i, [i, ctypes.POINTER(BatchStack), i, pvp]),
        "libsmm_acc_batch_create_ex": (i, [i, ctypes.POINTER(BatchStack), i, ctypes.POINTER(BatchParams), pvp]),
        "libsmm_acc_batch_tune": (i, [i, ctypes.POINTER(BatchStack), i, ctypes.POINTER(BatchParams),
                                      ctypes.POINTER(ctypes.c_double)]),
        "libsmm_acc_batch_process": (i, [vp, vp]),
        "libsmm_acc_batch_replay": (i, [vp, vp]),
        "libsmm_acc_batch_info": (i, [vp, pi, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "libsmm_acc_batch_destroy": (i, [vp]),
        "c_dbcsr_acc_init": (i, []),
        "c_dbcsr_acc_finalize": (i, []),
        "c_dbcsr_acc_get_ndevices": (i, [pi]),
        "c_dbcsr_acc_set_active_device": (i, [i]),
        "c_calculate_norms": (i, [vp, i, vp, vp, vp, vp]),
        "libsmm_acc_mlp_create": (i, [pvp, i, i, i, ctypes.c_ulonglong]),
        "libsmm_acc_mlp_destroy": (i, [vp]),
        "libsmm_acc_mlp_encoding": (i, [vp, vp, i]),
        "libsmm_acc_mlp_features": (i, [i, i, i, i, ctypes.POINTER(SmmConfig), vp]),
        "libsmm_acc_mlp_info": (i, [vp, pi, pi, vp, i]),
        "libsmm_acc_mlp_params": (i, [vp, i, vp, i]),
        "libsmm_acc_mlp_set_params": (i, [vp, i, vp, i]),
        "libsmm_acc_mlp_gradient": (i, [vp, i, vp, vp, vp, i, vp, ctypes.POINTER(ctypes.c_double)]),
        "libsmm_acc_mlp_train": (i, [vp, vp, vp, vp, i, i, ctypes.c_double, ctypes.c_double,
                                     ctypes.POINTER(ctypes.c_double)]),
        "libsmm_acc_mlp_forward": (i, [vp, vp, i, vp]),
        "libsmm_acc_mlp_forward_host": (i, [vp, vp, i, vp]),
        "libsmm_acc_mlp_save": (i, [vp, ctypes.c_char_p]),
        "libsmm_acc_mlp_load": (i, [pvp, ctypes.c_char_p]),
        "libsmm_acc_mlp_predict_config": (i, [vp, i, i, i, i, ctypes.POINTER(SmmConfig),
                                              ctypes.POINTER(ctypes.c_double)]),
        "libsmm_acc_mlp_predict_topk": (i, [vp, i, i, i, i, i, ctypes.POINTER(SmmConfig),
                                            ctypes.POINTER(ctypes.c_double), pi]),
        "libsmm_acc_learned_set": (i, [vp, ctypes.c_double]),
        "ozaki_init": (i, [ctypes.POINTER(OzakiContext), i, i, i, i, i, i, i, i, i, i]),
        "ozaki_destroy": (None, [ctypes.POINTER(OzakiContext)]),
        "ozaki_scratch_size": (sz, [ctypes.POINTER(OzakiContext), ctypes.c_char, ctypes.c_char, i, i, i, i, i, i]),
        "ozaki_scratch_set": (i, [ctypes.POINTER(OzakiContext), vp, sz]),
        "ozaki_scratch_get": (vp, [ctypes.POINTER(OzakiContext), psz]),
        "ozaki_gemm": (i, [ctypes.POINTER(OzakiContext), vp, ctypes.c_char, ctypes.c_char, i, i, i, ctypes.c_double,
                           vp, i, vp, i, ctypes.c_double, vp, i, i]),
        "ozaki_gemm_noalloc": (i, [ctypes.POINTER(OzakiContext), vp, ctypes.c_char, ctypes.c_char, i, i, i,
                                   ctypes.c_double, vp, i, vp, i, ctypes.c_double, vp, i]),
        "ozaki_gemm_complex": (i, [ctypes.POINTER(OzakiContext), vp, ctypes.c_char, ctypes.c_char, i, i, i,
                                   ctypes.POINTER(ctypes.c_double), vp, i, vp, i, ctypes.POINTER(ctypes.c_double), vp,
                                   i]),
        "ozaki_invalidate_cache": (None, [ctypes.POINTER(OzakiContext), vp, vp]),
        "ozaki_stack": (i, [ctypes.POINTER(OzakiContext), vp, vp, i, i, i, i, vp, i, vp, i, vp]),
        "ozaki_dgemm_native": (i, [vp, ctypes.c_char, ctypes.c_char, i, i, i, ctypes.c_double, vp, i, vp, i,
                                   ctypes.c_double, vp, i]),
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

    def wait(self, event):
        """Later work on this stream waits for event (libxstream_stream_wait_event)."""
        check(lib().libxstream_stream_wait_event(self.handle, event.handle), "stream_wait_event")

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


class HostBuffer:
    """Pinned host memory (libxstream_mem_host_allocate) seen as a numpy array of dtype and count."""

    def __init__(self, dtype, count, stream=None):
        self.stream = stream
        self.nbytes = int(np.dtype(dtype).itemsize * count)
        p = ctypes.c_void_p()
        check(lib().libxstream_mem_host_allocate(ctypes.byref(p), max(self.nbytes, 1),
                                                 stream.handle if stream else None), "mem_host_allocate")
        self.ptr = p.value
        self.array = np.ctypeslib.as_array((ctypes.c_char * max(self.nbytes, 1)).from_address(self.ptr))
        self.array = self.array.view(np.uint8)[:self.nbytes].view(dtype)

    def copy_to(self, dev, stream, nbytes=None):
        """Asynchronous host-to-device copy on stream (pinned source, libxstream_mem_copy_h2d)."""
        n = self.nbytes if nbytes is None else int(nbytes)
        assert n <= self.nbytes and n <= dev.nbytes
        check(lib().libxstream_mem_copy_h2d(self.ptr, dev.ptr, n, stream.handle), "copy_h2d")

    def free(self):
        if self.ptr:
            self.array = None
            check(lib().libxstream_mem_host_deallocate(self.ptr, self.stream.handle if self.stream else None),
                  "mem_host_deallocate")
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


def pack_format(m, n, k):
    """The param_format libsmm_acc_process hands a DBM launch function (smm_acc.c:583)."""
    return m | (n << PFORMAT) | (k << (2 * PFORMAT))


def process_format(stack_dev, stack_size, tid, a_dev, b_dev, c_dev, m, n, k, param_format, stream=None,
                   event=None, max_kernel_dim=MAX_KERNEL_DIM, def_mnk=1):
    """opencl_libsmm_acc_process: param_format 0 reads (a, b, c) 1-based triplets, any other value DBM task
    records (m, n, k, a, b, c) with 0-based offsets; returns 0, or -1 for host fallback."""
    code = lib().opencl_libsmm_acc_process(None, stack_dev.ptr if stack_dev else None, stack_size, tid,
                                           a_dev.ptr if a_dev else None, b_dev.ptr if b_dev else None,
                                           c_dev.ptr if c_dev else None, m, n, k, max_kernel_dim, def_mnk,
                                           stream.handle if stream else None, None, param_format,
                                           event.handle if event else None)
    if code not in (0, -1):
        check(code, "opencl_libsmm_acc_process")
    return code


def set_dbm_launch_fn(fn):
    """Register a DBM_LAUNCH_FN (keep a reference to it while registered), or None to unregister."""
    lib().opencl_libsmm_acc_set_dbm_launch_fn(fn if fn is not None else DBM_LAUNCH_FN())


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


def default_config(tid, m, n, k):
    """the built-in heuristic's configuration (no registry, no predictor)"""
    cfg = SmmConfig()
    check(lib().libsmm_acc_config_heuristic(tid, m, n, k, ctypes.byref(cfg)), "config_heuristic")
    return cfg


def config_valid(tid, m, n, k, cfg):
    return lib().libsmm_acc_config_valid(tid, m, n, k, ctypes.byref(cfg)) == 0


def kernel_variants(tid):
    """Compiled runtime-shape stack kernels for a datatype: (variant, tm, tn, nt, kt) tuples."""
    cap = 256
    buf = (ctypes.c_int * (5 * cap))()
    count = lib().libsmm_acc_kernel_variants(tid, buf, cap)
    return [tuple(buf[5 * i + j] for j in range(5)) for i in range(min(count, cap))]


def kernel_candidates(tid, m, n, k):
    """Every compiled stack kernel usable for (m, n, k), runtime-shape or compiled for exactly this shape:
    (variant, tm, tn, nt, kt) tuples."""
    cap = 256
    buf = (ctypes.c_int * (5 * cap))()
    count = lib().libsmm_acc_kernel_candidates(tid, m, n, k, buf, cap)
    return [tuple(buf[5 * i + j] for j in range(5)) for i in range(min(count, cap))]


def profile_enable(on=True):
    lib().libxstream_profile_enable(1 if on else 0)


def profile_mem_enable(on=True):
    check(lib().libxstream_profile_mem_enable(1 if on else 0), "libxstream_profile_mem_enable")


def profile_stats(name):
    """{min_ms, median_ms, max_ms, median_rate, discarded} of one profiled kernel or transfer kind"""
    vals = [ctypes.c_double(0) for _ in range(4)]
    disc = ctypes.c_long(0)
    rc = lib().libxstream_profile_stats(os.fsencode(name), *[ctypes.byref(v) for v in vals], ctypes.byref(disc))
    if rc != 0:
        lib().libxstream_error_consume(None, 0)
        return None
    return {"min_ms": vals[0].value, "median_ms": vals[1].value, "max_ms": vals[2].value,
            "median_rate": vals[3].value, "discarded": disc.value}


def profile_print():
    """histograms of every profiled kernel and transfer kind, printed to stderr"""
    check(lib().libxstream_profile_print(None), "libxstream_profile_print")


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

    @staticmethod
    def _stacks(stacks):
        arr = (BatchStack * len(stacks))()
        for x, (st, size, m, n, k, a, b, c) in zip(arr, stacks):
            x.dev_param_stack = st.ptr.value if st and st.ptr else None
            x.stack_size, x.m, x.n, x.k = size, m, n, k
            x.dev_a_data, x.dev_b_data, x.dev_c_data = a.ptr.value, b.ptr.value, c.ptr.value
        return arr

    def __init__(self, tid, stacks, params=None):
        """stacks: list of (stack DeviceBuffer, stack_size, m, n, k, a DeviceBuffer, b DeviceBuffer, c DeviceBuffer);
        params: a BatchParams schedule (None: the default)."""
        arr = self._stacks(stacks)
        h = ctypes.c_void_p()
        check(lib().libsmm_acc_batch_create_ex(tid, arr, len(stacks), ctypes.byref(params) if params else None,
                                               ctypes.byref(h)), "libsmm_acc_batch_create_ex")
        self.handle = h

    @staticmethod
    def tune(tid, stacks):
        """libsmm_acc_batch_tune: (BatchParams of the fastest schedule, its ms per launch)"""
        arr = Batch._stacks(stacks)
        best, ms = BatchParams(), ctypes.c_double(0)
        check(lib().libsmm_acc_batch_tune(tid, arr, len(stacks), ctypes.byref(best), ctypes.byref(ms)),
              "libsmm_acc_batch_tune")
        return best, ms.value

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


class Ozaki:
    """FP64/FP32 GEMM emulated with int8 matrix cores (include/ozaki.h). kind 1: Ozaki scheme I (nslices
    digit slices, trim drops levels); kind 2: scheme II (CRT, at most nprimes moduli, trim drops 2 trim
    mantissa bits); kind 3: per call, whichever the cost model picks (nprimes for scheme II, default
    slices for scheme I), as ozaki_init resolves ndecomp in the reference."""

    def __init__(self, use_double=True, nslices=0, trim=0, flags=3, maxk=0, verbosity=0, kind=1, nprimes=0):
        self.ctx = OzakiContext()
        ndecomp = nslices if kind == 1 else nprimes
        check(lib().ozaki_init(ctypes.byref(self.ctx), 0, 0, 1 if use_double else 0, kind, verbosity, ndecomp, flags,
                               trim, 0, maxk), "ozaki_init")

    def gemm(self, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None, dev=True):
        """C = alpha op(A) op(B) + beta C (column-major). dev: a, b, c are DeviceBuffers (else numpy arrays)."""
        ptr = (lambda x: x.ptr) if dev else (lambda x: x.ctypes.data)
        check(lib().ozaki_gemm(ctypes.byref(self.ctx), stream.handle if stream else None, transa.encode(),
                               transb.encode(), m, n, k, alpha, ptr(a), lda, ptr(b), ldb, beta, ptr(c), ldc,
                               1 if dev else 0), "ozaki_gemm")

    def gemm_noalloc(self, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None):
        """gemm on DeviceBuffers with the library's static device scratch (no allocation; include/ozaki.h:
        ozaki_gemm_noalloc)."""
        check(lib().ozaki_gemm_noalloc(ctypes.byref(self.ctx), stream.handle if stream else None, transa.encode(),
                                       transb.encode(), m, n, k, alpha, a.ptr, lda, b.ptr, ldb, beta, c.ptr, ldc),
              "ozaki_gemm_noalloc")

    def gemm_complex(self, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None):
        """Complex C = alpha op(A) op(B) + beta C on DeviceBuffers of interleaved (re, im) values."""
        al = (ctypes.c_double * 2)(alpha.real, alpha.imag)
        be = (ctypes.c_double * 2)(beta.real, beta.imag)
        check(lib().ozaki_gemm_complex(ctypes.byref(self.ctx), stream.handle if stream else None, transa.encode(),
                                       transb.encode(), m, n, k, al, a.ptr, lda, b.ptr, ldb, be, c.ptr, ldc),
              "ozaki_gemm_complex")

    def stack(self, stack, stack_size, m, n, k, a, na, b, nb, c, stream=None):
        """C[c] += A[a] B[b] over a parameter stack (DeviceBuffers; B blocks in the kernel layout), each
        product by scheme I with this context's slices (include/ozaki.h: ozaki_stack)."""
        check(lib().ozaki_stack(ctypes.byref(self.ctx), stream.handle if stream else None, stack.ptr, stack_size, m, n,
                                k, a.ptr, na, b.ptr, nb, c.ptr), "ozaki_stack")

    def scratch_size(self, transa, transb, m, n, k, lda, ldb, ldc):
        return lib().ozaki_scratch_size(ctypes.byref(self.ctx), transa.encode(), transb.encode(), m, n, k, lda, ldb,
                                        ldc)

    def scratch_set(self, buf):
        check(lib().ozaki_scratch_set(ctypes.byref(self.ctx), buf.ptr if buf else None, buf.nbytes if buf else 0),
              "ozaki_scratch_set")

    def close(self):
        if self.ctx.impl:
            lib().ozaki_destroy(ctypes.byref(self.ctx))


def dgemm_native(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None):
    """Native FP64 matrix-core GEMM on DeviceBuffers (the Ozaki comparator)."""
    check(lib().ozaki_dgemm_native(stream.handle if stream else None, transa.encode(), transb.encode(), m, n, k, alpha,
                                   a.ptr, lda, b.ptr, ldb, beta, c.ptr, ldc), "ozaki_dgemm_native")


def mlp_encoding():
    """(segment starts of the configuration one-hot inputs incl. the end, class values, number of inputs)"""
    seg = np.zeros(64, dtype=np.int32)
    vals = np.zeros(64, dtype=np.int32)
    nin = lib().libsmm_acc_mlp_encoding(seg.ctypes.data, vals.ctypes.data, 64)
    seg = seg[:8].copy()
    return seg, vals[:seg[-1]].copy(), nin


def mlp_features(tid, m, n, k, cfg=None):
    """network inputs of one (shape, configuration) row (cfg None: the shape only)"""
    nin = lib().libsmm_acc_mlp_encoding(None, None, 0)
    f = np.zeros(nin, dtype=np.float32)
    check(lib().libsmm_acc_mlp_features(tid, m, n, k, ctypes.byref(cfg) if cfg is not None else None, f.ctypes.data),
          "libsmm_acc_mlp_features")
    return f


class Mlp:
    """The learned predictor's network ensemble (libsmm_acc_mlp_*): training and inference on the GPU."""

    def __init__(self, nnets=4, hidden=64, nhidden=3, seed=1, path=None):
        h = ctypes.c_void_p()
        if path is not None:
            check(lib().libsmm_acc_mlp_load(ctypes.byref(h), os.fsencode(path)), "libsmm_acc_mlp_load")
        else:
            check(lib().libsmm_acc_mlp_create(ctypes.byref(h), nnets, hidden, nhidden, seed), "libsmm_acc_mlp_create")
        self.handle = h

    def info(self):
        nn, nl = ctypes.c_int(0), ctypes.c_int(0)
        w = np.zeros(16, dtype=np.int32)
        npar = lib().libsmm_acc_mlp_info(self.handle, ctypes.byref(nn), ctypes.byref(nl), w.ctypes.data, 16)
        return {"nnets": nn.value, "nlayers": nl.value, "width": w[:nl.value + 1].tolist(), "nparams": npar}

    def params(self, net=0):
        n = lib().libsmm_acc_mlp_params(self.handle, net, None, 0)
        p = np.zeros(n, dtype=np.float32)
        lib().libsmm_acc_mlp_params(self.handle, net, p.ctypes.data, n)
        return p

    def set_params(self, p, net=0):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().libsmm_acc_mlp_set_params(self.handle, net, p.ctypes.data, p.size), "libsmm_acc_mlp_set_params")

    @staticmethod
    def _w(out_weights):
        return None if out_weights is None else np.ascontiguousarray(out_weights, dtype=np.float32)

    def gradient(self, x, t, out_weights=None, net=0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = self._w(out_weights)
        g = np.zeros(lib().libsmm_acc_mlp_params(self.handle, net, None, 0), dtype=np.float32)
        loss = ctypes.c_double(0)
        check(lib().libsmm_acc_mlp_gradient(self.handle, net, x.ctypes.data, t.ctypes.data,
                                            None if w is None else w.ctypes.data, x.shape[0], g.ctypes.data,
                                            ctypes.byref(loss)), "libsmm_acc_mlp_gradient")
        return loss.value, g

    def train(self, x, t, epochs, lr=3e-3, weight_decay=1e-5, out_weights=None):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = self._w(out_weights)
        loss = ctypes.c_double(0)
        check(lib().libsmm_acc_mlp_train(self.handle, x.ctypes.data, t.ctypes.data, None if w is None else w.ctypes.data,
                                         x.shape[0], epochs, lr, weight_decay, ctypes.byref(loss)),
              "libsmm_acc_mlp_train")
        return loss.value

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros((x.shape[0], 2), dtype=np.float32)
        check(lib().libsmm_acc_mlp_forward(self.handle, x.ctypes.data, x.shape[0], out.ctypes.data),
              "libsmm_acc_mlp_forward")
        return out

    def forward_host(self, x):
        """the same inference on the host (what prediction uses)"""
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros((x.shape[0], 2), dtype=np.float32)
        check(lib().libsmm_acc_mlp_forward_host(self.handle, x.ctypes.data, x.shape[0], out.ctypes.data),
              "libsmm_acc_mlp_forward_host")
        return out

    def predict_config(self, tid, m, n, k):
        cfg = SmmConfig()
        conf = ctypes.c_double(0)
        check(lib().libsmm_acc_mlp_predict_config(self.handle, tid, m, n, k, ctypes.byref(cfg), ctypes.byref(conf)),
              "libsmm_acc_mlp_predict_config")
        return cfg, conf.value

    def predict_topk(self, tid, m, n, k, topk):
        """the ensemble's topk configurations for the shape, best first, as [(SmmConfig, score)]"""
        cfgs = (SmmConfig * topk)()
        scores = (ctypes.c_double * topk)()
        count = ctypes.c_int(0)
        check(lib().libsmm_acc_mlp_predict_topk(self.handle, tid, m, n, k, topk, cfgs, scores, ctypes.byref(count)),
              "libsmm_acc_mlp_predict_topk")
        return [(cfgs[i], scores[i]) for i in range(count.value)]

    def save(self, path):
        check(lib().libsmm_acc_mlp_save(self.handle, os.fsencode(path)), "libsmm_acc_mlp_save")

    def install(self, min_confidence=0.0):
        check(lib().libsmm_acc_learned_set(self.handle, min_confidence), "libsmm_acc_learned_set")

    def close(self):
        if self.handle:
            lib().libsmm_acc_mlp_destroy(self.handle)
            self.handle = None


def learned_clear():
    check(lib().libsmm_acc_learned_set(None, 0.0), "libsmm_acc_learned_set")
