"""The DBM launch hook of libsmm_acc_process (samples/smm/smm_acc.c:579-585, smm_kernel.c:74-76).

With a launch function registered, every real_8 stack of one shape (def_mnk != 0) goes to it with
alpha 1 and the shape packed as m | n << 8 | k << 16; its result is libsmm_acc_process's result. No
device work happens on this path, so these tests run without a GPU (the pointers are never touched).
"""
import ctypes

import pytest

import libxstream_amd as lx


@pytest.fixture
def hook():
    calls = []

    def launch(stream, alpha, ntasks, param_format, params_host, params, a, b, c):
        calls.append((stream, alpha, ntasks, param_format, params_host, params, a, b, c))
        return 7 if ntasks == 13 else 0

    fn = lx.DBM_LAUNCH_FN(launch)
    lx.set_dbm_launch_fn(fn)
    try:
        yield calls
    finally:
        lx.set_dbm_launch_fn(None)


def _process(tid, S, m, n, k, def_mnk=1, host=0x1000, dev=0x2000, mkd=lx.MAX_KERNEL_DIM):
    L = lx.lib()
    return L.libsmm_acc_process(ctypes.c_void_p(host), ctypes.c_void_p(dev), S, tid, ctypes.c_void_p(0xA000),
                                ctypes.c_void_p(0xB000), ctypes.c_void_p(0xC000), m, n, k, mkd,
                                def_mnk, ctypes.c_void_p(0x5000), None)


def test_pack_format():
    assert lx.pack_format(23, 23, 23) == 23 | 23 << 8 | 23 << 16
    assert lx.pack_format(4, 9, 22) == 0x160904


def test_routes_fp64(hook):
    assert _process(lx.REAL_8, 100, 4, 9, 22) == 0
    assert len(hook) == 1
    stream, alpha, ntasks, pf, ph, pd, a, b, c = hook[0]
    assert (stream, alpha, ntasks) == (0x5000, 1.0, 100)
    assert pf == lx.pack_format(4, 9, 22)
    assert (ph, pd, a, b, c) == (0x1000, 0x2000, 0xA000, 0xB000, 0xC000)


def test_result_propagates(hook):
    assert _process(lx.REAL_8, 13, 5, 5, 5) == 7


def test_unsuitable_shapes_still_routed(hook):
    """The reference routes before its suitability check: a stack larger than max_kernel_dim goes to DBM."""
    assert _process(lx.REAL_8, 10, 100, 100, 100) == 0
    assert hook[-1][3] == lx.pack_format(100, 100, 100)


def test_not_routed(hook):
    # FP32 and inhomogeneous stacks take the LIBSMM path: an empty stack succeeds, an inhomogeneous one
    # is a host fallback (-1); neither touches the device.
    assert _process(lx.REAL_4, 0, 23, 23, 23) == 0
    assert _process(lx.REAL_8, 100, 23, 23, 23, def_mnk=0) == -1
    # extents of 128 and above do not fit the packed format: the stack is not handed over (here it is then
    # unsuitable for LIBSMM too, m n > max_kernel_dim^2, so the result is the host fallback)
    assert _process(lx.REAL_8, 100, 128, 1, 1, mkd=11) == -1
    assert _process(lx.REAL_8, 100, 1, 1, 128, mkd=0) == -1
    assert hook == []


def test_unregister():
    calls = []
    fn = lx.DBM_LAUNCH_FN(lambda *args: calls.append(args) or 0)
    lx.set_dbm_launch_fn(fn)
    lx.set_dbm_launch_fn(None)
    assert _process(lx.REAL_8, 100, 23, 23, 23, def_mnk=0) == -1
    assert _process(lx.REAL_8, 0, 23, 23, 23) == 0
    assert calls == []
