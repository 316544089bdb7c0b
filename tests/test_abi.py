"""C-ABI boundary and host logic (CPU only: no call here needs a GPU).

Covers: every function declared in include/*.h is exported by libxstream_amd.so; the
c_dbcsr_acc_* / libxstream_* subset of the reference ABI (.abi.txt) is present; the suitability
gate and the -1 host fallback (samples/smm/smm_acc.c:527-573, smm_kernel.c:524-526); transpose
no-ops (smm_trans.c:54-68); the tuned-parameter CSV reader on rows of the reference's own CSVs
(samples/smm/params/*.csv, excerpt under tests/golden/); the MI355X CSV round trip; configuration
selection and validation.
"""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import libxstream_amd as lx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# .abi.txt of the reference, minus the OpenCL-specific libxstream_opencl_* and the CP2K offload* shims
REFERENCE_ABI = """
c_dbcsr_acc_clear_errors c_dbcsr_acc_dev_mem_allocate c_dbcsr_acc_dev_mem_deallocate c_dbcsr_acc_dev_mem_info
c_dbcsr_acc_dev_mem_set_ptr c_dbcsr_acc_device_synchronize c_dbcsr_acc_event_create c_dbcsr_acc_event_destroy
c_dbcsr_acc_event_query c_dbcsr_acc_event_record c_dbcsr_acc_event_synchronize c_dbcsr_acc_finalize
c_dbcsr_acc_get_ndevices c_dbcsr_acc_host_mem_allocate c_dbcsr_acc_host_mem_deallocate c_dbcsr_acc_init
c_dbcsr_acc_memcpy_d2d c_dbcsr_acc_memcpy_d2h c_dbcsr_acc_memcpy_h2d c_dbcsr_acc_memset_zero
c_dbcsr_acc_set_active_device c_dbcsr_acc_stream_create c_dbcsr_acc_stream_destroy
c_dbcsr_acc_stream_priority_range c_dbcsr_acc_stream_sync c_dbcsr_acc_stream_wait_event
libxstream_device_count libxstream_device_set_active libxstream_device_sync libxstream_event_create
libxstream_event_destroy libxstream_event_query libxstream_event_record libxstream_event_sync libxstream_finalize
libxstream_init libxstream_init_config libxstream_init_config_default libxstream_mem_allocate libxstream_mem_copy_d2d
libxstream_mem_copy_d2h libxstream_mem_copy_h2d libxstream_mem_deallocate libxstream_mem_dev_allocate_hint
libxstream_mem_dev_deallocate_hint libxstream_mem_host_allocate libxstream_mem_host_deallocate libxstream_mem_info
libxstream_mem_offset libxstream_mem_zero libxstream_stream_create libxstream_stream_destroy
libxstream_stream_priority_range libxstream_stream_set_profiling libxstream_stream_sync libxstream_stream_wait_event
libsmm_acc_init libsmm_acc_finalize libsmm_acc_is_thread_safe libsmm_acc_process libsmm_acc_transpose
c_calculate_norms
""".split()


def declared_functions():
    names = set()
    for path in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = "\n".join(l for l in open(path).read().splitlines() if not l.lstrip().startswith("#"))
        for mo in re.finditer(r"LIBXSTREAM_API\s+[\w\s\*]+?\b(\w+)\s*\(", text):
            names.add(mo.group(1))
    return sorted(names)


def test_library_loads_and_exports_declared():
    L = lx.lib()
    names = declared_functions()
    assert len(names) >= 70
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_reference_abi_subset():
    L = lx.lib()
    missing = [n for n in REFERENCE_ABI if not hasattr(L, n)]
    assert not missing, missing
    assert set(REFERENCE_ABI) <= set(declared_functions())


def test_init_without_device():
    """Initialisation succeeds with zero devices (acc_bench.c:122-142 then reports no device)."""
    L = lx.lib()
    assert L.c_dbcsr_acc_init() == 0
    n = ctypes.c_int(-1)
    assert L.c_dbcsr_acc_get_ndevices(ctypes.byref(n)) == 0
    assert n.value >= 0
    assert L.libsmm_acc_init() == 0
    assert L.libsmm_acc_is_thread_safe() == 1
    assert L.libsmm_acc_finalize() == 0
    assert L.c_dbcsr_acc_finalize() == 0


@pytest.mark.parametrize("args,expect", [
    ((1, 3, 100, 23, 23, 23, 80), 1),
    ((1, 1, 100, 23, 23, 23, 80), 1),
    ((0, 3, 100, 23, 23, 23, 80), 0),   # inhomogeneous
    ((1, 7, 100, 23, 23, 23, 80), 0),   # complex_8
    ((1, 5, 100, 23, 23, 23, 80), 0),   # complex_4
    ((1, 3, 0, 23, 23, 23, 80), 0),     # empty stack
    ((1, 3, 100, 81, 80, 23, 80), 0),   # m*n > max_kernel_dim^2
    ((1, 3, 100, 80, 80, 500, 80), 1),  # k may exceed max_kernel_dim
    ((1, 3, 100, 0, 23, 23, 80), 0),
    ((1, 3, 100, 23, 23, 0, 80), 0),
])
def test_suitability(args, expect):
    assert lx.lib().libsmm_acc_process_suitable(*args) == expect


def test_process_fallback_without_device():
    L = lx.lib()
    # not suitable: -1 before any device work; empty stack: success
    assert L.libsmm_acc_process(None, None, 10, 3, None, None, None, 23, 23, 23, 80, 0, None, None) == -1
    assert L.libsmm_acc_process(None, None, 10, 7, None, None, None, 23, 23, 23, 80, 1, None, None) == -1
    assert L.libsmm_acc_process(None, None, 10, 3, None, None, None, 90, 90, 23, 80, 1, None, None) == -1
    assert L.libsmm_acc_process(None, None, 0, 3, None, None, None, 23, 23, 23, 80, 1, None, None) == 0
    assert L.c_calculate_norms(None, 0, None, None, None, None) == -1


def test_transpose_noops_without_device():
    L = lx.lib()
    assert L.libsmm_acc_transpose(None, 0, 0, None, 3, 23, 23, 80, None) == 0      # empty
    assert L.libsmm_acc_transpose(None, 0, 10, None, 7, 23, 23, 80, None) == 0     # complex
    assert L.libsmm_acc_transpose(None, 0, 10, None, 3, 1, 1, 80, None) == 0       # 1x1
    assert L.libsmm_acc_transpose(None, 0, 10, None, 3, 90, 90, 80, None) == 0     # too large
    assert L.libsmm_acc_transpose(None, 0, 10, None, 3, 0, 5, 80, None) != 0       # invalid shape


def test_mem_offset_and_priority_range():
    L = lx.lib()
    p = ctypes.c_void_p()
    L.libxstream_mem_offset.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.c_size_t]
    assert L.libxstream_mem_offset(ctypes.byref(p), 4096, 64) == 0 and p.value == 4160
    assert L.libxstream_mem_offset(ctypes.byref(p), None, 0) == 0 and not p.value
    assert L.libxstream_mem_offset(ctypes.byref(p), None, 8) != 0
    lo, hi = ctypes.c_int(7), ctypes.c_int(7)
    assert L.c_dbcsr_acc_stream_priority_range(ctypes.byref(lo), ctypes.byref(hi)) == 0
    assert lo.value == hi.value  # no priority range: every stream runs at the default priority
    assert L.libxstream_error_consume(None, 0) != 0  # the failed offset left an error ...
    assert L.libxstream_error_consume(None, 0) == 0  # ... which is consumed once


def read_ref(line, with_device=1, with_perf=1):
    key = (ctypes.c_int * 4)()
    params = (ctypes.c_int * 16)()
    perf = (ctypes.c_double * 2)()
    dev = ctypes.create_string_buffer(128)
    rc = lx.lib().libsmm_acc_read_ref_params(line.encode(), with_device, with_perf, dev, 128, key, params, perf)
    return rc, dev.value.decode(), list(key), list(params), list(perf)


def test_reference_csv_rows():
    rows = open(os.path.join(ROOT, "tests", "golden", "tune_multiply_excerpt.csv")).read().splitlines()
    assert rows[0].startswith("DEVICE;TYPEID;M;N;K;S;GFLOPS;BS;BM;BN;BK")
    parsed = [read_ref(r) for r in rows[1:]]
    assert all(p[0] == 0 for p in parsed)
    # samples/smm/params/tune_multiply_Mi250.csv:355 (23^3 on gfx90a)
    rc, dev, key, params, perf = read_ref("gfx90a [0x989f];3;23;23;23;30000;0;20;23;1;10;23;-2;-2;0;1;1;1;1;2;2;0;0")
    assert dev == "gfx90a [0x989f]" and key == [3, 23, 23, 23]
    assert params[:15] == [20, 23, 1, 10, 23, -2, -2, 0, 1, 1, 1, 1, 2, 2, 0] and perf == [30000.0, 0.0]
    # samples/smm/params/tune_multiply_V100.csv:117
    rc, dev, key, params, perf = read_ref(
        "Tesla V100-PCIE-16GB [0xb271];3;23;23;23;30000;2459.7;50;23;1;14;0;0;0;1;0;1;1;1;1;2;0;0")
    assert rc == 0 and perf[1] == 2459.7 and params[0] == 50
    # optional FLAGS column, and the parameter-string form without device/perf columns
    assert read_ref("3;5;5;5;8;5;1;1;1;0;0;0;0;0;0;0;0;0;0;1", 0, 0)[0] == 0
    # malformed rows
    assert read_ref("gfx90a;3;23;23")[0] != 0
    assert read_ref("gfx90a;4;23;23;23;30000;0;20;23;1;10;23;-2;-2;0;1;1;1;1;2;2;0;0")[0] != 0  # bad type
    assert read_ref("gfx90a;3;0;23;23;30000;0;20;23;1;10;23;-2;-2;0;1;1;1;1;2;2;0;0")[0] != 0  # m = 0


def test_params_load_reference_and_roundtrip(tmp_path, monkeypatch):
    L = lx.lib()
    monkeypatch.setenv("OPENCL_LIBSMM_SMM_PARAMS", "0")  # start from an empty registry (no shipped CSV)
    assert L.libsmm_acc_init() == 0
    try:
        n = L.libsmm_acc_params_load(os.path.join(ROOT, "tests", "golden", "tune_multiply_excerpt.csv").encode())
        assert n == 13
        cfg, src = lx.config_get(lx.REAL_8, 23, 23, 23)
        assert src == 1  # tuned
        assert cfg.bs == 50 and cfg.gflops == 2459.7  # fastest duplicate row wins (smm_acc.c:117-119)
        assert cfg.s == 30000  # the row's S: the stack size bs was tuned for (scaled down for shorter stacks)
        out = tmp_path / "mi355x.csv"
        rows = L.libsmm_acc_params_save(str(out).encode(), b"MI355X")
        assert rows == 5  # distinct shapes of the excerpt: 5^3, 13^3, 23^3, 32^3, 64^3
        text = out.read_text().splitlines()
        assert text[0] == "DEVICE;TYPEID;M;N;K;S;GFLOPS;VARIANT;TM;TN;NT;BS;KT;FLUSH;SPLIT;BSS;SPLITS"
        mine = lx.SmmConfig(variant=0, tm=3, tn=3, nt=128, bs=40, kt=0, flush=1, s=7777, split=3, bs_short=8,
                            split_short=1, gflops=1234.5)
        lx.config_set(lx.REAL_8, 23, 23, 23, mine)
        L.libsmm_acc_params_save(str(out).encode(), b"MI355X")
        lx.config_set(lx.REAL_8, 23, 23, 23, lx.SmmConfig(variant=0, tm=4, tn=4, nt=64, bs=8))
        assert L.libsmm_acc_params_load(str(out).encode()) == 5
        monkeypatch.delenv("OPENCL_LIBSMM_SMM_PARAMS")
        back, src = lx.config_get(lx.REAL_8, 23, 23, 23)
        assert (back.variant, back.tm, back.tn, back.nt, back.bs, back.flush, back.s, back.split, back.bs_short,
                back.split_short) == (0, 3, 3, 128, 40, 1, 7777, 3, 8, 1)
    finally:
        assert L.libsmm_acc_finalize() == 0


def test_params_load_device_match(tmp_path, monkeypatch):
    """A CSV with rows of several devices (the reference's multi-device files): the rows of the device that best
    matches the active one take precedence even when slower, other devices' rows fill the remaining shapes with
    their fastest row, and without a device to match the fastest row of every shape wins (smm_acc.c:98-235)."""
    L = lx.lib()
    head = "DEVICE;TYPEID;M;N;K;S;GFLOPS;BS;BM;BN;BK;WS;WG;LU;NZ;AL;TB;TC;AP;AA;AB;AC\n"
    tail = ";23;1;10;23;-2;-2;0;1;1;1;1;2;2;0;0\n"
    rows = ["NVIDIA A100 80GB PCIe [0x1f79];3;23;23;23;30000;3000.0;50" + tail,
            "AMD Instinct MI355X [0x75a3];3;23;23;23;30000;1000.0;20" + tail,
            "AMD Instinct MI250X [0x740c];3;23;23;23;30000;2000.0;30" + tail,
            "NVIDIA A100 80GB PCIe [0x1f79];3;13;13;13;30000;900.0;41" + tail.replace(";23;", ";13;", 1),
            "AMD Instinct MI250X [0x740c];3;13;13;13;30000;800.0;42" + tail.replace(";23;", ";13;", 1)]
    csv = tmp_path / "multi.csv"
    csv.write_text(head + "".join(rows))
    monkeypatch.setenv("OPENCL_LIBSMM_SMM_PARAMS", "0")
    for device, bs23, bs13 in ((b"AMD Instinct MI355X gfx950:sramecc+:xnack-", 20, 41),
                               (b"AMD Instinct MI250X gfx90a", 30, 42), (b"", 50, 41)):
        assert L.libsmm_acc_init() == 0
        try:
            assert L.libsmm_acc_params_load_device(str(csv).encode(), device) == 5
            assert lx.config_get(lx.REAL_8, 23, 23, 23)[0].bs == bs23, device
            assert lx.config_get(lx.REAL_8, 13, 13, 13)[0].bs == bs13, device
        finally:
            assert L.libsmm_acc_finalize() == 0


@pytest.mark.parametrize("shape", [(23, 23, 23), (5, 5, 5), (64, 64, 64), (45, 45, 13), (4, 9, 22), (80, 80, 80),
                                   (80, 80, 500), (7, 11, 3), (1, 1, 1), (17, 33, 2)])
@pytest.mark.parametrize("tid", [lx.REAL_8, lx.REAL_4])
def test_default_config_valid(shape, tid):
    assert lx.config_valid(tid, *shape, lx.default_config(tid, *shape))
    cfg, src = lx.config_get(tid, *shape)  # tuned row, learned prediction (with a device) or the heuristic
    assert src in (0, 1, 3) and lx.config_valid(tid, *shape, cfg)


@pytest.mark.parametrize("shape,tid,want", [
    ((5, 5, 5), lx.REAL_8, (4, 2, 2, 256)),      # FP64 up to 16x16: 4x4 blocks splitting k, exact block grid
    ((4, 9, 22), lx.REAL_8, (4, 1, 3, 256)),
    ((16, 16, 16), lx.REAL_8, (4, 4, 4, 256)),
    ((17, 5, 5), lx.REAL_8, (2, 2, 1, 512)),     # beyond 16 rows: 16x16 tiles, 8 wavefronts up to 4 tiles
    ((5, 5, 5), lx.REAL_4, (2, 1, 1, 512)),      # FP32: 16x16 tiles
    ((30, 60, 9), lx.REAL_4, (2, 2, 4, 256)),    # 8 tiles: 4 wavefronts per run
    ((40, 40, 9), lx.REAL_4, (2, 3, 3, 128)),    # 9 tiles: 2 wavefronts per run
    ((64, 64, 64), lx.REAL_8, (2, 4, 4, 128)),   # 16 tiles: 2 wavefronts per run
    ((80, 80, 80), lx.REAL_8, (1, 5, 5, 320))])  # more than 16 tiles: one wavefront per run, a wavefront per tile row
def test_default_heuristic_choices(shape, tid, want):
    """the built-in choice for untuned shapes (libsmm_acc_config_heuristic): entry-parallel kernels up to 16 tiles"""
    cfg = lx.default_config(tid, *shape)
    assert (cfg.variant, cfg.tm, cfg.tn, cfg.nt) == want, (shape, cfg.variant, cfg.tm, cfg.tn, cfg.nt)
    assert cfg.bs == (16 if tid == lx.REAL_4 and cfg.variant == 2 else 32) or cfg.variant not in (2, 4)


def test_config_validation():
    tid = lx.REAL_8
    assert not lx.config_valid(tid, 23, 23, 23, lx.SmmConfig(variant=0, tm=3, tn=3, nt=32, bs=32))   # too few threads
    assert not lx.config_valid(tid, 23, 23, 23, lx.SmmConfig(variant=0, tm=7, tn=7, nt=64, bs=32))   # not compiled
    assert not lx.config_valid(tid, 23, 23, 23, lx.SmmConfig(variant=2, tm=1, tn=1, nt=64, bs=32))   # unknown variant
    assert not lx.config_valid(tid, 40, 23, 23, lx.SmmConfig(variant=1, tm=2, tn=1, nt=64, bs=32))   # tiles do not cover n
    assert lx.config_valid(tid, 23, 23, 23, lx.SmmConfig(variant=1, tm=2, tn=2, nt=64, bs=32))
    assert lx.config_valid(tid, 80, 80, 2000, lx.SmmConfig(variant=0, tm=5, tn=5, nt=256, bs=32, kt=40))
    assert not lx.config_valid(tid, 80, 80, 2000, lx.SmmConfig(variant=0, tm=5, tn=5, nt=256, bs=32, kt=2000))
    variants = lx.kernel_variants(tid)
    assert any(v[0] == 0 for v in variants) and any(v[0] == 1 for v in variants)


def test_shipped_params(monkeypatch):
    """libsmm_acc_init loads the tuned MI355X CSV that ships next to the library."""
    monkeypatch.delenv("OPENCL_LIBSMM_SMM_PARAMS", raising=False)
    L = lx.lib()
    assert L.libsmm_acc_init() == 0
    try:
        for tid in (lx.REAL_8, lx.REAL_4):
            for shape in [(5, 5, 5), (13, 13, 13), (23, 23, 23), (32, 32, 32), (64, 64, 64), (4, 9, 22), (45, 45, 13)]:
                cfg, src = lx.config_get(tid, *shape)
                assert src == 1 and cfg.gflops > 0 and lx.config_valid(tid, *shape, cfg), (tid, shape)
    finally:
        assert L.libsmm_acc_finalize() == 0


def test_env_overrides(monkeypatch):
    monkeypatch.setenv("OPENCL_LIBSMM_SMM_BS", "17")
    cfg, src = lx.config_get(lx.REAL_8, 23, 23, 23)
    assert cfg.bs == 17 and src == 2
    monkeypatch.setenv("OPENCL_LIBSMM_SMM_PARAMS", "0")  # suppresses the 2-character keys
    cfg, src = lx.config_get(lx.REAL_8, 23, 23, 23)
    assert cfg.bs != 17
    monkeypatch.delenv("OPENCL_LIBSMM_SMM_PARAMS")
    monkeypatch.setenv("OPENCL_LIBSMM_SMM_NT", "48")  # invalid: ignored
    cfg, src = lx.config_get(lx.REAL_8, 23, 23, 23)
    assert cfg.nt != 48


def test_workload_generator():
    from libxstream_amd.workload import make_stack_workload

    w = make_stack_workload(23, 23, 23, 30000, nc=1000, na=2000, nb=2000, dtype=np.float64)
    s = w.stack
    assert s.shape == (30000, 3) and s.dtype == np.int32
    assert np.all(np.diff(s[:, 2]) >= 0)  # sorted by C
    runs = np.diff(np.flatnonzero(np.diff(np.concatenate([[-1], s[:, 2], [-2]]))))
    assert 26 <= np.median(runs) <= 34
    assert s[:, 0].min() >= 1 and s[:, 0].max() <= (2000 - 1) * 529 + 1
    assert np.all((s - 1) % 529 == 0)
    assert w.bytes_alg() == 30000 * 8 * (529 * 4)
    again = make_stack_workload(23, 23, 23, 30000, nc=1000, na=2000, nb=2000, dtype=np.float64)
    assert np.array_equal(again.stack, s) and np.array_equal(again.a, w.a)
