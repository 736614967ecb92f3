"""The C ABI library: it loads, exports every function include/qk.h declares,
and its host-only calls behave without a GPU (no kernel is launched here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "qk.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(qk_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("qk_ctx_create", "qk_price", "qk_price_payoffs", "qk_price_lanes", "qk_price_greeks",
                 "qk_price_american", "qk_simulate_paths", "qk_last_timing", "qk_footprint_bytes"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from quantkos_amd import _lib
    L = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts_match_header():
    from quantkos_amd import _lib
    # qk_contract: 4 int32, int64, uint64, 13 doubles
    assert C.sizeof(_lib.Contract) == 4 * 4 + 8 + 8 + 13 * 8
    assert C.sizeof(_lib.Payoff) == 24
    assert C.sizeof(_lib.Stats) == 5 * 8 + 2 * 8
    assert C.sizeof(_lib.Greeks) == 9 * 8 + 8
    assert C.sizeof(_lib.Bumps) == 6 * 4 + 4 * 8


def test_strerror_and_footprint_are_host_only():
    from quantkos_amd import _lib
    import quantkos_amd as qk
    L = _lib.load()
    assert L.qk_strerror(0) == b"ok"
    assert L.qk_strerror(-2) == b"out of memory budget"
    c = qk.contract(n_paths=500_000, n_steps=50, precision=qk.FP32)
    eu = L.qk_footprint_bytes(C.byref(c), 1, 1, 0)
    am = L.qk_footprint_bytes(C.byref(c), 1, 1, 1)
    assert am - eu >= 500_000 * 50 * 4          # the FP32 path store [date][path]
    c.precision = qk.FP64
    assert L.qk_footprint_bytes(C.byref(c), 1, 1, 1) - am >= 500_000 * 50 * 4


def test_ctx_create_reports_missing_device_as_error():
    from quantkos_amd import _lib
    L = _lib.load()
    n = C.c_int(0)
    h = C.c_void_p()
    st = L.qk_ctx_create(0, 0, C.byref(h))
    if st == 0:
        L.qk_ctx_destroy(h)
        pytest.skip("a GPU is present")
    assert st in (-3, -1) and not h.value


def test_footprint_saturates_instead_of_wrapping():
    """A path store of n_paths * n_steps * 8 bytes beyond 2^64 reads as SIZE_MAX
    (no budget admits it) instead of wrapping to a small size; so do n_paths
    or n_steps below 1."""
    from quantkos_amd import _lib
    import quantkos_amd as qk
    L = _lib.load()
    size_max = 2 ** (8 * C.sizeof(C.c_size_t)) - 1
    c = qk.contract(n_paths=1 << 31, n_steps=1 << 30, precision=qk.FP64)
    assert L.qk_footprint_bytes(C.byref(c), 1, 1, 1) == size_max
    c = qk.contract(n_paths=1 << 40, n_steps=1 << 30, precision=qk.FP32)
    assert L.qk_footprint_bytes(C.byref(c), 1, 1, 1) == size_max
    c = qk.contract(n_paths=0, n_steps=10)
    assert L.qk_footprint_bytes(C.byref(c), 1, 1, 0) == size_max
    c = qk.contract(n_paths=1 << 20, n_steps=100, precision=qk.FP32)
    assert L.qk_footprint_bytes(C.byref(c), 1, 1, 1) < 2 ** 30


def test_package_raises_the_hardware_queue_count():
    """Importing the package raises GPU_MAX_HW_QUEUES to 12 when it is unset or
    lower (each of the request server's streams needs a hardware queue of its
    own), keeps a larger setting, and leaves it alone with QK_KEEP_HW_QUEUES=1."""
    import subprocess
    import sys
    code = "import os, quantkos_amd; print(os.environ['GPU_MAX_HW_QUEUES'])"
    for before, keep, after in (("4", None, "12"), (None, None, "12"), ("16", None, "16"), ("4", "1", "4")):
        env = dict(os.environ)
        env.pop("GPU_MAX_HW_QUEUES", None)
        env.pop("QK_KEEP_HW_QUEUES", None)
        if before is not None:
            env["GPU_MAX_HW_QUEUES"] = before
        if keep is not None:
            env["QK_KEEP_HW_QUEUES"] = keep
        out = subprocess.run([sys.executable, "-c", code], env=env, cwd=ROOT, capture_output=True, text=True,
                             check=True).stdout.strip()
        assert out == after, (before, keep, out)
