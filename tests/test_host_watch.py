"""The host write detection behind the drop-in module's device mirrors
(csrc/host/HostWriteWatch.hpp).  Needs no GPU: it only protects and compares
host memory."""
import os
import subprocess
import sys

import numpy as np
import pytest


@pytest.fixture(scope="module")
def ops():
    from lightning_unification_amd import lightning_qubit_ops
    return lightning_qubit_ops


def test_reads_do_not_count(ops):
    a = np.arange(1 << 18, dtype=np.complex128)
    w = ops._HostWriteWatch()
    w.watch(a)
    assert not w.written()
    assert float(np.sum(a.real)) == float(np.sum(np.arange(1 << 18)))
    b = a.copy()
    assert not w.written()
    assert np.array_equal(a, b)
    w.release()


@pytest.mark.parametrize("where", ["first", "middle", "last", "slice", "ufunc", "copyto"])
def test_writes_are_seen_and_take_effect(ops, where):
    a = np.zeros(1 << 17, dtype=np.complex128)
    w = ops._HostWriteWatch()
    w.watch(a)
    if where == "first":
        a[0] = 1 + 2j
    elif where == "middle":
        a[a.size // 2] = 3.0
    elif where == "last":
        a[-1] = -1j
    elif where == "slice":
        a[1000:1010] = 7
    elif where == "ufunc":
        np.add(a, 1, out=a)
    else:
        np.copyto(a, np.ones_like(a))
    assert w.written()
    # the write completed in full
    if where == "first":
        assert a[0] == 1 + 2j
    elif where == "middle":
        assert a[a.size // 2] == 3.0
    elif where == "last":
        assert a[-1] == -1j
    elif where == "slice":
        assert np.all(a[1000:1010] == 7)
    else:
        assert np.all(a == 1)
    w.release()
    a[5] = 9  # writable again after release
    assert a[5] == 9


def test_edges_of_unaligned_buffers(ops):
    raw = np.zeros((1 << 16) * 16 + 64, np.uint8)
    for off in (0, 16, 48):
        a = raw[off:off + (1 << 16) * 16].view(np.complex128)
        w = ops._HostWriteWatch()
        w.watch(a)
        assert not w.written()
        raw[off - 1 if off else -1] = 5  # a byte next to the buffer, not in it
        assert not w.written()
        a[0] = 1
        assert w.written()
        w.release()
        w.watch(a)
        a[-1] = 2
        assert w.written()
        w.release()


def test_small_buffer_without_whole_pages(ops):
    a = np.zeros(8, np.complex128)
    w = ops._HostWriteWatch()
    w.watch(a)
    assert not w.written()
    a[3] = 1
    assert w.written()
    w.release()


def test_overlapping_watches(ops):
    a = np.zeros(1 << 16, np.complex128)
    w1, w2 = ops._HostWriteWatch(), ops._HostWriteWatch()
    w1.watch(a)
    w2.watch(a)
    w1.release()  # the pages become writable: w2 can no longer see writes
    assert w2.written()
    w2.release()
    w1.watch(a)
    w2.watch(a[: a.size // 2])
    a[10] = 1
    assert w1.written() and w2.written()
    w1.release()
    w2.release()


def test_foreign_faults_still_crash():
    """A fault outside every watched range still takes the default action."""
    code = ("import ctypes, numpy as np\n"
            "from lightning_unification_amd import lightning_qubit_ops as o\n"
            "a = np.zeros(1 << 16, np.complex128); w = o._HostWriteWatch(); w.watch(a)\n"
            "a[100] = 1\n"
            "assert w.written()\n"
            "print('watched write ok', flush=True)\n"
            "ctypes.string_at(8)\n")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True,
                       timeout=120)
    assert "watched write ok" in r.stdout
    assert r.returncode in (-11, 139), (r.returncode, r.stderr[-500:])
