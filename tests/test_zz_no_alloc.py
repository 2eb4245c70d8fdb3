"""s21_batch_op_noalloc (the operating point that allocates nothing) against s21_batch_op, through the C
ABI, bit for bit: two calls in a row on one GPU, the second on a smaller batch, so the second runs on the
scratch array the first left dirty; and the errors for a batch without structure and one whose state does
not fit."""
import ctypes as C

import numpy as np
import pytest

from spice21_amd import _abi as A
from spice21_amd import workloads as W
from spice21_amd._lib import lib
from spice21_amd.batch import FP64, MIXED, Batch, Group

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


class DeviceBuffers:
    """Device buffers allocated by the test through the HIP runtime the engine library links (the entry
    point under test allocates nothing; its caller does)."""

    def __init__(self, *sizes):
        self.hip = C.CDLL("libamdhip64.so")
        self.ptrs = []
        for nbytes in sizes:
            p = C.c_void_p()
            assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 8))) == 0
            self.ptrs.append(p)

    def upload(self, k, arr):
        arr = np.ascontiguousarray(arr)
        assert self.hip.hipMemcpy(self.ptrs[k], arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1) == 0

    def download(self, k, arr):
        assert self.hip.hipMemcpy(arr.ctypes.data_as(C.c_void_p), self.ptrs[k], C.c_size_t(arr.nbytes), 2) == 0
        return arr

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def call_noalloc(b, bufs):
    return lib().s21_batch_op_noalloc(b.h, C.cast(bufs.ptrs[0], C.POINTER(C.c_double)),
                                       C.cast(bufs.ptrs[1], C.POINTER(C.c_int32)),
                                       C.cast(bufs.ptrs[2], C.POINTER(C.c_int64)))


def noalloc_op(b, n_inst, n_vars):
    """Runs s21_batch_op_noalloc into freshly allocated device buffers filled with garbage; returns
    (status code, x, status, iterations) copied back to the host."""
    bufs = DeviceBuffers(n_inst * n_vars * 8, n_inst * 4, n_inst * 8)
    try:
        bufs.upload(0, np.full(n_inst * n_vars, 7.0))  # garbage the call must clear
        bufs.upload(1, np.full(n_inst, 99, dtype=np.int32))
        bufs.upload(2, np.full(n_inst, -5, dtype=np.int64))
        rc = call_noalloc(b, bufs)
        x = bufs.download(0, np.zeros(n_inst * n_vars)).reshape(n_inst, n_vars)
        st = bufs.download(1, np.zeros(n_inst, dtype=np.int32))
        it = bufs.download(2, np.zeros(n_inst, dtype=np.int64))
        return rc, x, st, it
    finally:
        bufs.free()


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_noalloc_op_matches_op_twice_in_a_row(precision):
    runs = []
    for n_inst in (96, 24):  # the second call is smaller: it reuses the first call's (dirty) scratch
        nl, over = W.chain_mc(n_inst=n_inst, seed=0xC2 + n_inst)
        b = Batch(Group(nl.to_c(), n_inst, over), precision=precision)
        log("noalloc", precision, n_inst, "reference call")
        b.op()  # builds and uploads the structure (allocating), the result to compare with
        x_ref, (st_ref, it_ref) = b.x(), b.status()
        assert np.all(st_ref == 0)
        log("noalloc", precision, n_inst, "allocation-free call")
        rc, x, st, it = noalloc_op(b, n_inst, b.n_vars())
        assert rc == A.S21_OK
        assert np.array_equal(st, st_ref) and np.array_equal(it, it_ref)
        assert np.array_equal(x, x_ref)
        # the batch's own results are untouched
        assert np.array_equal(b.x(), x_ref)
        runs.append(b)
    # two groups in one batch: outputs laid out group after group
    nl1, over1 = W.chain_mc(n_inst=8, seed=5)
    nl2 = W.ring5()
    b = Batch([Group(nl1.to_c(), 8, over1), Group(nl2.to_c(), 3)], precision=precision)
    log("noalloc", precision, "two groups")
    b.op()
    n1, n2 = b.n_vars(0), b.n_vars(1)
    bufs = DeviceBuffers((8 * n1 + 3 * n2) * 8, 11 * 4, 11 * 8)
    try:
        assert call_noalloc(b, bufs) == A.S21_OK
        xh = bufs.download(0, np.zeros(8 * n1 + 3 * n2))
        sth = bufs.download(1, np.zeros(11, dtype=np.int32))
        ith = bufs.download(2, np.zeros(11, dtype=np.int64))
    finally:
        bufs.free()
    assert np.array_equal(xh[: 8 * n1].reshape(8, n1), b.x(0))
    assert np.array_equal(xh[8 * n1:].reshape(3, n2), b.x(1))
    assert np.array_equal(sth, np.concatenate([b.status(0)[0], b.status(1)[0]]))
    assert np.array_equal(ith, np.concatenate([b.status(0)[1], b.status(1)[1]]))


def test_noalloc_op_rejects_unprepared_and_oversized_batches():
    nl, over = W.chain_mc(n_inst=16)
    b = Batch(Group(nl.to_c(), 16, over))
    log("noalloc unprepared")
    rc, _, _, _ = noalloc_op(b, 16, nl.to_c().n_vars)
    assert rc == A.S21_INVALID
    # 128 Mos1 x 14 state fields x 8 B per instance: 5,000 instances need about 72 MB > 64 MiB
    nl, over = W.chain_mc(n_inst=5000)
    b = Batch(Group(nl.to_c(), 5000, over))
    log("noalloc oversized")
    b.op()
    rc, _, _, _ = noalloc_op(b, 5000, b.n_vars())
    assert rc == A.S21_INVALID
    buf = C.create_string_buffer(256)
    lib().s21_last_error(buf, 256)
    assert b"scratch" in buf.value


def test_noalloc_op_with_bsim4_devices_matches_op():
    """BSIM4 devices keep their points in the batch's own buffers (allocated by the preparing s21_batch_op
    call): the allocation-free call clears them and reproduces s21_batch_op bit for bit, twice in a row."""
    from tests.test_gpu_bsim4 import varied_instances

    nl = varied_instances()
    for n_inst in (12, 6):
        over = [(k, "inst", "delvto", 0.02 * np.linspace(-1, 1, n_inst))
                for k, c in enumerate(nl.comps) if c[0] == A.S21_COMP_BSIM4]
        b = Batch(Group(nl.to_c(), n_inst, over))
        log("noalloc bsim4", n_inst)
        b.op()
        x_ref, (st_ref, it_ref) = b.x(), b.status()
        assert np.all(st_ref == 0)
        for _ in range(2):
            rc, x, st, it = noalloc_op(b, n_inst, b.n_vars())
            assert rc == A.S21_OK
            assert np.array_equal(st, st_ref) and np.array_equal(it, it_ref)
            assert np.array_equal(x, x_ref)
