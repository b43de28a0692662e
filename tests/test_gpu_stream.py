"""The caller's stream (fla_hip_set_stream): the library's work is ordered
after the work already queued on that stream, and the caller's later work on
it after the library's.  The stream replaces the rocblas_handle that the
reference's FLA_*_external_hip wrappers take
(src/base/flamec/supermatrix/hip/main/FLASH_Queue_hip.c:297-318).

Each case queues a long device-to-device copy and then the upload of the
operands on the caller's stream, calls the library, and queues the download
of the result on the same stream.  No host synchronisation happens between
them, so without the ordering the library would read the operands before
they arrive."""
import ctypes as C

import numpy as np
import pytest

import libflame_amd as fla

pytestmark = pytest.mark.gpu

H2D, D2H, D2D = 1, 2, 3  # hipMemcpyKind


@pytest.fixture(scope="module")
def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    h.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
    h.hipHostFree.argtypes = [C.c_void_p]
    return h


@pytest.fixture(scope="module")
def stream(hip):
    fla.FLA_Init()
    s = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(s)) == 0
    L = fla.lib()
    assert L.fla_hip_set_stream(s) == 0
    assert L.fla_hip_get_stream() == s.value
    yield s
    assert L.fla_hip_set_stream(None) == 0
    assert L.fla_hip_get_stream() is None
    hip.hipStreamSynchronize(s)
    hip.hipStreamDestroy(s)


class Pinned:
    """Pinned host array (an asynchronous copy from it is ordered on the stream)."""

    def __init__(self, hip, shape):
        self.hip = hip
        n = int(np.prod(shape)) * 8
        self.p = C.c_void_p()
        assert hip.hipHostMalloc(C.byref(self.p), n, 0) == 0
        self.a = np.frombuffer((C.c_double * int(np.prod(shape))).from_address(self.p.value),
                               dtype=np.float64).reshape(shape, order="F")

    def free(self):
        self.hip.hipHostFree(self.p)


def _busy(hip, s, big_src, big_dst):
    """About a millisecond of device-to-device copying on the caller's stream."""
    for _ in range(4):
        assert hip.hipMemcpyAsync(big_dst.ptr, big_src.ptr, big_src.nbytes, D2D, s) == 0


def test_dgemm_is_ordered_on_the_callers_stream(hip, stream):
    L = fla.lib()
    m, n, k = 512, 384, 256
    rng = np.random.default_rng(5)
    ha, hb, hc = Pinned(hip, (m, k)), Pinned(hip, (k, n)), Pinned(hip, (m, n))
    ha.a[:] = rng.integers(-4, 5, size=(m, k))
    hb.a[:] = rng.integers(-4, 5, size=(k, n))
    hc.a[:] = 0.0
    dA, dB, dC = fla.DeviceMatrix(m, k), fla.DeviceMatrix(k, n), fla.DeviceMatrix(m, n)
    big_src, big_dst = fla.DeviceBuffer(1 << 30), fla.DeviceBuffer(1 << 30)
    for rep in range(3):
        # zero the operands first (synchronously), then queue the real upload behind busy work
        dA.upload(np.zeros((m, k), order="F"))
        dB.upload(np.zeros((k, n), order="F"))
        _busy(hip, stream, big_src, big_dst)
        assert hip.hipMemcpyAsync(dA.ptr, ha.p, m * k * 8, H2D, stream) == 0
        assert hip.hipMemcpyAsync(dB.ptr, hb.p, k * n * 8, H2D, stream) == 0
        fla.check(L.fla_hip_dgemm(fla.FLA_NO_TRANSPOSE, fla.FLA_NO_TRANSPOSE, m, n, k, 1.0,
                                  dA.ptr, m, dB.ptr, k, 0.0, dC.ptr, m))
        assert hip.hipMemcpyAsync(hc.p, dC.ptr, m * n * 8, D2H, stream) == 0
        assert hip.hipStreamSynchronize(stream) == 0
        assert np.array_equal(hc.a, ha.a @ hb.a), rep
    for p in (ha, hb, hc):
        p.free()


def test_cholesky_is_ordered_on_the_callers_stream(hip, stream):
    L = fla.lib()
    n = 640
    rng = np.random.default_rng(6)
    F = np.tril(rng.integers(-3, 4, size=(n, n)).astype(np.float64), -1) + np.diag(
        rng.integers(1, 4, size=n).astype(np.float64))
    ha, hl = Pinned(hip, (n, n)), Pinned(hip, (n, n))
    ha.a[:] = F @ F.T
    dA = fla.DeviceMatrix(n, n)
    big_src, big_dst = fla.DeviceBuffer(1 << 30), fla.DeviceBuffer(1 << 30)
    info = C.c_int(0)
    # a matrix that fails at column 0 sits there until the upload lands
    dA.upload(np.asfortranarray(-np.eye(n)))
    _busy(hip, stream, big_src, big_dst)
    assert hip.hipMemcpyAsync(dA.ptr, ha.p, n * n * 8, H2D, stream) == 0
    fla.check(L.fla_hip_dpotrf(fla.FLA_LOWER_TRIANGULAR, n, dA.ptr, n, C.byref(info)))
    assert info.value == fla.FLA_SUCCESS
    assert hip.hipMemcpyAsync(hl.p, dA.ptr, n * n * 8, D2H, stream) == 0
    assert hip.hipStreamSynchronize(stream) == 0
    assert np.array_equal(np.tril(hl.a), F)
    for p in (ha, hl):
        p.free()
