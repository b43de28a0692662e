"""fla_hip_dtrsm_noalloc: the allocation-free dtrsm against fla_hip_dtrsm.

The no-allocation entry point runs the same kernels with its workspace in a
file-scope device array, so its results must equal fla_hip_dtrsm's bit for
bit.  Two calls in a row (the second on a smaller problem, so the array
keeps the first call's contents), the narrow (<= 32 right-hand sides) and the
blocked path, left and right side, and the size limit (order > 8192 is
rejected before any GPU work).
"""
import numpy as np
import pytest

import libflame_amd as fla
from tests.gpu_util import to_dev

pytestmark = pytest.mark.gpu
Lo, Up = fla.FLA_LOWER_TRIANGULAR, fla.FLA_UPPER_TRIANGULAR
N, T = fla.FLA_NO_TRANSPOSE, fla.FLA_TRANSPOSE
NU, UN = fla.FLA_NONUNIT_DIAG, fla.FLA_UNIT_DIAG
LIMIT = 8192


@pytest.fixture(scope="module", autouse=True)
def _init():
    fla.FLA_Init()
    yield


def _triangle(n, seed, upper):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1.0, 1.0, size=(n, n)) / n
    A[np.diag_indices(n)] = rng.uniform(1.0, 2.0, size=n) * rng.choice([-1.0, 1.0], size=n)
    # the other triangle holds values that neither entry point may read
    other = np.tril(np.ones((n, n), bool), -1) if upper else np.triu(np.ones((n, n), bool), 1)
    return np.asfortranarray(np.where(other, rng.uniform(-1e3, 1e3, size=(n, n)), A))


def _run(fn, side, uplo, trans, diag, A, B, alpha):
    dA, dB = to_dev(A), to_dev(B)
    m, n = B.shape
    fla.check(fn(side, uplo, trans, diag, m, n, alpha, dA.ptr, dA.ld, dB.ptr, dB.ld), "dtrsm")
    fla.synchronize()
    return dB.download()


CASES = [
    # side, uplo, trans, diag, (order, other dim) first call, then second call
    (fla.FLA_LEFT, Lo, N, NU, (3000, 16), (700, 5)),
    (fla.FLA_LEFT, Up, T, UN, (LIMIT, 8), (129, 3)),
    (fla.FLA_LEFT, Lo, T, NU, (1000, 200), (300, 40)),
    (fla.FLA_RIGHT, Up, N, NU, (2000, 64), (257, 33)),
]


@pytest.mark.parametrize("side,uplo,trans,diag,first,second", CASES)
def test_noalloc_equals_dtrsm_twice(side, uplo, trans, diag, first, second):
    lib = fla.lib()
    for k, (na, other) in enumerate((first, second)):
        print(f"case side={side} uplo={uplo} trans={trans} diag={diag} order={na} "
              f"other={other} call={k}", flush=True)
        A = _triangle(na, seed=na + k, upper=(uplo == Up))
        rng = np.random.default_rng(7 + na)
        B = np.asfortranarray(rng.standard_normal((na, other) if side == fla.FLA_LEFT
                                                  else (other, na)))
        alpha = 1.0 if k == 0 else -0.5
        want = _run(lib.fla_hip_dtrsm, side, uplo, trans, diag, A, B, alpha)
        got = _run(lib.fla_hip_dtrsm_noalloc, side, uplo, trans, diag, A, B, alpha)
        assert np.all(np.isfinite(want))
        assert np.array_equal(got, want)


def test_noalloc_rejects_order_beyond_scratch():
    lib = fla.lib()
    na = LIMIT + 1
    dA = fla.DeviceMatrix(na, na)
    dB = fla.DeviceMatrix(na, 2)
    st = lib.fla_hip_dtrsm_noalloc(fla.FLA_LEFT, Lo, N, NU, na, 2, 1.0, dA.ptr, dA.ld, dB.ptr,
                                   dB.ld)
    assert st == fla.FLA_HIP_ESCRATCH
    # invalid arguments are still reported as such
    st = lib.fla_hip_dtrsm_noalloc(fla.FLA_LEFT, Lo, N, NU, 10, 2, 1.0, dA.ptr, 5, dB.ptr, 10)
    assert st == fla.FLA_HIP_EARG
