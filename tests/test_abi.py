"""CPU-side checks of the drop-in boundary: the C-ABI library loads and
exports every function that include/*.h declares, the constants keep the
reference's numeric values, and parameter errors abort as FLA_Check_error_code
does (no GPU call is made in this file)."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "libflame_amd", "libflame_amd.so")


def declared_functions():
    names = set()
    for h in ("FLAME.h", "FLA_hip.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return names


def test_library_built():
    assert os.path.exists(LIB), "run __graft_entry__.build() first"


def test_exports_every_declared_function():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(n for n in declared_functions() if n not in exported)
    assert not missing, f"declared but not exported: {missing}"
    # the FLA_Obj constants are data symbols
    for sym in ("FLA_ONE", "FLA_ZERO", "FLA_MINUS_ONE"):
        assert sym in exported


def test_constants_match_reference_values():
    text = open(os.path.join(ROOT, "include", "FLAME.h")).read()
    want = {"FLA_SUCCESS": "(-1)", "FLA_FAILURE": "(-2)", "FLA_LOWER_TRIANGULAR": "300",
            "FLA_UPPER_TRIANGULAR": "301", "FLA_NATIVE_PIVOTS": "700", "FLA_LAPACK_PIVOTS": "701",
            "FLA_DOUBLE": "101", "FLA_INT": "104", "FLA_LEFT": "210", "FLA_NO_TRANSPOSE": "400",
            "FLA_UNIT_DIAG": "500", "FLA_INVALID_UPLO": "(-11)",
            "FLA_OBJECT_NOT_SQUARE": "(-24)"}
    for k, v in want.items():
        assert re.search(rf"#define {k} {re.escape(v)}(?![\w(])", text), k


def test_python_binding_loads_without_gpu():
    import libflame_amd as fla
    L = fla.lib()
    L.FLA_Init()
    assert L.FLA_Initialized() == 1
    assert L.fla_hip_get_stream() is None  # no caller's stream until one is set
    L.FLA_Finalize()


_ABORT_SNIPPET = r"""
import ctypes as C, sys
sys.path.insert(0, {root!r})
import numpy as np
import libflame_amd as fla
L = fla.lib()
L.FLA_Init()
o = fla.FLA_Obj()
L.FLA_Obj_create_without_buffer(fla.FLA_DOUBLE, 4, 3, C.byref(o))
buf = np.zeros((4, 3), order="F")
o.base.contents.buffer = buf.ctypes.data
o.base.contents.rs, o.base.contents.cs = 1, 4
{call}
print("returned")
"""


@pytest.mark.parametrize("call", ["L.FLA_Chol(fla.FLA_LOWER_TRIANGULAR, o)",  # not square
                                  "L.FLA_Chol(999, o)"])                     # bad uplo
def test_parameter_errors_abort(call):
    """FLA_Chol_check fails -> message + abort() (SIGABRT), before any GPU use."""
    code = _ABORT_SNIPPET.format(root=ROOT, call=call)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == -6, (r.returncode, r.stdout, r.stderr)
    assert "libflame:" in r.stderr
    assert "returned" not in r.stdout


def test_error_checking_disabled_skips_abort():
    """With FLA_NO_ERROR_CHECKING the check is skipped (FLA_Chol.c:20-22); a
    zero-size object then returns FLA_SUCCESS without touching the GPU."""
    code = r"""
import ctypes as C, sys
sys.path.insert(0, %r)
import libflame_amd as fla
L = fla.lib()
L.FLA_Init()
L.FLA_Check_error_level_set(fla.FLA_NO_ERROR_CHECKING)
o = fla.FLA_Obj()
L.FLA_Obj_create_without_buffer(fla.FLA_DOUBLE, 0, 0, C.byref(o))
print(L.FLA_Chol(999, o))
""" % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "-1"


def test_inverse_lapack_argument_errors(capfd):
    """dtrtri_ / dlauum_ / dpotri_ argument checks (netlib order) return
    before any device work."""
    import ctypes as C

    import numpy as np

    import libflame_amd as fla

    L = fla.lib()
    a = np.zeros((4, 4), order="F")
    ap = a.ctypes.data_as(C.POINTER(C.c_double))
    n, bad, info = (np.array([v], np.int32) for v in (4, 2, 0))
    ip = lambda x: x.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    L.dtrtri_(C.c_char_p(b"X"), C.c_char_p(b"N"), ip(n), ap, ip(n), ip(info))
    assert info[0] == -1
    L.dtrtri_(C.c_char_p(b"L"), C.c_char_p(b"Q"), ip(n), ap, ip(n), ip(info))
    assert info[0] == -2
    L.dtrtri_(C.c_char_p(b"L"), C.c_char_p(b"N"), ip(n), ap, ip(bad), ip(info))
    assert info[0] == -5
    L.dlauum_(C.c_char_p(b"L"), ip(np.array([-1], np.int32)), ap, ip(n), ip(info))
    assert info[0] == -2
    L.dpotri_(C.c_char_p(b"U"), ip(n), ap, ip(bad), ip(info))
    assert info[0] == -4
    assert "DTRTRI" in capfd.readouterr().err


def test_lu_blocksize_values():
    """lu_nb: multiples of 32 up to 128 (one panel per step) or 256 (pairs of
    128-column panels); anything else is an argument error.  Host-only."""
    import libflame_amd as fla

    L = fla.lib()
    for ok in (32, 64, 96, 128, 256):
        assert L.fla_hip_set_blocksizes(0, ok) == 0
    for bad in (16, 160, 192, 224, 288, 512):
        assert L.fla_hip_set_blocksizes(0, bad) != 0
    assert L.fla_hip_set_blocksizes(0, 256) == 0
    # -1 restores the defaults (Cholesky block chosen from n, lu_nb = 256)
    assert L.fla_hip_set_blocksizes(512, 64) == 0
    assert L.fla_hip_chol_blocksize(32768) == 512
    assert L.fla_hip_set_blocksizes(-1, -1) == 0
    assert L.fla_hip_chol_blocksize(32768) == 1024
    assert L.fla_hip_chol_blocksize(8192) == 256
    cb, lb = C.c_int(), C.c_int()
    L.fla_hip_get_blocksizes(C.byref(cb), C.byref(lb))
    assert (cb.value, lb.value) == (0, 256)
    assert L.fla_hip_set_blocksizes(-2, 0) != 0


def test_driver_lapack_argument_errors(capfd):
    """dposv_ / dgesv_ argument checks (dposv.c:153-175, dgesv.c:145-166:
    -1 uplo/n, ... -7 ldb) return before any device work; dlaswp_ with
    incx == 0 or k2 < k1 does nothing (dlaswp.c:130-145)."""
    import ctypes as C

    import numpy as np

    import libflame_amd as fla

    L = fla.lib()
    a = np.arange(16.0).reshape(4, 4, order="F")
    b = np.ones((4, 2), order="F")
    ap = a.ctypes.data_as(C.POINTER(C.c_double))
    bp = b.ctypes.data_as(C.POINTER(C.c_double))
    ip = lambda x: x.ctypes.data_as(C.POINTER(C.c_int))  # noqa: E731
    v = lambda x: np.array([x], np.int32)  # noqa: E731
    info = v(0)
    piv = np.array([4, 3, 3, 4], np.int32)
    for args, want in [((b"X", 4, 2, 4, 4), -1), ((b"L", -1, 2, 4, 4), -2),
                       ((b"L", 4, -1, 4, 4), -3), ((b"L", 4, 2, 3, 4), -5),
                       ((b"L", 4, 2, 4, 3), -7)]:
        u, n, nr, lda, ldb = args
        L.dposv_(C.c_char_p(u), ip(v(n)), ip(v(nr)), ap, ip(v(lda)), bp, ip(v(ldb)), ip(info))
        assert info[0] == want
    for args, want in [((-1, 2, 4, 4), -1), ((4, -1, 4, 4), -2), ((4, 2, 3, 4), -4),
                       ((4, 2, 4, 3), -7)]:
        n, nr, lda, ldb = args
        L.dgesv_(ip(v(n)), ip(v(nr)), ap, ip(v(lda)), ip(piv), bp, ip(v(ldb)), ip(info))
        assert info[0] == want
    before = a.copy()
    L.dlaswp_(ip(v(4)), ap, ip(v(4)), ip(v(1)), ip(v(4)), ip(piv), ip(v(0)))
    L.dlaswp_(ip(v(4)), ap, ip(v(4)), ip(v(3)), ip(v(2)), ip(piv), ip(v(1)))
    assert np.array_equal(a, before) and np.array_equal(b, np.ones((4, 2)))
    assert "DPOSV" in capfd.readouterr().err
