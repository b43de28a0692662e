"""Writes the full-size parity fixtures under tests/golden/ from the CPU
oracle (oracle/liboracle.so, the restatement of libflame's FLA_LU_piv control
tree: blocked var5 at 128 -> var5 at 16 -> Crout opt_var4, f2c BLAS order,
idamax's first-max rule).  The reference's own tests hold no golden vectors
(test/src/test_lu_piv.c:27-30 is residual-only), so the fixtures pin the
configured sizes against the oracle:

  lu_c2_n16384_seed7.npz   FLA_LU_piv on the n = 16384 well-separated-pivot
                           matrix (U[-1,1], row i scaled by 10^(-i/n), seed 7;
                           or_gen_lu_c2): the FLAME-native pivot vector, the
                           return code, the dget01 residual of the oracle's
                           factors, and a SHA-256 of the input bytes (column
                           major) so that a change of generator is caught.

Run from the repository root:  python tests/golden/make_fixtures.py
(about 15-30 minutes on one core: the oracle is single-threaded f2c code).
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402


def lu_c2(n=16384, seed=7):
    t0 = time.time()
    A = orc.gen_lu_c2(seed, n, n)
    digest = hashlib.sha256(np.asfortranarray(A).tobytes(order="F")).hexdigest()
    print(f"generated n={n} in {time.time() - t0:.1f} s, sha256 {digest}", flush=True)
    t0 = time.time()
    F, p, r = orc.lu_piv(A)
    print(f"oracle FLA_LU_piv in {time.time() - t0:.1f} s, r={r}", flush=True)
    # dget01 ratio on a column sample would be partial; the full O(n^3) check
    # in long double is too slow here, so record the factor's growth instead.
    growth = float(np.max(np.abs(np.triu(F))) / np.max(np.abs(A)))
    out = os.path.join(HERE, f"lu_c2_n{n}_seed{seed}.npz")
    np.savez_compressed(out, pivots=p.astype(np.int32), ret=np.int32(r), n=np.int32(n),
                        seed=np.int32(seed), input_sha256=np.array(digest),
                        growth=np.float64(growth),
                        u_diag=np.diag(F).astype(np.float64))
    print(f"wrote {out}", flush=True)


if __name__ == "__main__":
    lu_c2()
