"""The oracle's OpenMP site loops give the same numbers on any thread count
(each site is computed by the same operations; the energy sum stays serial),
so the multi-core CPU baseline runs the checked restatement itself."""
import ctypes

import numpy as np

from uppasd_amd import configs as CF


def _run(O, sysd, threads, alg, noise):
    L = O.lib()
    L.or_set_threads(threads)
    N = sysd["lattice"].natom
    ham = O.OracleHam(N, sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    rng = np.random.default_rng(4)
    e = rng.normal(size=(1, N, 3))
    e = np.ascontiguousarray(e / np.linalg.norm(e, axis=2, keepdims=True))
    mm = np.ascontiguousarray(np.broadcast_to(sysd["mmom"], (1, N)))
    eM = np.ascontiguousarray(e * mm[..., None])
    lam, T, lg = np.full(N, 0.1), np.full(N, 300.0), np.ones(N)
    L.or_sd_run(ctypes.byref(ham.h), N, 1, alg, 5, 0, 1e-16, lam, T, lg, mm, e, eM, 11, noise, None)
    b, en = ham.field(eM)
    return e, b, en


def test_thread_count_does_not_change_results(O):
    small = CF.bcc_shells(10, [1.34, 0.77], moment=2.2)     # 2000 sites: serial loops
    big = CF.bcc_shells(14, [1.34, 0.77], moment=2.2)       # 5488 sites: parallel loops
    prev = O.lib().or_threads()
    try:
        for s in (small, big):
            for alg, noise in ((5, 2), (1, 2), (5, 1)):
                e1, b1, en1 = _run(O, s, 1, alg, noise)
                e8, b8, en8 = _run(O, s, 8, alg, noise)
                assert np.array_equal(e1, e8) and np.array_equal(b1, b8) and en1 == en8
    finally:
        O.lib().or_set_threads(prev)
