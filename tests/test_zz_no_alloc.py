"""uasd_energy_terms_noalloc, the variant of uasd_energy_terms that allocates
no device memory (its partial sums sit in a fixed per-device array), called
through the C-ABI and compared with uasd_energy_terms: identical kernels and
summation order, so the results are bit-identical.  Two calls in a row, the
second on a smaller lattice, and a third size beyond the fixed array, which
must be refused with an argument error."""

import numpy as np
import pytest

import uppasd_amd as U
from uppasd_amd import _lib as B
from uppasd_amd import batch as BT
from uppasd_amd import configs as CF

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = 1 << 19


def _system(rng, L):
    s = CF.triangular_dmi(L)
    N = s["lattice"].natom
    e = rng.normal(size=(N, 3))
    s.update(taniso=rng.choice([1, 2, 7], size=N).astype(np.int32),
             eaniso=e / np.linalg.norm(e, axis=1)[:, None],
             kaniso=rng.normal(size=(N, 2)) * 5.0, sb=rng.uniform(0.2, 1.5, size=N),
             mmom=rng.uniform(0.8, 2.4, size=N))
    return s


def _both(eng, nseg, M):
    """(uasd_energy_terms, uasd_energy_terms_noalloc) of one context, each
    called through the C-ABI into its own (4, M, nseg) array."""
    L = B.lib()
    ref = np.full((nseg, M, 4), np.nan)
    got = np.full((nseg, M, 4), np.nan)
    B.check(L.uasd_energy_terms(eng._h, B.ptr(ref)), "uasd_energy_terms")
    B.check(L.uasd_energy_terms_noalloc(eng._h, B.ptr(got)), "uasd_energy_terms_noalloc")
    return ref, got


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_energy_terms_equal_the_allocating_entry_point(precision):
    rng = np.random.default_rng(31)
    # first call: three replicas of uneven size, 5 ensembles, uniform field
    systems = [_system(rng, L) for L in (24, 40, 16)]
    j = BT.join(systems)
    M = 5
    print(f"noalloc {precision} call 1: offsets={list(j.offsets)} M={M}", flush=True)
    eng = BT.engine_for(j, M, precision)
    eng.set_external_field(np.array([0.3, -1.2, 2.0]))
    eng.set_moments(CF.random_unit(j.natom, rng, M), j.mmom)
    ref, got = _both(eng, 3, M)
    assert np.all(np.isfinite(ref)) and np.any(ref != 0.0)
    assert np.array_equal(got, ref), np.max(np.abs(got - ref))
    # a second no-alloc call on the same state gives the same bytes again
    again = np.full_like(got, np.nan)
    B.check(B.lib().uasd_energy_terms_noalloc(eng._h, B.ptr(again)), "uasd_energy_terms_noalloc")
    assert np.array_equal(again, ref)
    del eng

    # second call: a smaller single lattice, 2 ensembles, per-site field;
    # the scratch still holds the first call's larger contents
    s = _system(rng, 12)
    N, M2 = s["lattice"].natom, 2
    print(f"noalloc {precision} call 2: N={N} M={M2}", flush=True)
    eng2 = U.Engine(N, M2, precision)
    eng2.set_exchange(s["nlist"], s["nlistsize"], s["ncoup"])
    eng2.set_dm(s["dmlist"], s["dmlistsize"], s["dm_vect"])
    eng2.set_anisotropy(s["taniso"], s["eaniso"], s["kaniso"], s["sb"])
    eng2.set_external_field(np.ascontiguousarray(rng.normal(size=(M2, N, 3))))
    eng2.set_moments(CF.random_unit(N, rng, M2), s["mmom"])
    # the no-alloc call first: the class tables are built by uploads here only
    # when an allocating call has run, so this one must say so, not allocate
    out = np.zeros((1, M2, 4))
    r = B.lib().uasd_energy_terms_noalloc(eng2._h, B.ptr(out))
    assert r == -5  # UASD_ERR_STATE
    ref2, got2 = _both(eng2, 1, M2)
    assert np.all(np.isfinite(ref2)) and np.any(ref2 != 0.0)
    assert np.array_equal(got2, ref2), np.max(np.abs(got2 - ref2))


def test_noalloc_energy_terms_refuse_sizes_beyond_the_scratch():
    """65536 one-site segments with 2 ensembles need (65536 + 65536) * 2 * 4
    doubles, twice the fixed array: an argument error, and the allocating
    entry point still works on the same context."""
    rng = np.random.default_rng(32)
    s = CF.triangular_dmi(256)
    N, M = s["lattice"].natom, 2
    assert N == 65536
    need = (N + N) * M * 4
    assert need > SCRATCH_DOUBLES
    print(f"noalloc refusal: N={N} nseg={N} M={M}", flush=True)
    eng = U.Engine(N, M)
    eng.set_exchange(s["nlist"], s["nlistsize"], s["ncoup"])
    eng.set_segments(np.arange(N + 1, dtype=np.int32))
    eng.set_moments(CF.random_unit(N, rng, M), s["mmom"])
    ref = np.zeros((N, M, 4))
    B.check(B.lib().uasd_energy_terms(eng._h, B.ptr(ref)), "uasd_energy_terms")
    out = np.zeros((N, M, 4))
    r = B.lib().uasd_energy_terms_noalloc(eng._h, B.ptr(out))
    assert r == -1  # UASD_ERR_ARG
    msg = B.lib().uasd_last_error().decode()
    assert "scratch" in msg and str(need) in msg, msg
    assert not np.any(out)
