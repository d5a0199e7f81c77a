"""Where every site's DM bonds are its exchange bonds (the triangular
monolayer and the kagome lattice list the same bonds in jfile and dmfile),
the list kernels gather each neighbour once for both sums and add the terms
in the order of the separate sums.  Fields, LLG trajectories and MC sweeps
are bit-identical to the separate gathers (tuning flag 0x800000)."""
import numpy as np
import pytest

import uppasd_amd as U
from uppasd_amd import batch as Bt
from uppasd_amd import configs as CF

pytestmark = pytest.mark.gpu
SEPARATE = 0x800000


def _engine(s, M, precision, separate, N=None):
    N = s["lattice"].natom if N is None else N
    eng = U.Engine(N, M, precision)
    eng.set_exchange(s["nlist"], s["nlistsize"], s["ncoup"])
    eng.set_dm(s["dmlist"], s["dmlistsize"], s["dm_vect"])
    if "taniso" in s:
        eng.set_anisotropy(s["taniso"], s["eaniso"], s["kaniso"], s["sb"])
    if "ext" in s:
        eng.set_external_field(s["ext"])
    for kind in (eng.TUNE_LLG, eng.TUNE_FIELD, eng.TUNE_MC):
        eng.set_tuning(kind, 4, 256 | (SEPARATE if separate else 0))
    return eng


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fused_dm_gather_triangular_bit_identical(precision):
    s = CF.triangular_dmi(16)
    N, M = s["lattice"].natom, 2
    rng = np.random.default_rng(81)
    e0 = np.ascontiguousarray(CF.random_unit(N, rng, M))
    print(f"fused DM triangular {precision} N={N} M={M}", flush=True)
    out = []
    for separate in (False, True):
        eng = _engine(s, M, precision, separate)
        assert (eng.tuning(eng.TUNE_LLG)[1] >> 23) & 1 == int(separate)
        eng.set_moments(e0, np.broadcast_to(s["mmom"], (M, N)))
        b, en = eng.field(with_energy=True)
        eng.set_llg(0.05, 50.0, 1e-16)
        eng.set_noise(U.UASD_NOISE_PHILOX, 17)
        eng.sd_steps(5, 40)
        e, _ = eng.get_moments()
        out.append((b, en, e))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_fused_dm_gather_kagome_batch_bit_identical():
    reps = [CF.kagome_afm(L, J=-1.0, Dz=0.08) for L in (6, 9, 13)]
    j = Bt.join(reps)
    N = j.natom
    rng = np.random.default_rng(82)
    e0 = np.ascontiguousarray(CF.random_unit(N, rng, 1))
    print(f"fused DM kagome batch N={N}", flush=True)
    s = dict(nlist=j.nlist, nlistsize=j.nlistsize, ncoup=j.ncoup, dmlist=j.extra["dmlist"],
             dmlistsize=j.extra["dmlistsize"], dm_vect=j.extra["dm_vect"])
    out = []
    for separate in (False, True):
        eng = _engine(s, 1, "fp32", separate, N=N)
        eng.set_moments(e0, j.mmom[None])
        eng.set_llg(0.2, 20.0, 5e-17)
        eng.set_noise(U.UASD_NOISE_PHILOX, 5)
        eng.sd_steps(5, 30)
        temps = np.array([40.0])
        eng.mc_sweeps("H", temps, 3, 9)
        out.append(eng.get_moments()[0])
    assert np.array_equal(out[0], out[1])
