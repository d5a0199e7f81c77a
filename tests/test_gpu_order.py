"""Basis-major device site order: every result equals the reference-order
run bit for bit (LLG with Philox noise keyed by the reference site, MC with
colour classes built in reference order, fields and reductions)."""
import numpy as np
import pytest

import uppasd_amd as U
from uppasd_amd import configs as CF

pytestmark = pytest.mark.gpu


def _eng(sysd, M, order, precision="fp64"):
    N = sysd["lattice"].natom
    eng = U.Engine(N, M, precision, site_order=order)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    if "dmlist" in sysd:
        eng.set_dm(sysd["dmlist"], sysd["dmlistsize"], sysd["dm_vect"])
    if "taniso" in sysd:
        eng.set_anisotropy(sysd["taniso"], sysd["eaniso"], sysd["kaniso"], sysd["sb"])
    return eng


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("alg", [1, 5])
def test_llg_same_in_both_orders(precision, alg):
    rng = np.random.default_rng(80)
    sysd = CF.bcc_shells(6, CF.bcc5_couplings())
    N, M = sysd["lattice"].natom, 1
    e0 = np.ascontiguousarray(CF.random_unit(N, rng, M))
    lam = rng.uniform(0.05, 0.2, N)     # per-site damping: a per-site array in both orders
    out = []
    for order in (1, 2):
        eng = _eng(sysd, M, order, precision)
        assert eng.index_bytes()[0] == 2      # packed int16 entries in both orders
        eng.set_moments(e0, sysd["mmom"][None])
        eng.set_llg(lam, 300.0, 1e-16)
        eng.set_noise(U.UASD_NOISE_PHILOX, 8)
        eng.sd_steps(alg, 25)
        beff, en = eng.field()
        out.append((eng.get_moments()[0], beff, en))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert abs(out[0][2] - out[1][2]) <= 1e-13 * abs(out[0][2])


def test_dropin_and_kagome_same_in_both_orders():
    rng = np.random.default_rng(81)
    sysd = CF.kagome_afm(12, J=-1.0, Dz=0.08)
    N, M = sysd["lattice"].natom, 2
    emom = np.ascontiguousarray(CF.random_unit(N, rng, M))
    emomM = np.ascontiguousarray(emom * sysd["mmom"][None, :, None])
    res = []
    for order in (1, 3):
        eng = _eng(sysd, M, order)
        assert eng.index_bytes() == (2, 2)
        b, e, b1, b2 = eng.effective_field(emomM, parts=True)
        eng.set_moments(emom, np.broadcast_to(sysd["mmom"], (M, N)))
        eng.set_llg(0.2, 50.0, 5e-17)
        normals = np.random.default_rng(5).normal(size=(M, N, 3))
        eng.set_noise(U.UASD_NOISE_INJECTED, 0, normals)
        eng.sd_steps(5, 3)
        res.append((b, b1, b2, eng.get_moments()[0], e))
    for x, y in zip(res[0][:4], res[1][:4]):
        assert np.array_equal(x, y)
    assert abs(res[0][4] - res[1][4]) <= 1e-13 * abs(res[0][4])


@pytest.mark.parametrize("mode", ["M", "H"])
def test_mc_same_in_both_orders(mode):
    rng = np.random.default_rng(82)
    sysd = CF.bcc_shells(6, CF.bcc5_couplings())
    N, M = sysd["lattice"].natom, 2
    e0 = np.ascontiguousarray(CF.random_unit(N, rng, M))
    out = []
    for order in (1, 2):
        eng = _eng(sysd, M, order)
        eng.set_moments(e0, np.broadcast_to(sysd["mmom"], (M, N)))
        eng.mc_sweeps(mode, [400.0, 900.0], 6, 12)
        ncol, col = eng.coloring()
        out.append((eng.get_moments()[0], ncol, col))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1] and np.array_equal(out[0][2], out[1][2])


def test_site_order_rules():
    from uppasd_amd import _lib as B
    sysd = CF.bcc_shells(3, CF.bcc5_couplings())
    N = sysd["lattice"].natom
    with pytest.raises(U.UasdError):
        U.Engine(N, 1, site_order=5)           # does not divide natom
    eng = _eng(sysd, 1, 1)                     # exchange already uploaded
    with pytest.raises(U.UasdError):
        B.check(B.lib().uasd_set_site_order(eng._h, 2), "uasd_set_site_order")
