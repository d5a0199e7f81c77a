"""LDS-tiled stencil kernels: identical results to the list-gather kernels for
every tile shape that fits, in both device site orders, for 3D and 2D
lattices with exchange, DMI, anisotropy and Zeeman terms, with halos that
wrap around small lattices and tiles that overhang the lattice edge; lists
that are not a stencil (vacuum boundaries) are detected and keep the list
kernels."""
import numpy as np
import pytest

import uppasd_amd as U
from uppasd_amd import configs as CF

pytestmark = pytest.mark.gpu

TILE = lambda k, pipe=0: 256 | (k << 18) | (pipe << 22)


def _eng(sysd, order, precision="fp64", M=1):
    lat = sysd["lattice"]
    eng = U.Engine(lat.natom, M, precision, site_order=order)
    eng.set_lattice_dims(lat.NA, *lat.n)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    if "dmlist" in sysd:
        eng.set_dm(sysd["dmlist"], sysd["dmlistsize"], sysd["dm_vect"])
    if "taniso" in sysd:
        eng.set_anisotropy(sysd["taniso"], sysd["eaniso"], sysd["kaniso"], sysd["sb"])
    if "ext" in sysd:
        eng.set_external_field(sysd["ext"])
    return eng


def _run(sysd, e0, order, tile, alg, precision="fp64", steps=12, pipe=0):
    M = e0.shape[0]
    eng = _eng(sysd, order, precision, M)
    eng.set_moments(e0, np.broadcast_to(sysd["mmom"], (M, sysd["lattice"].natom)))
    eng.set_llg(0.1, 200.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_PHILOX, 17)
    eng.set_tuning(eng.TUNE_LLG, 4, TILE(tile, pipe) if tile else 256)
    eng.sd_steps(alg, steps)
    return eng.get_moments()[0], eng


CASES = {
    "bcc5_L5": lambda: CF.bcc_shells(5, CF.bcc5_couplings()),      # halo wraps (R=1, L=5)
    "bcc5_L10": lambda: CF.bcc_shells(10, CF.bcc5_couplings()),    # partial tiles
    "tri_dmi": lambda: CF.triangular_dmi(20),                      # 2D, DM + K + B
    "kagome": lambda: CF.kagome_afm(9, J=-1.0, Dz=0.08),           # NA = 3, DM
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("alg", [1, 4, 5, 6, 7])
def test_tile_equals_list_kernels(case, alg):
    sysd = CASES[case]()
    lat = sysd["lattice"]
    rng = np.random.default_rng(90)
    e0 = np.ascontiguousarray(CF.random_unit(lat.natom, rng, 1))
    ref, eng = _run(sysd, e0, 1, 0, alg)
    has, radius = eng.stencil()
    assert has, case
    shapes = [k for k, sh in enumerate(U.engine.TILE_SHAPES, 1) if (lat.n[2] > 1) == (sh[2] > 1)]
    for order in (1, lat.NA):
        for k in shapes:
            for pipe in (0, 1):
                got, _ = _run(sysd, e0, order, k, alg, pipe=pipe)
                assert np.array_equal(got, ref), (case, order, k, pipe)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("M", [2, 3])
@pytest.mark.parametrize("alg", [1, 4, 5, 6, 7])
def test_tile_several_ensembles_equals_list_kernels(alg, M, precision):
    """M ensembles of one lattice: the tile kernel runs (tile, ensemble)
    units, staging each ensemble's halo from the interleaved quads, with the
    per-ensemble noise and fields; every shape, one unit per workgroup and
    the resident loop (whose runs of units cross from one ensemble and tile
    to the next), both site orders, give the list kernels' moments bit for
    bit."""
    for case in ("bcc5_L10", "tri_dmi"):
        sysd = CASES[case]()
        lat = sysd["lattice"]
        rng = np.random.default_rng(95 + M)
        e0 = np.ascontiguousarray(CF.random_unit(lat.natom, rng, M))
        ref, eng = _run(sysd, e0, 1, 0, alg, precision)
        assert eng.stencil()[0], case
        assert np.abs(ref[0] - ref[1]).max() > 1e-3      # the ensembles differ
        shapes = [k for k, sh in enumerate(U.engine.TILE_SHAPES, 1)
                  if (lat.n[2] > 1) == (sh[2] > 1)][:3]
        for order in (1, lat.NA):
            for k in shapes:
                for pipe in (0, 1):
                    got, e = _run(sysd, e0, order, k, alg, precision, pipe=pipe)
                    assert e.llg_tile() == k      # the tile kernel ran
                    assert np.array_equal(got, ref), (case, order, k, pipe)


@pytest.mark.parametrize("alg", [1, 4, 5, 6, 7])
def test_resident_workgroups_loop_over_tiles(alg):
    """A lattice with more tiles than resident workgroups (bcc 70^3 x 2,
    partial tiles at the upper edges): with the resident-workgroup launch
    every workgroup runs through several tiles, prefetching the next halo
    while it computes; the result equals the list-gather kernels bit for bit."""
    sysd = CF.bcc_shells(70, CF.bcc5_couplings())
    lat = sysd["lattice"]
    rng = np.random.default_rng(94)
    e0 = np.ascontiguousarray(CF.random_unit(lat.natom, rng, 1))
    ref, _ = _run(sysd, e0, lat.NA, 0, alg, steps=3)
    for k in (1, 4):                       # (8,8,4): 1458 tiles, (16,4,4): 1620 tiles
        got, _ = _run(sysd, e0, lat.NA, k, alg, steps=3, pipe=1)
        assert np.array_equal(got, ref), k


@pytest.mark.parametrize("alg", [1, 4, 5, 6, 7])
def test_resident_loop_over_units_of_several_ensembles(alg):
    """Three ensembles of bcc 40^3 x 2 (250 tiles of (8,8,4) cells or 300
    of (16,4,4), partial at the upper x edge; 750 or 900 (tile, ensemble)
    units against 512 resident workgroups, two per CU as the FP64 tiles'
    LDS allows): runs of two consecutive units, so a workgroup moves from
    one ensemble to the next of a tile, or from a tile's last ensemble to
    the next tile's first, prefetching that unit's halo; the moments equal
    the list kernels' bit for bit."""
    sysd = CF.bcc_shells(40, CF.bcc5_couplings())
    lat = sysd["lattice"]
    rng = np.random.default_rng(96)
    e0 = np.ascontiguousarray(CF.random_unit(lat.natom, rng, 3))
    ref, _ = _run(sysd, e0, lat.NA, 0, alg, steps=3)
    for k in (1, 4):
        got, e = _run(sysd, e0, lat.NA, k, alg, steps=3, pipe=1)
        assert e.llg_tile() == k
        assert np.array_equal(got, ref), k


@pytest.mark.parametrize("alg", [1, 4, 5, 6, 7])
def test_tile_fp32_equals_list_fp32(alg):
    sysd = CF.triangular_dmi(24)
    rng = np.random.default_rng(91)
    e0 = np.ascontiguousarray(CF.random_unit(sysd["lattice"].natom, rng, 1))
    ref, _ = _run(sysd, e0, 1, 0, alg, "fp32")
    got, _ = _run(sysd, e0, 1, 5, alg, "fp32")
    assert np.array_equal(got, ref)


def test_no_stencil_with_vacuum_boundary():
    sysd = CF.bcc_shells(6, CF.bcc5_couplings(), bc=("P", "P", "0"))
    rng = np.random.default_rng(92)
    e0 = np.ascontiguousarray(CF.random_unit(sysd["lattice"].natom, rng, 1))
    ref, eng = _run(sysd, e0, 1, 0, 1)
    assert not eng.stencil()[0]
    got, e = _run(sysd, e0, 1, 1, 1)   # tile requested: falls back to the list kernels
    assert e.llg_tile() == 0
    assert np.array_equal(got, ref)


def test_autotune_may_pick_tile_and_keeps_results():
    sysd = CF.bcc_shells(16, CF.bcc5_couplings())
    rng = np.random.default_rng(93)
    e0 = np.ascontiguousarray(CF.random_unit(sysd["lattice"].natom, rng, 1))
    ref, _ = _run(sysd, e0, 2, 0, 1, steps=6)
    eng = _eng(sysd, 2)
    eng.set_moments(e0, sysd["mmom"][None])
    eng.set_llg(0.1, 200.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_PHILOX, 17)
    eng.autotune_llg(1, steps=2)
    eng.sd_steps(1, 6)
    assert np.array_equal(eng.get_moments()[0], ref)
