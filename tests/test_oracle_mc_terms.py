"""Pins the oracle's Monte Carlo energy terms beyond exchange by hand-computed
cases, and the anisotropy set-up by what its scaling must preserve.

* DM (calculate_energy, montecarlo_common.f90:609-625): two sites, one DM
  bond i -> j.  With unit moments the current-state term is -D.(m_i x m_j).
  With |m_i| = 2, |m_j| = 3 the reference mixes the unit vector emom into
  two of the three current-state products (lines 612 and 615: emom(3,i) in
  the D_x term, emom(1,i) in the D_z term), so E_c differs from
  -D.(m_i x m_j) by an amount worked out below by hand for each component;
  the D_y term has no such mix.  The trial term uses the full trial moment.
* Zeeman (858-863): dE = -mub (B.(t - c)) for trial moment t, current c.
* setup_anisotropies (hamiltonianinit.f90:857-976) scales the input K1, K2
  (mRy) by fc/m^2, fc/m^4 (uniaxial) and fc/m^4, fc/m^6 (cubic), fc = mry/mub,
  and normalises the axis.  The MC energy (a separate routine) of the moment
  m e then equals K1 (e.a)^2 + K2 (e.a)^4, resp. -K1 (ex^2 ey^2 + ...) -
  K2 ex^2 ey^2 ez^2, in mRy, whatever |m| is: a wrong power of m fails at
  |m| != 1.  Type 7 scales both parts as uniaxial (m^-2, m^-4), and its MC
  energy adds the cubic part with sb and the opposite sign of type 2
  (montecarlo_common.f90:571-590), so that part is sb m^2 (K1 c4 + K2 c6) mRy
  with c4 = ex^2 ey^2 + ey^2 ez^2 + ez^2 ex^2, c6 = ex^2 ey^2 ez^2: the
  reference's own scaling, pinned so that a change in either routine shows.
"""
import ctypes

import numpy as np
import pytest

MUB = 9.274009994e-24
MRY = 2.179872325e-21


def _dm_pair_ham(O, D):
    nl = np.array([[2], [1]], dtype=np.int32)
    ns = np.zeros(2, dtype=np.int32)               # no exchange
    nc = np.zeros((2, 1))
    dml = np.array([[2], [1]], dtype=np.int32)
    dms = np.array([1, 0], dtype=np.int32)         # only the bond i=1 -> j=2
    dmv = np.zeros((2, 1, 3))
    dmv[0, 0] = D
    return O.OracleHam(2, nl, ns, nc, dmlist=dml, dmlistsize=dms, dm_vect=dmv)


def _de(O, ham, e, m, newmom, ext=(0.0, 0.0, 0.0), site=0):
    emom = np.ascontiguousarray(np.asarray(e, dtype=np.float64)[None])
    mm = np.ascontiguousarray(np.asarray(m, dtype=np.float64)[None])
    emomM = np.ascontiguousarray(emom * mm[..., None])
    return O.lib().or_mc_energy_change(ctypes.byref(ham.h), emom.shape[1], emomM, emom, mm, site,
                                       0, np.ascontiguousarray(newmom, dtype=np.float64),
                                       np.ascontiguousarray(ext, dtype=np.float64))


d = 0.37


# (D, e_i, e_j, trial direction, dE/mub with unit moments, dE/mub with |m_i| = 2, |m_j| = 3)
DM_CASES = [
    # D along x: E_c = -d (c_y n_z - u_z n_y).
    #   unit: c = u = (0,.6,.8), n = (0,.8,-.6): -d(.6*-.6 - .8*.8) = d; trial -e: E_t = -d; dE = -2d
    #   m: c = (0,1.2,1.6), u_z = .8, n = (0,2.4,-1.8): E_c = -d(1.2*-1.8 - .8*2.4) = 4.08 d,
    #      t = (0,-1.2,-1.6): E_t = -d(-1.2*-1.8 - (-1.6)*2.4) = -6 d; dE = -10.08 d
    ((d, 0, 0), (0, .6, .8), (0, .8, -.6), (0, -.6, -.8), -2 * d, -10.08 * d),
    # D along y: E_c = -d (c_z n_x - c_x n_z), no unit-vector mix.
    #   unit: c = (.6,0,.8), n = (.8,0,-.6): -d(.64 + .36) = -d; trial -e: +d; dE = 2d
    #   m: c = (1.2,0,1.6), n = (2.4,0,-1.8): -d(3.84 + 2.16) = -6d; E_t = 6d; dE = 12 d
    ((0, d, 0), (.6, 0, .8), (.8, 0, -.6), (-.6, 0, -.8), 2 * d, 12 * d),
    # D along z: E_c = -d (u_x n_y - c_y n_x).
    #   unit: c = u = (.6,.8,0), n = (.8,-.6,0): -d(-.36 - .64) = d; trial -e: -d; dE = -2d
    #   m: u_x = .6, c = (1.2,1.6,0), n = (2.4,-1.8,0): E_c = -d(.6*-1.8 - 1.6*2.4) = 4.92 d,
    #      t = (-1.2,-1.6,0): E_t = -d(-1.2*-1.8 - (-1.6)*2.4) = -6 d; dE = -10.92 d
    ((0, 0, d), (.6, .8, 0), (.8, -.6, 0), (-.6, -.8, 0), -2 * d, -10.92 * d),
]


@pytest.mark.parametrize("case", range(len(DM_CASES)))
def test_mc_dm_energy_hand_cases(O, case):
    D, ei, ej, tr, want_unit, want_m = DM_CASES[case]
    ham = _dm_pair_ham(O, np.array(D, dtype=np.float64))
    e = np.array([ei, ej], dtype=np.float64)
    got1 = _de(O, ham, e, [1.0, 1.0], tr) / MUB
    assert got1 == pytest.approx(want_unit, rel=1e-13)
    gotm = _de(O, ham, e, [2.0, 3.0], tr) / MUB
    assert gotm == pytest.approx(want_m, rel=1e-13)
    # a perpendicular trial through the other branch of the cross products
    tr2 = np.cross(ei, ej)
    got2 = _de(O, ham, e, [1.0, 1.0], tr2) / MUB
    # unit moments: E = -D.(m_i x m_j) for both states
    want2 = -np.dot(D, np.cross(tr2, ej)) + np.dot(D, np.cross(ei, ej))
    assert got2 == pytest.approx(want2, rel=1e-13, abs=1e-15)


def test_mc_zeeman_energy_hand_case(O):
    nl = np.zeros((1, 1), dtype=np.int32)
    ham = O.OracleHam(1, nl, np.zeros(1, dtype=np.int32), np.zeros((1, 1)))
    B = (0.5, -2.0, 3.0)
    e = np.array([[0.0, 0.6, 0.8]])
    tr = (1.0, 0.0, 0.0)
    # c = 2.5 (0,.6,.8) = (0,1.5,2); t = (2.5,0,0): E_c = -(-3 + 6) = -3, E_t = -1.25; dE = 1.75
    assert _de(O, ham, e, [2.5], tr, ext=B) / MUB == pytest.approx(1.75, rel=1e-14)


def _aniso_energy_mry(O, taniso, eaniso, kaniso, sb, m, e_dir):
    """MC energy of moment m*e_dir relative to a state with zero anisotropy
    energy, in mRy: dE(trial e_dir from a reference direction) / mry."""
    ham = O.OracleHam(1, np.zeros((1, 1), dtype=np.int32), np.zeros(1, dtype=np.int32),
                      np.zeros((1, 1)), taniso=np.array([taniso], dtype=np.int32),
                      eaniso=np.ascontiguousarray(eaniso.reshape(1, 3)),
                      kaniso=np.ascontiguousarray(kaniso.reshape(1, 2)),
                      sb=np.array([sb]))
    # from e0 = x: the cubic terms vanish there, and so does the uniaxial term
    # for an axis perpendicular to x
    e0 = np.array([[1.0, 0.0, 0.0]])
    return _de(O, ham, e0, [m], e_dir) / MRY


@pytest.mark.parametrize("m", [0.7, 1.0, 2.3])
def test_setup_anisotropy_scaling_gives_energy_in_mry(O, m):
    """or_setup_anisotropies + or_mc_energy_change: the anisotropy energy of
    the moment m e in mRy is the input's K1, K2 expression, independent of m
    (types 1 and 2); type 7 adds the reference's m^2-scaled cubic part."""
    lat = O.LatticeSpec(np.eye(3), [[0, 0, 0]], [1], (1, 1, 1))
    axis_in = np.array([0.0, 3.0, 4.0])                 # not normalised in the input
    a = axis_in / 5.0
    K1, K2, sbv = -0.37, 0.11, 1.9
    rng = np.random.default_rng(5)
    for ty in (1, 2, 7):
        aniso = np.array([[K1, K2, *axis_in, sbv]])
        taniso, eaniso, kaniso, sb = O.setup_anisotropies(lat, [ty], aniso, [m])
        assert np.allclose(eaniso[0], a / np.sqrt(1.0 + 1e-15 / 25.0), rtol=1e-15)
        for _ in range(4):
            ed = rng.normal(size=3)
            ed /= np.linalg.norm(ed)
            got = _aniso_energy_mry(O, taniso[0], eaniso[0], kaniso[0], sb[0], m, ed)
            t, t0 = ed.dot(eaniso[0]), eaniso[0][0]           # e0 = x: e0.a = a_x = 0
            c4 = ed[0] ** 2 * ed[1] ** 2 + ed[1] ** 2 * ed[2] ** 2 + ed[2] ** 2 * ed[0] ** 2
            c6 = ed[0] ** 2 * ed[1] ** 2 * ed[2] ** 2
            uni = K1 * t ** 2 + K2 * t ** 4 - (K1 * t0 ** 2 + K2 * t0 ** 4)
            if ty == 1:
                want = uni
            elif ty == 2:
                want = -K1 * c4 - K2 * c6
            else:
                want = uni + sbv * m ** 2 * (K1 * c4 + K2 * c6)
            assert got == pytest.approx(want, rel=1e-12, abs=1e-15), (ty, m)


def test_fp32_trial_moves_from_one_philox_block(O):
    """The FP32 engine's Metropolis update draws everything from one Philox
    block (or_trial_move_f32 restates it).  Over 60,000 (site, ensemble)
    draws: the three moves come in thirds; the uniform-direction move is
    uniform on the sphere (<z> = 0, <z^2> = 1/3, azimuth uniform); the
    Gaussian step from e = 0 points in an isotropic direction (<n> = 0,
    <n_i n_j> = delta_ij / 3, so g2 with its 16-bit angle matches g0 and g1);
    the flip reverses; the acceptance uniform v3 is uniform within each move
    type (mean 1/2, variance 1/12), i.e. independent of the move choice."""
    L = O.lib()
    n = 60000
    e = np.array([0.6, 0.0, 0.8])
    kinds = {0: [], 1: [], 2: []}
    acc = {0: [], 1: [], 2: []}
    w = np.zeros(4, np.uint32)
    out = np.zeros(3)
    for t in range(n):
        L.or_counter_words(77, 0x21, 5, t // 8, t % 8, w)
        b = (int(w[0] & 0xff) << 8) | int(w[1] & 0xff)
        k = (3 * b) >> 16
        L.or_trial_move_f32(np.zeros(3) if k == 1 else e, 1.0, w, out)
        kinds[k].append(out.copy())
        acc[k].append(L.or_u24(int(w[3])))
    counts = np.array([len(kinds[k]) for k in range(3)])
    assert np.all(np.abs(counts - n / 3) < 4 * np.sqrt(n * (1 / 3) * (2 / 3))), counts
    u = np.array(kinds[0])
    sig = 4 / np.sqrt(len(u))
    assert np.allclose(np.linalg.norm(u, axis=1), 1.0, atol=1e-6)
    assert abs(u[:, 2].mean()) < sig and abs((u[:, 2] ** 2).mean() - 1 / 3) < sig
    phi = np.arctan2(u[:, 1], u[:, 0])
    assert abs(np.cos(phi).mean()) < sig and abs(np.sin(phi).mean()) < sig
    g = np.array(kinds[1])
    assert np.allclose(np.linalg.norm(g, axis=1), 1.0, atol=1e-6)
    assert np.all(np.abs(g.mean(axis=0)) < sig)
    cov = g.T @ g / len(g)
    assert np.allclose(cov, np.eye(3) / 3, atol=sig), cov
    f = np.array(kinds[2])
    assert np.allclose(f, -e[None, :], atol=1e-7)
    for k in range(3):
        a = np.array(acc[k])
        s = 4 / np.sqrt(len(a))
        assert abs(a.mean() - 0.5) < s and abs(a.var() - 1 / 12) < s, (k, a.mean(), a.var())
