"""Pins the oracle's term-resolved energies (or_energy_terms, restating
calc_energy, energy.f90:183-366) against a separately restated routine and
the mathematics.

- The field routine's energy (hamiltonianactions.f90:245-250, pinned by the
  golden runs) equals the sum of the terms when no site has type-7
  anisotropy; with type 7 the field routine counts the sb*cubic part with
  factor 1 and calc_energy with 0.5, so they differ by exactly
  0.5 * sum m.beff_q.
- With every other term switched off, each term alone equals the field
  routine's energy, so a term booked in the wrong column fails.
- The exchange term is -1/2 sum_ij J_ij m_i.m_j over the lists (numpy), the
  Zeeman term -sum m.B.
The measurement regression case (tests/test_output.py) pins Exc and Zeeman
against the reference's printed values.
"""
import numpy as np

MUB = 9.274009994e-24
MRY = 2.179872325e-21
FC = MUB / MRY


def _lists(rng, N, W):
    nl = np.zeros((N, W), dtype=np.int32)
    for i in range(N):
        nl[i] = rng.choice(np.delete(np.arange(N), i), size=W, replace=False) + 1
    return nl, np.full(N, W, dtype=np.int32)


def _ham(O, rng, N, J=True, dm=True, aniso=(1, 2), W=6):
    nl, ns = _lists(rng, N, W)
    nc = rng.normal(size=(N, W)) if J else np.zeros((N, W))
    kw = {}
    if dm:
        dl, ds = _lists(rng, N, 4)
        kw.update(dmlist=dl, dmlistsize=ds, dm_vect=np.ascontiguousarray(rng.normal(size=(N, 4, 3))))
    if aniso:
        ea = rng.normal(size=(N, 3))
        ea /= np.linalg.norm(ea, axis=1, keepdims=True)
        kw.update(taniso=np.asarray(rng.choice(aniso, size=N), dtype=np.int32),
                  eaniso=np.ascontiguousarray(ea),
                  kaniso=np.ascontiguousarray(rng.normal(size=(N, 2))),
                  sb=np.ascontiguousarray(rng.uniform(0.2, 1.5, size=N)))
    return O.OracleHam(N, nl, ns, nc, **kw), nl, nc


def _state(rng, M, N):
    e = rng.normal(size=(M, N, 3))
    e /= np.linalg.norm(e, axis=2, keepdims=True)
    return np.ascontiguousarray(e * rng.uniform(0.5, 2.5, size=(M, N, 1)))


def test_terms_sum_to_field_energy(O):
    rng = np.random.default_rng(3)
    N, M = 40, 2
    ham, _, _ = _ham(O, rng, N)
    for k in range(M):
        emomM = _state(rng, 1, N)
        ext = rng.normal(size=(1, N, 3))
        t = ham.energy_terms(emomM, ext)
        _, en = ham.field(emomM, ext)
        assert np.isclose(t.sum() * N * FC, en, rtol=1e-12)


def test_each_term_alone(O):
    rng = np.random.default_rng(4)
    N = 32
    emomM = _state(rng, 1, N)
    cases = {0: dict(J=True, dm=False, aniso=()), 1: dict(J=False, dm=False, aniso=(1, 2)),
             2: dict(J=False, dm=True, aniso=())}
    for col, kw in cases.items():
        ham, _, _ = _ham(O, rng, N, **kw)
        t = ham.energy_terms(emomM)[0]
        _, en = ham.field(emomM)
        assert np.isclose(t[col] * N * FC, en, rtol=1e-12), col
        assert np.all(np.delete(t, col) == 0.0), col
    ham, _, _ = _ham(O, rng, N, J=False, dm=False, aniso=())
    ext = rng.normal(size=(1, N, 3))
    t = ham.energy_terms(emomM, ext)[0]
    assert np.isclose(t[3], -np.sum(emomM * ext) / N, rtol=1e-13)
    assert np.all(t[:3] == 0.0)


def test_exchange_term_is_pair_sum(O):
    rng = np.random.default_rng(5)
    N, M = 50, 3
    ham, nl, nc = _ham(O, rng, N, dm=False, aniso=())
    emomM = _state(rng, M, N)
    t = ham.energy_terms(emomM)
    for k in range(M):
        m = emomM[k]
        want = -0.5 * np.sum(nc * np.einsum("ic,iwc->iw", m, m[nl - 1])) / N
        assert np.isclose(t[k, 0], want, rtol=1e-12)


def test_type7_counts_half_of_the_cubic_part(O):
    rng = np.random.default_rng(6)
    N = 30
    ham, _, _ = _ham(O, rng, N, J=False, dm=False, aniso=(7,))
    emomM = _state(rng, 1, N)
    t = ham.energy_terms(emomM)[0]
    _, en, _, b2 = ham.field_parts(emomM)          # beff2 = sb * cubic (no field)
    bq = float(np.sum(emomM * b2))
    assert abs(bq) > 1e-3
    assert np.isclose(en, (t[1] * N - 0.5 * bq) * FC, rtol=1e-12)
