"""Host side of batches: the sublattice-major renumbering keeps every bond,
its coupling and its DM vector, and the joined lattice's offsets."""
import numpy as np

from uppasd_amd import batch as Bt
from uppasd_amd import configs as CF


def test_sublattice_major_keeps_every_bond():
    d = CF.kagome_afm(5)
    r = Bt.sublattice_major(d, 3)
    N = len(d["nlistsize"])
    nc = N // 3

    def new(i):
        return (i % 3) * nc + i // 3
    for i in range(N):
        n = d["nlistsize"][i]
        assert r["nlistsize"][new(i)] == n
        assert list(r["nlist"][new(i), :n]) == [new(x - 1) + 1 for x in d["nlist"][i, :n]]
        assert np.array_equal(r["ncoup"][new(i)], d["ncoup"][i])
        assert list(r["dmlist"][new(i), :n]) == [new(x - 1) + 1 for x in d["dmlist"][i, :n]]
        assert np.array_equal(r["dm_vect"][new(i)], d["dm_vect"][i])
    # sublattice b is the range [b*nc, (b+1)*nc)
    for b in range(3):
        assert np.all(np.arange(b * nc, (b + 1) * nc) == [new(c * 3 + b) for c in range(nc)])


def test_join_offsets_and_shifted_lists():
    a, b = CF.kagome_afm(3), CF.kagome_afm(4)
    j = Bt.join([a, b])
    na = len(a["nlistsize"])
    assert list(j.offsets) == [0, na, na + len(b["nlistsize"])]
    w = b["nlist"].shape[1]
    assert np.array_equal(j.nlist[na:, :w], np.where(b["nlist"] > 0, b["nlist"] + na, 0))
