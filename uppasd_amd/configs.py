"""Standard workloads of the spin-update path, built with the reference's
neighbour-map rules (build_pair_list).  Couplings are given in mRy and
moments in Bohr magnetons, as in UppASD input files.

  bcc_shells   bcc ferromagnet, cubic cell with a 2-atom basis, shells at
               (1/2,1/2,1/2), (1,0,0), (1,1,0), (3/2,1/2,1/2), (1,1,1)
               (8/6/12/24/8 neighbours) under the 48 cubic operations.
  triangular_dmi  triangular monolayer, J to 6 NN, interfacial DMI
               D_ij = D (z x r_ij), uniaxial K along z, field along z.
  bcc_alloy    random binary bcc alloy, NN couplings per species pair.
  kagome_afm   kagome lattice (3 sublattices), NN J and out-of-plane DMI.
"""
from __future__ import annotations

import numpy as np

from .lattice import BCC_SHELLS, Lattice, Shells, bcc, build_pair_list, kagome, triangular

MRY_PER_MUB = 2.179872325e-21 / 9.274009994e-24   # Tesla per (mRy/muB)


def bcc_shells(L, J, moment=2.2, bc=("P", "P", "P")):
    """J: couplings of the first len(J) bcc shells (mRy)."""
    lat = bcc((L, L, L) if np.isscalar(L) else L, bc)
    lines = [(1, 1, 1, 1, *BCC_SHELLS[k], float(J[k])) for k in range(len(J))]
    sh = Shells.from_lines(lat, lines)
    nlist, nsize, ncoup = build_pair_list(lat, sh, 1, np.full((2, 1), moment))
    return dict(lattice=lat, nlist=nlist, nlistsize=nsize, ncoup=ncoup,
                mmom=np.full(lat.natom, float(moment)), lines=lines, sym=1)


def bcc5_couplings():
    """J_k = 1.34 exp(-1.2 (k-1)) mRy, sign alternating from shell 3 on."""
    J = [1.34 * np.exp(-1.2 * k) for k in range(5)]
    return [J[0], J[1], -J[2], J[3], -J[4]]


TRI_BONDS = ((1.0, 0.0, 0.0), (-1.0, 0.0, 0.0), (0.5, np.sqrt(3.0) / 2, 0.0),
             (-0.5, -np.sqrt(3.0) / 2, 0.0), (-0.5, np.sqrt(3.0) / 2, 0.0),
             (0.5, -np.sqrt(3.0) / 2, 0.0))


def triangular_dmi(L, J=1.0, D=0.2, K=0.05, B=0.3, moment=1.0):
    """D, K in mRy; B in mRy/muB (converted to Tesla)."""
    lat = triangular((L, L, 1) if np.isscalar(L) else L)
    jl = [(1, 1, 1, 1, *b, J) for b in TRI_BONDS]
    sh = Shells.from_lines(lat, jl)
    nlist, nsize, ncoup = build_pair_list(lat, sh, 0, np.array([[moment]]))
    dml = [(1, 1, 1, 1, *b, *(D * np.cross([0.0, 0.0, 1.0], b))) for b in TRI_BONDS]
    dsh = Shells.from_lines(lat, dml, hdim=3)
    dlist, dsize, dvec = build_pair_list(lat, dsh, 0, np.array([[moment]]), use_nntype=False)
    N = lat.natom
    fc = MRY_PER_MUB
    taniso = np.ones(N, dtype=np.int32)
    eaniso = np.tile([0.0, 0.0, 1.0], (N, 1))
    kaniso = np.tile([fc * K / (moment * moment), 0.0], (N, 1))
    sb = np.zeros(N)
    return dict(lattice=lat, nlist=nlist, nlistsize=nsize, ncoup=ncoup, dmlist=dlist,
                dmlistsize=dsize, dm_vect=dvec, taniso=taniso, eaniso=eaniso, kaniso=kaniso,
                sb=sb, ext=np.array([0.0, 0.0, B * fc]), mmom=np.full(N, float(moment)),
                jlines=jl, dmlines=dml, sym=0)


def bcc_alloy(L, occupancy, Jaa=1.2, Jab=2.0, Jbb=0.4, ma=2.6, mb=1.7):
    """occupancy: (Natom,) species 1 or 2."""
    lat = bcc((L, L, L) if np.isscalar(L) else L)
    lines = [(1, 1, 1, 1, 0.5, 0.5, 0.5, Jaa), (1, 1, 1, 2, 0.5, 0.5, 0.5, Jab),
             (1, 1, 2, 1, 0.5, 0.5, 0.5, Jab), (1, 1, 2, 2, 0.5, 0.5, 0.5, Jbb)]
    sh = Shells.from_lines(lat, lines, nchmax=2, alloy=True)
    ammom = np.array([[ma, mb], [ma, mb]])
    occ = np.asarray(occupancy, dtype=np.int32)
    nlist, nsize, ncoup = build_pair_list(lat, sh, 1, ammom, achem=occ)
    mm = np.where(occ == 1, ma, mb).astype(np.float64)
    return dict(lattice=lat, nlist=nlist, nlistsize=nsize, ncoup=ncoup, mmom=mm, lines=lines,
                achem=occ, ammom=ammom, sym=1)


def kagome_afm(L, J=-1.0, Dz=0.08, moment=1.0):
    lat0 = kagome((L, L, 1) if np.isscalar(L) else L)
    lat = Lattice(lat0.cell, lat0.bas, [1, 2, 3], lat0.n, lat0.bc)
    h = np.sqrt(3.0) / 4.0
    nb = {1: [(2, (0.5, 0, 0)), (2, (-0.5, 0, 0)), (3, (0.25, h, 0)), (3, (-0.25, -h, 0))],
          2: [(1, (0.5, 0, 0)), (1, (-0.5, 0, 0)), (3, (-0.25, h, 0)), (3, (0.25, -h, 0))],
          3: [(1, (0.25, h, 0)), (1, (-0.25, -h, 0)), (2, (-0.25, h, 0)), (2, (0.25, -h, 0))]}
    jl = [(i, j, 1, 1, *r, J) for i in (1, 2, 3) for j, r in nb[i]]
    sh = Shells.from_lines(lat, jl)
    ones = np.full((3, 1), moment)
    nlist, nsize, ncoup = build_pair_list(lat, sh, 0, ones)
    # out-of-plane DMI with alternating sign around each triangle
    dl = []
    for i in (1, 2, 3):
        for j, r in nb[i]:
            s = 1.0 if (j - i) % 3 == 1 else -1.0
            dl.append((i, j, 1, 1, *r, 0.0, 0.0, s * Dz))
    dsh = Shells.from_lines(lat, dl, hdim=3)
    dlist, dsize, dvec = build_pair_list(lat, dsh, 0, ones, use_nntype=False)
    return dict(lattice=lat, nlist=nlist, nlistsize=nsize, ncoup=ncoup, dmlist=dlist,
                dmlistsize=dsize, dm_vect=dvec, mmom=np.full(lat.natom, float(moment)),
                jlines=jl, dmlines=dl, sym=0)


def random_unit(n, rng, m=1):
    v = rng.normal(size=(m, n, 3))
    return v / np.linalg.norm(v, axis=2, keepdims=True)
