"""Lattice and pair-Hamiltonian set-up on the host side.

``Shells.from_lines`` restates read_exchange / read_dmdata
(source/Input/inputhandler_ext.f90:444-618, 1095-1251): bond vectors closer
than 1e-5 (squared) to an earlier one of the same atom type merge into one
shell and the later value replaces the earlier.  ``build_pair_list`` calls the
library's bit-exact restatement of setup_nm + setup_neighbour_hamiltonian.

Array conventions: numpy C-order with the Fortran dimensions reversed, e.g.
nlist (Natom, max_no_neigh) == ham%nlist(max_no_neigh, Natom), 1-based.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as B


@dataclass
class Lattice:
    cell: np.ndarray          # (3,3) rows C1, C2, C3 in units of alat
    bas: np.ndarray           # (NA,3) Cartesian basis positions
    atype: np.ndarray         # (NA,) 1-based atom types
    n: tuple                  # (N1, N2, N3) repetitions
    bc: tuple = ("P", "P", "P")

    def __post_init__(self):
        self.cell = np.asarray(self.cell, dtype=np.float64).reshape(3, 3)
        self.bas = np.ascontiguousarray(np.asarray(self.bas, dtype=np.float64).reshape(-1, 3))
        self.atype = np.ascontiguousarray(np.asarray(self.atype, dtype=np.int32))
        self.n = tuple(int(x) for x in self.n)

    @property
    def NA(self) -> int:
        return self.bas.shape[0]

    @property
    def NT(self) -> int:
        return int(self.atype.max())

    @property
    def natom(self) -> int:
        return self.NA * self.n[0] * self.n[1] * self.n[2]


@dataclass
class Shells:
    nn: np.ndarray        # (NT,)
    redcoord: np.ndarray  # (NT, max_no_shells, 3)
    nntype: np.ndarray    # (NT, max_no_shells)
    jc: np.ndarray        # (nchmax, nchmax, max_no_shells, NT, hdim) == Fortran jc(hdim,NT,shell,ichem,jchem)
    hdim: int
    nchmax: int

    @classmethod
    def from_lines(cls, lat: Lattice, lines, hdim: int = 1, nchmax: int = 1,
                   alloy: bool = False, maptype: int = 1) -> "Shells":
        """lines: (isite, jsite, ichem, jchem, rx, ry, rz, value[, value2, value3]).

        maptype 1: (rx, ry, rz) is the Cartesian bond vector.  maptype 2
        ("bgfm"): it counts cell vectors, and the bond vector is
        bas_j - bas_i + C1*rx + C2*ry + C3*rz, evaluated component by
        component in that order (read_exchange_getNeighVec,
        inputhandler_ext.f90:1039-1043)."""
        if maptype not in (1, 2):
            raise ValueError("maptype must be 1 or 2")
        tol = 1.0e-5
        NT = lat.NT
        red = [[] for _ in range(NT)]
        ntyp = [[] for _ in range(NT)]
        vals = [[] for _ in range(NT)]   # list of dict (ichem,jchem) -> value tuple
        for ln in lines:
            isite, jsite = int(ln[0]), int(ln[1])
            ic, jcm = (int(ln[2]), int(ln[3])) if alloy else (1, 1)
            r = np.array([float(x) for x in ln[4:7]])
            if maptype == 2:
                c, bi, bj = lat.cell, lat.bas[isite - 1], lat.bas[jsite - 1]
                r = np.array([bj[a] - bi[a] + c[0, a] * r[0] + c[1, a] * r[1] + c[2, a] * r[2]
                              for a in range(3)])
            v = tuple(float(x) for x in ln[7:7 + hdim])
            it = int(lat.atype[isite - 1]) - 1
            unique = True
            for s, rr in enumerate(red[it]):
                d = r - rr
                if d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < tol:
                    unique = False
                    vals[it][s][(ic, jcm)] = v
            if unique:
                red[it].append(r)
                ntyp[it].append(int(lat.atype[jsite - 1]))
                vals[it].append({(ic, jcm): v})
        nn = np.array([len(x) for x in red], dtype=np.int32)
        ms = max(1, int(nn.max()))
        redcoord = np.zeros((NT, ms, 3))
        nntype = np.zeros((NT, ms), dtype=np.int32)
        jc = np.zeros((nchmax, nchmax, ms, NT, hdim))
        for t in range(NT):
            for s in range(nn[t]):
                redcoord[t, s] = red[t][s]
                nntype[t, s] = ntyp[t][s]
                for (ic, jcm), v in vals[t][s].items():
                    jc[jcm - 1, ic - 1, s, t, :] = v
        return cls(nn=nn, redcoord=redcoord, nntype=nntype, jc=jc, hdim=hdim, nchmax=nchmax)


def build_pair_list(lat: Lattice, shells: Shells, sym: int, ammom, achem=None,
                    use_nntype: bool = True):
    """Neighbour list and couplings exactly as the reference builds them.

    ammom: (NA, nchmax) moment magnitudes; achem: (Natom,) 1-based chemistry
    for random alloys (None otherwise).  Returns (nlist (N,mnn) 1-based,
    nlistsize (N,), ncoup (N,mnn) or (N,mnn,3) for hdim=3).
    """
    L = B.lib()
    c1, c2, c3 = (np.ascontiguousarray(lat.cell[k]) for k in range(3))
    bc = (("".join(lat.bc)) + "\0").encode()
    ms = shells.redcoord.shape[1]
    # Fortran (NT, max_no_shells, 3) -> flat column-major
    red_f = np.ascontiguousarray(shells.redcoord.transpose(2, 1, 0)).reshape(-1)
    nnt_f = np.ascontiguousarray(shells.nntype.T).reshape(-1) if use_nntype else None
    jc_f = np.ascontiguousarray(shells.jc).reshape(-1)
    am = np.ascontiguousarray(np.asarray(ammom, dtype=np.float64).reshape(lat.NA, shells.nchmax).T).reshape(-1)
    ach = B.i32(achem)
    mnn = C.c_int(0)
    args = [lat.NA, lat.NT, *lat.n, B.ptr(c1), B.ptr(c2), B.ptr(c3), B.ptr(lat.bas),
            B.ptr(lat.atype), C.c_char_p(bc), sym, ms, B.ptr(shells.nn), B.ptr(red_f),
            B.ptr(nnt_f), shells.hdim, shells.nchmax, B.ptr(jc_f), B.ptr(am), B.ptr(ach)]
    B.check(L.uasd_build_pair_list(*args, C.byref(mnn), None, None, None), "uasd_build_pair_list")
    N, W = lat.natom, mnn.value
    nlist = np.zeros((N, W), dtype=np.int32)
    nlistsize = np.zeros(N, dtype=np.int32)
    ncoup = np.zeros((N, W, shells.hdim))
    B.check(L.uasd_build_pair_list(*args, C.byref(mnn), B.ptr(nlist), B.ptr(nlistsize),
                                   B.ptr(ncoup)), "uasd_build_pair_list")
    if shells.hdim == 1:
        ncoup = ncoup[:, :, 0].copy()
    return nlist, nlistsize, ncoup


# ------------------------------------------------------- standard geometries
def bcc(n, bc=("P", "P", "P")) -> Lattice:
    return Lattice(np.eye(3), [[0, 0, 0], [0.5, 0.5, 0.5]], [1, 1], n, bc)


def triangular(n, bc=("P", "P", "0")) -> Lattice:
    cell = [[1.0, 0.0, 0.0], [-0.5, np.sqrt(3.0) / 2.0, 0.0], [0.0, 0.0, 1.0]]
    return Lattice(cell, [[0, 0, 0]], [1], n, bc)


def kagome(n, bc=("P", "P", "0")) -> Lattice:
    cell = [[1.0, 0.0, 0.0], [0.5, np.sqrt(3.0) / 2.0, 0.0], [0.0, 0.0, 1.0]]
    bas = [[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [0.25, np.sqrt(3.0) / 4.0, 0.0]]
    return Lattice(cell, bas, [1, 1, 1], n, bc)


BCC_SHELLS = ((0.5, 0.5, 0.5), (1.0, 0.0, 0.0), (1.0, 1.0, 0.0), (1.5, 0.5, 0.5), (1.0, 1.0, 1.0))
