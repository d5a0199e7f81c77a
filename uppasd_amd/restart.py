"""Moment configuration files in the reference's formats.

  restart<suffix>.<simid>.out   prn_mag_conf_iter, type 'R'
                                (System/restart.f90:186-255): the state to
                                resume from, one file, rewritten
  moment<suffix>.<simid>.out    the same routine, type 'M': snapshots appended
                                at the given iterations
  read_restart                  read_mag_conf_std (System/restart.f90:310-360):
                                the 7-line header skipped, moments normalised,
                                or the legacy layout (one step line, no header)

Rows are (iter, ensemble, atom, |m|, m_x, m_y, m_z) with the edit
descriptors (i8, i8, i8, 2x, 4es16.8), ensembles outer and atoms inner,
1-based, so files written here are read by the reference and vice versa.
"""
from __future__ import annotations

import os

import numpy as np

from .output import fmt_a, fmt_es, fmt_i


def _header(kind: str, mode: str, natom: int, mensemble: int) -> str:
    return "\n".join([
        "#" * 80,
        f"# File type: {kind}",
        f"# Simulation type: {mode}",
        f"# Number of atoms:  {fmt_i(natom)}",
        f"# Number of ensembles:  {fmt_i(mensemble)}",
        "#" * 80,
        fmt_a("# iter", 8) + fmt_a("ens", 8) + fmt_a("iatom", 8) +
        "".join(fmt_a(h, 16) for h in ("|Mom|", "M_x", "M_y", "M_z")),
    ]) + "\n"


def _rows(mstep: int, mmom: np.ndarray, emom: np.ndarray) -> str:
    M, N = mmom.shape
    out = []
    for k in range(M):
        for i in range(N):
            out.append(fmt_i(mstep) + fmt_i(k + 1) + fmt_i(i + 1) + "  " + fmt_es(mmom[k, i]) +
                       fmt_es(emom[k, i, 0]) + fmt_es(emom[k, i, 1]) + fmt_es(emom[k, i, 2]))
    return "\n".join(out) + "\n"


def _check(mmom, emom):
    mmom = np.asarray(mmom, dtype=np.float64)
    emom = np.asarray(emom, dtype=np.float64)
    if mmom.ndim != 2 or emom.shape != mmom.shape + (3,):
        raise ValueError("mmom must be (M, N) and emom (M, N, 3)")
    return mmom, emom


def write_restart(simid: str, mstep: int, mmom, emom, mode: str = "S", suffix: str = "",
                  outdir: str = ".") -> str:
    """restart<suffix>.<simid>.out with the state (mmom (M, N), emom (M, N, 3))."""
    mmom, emom = _check(mmom, emom)
    path = os.path.join(outdir, f"restart{suffix}.{simid.strip()}.out")
    with open(path, "w") as fh:
        fh.write(_header("R", mode, mmom.shape[1], mmom.shape[0]))
        fh.write(_rows(mstep, mmom, emom))
    return path


def write_moments(simid: str, mstep: int, mmom, emom, mode: str = "S", suffix: str = "",
                  outdir: str = ".") -> str:
    """Append a snapshot to moment<suffix>.<simid>.out; iteration 0 starts the
    file with its header."""
    mmom, emom = _check(mmom, emom)
    path = os.path.join(outdir, f"moment{suffix}.{simid.strip()}.out")
    with open(path, "w" if mstep == 0 else "a") as fh:
        if mstep == 0:
            fh.write(_header("M", mode, mmom.shape[1], mmom.shape[0]))
        fh.write(_rows(mstep, mmom, emom))
    return path


def read_restart(path: str, natom: int, mensemble: int, legacy: bool = False):
    """(rstep, mmom (M, N), emom (M, N, 3)) from a restart file.

    The standard layout skips 7 header lines and normalises every moment
    direction (f_normalize_vec); the legacy layout starts with the step and
    takes the directions as written."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    if legacy:
        rstep = int(lines[0].split()[0])
        body = lines[1:]
        cols = slice(2, 6)
    else:
        body = lines[7:]
        cols = slice(3, 7)
    n = natom * mensemble
    rows = [ln.split() for ln in body if ln.strip()][:n]
    if len(rows) < n:
        raise ValueError(f"{path}: {len(rows)} moment lines, {n} expected")
    vals = np.array([[float(x.replace("D", "E").replace("d", "e")) for x in r[cols]]
                     for r in rows], dtype=np.float64)
    if not legacy:
        rstep = int(rows[-1][0])
    mmom = vals[:, 0].reshape(mensemble, natom)
    emom = vals[:, 1:4].reshape(mensemble, natom, 3)
    if not legacy:
        emom = emom / np.linalg.norm(emom, axis=2, keepdims=True)
    return rstep, mmom, np.ascontiguousarray(emom)
