"""Batches: many independent systems (replicas of uneven size, each with its
own neighbour list) joined into one lattice for one context.

The joined lattice is the disjoint union of the systems: the sites of system
r are numbered after those of systems 0..r-1 and its neighbour indices are
shifted by the same offset, so no bond crosses systems and one kernel launch
advances all of them.  Ensembles (the context's M) run over every system at
once, e.g. one temperature per ensemble for a parallel scan.  Per-system
observables come from segments (Engine.set_segments with `offsets`).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Joined:
    nlist: np.ndarray       # (Ntot, W) int32, 1-based
    nlistsize: np.ndarray   # (Ntot,) int32
    ncoup: np.ndarray       # (Ntot, W) float64
    mmom: np.ndarray        # (Ntot,) float64
    offsets: np.ndarray     # (R+1,) int32 site offsets of the systems
    extra: dict             # joined optional arrays (dm lists, anisotropy, per-site values)

    @property
    def natom(self) -> int:
        return int(self.offsets[-1])


def _join_lists(lists, sizes, coups, offsets, hdim):
    W = max(int(l.shape[1]) for l in lists)
    N = int(offsets[-1])
    nl = np.zeros((N, W), dtype=np.int32)
    ns = np.zeros(N, dtype=np.int32)
    nc = np.zeros((N, W, hdim)) if hdim > 1 else np.zeros((N, W))
    for r, (l, s, c) in enumerate(zip(lists, sizes, coups)):
        a, b = int(offsets[r]), int(offsets[r + 1])
        w = l.shape[1]
        # 1-based indices; 0 entries (past nlistsize) stay 0
        nl[a:b, :w] = np.where(l > 0, l + a, 0)
        ns[a:b] = s
        c = np.asarray(c, dtype=np.float64)
        if hdim > 1:
            nc[a:b, :w, :] = c.reshape(b - a, w, hdim)
        else:
            nc[a:b, :w] = c.reshape(b - a, w)
    return nl, ns, nc


def join(systems, per_site=()) -> Joined:
    """systems: dicts with nlist (N_r, W_r) 1-based, nlistsize, ncoup, mmom and
    optionally dmlist/dmlistsize/dm_vect, taniso/eaniso/kaniso/sb and the
    per-site arrays named in per_site (e.g. "lambda", "temp")."""
    sizes = [int(s["nlistsize"].shape[0]) for s in systems]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    nl, ns, nc = _join_lists([s["nlist"] for s in systems], [s["nlistsize"] for s in systems],
                             [s["ncoup"] for s in systems], offsets, 1)
    mmom = np.concatenate([np.asarray(s["mmom"], dtype=np.float64) for s in systems])
    extra = {}
    if all("dmlist" in s for s in systems):
        dl, ds, dv = _join_lists([s["dmlist"] for s in systems], [s["dmlistsize"] for s in systems],
                                 [s["dm_vect"] for s in systems], offsets, 3)
        extra.update(dmlist=dl, dmlistsize=ds, dm_vect=dv)
    if all("taniso" in s for s in systems):
        extra.update(taniso=np.concatenate([s["taniso"] for s in systems]).astype(np.int32),
                     eaniso=np.concatenate([s["eaniso"] for s in systems]),
                     kaniso=np.concatenate([s["kaniso"] for s in systems]),
                     sb=np.concatenate([s["sb"] for s in systems]))
    for k in per_site:
        extra[k] = np.concatenate([np.broadcast_to(np.asarray(s[k], dtype=np.float64),
                                                   (n,)) for s, n in zip(systems, sizes)])
    return Joined(nl, ns, nc, mmom, offsets, extra)


def sublattice_major(system: dict, na: int) -> dict:
    """The system with its sites renumbered basis-major: reference site
    i = cell*na + b becomes b*ncell + cell, so each of the na sublattices is
    a contiguous range (a segment of a batch: per-sublattice FP64 sums and
    accumulators on the device).  Lists, couplings and per-site arrays are
    permuted; every site keeps its neighbours and their order."""
    N = int(system["nlistsize"].shape[0])
    if N % na:
        raise ValueError("na must divide the number of sites")
    nc = N // na
    i = np.arange(N)
    new = (i % na) * nc + i // na            # new index of old site i
    old = np.empty(N, dtype=np.int64)
    old[new] = i                             # old site at new index
    out = dict(system)

    def relist(nl):
        nl = np.asarray(nl)[old]
        return np.where(nl > 0, new[np.maximum(nl, 1) - 1] + 1, 0).astype(np.int32)

    out["nlist"], out["nlistsize"] = relist(system["nlist"]), np.asarray(system["nlistsize"])[old]
    out["ncoup"] = np.asarray(system["ncoup"])[old]
    if "dmlist" in system:
        out["dmlist"] = relist(system["dmlist"])
        out["dmlistsize"] = np.asarray(system["dmlistsize"])[old]
        out["dm_vect"] = np.asarray(system["dm_vect"])[old]
    for k in ("mmom", "taniso", "eaniso", "kaniso", "sb"):
        if k in system:
            out[k] = np.asarray(system[k])[old]
    return out


def engine_for(joined: Joined, M: int, precision: str = "fp64"):
    """A context holding the joined lattice, with one segment per system."""
    from .engine import Engine
    eng = Engine(joined.natom, M, precision)
    eng.set_exchange(joined.nlist, joined.nlistsize, joined.ncoup)
    x = joined.extra
    if "dmlist" in x:
        eng.set_dm(x["dmlist"], x["dmlistsize"], x["dm_vect"])
    if "taniso" in x:
        eng.set_anisotropy(x["taniso"], x["eaniso"], x["kaniso"], x["sb"])
    eng.set_segments(joined.offsets)
    return eng
