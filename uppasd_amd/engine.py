"""Host-side mirror of the reference's hot-path routines over libuasd_hip.so.

``Engine`` holds one device context.  Its methods keep the names, argument
meaning and in/out behaviour of the reference routines they replace:

  effective_field   effective_field_full  (Hamiltonian/hamiltonianactions.f90:108-252)
  evolve_first      evolve_first          (Evolution/evolution.f90:38-273), SDEalgh 1 & 5
  evolve_second     evolve_second         (Evolution/evolution.f90:280-406)
  moment_update     moment_update / copym (Evolution/updatemoments.f90:19-95)
  mc_sweeps         mc_evolve             (MonteCarlo/montecarlo.f90:44-273), coloured
  sd_steps          sd_mphase loop body   (sd_driver.f90:668-765), device resident

Arrays are numpy, C-order with the Fortran dimensions reversed: a reference
(3,Natom,Mensemble) array is shape (Mensemble, Natom, 3) here.  Errors raise
``UasdError`` with the library's message; nothing falls back to the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as B
from ._lib import (UASD_FP32, UASD_FP64, UASD_NOISE_INJECTED, UASD_NOISE_NONE,
                   UASD_NOISE_PHILOX, UASD_SDE_DEPONDT, UASD_SDE_HEUN, UASD_SDE_MIDPOINT,
                   UASD_SDE_SIB, UASD_SDE_SPHMID, UasdError)

# LDS tile shapes (cells x, y, z) selectable with set_tuning(TUNE_LLG, g, 256 | (k << 18))
TILE_SHAPES = ((8, 8, 4), (16, 8, 2), (32, 4, 2), (16, 4, 4), (16, 16, 1), (32, 8, 1), (64, 4, 1),
               (8, 4, 4), (4, 4, 4), (8, 8, 2), (16, 8, 1), (8, 8, 1), (16, 8, 4), (8, 16, 4),
               (8, 8, 8))

__all__ = ["Engine", "moment_update", "UasdError", "UASD_NOISE_NONE", "UASD_NOISE_PHILOX",
           "UASD_NOISE_INJECTED", "UASD_SDE_MIDPOINT", "UASD_SDE_HEUN", "UASD_SDE_DEPONDT",
           "UASD_SDE_SIB", "UASD_SDE_SPHMID"]


def device_count() -> int:
    n = C.c_int(0)
    B.check(B.lib().uasd_device_count(C.byref(n)), "uasd_device_count")
    return n.value


class Engine:
    def __init__(self, natom: int, mensemble: int = 1, precision: str = "fp64",
                 site_order: int = 1):
        """site_order = NA stores each basis atom's sites contiguously on the
        device (host arrays keep the reference order; results are the same)."""
        prec = {"fp64": UASD_FP64, "fp32": UASD_FP32}[precision]
        self.N, self.M, self.precision = int(natom), int(mensemble), precision
        h = C.c_void_p()
        B.check(B.lib().uasd_ctx_create(C.byref(h), self.N, self.M, prec), "uasd_ctx_create")
        self._h = h
        if site_order != 1:
            B.check(B.lib().uasd_set_site_order(self._h, int(site_order)), "uasd_set_site_order")

    # -------------------------------------------------------------- lifetime
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            B.lib().uasd_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_stream(self, hip_stream: int | None):
        B.check(B.lib().uasd_ctx_set_stream(self._h, hip_stream), "uasd_ctx_set_stream")

    def stream(self) -> int:
        return B.lib().uasd_ctx_stream(self._h)

    def sync(self):
        B.check(B.lib().uasd_ctx_sync(self._h), "uasd_ctx_sync")

    # ----------------------------------------------------------- Hamiltonian
    def set_exchange(self, nlist, nlistsize, ncoup, aham=None):
        nl, ns, nc = B.i32(nlist), B.i32(nlistsize), B.f64(ncoup)
        mnn = nl.shape[1]
        B.check(B.lib().uasd_set_exchange(self._h, mnn, ns.shape[0], B.ptr(nl), B.ptr(ns),
                                          B.ptr(nc), B.ptr(B.i32(aham))), "uasd_set_exchange")

    def set_dm(self, dmlist, dmlistsize, dm_vect, aham=None):
        if dmlist is None:
            B.check(B.lib().uasd_set_dm(self._h, 0, 0, None, None, None, None), "uasd_set_dm")
            return
        dl, ds, dv = B.i32(dmlist), B.i32(dmlistsize), B.f64(dm_vect)
        B.check(B.lib().uasd_set_dm(self._h, dl.shape[1], ds.shape[0], B.ptr(dl), B.ptr(ds),
                                    B.ptr(dv), B.ptr(B.i32(aham))), "uasd_set_dm")

    def set_anisotropy(self, taniso, eaniso, kaniso, sb):
        B.check(B.lib().uasd_set_anisotropy(self._h, B.ptr(B.i32(taniso)), B.ptr(B.f64(eaniso)),
                                            B.ptr(B.f64(kaniso)), B.ptr(B.f64(sb))),
                "uasd_set_anisotropy")

    def set_external_field(self, field):
        """field: (M,N,3) array, a 3-vector applied everywhere, or None."""
        if field is not None:
            f = np.asarray(field, dtype=np.float64)
            if f.shape == (3,):
                f = np.broadcast_to(f, (self.M, self.N, 3))
            field = np.ascontiguousarray(f)
        B.check(B.lib().uasd_set_external_field(self._h, B.ptr(field)), "uasd_set_external_field")

    # ------------------------------------------------------------ parameters
    def set_llg(self, lambda1, temp, delta_t, landeg=None, temprescale=1.0, bn=1.0):
        lam = B.f64(np.broadcast_to(np.asarray(lambda1, dtype=np.float64), (self.N,)))
        tmp = B.f64(np.broadcast_to(np.asarray(temp, dtype=np.float64), (self.N,)))
        lg = None if landeg is None else B.f64(np.broadcast_to(np.asarray(landeg, dtype=np.float64), (self.N,)))
        B.check(B.lib().uasd_set_llg(self._h, B.ptr(lam), B.ptr(tmp), B.ptr(lg), float(delta_t),
                                     float(temprescale), float(bn)), "uasd_set_llg")

    def set_noise(self, mode: int, seed: int = 0, normals=None):
        B.check(B.lib().uasd_set_noise(self._h, int(mode), int(seed), B.ptr(B.f64(normals))),
                "uasd_set_noise")

    # ---------------------------------------------------------------- state
    def set_moments(self, emom, mmom):
        """emom (M,N,3) unit vectors, mmom (M,N) magnitudes."""
        e, m = B.f64(emom), B.f64(np.broadcast_to(np.asarray(mmom, dtype=np.float64), (self.M, self.N)))
        B.check(B.lib().uasd_set_moments(self._h, B.ptr(e), B.ptr(m)), "uasd_set_moments")

    def get_moments(self):
        e = np.zeros((self.M, self.N, 3))
        eM = np.zeros((self.M, self.N, 3))
        B.check(B.lib().uasd_get_moments(self._h, B.ptr(e), B.ptr(eM)), "uasd_get_moments")
        return e, eM

    # -------------------------------------------------- drop-in (host arrays)
    def effective_field(self, emomM, parts: bool = False):
        """beff and the field-routine energy (mRy) of emomM, as effective_field_full;
        with parts=True also (beff1, beff2) as that routine splits them."""
        eM = B.f64(emomM)
        beff = np.zeros_like(eM)
        b1 = np.zeros_like(eM) if parts else None
        b2 = np.zeros_like(eM) if parts else None
        en = C.c_double(0.0)
        B.check(B.lib().uasd_effective_field_host(self._h, B.ptr(eM), B.ptr(beff), B.ptr(b1),
                                                  B.ptr(b2), C.byref(en)),
                "uasd_effective_field_host")
        if parts:
            return beff, en.value, b1, b2
        return beff, en.value

    def evolve_first(self, sdealgh, beff, emom, emom2, emomM, b2eff=None):
        """In place on emom, emom2, emomM (and b2eff), like evolve_first."""
        for a in (emom, emom2, emomM) + ((b2eff,) if b2eff is not None else ()):
            assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        B.check(B.lib().uasd_evolve_first_host(self._h, int(sdealgh), B.ptr(B.f64(beff)),
                                               B.ptr(emom), B.ptr(emom2), B.ptr(emomM),
                                               B.ptr(b2eff)), "uasd_evolve_first_host")

    def evolve_second(self, sdealgh, beff, b2eff, emom, emom2):
        """In place on emom and emom2, like evolve_second."""
        for a in (emom, emom2):
            assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
        B.check(B.lib().uasd_evolve_second_host(self._h, int(sdealgh), B.ptr(B.f64(beff)),
                                                B.ptr(B.f64(b2eff)), B.ptr(emom), B.ptr(emom2)),
                "uasd_evolve_second_host")

    # -------------------------------------------------------- resident path
    def sd_steps(self, sdealgh: int, nsteps: int):
        B.check(B.lib().uasd_sd_steps(self._h, int(sdealgh), int(nsteps)), "uasd_sd_steps")

    def field(self, with_energy: bool = True):
        beff = np.zeros((self.M, self.N, 3))
        en = C.c_double(0.0)
        B.check(B.lib().uasd_effective_field(self._h, B.ptr(beff),
                                             C.byref(en) if with_energy else None),
                "uasd_effective_field")
        return beff, en.value

    def mc_sweeps(self, mode: str, temps, nsweeps: int, seed: int):
        t = B.f64(np.broadcast_to(np.asarray(temps, dtype=np.float64), (self.M,)))
        B.check(B.lib().uasd_mc_sweeps(self._h, mode.encode()[:1], B.ptr(t), int(nsweeps),
                                       int(seed)), "uasd_mc_sweeps")

    def measure(self, with_energy: bool = False):
        """(msum (M,3) = sum_i emomM, energy (M,) in mRy or None)."""
        ms = np.zeros((self.M, 3))
        en = np.zeros(self.M) if with_energy else None
        B.check(B.lib().uasd_measure(self._h, B.ptr(ms), B.ptr(en)), "uasd_measure")
        return ms, en

    def energy_terms(self):
        """calc_energy's terms per ensemble, energy per site in mRy: (M, 4)
        with columns exchange, anisotropy, DM, Zeeman; (nseg, M, 4) when the
        lattice holds several segments."""
        nseg = getattr(self, "nseg", 1)
        out = np.zeros((nseg, self.M, 4))
        B.check(B.lib().uasd_energy_terms(self._h, B.ptr(out)), "uasd_energy_terms")
        return out[0] if nseg == 1 else out

    def coloring(self):
        nc = C.c_int(0)
        col = np.zeros(self.N, dtype=np.int32)
        B.check(B.lib().uasd_get_coloring(self._h, C.byref(nc), B.ptr(col)), "uasd_get_coloring")
        return nc.value, col

    def set_kernel_timing(self, enable: bool, every: int = 1, calls_only: bool = False):
        """HIP events around one launch in `every` of each kernel kind and
        around every sd_steps / mc_sweeps call (kind 6); calls_only times the
        calls alone, with no events between kernels."""
        code = (-1 if calls_only else int(every)) if enable else 0
        B.check(B.lib().uasd_set_kernel_timing(self._h, code), "uasd_set_kernel_timing")

    def kernel_timing(self, kind: int):
        n = C.c_int64(0)
        ms = C.c_double(0.0)
        B.check(B.lib().uasd_get_kernel_timing(self._h, kind, C.byref(n), C.byref(ms)),
                "uasd_get_kernel_timing")
        return n.value, ms.value

    def set_lattice_dims(self, na: int, n1: int, n2: int, n3: int):
        """Site i = b + na*(x + n1*(y + n2*z)); enables the LDS-tiled stencil
        kernel when the lists are translation invariant."""
        B.check(B.lib().uasd_set_lattice_dims(self._h, int(na), int(n1), int(n2), int(n3)),
                "uasd_set_lattice_dims")

    def stencil(self):
        """(has_stencil, (rx, ry, rz))."""
        has = C.c_int(0)
        r = np.zeros(3, dtype=np.int32)
        B.check(B.lib().uasd_has_stencil(self._h, C.byref(has), B.ptr(r)), "uasd_has_stencil")
        return bool(has.value), tuple(int(x) for x in r)

    def index_bytes(self):
        """(exchange, dm) bytes per neighbour entry read by the kernels (2 or 4)."""
        a, b = C.c_int(0), C.c_int(0)
        B.check(B.lib().uasd_index_bytes(self._h, C.byref(a), C.byref(b)), "uasd_index_bytes")
        return a.value, b.value

    def llg_tile(self):
        """Tile shape k the fused LLG step launches under the current tuning,
        0 for the list-gather kernels."""
        k = C.c_int(0)
        B.check(B.lib().uasd_llg_tile(self._h, C.byref(k)), "uasd_llg_tile")
        return k.value

    def set_graph_steps(self, k: int):
        """Replay hipGraphs of k LLG steps / MC sweeps (0 = launch one by one)."""
        B.check(B.lib().uasd_set_graph_steps(self._h, int(k)), "uasd_set_graph_steps")

    # --------------------------------------------------------------- batches
    def set_segments(self, offsets):
        """Replica site ranges: segment s = sites [offsets[s], offsets[s+1])."""
        off = B.i32(offsets)
        B.check(B.lib().uasd_set_segments(self._h, off.shape[0] - 1, B.ptr(off)),
                "uasd_set_segments")
        self.nseg = off.shape[0] - 1

    def measure_segments(self, with_energy: bool = False):
        """(msum (nseg, M, 3), energy (nseg, M) in mRy or None), FP64."""
        nseg = getattr(self, "nseg", 1)
        ms = np.zeros((nseg, self.M, 3))
        en = np.zeros((nseg, self.M)) if with_energy else None
        B.check(B.lib().uasd_measure_segments(self._h, B.ptr(ms), B.ptr(en)),
                "uasd_measure_segments")
        return ms, en

    def accumulate(self, with_energy: bool = True):
        """One device-side cumulant sample per (segment, ensemble); no host sync."""
        B.check(B.lib().uasd_accumulate(self._h, 1 if with_energy else 0), "uasd_accumulate")

    def accumulators(self):
        """dict of (nseg, M) arrays: W, n, m, m2, m4, e, e2 and U4; plus samples."""
        nseg = getattr(self, "nseg", 1)
        a = np.zeros((nseg, self.M, 8))
        n = C.c_int64(0)
        B.check(B.lib().uasd_get_accumulators(self._h, B.ptr(a), C.byref(n)),
                "uasd_get_accumulators")
        out = {k: a[..., i] for i, k in enumerate(("W", "n", "m", "m2", "m4", "e", "e2"))}
        with np.errstate(divide="ignore", invalid="ignore"):
            out["U4"] = 1.0 - out["m4"] / (3.0 * out["m2"] ** 2)
        out["samples"] = n.value
        return out

    def reset_accumulators(self):
        B.check(B.lib().uasd_reset_accumulators(self._h), "uasd_reset_accumulators")

    # kinds of launch shape
    TUNE_LLG, TUNE_FIELD, TUNE_MC = 0, 1, 2

    def set_tuning(self, kind: int, gather: int, block: int):
        """Launch shape of a gathering kernel (results do not depend on it)."""
        B.check(B.lib().uasd_set_tuning(self._h, kind, gather, block), "uasd_set_tuning")

    def tuning(self, kind: int):
        g, b = C.c_int(0), C.c_int(0)
        B.check(B.lib().uasd_get_tuning(self._h, kind, C.byref(g), C.byref(b)), "uasd_get_tuning")
        return g.value, b.value

    def autotune_llg(self, sdealgh: int, steps: int = 3) -> float:
        """Times every launch shape on this lattice, keeps the fastest; the
        state and the step counter are restored.  Returns the best time (ms)."""
        ms = C.c_double(0.0)
        B.check(B.lib().uasd_autotune_llg(self._h, sdealgh, steps, C.byref(ms)),
                "uasd_autotune_llg")
        return ms.value

    def autotune_mc(self, mode: str, temps, sweeps: int = 2) -> float:
        t = B.f64(np.broadcast_to(np.asarray(temps, dtype=np.float64), (self.M,)))
        ms = C.c_double(0.0)
        B.check(B.lib().uasd_autotune_mc(self._h, mode.encode()[:1], B.ptr(t), sweeps, C.byref(ms)),
                "uasd_autotune_mc")
        return ms.value


def set_tune_cache(enable: bool = True, path: str | None = None):
    """Keep autotuner winners per workload for the process (and in `path`,
    loaded now and appended to); later autotune calls on the same workload
    reuse them without timing.  enable=False switches the cache off."""
    B.check(B.lib().uasd_set_tune_cache(1 if enable else 0,
                                        None if path is None else str(path).encode()),
            "uasd_set_tune_cache")


def moment_update(mmom, emom, emom2, emomM):
    """copym (updatemoments.f90:48-95): emom <- emom2, emomM <- emom2*mmom (in place)."""
    emom[...] = emom2
    emomM[...] = emom2 * np.asarray(mmom)[..., None]
