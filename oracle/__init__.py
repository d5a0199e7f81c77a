"""CPU oracle for the UppASD spin-update hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker, never the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  It wraps ``liboracle.so`` (plain C, FP64, built with
-ffp-contract=off) whose functions restate the reference routines term for
term; each C function cites the reference file:line it follows.

Pinning (what checks each restated function, see tests/):
  MT19937 / ziggurat        reference golden values of tests/Solvers and
                            tests/bccFe (regulartests.yaml), stream-identical
  setup_nm & co.            shell counts / closed forms, PBC invariance,
                            brute-force distance search, golden runs
  field, Depondt, midpoint  Solvers golden values (T=0 chain), closed-form
                            Larmor precession, brute-force pair sums
  MC (reference order)      bccFe M/H golden values (statistical), exact
                            enumeration / single-spin heat-bath averages
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (gcc)."""
    args = ["make", "-s", "-C", _HERE]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return _LIB_PATH


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


# ----------------------------------------------------------------- ctypes types
dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class MT(C.Structure):
    _fields_ = [("mti", C.c_int64), ("mt", C.c_uint64 * 624)]


class Zig(C.Structure):
    _fields_ = [("kn", C.c_int32 * 128), ("fn", C.c_double * 128),
                ("wn", C.c_double * 128), ("st", MT)]


class Lattice(C.Structure):
    _fields_ = [("NA", C.c_int), ("NT", C.c_int), ("N1", C.c_int), ("N2", C.c_int),
                ("N3", C.c_int), ("C1", C.c_double * 3), ("C2", C.c_double * 3),
                ("C3", C.c_double * 3), ("bas", C.POINTER(C.c_double)),
                ("atype_inp", C.POINTER(C.c_int)), ("BC", C.c_char * 3)]


class Shells(C.Structure):
    _fields_ = [("NT", C.c_int), ("max_no_shells", C.c_int), ("nchmax", C.c_int),
                ("nn", C.POINTER(C.c_int)), ("redcoord", C.POINTER(C.c_double)),
                ("nntype", C.POINTER(C.c_int)), ("jc", C.POINTER(C.c_double)),
                ("hdim", C.c_int)]


class NMap(C.Structure):
    _fields_ = [("Natom", C.c_int), ("max_no_shells", C.c_int), ("max_no_equiv", C.c_int),
                ("max_no_neigh", C.c_int), ("nm", C.POINTER(C.c_int)),
                ("nmdim", C.POINTER(C.c_int))]


class PairList(C.Structure):
    _fields_ = [("Natom", C.c_int), ("hdim", C.c_int), ("max_no_neigh", C.c_int),
                ("nlist", C.POINTER(C.c_int)), ("nlistsize", C.POINTER(C.c_int)),
                ("ncoup", C.POINTER(C.c_double))]


class Ham(C.Structure):
    _fields_ = [("Natom", C.c_int),
                ("max_no_neigh", C.c_int), ("nlist", C.c_void_p), ("nlistsize", C.c_void_p),
                ("ncoup", C.c_void_p),
                ("do_dm", C.c_int), ("max_no_dmneigh", C.c_int), ("dmlist", C.c_void_p),
                ("dmlistsize", C.c_void_p), ("dm_vect", C.c_void_p),
                ("do_anisotropy", C.c_int), ("taniso", C.c_void_p), ("eaniso", C.c_void_p),
                ("kaniso", C.c_void_p), ("sb", C.c_void_p)]


def _declare(L):
    vp = C.c_void_p
    L.or_mt_init.argtypes = [C.POINTER(MT), C.c_int32]
    L.or_mt_rand64.argtypes = [C.POINTER(MT)]
    L.or_mt_rand64.restype = C.c_uint64
    L.or_mt_real2.argtypes = [C.POINTER(MT)]
    L.or_mt_real2.restype = C.c_double
    L.or_mt_real1.argtypes = [C.POINTER(MT)]
    L.or_mt_real1.restype = C.c_double
    L.or_zig_setup.argtypes = [C.POINTER(Zig), C.c_int32]
    L.or_zig_nor.argtypes = [C.POINTER(Zig)]
    L.or_zig_nor.restype = C.c_double
    L.or_mt_fill_real2.argtypes = [C.POINTER(MT), dp, C.c_long]
    L.or_zig_fill.argtypes = [C.POINTER(Zig), dp, C.c_long]
    L.or_noise3_f64.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, dp]
    L.or_thermal3.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, fp]
    L.or_bm_ln.argtypes = [C.c_float]
    L.or_bm_ln.restype = C.c_float
    L.or_bm_sincos.argtypes = [C.c_float, fp]
    L.or_uniform4.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32, dp]
    L.or_philox4x32_10.argtypes = [C.POINTER(C.c_uint32 * 4), C.POINTER(C.c_uint32 * 2),
                                   C.POINTER(C.c_uint32 * 4)]
    L.or_read_exchange_lines.argtypes = [C.POINTER(Lattice), C.c_int, C.c_int, C.c_int, C.c_int,
                                         ip, ip, ip, ip, dp, dp, C.POINTER(Shells)]
    L.or_shells_free.argtypes = [C.POINTER(Shells)]
    L.or_setup_nm.argtypes = [C.POINTER(Lattice), C.c_int, C.POINTER(Shells), C.c_int,
                              C.POINTER(NMap)]
    L.or_nmap_free.argtypes = [C.POINTER(NMap)]
    L.or_setup_neighbour_hamiltonian.argtypes = [C.POINTER(Lattice), C.POINTER(Shells),
                                                 C.POINTER(NMap), dp, vp, C.c_int, C.c_int,
                                                 C.POINTER(PairList)]
    L.or_pairlist_free.argtypes = [C.POINTER(PairList)]
    L.or_setup_anisotropies.argtypes = [C.POINTER(Lattice), ip, dp, dp, ip, dp, dp, dp]
    L.or_effective_field.argtypes = [C.POINTER(Ham), C.c_int, C.c_int, dp, vp, dp, vp, vp,
                                     C.POINTER(C.c_double)]
    L.or_effective_field_single.argtypes = [C.POINTER(Ham), C.c_int, dp, vp, C.c_int, C.c_int, dp]
    L.or_energy_terms.argtypes = [C.POINTER(Ham), C.c_int, C.c_int, dp, vp, dp]
    L.or_pt_exchange.argtypes = [C.c_int, dp, dp, C.c_int, dp, C.c_double,
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.or_pt_select.argtypes = [C.c_int, dp]
    L.or_pt_select.restype = C.c_int
    L.or_depondt_first.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double, C.c_double, dp, dp,
                                   dp, dp, dp, dp, dp, vp]
    L.or_depondt_second.argtypes = [C.c_int, C.c_int, dp, C.c_double, dp, dp, dp, dp, dp]
    L.or_rannum_scale.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_double, dp, dp,
                                  C.c_double, dp]
    L.or_midpoint_first.argtypes = [C.c_int, C.c_int, dp, C.c_double, dp, dp, dp, dp, dp, dp,
                                    C.c_double, dp, vp]
    L.or_heun_first.argtypes = [C.c_int, C.c_int, dp, dp, dp, C.c_double, C.c_double, dp, dp,
                                dp, dp, dp, dp, dp]
    L.or_sib_first.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_double, dp, dp, dp, dp, dp,
                               dp, C.c_double, dp]
    L.or_sib_second.argtypes = [C.c_int, C.c_int, C.c_int, dp, C.c_double, dp, dp, dp, dp,
                                C.c_double, dp]
    L.or_heun_second.argtypes = [C.c_int, C.c_int, dp, dp, C.c_double, dp, dp, dp, dp, dp]
    L.or_midpoint_second.argtypes = [C.c_int, C.c_int, dp, C.c_double, dp, dp, dp, dp,
                                     C.c_double, dp]
    L.or_moment_update.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, vp]
    L.or_sd_run.argtypes = [C.POINTER(Ham), C.c_int, C.c_int, C.c_int, C.c_long, C.c_uint64,
                            C.c_double, dp, dp, dp, dp, dp, dp, C.c_uint64, C.c_int, vp]
    L.or_mc_delta.argtypes = [C.c_double]
    L.or_mc_delta.restype = C.c_double
    L.or_mc_energy_change.argtypes = [C.POINTER(Ham), C.c_int, dp, dp, dp, C.c_int, C.c_int,
                                      dp, dp]
    L.or_mc_energy_change.restype = C.c_double
    L.or_flip_h.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.c_double, dp, C.c_double]
    L.or_choose_random_atom_x.argtypes = [C.POINTER(MT), C.c_int, ip]
    L.or_mc_sweep_ref.argtypes = [C.POINTER(Ham), C.c_int, C.c_int, C.c_double, C.c_double,
                                  C.c_char, C.POINTER(MT), C.POINTER(Zig), ip, dp, dp, dp, dp]
    L.or_mc_sweep_colored.argtypes = [C.POINTER(Ham), C.c_int, C.c_int, dp, C.c_char,
                                      C.c_uint64, C.c_uint64, C.c_int, ip, ip, dp, dp, dp, dp]
    L.or_mc_sweep_colored_ex.argtypes = [C.POINTER(Ham), C.c_int, C.c_int, dp, C.c_char,
                                         C.c_uint64, C.c_uint64, C.c_int, ip, ip, dp, dp, dp,
                                         dp, C.c_int]
    L.or_trial_move_f32.argtypes = [dp, C.c_double, np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"), dp]
    L.or_counter_words.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32,
                                   np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")]
    L.or_u24.argtypes = [C.c_uint32]
    L.or_u24.restype = C.c_float
    L.or_set_threads.argtypes = [C.c_int]
    L.or_threads.restype = C.c_int
    L.or_thermal3_f64.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, dp]
    L.or_calc_mavrg.argtypes = [C.c_int, C.c_int, dp]
    L.or_calc_mavrg.restype = C.c_double
    L.or_sum_moments.argtypes = [C.c_int, C.c_int, dp, dp]


# --------------------------------------------------------------------- helpers
def _np_from(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class LatticeSpec:
    """Geometry of a run: cell vectors, basis (Cartesian), types, repetitions."""

    def __init__(self, cell, bas, atype, n, bc=("P", "P", "P")):
        self.cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
        self.bas = np.ascontiguousarray(np.asarray(bas, dtype=np.float64).reshape(-1, 3))
        self.atype = np.ascontiguousarray(np.asarray(atype, dtype=np.int32))
        self.n = tuple(int(x) for x in n)
        self.bc = tuple(bc)
        self.NA = self.bas.shape[0]
        self.NT = int(self.atype.max())
        self.natom = self.NA * self.n[0] * self.n[1] * self.n[2]

    def c_struct(self) -> Lattice:
        s = Lattice()
        s.NA, s.NT = self.NA, self.NT
        s.N1, s.N2, s.N3 = self.n
        for k, name in enumerate(("C1", "C2", "C3")):
            getattr(s, name)[:] = list(self.cell[k])
        s.bas = self.bas.ctypes.data_as(C.POINTER(C.c_double))
        s.atype_inp = self.atype.ctypes.data_as(C.POINTER(C.c_int))
        s.BC = "".join(self.bc).encode()
        return s


def build_pairs(lat: LatticeSpec, lines, sym: int, hdim: int = 1, use_nntype: bool = True,
                ammom=None, achem=None, nchmax: int = 1, do_ralloy: int = 0):
    """read_exchange + setup_nm + setup_neighbour_hamiltonian.

    ``lines``: list of (isite, jsite, ichem, jchem, rx, ry, rz, *values).
    Returns dict(nlist (Natom,mnn) 1-based, nlistsize, ncoup (Natom,mnn[,3]), mnn, nm, nmdim).
    """
    L = lib()
    ls = lat.c_struct()
    nl = len(lines)
    isite = np.ascontiguousarray([int(x[0]) for x in lines], dtype=np.int32)
    jsite = np.ascontiguousarray([int(x[1]) for x in lines], dtype=np.int32)
    ichem = np.ascontiguousarray([int(x[2]) for x in lines], dtype=np.int32)
    jchem = np.ascontiguousarray([int(x[3]) for x in lines], dtype=np.int32)
    r = np.ascontiguousarray([[float(v) for v in x[4:7]] for x in lines], dtype=np.float64).reshape(-1)
    val = np.ascontiguousarray([[float(v) for v in x[7:7 + hdim]] for x in lines],
                               dtype=np.float64).reshape(-1)
    sh = Shells()
    if L.or_read_exchange_lines(C.byref(ls), do_ralloy, nchmax, hdim, nl, isite, jsite, ichem,
                                jchem, r, val, C.byref(sh)) != 0:
        raise RuntimeError("or_read_exchange_lines failed")
    nm = NMap()
    L.or_setup_nm(C.byref(ls), sym, C.byref(sh), 1 if use_nntype else 0, C.byref(nm))
    if ammom is None:
        ammom = np.ones((lat.NA, nchmax))
    am = np.ascontiguousarray(np.asarray(ammom, dtype=np.float64).reshape(lat.NA, nchmax).T).reshape(-1)
    ach = None
    if achem is not None:
        ach = np.ascontiguousarray(achem, dtype=np.int32)
    pl = PairList()
    L.or_setup_neighbour_hamiltonian(C.byref(ls), C.byref(sh), C.byref(nm), am,
                                     ach.ctypes.data if ach is not None else None,
                                     1, 0, C.byref(pl))
    N, mnn = pl.Natom, pl.max_no_neigh
    out = {
        "mnn": mnn,
        "nlist": _np_from(pl.nlist, N * mnn, np.int32).reshape(N, mnn),
        "nlistsize": _np_from(pl.nlistsize, N, np.int32),
        "ncoup": _np_from(pl.ncoup, N * mnn * hdim, np.float64).reshape(N, mnn, hdim),
        "nm": _np_from(nm.nm, N * nm.max_no_shells * nm.max_no_equiv, np.int32).reshape(
            nm.max_no_equiv, nm.max_no_shells, N),
        "nmdim": _np_from(nm.nmdim, N * nm.max_no_shells, np.int32).reshape(N, nm.max_no_shells),
        "nn": _np_from(sh.nn, sh.NT, np.int32),
    }
    if hdim == 1:
        out["ncoup"] = out["ncoup"][:, :, 0]
    L.or_pairlist_free(C.byref(pl))
    L.or_nmap_free(C.byref(nm))
    L.or_shells_free(C.byref(sh))
    return out


def setup_anisotropies(lat: LatticeSpec, types, aniso, ammom):
    """types (NA,), aniso (NA,6) rows K1,K2,ex,ey,ez,ratio; ammom (NA,)."""
    L = lib()
    ls = lat.c_struct()
    N = lat.natom
    ty = np.ascontiguousarray(types, dtype=np.int32)
    an = np.ascontiguousarray(np.asarray(aniso, dtype=np.float64).reshape(lat.NA, 6).T).reshape(-1)
    am = np.ascontiguousarray(ammom, dtype=np.float64)
    taniso = np.zeros(N, np.int32)
    eaniso = np.zeros(3 * N)
    kaniso = np.zeros(2 * N)
    sb = np.zeros(N)
    L.or_setup_anisotropies(C.byref(ls), ty, an, am, taniso, eaniso, kaniso, sb)
    return taniso, eaniso.reshape(N, 3), kaniso.reshape(N, 2), sb


class OracleHam:
    """Keeps numpy arrays alive and exposes an or_ham struct."""

    def __init__(self, natom, nlist, nlistsize, ncoup, dmlist=None, dmlistsize=None,
                 dm_vect=None, taniso=None, eaniso=None, kaniso=None, sb=None):
        self.N = int(natom)
        self.nlist = np.ascontiguousarray(nlist, dtype=np.int32)
        self.nlistsize = np.ascontiguousarray(nlistsize, dtype=np.int32)
        self.ncoup = np.ascontiguousarray(ncoup, dtype=np.float64)
        self.mnn = self.nlist.shape[1]
        h = Ham()
        h.Natom = self.N
        h.max_no_neigh = self.mnn
        h.nlist, h.nlistsize, h.ncoup = (self.nlist.ctypes.data, self.nlistsize.ctypes.data,
                                         self.ncoup.ctypes.data)
        if dmlist is not None:
            self.dmlist = np.ascontiguousarray(dmlist, dtype=np.int32)
            self.dmlistsize = np.ascontiguousarray(dmlistsize, dtype=np.int32)
            self.dm_vect = np.ascontiguousarray(dm_vect, dtype=np.float64)
            h.do_dm = 1
            h.max_no_dmneigh = self.dmlist.shape[1]
            h.dmlist, h.dmlistsize, h.dm_vect = (self.dmlist.ctypes.data,
                                                 self.dmlistsize.ctypes.data,
                                                 self.dm_vect.ctypes.data)
        if taniso is not None:
            self.taniso = np.ascontiguousarray(taniso, dtype=np.int32)
            self.eaniso = np.ascontiguousarray(eaniso, dtype=np.float64)
            self.kaniso = np.ascontiguousarray(kaniso, dtype=np.float64)
            self.sb = np.ascontiguousarray(sb, dtype=np.float64)
            h.do_anisotropy = 1
            h.taniso, h.eaniso, h.kaniso, h.sb = (self.taniso.ctypes.data, self.eaniso.ctypes.data,
                                                  self.kaniso.ctypes.data, self.sb.ctypes.data)
        self.h = h

    # arrays are (M, N, 3) C-order == Fortran (3, N, M)
    def field(self, emomM, ext=None):
        M = emomM.shape[0]
        beff = np.zeros_like(emomM)
        b1 = np.zeros_like(emomM)
        b2 = np.zeros_like(emomM)
        en = C.c_double()
        ext_arr = None if ext is None else np.ascontiguousarray(ext, dtype=np.float64)
        lib().or_effective_field(C.byref(self.h), self.N, M, np.ascontiguousarray(emomM),
                                 ext_arr.ctypes.data if ext_arr is not None else None,
                                 beff, b1.ctypes.data, b2.ctypes.data, C.byref(en))
        self.last_parts = (b1, b2)
        return beff, en.value

    def field_parts(self, emomM, ext=None):
        """(beff, energy, beff1, beff2) as effective_field_full returns them."""
        beff, en = self.field(emomM, ext)
        return beff, en, self.last_parts[0], self.last_parts[1]

    def energy_terms(self, emomM, ext=None):
        """calc_energy terms per ensemble: (M, 4) per atom in field units,
        columns exchange, anisotropy, DM, external field (times mub/mry for
        mRy)."""
        M = emomM.shape[0]
        out = np.zeros((M, 4))
        ext_arr = None if ext is None else np.ascontiguousarray(ext, dtype=np.float64)
        lib().or_energy_terms(C.byref(self.h), self.N, M, np.ascontiguousarray(emomM),
                              ext_arr.ctypes.data if ext_arr is not None else None, out)
        return out

    def field_single(self, emomM, i, k, ext_uniform=None):
        out = np.zeros(3)
        e = None if ext_uniform is None else np.ascontiguousarray(ext_uniform, dtype=np.float64)
        lib().or_effective_field_single(C.byref(self.h), self.N, np.ascontiguousarray(emomM),
                                        e.ctypes.data if e is not None else None, i, k, out)
        return out


def zig(seed: int) -> Zig:
    z = Zig()
    lib().or_zig_setup(C.byref(z), seed)
    return z


def mt(seed: int) -> MT:
    s = MT()
    lib().or_mt_init(C.byref(s), seed)
    return s


def thermal3(seed, tag, step, site, ens):
    """The LLG thermal deviates of (site, ens, step): three FP32 values."""
    out = np.zeros(3, np.float32)
    lib().or_thermal3(seed, tag, step, site, ens, out)
    return out


def thermal3_f64(seed, step, site, ens):
    """The FP64 engine's LLG thermal deviates of (site, ens, step)."""
    out = np.zeros(3)
    lib().or_thermal3_f64(seed, step, site, ens, out)
    return out


def noise3_f64(seed, tag, step, site, ens):
    out = np.zeros(3)
    lib().or_noise3_f64(seed, tag, step, site, ens, out)
    return out


# ------------------------------------------------------- the reference itself
# _ref/libuppasd_ref.so: the reference's own field / Depondt / midpoint / RNG
# / moment-update routines compiled from its sources (make target `ref`,
# amdflang), with the C entry points of ref_driver.f90.  Built here when the
# reference sources are present; the built library travels with the tree.
_REF_PATH = os.path.join(_HERE, "_ref", "libuppasd_ref.so")
REF_SOURCES = "/root/reference/source"
_ref = None


def build_ref(force: bool = False):
    """Compile _ref/libuppasd_ref.so; None when the reference sources or
    amdflang are absent (the tests that need it then skip)."""
    if not os.path.isdir(REF_SOURCES) or not os.path.exists("/opt/rocm/bin/amdflang"):
        return None
    args = ["make", "-s", "-C", _HERE, "ref"]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return _REF_PATH


def ref_lib():
    """The reference's compiled routines, or None when not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(_REF_PATH):
            return None
        L = C.CDLL(_REF_PATH)
        vp = C.c_void_p
        L.ref_set_threads.argtypes = [C.c_int]
        L.ref_threads.restype = C.c_int
        L.ref_setup.argtypes = [C.c_int, C.c_int, C.c_int, ip, ip, dp, C.c_int, ip, ip, dp,
                                C.c_int, ip, dp, dp, dp]
        L.ref_set_external_field.argtypes = [vp]
        L.ref_field.argtypes = [dp, dp, dp, C.POINTER(C.c_double)]
        L.ref_seed.argtypes = [C.c_int, C.c_int]
        L.ref_sd_run.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp, dp, dp, dp]
        L.ref_sd_run.restype = C.c_double
        _ref = L
    return _ref


class RefSD:
    """One system in the reference's own module state (one at a time: the
    reference keeps ham_t and its work arrays in modules).  Arrays as for
    OracleHam: (M, N, 3) C order == Fortran (3, N, M); lists (N, W)."""

    def __init__(self, natom, mensemble, nlist, nlistsize, ncoup, dmlist=None, dmlistsize=None,
                 dm_vect=None, taniso=None, eaniso=None, kaniso=None, sb=None):
        L = ref_lib()
        if L is None:
            raise RuntimeError(f"{_REF_PATH} is not built (oracle.build_ref())")
        self.L, self.N, self.M = L, int(natom), int(mensemble)
        nl = np.ascontiguousarray(nlist, dtype=np.int32)
        zi, zd = np.zeros(1, np.int32), np.zeros(3)
        if dmlist is not None:
            dl = np.ascontiguousarray(dmlist, dtype=np.int32)
            ds = np.ascontiguousarray(dmlistsize, dtype=np.int32)
            dv = np.ascontiguousarray(dm_vect, dtype=np.float64)
            mdm = dl.shape[1]
        else:
            dl, ds, dv, mdm = zi, zi, zd, 0
        if taniso is not None:
            ta = np.ascontiguousarray(taniso, dtype=np.int32)
            ea = np.ascontiguousarray(eaniso, dtype=np.float64)
            ka = np.ascontiguousarray(kaniso, dtype=np.float64)
            sbv = np.ascontiguousarray(sb, dtype=np.float64)
        else:
            ta, ea, ka, sbv = zi, zd, zd, zd
        L.ref_setup(self.N, self.M, nl.shape[1], nl,
                    np.ascontiguousarray(nlistsize, dtype=np.int32),
                    np.ascontiguousarray(ncoup, dtype=np.float64), mdm, dl, ds, dv,
                    1 if taniso is not None else 0, ta, ea, ka, sbv)
        L.ref_set_external_field(None)

    def set_external_field(self, ext):
        """(M, N, 3) external_field, or None."""
        self._ext = None if ext is None else np.ascontiguousarray(
            np.broadcast_to(ext, (self.M, self.N, 3)), dtype=np.float64)
        self.L.ref_set_external_field(None if self._ext is None else self._ext.ctypes.data)

    def field(self, emomM, mmom):
        """effective_field_full: beff (M, N, 3) and the energy in mRy."""
        b = np.zeros((self.M, self.N, 3))
        en = C.c_double(0.0)
        self.L.ref_field(np.ascontiguousarray(emomM, dtype=np.float64),
                         np.ascontiguousarray(mmom, dtype=np.float64), b, C.byref(en))
        return b, en.value

    def seed(self, tseed: int, ziggurat: bool = True):
        self.L.ref_seed(int(tseed), 1 if ziggurat else 0)

    def run(self, alg, nsteps, dt, lam, temp, mmom, emom, emomM):
        """nsteps of the sd_mphase body (SDEalgh 1 or 5); emom, emomM and mmom
        ((M, N, 3), (M, N)) are advanced in place.  Returns the last field
        energy (mRy)."""
        lam = np.ascontiguousarray(np.broadcast_to(lam, (self.N,)), dtype=np.float64)
        temp = np.ascontiguousarray(np.broadcast_to(temp, (self.N,)), dtype=np.float64)
        return self.L.ref_sd_run(int(alg), int(nsteps), float(dt), lam, temp, mmom, emom, emomM)


def shells_and_map(lat: LatticeSpec, lines, sym: int, hdim: int = 1, use_nntype: bool = True,
                   nchmax: int = 1, do_ralloy: int = 0) -> dict:
    """read_exchange (restated) and the oracle's setup_nm: the shell arrays
    (nn, redcoord (NT,max_no_shells,3) Fortran order, nntype) and the map nm
    (max_no_equiv, max_no_shells, Natom) C order, nmdim (Natom,
    max_no_shells) C order, with max_no_neigh and max_no_equiv."""
    L = lib()
    ls = lat.c_struct()
    isite = np.ascontiguousarray([int(x[0]) for x in lines], dtype=np.int32)
    jsite = np.ascontiguousarray([int(x[1]) for x in lines], dtype=np.int32)
    ichem = np.ascontiguousarray([int(x[2]) for x in lines], dtype=np.int32)
    jchem = np.ascontiguousarray([int(x[3]) for x in lines], dtype=np.int32)
    r = np.ascontiguousarray([[float(v) for v in x[4:7]] for x in lines]).reshape(-1)
    val = np.ascontiguousarray([[float(v) for v in x[7:7 + hdim]] for x in lines]).reshape(-1)
    sh = Shells()
    if L.or_read_exchange_lines(C.byref(ls), do_ralloy, nchmax, hdim, len(lines), isite, jsite,
                                ichem, jchem, r, val, C.byref(sh)) != 0:
        raise RuntimeError("or_read_exchange_lines failed")
    nm = NMap()
    L.or_setup_nm(C.byref(ls), sym, C.byref(sh), 1 if use_nntype else 0, C.byref(nm))
    N, S, E = nm.Natom, nm.max_no_shells, nm.max_no_equiv
    out = {"NT": sh.NT, "max_no_shells": sh.max_no_shells,
           "nn": _np_from(sh.nn, sh.NT, np.int32).copy(),
           "redcoord": _np_from(sh.redcoord, sh.NT * sh.max_no_shells * 3, np.float64).copy(),
           "nntype": _np_from(sh.nntype, sh.NT * sh.max_no_shells, np.int32).copy(),
           "nm": _np_from(nm.nm, N * S * E, np.int32).reshape(E, S, N).copy(),
           "nmdim": _np_from(nm.nmdim, N * S, np.int32).reshape(N, S).copy(),
           "max_no_neigh": nm.max_no_neigh, "max_no_equiv": E}
    L.or_nmap_free(C.byref(nm))
    L.or_shells_free(C.byref(sh))
    return out


def ref_setup_nm(lat: LatticeSpec, shells: dict, sym: int, use_nntype: bool = True) -> dict:
    """The reference's own setup_nm (oracle/_ref) on the same lattice and
    shell arrays; nm and nmdim in the layouts of shells_and_map."""
    L = ref_lib()
    if L is None:
        raise RuntimeError(f"{_REF_PATH} is not built (oracle.build_ref())")
    L.ref_setup_nm.argtypes = [C.c_int] * 6 + [dp, dp, dp, C.c_char_p, ip, dp, C.c_int, ip, dp,
                                               C.c_int, C.c_int, ip, C.POINTER(C.c_int),
                                               C.POINTER(C.c_int)]
    L.ref_nm_shape.argtypes = [ip]
    L.ref_nm_copy.argtypes = [ip, ip]
    N = lat.natom
    cells = N // lat.NA
    atype = np.ascontiguousarray(np.tile(lat.atype, cells), dtype=np.int32)   # i = b + NA*cell
    bas = np.ascontiguousarray(lat.bas, dtype=np.float64)                     # (NA,3) = (3,NA) F
    mnn, meq = C.c_int(0), C.c_int(0)
    L.ref_setup_nm(N, shells["NT"], lat.NA, *lat.n, np.ascontiguousarray(lat.cell[0]),
                   np.ascontiguousarray(lat.cell[1]), np.ascontiguousarray(lat.cell[2]),
                   "".join(lat.bc).encode(), atype, bas, shells["max_no_shells"], shells["nn"],
                   shells["redcoord"], sym, 1 if use_nntype else 0, shells["nntype"],
                   C.byref(mnn), C.byref(meq))
    d = np.zeros(5, np.int32)
    L.ref_nm_shape(d)
    nm = np.zeros(int(d[0]) * int(d[1]) * int(d[2]), np.int32)
    nmdim = np.zeros(int(d[3]) * int(d[4]), np.int32)
    L.ref_nm_copy(nm, nmdim)
    return {"nm": nm.reshape(int(d[2]), int(d[1]), int(d[0])),
            "nmdim": nmdim.reshape(int(d[4]), int(d[3])),
            "max_no_neigh": mnn.value, "max_no_equiv": meq.value}
