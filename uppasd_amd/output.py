"""Measurement files in the reference's formats, fed by the engine.

  averages.<simid>.out    prn_avrg (prn_averages.f90:569-656)
  cumulants.<simid>.out   calc_and_print_cumulant (prn_averages.f90:919-1095)
  cumulants.<simid>.json
  totenergy.<simid>.out   calc_energy (energy.f90:353-455)
  stdenergy.<simid>.out
  projavgs.<simid>.out    prn_proj_avrg (prn_averages.f90:660-800)

Columns, Fortran edit descriptors and the running means follow the reference,
so the reference's own comparisons (fixed column positions) read these files
unchanged.  MeasurementPhase drives the resident engine with the measurement
cadence of sd_mphase / mc_mphase (sd_driver.f90:517-836): averages and
energies at mod(mstep-1, avrg_step) == 0, cumulants at mod(mstep, cumu_step)
== 0, the energy of a cumulant sample being the last one computed.
"""
from __future__ import annotations

import os

import numpy as np

MUB = 9.274009994e-24      # Parameters/constants.f90
KBOLT = 1.38064852e-23
MRY = 2.179872325e-21
DBL_TOLERANCE = float(np.float32(1e-14))


def fmt_es(x: float, w: int = 16, d: int = 8) -> str:
    """Fortran ESw.d."""
    s = f"{x:.{d}E}"
    mant, exp = s.split("E")
    e = int(exp)
    s = f"{mant}E{'+' if e >= 0 else '-'}{abs(e):02d}" if abs(e) < 100 else f"{mant}{e:+04d}"
    return s.rjust(w)


def fmt_i(i: int, w: int = 8) -> str:
    return str(int(i)).rjust(w)


def fmt_a(s: str, w: int) -> str:
    """Fortran Aw on output: right-justified when shorter, truncated when longer."""
    return s.rjust(w) if len(s) <= w else s[:w]


class Averages:
    """prn_avrg: per measurement, <M> vector and |M| averaged over ensembles
    and the ensemble standard deviation of |M|."""

    def __init__(self, simid: str, outdir: str = "."):
        self.path = os.path.join(outdir, f"averages.{simid.strip()}.out")
        self.started = False

    @staticmethod
    def row(msum: np.ndarray, natom: int):
        """msum (M, 3) = sum_i emomM per ensemble -> (<M>x, <M>y, <M>z, <M>, M_stdv)."""
        mv = np.asarray(msum, dtype=np.float64) / natom
        m = np.sqrt(np.sum(mv * mv, axis=1))
        M = mv.shape[0]
        mvm = mv.sum(axis=0) / M
        mm = m.sum() / M
        ms = (m * m).sum() / M - mm * mm
        ms = 0.0 if ms < DBL_TOLERANCE else float(np.sqrt(ms))
        return float(mvm[0]), float(mvm[1]), float(mvm[2]), float(mm), ms

    def write(self, it: int, msum: np.ndarray, natom: int):
        vals = self.row(msum, natom)
        with open(self.path, "a") as fh:
            if not self.started:
                fh.write(fmt_a("#Iter", 8) + "".join(fmt_a(h, 16) for h in
                         ("<M>_x", "<M>_y", "<M>_z", "<M>", "M_{stdv}")) + "\n")
                self.started = True
            fh.write(fmt_i(it) + "".join(fmt_es(v) for v in vals) + "\n")
        return vals


class Cumulants:
    """calc_and_print_cumulant: running means of |M|, |M|^2, |M|^4 (and of the
    energy per atom) with linearly increasing weights, pooled over ensembles
    in ensemble order; U4, susceptibility and specific heat."""

    def __init__(self, simid: str, temp: float, natom: int, outdir: str = ".",
                 plotenergy: bool = True, temprescale: float = 1.0,
                 temprescalegrad: float = 0.0, cumu_buff: int = 1):
        self.simid = simid.strip()
        self.path = os.path.join(outdir, f"cumulants.{self.simid}.out")
        self.json_path = os.path.join(outdir, f"cumulants.{self.simid[:8].ljust(8)}.json")
        self.temp, self.natom = float(temp), int(natom)
        self.plotenergy, self.trs, self.trg = plotenergy, temprescale, temprescalegrad
        self.cumu_buff = cumu_buff
        self.cumuw = self.cumutotw = 0.0
        self.m = self.m2 = self.m4 = self.e = self.e2 = self.et = 0.0
        self.navrgcum = 0
        self.last = None

    def sample(self, msum: np.ndarray, energy_per_atom=None, write: bool = True,
               exchange_per_atom=None):
        """msum (M, 3); energy_per_atom (M,) total energy per atom in mRy (or
        None); exchange_per_atom (M,) the exchange term alone, the running
        mean printed as <E_{exc}> (the total when None)."""
        ms = np.asarray(msum, dtype=np.float64)
        M, N, T = ms.shape[0], self.natom, self.temp
        for k in range(M):
            me = float(np.sqrt(np.sum(ms[k] * ms[k]))) / N
            m2 = me ** 2
            m4 = m2 ** 2
            self.cumuw += 1.0
            w, tot = self.cumuw, self.cumutotw
            mt = (self.m * tot + me * w) / (tot + w)
            mt2 = (self.m2 * tot + m2 * w) / (tot + w)
            mt4 = (self.m4 * tot + m4 * w) / (tot + w)
            u4 = 1 - (mt4 / 3 / mt2 ** 2)
            self.m, self.m2, self.m4 = mt, mt2, mt4
            if T > 0.0:
                chi = (mt2 - mt ** 2) * MUB ** 2 * N / (KBOLT ** 2) / T
            else:
                chi = (mt2 - mt ** 2) * N * MUB ** 2 / KBOLT
            cv = en = ent = 0.0
            if self.plotenergy and energy_per_atom is not None:
                ek = float(energy_per_atom[k])
                xk = ek if exchange_per_atom is None else float(exchange_per_atom[k])
                en = (self.e * tot + ek * w) / (tot + w)
                en2 = (self.e2 * tot + ek ** 2 * w) / (tot + w)
                ent = (self.et * tot + xk * w) / (tot + w)
                if T > 0.0:
                    cv = (en2 - en ** 2) * MRY ** 2 * N / (KBOLT ** 2) / T ** 2 * \
                        (self.trg * T + self.trs) / self.trs ** 2
                self.e, self.e2, self.et = en, en2, ent
            self.navrgcum += 1
            self.cumutotw += w
            self.last = (mt, mt2, mt4, u4, chi, cv, en, ent)
        if write:
            self._write(M)
        return self.last

    def _write(self, M: int):
        it = self.navrgcum // M
        if (it - 1) % self.cumu_buff != 0:
            return
        mt, mt2, mt4, u4, chi, cv, en, ent = self.last
        with open(self.path, "a") as fh:
            if it == 1:
                fh.write(fmt_a("#Iter", 8) + "".join(fmt_a(h, 16) for h in (
                    "<M>", "<M^2>", "<M^4>", "U_{Binder}", "\\chi", "C_v(tot)", "<E>",
                    "<E_{exc}>", "<E_{lsf}>")) + "\n")
            fh.write(fmt_i(it) + "".join(fmt_es(v) for v in (mt, mt2, mt4, u4, chi, cv, en, ent,
                                                             0.0)) + "\n")
        f = lambda x: f"{x:16.8f}"
        with open(self.json_path, "w") as fh:
            fh.write("{\n")
            fh.write(f'    "temperature"     : {f(self.temp)} ,\n')
            fh.write(f'    "magnetization"   : {f(mt)} ,\n')
            fh.write(f'    "binder_cumulant" : {f(u4)} ,\n')
            fh.write(f'    "energy"          : {f(en)} ,\n' if en != 0.0 else
                     '    "energy"          :  null ,\n')
            fh.write('    "skyrmion_num"    :  null ,\n')
            fh.write('    "skyrmion_std"    :  null ,\n')
            fh.write(f'    "susceptibility"  : {f(chi)} ,\n')
            fh.write(f'    "specific_heat"   : {f(cv)}\n')
            fh.write("}\n")


class TotEnergy:
    """calc_energy's files (energy.f90:353-455) from the term-resolved
    energies per ensemble (Engine.energy_terms: columns exchange, anisotropy,
    DM, Zeeman, per atom in mRy).

    totenergy holds the ensemble mean of each term and of their total; with
    Mensemble > 1 stdenergy holds what calculate_mean_and_deviation
    (energy.f90:577-602) calls the deviation: the sample variance of the
    field-unit energies times mub/mry, i.e. the variance in mRy^2 divided by
    mub/mry.  Terms outside the accelerated Hamiltonian are printed as 0."""

    COLS = ("Tot", "Exc", "Ani", "DM", "PD", "BiqDM", "BQ", "Dip", "Zeeman", "LSF", "Chir",
            "Ring", "SA")

    def __init__(self, simid: str, outdir: str = ".", real_time: bool = False,
                 delta_t: float = 0.0):
        sid = simid.strip()
        self.path = os.path.join(outdir, f"totenergy.{sid}.out")
        self.std_path = os.path.join(outdir, f"stdenergy.{sid}.out")
        self.real_time, self.delta_t = real_time, float(delta_t)

    @staticmethod
    def per_ensemble(terms: np.ndarray) -> np.ndarray:
        """(M, 4) terms -> (M, 13) in column order; Tot summed in the order of
        ene%energy (energy.f90:368-371): xc+dm+pd+bq+ring+ext+ani+dip+bqdm+lsf+chir+sa."""
        t = np.asarray(terms, dtype=np.float64).reshape(-1, 4)
        M = t.shape[0]
        out = np.zeros((M, 13))
        for k in range(M):
            xc, ani, dm, ext = (float(v) for v in t[k])
            tot = xc + dm + 0.0 + 0.0 + 0.0 + ext + ani + 0.0 + 0.0 + 0.0 + 0.0 + 0.0
            out[k] = (tot, xc, ani, dm, 0.0, 0.0, 0.0, 0.0, ext, 0.0, 0.0, 0.0, 0.0)
        return out

    @staticmethod
    def mean_dev(v: np.ndarray):
        """calculate_mean_and_deviation on values already in mRy."""
        v = np.asarray(v, dtype=np.float64)
        n = v.shape[0]
        if n > 1:
            mean = float(np.sum(v)) / n
            dev = float(np.sum((v - mean) ** 2)) / (n - 1.0) / (MUB / MRY)
        else:
            mean, dev = float(v[0]), 0.0
        return mean, dev

    def row(self, terms: np.ndarray):
        e = self.per_ensemble(terms)
        md = [self.mean_dev(e[:, j]) for j in range(13)]
        return [m for m, _ in md], [d for _, d in md]

    def _line(self, fh, it, vals):
        if self.real_time:
            fh.write(fmt_es(it * self.delta_t, 12, 4) + "".join(fmt_es(v) for v in vals) + "\n")
        else:
            fh.write(fmt_i(it) + "".join(fmt_es(v) for v in vals) + "\n")

    def write(self, it: int, terms: np.ndarray):
        """One measurement at iteration it (= mstep-1); header at it == 0."""
        means, devs = self.row(terms)
        first = "#Time" if self.real_time else "#Iter"
        with open(self.path, "a") as fh:
            if it == 0:
                fh.write(fmt_a(first, 8) + "".join(fmt_a(h, 20) for h in self.COLS) + "\n")
            self._line(fh, it, means)
        if np.asarray(terms).reshape(-1, 4).shape[0] > 1:
            with open(self.std_path, "a") as fh:
                if it == 0:
                    fh.write(fmt_a(first, 8) + "".join(fmt_a(h, 20) for h in self.COLS) + "\n")
                self._line(fh, it, devs)
        return means, devs


class ProjAverages:
    """prn_proj_avrg (prn_averages.f90:660-800) over the per-basis-atom sums
    of emomM that calc_mavrg_proj buffers (prn_averages.f90:480-512).

    mode 'Y' projects on atom types (atype (NA,) 1-based), 'A' on basis atoms.
    Per ensemble the projection's average moment per cell is the sum, in
    basis-atom order, of each basis atom's sum over the cells divided by the
    number of cells; the row holds the ensemble means of its norm and
    components and the ensemble deviation of the norm."""

    def __init__(self, simid: str, atype, ncell: int, mode: str = "Y", outdir: str = "."):
        if mode not in ("Y", "A"):
            raise ValueError("mode must be 'Y' (types) or 'A' (basis atoms)")
        self.path = os.path.join(outdir, f"projavgs.{simid.strip()}.out")
        self.atype = np.asarray(atype, dtype=np.int64)
        self.NA, self.ncell, self.mode = self.atype.shape[0], int(ncell), mode

    def rows(self, emomM: np.ndarray):
        """emomM (M, N, 3) in reference site order -> [(k, <M>, stdv, x, y, z)]."""
        e = np.asarray(emomM, dtype=np.float64)
        M = e.shape[0]
        avg = e.reshape(M, self.ncell, self.NA, 3).sum(axis=1)      # (M, NA, 3)
        if self.mode == "Y":
            projs = [[b for b in range(self.NA) if self.atype[b] == k]
                     for k in range(1, int(self.atype.max()) + 1)]
        else:
            projs = [[b] for b in range(self.NA)]
        out = []
        for k, members in enumerate(projs, 1):
            v = np.zeros((M, 3))
            for b in members:
                v = v + avg[:, b] / self.ncell
            m = np.sqrt(np.sum(v * v, axis=1))
            mm = float(np.sum(m)) / M
            ms = float(np.sum(m * m)) / M - mm * mm
            ms = 0.0 if ms < 0 else float(np.sqrt(ms))
            xm, ym, zm = (float(np.sum(v[:, c])) / M for c in range(3))
            out.append((k, mm, ms, xm, ym, zm))
        return out

    def write(self, it: int, emomM: np.ndarray):
        rows = self.rows(emomM)
        with open(self.path, "a") as fh:
            if it == 0:
                fh.write(fmt_a("#Iter", 8) + "Proj" + "".join(fmt_a(h, 16) for h in
                         ("<M>", "M_{stdv}", "<M>_x", "<M>_y", "<M>_z")) + "\n")
            for k, *vals in rows:
                fh.write(fmt_i(it) + fmt_i(k) + "".join(fmt_es(v) for v in vals) + "\n")
        return rows


class MeasurementPhase:
    """Resident measurement phase with the reference's cadence.

    step(n) advances the engine n steps (LLG, SDEalgh `alg`) or sweeps (MC,
    mode 'M'/'H'); measurements are taken before the step of iteration mstep
    exactly where sd_mphase / mc_mphase take them.  With plotenergy the
    term-resolved energies (Engine.energy_terms) are written to totenergy and
    feed the cumulants' <E> and <E_{exc}>."""

    def __init__(self, eng, simid: str, temp: float, avrg_step: int = 100,
                 cumu_step: int = 50, plotenergy: bool = True, outdir: str = ".",
                 cumu_buff: int = 1, final_measure: bool = True, proj=None):
        self.eng, self.temp = eng, float(temp)
        self.avrg_step, self.cumu_step, self.plotenergy = avrg_step, cumu_step, plotenergy
        self.avg = Averages(simid, outdir)
        self.cum = Cumulants(simid, temp, eng.N, outdir, plotenergy, cumu_buff=cumu_buff)
        self.tot = TotEnergy(simid, outdir) if plotenergy else None
        self.energy = self.exchange = None
        self.final_measure = final_measure
        self.proj = proj    # ProjAverages, written with the averages (engine's get_moments)

    def run(self, nstep: int, advance):
        """advance(n) moves the state n steps; iterations 1..nstep."""
        mstep = 1
        while mstep <= nstep:
            if (mstep - 1) % self.avrg_step == 0:
                ms, _ = self.eng.measure()
                self.avg.write(mstep - 1, ms, self.eng.N)
                if self.proj is not None:
                    self.proj.write(mstep - 1, self.eng.get_moments()[1])
            if mstep % self.cumu_step == 0:
                ms, _ = self.eng.measure()
                self.cum.sample(ms, self.energy, exchange_per_atom=self.exchange)
            if self.plotenergy and (mstep - 1) % self.avrg_step == 0:
                terms = self.eng.energy_terms()
                self.tot.write(mstep - 1, terms)
                e = TotEnergy.per_ensemble(terms)
                self.energy, self.exchange = e[:, 0], e[:, 1]
            # next iteration that measures anything
            nxt = nstep + 1
            for m in range(mstep + 1, min(nstep, mstep + max(self.avrg_step, self.cumu_step)) + 1):
                if (m - 1) % self.avrg_step == 0 or m % self.cumu_step == 0:
                    nxt = m
                    break
            advance(nxt - mstep)
            mstep = nxt
        # sd_mphase measures once more after the loop (sd_driver.f90:839),
        # at mstep = nstep + 1, without an energy evaluation; mc_mphase does
        # not (mc_driver.f90:430-440): final_measure=False
        if not self.final_measure:
            return
        if (mstep - 1) % self.avrg_step == 0:
            ms, _ = self.eng.measure()
            self.avg.write(mstep - 1, ms, self.eng.N)
            if self.proj is not None:
                self.proj.write(mstep - 1, self.eng.get_moments()[1])
        if mstep % self.cumu_step == 0:
            ms, _ = self.eng.measure()
            self.cum.sample(ms, self.energy, exchange_per_atom=self.exchange)
