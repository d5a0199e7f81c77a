"""Heat-bath parallel tempering (pt_iphase, pt_driver.f90:58-367) on the
batched engine.

The reference keeps one copy of the system per temperature and sweeps the
copies one after the other with mc_evolve in heat-bath mode.  Here the
replicas are ensembles of one context: replica r owns the ensembles
[r*mensemble, (r+1)*mensemble), one sweep of every replica is one launch per
colour (uasd_mc_sweeps with a temperature per ensemble), and the replica
energies come from the device reductions (uasd_measure).  The temperatures
only change at exchange attempts, so the pt_step sweeps between two attempts
run in one call.

Exchange (pt_driver.f90:205-232): every pt_step sweeps, while
ipmcstep/ipmcnstep < 0.9, one uniform draw per replica and two passes over
neighbouring pairs, (1,2), (3,4), ... then (2,3), (4,5), ...; a pair swaps
its temperatures when u <= min(1, exp((E2 - E1)(beta2 - beta1))), with E the
replica's field-routine energy per atom in J (mRy * mry / Natom) and
beta = 1/(kB (temprescale*T + 1e-15)).  The replicas keep their states; the
temperatures move between them.  At the end the first replica of lowest
energy is the one the reference continues with (pt_driver.f90:246-250).

The reference's exponent uses energies per atom, i.e. 1/Natom of the
replica-exchange exponent of the canonical ensembles, so its swaps are
accepted far more often than detailed balance allows and the states found at
one temperature are not canonical samples of it.  per_atom=True (the
default) keeps that rule; per_atom=False uses the replica's total energy,
for which every temperature keeps its canonical distribution.
"""
from __future__ import annotations

import math
import os

import numpy as np

from .output import KBOLT, MRY, fmt_es, fmt_i

__all__ = ["ParallelTempering", "exchange", "select_replica", "write_pt_restart"]


def exchange(temps, repene, natom: int, flipprob, temprescale: float = 1.0,
             per_atom: bool = True):
    """One exchange attempt over all neighbouring pairs.

    temps (nrep) temperatures of the replicas, repene (nrep) their energies
    in mRy (summed over the replica's ensembles), flipprob (nrep) uniform
    draws; per_atom divides the energies by natom as the reference does.
    Returns (new temperatures, hits, trials)."""
    t = [float(x) for x in temps]
    nrep = len(t)
    hits = trials = 0
    for ishift in (0, 1):
        for irep in range(ishift, nrep - 1, 2):
            trials += 1
            beta_1 = 1.0 / KBOLT / (temprescale * t[irep] + 1.0e-15)
            beta_2 = 1.0 / KBOLT / (temprescale * t[irep + 1] + 1.0e-15)
            ene_1 = float(repene[irep]) * MRY
            ene_2 = float(repene[irep + 1]) * MRY
            if per_atom:
                ene_1, ene_2 = ene_1 / natom, ene_2 / natom
            try_pow = (ene_2 - ene_1) * (beta_2 - beta_1)
            # exp overflows only where the minimum is 1 anyway
            trial_prob = 1.0 if try_pow >= 0.0 else min(1.0, math.exp(try_pow))
            if float(flipprob[irep]) <= trial_prob:
                t[irep], t[irep + 1] = t[irep + 1], t[irep]
                hits += 1
    return t, hits, trials


def select_replica(repene) -> int:
    """The first replica whose energy is the lowest (0-based)."""
    e = [float(x) for x in repene]
    mn = min(e)
    irep = 0
    while e[irep] > mn:
        irep += 1
    return irep


def write_pt_restart(simid: str, mstep: int, temps, mmom, emom, outdir: str = ".",
                     ftype: str = "X", mode: str = "H") -> str:
    """pt_prn_mag_conf_clone (pt_driver.f90:370-417): one block per replica,
    rows (i8, f8.3, i8, 2x, 4es16.8) of iteration, temperature, atom, |m|
    and the direction; mmom (nrep, N), emom (nrep, N, 3)."""
    path = os.path.join(outdir, f"pt_restart.{simid.strip()}.out")
    mmom = np.asarray(mmom, dtype=np.float64)
    emom = np.asarray(emom, dtype=np.float64)
    nrep, natom = mmom.shape
    with open(path, "w") as fh:
        fh.write("#" * 80 + "\n")
        fh.write(f"# File type: {ftype}\n")
        fh.write(f"# Simulation type: {mode}\n")
        fh.write(f"# Number of atoms:  {natom:8d}\n")
        fh.write(f"# Number of ensembles:  {nrep:8d}\n")
        fh.write("#" * 80 + "\n")
        fh.write("   #iterTemp   iatom" + "".join(h.rjust(16) for h in
                                              ("|Mom|", "M_x", "M_y", "M_z")) + "\n")
        for r in range(nrep):
            for j in range(natom):
                fh.write(fmt_i(mstep) + f"{float(temps[r]):8.3f}" + fmt_i(j + 1) + "  " +
                         "".join(fmt_es(v) for v in (mmom[r, j], *emom[r, j])) + "\n")
    return path


class ParallelTempering:
    """pt_iphase on one context of nrep*mensemble ensembles.

    eng: an Engine holding the Hamiltonian and the start moments of every
    ensemble; temps: the nrep start temperatures (ipTemp); uniform(n): the
    source of the exchange draws (rng_uniform in the reference)."""

    def __init__(self, eng, temps, mensemble: int = 1, pt_step: int = 10, seed: int = 1,
                 uniform=None, mode: str = "H", temprescale: float = 1.0,
                 per_atom: bool = True):
        self.eng = eng
        self.temps = [float(x) for x in temps]
        self.nrep, self.mensemble = len(self.temps), int(mensemble)
        if self.nrep * self.mensemble != eng.M:
            raise ValueError("the context must hold nrep*mensemble ensembles")
        if pt_step < 1:
            raise ValueError("pt_step must be >= 1")
        self.pt_step, self.seed, self.mode = int(pt_step), int(seed), mode
        self.temprescale, self.per_atom = float(temprescale), bool(per_atom)
        if uniform is None:
            gen = np.random.default_rng(seed)
            uniform = gen.random
        self.uniform = uniform
        self.n_hit = self.n_trial = 0
        self.repene = None
        self.history = []   # (sweep, temperatures, replica energies, draws) per attempt

    def ensemble_temps(self) -> np.ndarray:
        return np.repeat(np.asarray(self.temps, dtype=np.float64), self.mensemble)

    def replica_energies(self) -> np.ndarray:
        """Field-routine energy of each replica in mRy, summed over its ensembles."""
        _, en = self.eng.measure(with_energy=True)
        return np.asarray(en, dtype=np.float64).reshape(self.nrep, self.mensemble).sum(axis=1)

    def run(self, nsteps: int) -> int:
        """ipmcnstep sweeps with exchange attempts; returns the replica the
        reference would continue with."""
        step = 0
        while step < nsteps:
            n = min(self.pt_step - step % self.pt_step, nsteps - step)
            self.eng.mc_sweeps(self.mode, self.ensemble_temps(), n, self.seed)
            step += n
            if step % self.pt_step == 0:
                self.repene = self.replica_energies()
                u = None
                if step / (1.0 * nsteps) < 0.90:
                    u = np.asarray(self.uniform(self.nrep), dtype=np.float64)
                self.history.append((step, list(self.temps), self.repene.copy(), u))
                if u is not None:
                    self.temps, h, t = exchange(self.temps, self.repene, self.eng.N, u,
                                                self.temprescale, self.per_atom)
                    self.n_hit += h
                    self.n_trial += t
        self.repene = self.replica_energies()
        return select_replica(self.repene)

    def hit_rate(self) -> float:
        return self.n_hit / self.n_trial if self.n_trial else 0.0
