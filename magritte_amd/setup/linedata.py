"""Line data of one line-producing species in SI units.

Restates linedata_from_LAMDA_file (magritte/setup/setup.py:370-507):
energies from cm^-1 to J (x 100 h c), Cd from cm^3 s^-1 to m^3 s^-1,
Ce by detailed balance at the table temperatures,
Ce = Cd g_i / g_j exp(-(E_i - E_j)/(k T)), line frequency (E_i - E_j)/h,
Bs = A c^2 / (2 h nu^3), Ba = g_i / g_j Bs.  Levels are 0-based.
"""
from dataclasses import dataclass, field
from typing import List

import numpy as np

# src/Tools/constants.hpp:34-37
CC = 2.99792458E+8
HH = 6.62607004E-34
KB = 1.38064852E-23
AMU = 1.66053904E-27


@dataclass
class CollisionPartner:
    num: int                      # species index of the partner
    orth_or_para: str             # 'n', 'o' or 'p'
    icol: np.ndarray              # [ncol] upper level
    jcol: np.ndarray              # [ncol] lower level
    tmp: np.ndarray               # [ntmp] K
    Cd: np.ndarray                # [ntmp, ncol] m^3 s^-1
    Ce: np.ndarray                # [ntmp, ncol] m^3 s^-1


@dataclass
class Linedata:
    num: int
    sym: str
    inverse_mass: float
    energy: np.ndarray            # [nlev] J
    weight: np.ndarray            # [nlev]
    irad: np.ndarray              # [nrad]
    jrad: np.ndarray              # [nrad]
    A: np.ndarray                 # [nrad] s^-1
    colpar: List[CollisionPartner] = field(default_factory=list)

    @property
    def nlev(self):
        return int(self.energy.shape[0])

    @property
    def nrad(self):
        return int(self.irad.shape[0])

    @property
    def frequency(self):
        return np.array([(self.energy[i] - self.energy[j]) / HH
                         for i, j in zip(self.irad, self.jrad)], dtype=np.float64)

    @property
    def Bs(self):
        f = self.frequency
        return np.array([self.A[k] * CC ** 2 / (2.0 * HH * (f[k]) ** 3)
                         for k in range(self.nrad)], dtype=np.float64)

    @property
    def Ba(self):
        bs = self.Bs
        return np.array([self.weight[i] / self.weight[j] * bs[k]
                         for k, (i, j) in enumerate(zip(self.irad, self.jrad))],
                        dtype=np.float64)


def detailed_balance_Ce(Cd, icol, jcol, weight, energy, tmp):
    """Ce[t][k] = Cd[t][k] g_i/g_j exp(-(E_i - E_j)/(k T_t)) (setup.py:447-452)."""
    Ce = np.zeros_like(Cd)
    for t in range(Cd.shape[0]):
        for k in range(Cd.shape[1]):
            i, j = icol[k], jcol[k]
            Ce[t][k] = Cd[t][k] * weight[i] / weight[j] * np.exp(
                -(energy[i] - energy[j]) / (KB * tmp[t]))
    return Ce


def read_lamda(path: str, species_index, num: int = 1) -> Linedata:
    """Parse a LAMDA file.  ``species_index`` maps a collision-partner name
    (after stripping the o-/p- prefix) to its species number."""
    with open(path) as fh:
        lines = [ln.rstrip("\n") for ln in fh]

    def col(start, n, c, typ):
        return [typ(lines[start + q].split()[c]) for q in range(n)]

    sym = lines[1].split()[0]
    mass = float(lines[3].split()[0])
    nlev = int(lines[5].split()[0])
    energy = np.array(col(7, nlev, 1, float)) * 1.0E+2 * HH * CC
    weight = np.array(col(7, nlev, 2, float))
    nrad = int(lines[8 + nlev].split()[0])
    irad = np.array(col(10 + nlev, nrad, 1, int)) - 1
    jrad = np.array(col(10 + nlev, nrad, 2, int)) - 1
    A = np.array(col(10 + nlev, nrad, 3, float))
    nlr = nlev + nrad
    ncolpar = int(lines[11 + nlr].split()[0])
    ind = 13 + nlr
    colpars = []
    for _ in range(ncolpar):
        text = lines[ind]
        words = text.replace("-", " ").replace("+", " ").split()
        partner = words[1] if len(words) > 1 else words[0]
        op = "n"
        if partner[:1] in ("o", "p") and partner[1:2].upper() == "H":
            op, partner = partner[0], partner[1:]
        ncol = int(lines[ind + 2].split()[0])
        ntmp = int(lines[ind + 4].split()[0])
        tmp = np.array([float(x) for x in lines[ind + 6].split()[:ntmp]])
        icol = np.array(col(ind + 8, ncol, 1, int)) - 1
        jcol = np.array(col(ind + 8, ncol, 2, int)) - 1
        Cd = np.array([col(ind + 8, ncol, 3 + t, float) for t in range(ntmp)]) * 1.0E-6
        Ce = detailed_balance_Ce(Cd, icol, jcol, weight, energy, tmp)
        colpars.append(CollisionPartner(species_index(partner), op, icol, jcol, tmp, Cd, Ce))
        ind += 9 + ncol
    return Linedata(num=num, sym=sym, inverse_mass=float(1.0 / mass), energy=energy,
                    weight=weight, irad=irad, jrad=jrad, A=A, colpar=colpars)


def two_level(nu=115.27e9, A=7.2e-8, g=(1.0, 3.0), Cd_cgs=3e-11, T_table=(20.0,),
              mass=28.0, num=1, partner_num=2) -> Linedata:
    """Two-level species: E0 = 0, E1 = h nu, one collision partner with a
    constant de-excitation rate (cm^3 s^-1) on the given temperature table."""
    energy = np.array([0.0, HH * nu])
    weight = np.array(g, dtype=np.float64)
    irad, jrad = np.array([1]), np.array([0])
    tmp = np.array(T_table, dtype=np.float64)
    Cd = np.full((len(tmp), 1), Cd_cgs * 1.0E-6)
    Ce = detailed_balance_Ce(Cd, irad, jrad, weight, energy, tmp)
    cp = CollisionPartner(partner_num, "n", irad.copy(), jrad.copy(), tmp, Cd, Ce)
    return Linedata(num=num, sym="CO", inverse_mass=1.0 / mass, energy=energy, weight=weight,
                    irad=irad, jrad=jrad, A=np.array([A]), colpar=[cp])


def linear_rotor(nlev=7, B=57.6e9, A1=7.2e-8, mass=28.0, num=1, partner_num=2,
                 T_table=None, C0_cgs=1e-10) -> Linedata:
    """Rigid linear rotor: E_J = h B J (J+1), g = 2J+1, lines J -> J-1 at 2BJ
    with A_J = A1 J^3; collisional de-excitation for every pair i > j,
    Cd = C0 exp(-(E_i - E_j)/(k T)) on the table, Ce by detailed balance."""
    J = np.arange(nlev)
    energy = HH * B * J * (J + 1.0)
    weight = 2.0 * J + 1.0
    irad = np.arange(1, nlev)
    jrad = np.arange(0, nlev - 1)
    A = A1 * (irad.astype(np.float64) ** 3)
    if T_table is None:
        T_table = np.logspace(1.0, np.log10(3000.0), 12)
    tmp = np.asarray(T_table, dtype=np.float64)
    pairs = [(i, j) for i in range(1, nlev) for j in range(i)]
    icol = np.array([p[0] for p in pairs])
    jcol = np.array([p[1] for p in pairs])
    Cd = np.zeros((len(tmp), len(pairs)))
    for t, T in enumerate(tmp):
        for k, (i, j) in enumerate(pairs):
            Cd[t, k] = C0_cgs * 1.0E-6 * np.exp(-(energy[i] - energy[j]) / (KB * T))
    Ce = detailed_balance_Ce(Cd, icol, jcol, weight, energy, tmp)
    cp = CollisionPartner(partner_num, "n", icol, jcol, tmp, Cd, Ce)
    return Linedata(num=num, sym="CO", inverse_mass=1.0 / mass, energy=energy, weight=weight,
                    irad=irad, jrad=jrad, A=A, colpar=[cp])


def split_rotor(nlev_J=4, B=57.6e9, eps=2.9e5, A1=7.2e-8, mass=28.0, num=1, partner_num=2,
                T_table=None, C0_cgs=1e-10) -> Linedata:
    """Linear rotor whose levels J >= 1 are split in two, E = h (B J (J+1) -/+ eps J),
    g = 2J+1 each: the lines J -> J-1 form blends of 2 (J = 1) or 4 (J >= 2)
    lines a few line widths apart at 2BJ (eps = 290 kHz: about one thermal width
    of CO at 100 K at 230 GHz), A = A1 J^3 / 2 each; collisional
    de-excitation for every pair i > j as linear_rotor."""
    levels = [(0, 0)] + [(J, s) for J in range(1, nlev_J) for s in (-1, 1)]
    energy = np.array([HH * (B * J * (J + 1.0) + s * eps * J) for J, s in levels])
    weight = np.array([2.0 * J + 1.0 for J, _ in levels])
    rad = [(i, j) for i, (Ji, _) in enumerate(levels) for j, (Jj, _) in enumerate(levels) if Jj == Ji - 1]
    irad = np.array([r[0] for r in rad])
    jrad = np.array([r[1] for r in rad])
    A = np.array([A1 * levels[i][0] ** 3 / 2.0 for i in irad])
    if T_table is None:
        T_table = np.logspace(1.0, np.log10(3000.0), 12)
    tmp = np.asarray(T_table, dtype=np.float64)
    nl = len(levels)
    pairs = [(i, j) for i in range(1, nl) for j in range(i)]
    icol = np.array([q[0] for q in pairs])
    jcol = np.array([q[1] for q in pairs])
    Cd = np.zeros((len(tmp), len(pairs)))
    for t, T in enumerate(tmp):
        for k, (i, j) in enumerate(pairs):
            Cd[t, k] = C0_cgs * 1.0E-6 * np.exp(-(energy[i] - energy[j]) / (KB * T))
    Ce = detailed_balance_Ce(Cd, icol, jcol, weight, energy, tmp)
    cp = CollisionPartner(partner_num, "n", icol, jcol, tmp, Cd, Ce)
    return Linedata(num=num, sym="CO", inverse_mass=1.0 / mass, energy=energy, weight=weight,
                    irad=irad, jrad=jrad, A=A, colpar=[cp])
