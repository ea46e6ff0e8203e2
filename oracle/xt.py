"""CPU restatement of eduPIC's spatio-temporal (XT) diagnostics (test
infrastructure).

eduPIC's measurement mode (C/eduPIC.cc:126-145 arrays, 514-516 time bin,
558-575 electrons, 603-614 ions, 705-715 grid, 864-903 normalisation)
accumulates, per time bin of N_BIN steps of the RF period:
  grid       pot, E, n_e, n_i of the step                  (:710-713)
  electrons  CIC weight c, c u, c energy, c sigma_ion(energy) |v| dt_e n_gas
             of the mid-step velocity u = vx - 0.5 E(x) dt_e e/m (:561-575)
  ions       c, c u, c energy with u = vx + 0.5 E(x) dt_i e/M (:606-613)
and normalises: grid sums / steps, u and energy / weight, j_e = -u_e n_e e,
j_i = u_i n_i e, power = j E, ionisation rate * W / (A dx) / time (:864-903).
The engine keeps the measured steps of each bin where eduPIC assumes
no_of_cycles * N_BIN; with whole cycles they are the same.

FP64 numpy, loops vectorised with np.add.at; only the checker of the GPU
maps and of their normalisation.  Pinned by closed forms in
tests/test_oracle_xt.py (particles at nodes and between them, a uniform
drift, the half-step correction of a uniform field, the ionisation
threshold).
"""
from __future__ import annotations

import numpy as np

E_CHARGE = 1.60217662e-19
E_MASS = 9.10938356e-31
AR_MASS = 6.63352090e-26
NAMES = ("pot", "efield", "ne", "ni", "je", "ji", "powere", "poweri", "meanee", "meanei", "ioniz")
# raw planes, in the engine's order
POT, EF, NE, NI, CE, UE, EE, IZ, CI, UI, EI = range(11)
DE_CS = 1.0e-3  # eV, C/eduPIC.cc:105


def new_raw(n_g, n_xt):
    return np.zeros((11, n_g, n_xt)), np.zeros(n_xt, np.int64)


def add_grid(raw, count, tb, pot, efield, ne, ni):
    """One measured step's grid quantities into bin tb (C/eduPIC.cc:705-715)."""
    raw[POT][:, tb] += pot
    raw[EF][:, tb] += efield
    raw[NE][:, tb] += ne
    raw[NI][:, tb] += ni
    count[tb] += 1


def add_particles(raw, tb, x, vx, vy, vz, efield, dx, dt, electrons, sigma_ion=None, gas_density=0.0):
    """The CIC-weighted mid-step quantities of one species into bin tb
    (C/eduPIC.cc:553-575 electrons, 596-614 ions); holes (x < 0) skipped."""
    x, vx, vy, vz = (np.asarray(a, np.float64) for a in (x, vx, vy, vz))
    live = x >= 0.0
    x, vx, vy, vz = x[live], vx[live], vy[live], vz[live]
    n_g = raw.shape[1]
    efield = np.asarray(efield, np.float64)
    c0 = x / dx
    p = np.minimum(c0.astype(np.int64), n_g - 2)
    c1 = p + 1.0 - c0
    c2 = c0 - p
    e_x = c1 * efield[p] + c2 * efield[p + 1]
    mass = E_MASS if electrons else AR_MASS
    kick = dt * E_CHARGE / mass
    u = vx - 0.5 * e_x * kick if electrons else vx + 0.5 * e_x * kick
    v2 = u * u + vy * vy + vz * vz
    energy = 0.5 * mass * v2 / E_CHARGE
    planes = (CE, UE, EE, IZ) if electrons else (CI, UI, EI)
    vals = [np.ones_like(u), u, energy]
    if electrons:
        idx = np.minimum((energy / DE_CS + 0.5).astype(np.int64), len(sigma_ion) - 1)
        vals.append(np.asarray(sigma_ion, np.float64)[idx] * np.sqrt(v2) * dt * gas_density)
    for q, v in zip(planes, vals):
        np.add.at(raw[q][:, tb], p, c1 * v)
        np.add.at(raw[q][:, tb], p + 1, c2 * v)


def normalise(raw, count, weight, area, dx, dt_e):
    """norm_all_xt (C/eduPIC.cc:864-903) with per-bin measured steps;
    returns {name: [n_g, n_xt]}."""
    cnt = count.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        f1 = np.where(cnt > 0, 1.0 / cnt, 0.0)
        f2 = np.where(cnt > 0, weight / (area * dx) / (cnt * dt_e), 0.0)
        pot, ef, ne, ni = (raw[q] * f1 for q in (POT, EF, NE, NI))
        ce, ci = raw[CE], raw[CI]
        has_e, has_i = ce > 0, ci > 0
        ue = np.where(has_e, raw[UE] / np.where(has_e, ce, 1.0), 0.0)
        ui = np.where(has_i, raw[UI] / np.where(has_i, ci, 1.0), 0.0)
        je = np.where(has_e, -ue * ne * E_CHARGE, 0.0)
        ji = np.where(has_i, ui * ni * E_CHARGE, 0.0)
        meanee = np.where(has_e, raw[EE] / np.where(has_e, ce, 1.0), 0.0)
        meanei = np.where(has_i, raw[EI] / np.where(has_i, ci, 1.0), 0.0)
        ioniz = np.where(has_e, raw[IZ] * f2, 0.0)
    return dict(zip(NAMES, (pot, ef, ne, ni, je, ji, je * ef, ji * ef, meanee, meanei, ioniz)))
