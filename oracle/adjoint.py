"""FP64 CPU restatement of the frozen-event adjoint (test infrastructure).

The engine's adjoint (cudapic_amd/csrc/pic_adjoint.hip) differentiates the
RF step with its discrete events replayed from a tape; the reference has no
differentiation, so there is no reference code to follow.  This module is
the checker: the same model in FP64 numpy -- the replay forward step (CIC
deposit, eduPIC's Thomas solve main.rs:408-494 / C/eduPIC.cc solve_Poisson,
the field formulas of main.rs:471-493, the leapfrog push main.rs:597-603,
the tape) and its adjoint written out by hand, with the whole trajectory
kept in memory (small cases only).

Pinned by: central finite differences of its own FP64 forward (tests/
test_oracle_adjoint.py, several waveform coefficients, with collisions,
absorptions and ionisations on a random tape), and the identity
dJ/dV(t) = sum_k dJ/dE_k(t) / L (the boundary voltage adds the uniform field
V/L at every node, interior central differences and the one-sided boundary
formulas alike).
"""
from __future__ import annotations

import numpy as np

EPSILON0 = 8.85418781e-12
E_CHARGE = 1.60217662e-19


class Model:
    """Geometry and factors of the frozen-event model (FP64 copies of the
    engine's f32 factors)."""

    def __init__(self, n_g, dx, wf, factor_e, factor_i, dt_e, dt_i, n_sub, n_t, t0=0):
        self.n_g, self.dx, self.wf = int(n_g), float(dx), float(wf)
        self.inv_dx = 1.0 / self.dx
        self.L = self.dx * (self.n_g - 1)
        self.c = {"e": float(factor_e), "i": float(factor_i)}
        self.dt = {"e": float(dt_e), "i": float(dt_i)}
        self.n_sub, self.n_t, self.t0 = int(n_sub), int(n_t), int(t0)
        self.node_f = np.ones(self.n_g)
        self.node_f[0] = self.node_f[-1] = 2.0

    @classmethod
    def from_params(cls, p, t0=0):
        return cls(p.n_g, p.dx, p.weight_factor, p.factor_e, p.factor_i, p.dt_e_f, p.dt_i_f, p.n_sub, p.n_t, t0)

    def phase(self, t):
        return 2.0 * np.pi * ((self.t0 + t) % self.n_t) / self.n_t

    def voltage(self, t, wc, ws):
        ph = self.phase(t)
        h = np.arange(1, len(wc) + 1)
        return float(np.dot(wc, np.cos(h * ph)) + np.dot(ws, np.sin(h * ph)))

    def cell(self, x):
        """Cell, weight of the upper node and whether x is inside the gap.
        Positions are clamped to the gap: a particle that leaves it without
        a taped absorption (possible when the waveform differs from the
        record pass's) deposits and feels the field at the wall, and its
        position derivative there is zero."""
        raw = x * self.inv_dx
        pos = np.clip(raw, 0.0, self.n_g - 1.0)
        q = np.minimum(np.floor(pos).astype(np.int64), self.n_g - 2)
        return q, pos - q, (raw > 0.0) & (raw < self.n_g - 1.0)

    def deposit(self, x):
        q, r, _ = self.cell(x)
        n = np.bincount(q, 1.0 - r, self.n_g) + np.bincount(q + 1, r, self.n_g)
        return n * self.wf * self.node_f

    def field(self, n_e, n_i, v):
        ng, dx = self.n_g, self.dx
        rho = E_CHARGE * (n_i - n_e)
        f = -dx * dx / EPSILON0 * rho[1:-1]
        f[0] -= v
        pot = np.zeros(ng)
        pot[0] = v
        pot[1:-1] = thomas(f)
        e = np.empty(ng)
        e[1:-1] = (pot[:-2] - pot[2:]) / (2.0 * dx)
        e[0] = (pot[0] - pot[1]) / dx - rho[0] * dx / (2.0 * EPSILON0)
        e[-1] = (pot[-2] - pot[-1]) / dx + rho[-1] * dx / (2.0 * EPSILON0)
        return e

    def field_adjoint(self, lE):
        """(lambda_rho, dJ/dV) from lambda_E (the transpose of field())."""
        ng, dx = self.n_g, self.dx
        lphi = np.zeros(ng)
        lrho = np.zeros(ng)
        s = 0.5 / dx
        lphi[:-2] += lE[1:-1] * s
        lphi[2:] -= lE[1:-1] * s
        lphi[0] += lE[0] / dx
        lphi[1] -= lE[0] / dx
        lrho[0] = -lE[0] * dx / (2.0 * EPSILON0)
        lphi[-2] += lE[-1] / dx
        lphi[-1] -= lE[-1] / dx
        lrho[-1] = lE[-1] * dx / (2.0 * EPSILON0)
        u = thomas(lphi[1:-1])       # A is symmetric
        lrho[1:-1] = -dx * dx / EPSILON0 * u
        return lrho, lphi[0] - u[0]


def thomas(f):
    """Solves tridiag(1, -2, 1) y = f (Dirichlet zero ends folded into f)."""
    n = len(f)
    w = np.empty(n)
    g = np.empty(n)
    w[0] = 1.0 / -2.0
    g[0] = f[0] / -2.0
    for i in range(1, n):
        den = -2.0 - w[i - 1]
        w[i] = 1.0 / den
        g[i] = (f[i] - g[i - 1]) / den
    y = np.empty(n)
    y[-1] = g[-1]
    for i in range(n - 2, -1, -1):
        y[i] = g[i] - w[i] * y[i + 1]
    return y


def interp(m, efield, x):
    q, r, inside = m.cell(x)
    return r * efield[q + 1] + (1.0 - r) * efield[q], np.where(inside, (efield[q + 1] - efield[q]) * m.inv_dx, 0.0), q, r


def forward(m, state, tape, wc, ws, target, steps, keep=False):
    """Frozen-event forward over `steps` steps.  state = dict(xe, vxe, xi,
    vxi) sized to the slot capacities; tape as Adjoint.tape() returns it.
    Returns (J, nbar[, trajectory])."""
    xe, vxe = state["xe"].astype(np.float64).copy(), state["vxe"].astype(np.float64).copy()
    xi, vxi = state["xi"].astype(np.float64).copy(), state["vxi"].astype(np.float64).copy()
    ide, idi = np.arange(len(xe)), np.arange(len(xi))
    de, di = tape["death_e"], tape["death_i"]
    nbar = np.zeros(m.n_g)
    n_i = np.zeros(m.n_g)
    traj = []
    for t in range(steps):
        ion = t % m.n_sub == 0
        ae = (ide < tape["slots_e"][t]) & (de[:len(xe)] > t)
        n_e = m.deposit(xe[ae])
        if ion:
            ai = (idi < tape["slots_i"][t]) & (di[:len(xi)] > t)
            n_i = m.deposit(xi[ai])
            nbar += n_i
        e = m.field(n_e, n_i, m.voltage(t, wc, ws))
        if keep:
            traj.append({"xe": xe.copy(), "ae": ae, "xi": xi.copy() if ion else None,
                         "ai": ai if ion else None, "E": e.copy()})
        if ion:
            ei = interp(m, e, xi[ai])[0]
            vxi[ai] += m.c["i"] * ei
            xi[ai] += vxi[ai] * m.dt["i"]
        ee = interp(m, e, xe[ae])[0]
        vxe[ae] += m.c["e"] * ee
        xe[ae] += vxe[ae] * m.dt["e"]
        sl = slice(tape["off_e"][t], tape["off_e"][t + 1])
        vxe[tape["slot_e"][sl]] = tape["v_e"][sl]
        if ion:
            sl = slice(tape["off_i"][t], tape["off_i"][t + 1])
            vxi[tape["slot_i"][sl]] = tape["v_i"][sl]
        sl = slice(tape["off_ion"][t], tape["off_ion"][t + 1])
        p, q, r = tape["parent"][sl], tape["new_e"][sl], tape["new_i"][sl]
        xe[q] = xe[p]
        vxe[q] = tape["v_new_e"][sl]
        xi[r] = xe[p]
        vxi[r] = tape["v_new_i"][sl]
    M = (steps + m.n_sub - 1) // m.n_sub
    nbar /= M
    J = float(np.sum((nbar - target) ** 2) / np.sum(target ** 2))
    return (J, nbar, traj) if keep else (J, nbar)


def gradient(m, state, tape, wc, ws, target, steps):
    """(J, dJ/dwave_cos, dJ/dwave_sin, dJ/dV(t)) by the hand-written adjoint."""
    J, nbar, traj = forward(m, state, tape, wc, ws, target, steps, keep=True)
    M = (steps + m.n_sub - 1) // m.n_sub
    lobj = 2.0 * (nbar - target) / (M * np.sum(target ** 2))
    lxe, lve = np.zeros(len(state["xe"])), np.zeros(len(state["xe"]))
    lxi, lvi = np.zeros(len(state["xi"])), np.zeros(len(state["xi"]))
    lheld = np.zeros(m.n_g)
    dJdV = np.zeros(steps)
    dJdV_sum = np.zeros(steps)
    for t in range(steps - 1, -1, -1):
        ion = t % m.n_sub == 0
        tr = traj[t]
        # tape
        sl = slice(tape["off_e"][t], tape["off_e"][t + 1])
        lve[tape["slot_e"][sl]] = 0.0
        if ion:
            sl = slice(tape["off_i"][t], tape["off_i"][t + 1])
            lvi[tape["slot_i"][sl]] = 0.0
        sl = slice(tape["off_ion"][t], tape["off_ion"][t + 1])
        p, q, r = tape["parent"][sl], tape["new_e"][sl], tape["new_i"][sl]
        np.add.at(lxe, p, lxe[q] + lxi[r])
        lxe[q] = lve[q] = 0.0
        lxi[r] = lvi[r] = 0.0
        # push adjoints
        lE = np.zeros(m.n_g)
        for sp, x_all, alive, lx, lv in (("e", tr["xe"], tr["ae"], lxe, lve),
                                          ("i", tr["xi"], tr["ai"], lxi, lvi)):
            if alive is None:
                continue
            x = x_all[alive]
            _, de, qc, rc = interp(m, tr["E"], x)
            mu = lv[alive] + m.dt[sp] * lx[alive]
            lx[alive] += mu * m.c[sp] * de
            lv[alive] = mu
            lE += np.bincount(qc, mu * m.c[sp] * (1.0 - rc), m.n_g) + np.bincount(qc + 1, mu * m.c[sp] * rc, m.n_g)
        lrho, dv = m.field_adjoint(lE)
        dJdV[t] = dv
        dJdV_sum[t] = lE.sum() / m.L
        lr = E_CHARGE * lrho
        ge = -lr * m.wf * m.node_f
        lheld += lr
        # deposit adjoints
        x = tr["xe"][tr["ae"]]
        qc, _, ins = m.cell(x)
        lxe[tr["ae"]] += np.where(ins, m.inv_dx * (ge[qc + 1] - ge[qc]), 0.0)
        if ion:
            gi = (lheld + lobj) * m.wf * m.node_f
            lheld[:] = 0.0
            x = tr["xi"][tr["ai"]]
            qc, _, ins = m.cell(x)
            lxi[tr["ai"]] += np.where(ins, m.inv_dx * (gi[qc + 1] - gi[qc]), 0.0)
    h = np.arange(1, len(wc) + 1)
    ph = np.array([m.phase(t) for t in range(steps)])
    gc = np.array([np.dot(dJdV, np.cos(k * ph)) for k in h])
    gs = np.array([np.dot(dJdV, np.sin(k * ph)) for k in h])
    return J, gc, gs, dJdV, dJdV_sum


def random_tape(m, state, wc, ws, steps, seed=0, p_coll=0.01, p_ion=0.002, v_e=6e5, v_i=300.0, cap_e=None,
                cap_i=None):
    """A tape with collisions, ionisations and absorptions for a state, made
    by an FP64 forward that draws its events from a seeded generator (a
    stand-in for the record pass: the adjoint does not care how the events
    were chosen).  Absorptions are the particles that leave the gap."""
    rng = np.random.default_rng(seed)
    ne0, ni0 = len(state["xe"]), len(state["xi"])
    cap_e = cap_e or ne0 * 2
    cap_i = cap_i or ni0 * 2
    xe = np.zeros(cap_e)
    vxe = np.zeros(cap_e)
    xi = np.zeros(cap_i)
    vxi = np.zeros(cap_i)
    xe[:ne0], vxe[:ne0] = state["xe"], state["vxe"]
    xi[:ni0], vxi[:ni0] = state["xi"], state["vxi"]
    de = np.zeros(cap_e, np.uint32)
    di = np.zeros(cap_i, np.uint32)
    de[:ne0] = 0xffffffff
    di[:ni0] = 0xffffffff
    T1 = steps + 1
    tape = {k: [] for k in ("slot_e", "v_e", "slot_i", "v_i", "parent", "new_e", "new_i", "v_new_e", "v_new_i")}
    off = {k: np.zeros(T1, np.uint32) for k in ("off_e", "off_i", "off_ion", "slots_e", "slots_i")}
    se, si = ne0, ni0
    n_i = np.zeros(m.n_g)
    for t in range(steps):
        ion = t % m.n_sub == 0
        off["off_e"][t], off["off_i"][t], off["off_ion"][t] = len(tape["slot_e"]), len(tape["slot_i"]), len(tape["parent"])
        off["slots_e"][t], off["slots_i"][t] = se, si
        ae = (np.arange(cap_e) < se) & (de > t)
        n_e = m.deposit(xe[ae])
        if ion:
            ai = (np.arange(cap_i) < si) & (di > t)
            n_i = m.deposit(xi[ai])
        e = m.field(n_e, n_i, m.voltage(t, wc, ws))
        if ion:
            idx = np.nonzero(ai)[0]
            vxi[idx] += m.c["i"] * interp(m, e, xi[idx])[0]
            xi[idx] += vxi[idx] * m.dt["i"]
            gone = idx[(xi[idx] < 0) | (xi[idx] > m.L)]
            di[gone] = t + 1
            live = idx[(xi[idx] >= 0) & (xi[idx] <= m.L)]
            hit = live[rng.random(len(live)) < p_coll * 5]
            vxi[hit] = rng.normal(0.0, v_i, len(hit))
            tape["slot_i"] += hit.tolist()
            tape["v_i"] += vxi[hit].tolist()
        idx = np.nonzero(ae)[0]
        vxe[idx] += m.c["e"] * interp(m, e, xe[idx])[0]
        xe[idx] += vxe[idx] * m.dt["e"]
        gone = idx[(xe[idx] < 0) | (xe[idx] > m.L)]
        de[gone] = t + 1
        live = idx[(xe[idx] >= 0) & (xe[idx] <= m.L)]
        u = rng.random(len(live))
        hit = live[u < p_coll]
        vxe[hit] = rng.normal(0.0, v_e, len(hit))
        tape["slot_e"] += hit.tolist()
        tape["v_e"] += vxe[hit].tolist()
        par = live[u < p_ion]
        for pp in par:
            if se >= cap_e or si >= cap_i:
                break
            ve_, vi_ = rng.normal(0.0, v_e), rng.normal(0.0, v_i)
            xe[se], vxe[se], de[se] = xe[pp], ve_, 0xffffffff
            xi[si], vxi[si], di[si] = xe[pp], vi_, 0xffffffff
            tape["parent"].append(pp)
            tape["new_e"].append(se)
            tape["new_i"].append(si)
            tape["v_new_e"].append(ve_)
            tape["v_new_i"].append(vi_)
            se += 1
            si += 1
    off["off_e"][steps], off["off_i"][steps], off["off_ion"][steps] = len(tape["slot_e"]), len(tape["slot_i"]), len(tape["parent"])
    off["slots_e"][steps], off["slots_i"][steps] = se, si
    out = {k: np.array(v, np.int64 if k in ("slot_e", "slot_i", "parent", "new_e", "new_i") else np.float64)
           for k, v in tape.items()}
    out.update(off)
    out["death_e"], out["death_i"] = de, di
    full = {"xe": np.zeros(cap_e), "vxe": np.zeros(cap_e), "xi": np.zeros(cap_i), "vxi": np.zeros(cap_i)}
    full["xe"][:ne0], full["vxe"][:ne0] = state["xe"], state["vxe"]
    full["xi"][:ni0], full["vxi"][:ni0] = state["xi"], state["vxi"]
    return out, full
