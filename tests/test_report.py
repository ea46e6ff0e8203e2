"""pic_report, eduPIC's run report (C/eduPIC.cc:926-978), on the CPU: the
lines that depend only on the cross-section tables and the discharge
parameters (the collision frequency maxima, the CFL energy) and the plasma
frequency of a given density against the report the reference's own eduPIC
(compiled from C/eduPIC.cc into oracle/_ref) writes; the Debye length and
plasma frequency against their textbook forms; the maxima's energy grid and
masses on constant tables; the condition flags."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EDUPIC = os.path.join(ROOT, "oracle", "_ref", "edupic")
E_CHARGE, E_MASS, AR_MASS = 1.60217662e-19, 9.10938356e-31, 6.63352090e-26


@pytest.fixture(scope="module")
def ca():
    import cudapic_amd
    cudapic_amd.lib()
    return cudapic_amd


def _edupic_info(tmp_path):
    for args in (["0"], ["1", "m"]):
        r = subprocess.run([EDUPIC, *args], cwd=tmp_path, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
    text = (tmp_path / "info.txt").read_text()

    def val(label):
        m = re.search(re.escape(label) + r"\s*=\s*(\S+)", text)
        assert m, label
        return m.group(1)
    return val


@pytest.mark.skipif(not os.path.exists(EDUPIC), reason="oracle/_ref/edupic not built")
def test_table_lines_and_plasma_frequency_match_edupic(ca, tmp_path):
    val = _edupic_info(tmp_path)
    p = ca.params()
    density = float(val("Electron density @ center"))
    r = ca.kernels.report(p, density, 4.0, 1.0 / p.freq, 1, 1, 1, 1)
    assert f"{r['nu_max_e_dt_e']:.3f}" == val("Max. electron coll. frequency * DT_E")
    assert f"{r['nu_max_i_dt_i']:.3f}" == val("Max. ion coll. frequency * DT_I")
    assert f"{r['cfl_energy']:.3f}" == val("Max e- energy for CFL condition")
    # the density is printed to four figures: the frequency goes as its root
    assert abs(r["plasma_freq"] / float(val("Plasma frequency @ center")) - 1.0) < 2e-4
    assert f"{r['wp_dt_e']:.3f}" == val("Plasma frequency @ center * DT_E")
    assert r["conditions_ok"]


def test_textbook_forms(ca):
    """omega_pe = 56.41 sqrt(n) rad/s and lambda_D = 7434 sqrt(T_e / n) m
    (T_e in eV, n in m^-3) with T_e = 2/3 of the mean energy."""
    p = ca.params()
    tables = (np.ones(1_000_000, np.float32), np.ones(1_000_000, np.float32))
    for n, mean_ev in ((1.0e16, 1.5), (3.0e15, 6.0), (2.0e17, 0.3)):
        r = ca.kernels.report(p, n, mean_ev, 2.0e-6, 3000, 50, 1000, 500, tables=tables)
        assert abs(r["plasma_freq"] / (56.41 * math.sqrt(n)) - 1.0) < 2e-4
        assert abs(r["debye_length"] / (7434.0 * math.sqrt(2.0 * mean_ev / 3.0 / n)) - 1.0) < 2e-4
        assert r["dx_over_debye"] == p.dx / r["debye_length"]
        assert abs(r["coll_freq_e"] - 3000 / 2.0e-6 / 1000) < 1e-6 * r["coll_freq_e"]
        assert abs(r["coll_freq_i"] - 50 / 2.0e-6 / 500) < 1e-6 * r["coll_freq_i"]


def test_collision_maxima_on_constant_tables(ca):
    """Constant macroscopic cross sections: the maxima sit at the table's top
    energy (1e6 - 1) meV, with the electron mass for electrons and the
    Ar+/Ar reduced mass for ions."""
    p = ca.params()
    c = 50.0
    tables = (np.full(1_000_000, c, np.float32), np.full(1_000_000, c, np.float32))
    r = ca.kernels.report(p, 1e15, 3.0, 1e-6, 1, 1, 1, 1, tables=tables)
    e_top = 999_999 * 1e-3 * E_CHARGE
    want_e = math.sqrt(2.0 * e_top / E_MASS) * c * p.dt_e
    want_i = math.sqrt(2.0 * e_top / (AR_MASS / 2.0)) * c * p.dt_i
    assert abs(r["nu_max_e_dt_e"] / want_e - 1.0) < 1e-12
    assert abs(r["nu_max_i_dt_i"] / want_i - 1.0) < 1e-12
    # a peak inside the table wins over the top energy
    se = np.full(1_000_000, 1e-3, np.float32)
    se[250_000] = c
    r = ca.kernels.report(p, 1e15, 3.0, 1e-6, 1, 1, 1, 1, tables=(se, tables[1]))
    assert abs(r["nu_max_e_dt_e"] / (math.sqrt(2.0 * 250.0 * E_CHARGE / E_MASS) * c * p.dt_e) - 1.0) < 1e-12


def test_condition_flags(ca):
    p = ca.params()
    _, se, si = ca.kernels.cross_sections(p.gas_density)
    ok = ca.kernels.report(p, 1e15, 3.0, 1e-6, 1, 1, 1, 1, tables=(se, si))
    assert ok["conditions_ok"]
    # dense plasma: omega_pe DT_E > 0.2
    assert not ca.kernels.report(p, 1e21, 3.0, 1e-6, 1, 1, 1, 1, tables=(se, si))["conditions_ok"]
    # cold dense plasma: DX / lambda_D > 1 with omega_pe DT_E still small
    r = ca.kernels.report(p, 3e16, 1e-3, 1e-6, 1, 1, 1, 1, tables=(se, si))
    assert r["wp_dt_e"] < 0.2 and r["dx_over_debye"] > 1.0 and not r["conditions_ok"]
    # no electrons measured at the centre: NaN passes, as in eduPIC
    r = ca.kernels.report(p, 0.0, float("nan"), 1e-6, 1, 1, 1, 1, tables=(se, si))
    assert math.isnan(r["debye_length"]) and r["conditions_ok"]
    # a centre density but an empty EEPF window: the mean energy is eduPIC's
    # 0/0 = NaN (not 0 eV, which would make the Debye length 0 and DX / Debye
    # infinite, a violation eduPIC does not report)
    r = ca.kernels.report(p, 7.4e15, float("nan"), 1e-6, 1, 1, 1, 1, tables=(se, si))
    assert math.isnan(r["debye_length"]) and math.isnan(r["dx_over_debye"]) and r["conditions_ok"]
    assert r["plasma_freq"] > 0
    with pytest.raises(ca.PicError):
        ca.kernels.report(p, 1e15, 3.0, 0.0, 1, 1, 1, 1, tables=(se, si))
