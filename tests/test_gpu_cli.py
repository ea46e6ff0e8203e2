"""The C++ host CLI: the reference's run interface and result files
(Rust-CUDA/src/main.rs:1498-1627)."""
import glob
import re
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cudapic_amd", "cudapic")


@pytest.mark.parametrize("path", ["fused", "reference"])
def test_cli_run_writes_reference_outputs(tmp_path, path):
    out = tmp_path / "results"
    ck = tmp_path / "state.ckpt"
    r = subprocess.run([CLI, "3", "--measure", "2", "--path", path, "--out", str(out),
                        "--save", str(ck)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "final particles" in r.stdout
    d = glob.glob(str(out / "*"))
    assert len(d) == 1
    names = sorted(os.path.basename(f) for f in glob.glob(d[0] + "/*"))
    assert any(n.startswith("0003_") and n.endswith("_electrons.csv") for n in names)
    assert any(n.startswith("0003_") and n.endswith("_ions.csv") for n in names)
    for prefix in ("particle_growth_", "density_avg_", "eepf_", "ifed_", "info_"):
        assert any(n.startswith(prefix) for n in names), (prefix, names)
    dens = [f for f in glob.glob(d[0] + "/density_avg_*")][0]
    head = open(dens).readline().strip()
    assert head == "x,n_e,n_i"
    a = np.loadtxt(dens, delimiter=",", skiprows=1)
    assert a.shape == (400, 3) and (a[:, 1] >= 0).all() and a[:, 2].sum() > 0
    eepf = np.loadtxt(glob.glob(d[0] + "/eepf_*")[0], delimiter=",", skiprows=1)
    assert eepf.shape == (2000, 2)
    e = glob.glob(d[0] + "/*_electrons.csv")[0]
    assert open(e).readline().strip() == "x,vx,vy,vz"
    assert ck.exists()
    # resume from the checkpoint
    r2 = subprocess.run([CLI, "1", "--path", path, "--out", str(tmp_path / "r2"), "--load", str(ck)],
                        capture_output=True, text=True, timeout=300)
    assert r2.returncode == 0, r2.stdout + r2.stderr
    assert "state loaded" in r2.stdout


def test_cli_writes_xt_maps(tmp_path):
    """--xt 20 writes eduPIC's eleven spatio-temporal maps (N_G rows of
    N_T / 20 time bins) for the measured cycle; averaged over the bins, the
    density maps are the density_avg profiles of the same cycle (every bin
    holds 20 steps and one ion step)."""
    out = tmp_path / "results"
    r = subprocess.run([CLI, "2", "--measure", "1", "--xt", "20", "--out", str(out)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    d = glob.glob(str(out / "*"))[0]
    for name in ("pot", "efield", "ne", "ni", "je", "ji", "powere", "poweri", "meanee", "meanei", "ioniz"):
        a = np.loadtxt(os.path.join(d, name + "_xt.dat"))
        assert a.shape == (400, 200), (name, a.shape)
        assert np.isfinite(a).all()
    avg = np.loadtxt(glob.glob(os.path.join(d, "density_avg_*"))[0], delimiter=",", skiprows=1)
    for name, col in (("ne", 1), ("ni", 2)):
        m = np.loadtxt(os.path.join(d, name + "_xt.dat"))
        assert (m >= 0).all() and avg[:, col].sum() > 0
        # %e keeps 7 significant digits
        np.testing.assert_allclose(m.mean(axis=1), avg[:, col], rtol=2e-6, atol=1e-6 * avg[:, col].max())
    # the run report (C/eduPIC.cc:926-1025): the table-only lines as the
    # reference's eduPIC prints them for the reference case, the centre
    # density of density_avg and the <j*E> averages of the power maps
    info = open(glob.glob(os.path.join(d, "info_*"))[0]).read()

    def val(label):
        m = re.search(re.escape(label) + r"\s*=\s*(\S+)", info)
        assert m, (label, info)
        return m.group(1)
    assert val("Max. electron coll. frequency * DT_E") == "0.013"
    assert val("Max. ion coll. frequency * DT_I") == "0.020"
    assert val("Max e- energy for CFL condition") == "32.834"
    assert "VIOLATED" not in info
    assert abs(float(val("Electron density @ center")) / avg[200, 1] - 1.0) < 1e-3
    assert float(val("Debye length @ center")) > 0 and float(val("Electron collision frequency")) > 0
    for label, name in (("Electron power density (average)", "powere"), ("Ion power density (average)", "poweri")):
        m = np.loadtxt(os.path.join(d, name + "_xt.dat"))
        # %12.3e keeps 4 figures; the map's %e rounding moves the mean by far less
        assert abs(float(val(label)) - m.mean()) <= 1e-3 * abs(m.mean()) + 1e-6 * np.abs(m).mean(), label


def test_cli_usage_error():
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr


def test_cli_imports_and_exports_edupic_picdata(tmp_path):
    """A converged reference-case state in eduPIC's picdata.bin goes in, one
    cycle runs, and picdata.bin comes out with the cycle and time advanced."""
    from cudapic_amd import edupic_io, synthetic
    E, I = synthetic.converged_state()
    period = 1.0 / 13.56e6
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    edupic_io.write_picdata(src, 1500 * period, 1500, E, I)
    r = subprocess.run([CLI, "1", "--capacity", "400000", "--import-edupic", str(src), "--export-edupic", str(dst),
                        "--out", str(tmp_path / "res")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"eduPIC data loaded: {len(E[0])} electrons {len(I[0])} ions, 1500 cycles" in r.stdout
    d = edupic_io.read_picdata(dst)
    assert d["cycles"] == 1501 and abs(d["time"] - 1501 * period) < 1e-6 * period
    # one cycle of a converged discharge changes the populations by ~1 %
    assert abs(len(d["electrons"][0]) / len(E[0]) - 1.0) < 0.05
    assert abs(len(d["ions"][0]) / len(I[0]) - 1.0) < 0.05
    x = d["ions"][0]
    assert (x >= 0).all() and (x <= 0.025).all()


def test_cli_reference_case_ten_cycles():
    """The plain eduPIC reference case from the CLI: 1000 electrons + 1000
    ions at rest, 10 RF cycles, 8 seeds.  The final counts are compared
    with the CPU oracle's (the C restatement of the reference step, 16
    seeds, tests/golden/oracle_conv10.json) within 3.5 combined standard
    errors of the means, and with eduPIC's published run (Figure9_conv.dat,
    cycle 10: 6091 electrons, 9444 ions) within 3.5 seed spreads of a single
    run."""
    import json
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "oracle_conv10.json")))
    counts = []
    for seed in range(8):
        r = subprocess.run([CLI, "10", "--seed", str(seed)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        m = re.search(r"final particles: (\d+) electrons, (\d+) ions", r.stdout)
        assert m, r.stdout
        counts.append((int(m.group(1)), int(m.group(2))))
    c = np.array(counts, np.float64)
    print("10 cycles, 8 seeds:", counts)
    for k, (name, pub) in enumerate((("e", 6091.0), ("i", 9444.0))):
        mean, sd = c[:, k].mean(), c[:, k].std(ddof=1)
        o_mean, o_sd = ref[f"mean_{name}"][-1], ref[f"sd_{name}"][-1]
        se = np.hypot(sd / np.sqrt(len(c)), o_sd / np.sqrt(ref["seeds"]))
        print(f"{name}: engine {mean:.0f} +- {sd:.0f}, oracle {o_mean:.0f} +- {o_sd:.0f}, published {pub:.0f}")
        assert abs(mean - o_mean) <= 3.5 * se, (name, mean, o_mean, se)
        assert abs(mean - pub) <= 3.5 * np.hypot(o_sd, o_sd / np.sqrt(len(c))), (name, mean, pub)
