"""eduPIC's picdata.bin (C/eduPIC.cc:726-781): the writer and reader round
trip in both header variants, and the reference's own eduPIC (compiled from
C/eduPIC.cc into oracle/_ref) reads a file this package wrote, runs a cycle
and writes one this package reads back."""
import os
import subprocess

import numpy as np
import pytest

from cudapic_amd import edupic_io, synthetic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EDUPIC = os.path.join(ROOT, "oracle", "_ref", "edupic")


@pytest.mark.parametrize("int_headers", [False, True])
def test_round_trip(tmp_path, int_headers):
    E, I = synthetic.state(1000, 1100, seed=2)
    p = tmp_path / "picdata.bin"
    edupic_io.write_picdata(p, 1.25e-5, 17, E, I, int_headers=int_headers)
    d = edupic_io.read_picdata(p, int_headers=int_headers)
    assert d["time"] == 1.25e-5 and d["cycles"] == 17
    for got, want in ((d["electrons"], E), (d["ions"], I)):
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g, np.asarray(w, np.float64))
    hdr = 8 if not int_headers else 4
    assert os.path.getsize(p) == 8 + hdr + 2 * hdr + 32 * (1000 + 1100)


@pytest.mark.skipif(not os.path.exists(EDUPIC), reason="oracle/_ref/edupic not built")
def test_reference_edupic_reads_and_writes_the_format(tmp_path):
    E, I = synthetic.state(2000, 2100, seed=4)
    period = 1.0 / 13.56e6
    edupic_io.write_picdata(tmp_path / "picdata.bin", 3 * period, 3, E, I)
    r = subprocess.run([EDUPIC, "1"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "data loaded : 2000 electrons 2100 ions, 3 cycles completed" in r.stdout
    d = edupic_io.read_picdata(tmp_path / "picdata.bin")
    assert d["cycles"] == 4
    assert abs(d["time"] - 4 * period) < 1e-3 * period
    n_e, n_i = len(d["electrons"][0]), len(d["ions"][0])
    assert 0 < n_e < 20_000 and 0 < n_i < 20_000
    x = d["electrons"][0]
    assert (x >= 0).all() and (x <= 0.025).all()
