"""matchSeismicSeries' front end on the CPU (it runs before any device call): the initial average made
by interleaving the reference's rows with the template's (justStackIniAverage), identical to the
reference program's (tests/golden/seismic.npz), and the refusals of what this build does not run."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "match-series_amd", "bin", "matchSeismicSeries")
GOLD = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, GOLD)
import make_golden as G  # noqa: E402


@pytest.fixture(scope="module")
def g():
    if not os.path.exists(EXE):
        pytest.skip("bin/matchSeismicSeries not built")
    with np.load(os.path.join(GOLD, "seismic.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def setup(tmp_path, g, extra="", frames=None):
    (tmp_path / "in").mkdir()
    for i, f in enumerate(g["frames"] if frames is None else frames):
        G.write_p9(str(tmp_path / "in" / f"frame_{i:02d}.dat.bz2"), f)
    par = bytes(g["par"]).decode()
    for line in extra.splitlines():
        key = line.split()[0]
        par = "\n".join(ln for ln in par.splitlines() if not ln.startswith(key + " ")) + "\n"
        par += line + "\n" if len(line.split()) > 1 else ""
    (tmp_path / "run.par").write_text(par)


def run(tmp_path):
    return subprocess.run([EXE, "run.par"], cwd=tmp_path, capture_output=True, text=True, timeout=60,
                          env=dict(os.environ, MSG_STRICT="1"))


def test_initial_average_by_stacking(tmp_path, g):
    setup(tmp_path, g)
    run(tmp_path)  # written before any device call (without a GPU the run stops at the first one)
    np.testing.assert_array_equal(G.read_p9(str(tmp_path / "out" / "initialAverage.q2bz")), g["initial_average"])


@pytest.mark.parametrize("extra,msg", [
    ("dyadicDepth 5", "dyadic path only"),
    ("forceNonDyadicXShiftYConst 1", "full-resolution step"),
    ("laplaceWeight 0.5", "laplaceWeight"),
    ("justStackIniAverage 0", "aligned interleaving"),
    ("refLineDistance 5", "refLineDistance too big"),
])
def test_refusals(tmp_path, g, extra, msg):
    setup(tmp_path, g, extra)
    r = run(tmp_path)
    assert r.returncode == 1 and msg in r.stderr, r.stderr[-1000:]


def test_refuses_frames_with_empty_rows(tmp_path, g):
    fr = g["frames"].copy()
    fr[2, 10] = 0
    setup(tmp_path, g, frames=fr)
    r = run(tmp_path)
    assert r.returncode == 1 and "empty row" in r.stderr, r.stderr[-1000:]
