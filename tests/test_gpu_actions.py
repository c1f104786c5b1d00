"""matchSeries ACTION 2 (analyzeDeformations) and ACTION 3 (applyDeformation) on the GPU against the
reference program's own outputs for the same deformation files (tests/golden/actions64.npz, written
by tests/golden/make_actions.py): the printed report line for line, the deformed templates, their
PNG previews and the x-derivative norms of ACTION 2, and the templates deformed by the summed
deformations of ACTION 3 (extended with the mean, and nearest neighbour)."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "match-series_amd", "bin", "matchSeries")
GOLD = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, GOLD)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from pngread import read_png_gray8  # noqa: E402
from test_gpu_cli import read_p9, write_p9  # noqa: E402


@pytest.fixture(scope="module")
def gold():
    with np.load(os.path.join(GOLD, "actions64.npz")) as z:
        return {k: z[k] for k in z.files}


def make_tree(tmp_path, g, precision, extra):
    """The frames and the reference program's stage-1/2 deformations and stage-1 median, laid out
    as its ACTION 0 left them; the parameter file with `extra` appended."""
    (tmp_path / "in").mkdir()
    for i, f in enumerate(g["frames"]):
        write_p9(str(tmp_path / "in" / f"frame_{i:02d}.dat.bz2"), f)
    for k, v in g.items():
        if k.startswith(("s1_def_", "s2_def_")):
            st, sub = k[1], k[len("s1_def_"):].replace("_", "-")
            d = tmp_path / "out" / f"stage{st}" / sub
            d.mkdir(parents=True, exist_ok=True)
            for c in range(2):
                write_p9(str(d / f"deformation_06_{c}.dat.bz2"), v[c])
    write_p9(str(tmp_path / "out" / "stage1" / "median.q2bz"), g["s1_median"])
    par = bytes(g["par"]).decode() + f"gpuFieldPrecision {precision}\n" + extra
    (tmp_path / "run.par").write_text(par)


def run(tmp_path, action):
    print(f"matchSeries ACTION {action} in {tmp_path}", flush=True)
    r = subprocess.run([EXE, "run.par", str(action)], cwd=tmp_path, capture_output=True, text=True,
                       env=dict(os.environ, MSG_STRICT="1"), timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stderr


@pytest.mark.parametrize("precision", [64, 32])
def test_analyze_deformations(tmp_path, gold, precision):
    if not os.path.exists(EXE):
        pytest.skip("bin/matchSeries not built")
    g = gold
    make_tree(tmp_path, g, precision, bytes(g["a2_par"]).decode())
    err = run(tmp_path, 2)
    report = [ln for ln in err.splitlines() if ln.startswith(("Translation", "average", "ratio"))]
    # the report only reads the deformation files: identical in either field precision
    assert report == bytes(g["a2_report"]).decode().splitlines()
    for st, n in ((1, 3), (2, 4)):
        sd = tmp_path / "out" / f"stage{st}"
        for i in range(n):
            np.testing.assert_array_equal(read_p9(str(sd / f"def_{i:03d}.q2bz")), g[f"a2_s{st}_xderiv_{i}"])
            ref = g[f"a2_s{st}_defTempl_{i}"]
            mine = read_p9(str(sd / f"defTempl_{i:03d}.dat.bz2"))
            assert np.array_equal(np.isinf(mine), np.isinf(ref))
            fin = np.isfinite(ref)
            if precision == 64:
                np.testing.assert_array_equal(mine[fin], ref[fin])
            else:  # FP32 frames and warp arithmetic
                assert np.abs(mine[fin] - ref[fin]).max() < 2e-6 * np.abs(ref[fin]).max()
            name = bytes(g[f"a2_s{st}_png_name_{i}"]).decode()
            pngs = [f for f in os.listdir(sd) if f.startswith(f"defTempl_{i:03d}-") and f.endswith(".png")]
            assert len(pngs) == 1
            if precision == 64:
                assert pngs[0] == name
            png = read_png_gray8(str(sd / pngs[0])).astype(int)
            diff = np.abs(png - g[f"a2_s{st}_png_{i}"].astype(int))
            assert diff.max() <= (0 if precision == 64 else 1)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("tag", ["mean", "nn"])
def test_apply_deformation(tmp_path, gold, precision, tag):
    if not os.path.exists(EXE):
        pytest.skip("bin/matchSeries not built")
    g = gold
    make_tree(tmp_path, g, precision, bytes(g[f"a3_{tag}_par"]).decode())
    run(tmp_path, 3)
    for st, n in ((1, 3), (2, 4)):
        sd = tmp_path / "out" / f"stage{st}"
        # stage 1 puts the reference first and keeps the template count: frame_03 is not written
        written = sorted(f for f in os.listdir(sd) if f.startswith("frame_") and f.endswith(".q2bz"))
        assert written == [f"frame_{i:02d}.dat.q2bz" for i in range(n)]
        for i in range(n):
            ref = g[f"a3_{tag}_s{st}_{i}"]
            mine = read_p9(str(sd / f"frame_{i:02d}.dat.q2bz"))
            if precision == 64:
                np.testing.assert_array_equal(mine, ref)
            elif tag == "nn":  # nearest neighbour: a sample point on a pixel boundary may pick the other pixel
                assert (np.abs(mine - ref) > 2e-6 * np.abs(ref).max()).mean() < 0.002
            else:
                assert np.abs(mine - ref).max() < 2e-6 * np.abs(ref).max()
