"""The reference's own matchSeries driver with msg::GpuRegistrationMultilevelDescent as its
registration object (match-series_amd/adapter/gpuRegistration.h, built into oracle/_ref/matchSeriesGpu
from the reference's sources by oracle/Makefile's "adapter" target): every stage-1 file against the
reference program's run of the same series (tests/golden/e2e64.npz, e2e64_ncc.npz), FP64 fields.

The reference code does everything but the level solves here (loading, normalisation, pyramids,
prolongation, composition, averaging, output), so this checks the adapter's boundary: images,
deformation and parameters in, deformation, energy and per-iteration record out."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "matchSeriesGpu")
GOLD = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_gpu_cli import check_stage, write_p9  # noqa: E402


@pytest.mark.parametrize("fixture", ["e2e64", "e2e64_ncc"])
def test_reference_driver_over_gpu_level_solves(tmp_path, fixture):
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/matchSeriesGpu not built (needs the reference's sources)")
    with np.load(os.path.join(GOLD, f"{fixture}.npz")) as z:
        g = {k: z[k] for k in z.files}
    (tmp_path / "in").mkdir()
    for i, f in enumerate(g["frames"]):
        write_p9(str(tmp_path / "in" / f"frame_{i:02d}.dat.bz2"), f)
    (tmp_path / "run.par").write_text(bytes(g["par"]).decode() + "gpuFieldPrecision 64\n")
    print(f"matchSeriesGpu {fixture} in {tmp_path}", flush=True)
    r = subprocess.run([EXE, "run.par"], cwd=tmp_path, capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, QUOC_NO_SYSTEM_PAUSE="ON"))
    assert r.returncode == 0 and "Exception" not in r.stderr, r.stderr[-2000:]
    check_stage(tmp_path / "out", g, 1, 64, 1e-6, 1e-6)
