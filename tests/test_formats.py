"""Image formats on the CPU: PNG, TIFF, DM3/DM4 and MetaImage files read by bin/msgio (host/
image_formats.cpp, the loader bin/matchSeries reads its frames with) against what the reference's
qc::ScalarArray<QC_2D> loader reads from the same files (tests/golden/formats.npz, written by
tests/golden/make_formats.py), bit for bit, and refused where the reference refuses them."""
import os
import subprocess

import numpy as np
import pytest

from test_host import msgio, read_p9

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "formats.npz")


def load_fixture():
    with np.load(GOLD) as z:
        return {k: z[k] for k in z.files}


FIX = load_fixture()
NAMES = sorted(k[len("file:"):] for k in FIX if k.startswith("file:") and not k.endswith(".raw"))


def write_files(tmp_path):
    for k, v in FIX.items():
        if k.startswith("file:"):
            (tmp_path / k[len("file:"):]).write_bytes(v.tobytes())


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("formats")
    write_files(d)
    return d


@pytest.mark.parametrize("name", NAMES)
def test_reads_like_the_reference(files, tmp_path, name):
    out = str(tmp_path / "out.dat")
    r = msgio("convert", str(files / name), out)
    if f"error:{name}" in FIX:
        assert r.returncode != 0, f"{name}: the reference refuses this file"
        assert "Exception" in r.stderr or "not" in r.stderr, r.stderr
        return
    assert r.returncode == 0, r.stderr
    np.testing.assert_array_equal(read_p9(out), FIX[f"array:{name}"])


def test_fixture_covers_every_format():
    kinds = {n.rsplit(".", 1)[1] for n in NAMES}
    assert kinds == {"png", "tif", "dm3", "dm4", "mhd"}
    assert sum(n.endswith(".png") for n in NAMES) >= 25 and sum(n.endswith(".tif") for n in NAMES) >= 40


def test_values_follow_the_formats_definition(files):
    """Where the decoded value follows from the encoder's input alone: the 16-bit TIFF of the
    fixture read by numpy from its uncompressed strip, a grey 8-bit PNG's bytes and an RGB PNG's mean
    (r + g + b) / 3 computed by PIL's decoder (a decoder independent of both programs)."""
    from PIL import Image
    raw = FIX["array:tif_u2_raw.tif"]
    b = FIX["file:tif_u2_raw.tif"].tobytes()
    assert b[:4] == b"II*\x00"
    np.testing.assert_array_equal(np.frombuffer(b[8:8 + raw.size * 2], "<u2").reshape(raw.shape), raw)
    for name in ("png_gray8.png", "png_gray8_adam7.png", "png_pil_L.png"):
        im = np.asarray(Image.open(files / name).convert("L"), dtype=float)
        np.testing.assert_array_equal(im, FIX[f"array:{name}"])
    for name in ("png_rgb8.png", "png_rgba8_adam7.png", "png_pil_RGB.png"):
        rgb = np.asarray(Image.open(files / name).convert("RGB"), dtype=float)
        np.testing.assert_array_equal(rgb.sum(axis=2) / 3.0, FIX[f"array:{name}"])


def test_matchseries_reads_png_frames(tmp_path):
    """The CLI's frame loader dispatches on the suffix like the reference: a missing PNG frame is
    reported with the PNG reader's message, not the array reader's."""
    exe = os.path.join(ROOT, "match-series_amd", "bin", "msgio")
    if not os.path.exists(exe):
        pytest.skip("bin/msgio not built")
    (tmp_path / "a.png").write_bytes(b"GIF89a....")
    r = subprocess.run([exe, "info", str(tmp_path / "a.png")], capture_output=True, text=True)
    assert r.returncode != 0 and "File is not a png" in r.stderr
