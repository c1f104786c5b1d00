"""Host layer on the CPU: the reference's array file formats (bin/msgio over host/array_io.cpp),
the .par grammar and the matchSeries error convention (bin/matchSeries, no device work)."""
import bz2
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "match-series_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")


def msgio(*args):
    exe = os.path.join(BIN, "msgio")
    if not os.path.exists(exe):
        pytest.skip("bin/msgio not built")
    return subprocess.run([exe] + list(args), capture_output=True, text=True)


def read_p9(path):
    b = open(path, "rb").read()
    if path.endswith((".bz2", ".q2bz")):
        b = bz2.decompress(b)
    fields, pos = [], 0
    while len(fields) < 3:
        e = b.index(b"\n", pos)
        line, pos = b[pos:e], e + 1
        if not line.startswith(b"#"):
            fields.append(line)
    W, H = map(int, fields[1].split())
    return np.frombuffer(b[pos:pos + 8 * W * H], dtype="<f8").reshape(H, W)


def write_pnm(path, magic, arr, maxval, payload):
    hdr = f"{magic}\n# test file\n{arr.shape[1]} {arr.shape[0]}\n{maxval}\n".encode()
    data = hdr + payload
    with open(path, "wb") as f:
        f.write(bz2.compress(data) if path.endswith(".bz2") else data)


def test_reads_the_reference_image_fixture(tmp_path):
    # the reference's own test image (examples/testdata/image_129.pgm.bz2) as its loader reads it
    out = str(tmp_path / "img.dat.bz2")
    r = msgio("convert", os.path.join(GOLD, "image_129.pgm.bz2"), out)
    assert r.returncode == 0, r.stderr
    with np.load(os.path.join(GOLD, "image129.npz")) as z:
        np.testing.assert_array_equal(read_p9(out), z["data"])


@pytest.mark.parametrize("kind", ["P9", "P8", "P5", "P5w", "P2", "Pa"])
@pytest.mark.parametrize("compress", [False, True])
def test_format_roundtrip(tmp_path, kind, compress):
    rng = np.random.default_rng(1)
    a = rng.integers(0, 250, size=(7, 11)).astype(np.float64)
    src = str(tmp_path / ("in.pgm" + (".bz2" if compress else "")))
    if kind == "P9":
        a = rng.random((7, 11))
        write_pnm(src, "P9", a, 1, a.astype("<f8").tobytes())
    elif kind == "P8":
        a = rng.random((7, 11)).astype(np.float32).astype(np.float64)
        write_pnm(src, "P8", a, 1, a.astype("<f4").tobytes())
    elif kind == "P5":
        write_pnm(src, "P5", a, 255, a.astype(np.uint8).tobytes())
    elif kind == "P5w":
        a = a * 200
        write_pnm(src, "P5", a, 65535, a.astype(">u2").tobytes())
    elif kind == "P2":
        write_pnm(src, "P2", a, 255, " ".join(str(int(v)) for v in a.ravel()).encode())
    else:
        write_pnm(src, "Pa", a, 65535, a.astype("<u2").tobytes())
    out = str(tmp_path / "out.q2bz")
    r = msgio("convert", src, out)
    assert r.returncode == 0, r.stderr
    np.testing.assert_array_equal(read_p9(out), a)


def cli(tmp_path, par_text, strict=False):
    exe = os.path.join(BIN, "matchSeries")
    if not os.path.exists(exe):
        pytest.skip("bin/matchSeries not built")
    par = tmp_path / "run.par"
    par.write_text(par_text)
    env = dict(os.environ, MSG_STRICT="1" if strict else "0")
    return subprocess.run([exe, str(par)], capture_output=True, text=True, cwd=tmp_path, env=env)


def test_cli_reports_errors_like_the_reference(tmp_path):
    # unsupported model: message on stderr, exit code 0 (matchSeries.cpp:204-208); strict mode: 1
    r = cli(tmp_path, "deformationModel 3\n")
    assert r.returncode == 0 and "deformationModel 3" in r.stderr
    r = cli(tmp_path, "deformationModel 3\n", strict=True)
    assert r.returncode == 1


@pytest.mark.parametrize("text,msg", [
    ("lambda 1\nlambda 2\n", "already defined"),
    ("lambda\n", "Unexpected end of line"),
    ("name \"unterminated\n", "quotation mark"),
    ("lambda 1 2\n", "Unexpected data after the variable value"),
])
def test_par_grammar_errors(tmp_path, text, msg):
    r = cli(tmp_path, text, strict=True)
    assert r.returncode == 1 and msg in r.stderr


def test_par_grammar_accepts_comments_quotes_lists(tmp_path):
    text = ("# comment line\n\n  deformationModel 5   \n"
            "saveDirectory \"with space/\"\ntemplateSkipNums { 1 9 10 }\nnested { { 1 2 } { 3 4 } }\n")
    r = cli(tmp_path, text, strict=True)
    # parsing succeeds; the run then stops at the unsupported model
    assert r.returncode == 1 and "deformationModel 5" in r.stderr


@pytest.mark.parametrize("name", ["median", "average", "numSamples"])
def test_png_preview_matches_the_reference(tmp_path, name):
    """The reconstruction preview: the reference's own PNG of its stage-1 output (fixture), rebuilt from
    the same array."""
    import sys
    sys.path.insert(0, GOLD)
    from pngread import read_png_gray8
    with np.load(os.path.join(GOLD, "e2e64.npz")) as z:
        arr, png = z[f"s1_{name}"], z[f"s1_{name}_png"]
    src = str(tmp_path / "a.dat")
    with open(src, "wb") as f:
        f.write(f"P9\n{arr.shape[1]} {arr.shape[0]}\n1\n".encode() + np.ascontiguousarray(arr, "<f8").tobytes())
    out = str(tmp_path / "a.png")
    r = msgio("png", src, out)
    assert r.returncode == 0, r.stderr
    np.testing.assert_array_equal(read_png_gray8(out), png)



def test_cli_loads_initial_deformation(tmp_path):
    # "initial-deformation" is read by the registration object and then unused by the series loop
    # (registration.h:495-500, matchSeries.h:111): a missing file is an error, an existing one is not
    with np.load(os.path.join(GOLD, "e2e64.npz")) as z:
        par = bytes(z["par"]).decode()
    r = cli(tmp_path, par + "initial-deformation { nothere_0.dat.bz2 nothere_1.dat.bz2 }\n", strict=True)
    assert r.returncode == 1 and "nothere_0.dat.bz2" in r.stderr


def test_cli_translation_start_needs_max_corr_shift(tmp_path):
    # useCorrelationToInitTranslation reads "maxCorrShift" with getInt, which fails when it is absent
    with np.load(os.path.join(GOLD, "e2e64.npz")) as z:
        par = bytes(z["par"]).decode()
    r = cli(tmp_path, par + "useCorrelationToInitTranslation 1\n", strict=True)
    assert r.returncode == 1 and "maxCorrShift" in r.stderr
