"""Build every native artefact in-tree: the gfx950 HIP library, the matchSeries CLI, the oracle.

    python match-series_amd/msamd/build.py
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))   # match-series_amd/
ROOT = os.path.dirname(PKG)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = ["-std=c++17", "-O3", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off",
             # packed FP32 instructions stay off (hipcc would otherwise form them from scalar code)
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def build_lib():
    os.makedirs(os.path.join(PKG, "lib"), exist_ok=True)
    out = os.path.join(PKG, "lib", "libmsgpu.so")
    tmp = out + ".tmp"
    run([HIPCC] + HIP_FLAGS + ["-shared", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "csrc"),
                               os.path.join(PKG, "csrc", "ms_gpu.hip"), os.path.join(PKG, "csrc", "ms_unet.hip"),
                               "-o", tmp])
    os.replace(tmp, out)
    return out


def build_cli():
    """bin/matchSeries (the drop-in CLI) and bin/msgio (array file utility), sharing params/array_io."""
    host = os.path.join(PKG, "host")
    common = [os.path.join(host, f) for f in ("params.cpp", "array_io.cpp", "image_formats.cpp")]
    os.makedirs(os.path.join(PKG, "bin"), exist_ok=True)
    libdir = os.path.join(PKG, "lib")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", host, "-I", "/opt/rocm/include",
           "-idirafter", "/opt/conda/include",  # bzlib.h (the library itself is the system libbz2)
           "-D__HIP_PLATFORM_AMD__"]
    bz2 = "/usr/lib/x86_64-linux-gnu/libbz2.so.1.0"
    zlib = "/usr/lib/x86_64-linux-gnu/libz.so.1"
    outs = []
    out = os.path.join(PKG, "bin", "matchSeries")
    run(["g++", "-std=c++17", "-O2", "-Wall"] + inc + common + [os.path.join(host, "matchSeries.cpp"), "-o", out,
         "-L", libdir, "-lmsgpu", "-Wl,-rpath,$ORIGIN/../lib", "-L/opt/rocm/lib", "-lamdhip64",
         "-Wl,-rpath,/opt/rocm/lib", bz2, zlib])
    outs.append(out)
    # matchSeismicSeries: the same driver with the row-wise x-shift registration (dyadic path)
    out = os.path.join(PKG, "bin", "matchSeismicSeries")
    run(["g++", "-std=c++17", "-O2", "-Wall", "-DMS_SEISMIC"] + inc + common +
        [os.path.join(host, "matchSeries.cpp"), "-o", out, "-L", libdir, "-lmsgpu", "-Wl,-rpath,$ORIGIN/../lib",
         "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", bz2, zlib])
    outs.append(out)
    out = os.path.join(PKG, "bin", "msgio")
    run(["g++", "-std=c++17", "-O2", "-Wall"] + inc + common + [os.path.join(host, "msgio.cpp"), "-o", out, bz2, zlib])
    outs.append(out)
    return outs


def build_oracle(with_reference=True):
    run(["make", "-C", os.path.join(ROOT, "oracle"), "-B", "lib"])
    if with_reference and os.path.isdir("/root/reference/quocmesh"):
        run(["make", "-C", os.path.join(ROOT, "oracle"), "-j8", "ref"])
        # the reference's driver over the GPU registration object (build check of the adapter)
        run(["make", "-C", os.path.join(ROOT, "oracle"), "-B", "adapter"])


def build_all(with_reference=True):
    build_lib()
    build_cli()
    build_oracle(with_reference)


if __name__ == "__main__":
    build_all(with_reference="--no-ref" not in sys.argv)
