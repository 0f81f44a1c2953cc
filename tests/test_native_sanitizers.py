"""Sanitizer builds of the CPU code (no GPU): the FP64 oracle, the host code
that parses input files (.lbmc reader, LBMCFLD1 field files, the CGNS/HDF5
reader and writer) and the kernels' node index arithmetic at the node limit,
each compiled with AddressSanitizer + UndefinedBehaviorSanitizer
(-fno-sanitize-recover=all: the first finding aborts the run) and driven by
a program under tests/native/ over valid and malformed inputs."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
HOST = os.path.join(ROOT, "lbm-c_amd", "host")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
HDF5 = "/opt/conda"
HAVE_HDF5 = os.path.exists(os.path.join(HDF5, "include", "hdf5.h")) and os.path.exists(
    os.path.join(HDF5, "lib", "libhdf5.so"))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def run(cmd, cwd, **kw):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=300, **kw)
    return r


def build(cmd, cwd):
    r = run(cmd, cwd)
    assert r.returncode == 0, r.stderr[-3000:]


def hdf5_dir(tmp_path):
    """libhdf5 through a directory that holds only it: the conda library
    directory also holds older sanitizer runtimes, which a plain -L / rpath
    there would pull in beside the compiler's own."""
    d = tmp_path / "h5"
    d.mkdir()
    lib = os.path.join(HDF5, "lib")
    for name in os.listdir(lib):
        if name.startswith("libhdf5.so"):
            os.symlink(os.path.realpath(os.path.join(lib, name)), d / name)
    return str(d)


def test_index_limits(tmp_path):
    exe = str(tmp_path / "index_limits")
    build(["g++", "-std=c++17", *SAN, "-o", exe, os.path.join(NATIVE, "index_limits.cpp")], tmp_path)
    r = run([exe], tmp_path, env=ENV)
    assert r.returncode == 0 and "index limits: ok" in r.stdout, r.stdout + r.stderr[-3000:]


def test_index_limits_check_bites(tmp_path):
    """The closure loop's former int arithmetic (n0 = b * 256 + lane)
    overflows near 2^31 nodes; the same check must stop on it."""
    exe = str(tmp_path / "index_limits_signed")
    build(["g++", "-std=c++17", *SAN, "-DSIGNED_TILES", "-o", exe, os.path.join(NATIVE, "index_limits.cpp")],
          tmp_path)
    r = run([exe], tmp_path, env=ENV)
    assert r.returncode != 0 and "signed integer overflow" in r.stderr, r.stdout + r.stderr[-2000:]


def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_sanitize")
    build(["gcc", "-std=c11", *SAN, "-ffp-contract=off", "-fopenmp", "-o", exe,
           os.path.join(NATIVE, "oracle_sanitize.c"), os.path.join(ROOT, "oracle", "lbmc_oracle.c"), "-lm"],
          tmp_path)
    r = run([exe], tmp_path, env=dict(ENV, OMP_NUM_THREADS="2"))
    assert r.returncode == 0 and "oracle sanitize: ok" in r.stdout, r.stdout + r.stderr[-3000:]


@pytest.mark.skipif(not HAVE_HDF5, reason="no HDF5 headers/library")
def test_cgns_io_under_sanitizers(tmp_path):
    h5 = hdf5_dir(tmp_path)
    exe = str(tmp_path / "cgns_sanitize")
    build(["gcc", "-std=c99", *SAN, f"-I{HDF5}/include", "-o", exe, os.path.join(NATIVE, "cgns_sanitize.c"),
           os.path.join(HOST, "cgns_io.c"), os.path.join(h5, "libhdf5.so"), f"-Wl,-rpath,{h5}"], tmp_path)
    r = run([exe, str(tmp_path)], tmp_path, env=ENV)
    assert r.returncode == 0 and "cgns sanitize: ok" in r.stdout, r.stdout + r.stderr[-3000:]


@pytest.mark.skipif(not HAVE_HDF5, reason="no HDF5 headers/library")
def test_host_readers_under_sanitizers(tmp_path):
    h5 = hdf5_dir(tmp_path)
    # the driver loads libcgnsio.so from its own directory: a sanitized copy
    build(["gcc", "-std=c99", "-fPIC", "-shared", *SAN, f"-I{HDF5}/include", "-o", "libcgnsio.so",
           os.path.join(HOST, "cgns_io.c"), os.path.join(h5, "libhdf5.so"), f"-Wl,-rpath,{h5}"], tmp_path)
    exe = str(tmp_path / "host_sanitize")
    build(["g++", "-std=c++17", *SAN, "-o", exe, os.path.join(NATIVE, "host_sanitize.cpp"),
           os.path.join(HOST, "infile_reader.cpp"), os.path.join(HOST, "field_file.cpp"), "-ldl"], tmp_path)
    r = run([exe, str(tmp_path)], tmp_path, env=ENV)
    assert r.returncode == 0 and "host sanitize: ok" in r.stdout, r.stdout + r.stderr[-3000:]


def test_sanitizer_toolchain_present():
    assert shutil.which("gcc") and shutil.which("g++")
