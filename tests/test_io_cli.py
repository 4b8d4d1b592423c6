"""The reference driver's text formats and command line (CPU side).

  * include/gse/io.hpp (C++) and gse_amd/io.py write identical saveForce
    files, in C++'s default stream format ("%g", six significant digits),
    and each reads the other's; loadArray rejects a short file;
  * Erms is the relative RMS norm of cpu_gse.cpp:131-146;
  * gse_cli checks its options as cpu_gse.cpp:100-103 does: the usage line
    and exit status 1, before any device work.
"""
import os
import subprocess

import numpy as np
import pytest

from gse_amd import io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gse_amd", "gse_cli")

PROG = r"""
#include "gse/io.hpp"
#include <cstdio>
using namespace gse_amd;
int main(int argc, char** argv) {
  const int n = 5;
  double fx[n] = {1.0, -2.5e-7, 123456789.0, 0.1, -0.0};
  double fy[n] = {3.14159265358979, 1e300, -1e-300, 2.0/3.0, 42.0};
  double fz[n] = {0.5, 0.25, -7.0, 1e-5, 99999.95};
  saveForce<double>(argv[1], n, fx, fy, fz);
  double rx[n], ry[n], rz[n];
  readForce<double>(argv[2], n, rx, ry, rz);
  for (int i = 0; i < n; i++) std::printf("%.17g %.17g %.17g\n", rx[i], ry[i], rz[i]);
  double a[8];
  try { loadArray<double>(4, 2, argv[3], a); std::printf("loaded\n"); }
  catch (const Error& e) { std::printf("short file rejected\n"); }
  return 0;
}
"""


@pytest.fixture(scope="module")
def ioprog(tmp_path_factory):
    d = tmp_path_factory.mktemp("io")
    src = d / "io.cpp"
    src.write_text(PROG)
    exe = d / "io"
    lib = os.path.join(ROOT, "gse_amd")
    if not os.path.exists(os.path.join(lib, "libgse_hip.so")):
        pytest.skip("libgse_hip.so not built")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", str(exe), "-L", lib, "-lgse_hip", f"-Wl,-rpath,{lib}",
                    "-Wl,-rpath,/opt/rocm/lib"], check=True)
    return exe, d


def test_saveforce_format_matches_cpp_stream_default(ioprog):
    exe, d = ioprog
    fx = [1.0, -2.5e-7, 123456789.0, 0.1, -0.0]
    fy = [3.14159265358979, 1e300, -1e-300, 2.0 / 3.0, 42.0]
    fz = [0.5, 0.25, -7.0, 1e-5, 99999.95]
    io.save_force(str(d / "py.txt"), fx, fy, fz)
    (d / "short.txt").write_text("1 2 3 4\n5 6 7\n")
    out = subprocess.run([str(exe), str(d / "cpp.txt"), str(d / "py.txt"), str(d / "short.txt")],
                         capture_output=True, text=True, check=True).stdout.split("\n")
    assert (d / "cpp.txt").read_text() == (d / "py.txt").read_text()
    # C++ read back what Python wrote: six significant digits of each value
    back = np.array([[float(v) for v in ln.split()] for ln in out[:5]])
    ref = np.array([fx, fy, fz]).T
    assert np.allclose(back, ref, rtol=5e-6, atol=0)
    assert out[5] == "short file rejected"
    rx, ry, rz = io.read_force(str(d / "cpp.txt"), 5)
    assert np.allclose(np.array([rx, ry, rz]).T, ref, rtol=5e-6, atol=0)
    with pytest.raises(ValueError):
        io.load_array(4, 2, str(d / "short.txt"))


def test_erms_definition():
    ref = np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0]])
    x = ref + np.array([[0.1, 0.0, 0.0], [0.0, 0.0, 0.2]])
    assert abs(io.erms(ref, x) - np.sqrt(0.05 / 5.0)) < 1e-15


@pytest.mark.parametrize("args", [["-order", "5"], ["-bogus"], ["-L", "-3"], ["-r"],
                                  ["-test", "file"], ["-M", "0", "-L", "0.5"]])
def test_cli_rejects_bad_options(args):
    if not os.path.exists(CLI):
        pytest.skip("gse_cli not built")
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "Usage: cpu_gse -r r -L L -M M -sigma sigma -order order -seed seed" in r.stdout
