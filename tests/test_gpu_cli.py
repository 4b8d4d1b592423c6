"""gse_cli, the reference driver's front-end over this library, run end to
end on the GPU and checked against the FP64 oracle at the precision each
printed line or file carries.

  * -test pair (cpu_gse.cpp:638-893): energy_GSEk / energy_SPME / energy_ew
    (six significant digits), the GSEk / SPME / EWALD force lines (twelve
    decimals), phi.txt and ExGSEk.txt;
  * -test random (cpu_gse.cpp:423-633): the particles come from
    srand(seed)/rand() as in the reference, reproduced here through the
    same C library; forceGSEk.txt, forceSPME.txt and forceEW.txt;
  * -test file (cpu_gse.cpp:216-418 on a file written here): with the
    reference's exclusion files, forceEWdir.txt against the oracle's
    CpuEwaldDirect restatement.
"""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

from gse_amd import io, systems
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gse_amd", "gse_cli")


def run(args, cwd):
    if not os.path.exists(CLI):
        pytest.fail(f"{CLI} is missing: run __graft_entry__.build()")
    print("running gse_cli", " ".join(args), flush=True)
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120, cwd=cwd)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def value(out, key):
    for ln in out.splitlines():
        if ln.startswith(key + " "):
            return float(ln.split()[-1])
    raise AssertionError(f"{key} not printed")


def forces(out, tag):
    for ln in out.splitlines():
        if ln.startswith(tag + ":"):
            a, b = ln.split(":", 1)[1].split("|")
            return np.array([[float(v) for v in a.split()], [float(v) for v in b.split()]])
    raise AssertionError(f"{tag} line not printed")


def test_cli_pair(tmp_path):
    r, L, M, s = 1.0, 20.0, 20, 2.121320344
    out = run(["-r", str(r), "-L", str(L), "-M", str(M), "-sigma", str(s)], str(tmp_path))
    x = systems.pair(r, L, orientation="diagonal")
    p = orc.Params(M, L, s)
    g = orc.gse_step(p, x, want_grids=True)
    kappa = 1.0 / (s * math.sqrt(2.0))
    e_pme, _, f_pme = orc.pme_step((M, M, M), 4, x, (L,) * 3, kappa)
    e_ew, f_ew = orc.ewald_recip(s, int(32 * M / L), (L,) * 3, x)
    for key, ref in (("energy_GSEk", g.energy), ("energy_SPME", e_pme), ("energy_ew", e_ew)):
        assert abs(value(out, key) - ref) <= 5e-6 * abs(ref), key
    for tag, ref in (("GSEk", g.force), ("SPME", f_pme), ("EWALD", f_ew)):
        got = forces(out, tag)
        assert np.max(np.abs(got - ref)) <= 1e-11 + 1e-10 * np.max(np.abs(ref)), tag
    phi = np.loadtxt(tmp_path / "phi.txt")
    assert np.max(np.abs(phi - g.phi.ravel())) <= 1e-5 * np.max(np.abs(g.phi))
    ex, _, _ = orc.efield((M, M, M), (L,) * 3, g.phi)
    got = np.loadtxt(tmp_path / "ExGSEk.txt")
    assert np.max(np.abs(got - ex.ravel())) <= 1e-5 * np.max(np.abs(ex))


def test_cli_random(tmp_path):
    n, L, M, s, order, seed = 60, 16.0, 16, 1.5, 6, 7
    out = run(["-test", "random", "-n", str(n), "-L", str(L), "-M", str(M), "-sigma", str(s),
               "-order", str(order), "-seed", str(seed)], str(tmp_path))
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(seed))
    rmax = 2147483647.0  # RAND_MAX of glibc
    x = np.empty((n, 4))
    for i in range(n):
        for c in range(3):
            x[i, c] = L * (libc.rand() / (rmax + 1.0))
        x[i, 3] = (i % 2) * 2 - 1
    p = orc.Params(M, L, s)
    g = orc.gse_step(p, x)
    _, _, f_pme = orc.pme_step((M, M, M), order, x, (L,) * 3, 1.0 / (s * math.sqrt(2.0)))
    e_ew, f_ew = orc.ewald_recip(s, 12, (L,) * 3, x)
    assert abs(value(out, "energy_GSEk =") - g.energy) <= 5e-6 * abs(g.energy)
    assert abs(value(out, "energy_ew =") - e_ew) <= 5e-6 * abs(e_ew)
    for fname, ref in (("forceGSEk.txt", g.force), ("forceSPME.txt", f_pme),
                       ("forceEW.txt", f_ew)):
        got = np.array(io.read_force(str(tmp_path / fname), n)).T
        assert io.erms(ref, got) <= 1e-5, fname
    assert "rms(GSEk) = " in out and "rms(SPME) = " in out


def test_cli_file_with_exclusions(tmp_path):
    x = systems.water_box(300, 22.0, seed=5)
    n = len(x)
    np.savetxt(tmp_path / "xyzq.txt", x, fmt="%.17g")
    ib, nb = [], []
    for j in range(n):
        t = j // 3
        for i in range(3 * t, 3 * t + 3):
            if i > j:
                nb.append(i + 1)
        ib.append(len(nb))
    np.savetxt(tmp_path / "iblo14.txt", np.array(ib), fmt="%d")
    np.savetxt(tmp_path / "inb14.txt", np.array(nb), fmt="%d")
    out = run(["-test", "file", "-xyzq", "xyzq.txt", "-n", str(n), "-L", "22", "-M", "24",
               "-sigma", "2.12", "-iblo14", "iblo14.txt", "-inb14", "inb14.txt", "-kcut", "10"],
              str(tmp_path))
    _, g = orc.ewald_direct(2.12, 9.0, (22.0,) * 3, x, ib, nb)
    got = np.array(io.read_force(str(tmp_path / "forceEWdir.txt"), n)).T
    assert io.erms(g, got) <= 1e-5
    assert out.count("rms(GSEk) = ") == 2
