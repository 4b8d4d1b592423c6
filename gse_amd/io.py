"""Text formats of the reference's driver program (cpu_gse.cpp), in Python.

  load_array(width, num_line, filename)   cpu_gse.cpp:152-171 loadArray
  load_xyzq(filename, n)                  loadArray with width 4 (x y z q)
  save_force(filename, fx, fy, fz)        cpu_gse.cpp:173-183 saveForce
  read_force(filename, n)                 cpu_gse.cpp:185-195 readForce
  erms(ref, x)                            cpu_gse.cpp:131-146 Erms

save_force writes what C++'s default stream formatting writes (six
significant digits, "%g"), so files match those of include/gse/io.hpp and of
the reference program.
"""
import numpy as np


def load_array(width, num_line, filename, dtype=np.float64):
    """num_line * width whitespace-separated values; raises ValueError when
    the file holds fewer (the reference prints and exits)."""
    with open(filename) as fh:
        vals = fh.read().split()
    need = width * num_line
    if len(vals) < need:
        raise ValueError(f"error reading {filename}: {len(vals)} values, {need} needed")
    return np.array(vals[:need], dtype=dtype).reshape(num_line, width)


def load_xyzq(filename, n):
    return load_array(4, n, filename)


def _g(v):
    s = "%g" % v
    return s


def save_force(filename, fx, fy, fz):
    with open(filename, "w") as fh:
        for a, b, c in zip(fx, fy, fz):
            fh.write(f"{_g(a)} {_g(b)} {_g(c)}\n")


def read_force(filename, n):
    a = load_array(3, n, filename)
    return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy()


def erms(ref, x):
    ref = np.asarray(ref, np.float64)
    x = np.asarray(x, np.float64)
    return float(np.sqrt(np.sum((x - ref) ** 2) / np.sum(ref ** 2)))
