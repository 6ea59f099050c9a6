#!/usr/bin/env python3
"""Writes tests/golden/example_init.npz: the velocity records of the
reference's examples/init.h5, dumped with h5dump (binary, memory layout),
the fixture the velocity reader is checked against.  Runs only where the
reference's examples and h5dump are present:
  python tests/golden/make_io_golden.py
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EXAMPLE_INIT = "/root/reference/examples/init.h5"
H5DUMP = "/opt/conda/bin/h5dump"


def main():
    with tempfile.TemporaryDirectory() as d:
        par = os.path.join(d, "par.bin")
        subprocess.run([H5DUMP, "-b", "MEMORY", "-d", "/Parameters", "-o",
                        par, EXAMPLE_INIT], check=True,
                       stdout=subprocess.DEVNULL)
        n, nel = np.fromfile(par, dtype="<i4")[:2]
        raw = os.path.join(d, "el.bin")
        subprocess.run([H5DUMP, "-b", "MEMORY", "-d", "/Element array", "-o",
                        raw, EXAMPLE_INIT], check=True,
                       stdout=subprocess.DEVNULL)
        n3 = int(n) ** 3
        rec = np.dtype([("U", "<f8", (n3,)), ("V", "<f8", (n3,)),
                        ("W", "<f8", (n3,))])
        a = np.fromfile(raw, dtype=rec)
    assert len(a) == nel
    np.savez_compressed(os.path.join(HERE, "example_init.npz"), n=np.int32(n),
                        u=a["U"], v=a["V"], w=a["W"])


if __name__ == "__main__":
    main()
