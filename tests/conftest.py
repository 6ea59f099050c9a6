import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = ["box2_n5", "cube2_n8", "periodic3_n6", "channel_n6", "cube2_n12"]


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X GPU (select with -m gpu)")


def golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"))


def golden_mesh(Z):
    from ssenss_amd.mesh import Mesh
    n = int(Z["n"])
    nel = Z["neigh"].shape[0]
    sh = (nel, n, n, n)
    return Mesh(n, Z["x"].reshape(sh), Z["y"].reshape(sh), Z["z"].reshape(sh),
                Z["neigh"], Z["edge_neigh"], Z["edge_size"], Z["corn_neigh"],
                Z["corn_size"], dt=float(Z["dt"]), visc=float(Z["visc"]))


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb > 0 else 1)


@pytest.fixture(params=CASES)
def case(request):
    Z = golden(request.param)
    return request.param, Z, golden_mesh(Z)
