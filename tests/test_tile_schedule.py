"""The tile walk's schedule (csrc/tile_sched.h), on the CPU: tests/tile_check.cpp
builds it for the matrix pattern of a mesh and replays the kernel's waves
sequentially (staging of neighbours from other tiles, steps in order, every
neighbour read from its LDS word), which must find each word written before
use and give the plain row-by-row DILU update and sweeps bit for bit.  The
pattern is Equation(mesh)'s: each cell and its face neighbours, ascending
(Discretise_1_schemes.jl builds one entry per owner/neighbour pair)."""
import os
import subprocess

import numpy as np
import pytest

import fvmhip as F

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "xcalibre_tw.jl_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("tile") / "tile_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC,
                    os.path.join(HERE, "tile_check.cpp"), "-o", exe], check=True)
    return exe


def write_pattern(mesh, path):
    fo = mesh.face_owners
    a, b = fo[:, 0], fo[:, 1]
    keep = a != b
    n = mesh.ncells
    r = np.concatenate([a[keep], b[keep], np.arange(n)])
    c = np.concatenate([b[keep], a[keep], np.arange(n)])
    o = np.lexsort((c, r))
    r, c = r[o], c[o]
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, r + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    with open(path, "wb") as f:
        np.array([n, c.size], np.int32).tofile(f)
        rp.tofile(f)
        c.astype(np.int32).tofile(f)


@pytest.mark.parametrize("tr,lag", [(64, 1), (256, 12), (1024, 3), (4096, 12)])
def test_tile_schedule_delaunay(checker, tmp_path, tr, lag):
    mesh = F.delaunay_unit_square(20_000, seed=3)
    path = str(tmp_path / "pattern.bin")
    write_pattern(mesh, path)
    out = subprocess.run([checker, path, str(tr), str(lag)], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr


def test_tile_schedule_quads_and_one_tile(checker, tmp_path):
    """A structured grid (wide levels: steps split at 64 rows, chunks at 64
    rows or 64 staged neighbours) and a mesh smaller than one tile."""
    path = str(tmp_path / "pattern.bin")
    for mesh, tr in ((F.cavity(40), 256), (F.cavity(40), 64), (F.cavity(12), 4096)):
        write_pattern(mesh, path)
        out = subprocess.run([checker, path, str(tr), "12"], capture_output=True, text=True)
        print(out.stdout)
        assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr
