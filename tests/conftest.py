import gzip
import os
import shutil
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "xcalibre_tw.jl_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: full-size case (minutes on CPU)")


@pytest.fixture(scope="session")
def quad40_path(tmp_path_factory):
    """The reference's own sample mesh unv_sample_meshes/quad40.unv (fixture data)."""
    d = tmp_path_factory.mktemp("unv")
    out = d / "quad40.unv"
    with gzip.open(os.path.join(GOLDEN, "quad40.unv.gz"), "rb") as f, open(out, "wb") as g:
        shutil.copyfileobj(f, g)
    return str(out)


def _unzip(tmp_path_factory, name):
    d = tmp_path_factory.mktemp("unv")
    out = d / name
    with gzip.open(os.path.join(GOLDEN, name + ".gz"), "rb") as f, open(out, "wb") as g:
        shutil.copyfileobj(f, g)
    return str(out)


@pytest.fixture(scope="session")
def bfs10_path(tmp_path_factory):
    """unv_sample_meshes/backwardFacingStep_10mm.unv (mesh of CASE_UNV_BFS.jl)."""
    return _unzip(tmp_path_factory, "backwardFacingStep_10mm.unv")


@pytest.fixture(scope="session")
def trig40_path(tmp_path_factory):
    return _unzip(tmp_path_factory, "trig40.unv")


REFERENCE_MESHES = ("quad40", "trig40", "quad100", "trig100", "backwardFacingStep_10mm", "backwardFacingStep_5mm")


@pytest.fixture(scope="session")
def ref_mesh(tmp_path_factory):
    """build_mesh of one of the reference's unv_sample_meshes (fixture data),
    scaled by 0.001 as the reference's case scripts do; cached per name."""
    cache = {}

    def get(name):
        if name not in cache:
            from fvmhip import build_mesh

            cache[name] = build_mesh(_unzip(tmp_path_factory, name + ".unv"), scale=0.001)
        return cache[name]

    return get


@pytest.fixture(scope="session")
def quad40(quad40_path):
    from fvmhip import build_mesh

    return build_mesh(quad40_path, scale=0.001)


def has_gpu():
    try:
        from fvmhip import device_count

        return device_count() > 0
    except Exception:
        return False
