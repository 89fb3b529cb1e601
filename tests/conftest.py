import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_package():
    """The product package (its directory name is not a Python identifier)."""
    if "vgi_amd" in sys.modules:
        return sys.modules["vgi_amd"]
    path = os.path.join(ROOT, "global-illumination_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location("vgi_amd", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["vgi_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def vgi():
    return load_package()


@pytest.fixture(scope="session")
def orc():
    import oracle_lib

    return oracle_lib


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    # A GPU test on a box without a GPU is an error, not a skip: the driver
    # runs -m gpu only where a device exists.
    pass


def make_scene(vgi, orc, workload, seed, edge=0):
    """Grid, optics and faces of a workload, plus its oracle scene/tracer."""
    grid = vgi.generate(workload, seed, edge)
    diffuse, emission, lights, env = vgi.optics(workload, seed)
    coords, pages = vgi.pages_of(grid)
    scene = orc.Scene()
    scene.load_pages(coords, pages)
    tracer = orc.Tracer(scene)
    tracer.set_materials(diffuse, emission)
    tracer.set_lights(lights)
    return dict(grid=grid, diffuse=diffuse, emission=emission, lights=lights, env=env, coords=coords, pages=pages,
                scene=scene, tracer=tracer)


def gpu_tracer(vgi, setup, resident=False):
    g = vgi.GpuTracer(0)
    if resident:
        g.enable_resident()
    g.upload_pages(setup["coords"], setup["pages"])
    g.set_materials(setup["diffuse"], setup["emission"])
    g.set_lights(setup["lights"])
    return g


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
