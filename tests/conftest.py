import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        O.build()
    O.set_libm(False)
    return O


@pytest.fixture(scope="session")
def torus_meshes():
    from puzzletunneldiscovery_amd import workloads as W
    return W.torus_scene(0)


@pytest.fixture(scope="session")
def torus_oracle(oracle_mod, torus_meshes):
    env, rob = torus_meshes
    return oracle_mod.OracleWorld().setup(*env, *rob)


@pytest.fixture(scope="session")
def gpu_world(torus_meshes):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from puzzletunneldiscovery_amd import pyosr
    env, rob = torus_meshes
    return pyosr.create_unit_world_from_arrays(env, rob)
