import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "asd-kernel-fusion_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: longer CPU oracle checks")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ref_fusion():
    import oracle
    if not oracle.RefFusion.available("parity"):
        pytest.skip("reference kernel_fusion build (oracle/_ref) not present")
    return oracle.RefFusion("parity")


@pytest.fixture(scope="session")
def gctx():
    """One libasd context on cuda:0 for the GPU tests."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import asd_amd
    c = asd_amd.Context(0)
    yield c
    c.close()
