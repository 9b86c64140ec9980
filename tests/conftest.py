import os
import sys
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    import nsffi
    if not os.path.exists(nsffi.ORACLE_SO):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"], check=True)
    return nsffi.Oracle()


@pytest.fixture(scope="session")
def ref():
    import nsffi
    if not nsffi.have_ref():
        pytest.skip("compiled reference (oracle/_ref) not present")
    return nsffi.Reference()


@pytest.fixture(scope="session")
def product():
    """the product library on a GPU; fails loudly when the HIP library is missing"""
    import nsffi
    if not os.path.exists(nsffi.PRODUCT_SO):
        raise RuntimeError("libnasw_hip.so is not built (run __graft_entry__.build())")
    p = nsffi.Product()
    if not p.lib.ns_gpu_init(0):
        raise RuntimeError("GPU not usable: " + p.error())
    return p
