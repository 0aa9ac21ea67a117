import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "gmap-gsnap_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "ref: needs the reference build under oracle/_ref or oracle/_ref_std")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    if not os.path.exists(pyoracle.ORACLE_SO):
        pytest.fail("oracle/liboracle.so missing: run __graft_entry__.build()")
    return pyoracle.Oracle()


@pytest.fixture(scope="session")
def reference():
    import pyoracle
    if not os.path.exists(pyoracle.REF_SO):
        pytest.skip("reference build oracle/_ref/libgmapref.so not present on this host")
    return pyoracle.Reference()


@pytest.fixture(scope="session")
def reference_std():
    import pyoracle
    if not os.path.exists(pyoracle.REF_STD_SO):
        pytest.skip("reference build oracle/_ref_std/libgmapref_std.so not present on this host")
    return pyoracle.ReferenceStd()
