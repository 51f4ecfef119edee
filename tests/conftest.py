import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def my_matrix():
    """The reference's only data fixture: examples/my_matrix.mtx (986-vertex
    unit-weight Laplacian, stored symmetric), as a Graph."""
    import scipy.io
    from ligmg_amd import graphs
    A = scipy.io.mmread(os.path.join(ROOT, "tests", "golden", "my_matrix.mtx")).tocsr()
    return graphs.from_scipy(A)


@pytest.fixture(scope="session")
def golden_dir():
    import pathlib
    return pathlib.Path(ROOT) / "tests" / "golden"
