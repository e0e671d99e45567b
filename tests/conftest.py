import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def cylinder_fixture():
    """owner/neighbour/boundary of examples/cylinder (see golden/make_cylinder_fixture.py)."""
    z = np.load(GOLDEN / "cylinder_faces.npz")
    owner = np.cumsum(z["owner_delta"]).astype(np.int32)
    neigh = np.cumsum(z["neigh_delta"]).astype(np.int32)
    patches = list(zip([str(s) for s in z["patch_names"]], z["patch_start"].tolist(),
                       z["patch_nfaces"].tolist()))
    return dict(owner=owner, neigh=neigh, patches=patches, nCells=int(z["nCells"]),
                nnz=int(z["nnz"]), csr_crc=int(z["csr_crc"]), row_len_hist=z["row_len_hist"])


@pytest.fixture(scope="session")
def gpu():
    from anabasis_amd import _lib
    if not _lib.gpu_available():
        pytest.fail("GPU test selected but no GPU is visible")
    _lib.check(_lib.lib().an_set_device(0))
    return True
