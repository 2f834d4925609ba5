import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from sp3mm4amg_amd.csr import CSR  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: full-size workloads")


def unpack(z, prefix) -> CSR:
    M, N = (int(x) for x in z[prefix + "_shape"])
    return CSR(M, N, z[prefix + "_irp"], z[prefix + "_ja"].astype(np.int64), z[prefix + "_vals"])


@pytest.fixture(scope="session")
def fixture_2744():
    """The reference's known-answer group code/test/{r,ac,p,ac_next}."""
    z = np.load(os.path.join(GOLDEN, "fixture_2744.npz"))
    return {k: unpack(z, k) for k in ("r", "ac", "p", "ac_next")}


@pytest.fixture(scope="session")
def ref_outputs():
    out = {}
    for case in ("fixture", "poisson2d64"):
        z = np.load(os.path.join(GOLDEN, f"ref_outputs_{case}.npz"))
        d = {k: unpack(z, k) for k in ("pair_symbnum", "pair_2d", "merged", "pair_symbnum_1",
                                        "merged_1")}
        d["ub_ra"] = z["ub_ra"]
        d["symb_ra"] = z["symb_ra"]
        out[case] = d
    return out


@pytest.fixture(scope="session")
def poisson2d64():
    from sp3mm4amg_amd import problems
    p = problems.poisson2d_sa(64)
    return p["R"], p["A"], p["P"]
