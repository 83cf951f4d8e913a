import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs through the C-ABI)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs (set HM_SLOW=1)")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("HM_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="long oracle run; set HM_SLOW=1")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


def load_fixture(name):
    import json

    import numpy as np
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    return meta, z


def workload_from_meta(meta):
    from hypar_amd import decks
    return decks.Workload(meta["name"], tuple(meta["dims"]), interp=meta["interp"],
                          upwind=meta["upwind"], Re=meta["Re"], Minf=meta["Minf"], Pr=meta["Pr"],
                          gamma=meta["gamma"], weno_eps=meta["eps"], n_iter=meta["steps"],
                          seed=meta["seed"], field=meta["field"], dt=meta["dt"],
                          time_scheme_type=meta.get("time_scheme_type", "44"),
                          weno_type=meta.get("weno_type", "js"))


def within_print(got, printed, digits):
    """|got - printed| within half a unit of the printed value's last digit
    (plus rounding-level slack: the GPU state differs from the reference's
    at the 1e-14 level, which may move a value across a rounding boundary)."""
    import numpy as np
    ref = float(printed)
    e = np.floor(np.log10(abs(ref)))
    return abs(got - ref) <= 0.5 * 10.0 ** (e - digits) + 1e-12 * abs(ref)
