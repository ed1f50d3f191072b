import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("psieve-cuda_amd/python", "oracle"):
    path = os.path.join(REPO, sub)
    if path not in sys.path:
        sys.path.insert(0, path)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


def read_lines(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return [l.rstrip("\n") for l in f if l.strip()]


def windowed_run(tag):
    """(params, windows) of a windowed golden run: the search's parameters
    and the p-windows [a, b) whose primes the oracle walked completely."""
    import json

    with open(os.path.join(GOLDEN, "runs.json")) as f:
        r = json.load(f)[tag]
    return dict(r["params"]), [tuple(w) for w in r["windows"]]
