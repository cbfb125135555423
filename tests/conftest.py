"""Shared fixtures: the HIP library (product) and the CPU oracle (checker).

The oracle is test infrastructure: tests load it only to compare against the
HIP path, never as a fallback for it.
"""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "fivohydro.jl_amd"))
ORACLE_LIB = os.path.join(REPO, "oracle", "liboracle2d.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    # Under pytest-xdist every worker would start one OpenMP thread per core
    # for the oracle (and torch its own pool): share the cores out instead, or
    # 8 workers on 8 cores run 60+ busy threads and the suite takes 10x longer.
    workers = os.environ.get("PYTEST_XDIST_WORKER_COUNT")
    if workers and int(workers) > 1:
        per = max(1, (os.cpu_count() or 1) // int(workers))
        os.environ["OMP_NUM_THREADS"] = str(per)
        os.environ["MKL_NUM_THREADS"] = str(per)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    import fivo2d
    if not os.path.exists(ORACLE_LIB):
        pytest.fail("oracle/liboracle2d.so is not built (run __graft_entry__.build())")
    return fivo2d.Lib(ORACLE_LIB, "orc_")


@pytest.fixture(scope="session")
def hip():
    import fivo2d
    if not gpu_available():
        pytest.skip("no GPU")
    return fivo2d.load_hip()
