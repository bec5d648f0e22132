import importlib.util
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    m = importlib.util.module_from_spec(spec)
    sys.modules[name] = m
    spec.loader.exec_module(m)
    return m


@pytest.fixture(scope="session")
def sb():
    m = _load("sbamd", os.path.join(ROOT, "ipdps-crs-cmp-code_amd", "__init__.py"))
    m.lib()  # raises if the library is not built
    return m


@pytest.fixture(scope="session")
def orc():
    return _load("sb_oracle", os.path.join(ROOT, "oracle", "oracle.py"))


@pytest.fixture(scope="session")
def manifest():
    with open(os.path.join(GOLD, "manifest.json")) as f:
        return json.load(f)


def args_dict(args):
    return {args[i]: args[i + 1] for i in range(0, len(args), 2)}


def cmp_args(case):
    a = args_dict(case["args"])
    return dict(c0=float(np.float32(a["-c0"])), c1=float(np.float32(a["-c1"])), nc=int(a["-nc"]),
                tau=float(np.float32(a["-tau"])), aph=int(a["-aph"]))


def crs_args(case):
    a = args_dict(case["args"])
    return dict(a0=float(a["-a0"]), a1=float(a["-a1"]), na=int(a["-na"]), b0=float(a["-b0"]), b1=float(a["-b1"]),
                nb=int(a["-nb"]), c0=float(a["-c0"]), c1=float(a["-c1"]), nc=int(a["-nc"]), tau=float(a["-tau"]),
                aph=int(a["-aph"]), apm=int(a["-apm"]))


@pytest.fixture(scope="session")
def gpu_ctx(sb):
    ctx = sb.Context(0)
    yield ctx
    ctx.close()


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
