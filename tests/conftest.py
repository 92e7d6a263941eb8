import gzip
import importlib.util
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X, gfx950)")


def _load_pyvc():
    spec = importlib.util.spec_from_file_location("pyvc", os.path.join(ROOT, "volume-compressor_amd", "pyvc.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.fixture(scope="session")
def pyvc():
    return _load_pyvc()


@pytest.fixture(scope="session")
def oracle():
    from pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    from pyoracle import Ref, ref_variants
    if "strict" not in ref_variants():
        pytest.skip("reference build oracle/_ref not present (make -f oracle/ref.mk)")
    return Ref("strict")


@pytest.fixture(scope="session")
def golden_atoms():
    return dict(np.load(os.path.join(GOLDEN, "atoms.npz")))


@pytest.fixture(scope="session")
def golden_corrupt():
    return dict(np.load(os.path.join(GOLDEN, "corrupt.npz")))


@pytest.fixture(scope="session")
def golden_containers():
    meta = json.load(open(os.path.join(GOLDEN, "containers.json")))
    out = {}
    for name, m in meta.items():
        b = {}
        for tag in ("v1", "v2"):
            with gzip.open(os.path.join(GOLDEN, f"{name}.{tag}.vca.gz"), "rb") as f:
                b[tag] = f.read()
        out[name] = (m, b)
    return out


@pytest.fixture(scope="session")
def golden_repack():
    """container-rewrite cases: {name: (meta, {"v1": bytes, "v2": bytes})},
    v2 from tools/vc_repack.c on the reference writer's v1"""
    meta = json.load(open(os.path.join(GOLDEN, "repack.json")))
    out = {}
    for name, m in meta.items():
        b = {}
        for tag in ("v1", "v2"):
            with gzip.open(os.path.join(GOLDEN, f"{name}.{tag}.vca.gz"), "rb") as f:
                b[tag] = f.read()
        out[name] = (m, b)
    return out


def payloads_of(g, qi=None):
    """[(atom index, q index, payload bytes, recon sha)] from atoms.npz"""
    blob = g["blob"].tobytes()
    res = []
    for k in range(len(g["lens"])):
        if qi is not None and g["qidx"][k] != qi:
            continue
        o, n = int(g["offs"][k]), int(g["lens"][k])
        res.append((int(g["aidx"][k]), int(g["qidx"][k]), blob[o:o + n], str(g["rec_sha"][k])))
    return res


def container_volume(oracle, m):
    nz, ny, nx = m["shape_zyx"]
    vol = oracle.synth(m["seed"], nx, ny, nz, m["sheets"])
    if m["zero_slab"]:
        vol[:32, :32, :] = 0
    return vol
