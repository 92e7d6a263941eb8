"""GPU: the exporter's within-cell 2x downscale (vcx_downscale2x_device)
against the reference's tools/downscale.h (golden SHA-256 from its build) and
the oracle's restatement."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_ds2x_matches_reference(pyvc):
    sys.path.insert(0, GOLDEN)
    from make_golden import ds_input
    for c in json.load(open(os.path.join(GOLDEN, "ds.json"))):
        v = ds_input(tuple(c["shape_zyx"]), c["kind"], c["seed"])
        st, out = pyvc.downscale2x(v, c["method"], c["alpha"])
        assert st == 0 and sha(out) == c["sha"], c


def test_ds2x_large_matches_oracle(pyvc, oracle):
    """odd dims (one voxel per thread, edge replication) and even dims with x
    a multiple of 8 (four voxels per thread from 8-byte row loads)"""
    for dims in ((514, 300, 131), (256, 96, 64)):
        vol = pyvc.synth(9, dims, 12)
        for m, a in ((0, 0.0), (1, 0.5), (1, 1.0)):
            st, out = pyvc.downscale2x(vol, m, a)
            assert st == 0 and np.array_equal(out, oracle.ds2x(vol, m, a)), (dims, m, a)


def test_wide_index_paths_vs_oracle(pyvc, oracle):
    """the 64-bit output-index bodies of the 2x kernels (taken on their own
    only past ~4.27e9 output voxels), forced on small volumes: the downscale
    in both thread shapes, the reference's test downsample and a whole encode
    (its pyramid) against the oracle"""
    pyvc.lib().vcx_test_wide_index(1)
    try:
        for dims in ((514, 300, 131), (256, 96, 64)):
            vol = pyvc.synth(9, dims, 12)
            for m, a in ((0, 0.0), (1, 0.5)):
                st, out = pyvc.downscale2x(vol, m, a)
                assert st == 0 and np.array_equal(out, oracle.ds2x(vol, m, a)), (dims, m, a)
            assert np.array_equal(pyvc.test_downsample2x(vol), oracle.downsample2x(vol)), dims
        vol = pyvc.synth(4, (97, 131, 75), 6)
        st, arc = pyvc.encode(vol, 10.0)
        ost, oarc = oracle.encode(vol, 10.0)
        assert st == 0 and ost == 0 and arc == oarc
    finally:
        pyvc.lib().vcx_test_wide_index(0)


def test_ds2x_argument_errors(pyvc):
    import ctypes as C
    L = pyvc.lib()
    d = pyvc.dev_alloc(64)
    try:
        assert L.vcx_downscale2x_device(d, pyvc.Dims(4, 4, 4), 2, C.c_float(0.5), d, 1) == 2
        assert L.vcx_downscale2x_device(d, pyvc.Dims(0, 4, 4), 0, C.c_float(0.5), d, 1) == 2
        assert L.vcx_downscale2x_device(d, pyvc.Dims(4, 4, 4), 1, C.c_float(float("nan")), d, 1) == 2
        assert L.vcx_downscale2x_device(None, pyvc.Dims(4, 4, 4), 0, C.c_float(0.5), d, 1) == 2
    finally:
        pyvc.dev_free(d)
