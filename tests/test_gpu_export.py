"""GPU: the exporter cascade (vcx_export_volume) against the exporter's own
order restated in the oracle (tools/vc_export_stream.c, single compute
thread) driving the reference-built writer: the v1 files must be byte for
byte the same (pyramid, per-band flush order, empty-band stamps, prepass zero
regions, coarse tail from the decoded archive)."""
import os
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _first_diff(a, b):
    n = min(len(a), len(b))
    x = np.frombuffer(a[:n], np.uint8) != np.frombuffer(b[:n], np.uint8)
    return int(np.argmax(x)) if x.any() else n


@pytest.mark.parametrize("case", [
    dict(dims=(330, 250, 170), seed=4),
    dict(dims=(330, 250, 170), seed=4, tile=256, method=0, q0=2.5, q_falloff=0.6, chunk=(64, 64, 64)),
    dict(dims=(1100, 40, 60), seed=6, alpha=0.8, chunk=(256, 32, 16)),
    dict(dims=(200, 170, 90), seed=9, device=True),
])
def test_export_matches_reference_writer(pyvc, oracle, ref, case):
    assert pyvc.device_init() == 0
    kw = {k: case[k] for k in ("q0", "q_falloff", "method", "alpha", "tile", "chunk") if k in case}
    vol = pyvc.synth(case["seed"], case["dims"], 10)
    d = tempfile.mkdtemp()
    pg, pr = os.path.join(d, "gpu.vca"), os.path.join(d, "ref.vca")
    try:
        if case.get("device"):
            dev = pyvc.dev_alloc(vol.size)
            try:
                pyvc.h2d(dev, vol.ravel())
                assert pyvc.export_volume(dev, pg, dims=case["dims"], **kw) == 0
            finally:
                pyvc.dev_free(dev)
        else:
            assert pyvc.export_volume(vol, pg, **kw) == 0
        ref.export_volume(oracle, vol, pr, **kw)
        g, r = open(pg, "rb").read(), open(pr, "rb").read()
    finally:
        for f in (pg, pr):
            if os.path.exists(f):
                os.unlink(f)
        os.rmdir(d)
    assert len(g) == len(r) and g == r, (len(g), len(r), _first_diff(g, r))
    a = pyvc.Archive(g)
    assert a.version() == 1 and a.lod_dims(0)[1] == tuple((v + 1023) // 1024 * 1024 for v in case["dims"])
    a.close()


def test_export_rejects_bad_arguments(pyvc):
    assert pyvc.device_init() == 0
    vol = np.zeros((8, 8, 8), np.uint8)
    assert pyvc.export_volume(vol, "/tmp/x.vca", method=5) == pyvc.VC_ERR_RANGE
    assert pyvc.export_volume(vol, "/tmp/x.vca", q0=0.0) == pyvc.VC_ERR_RANGE
    assert pyvc.export_volume(vol, "/tmp/x.vca", chunk=(0, 8, 8)) == pyvc.VC_ERR_RANGE
    assert pyvc.export_volume(vol, "/nonexistent-dir/x.vca") == pyvc.VC_ERR_IO
