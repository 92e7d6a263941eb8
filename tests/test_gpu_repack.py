"""GPU: container rewrite into the canonical v2 layout (vcx_repack,
vcx_repack_device) through the C ABI.  The bytes must equal those
tools/vc_repack.c wrote for the golden v1 archives, and the oracle's
restatement (vco_repack, pinned to the same fixtures) on archives made here."""
import os
import struct
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VC_ERR_FORMAT, VC_ERR_RANGE = 1, 2


def say(*a):
    print(*a, flush=True)


def test_repack_equals_reference_tool(pyvc, golden_containers, golden_repack):
    cases = dict(golden_containers)
    cases.update(golden_repack)
    for name, (m, b) in cases.items():
        say("repack", name)
        for tag in ("v1", "v2"):
            a = pyvc.Archive(b[tag])
            st, out = a.repack()
            a.close()
            assert st == 0 and out == b["v2"], (name, tag)


def test_repack_device_sizes_and_bounds(pyvc, golden_repack):
    m, b = golden_repack["repack_multi"]
    a = pyvc.Archive(b["v1"])
    n = len(b["v2"])
    st, need = a.repack_device(None, 0)
    assert st == VC_ERR_RANGE and need == n
    dev = pyvc.dev_alloc(n + 4096)
    try:
        pyvc.h2d(dev, np.full(n + 4096, 0xAB, np.uint8))
        st, need = a.repack_device(dev, n - 1)          # one byte short: nothing written
        assert st == VC_ERR_RANGE and need == n
        assert np.all(pyvc.d2h(dev, n + 4096) == 0xAB)
        st, need = a.repack_device(dev, n + 4096)
        assert st == 0 and need == n
        got = pyvc.d2h(dev, n + 4096)
        assert got[:n].tobytes() == b["v2"]
        assert np.all(got[n:] == 0xAB)
    finally:
        pyvc.dev_free(dev)
        a.close()


def _shuffled_v1(pyvc, vol, qs, seed):
    """a v1 archive of vol from this library's writer, atoms appended in a
    shuffled order, about 10% left absent and 10% marked zero"""
    nz, ny, nx = vol.shape
    lods = [vol]
    while True:
        z, y, x = lods[-1].shape
        if max(z, y, x) <= 32:
            break
        lods.append(lods[-1][::2, ::2, ::2])     # any content will do: the rewrite copies payloads
    rng = np.random.default_rng(seed)
    jobs = []
    for lod, v in enumerate(lods):
        lz, ly, lx = v.shape
        for az in range((lz + 31) // 32):
            for ay in range((ly + 31) // 32):
                for ax in range((lx + 31) // 32):
                    u = rng.random()
                    if u < 0.1:
                        continue
                    jobs.append((lod, az, ay, ax, u < 0.2))
    fd, path = tempfile.mkstemp(suffix=".vca")
    os.close(fd)
    try:
        w = pyvc.Writer(path, (nx, ny, nz), 10.0)
        for l in range(len(lods)):
            w.set_base_q(l, qs[min(l, len(qs) - 1)])
        for k in rng.permutation(len(jobs)):
            lod, az, ay, ax, zero = jobs[k]
            if zero:
                assert w.mark_zero_atom(lod, az, ay, ax) == 0
                continue
            v = lods[lod]
            a = np.zeros((32, 32, 32), np.uint8)
            sl = v[az * 32:az * 32 + 32, ay * 32:ay * 32 + 32, ax * 32:ax * 32 + 32]
            a[:sl.shape[0], :sl.shape[1], :sl.shape[2]] = sl
            assert w.append_atom(lod, az, ay, ax, a) == 0
        w.close()
        return open(path, "rb").read()
    finally:
        os.unlink(path)


def test_repack_equals_oracle_on_library_archives(pyvc, oracle):
    """A v1 of several regions per LOD from this library's writer, shuffled,
    and the v2 this library's encoder writes: both rewrite to the oracle's
    image, and the encoder's canonical v2 rewrites to itself."""
    dims = (1100, 72, 40)
    vol = pyvc.synth(17, dims, 5)
    v1 = _shuffled_v1(pyvc, vol, [6.0, 4.0, 2.5, 1.5, 1.0, 0.6, 0.4, 0.25], 5)
    ost, want = oracle.repack(v1)
    assert ost == 0
    a = pyvc.Archive(v1)
    st, out = a.repack()
    a.close()
    assert st == 0 and out == want
    st, v2 = pyvc.encode(vol, 20.0)
    assert st == 0
    a = pyvc.Archive(v2)
    st, out = a.repack()
    a.close()
    assert st == 0 and out == v2
    say("rewrote", len(v1), "->", len(want), "bytes")


def test_repack_rejects_payload_outside_archive(pyvc, oracle, golden_repack):
    m, b = golden_repack["repack_multi"]
    buf = bytearray(b["v1"])
    a0 = pyvc.Archive(bytes(buf))
    # a present atom's slot: its (offset, length, flag 1) record in the v1 block
    c, off, ln = next(r for r in (a0.payload_range(0, 0, ay, ax) for ay in range(3) for ax in range(32))
                      if r[0] == 2)
    a0.close()
    rec = struct.pack("<QI", off, ln) + b"\x01"
    k = bytes(buf).find(rec)
    assert k > 0
    struct.pack_into("<I", buf, k + 8, len(buf) - off + 1)
    assert oracle.repack(bytes(buf))[0] == VC_ERR_FORMAT
    a = pyvc.Archive(bytes(buf))
    st, out = a.repack()
    a.close()
    assert st == VC_ERR_FORMAT and out is None
