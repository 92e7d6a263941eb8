"""Edge cases of the reference's own tests, restated on the GPU path and
checked against the CPU oracle:

- edge dims (tests/test_coverage.c:163-190): tall/thin 2048x32x32, primes
  97x131x193, 1x1x1 and one atom, through vc_encode / vc_decode_lod and
  through the writer (vc_create + vc_append_atom at q = 1);
- q-monotone size and quality (tests/test_coverage.c:80-110);
- 16 threads appending disjoint atoms into the same regions of one writer
  (tests/test_concurrent.c:12-49);
- streaming archives against the oracle (tests/test_streaming.c:48-64).
"""
import os
import tempfile
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EDGE_DIMS = [(2048, 32, 32), (97, 131, 193), (1, 1, 1), (32, 32, 32)]


def _nlod(a):
    n = 0
    while n < 8 and a.lod_dims(n)[0] == 0:
        n += 1
    return n


def _every_lod_equal(a, oa):
    for lod in range(_nlod(a)):
        _, dims = a.lod_dims(lod)
        st, dec = a.decode_lod(lod)
        ost, odec = oa.decode_lod(lod, dims)
        assert st == 0 and ost == 0 and np.array_equal(dec, odec), lod


@pytest.mark.parametrize("dims", EDGE_DIMS)
def test_edge_dims_encode_vs_oracle(pyvc, oracle, dims):
    assert pyvc.device_init() == 0
    vol = pyvc.synth(90 + dims[0] % 7, dims, 3)
    for ratio in (10.0, 50.0):
        st, arc = pyvc.encode(vol, ratio)
        ost, oarc = oracle.encode(vol, ratio)
        assert st == 0 and ost == 0 and arc == oarc, (dims, ratio)
        a = pyvc.Archive(arc)
        _every_lod_equal(a, oracle.open(arc))
        a.close()


def _fill(az, ay, ax):
    """the concurrency test's atom pattern (tests/test_concurrent.c:16-18)"""
    i = np.arange(32768, dtype=np.int64)
    return ((az * 7 + ay * 13 + ax * 17 + i) % 200 + 10).astype(np.uint8)


def _writer_file(pyvc, dims, ratio, atoms, threads=1):
    """v1 archive from vc_append_atom at q = 1 on every LOD; atoms: list of
    (az, ay, ax); returns the file bytes"""
    fd, path = tempfile.mkstemp(suffix=".vca")
    os.close(fd)
    try:
        w = pyvc.Writer(path, dims, ratio)
        for lod in range(8):
            w.set_base_q(lod, 1.0)
        errs = []

        def run(k):
            for az, ay, ax in atoms[k::threads]:
                if w.append_atom(0, az, ay, ax, _fill(az, ay, ax)) != 0:
                    errs.append((az, ay, ax))
        ts = [threading.Thread(target=run, args=(k,)) for k in range(threads)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        w.close()
        assert not errs, errs[:4]
        return open(path, "rb").read()
    finally:
        os.unlink(path)


def _check_atoms(pyvc, oracle, buf, atoms):
    a = pyvc.Archive(buf)
    oa = oracle.open(buf)
    for az, ay, ax in atoms:
        assert a.coverage(0, az, ay, ax) == 2, (az, ay, ax)
        cov, off, ln = a.payload_range(0, az, ay, ax)
        assert buf[off:off + ln] == oracle.encode_atom(_fill(az, ay, ax), 1.0), (az, ay, ax)
    q = np.array([(0, ax, ay, az) for az, ay, ax in atoms], np.int32)
    st, got, per = a.decode_atoms(q)
    assert st == 0 and not per.any()
    for k, (az, ay, ax) in enumerate(atoms):
        ost, ref = oa.decode_atom(0, ax, ay, az)
        assert ost == 0 and np.array_equal(got[k].reshape(32, 32, 32), ref), (az, ay, ax)
    a.close()


@pytest.mark.parametrize("dims", [(2048, 32, 32), (97, 131, 193)])
def test_edge_dims_writer(pyvc, oracle, dims):
    assert pyvc.device_init() == 0
    nx, ny, nz = ((d + 31) // 32 for d in dims)
    atoms = [(az, ay, ax) for az in range(nz) for ay in range(ny) for ax in range(nx)]
    buf = _writer_file(pyvc, dims, 20.0, atoms)
    _check_atoms(pyvc, oracle, buf, atoms)


def test_concurrent_appends_16_threads(pyvc, oracle):
    """16 threads stride over the atom grid of a 512x256x256 writer: disjoint
    atoms, the same regions (the exporter's pattern).  Every atom is present,
    its payload is the oracle's encode of it, and it decodes like the
    oracle's decode."""
    assert pyvc.device_init() == 0
    dims = (512, 256, 256)
    atoms = [(az, ay, ax) for az in range(8) for ay in range(8) for ax in range(16)]
    buf = _writer_file(pyvc, dims, 30.0, atoms, threads=16)
    _check_atoms(pyvc, oracle, buf, atoms)


def _psnr(a, b):
    d = a.astype(np.float64) - b.astype(np.float64)
    mse = float((d * d).mean())
    return 99.0 if mse == 0 else 10 * np.log10(255.0 ** 2 / mse)


def test_q_monotone_size_and_psnr(pyvc, oracle):
    """q up -> archive size down, sampled-atom PSNR down or equal (+1 dB
    slack), each archive byte-identical to the oracle's"""
    assert pyvc.device_init() == 0
    dims = (256, 256, 256)
    vol = pyvc.synth(12, dims, 40)
    prev_size, prev_psnr = None, None
    for q in (0.5, 1.0, 2.0, 4.0, 8.0, 16.0):
        qs = [q] * 8
        st, arc = pyvc.encode_q(vol, 20.0, qs)
        ost, oarc = oracle.encode_q(vol, 20.0, qs)
        assert st == 0 and ost == 0 and arc == oarc, q
        a = pyvc.Archive(arc)
        st, dec = a.decode_lod(0)
        a.close()
        assert st == 0
        p = float(np.mean([_psnr(vol[z:z + 32, y:y + 32, x:x + 32], dec[z:z + 32, y:y + 32, x:x + 32])
                           for z in range(0, 256, 64) for y in range(0, 256, 64) for x in range(0, 256, 64)]))
        if prev_size is not None:
            assert len(arc) <= prev_size, (q, len(arc), prev_size)
            assert p <= prev_psnr + 1.0, (q, p, prev_psnr)
        prev_size, prev_psnr = len(arc), p


def test_streaming_vs_oracle(pyvc, oracle):
    """vc_open_streaming (region blobs cached in HBM, and the per-atom read
    path with a cache too small for any blob) against the oracle's decode of
    the same bytes: every LOD and random unaligned boxes, on an archive of
    several LOD0 regions (tests/test_streaming.c:48-64)."""
    assert pyvc.device_init() == 0
    dims = (1100, 70, 40)                      # two LOD0 regions along x
    vol = pyvc.synth(21, dims, 4)
    st, arc = pyvc.encode(vol, 20.0)
    assert st == 0
    oa = oracle.open(arc)
    rng = np.random.default_rng(3)
    for cache in (0, 1 << 20):
        s = pyvc.StreamingArchive(arc)
        s.set_cache(cache)
        _every_lod_equal(s, oa)
        for _ in range(12):
            lod = int(rng.integers(0, 3))
            _, (nx, ny, nz) = s.lod_dims(lod)
            lo = [int(rng.integers(0, d)) for d in (nx, ny, nz)]
            hi = [int(rng.integers(l + 1, d + 1)) for l, d in zip(lo, (nx, ny, nz))]
            box = (lo[0], lo[1], lo[2], hi[0], hi[1], hi[2])
            st, got = s.decode_region(lod, box)
            ost, ref = oa.decode_region(lod, box)
            assert st == 0 and ost == 0 and np.array_equal(got, ref), (cache, lod, box)
        if cache == 0:
            assert s.stats()[0] > 0                # served from region blobs
        s.close()
