"""Parity at the benchmark workloads' full sizes, against the CPU oracle.

- 1024^3 (160 sheets, seed 2) at target ratios 10, 50 and 100: the archive
  vc_encode writes is byte-identical to the oracle's, and every atom of every
  LOD decoded by vc_decode_lod equals the oracle's decode of that archive
  (the reference's determinism and region == atom invariants,
  tests/test_coverage.c:66-77,113-142, at the size tests/test_coverage.c:
  163-190 names).
- 2048^3 (320 sheets, seed 3) at ratio 50, the serving archive: rate control
  and sampled atoms of LOD0 and LOD1 against the oracle's encoder, then
  open-loop Poisson traces of atom requests (1-4096 atoms, LOD 0 with p = 0.7)
  and region requests (log-uniform edges 17-600, any offset, LOD 0-7) through
  the continuous-batching server, every served byte against the oracle.
- 256 volumes with axes uniform in [40, 512] at ratios {10, 20, 50, 100}: one
  vcx_encode_batch call, every archive against the oracle's vc_encode.
"""
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the benchmark's trace generators)


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# 1024^3, ratios 10 / 50 / 100
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def vol1024(pyvc):
    assert pyvc.device_init() == 0
    return pyvc.synth(2, (1024, 1024, 1024), 160)


@pytest.mark.parametrize("ratio", [10.0, 50.0, 100.0])
def test_1024_encode_and_every_lod_vs_oracle(pyvc, oracle, vol1024, ratio):
    t0 = time.time()
    st, arc = pyvc.encode(vol1024, ratio)
    assert st == 0
    ost, oarc = oracle.encode(vol1024, ratio)
    assert ost == 0
    assert len(arc) == len(oarc), (len(arc), len(oarc))
    assert arc == oarc, f"archive differs at byte {np.flatnonzero(np.frombuffer(arc, np.uint8) != np.frombuffer(oarc, np.uint8))[0]}"
    a = pyvc.Archive(arc)
    oa = oracle.open(arc)
    for lod in range(8):
        st, dims = a.lod_dims(lod)
        assert st == 0
        st, dec = a.decode_lod(lod)
        ost, odec = oa.decode_lod(lod, dims)
        assert st == 0 and ost == 0
        assert np.array_equal(dec, odec), (ratio, lod)
    a.close()
    _say(f"1024^3 ratio {ratio:g}: {len(arc)} bytes, 8 LODs equal, {time.time() - t0:.1f} s")


# ---------------------------------------------------------------------------
# 2048^3 ratio 50: encoder samples, atom and region serving
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def serve2048(pyvc):
    assert pyvc.device_init() == 0
    n = 2048
    dvol = pyvc.dev_alloc(n ** 3)
    try:
        assert pyvc.synth_device(3, (n, n, n), 320, dvol) == 0
        st, arc = pyvc.encode_device(dvol, (n, n, n), 50.0)
        assert st == 0
        vol = pyvc.d2h(dvol, n ** 3).reshape(n, n, n)
    finally:
        pyvc.dev_free(dvol)
    yield vol, arc


def _hdr_q(arc):
    return np.frombuffer(arc, np.float32, 8, 72)


def _gather(vol, az, ay, ax):
    nz, ny, nx = vol.shape
    at = np.zeros((32, 32, 32), np.uint8)
    blk = vol[az * 32:min(nz, az * 32 + 32), ay * 32:min(ny, ay * 32 + 32), ax * 32:min(nx, ax * 32 + 32)]
    at[:blk.shape[0], :blk.shape[1], :blk.shape[2]] = blk
    return at


def test_2048_rate_control_and_sampled_payloads(pyvc, oracle, serve2048):
    vol, arc = serve2048
    q = _hdr_q(arc)
    q0 = oracle.rate_control(vol, 50.0)
    assert np.float32(q0) == q[0], (q0, q[0])
    assert np.array_equal(oracle.lod_schedule(q0, 8), q)
    a = pyvc.Archive(arc)
    rng = np.random.default_rng(41)
    lod1 = oracle.downsample2x(vol)
    for lod, v, k in ((0, vol, 192), (1, lod1, 96)):
        na = (v.shape[0] + 31) // 32
        for _ in range(k):
            az, ay, ax = (int(x) for x in rng.integers(0, na, 3))
            cov, off, ln = a.payload_range(lod, az, ay, ax)
            want = oracle.encode_atom(_gather(v, az, ay, ax), float(q[lod]))
            got = arc[off:off + ln] if cov == 2 else b""
            assert got == want, (lod, az, ay, ax, cov, ln, len(want))
    a.close()


def _serve(pyvc, a, reqs, qps, seed):
    """Open-loop Poisson submission through the server; returns the results
    (one array per request) and the statuses."""
    sizes = [bench.req_bytes(r) for r in reqs]
    offs = np.cumsum([0] + sizes)
    arena = pyvc.PinnedBuffer(int(offs[-1]))
    srv = pyvc.Server(a)
    gaps = np.random.default_rng(seed).exponential(1.0 / qps, len(reqs))
    ids, status = {}, {}
    try:
        t = time.perf_counter()
        for k, (kind, x) in enumerate(reqs):
            t += gaps[k]
            while time.perf_counter() < t:
                pass
            if kind == "atoms":
                i = srv.submit_atoms(x, arena.p + int(offs[k]))
            else:
                i = srv.submit_region(x[0], x[1], arena.p + int(offs[k]))
            assert i >= 0
            ids[int(i)] = k
        while len(status) < len(reqs):
            got, st, _ = srv.poll(wait_ms=100)
            for g, s_ in zip(got, st):
                status[ids[int(g)]] = int(s_)
    finally:
        srv.close()
    outs = [arena.a[int(offs[k]):int(offs[k + 1])].copy() for k in range(len(reqs))]
    arena.free()
    return outs, [status[k] for k in range(len(reqs))]


def test_2048_atom_serving_vs_oracle(pyvc, oracle, serve2048):
    _, arc = serve2048
    a = pyvc.Archive(arc)
    rng = np.random.default_rng(61)
    reqs = bench.atom_requests(a, 40, rng)
    # two full-size requests: 4096 LOD0 atoms, and 4096 LOD1 atoms (coarse,
    # several times the bins of an LOD0 atom)
    for lod in (0, 1):
        _, (nx, ny, nz) = a.lod_dims(lod)
        q = np.stack([np.full(4096, lod), rng.integers(0, nx // 32, 4096), rng.integers(0, ny // 32, 4096),
                      rng.integers(0, nz // 32, 4096)], 1).astype(np.int32)
        reqs.insert(5 + 10 * lod, ("atoms", q))
    t0 = time.time()
    outs, sts = _serve(pyvc, a, reqs, 400.0, 62)
    assert not any(sts)
    # the oracle's decode of every requested atom, LOD by LOD
    oa = oracle.open(arc)
    allq = np.concatenate([x for _, x in reqs])
    ref = np.zeros((len(allq), 32768), np.uint8)
    for lod in np.unique(allq[:, 0]):
        m = np.flatnonzero(allq[:, 0] == lod)
        zyx = allq[m][:, [3, 2, 1]].astype(np.uint32)
        _, out = oa.decode_atoms_omp(int(lod), zyx, 0, keep=True)
        ref[m] = out
    k = 0
    for r, (_, x) in enumerate(reqs):
        got = outs[r].reshape(len(x), 32768)
        assert np.array_equal(got, ref[k:k + len(x)]), r
        k += len(x)
    a.close()
    _say(f"2048^3 atom serving: {len(reqs)} requests, {len(allq)} atoms equal to the oracle, {time.time() - t0:.1f} s")


def test_2048_region_serving_vs_oracle(pyvc, oracle, serve2048):
    _, arc = serve2048
    a = pyvc.Archive(arc)
    rng = np.random.default_rng(71)
    reqs = bench.region_requests(a, 30, rng)
    # one large unaligned LOD0 box near the 600-voxel edge limit
    reqs.insert(7, ("region", (0, (101, 333, 517, 101 + 589, 333 + 577, 517 + 600))))
    t0 = time.time()
    outs, sts = _serve(pyvc, a, reqs, 400.0, 72)
    assert not any(sts)
    oa = oracle.open(arc)
    for r, (_, (lod, box)) in enumerate(reqs):
        ost, ref = oa.decode_region(lod, box)
        assert ost == 0
        assert np.array_equal(outs[r], ref.ravel()), (r, lod, box)
    a.close()
    _say(f"2048^3 region serving: {len(reqs)} boxes equal to the oracle, {time.time() - t0:.1f} s")


# ---------------------------------------------------------------------------
# 256 volumes of uneven dims, one batched encode
# ---------------------------------------------------------------------------
NVOL, CHUNK = 256, 32


@pytest.fixture(scope="module")
def batch256(pyvc):
    assert pyvc.device_init() == 0
    rng = np.random.default_rng(2024)       # the benchmark's volume set
    dims = [tuple(int(v) for v in rng.integers(40, 513, 3)) for _ in range(NVOL)]
    targets = [float((10, 20, 50, 100)[int(rng.integers(0, 4))]) for _ in range(NVOL)]
    vols = [pyvc.synth(5000 + k, d, max(1, round(20 * (d[0] * d[1] * d[2]) ** (1 / 3) / 128)))
            for k, d in enumerate(dims)]
    rc, res = pyvc.encode_batch(vols, targets)
    assert rc == 0
    return vols, targets, res


@pytest.mark.parametrize("chunk", range(NVOL // CHUNK))
def test_batch256_vs_oracle(oracle, batch256, chunk):
    vols, targets, res = batch256
    ks = range(chunk * CHUNK, (chunk + 1) * CHUNK)
    with ThreadPoolExecutor(4) as ex:
        ref = list(ex.map(lambda k: oracle.encode(vols[k], targets[k]), ks))
    for k, (ost, oarc) in zip(ks, ref):
        st, arc = res[k]
        assert st == 0 and ost == 0, (k, st, ost)
        assert arc == oarc, (k, vols[k].shape, targets[k], len(arc), len(oarc))
