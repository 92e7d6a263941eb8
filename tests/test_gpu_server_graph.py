"""GPU: the request server's batches run as persistent kernels replayed from
captured hipGraphs, with launch shapes from a table the autotuner fills.
Served bytes must equal the oracle's decode whatever the shape: forced
shapes where entropy waves take many rounds of 64 atoms and transform
workgroups claim many atoms each, and the shapes the autotuner picks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A3 = 32768


@pytest.fixture(scope="module")
def arc_pair(pyvc, oracle):
    assert pyvc.device_init() == 0
    dims = (288, 256, 224)                  # 9 x 8 x 7 LOD0 atoms, 7 LODs
    vol = pyvc.synth(44, dims, 30)
    st, arc = pyvc.encode(vol, 20.0)
    assert st == 0
    return arc, pyvc.Archive(arc), oracle.open(arc)


def _trace(a, rng, n):
    nlod = 0
    while nlod < 8 and a.lod_dims(nlod)[0] == 0:
        nlod += 1
    reqs = []
    for k in range(n):
        lod = 0 if rng.random() < 0.6 else int(rng.integers(1, nlod))
        _, (nx, ny, nz) = a.lod_dims(lod)
        if rng.random() < 0.7:
            m = int(round(np.exp(rng.uniform(0, np.log(300)))))
            q = np.stack([np.full(m, lod), rng.integers(0, (nx + 31) // 32, m), rng.integers(0, (ny + 31) // 32, m),
                          rng.integers(0, (nz + 31) // 32, m)], 1).astype(np.int32)
            reqs.append(("atoms", q))
        else:
            lo = [int(rng.integers(0, d)) for d in (nx, ny, nz)]
            hi = [int(min(d, l + 1 + rng.integers(0, 90))) for d, l in zip((nx, ny, nz), lo)]
            reqs.append(("region", (lod, (lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]))))
    return reqs


def _serve(pyvc, srv, reqs):
    bufs, ids, done = [], {}, {}
    try:
        for k, (kind, r) in enumerate(reqs):
            if kind == "atoms":
                b = pyvc.PinnedBuffer(len(r) * A3)
                i = srv.submit_atoms(r, b.p)
            else:
                lod, (x0, y0, z0, x1, y1, z1) = r
                b = pyvc.PinnedBuffer((x1 - x0) * (y1 - y0) * (z1 - z0))
                i = srv.submit_region(lod, (x0, y0, z0, x1, y1, z1), b.p)
            assert i >= 0
            bufs.append(b)
            ids[int(i)] = k
        for _ in range(4000):
            got, st, _ = srv.poll(wait_ms=50)
            for g, s in zip(got, st):
                done[ids[int(g)]] = int(s)
            if len(done) == len(reqs):
                break
        assert len(done) == len(reqs)
        return [b.a.copy() for b in bufs], done
    finally:
        for b in bufs:
            b.free()


def _check(oa, reqs, outs, done):
    for k, (kind, r) in enumerate(reqs):
        assert done[k] == 0, k
        if kind == "atoms":
            got = outs[k].reshape(len(r), A3)
            for m, (lod, ax, ay, az) in enumerate(r):
                st, ref = oa.decode_atom(int(lod), int(ax), int(ay), int(az))
                assert st == 0 and np.array_equal(got[m], ref.ravel()), (k, m)
        else:
            lod, box = r
            st, ref = oa.decode_region(lod, box)
            assert st == 0 and np.array_equal(outs[k], ref.ravel()), (k, lod, box)


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 3, 5), (4, 1, 256), (1, 2, 7)])
def test_forced_shapes_vs_oracle(pyvc, arc_pair, shape):
    _, a, oa = arc_pair
    srv = pyvc.Server(a, 600 * pyvc.Server.atom_bytes() * 3)     # 600 atoms per batch slot
    try:
        assert srv.set_shape(*shape) == 0
        reqs = _trace(a, np.random.default_rng(sum(shape)), 24)
        outs, done = _serve(pyvc, srv, reqs)
    finally:
        srv.close()
    _check(oa, reqs, outs, done)


def test_autotune_table_and_results(pyvc, arc_pair):
    _, a, oa = arc_pair
    srv = pyvc.Server(a, 4096 * pyvc.Server.atom_bytes() * 3)    # buckets 64, 512, 4096
    try:
        before = srv.tuning()
        assert all(r[2] == 4 for r in before)                     # built-in shape: four-wave workgroups
        assert srv.autotune() == 0
        rows = srv.tuning()
        assert len(rows) == 8
        for cls, bucket, w, ge, gt, ms_def, ms_tuned in rows:
            assert w in (1, 2, 4) and ge >= 1 and gt >= 1
            if bucket <= 4096:
                assert 0 < ms_tuned <= ms_def, (cls, bucket, ms_tuned, ms_def)
        reqs = _trace(a, np.random.default_rng(77), 30)
        outs, done = _serve(pyvc, srv, reqs)
    finally:
        srv.close()
    _check(oa, reqs, outs, done)


def test_oversize_requests_split_and_device_destination(pyvc, arc_pair):
    """Slots of 40 atoms: a 150-atom request and boxes of more atoms than a
    slot are served in parts (one completion each), a box whose single atom
    layer exceeds a slot is rejected, and results land in device memory when
    out is a device pointer; every byte equals the oracle's decode."""
    _, a, oa = arc_pair
    srv = pyvc.Server(a, 40 * pyvc.Server.atom_bytes() * 3)
    assert srv.capacity() == 40
    rng = np.random.default_rng(9)
    try:
        _, (nx, ny, nz) = a.lod_dims(0)
        q = np.stack([np.zeros(150, np.int64), rng.integers(0, (nx + 31) // 32, 150),
                      rng.integers(0, (ny + 31) // 32, 150), rng.integers(0, (nz + 31) // 32, 150)], 1).astype(np.int32)
        reqs = [("atoms", q),
                ("region", (0, (5, 17, 3, 101, 60, 220))),       # 4 x 2 atoms a layer, 7 layers: 2 parts
                ("region", (0, (0, 0, 0, 288, 64, 224))),        # 9 x 2 atoms a layer, 7 layers: 4 parts
                ("region", (1, (3, 5, 7, 140, 100, 111))),       # LOD1, 5 x 4 atoms a layer, 4 layers: 2 parts
                ("atoms", q[:7])]
        outs, done = _serve(pyvc, srv, reqs)
        _check(oa, reqs, outs, done)
        # one atom layer of 7 x 8 atoms > 40: rejected
        big = pyvc.PinnedBuffer(8)
        assert srv.submit_region(0, (0, 0, 0, 224, 256, 10), big.p) == -1
        big.free()
        # device destinations (4 and 6 parts)
        d1 = pyvc.dev_alloc(150 * A3)
        box = (33, 40, 10, 200, 190, 170)
        nb = (box[3] - box[0]) * (box[4] - box[1]) * (box[5] - box[2])
        d2 = pyvc.dev_alloc(nb)
        try:
            ids = {srv.submit_atoms(q, d1): 0, srv.submit_region(0, box, d2): 1}
            st = {}
            for _ in range(400):
                got, s_, _ = srv.poll(wait_ms=50)
                for g, x in zip(got, s_):
                    st[ids[int(g)]] = int(x)
                if len(st) == 2:
                    break
            assert st == {0: 0, 1: 0}
            _check(oa, [("atoms", q), ("region", (0, box))], [pyvc.d2h(d1, 150 * A3), pyvc.d2h(d2, nb)], {0: 0, 1: 0})
        finally:
            pyvc.dev_free(d1)
            pyvc.dev_free(d2)
    finally:
        srv.close()


def test_capacity_from_free_memory(pyvc, arc_pair):
    _, a, _ = arc_pair
    srv = pyvc.Server(a, 0)
    try:
        assert 0 < srv.capacity() <= 32768
    finally:
        srv.close()
