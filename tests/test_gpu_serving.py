"""GPU: the continuous-batching request loop (vcx_server_*).  Every request
served in a mix of atom and region requests of uneven sizes, split over
several in-flight batches, decodes to the same bytes (and status) as the same
request decoded alone through vc_decode_atom / vc_decode_region."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A3 = 32768


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def served(pyvc):
    dims = (160, 128, 96)
    vol = pyvc.synth(21, dims, 24)
    st, arc = pyvc.encode(vol, 10.0)
    assert st == 0
    return pyvc.Archive(arc)


def _random_requests(a, rng, n_atom, n_region):
    nlod = 0
    while a.lod_dims(nlod)[0] == 0:
        nlod += 1
    reqs = []
    for k in range(n_atom + n_region):
        lod = 0 if rng.random() < 0.7 else int(rng.integers(1, nlod))
        _, (nx, ny, nz) = a.lod_dims(lod)
        if k < n_atom:
            n = int(round(np.exp(rng.uniform(0, np.log(40)))))
            q = np.stack([np.full(n, lod),
                          rng.integers(0, (nx + 31) // 32, n),
                          rng.integers(0, (ny + 31) // 32, n),
                          rng.integers(0, (nz + 31) // 32, n)], 1).astype(np.int32)
            reqs.append(("atoms", q))
        else:
            lo = [int(rng.integers(0, d)) for d in (nx, ny, nz)]
            hi = [int(min(d, l + 1 + rng.integers(0, 70))) for d, l in zip((nx, ny, nz), lo)]
            reqs.append(("region", (lod, (lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]))))
    rng.shuffle(reqs)
    return reqs


def _serve(pyvc, a, reqs, device_bytes):
    srv = pyvc.Server(a, device_bytes)
    bufs, ids = [], {}
    try:
        for k, (kind, r) in enumerate(reqs):
            if kind == "atoms":
                b = pyvc.PinnedBuffer(len(r) * A3)
                i = srv.submit_atoms(r, b.p)
            else:
                lod, (x0, y0, z0, x1, y1, z1) = r
                b = pyvc.PinnedBuffer((x1 - x0) * (y1 - y0) * (z1 - z0))
                i = srv.submit_region(lod, (x0, y0, z0, x1, y1, z1), b.p)
            assert i >= 0, (kind, r)
            bufs.append(b)
            ids[int(i)] = k
        done = {}
        for _ in range(2000):
            got, st, t = srv.poll(wait_ms=50)
            for g, s in zip(got, st):
                done[ids[int(g)]] = int(s)
            if len(done) == len(reqs):
                break
        assert len(done) == len(reqs)
        return [b.a.copy() for b in bufs], done
    finally:
        srv.close()
        for b in bufs:
            b.free()


def test_server_mix_equals_alone(pyvc, served):
    a = served
    rng = np.random.default_rng(3)
    reqs = _random_requests(a, rng, 28, 8)
    # a small budget: 48 atoms per batch slot, so the mix spans many batches
    outs, status = _serve(pyvc, a, reqs, 3 * 48 * pyvc.Server.atom_bytes())
    say("served", len(reqs), "requests")
    for k, (kind, r) in enumerate(reqs):
        assert status[k] == 0, (k, kind)
        if kind == "atoms":
            for m, (lod, ax, ay, az) in enumerate(r):
                st, ref = a.decode_atom(int(lod), int(ax), int(ay), int(az))
                assert st == 0 and np.array_equal(outs[k][m * A3:(m + 1) * A3], ref.ravel()), (k, m)
        else:
            lod, box = r
            st, ref = a.decode_region(lod, box)
            assert st == 0 and np.array_equal(outs[k], ref.ravel()), (k, box)


def test_server_rejects_invalid(pyvc, served):
    srv = pyvc.Server(served, 3 * 8 * pyvc.Server.atom_bytes())
    try:
        b = pyvc.PinnedBuffer(16 * A3)
        assert srv.capacity() == 8
        # 160 x 128 voxels: one atom layer of 5 x 4 atoms exceeds the 8-atom slot
        assert srv.submit_region(0, (0, 0, 0, 160, 128, 3), b.p) == -1
        assert srv.submit_atoms(np.array([[0, 99, 0, 0]], np.int32), b.p) == -1  # outside the grid
        assert srv.submit_atoms(np.array([[9, 0, 0, 0]], np.int32), b.p) == -1   # no such LOD
        assert srv.submit_region(0, (5, 5, 5, 5, 9, 9), b.p) == -1               # empty box
        assert srv.submit_region(0, (0, 0, 0, 161, 1, 1), b.p) == -1             # past the LOD edge
        i = srv.submit_atoms(np.array([[0, 1, 1, 1]], np.int32), b.p)
        assert i >= 0
        got = []
        for _ in range(200):
            g, st, t = srv.poll(wait_ms=50)
            got += list(zip(g, st))
            if got:
                break
        assert got == [(i, 0)]
        assert np.array_equal(b.a[:A3], served.decode_atom(0, 1, 1, 1)[1].ravel())
        b.free()
    finally:
        srv.close()


def test_server_corrupt_atom(pyvc, oracle, golden_containers):
    """A request holding an atom whose slot points past the archive end gets
    VC_ERR_FORMAT and zeros for that atom; the other atoms and requests in the
    same batch are unaffected."""
    m, b = golden_containers["synth_64x64x64"]
    buf = bytearray(b["v2"])
    a0 = pyvc.Archive(bytes(buf))
    c, off, ln = a0.payload_range(0, 0, 0, 1)
    good = [a0.decode_atom(0, 0, 0, 0)[1].ravel(), a0.decode_atom(0, 1, 1, 1)[1].ravel()]
    a0.close()
    idx = bytes(buf).find(off.to_bytes(8, "little") + ln.to_bytes(4, "little") + b"\x01")
    assert idx > 0
    buf[idx:idx + 8] = (len(buf) - 2).to_bytes(8, "little")
    a = pyvc.Archive(bytes(buf))
    reqs = [("atoms", np.array([[0, 0, 0, 0]], np.int32)),
            ("atoms", np.array([[0, 1, 0, 0], [0, 1, 1, 1]], np.int32)),
            ("region", (0, (0, 0, 0, 64, 64, 32)))]
    outs, status = _serve(pyvc, a, reqs, 3 * 16 * pyvc.Server.atom_bytes())
    assert status == {0: 0, 1: 1, 2: 1}
    assert np.array_equal(outs[0], good[0])
    assert not outs[1][:A3].any() and np.array_equal(outs[1][A3:], good[1])
    st, ref = a.decode_region(0, (0, 0, 0, 64, 64, 32))
    assert st == 1 and np.array_equal(outs[2], ref.ravel())


def test_recycled_context_starts_clean(pyvc, golden_containers):
    """The server's worker decodes a corrupt atom (reported per request, so
    its thread's status word is never collected) and exits; the next host
    thread that starts decoding picks up that thread's context and must not
    see the stale status: a valid decode there returns VC_OK."""
    import threading
    m, b = golden_containers["synth_64x64x64"]
    buf = bytearray(b["v2"])
    a0 = pyvc.Archive(bytes(buf))
    c, off, ln = a0.payload_range(0, 0, 0, 1)
    a0.close()
    idx = bytes(buf).find(off.to_bytes(8, "little") + ln.to_bytes(4, "little") + b"\x01")
    assert idx > 0
    buf[idx:idx + 8] = (len(buf) - 2).to_bytes(8, "little")
    bad = pyvc.Archive(bytes(buf))
    outs, status = _serve(pyvc, bad, [("atoms", np.array([[0, 1, 0, 0]], np.int32))], 3 * 4 * pyvc.Server.atom_bytes())
    assert status == {0: 1}
    good = pyvc.Archive(b["v2"])
    want = good.decode_atom(0, 1, 1, 0)
    got = []
    t = threading.Thread(target=lambda: got.append(good.decode_atom(0, 1, 1, 0)))
    t.start()
    t.join(timeout=60)
    assert got and got[0][0] == 0 == want[0] and np.array_equal(got[0][1], want[1])


def test_server_poll_served_timing(pyvc, served):
    """vcx_server_poll_served: every request completes once, with its status,
    submit <= admit <= done on the vcx_now_ns clock, the class of the batch
    that completed it (LOD0-only requests in class-0 batches, coarser ones in
    class 1) and a batch at least as large as the request's atoms (for a
    request served in parts, the batch of its last part)."""
    a = served
    rng = np.random.default_rng(8)
    reqs = _random_requests(a, rng, 20, 6)
    srv = pyvc.Server(a, 3 * 32 * pyvc.Server.atom_bytes())
    bufs, ids, natoms = [], {}, {}
    try:
        cap = srv.capacity()
        t0 = pyvc.now_ns()
        for k, (kind, r) in enumerate(reqs):
            if kind == "atoms":
                b = pyvc.PinnedBuffer(len(r) * A3)
                i = srv.submit_atoms(r, b.p)
                natoms[k] = len(r)
            else:
                lod, (x0, y0, z0, x1, y1, z1) = r
                b = pyvc.PinnedBuffer((x1 - x0) * (y1 - y0) * (z1 - z0))
                i = srv.submit_region(lod, (x0, y0, z0, x1, y1, z1), b.p)
                natoms[k] = ((x1 - 1) // 32 - x0 // 32 + 1) * ((y1 - 1) // 32 - y0 // 32 + 1) * \
                    ((z1 - 1) // 32 - z0 // 32 + 1)
            assert i >= 0
            bufs.append(b)
            ids[int(i)] = k
        recs = {}
        for _ in range(2000):
            for rec in srv.poll_served(wait_ms=50):
                k = ids[int(rec[0])]
                assert k not in recs
                recs[k] = rec
            if len(recs) == len(reqs):
                break
        t1 = pyvc.now_ns()
        assert len(recs) == len(reqs)
        for k, (kind, r) in enumerate(reqs):
            _, st, cls, ts, ta, td, ba = recs[k]
            lod = int(r[0, 0]) if kind == "atoms" else r[0]
            coarse = bool((r[:, 0] != 0).any()) if kind == "atoms" else lod != 0
            assert st == 0 and cls == int(coarse), (k, cls, coarse)
            assert t0 <= ts <= ta <= td <= t1, (k, ts, ta, td)
            # a request served in parts reports the batch of its last part
            assert (ba >= natoms[k] if natoms[k] <= cap else ba >= 1) and ba <= cap, (k, ba, natoms[k], cap)
    finally:
        srv.close()
        for b in bufs:
            b.free()
