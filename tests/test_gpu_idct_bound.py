"""GPU: the inverse transform's two paths.  Atoms whose coefficient L2 norm
is within the matrix-core path's bound (31000) run as int8 products over an
i16 tile; larger ones are handed to the i32 path by a second kernel on the
same stream.  Payloads with levels up to 32767 (range-coded by the oracle)
are decoded through every entry point that launches a transform (payload
batches, per-call atom/list/box/LOD decodes, the request server's persistent
kernels, the allocation-free decode) and checked against the oracle's
decode of the same bytes."""
import ctypes as C
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A3 = 32768


def _levels(rng, oracle, q, outside):
    """a level field (scan order) whose dequantized L2 is far inside the
    bound, or past it for q >= 0.5 (one level of 25000..32767 at scan
    position 100..399, frequency 8..14)"""
    lv = np.zeros(A3, np.int16)
    k = int(rng.integers(5, 120))
    pos = np.unique(rng.integers(0, 3000, k))
    v = rng.integers(-60, 61, len(pos))
    lv[pos] = np.where(v == 0, 1, v)
    if outside:
        lv[int(rng.integers(100, 400))] = int(rng.choice([-1, 1])) * int(rng.integers(25000, 32768))
    return lv


def _payload(oracle, rng, q, outside):
    rc = oracle.entropy_encode(_levels(rng, oracle, q, outside))
    assert rc is not None
    return bytes([int(rng.integers(0, 256))]) + rc


def test_payload_batches_both_paths(pyvc, oracle):
    """1100 payloads, one in ten past the bound, in one batch (the second
    kernel scans three chunks of 512 items)."""
    assert pyvc.device_init() == 0
    rng = np.random.default_rng(8)
    q = 1.3
    pays = [_payload(oracle, rng, q, k % 10 == 3) for k in range(1100)]
    st, out, per = pyvc.decode_payloads(pays, q)
    assert st == 0 and not per.any()
    for k in range(len(pays)):
        ost, ref = oracle.decode_payload(pays[k], q)
        assert ost == 0 and np.array_equal(out[k], ref), k


def _patch(arc, oracle_arc, rng, oracle, q_of, atoms):
    """replace the payloads of `atoms` (lod, ax, ay, az, outside) in a v2
    archive with crafted ones no longer than the originals"""
    b = bytearray(arc)
    for lod, ax, ay, az, outside in atoms:
        cov, off, ln = oracle_arc.coverage(lod, az, ay, ax)
        assert cov == 2
        for _ in range(50):
            p = _payload(oracle, rng, q_of[lod], outside)
            if len(p) <= ln:
                break
        assert len(p) <= ln
        slot = struct.pack("<QI", off, ln)
        at = b.find(slot)
        assert at >= 0 and b.find(slot, at + 1) < 0
        b[at + 8:at + 12] = struct.pack("<I", len(p))
        b[at + 12] = 1                          # PRESENT
        b[at + 13] = p[0]                       # slot dc = payload dc
        b[off:off + len(p)] = p
    return bytes(b)


@pytest.fixture(scope="module")
def patched(pyvc, oracle):
    assert pyvc.device_init() == 0
    dims = (160, 128, 96)                       # 5 x 4 x 3 LOD0 atoms
    st, arc = pyvc.encode(pyvc.synth(12, dims, 8), 10.0)
    assert st == 0
    q_of = np.frombuffer(arc, np.float32, 8, 72)
    rng = np.random.default_rng(21)
    atoms = [(0, 1, 1, 1, True), (0, 4, 3, 2, True), (0, 0, 0, 0, False), (0, 2, 3, 0, True),
             (1, 1, 1, 0, True), (1, 0, 0, 1, False), (2, 0, 0, 0, True)]
    buf = _patch(arc, oracle.open(arc), rng, oracle, q_of, atoms)
    return buf, atoms


def test_archive_atoms_past_the_bound(pyvc, oracle, patched):
    buf, atoms = patched
    a = pyvc.Archive(buf)
    oa = oracle.open(buf)
    try:
        # every LOD through the box path
        for lod in range(3):
            _, dims = a.lod_dims(lod)
            st, dec = a.decode_lod(lod)
            ost, ref = oa.decode_lod(lod, dims)
            assert st == 0 and ost == 0 and np.array_equal(dec, ref), lod
        # single atoms and one list batch
        quads = np.array([(lod, ax, ay, az) for lod, ax, ay, az, _ in atoms] + [(0, 3, 2, 1), (1, 2, 1, 1)], np.int32)
        st, got, per = a.decode_atoms(quads)
        assert st == 0 and not per.any()
        for k, (lod, ax, ay, az) in enumerate(quads):
            ost, ref = oa.decode_atom(int(lod), int(ax), int(ay), int(az))
            assert ost == 0 and np.array_equal(got[k].reshape(32, 32, 32), ref), k
            st1, one = a.decode_atom(int(lod), int(ax), int(ay), int(az))
            assert st1 == 0 and np.array_equal(one, ref), k
    finally:
        a.close()


def test_server_and_noalloc_past_the_bound(pyvc, oracle, patched):
    buf, atoms = patched
    a = pyvc.Archive(buf)
    oa = oracle.open(buf)
    quads = np.array([(lod, ax, ay, az) for lod, ax, ay, az, _ in atoms] * 3, np.int32)
    try:
        # the server's persistent kernels: an atom request and a box over a patched atom
        srv = pyvc.Server(a, 64 * pyvc.Server.atom_bytes() * 3)
        ob = pyvc.PinnedBuffer(len(quads) * A3)
        box = (20, 20, 20, 100, 70, 90)
        nb = 80 * 50 * 70
        rb = pyvc.PinnedBuffer(nb)
        try:
            ids = {srv.submit_atoms(quads, ob.p): 0, srv.submit_region(0, box, rb.p): 1}
            done = {}
            for _ in range(400):
                got, st, _ = srv.poll(wait_ms=50)
                for g, s in zip(got, st):
                    done[ids[int(g)]] = int(s)
                if len(done) == 2:
                    break
            assert done == {0: 0, 1: 0}
            res = ob.a.reshape(len(quads), A3)
            for k, (lod, ax, ay, az) in enumerate(quads):
                ost, ref = oa.decode_atom(int(lod), int(ax), int(ay), int(az))
                assert np.array_equal(res[k].reshape(32, 32, 32), ref), k
            ost, ref = oa.decode_region(0, box)
            assert ost == 0 and np.array_equal(rb.a, ref.ravel())
        finally:
            srv.close()
            ob.free()
            rb.free()
        # the allocation-free decode (after a decode made the device copy)
        assert a.decode_atom(0, 0, 0, 0)[0] == 0
        L = pyvc.lib()
        hip = C.CDLL("libamdhip64.so")
        hip.hipStreamCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
        hip.hipStreamSynchronize.argtypes = [C.c_void_p]
        hip.hipStreamDestroy.argtypes = [C.c_void_p]
        stream = C.c_void_p()
        assert hip.hipStreamCreateWithFlags(C.byref(stream), 1) == 0
        n = len(atoms)
        dq = pyvc.dev_alloc(n * 16)
        dout = pyvc.dev_alloc(n * A3)
        dst = pyvc.dev_alloc(n * 4)
        try:
            pyvc.h2d(dq, np.ascontiguousarray(quads[:n]).view(np.uint8))
            assert L.vcx_decode_atoms_noalloc(a.h, n, dq, dout, dst, stream) == 0
            assert hip.hipStreamSynchronize(stream) == 0
            got = pyvc.d2h(dout, n * A3).reshape(n, A3)
            assert not pyvc.d2h(dst, n * 4).view(np.int32).any()
            for k in range(n):
                ost, ref = oa.decode_atom(*(int(v) for v in quads[k]))
                assert np.array_equal(got[k].reshape(32, 32, 32), ref), k
        finally:
            pyvc.dev_free(dq)
            pyvc.dev_free(dout)
            pyvc.dev_free(dst)
            hip.hipStreamDestroy(stream)
    finally:
        a.close()
