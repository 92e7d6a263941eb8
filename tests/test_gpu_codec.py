"""GPU parity: the HIP library (libvc.so, called through its C ABI) against the
reference's golden fixtures and the CPU oracle, bit for bit."""
import hashlib
import os
import tempfile

import numpy as np
import pytest

from conftest import container_volume, payloads_of
from test_oracle import eg_prefix_stream

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def say(*a):
    print(*a, flush=True)


def test_device_init(pyvc):
    assert pyvc.device_init() == 0


def test_atom_payloads_match_reference(pyvc, golden_atoms):
    atoms, qs = golden_atoms["atoms"], golden_atoms["qs"]
    for qi, q in enumerate(qs):
        say("encode atoms q", q)
        st, pays = pyvc.encode_atoms(atoms, float(q))
        assert st == 0
        exp = [p for (ai, qj, p, _) in sorted(payloads_of(golden_atoms, qi))]
        assert pays == exp, q


def test_payload_decode_matches_reference(pyvc, golden_atoms):
    qs = golden_atoms["qs"]
    for qi, q in enumerate(qs):
        items = sorted(payloads_of(golden_atoms, qi))
        st, out, per = pyvc.decode_payloads([p for (_, _, p, _) in items], float(q))
        assert st == 0 and not per.any()
        for k, (ai, _, p, rsha) in enumerate(items):
            assert sha(out[k]) == rsha, (q, ai)


def test_atom_rt_hook(pyvc, golden_atoms, oracle):
    atoms = golden_atoms["atoms"]
    for i in (0, 13, 17):
        n, rec = pyvc.test_atom_rt(atoms[i], 1.0)
        pay = oracle.encode_atom(atoms[i], 1.0)
        assert n == len(pay)
        if pay:
            assert np.array_equal(rec, oracle.decode_payload(pay, 1.0)[1])


def test_damaged_payloads_match_reference(pyvc, golden_corrupt):
    g = golden_corrupt
    blob = g["blob"].tobytes()
    for q in np.unique(g["qs"]):
        ks = [k for k in range(len(g["lens"])) if g["qs"][k] == q]
        pays = [blob[int(g["offs"][k]):int(g["offs"][k]) + int(g["lens"][k])] for k in ks]
        st, out, per = pyvc.decode_payloads(pays, float(q))
        assert st == 0
        for j, k in enumerate(ks):
            assert sha(out[j]) == str(g["rec_sha"][k])


def test_eg_prefix_cap_rejected(pyvc):
    bad = bytes([100]) + eg_prefix_stream(70)
    good = bytes([100]) + eg_prefix_stream(20)
    st, out, per = pyvc.decode_payloads([bad, good], 1.0)
    assert st == 1 and list(per) == [1, 0] and not out[0].any()


def test_downsample_matches(pyvc, oracle):
    rng = np.random.default_rng(3)
    for shp in [(1, 1, 1), (2, 3, 5), (33, 17, 5), (64, 64, 64), (97, 131, 65)]:
        v = rng.integers(0, 256, shp, dtype=np.uint8)
        assert np.array_equal(pyvc.test_downsample2x(v), oracle.downsample2x(v)), shp


def test_synth_matches_oracle(pyvc, oracle):
    for seed, (nx, ny, nz), ns in [(3, (96, 80, 70), 5), (7, (33, 17, 5), 2), (11, (128, 128, 128), 20)]:
        assert np.array_equal(pyvc.synth(seed, (nx, ny, nz), ns), oracle.synth(seed, nx, ny, nz, ns))


def test_encode_q_matches_reference_container(pyvc, oracle, golden_containers):
    for name, (m, b) in golden_containers.items():
        say("encode_q", name)
        vol = container_volume(oracle, m)
        st, arc = pyvc.encode_q(vol, m["target_ratio"], m["qs"])
        assert st == 0 and arc == b["v2"], name


def _lod_shapes(nx, ny, nz, nlod):
    out = []
    for _ in range(nlod):
        out.append((nx, ny, nz))
        nx, ny, nz = (nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2
    return out


def test_decode_matches_oracle(pyvc, oracle, golden_containers):
    rng = np.random.default_rng(5)
    for name, (m, b) in golden_containers.items():
        nz, ny, nx = m["shape_zyx"]
        for tag in ("v1", "v2"):
            say("decode", name, tag)
            a, o = pyvc.Archive(b[tag]), oracle.open(b[tag])
            for lod, (lx, ly, lz) in enumerate(_lod_shapes(nx, ny, nz, oracle.nlod(nx, ny, nz))):
                st, vol = a.decode_lod(lod)
                ost, ovol = o.decode_lod(lod, (lx, ly, lz))
                assert st == ost == 0 and np.array_equal(vol, ovol), (name, tag, lod)
                quads = [(lod, ax, ay, az) for az in range((lz + 31) // 32) for ay in range((ly + 31) // 32)
                         for ax in range((lx + 31) // 32)]
                st, out, per = a.decode_atoms(quads)
                assert st == 0 and not per.any()
                for k, (_, ax, ay, az) in enumerate(quads):
                    assert np.array_equal(out[k].reshape(32, 32, 32), o.decode_atom(lod, ax, ay, az)[1])
            st, one = a.decode_atom(0, 0, 0, 0)
            assert st == 0 and np.array_equal(one, o.decode_atom(0, 0, 0, 0)[1])
            for _ in range(6):
                x0, y0, z0 = (int(rng.integers(0, d)) for d in (nx, ny, nz))
                x1, y1, z1 = int(rng.integers(x0 + 1, nx + 1)), int(rng.integers(y0 + 1, ny + 1)), \
                    int(rng.integers(z0 + 1, nz + 1))
                box = (x0, y0, z0, x1, y1, z1)
                st, reg = a.decode_region(0, box)
                ost, oreg = o.decode_region(0, box)
                assert st == ost == 0 and np.array_equal(reg, oreg), box
            a.close()


def test_decode_range_errors(pyvc, golden_containers):
    m, b = golden_containers["synth_64x64x64"]
    a = pyvc.Archive(b["v2"])
    assert a.decode_atom(0, 2, 0, 0)[0] == 2 and a.decode_atom(0, 0, 0, -1)[0] == 2
    assert a.decode_atom(9, 0, 0, 0)[0] == 2
    assert a.decode_region(0, (0, 0, 0, 65, 10, 10))[0] == 2
    assert a.decode_region(0, (5, 0, 0, 5, 10, 10))[0] == 2
    assert a.decode_atoms([(0, 0, 0, 0), (0, 3, 0, 0)])[0] == 2


def test_corrupt_slot_reports_format(pyvc, oracle, golden_containers):
    """A PRESENT slot whose payload runs past the archive end decodes to zeros
    with VC_ERR_FORMAT (src/vc/vc.c:1413); its neighbours still decode."""
    m, b = golden_containers["synth_64x64x64"]
    buf = bytearray(b["v2"])
    a0 = pyvc.Archive(bytes(buf))
    c, off, ln = a0.payload_range(0, 0, 0, 1)
    a0.close()
    # find the slot of atom (0,0,1) and point it past the end
    idx = bytes(buf).find(off.to_bytes(8, "little") + ln.to_bytes(4, "little") + b"\x01")
    assert idx > 0
    buf[idx:idx + 8] = (len(buf) - 2).to_bytes(8, "little")
    a, o = pyvc.Archive(bytes(buf)), oracle.open(bytes(buf))
    st, out = a.decode_atom(0, 1, 0, 0)
    assert st == 1 and not out.any()
    st, vol = a.decode_lod(0)
    ost, ovol = o.decode_lod(0, (64, 64, 64))
    assert st == ost == 1 and np.array_equal(vol, ovol)


def test_streaming_matches_flat(pyvc, golden_containers):
    m, b = golden_containers["synth_70x80x96"]
    for tag in ("v1", "v2"):
        f, s = pyvc.Archive(b[tag]), pyvc.StreamingArchive(b[tag])
        for az in range(3):
            for ay in range(3):
                for ax in range(3):
                    s1, v1 = f.decode_atom(0, ax, ay, az)
                    s2, v2 = s.decode_atom(0, ax, ay, az)
                    assert s1 == s2 and np.array_equal(v1, v2)
        box = (3, 5, 7, 90, 77, 64)
        assert np.array_equal(f.decode_region(0, box)[1], s.decode_region(0, box)[1])
        assert s.reads > 0
    # a failing byte source reads as a format error with zeros (src/vc/vc.c:1413-1415)
    bad = pyvc.StreamingArchive(b["v2"], fail_after=1)
    st, out = bad.decode_atom(0, 1, 1, 1)
    assert st in (0, 1) and not out.any()


def test_vc_encode_matches_oracle_config1(pyvc, oracle):
    """vc_encode on the 128^3 synthetic volume (20 sheets) at ratio 10: the
    rate control picks the same q0 and the archive is byte-identical."""
    vol = oracle.synth(1, 128, 128, 128, 20)
    for ratio in (10.0, 50.0):
        q_gpu = pyvc.rate_control(vol, ratio)
        q_cpu = oracle.rate_control(vol, ratio)
        assert q_gpu == q_cpu, ratio
        st, arc = pyvc.encode(vol, ratio)
        ost, oarc = oracle.encode(vol, ratio)
        assert st == ost == 0 and arc == oarc, ratio
        a = pyvc.Archive(arc)
        st, dec = a.decode_lod(0)
        assert st == 0 and np.array_equal(dec, oracle.open(arc).decode_lod(0, (128, 128, 128))[1])


def test_vc_encode_ragged_dims(pyvc, oracle):
    for seed, (nx, ny, nz), ns, r in [(2, (97, 40, 65), 4, 20.0), (6, (33, 1, 7), 1, 10.0), (8, (1, 1, 1), 0, 100.0)]:
        vol = oracle.synth(seed, nx, ny, nz, ns)
        st, arc = pyvc.encode(vol, r)
        ost, oarc = oracle.encode(vol, r)
        assert st == ost == 0 and arc == oarc, (nx, ny, nz)


def test_vc_encode_argument_errors(pyvc):
    vol = np.zeros((4, 4, 4), np.uint8)
    assert pyvc.encode(vol, 0.5)[0] == 2
    assert pyvc.encode(np.zeros((0, 4, 4), np.uint8), 10.0)[0] == 2


def test_all_zero_volume(pyvc, oracle):
    vol = np.zeros((40, 50, 70), np.uint8)
    st, arc = pyvc.encode(vol, 10.0)
    ost, oarc = oracle.encode(vol, 10.0)
    assert st == ost == 0 and arc == oarc
    a = pyvc.Archive(arc)
    assert a.coverage(0, 0, 0, 0) == 1                  # KNOWN_ZERO
    assert not a.decode_lod(0)[1].any()


def test_encode_archive_outgrows_estimate(pyvc, oracle):
    """The library allocates each archive ahead at an estimate from the target
    ratio; an explicit fine base_q with a high target makes the archive several
    times that estimate, and the archive must still equal the oracle's."""
    vol = oracle.synth(3, 200, 176, 160, 12)
    nlod = oracle.nlod(200, 176, 160)
    qs = [0.05] * nlod
    say("encode_q past the estimate", vol.shape)
    st, arc = pyvc.encode_q(vol, 1000.0, qs)
    ost, oarc = oracle.encode_q(vol, 1000.0, qs)
    assert st == ost == 0
    est = 2 * vol.size // 1000 + 8 * 512 * 1024 + 4096 + 64 * 1024
    assert len(arc) > 2 * est, (len(arc), est)
    assert arc == oarc


def test_writer_matches_reference_file(pyvc, oracle, golden_containers):
    """The v1 writer, driven in the same single-threaded order as the golden
    archives, writes byte-identical files to the reference writer."""
    for name in ("synth_70x80x96", "zeros_40x33x17"):
        m, b = golden_containers[name]
        vol = container_volume(oracle, m)
        nz, ny, nx = m["shape_zyx"]
        nlod = oracle.nlod(nx, ny, nz)
        lods = [vol]
        for _ in range(1, nlod):
            lods.append(oracle.downsample2x(lods[-1]))
        fd, path = tempfile.mkstemp(suffix=".vca")
        os.close(fd)
        try:
            w = pyvc.Writer(path, (nx, ny, nz), m["target_ratio"])
            for l in range(nlod):
                w.set_base_q(l, m["qs"][l])
            for lod, v in enumerate(lods):
                lz, ly, lx = v.shape
                for az in range((lz + 31) // 32):
                    for ay in range((ly + 31) // 32):
                        for ax in range((lx + 31) // 32):
                            at = np.zeros((32, 32, 32), np.uint8)
                            s = v[az * 32:az * 32 + 32, ay * 32:ay * 32 + 32, ax * 32:ax * 32 + 32]
                            at[:s.shape[0], :s.shape[1], :s.shape[2]] = s
                            assert w.append_atom(lod, az, ay, ax, at) == 0
            # read-back before close: same bytes the reader decodes after close
            st, rb = w.decode_atom(0, 0, 0, 0)
            assert st == 0
            assert w.decode_atom(0, 0, 0, (nx + 31) // 32)[0] == 2     # out of the atom grid
            w.close()
            got = open(path, "rb").read()
        finally:
            os.unlink(path)
        assert got == b["v1"], name
        o = oracle.open(got)
        assert np.array_equal(rb, o.decode_atom(0, 0, 0, 0)[1])


def test_writer_box_zero_and_readback(pyvc, oracle):
    fd, path = tempfile.mkstemp(suffix=".vca")
    os.close(fd)
    try:
        vol = oracle.synth(4, 128, 64, 64, 4)
        w = pyvc.Writer(path, (2048, 64, 64), 20.0)
        assert w.append_box(0, (992, 0, 0, 1056, 64, 64), vol[:, :, :64]) == 2      # q unset
        for l in range(8):
            w.set_base_q(l, 1.0)
        assert w.append_box(0, (992, 0, 0, 1056, 64, 64), vol[:, :, :64]) == 0
        assert w.mark_zero_region(0, 0, 0, 0) == 0
        assert w.mark_zero_atom(0, 1, 1, 40) == 0
        assert w.coverage(0, 0, 0, 31) == 2 and w.coverage(0, 0, 0, 32) == 2
        assert w.coverage(0, 1, 1, 40) == 1 and w.coverage(0, 0, 0, 0) == 0
        st, out8, anyp = w.decode_2x2x2(0, 0, 0, 30)
        assert st == 0 and anyp == 1
        pay = oracle.encode_atom(np.ascontiguousarray(vol[:32, :32, :32]), 1.0)
        assert np.array_equal(out8[1], oracle.decode_payload(pay, 1.0)[1])
        w.close()
        arc = open(path, "rb").read()
    finally:
        os.unlink(path)
    a = pyvc.Archive(arc)
    st, at = a.decode_atom(0, 31, 0, 0)
    assert st == 0 and np.array_equal(at.ravel(), oracle.decode_payload(pay, 1.0)[1])


def test_encode_device_resident_256(pyvc, oracle):
    """Encode from device memory (256^3), then decode LOD0 on device; both
    bit-identical to the oracle."""
    n = 256
    d = pyvc.dev_alloc(n ** 3)
    try:
        assert pyvc.synth_device(13, (n, n, n), 40, d) == 0
        vol = pyvc.d2h(d, n ** 3).reshape(n, n, n)
        st, arc = pyvc.encode_device(d, (n, n, n), 10.0)
    finally:
        pyvc.dev_free(d)
    assert st == 0
    ost, oarc = oracle.encode(vol, 10.0)
    assert ost == 0 and arc == oarc
    a = pyvc.Archive(arc)
    out = pyvc.dev_alloc(n ** 3)
    try:
        assert a.decode_lod_device(0, out, 1) == 0
        dec = pyvc.d2h(out, n ** 3).reshape(n, n, n)
    finally:
        pyvc.dev_free(out)
    assert np.array_equal(dec, oracle.open(arc).decode_lod(0, (n, n, n))[1])
