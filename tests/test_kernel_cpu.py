"""The device codec code (vcg_codec.hpp, vcg_xform.hpp) compiled for the CPU
and run phase by phase with the kernels' thread layout, against the
reference's golden payloads/reconstructions and the oracle.  Catches kernel
logic errors (indexing, strides, pruning, overflow paths) without a GPU."""
import ctypes as C
import hashlib
import os
import subprocess
import time

import numpy as np
import pytest

from conftest import ROOT, payloads_of

KDIR = os.path.join(ROOT, "tests", "kernel_cpu")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def kc():
    subprocess.check_call(["make", "-s", "-C", KDIR])
    L = C.CDLL(os.path.join(KDIR, "libkernel_cpu.so"))
    vp = C.c_void_p
    L.kc_encode.restype = C.c_uint32
    L.kc_encode.argtypes = [vp, C.c_uint32, vp, vp, C.c_uint32, C.POINTER(C.c_int)]
    L.kc_encode_fsm.restype = C.c_uint32
    L.kc_encode_fsm.argtypes = [vp, C.c_uint32, vp, vp, C.c_uint32, C.POINTER(C.c_int), vp, C.c_uint32,
                                C.POINTER(C.c_int), C.c_int]
    L.kc_decode.argtypes = [vp, C.c_uint32, vp, vp, C.POINTER(C.c_uint32)]
    L.kc_decode_fsm.argtypes = [vp, C.c_uint32, C.c_uint32, vp, vp, C.POINTER(C.c_uint32), C.c_int]
    L.kc_fdct.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                          vp, vp, vp, C.c_float, vp, vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.kc_idct.argtypes = [vp, C.c_uint32, C.c_int, vp, vp, vp, vp, C.c_int, vp]
    L.kc_idct_m.argtypes = [vp, C.c_uint32, C.c_int, vp, vp, vp, vp, vp]
    L.kc_quant.argtypes = [vp, C.c_size_t, C.c_float, C.c_float, vp]
    L.kc_decode_wave.argtypes = [vp, vp, vp, C.c_int, vp, vp, vp, vp, C.c_int]
    return L


@pytest.fixture(scope="module")
def tabs(oracle):
    cm, gs, scan, band = oracle.tables()
    bs = np.zeros(9, np.uint32)
    for b in range(1, 8):
        bs[b] = np.argmax(band >= b)
    bs[8] = 32768
    return cm, gs, scan, bs


def k_fdct(kc, tabs, vol, ax, ay, az, q):
    cm, gs, scan, bs = tabs
    vol = np.ascontiguousarray(vol, np.uint8)
    nz, ny, nx = vol.shape
    lv = np.zeros(32768, np.int16)
    co = np.zeros(32768, np.int32)
    eob, dc = C.c_uint32(), C.c_uint32()
    z = kc.kc_fdct(vol.ctypes.data, nx, ny, nz, ax, ay, az, cm.ctypes.data, scan.ctypes.data, gs.ctypes.data,
                   C.c_float(q), lv.ctypes.data, co.ctypes.data, C.byref(eob), C.byref(dc))
    return z, lv, co, eob.value, dc.value


def k_payload(kc, tabs, atom, q):
    cm, gs, scan, bs = tabs
    z, lv, co, eob, dc = k_fdct(kc, tabs, atom, 0, 0, 0, q)
    if z:
        return b""
    out = np.zeros(32768 * 4 + 64, np.uint8)
    ovf = C.c_int()
    n = kc.kc_encode(lv.ctypes.data, eob, bs.ctypes.data, out.ctypes.data, out.size, C.byref(ovf))
    assert not ovf.value
    return bytes([dc]) + out[:n].tobytes()


def k_decode(kc, tabs, pay, q, oracle, force=0):
    cm, gs, scan, bs = tabs
    rc = np.frombuffer(pay[1:], np.uint8) if len(pay) > 1 else np.zeros(1, np.uint8)
    ent = np.zeros(32768, np.uint32)
    cnt = C.c_uint32()
    st = kc.kc_decode(rc.ctypes.data, len(pay) - 1, bs.ctypes.data, ent.ctypes.data, C.byref(cnt))
    if st:
        return st, np.zeros(32768, np.uint8), 0
    stp, bia = oracle.dq_tables(q)
    out = np.zeros(32768, np.uint8)
    path = kc.kc_idct(ent.ctypes.data, cnt.value, pay[0], stp.ctypes.data, bia.ctypes.data, cm.ctypes.data,
                      scan.ctypes.data, force, out.ctypes.data)
    return 0, out, path


def test_fdct_phases_match_oracle(kc, tabs, oracle, golden_atoms):
    atoms = golden_atoms["atoms"]
    for i in range(len(atoms)):
        z, lv, co, eob, dc = k_fdct(kc, tabs, atoms[i], 0, 0, 0, 1.0)
        s = int(atoms[i].astype(np.int64).sum())
        if s == 0:
            assert z == 1
            continue
        odc = (s + 16384) // 32768
        assert dc == odc
        assert np.array_equal(co, oracle.fdct(atoms[i], odc)), i


def test_fdct_gather_edges(kc, tabs, oracle):
    vol = oracle.synth(5, 70, 45, 33, 3)
    for (ax, ay, az) in [(0, 0, 0), (2, 1, 1), (1, 0, 1)]:
        z, lv, co, eob, dc = k_fdct(kc, tabs, vol, ax, ay, az, 2.0)
        at = np.zeros((32, 32, 32), np.uint8)
        s = vol[az * 32:az * 32 + 32, ay * 32:ay * 32 + 32, ax * 32:ax * 32 + 32]
        at[:s.shape[0], :s.shape[1], :s.shape[2]] = s
        odc = (int(at.astype(np.int64).sum()) + 16384) // 32768
        assert np.array_equal(co, oracle.fdct(at, odc))


def test_kernel_payloads_match_reference(kc, tabs, golden_atoms):
    atoms, qs = golden_atoms["atoms"], golden_atoms["qs"]
    t0 = time.time()
    for ai, qi, pay, _ in payloads_of(golden_atoms):
        assert k_payload(kc, tabs, atoms[ai], float(qs[qi])) == pay, (ai, qi)
    assert time.time() - t0 < 120


def test_kernel_decode_matches_reference(kc, tabs, oracle, golden_atoms):
    qs = golden_atoms["qs"]
    for ai, qi, pay, rsha in payloads_of(golden_atoms):
        if not pay:
            continue
        st, out, path = k_decode(kc, tabs, pay, float(qs[qi]), oracle)
        assert st == 0 and sha(out) == rsha, (ai, qi, path)


def test_kernel_decode_paths_agree(kc, tabs, oracle, golden_atoms):
    """The 24-bit path is taken only under the L1 bound; both paths agree there."""
    qs = golden_atoms["qs"]
    for ai, qi, pay, rsha in payloads_of(golden_atoms):
        if not pay:
            continue
        st, a, path = k_decode(kc, tabs, pay, float(qs[qi]), oracle, 0)
        st2, b, _ = k_decode(kc, tabs, pay, float(qs[qi]), oracle, 2)
        assert np.array_equal(a, b)


def _idct_both(kc, tabs, ent, dc, stp, bia):
    cm, gs, scan, bs = tabs
    ent = np.ascontiguousarray(ent, np.uint32)
    a = np.zeros(32768, np.uint8)
    b = np.zeros(32768, np.uint8)
    path = kc.kc_idct_m(ent.ctypes.data, len(ent), dc, stp.ctypes.data, bia.ctypes.data, cm.ctypes.data,
                        scan.ctypes.data, a.ctypes.data)
    kc.kc_idct(ent.ctypes.data, len(ent), dc, stp.ctypes.data, bia.ctypes.data, cm.ctypes.data, scan.ctypes.data,
               2, b.ctypes.data)
    return path, a, b


def test_kernel_idct_mfma_matches_reference(kc, tabs, oracle, golden_atoms):
    """The matrix-core inverse (int8 products over an i16 tile, emulated lane
    by lane) reproduces the reference's reconstructions of every golden atom,
    all of them within its coefficient bound."""
    cm, gs, scan, bs = tabs
    qs = golden_atoms["qs"]
    for ai, qi, pay, rsha in payloads_of(golden_atoms):
        if not pay:
            continue
        rc = np.frombuffer(pay[1:], np.uint8) if len(pay) > 1 else np.zeros(1, np.uint8)
        ent = np.zeros(32768, np.uint32)
        cnt = C.c_uint32()
        assert kc.kc_decode(rc.ctypes.data, len(pay) - 1, bs.ctypes.data, ent.ctypes.data, C.byref(cnt)) == 0
        stp, bia = oracle.dq_tables(float(qs[qi]))
        path, a, b = _idct_both(kc, tabs, ent[:cnt.value], pay[0], stp, bia)
        assert path == 3, (ai, qi)
        assert sha(a) == rsha and np.array_equal(a, b), (ai, qi)


def _dequant(ent, scan, stp, bia):
    """r per entry as the scatter computes it (src/vc/vc.c:327-338, i32 wrap)"""
    e = ent.astype(np.int64)
    t = (e + 11) & 0x1FFFF
    m = (t >> 1) & 0xFFFF
    lv = np.where(t & 1, -m, m).astype(np.int16).astype(np.int64)
    idx = scan[(e >> 17)].astype(np.int64)
    f = (idx >> 10) + ((idx >> 5) & 31) + (idx & 31)
    mag = np.abs(lv)
    r = ((mag * stp[f] + bia[f] + 128) & 0xFFFFFFFF).astype(np.uint32).view(np.int32) >> 8
    return np.where(lv < 0, -r, r)


def test_kernel_idct_mfma_bound_and_fallback(kc, tabs, oracle):
    """Random level fields (sparse, dense, low-frequency only) scaled to
    coefficient L2 norms up to and past the bound: inside it the matrix-core
    path equals the i32 path byte for byte, past it the atom takes the i32
    path; extreme levels (|lv| = 32767) included."""
    cm, gs, scan, bs = tabs
    rng = np.random.default_rng(17)
    stp, bia = oracle.dq_tables(1.7)
    seen = {3: 0, 2: 0, 1: 0}
    for case in range(60):
        kind = case % 4
        nmax = {0: 40, 1: 4000, 2: 30000, 3: 300}[kind]
        hi = 32768 if kind != 3 else 700           # kind 3: low frequencies only (pass pruning)
        pos = np.unique(rng.integers(0, hi, int(rng.integers(1, nmax))))
        mag = rng.integers(1, 40, len(pos)).astype(np.int64)
        neg = rng.integers(0, 2, len(pos))
        target = [500.0, 12000.0, 29000.0, 30900.0, 30999.0, 31001.0, 33000.0, 1e6][case % 8]
        ent = ((pos.astype(np.int64) << 17) | ((2 * mag - 11 + neg) & 0x1FFFF)).astype(np.uint32)
        l2 = np.sqrt((_dequant(ent, scan, stp, bia).astype(np.float64) ** 2).sum())
        mag = np.clip(np.round(mag * target / max(l2, 1.0)), 1, 32767).astype(np.int64)
        if case % 11 == 5:
            mag[0] = 32767
        ent = ((pos.astype(np.int64) << 17) | ((2 * mag - 11 + neg) & 0x1FFFF)).astype(np.uint32)
        r = _dequant(ent, scan, stp, bia).astype(np.int64)
        inside = np.abs(r).max() <= 31000 and (r * r).sum() <= 31000 ** 2
        dc = int(rng.integers(0, 256))
        path, a, b = _idct_both(kc, tabs, ent, dc, stp, bia)
        assert (path == 3) == inside, (case, path, np.sqrt((r * r).sum()))
        assert np.array_equal(a, b), (case, path)
        seen[path] += 1
    assert seen[3] >= 20 and seen[2] + seen[1] >= 10, seen


def test_kernel_decode_damaged(kc, tabs, oracle, golden_corrupt):
    g = golden_corrupt
    blob = g["blob"].tobytes()
    for k in range(len(g["lens"])):
        p = blob[int(g["offs"][k]):int(g["offs"][k]) + int(g["lens"][k])]
        st, out, _ = k_decode(kc, tabs, p, float(g["qs"][k]), oracle)
        assert st == 0 and sha(out) == str(g["rec_sha"][k])


def test_kernel_eg_cap(kc, tabs, oracle):
    from test_oracle import eg_prefix_stream
    st, out, _ = k_decode(kc, tabs, bytes([9]) + eg_prefix_stream(70), 1.0, oracle)
    assert st == 1
    st, out, _ = k_decode(kc, tabs, bytes([9]) + eg_prefix_stream(20), 1.0, oracle)
    assert st == 0


def test_kernel_quantizer_exhaustive(kc, tabs, oracle):
    cm, gs, scan, bs = tabs
    c = np.arange(-65535, 65536, dtype=np.int32)
    out = np.zeros(c.size, np.int32)
    for q in (0.05, 0.77, 13.3, 4096.0):
        for f in range(94):
            kc.kc_quant(c.ctypes.data, c.size, C.c_float(q), C.c_float(float(gs[f])), out.ctypes.data)
            assert np.array_equal(out, oracle.quant_levels(q, f, c)), (q, f)


def fsm_decode(kc, tabs, rc, mis, check=0):
    cm, gs, scan, bs = tabs
    b = np.frombuffer(rc, np.uint8) if len(rc) else np.zeros(1, np.uint8)
    ent = np.zeros(32768 + 64, np.uint32)         # the device list stride: flushes write up to 21 placeholders past the end
    cnt = C.c_uint32()
    st = kc.kc_decode_fsm(b.ctypes.data, len(rc), mis, bs.ctypes.data, ent.ctypes.data, C.byref(cnt), check)
    return st, ent[:cnt.value]


def entry_levels(ent):
    """Scan positions and int16 levels of list entries (position << 17 |
    (2 m - 11 + sign) mod 2^17 for magnitude m)."""
    ent = ent.astype(np.int64)
    t = ((ent & 0x1FFFF) + 11) & 0x1FFFF
    m = t >> 1
    lv = np.where(t & 1, -m, m)
    return ent >> 17, ((lv + 0x8000) & 0xFFFF) - 0x8000


def loop_decode(kc, tabs, rc):
    cm, gs, scan, bs = tabs
    b = np.frombuffer(rc, np.uint8) if len(rc) else np.zeros(1, np.uint8)
    ent = np.zeros(32768, np.uint32)
    cnt = C.c_uint32()
    st = kc.kc_decode(b.ctypes.data, len(rc), bs.ctypes.data, ent.ctypes.data, C.byref(cnt))
    return st, ent[:cnt.value]


def test_fsm_decoder_matches(kc, tabs, golden_atoms, golden_corrupt):
    """The kernels' state-machine decoder gives the same entries as the
    straight-line decoder at every byte alignment, on reference payloads,
    damaged payloads and over-long Exp-Golomb prefixes."""
    from test_oracle import eg_prefix_stream
    streams = [p[1:] for (_, _, p, _) in payloads_of(golden_atoms) if p]
    g = golden_corrupt
    blob = g["blob"].tobytes()
    streams += [blob[int(g["offs"][k]) + 1:int(g["offs"][k]) + int(g["lens"][k])] for k in range(len(g["lens"]))]
    streams += [eg_prefix_stream(70), eg_prefix_stream(20), b"", b"\x07", b"\xff" * 9]
    for rc in streams:
        ref = loop_decode(kc, tabs, rc)
        for mis in range(8):
            got = fsm_decode(kc, tabs, rc, mis, mis & 1)
            assert got[0] == ref[0] and np.array_equal(got[1], ref[1]), (len(rc), mis)


def wave_decode(kc, tabs, pool, offs, lens, variant=0):
    """k_entropy's loop over one wavefront, lanes in lockstep (kc_decode_wave):
    per lane (bad, entries)"""
    cm, gs, scan, bs = tabs
    nl = len(offs)
    buf = np.concatenate([np.frombuffer(pool, np.uint8), np.zeros(64, np.uint8)])   # the device copy's padding
    o, n = np.array(offs, np.uint32), np.array(lens, np.uint32)
    ent = np.zeros(nl * (32768 + 64), np.uint32)
    cnt, bad = np.zeros(nl, np.uint32), np.zeros(nl, np.int32)
    assert kc.kc_decode_wave(buf.ctypes.data, o.ctypes.data, n.ctypes.data, nl, bs.ctypes.data, ent.ctypes.data,
                             cnt.ctypes.data, bad.ctypes.data, variant) == 0
    return [(int(bad[k]), ent[k * (32768 + 64):k * (32768 + 64) + int(cnt[k])]) for k in range(nl)]


def test_fsm_wave_lockstep(kc, tabs, golden_atoms, golden_corrupt):
    """Lanes of a wavefront in lockstep over a pool of back-to-back payloads,
    as k_entropy runs them: a parked lane keeps stepping until the whole wave
    is done, on the bytes that follow its stream (the next payload's).  Every
    lane's entries equal the straight-line decoder's.  The refill variant that
    skips the stream-end mask when no active lane is near its end (it passed
    the one-lane-at-a-time check, which stops a lane once it parks, and failed
    on the GPU) fails here: a parked lane appends the next payload's bytes to
    its empty window, its code turns nonzero and its zero bins start decoding
    coefficients again."""
    for g in (golden_atoms, golden_corrupt):
        pool = g["blob"].tobytes()
        recs = [(int(o), int(n)) for o, n in zip(g["offs"], g["lens"]) if int(n) > 1]
        ref = [loop_decode(kc, tabs, pool[o + 1:o + n]) for o, n in recs]
        planted_differs = False
        for w in range(0, len(recs), 64):
            offs = [o for o, _ in recs[w:w + 64]]
            lens = [n for _, n in recs[w:w + 64]]
            for variant in (0, 1):
                got = wave_decode(kc, tabs, pool, offs, lens, variant)
                for k, (bad, ent) in enumerate(got):
                    r = ref[w + k]
                    same = bool(bad) == bool(r[0]) and (bad or np.array_equal(ent, r[1]))
                    if variant == 0:
                        assert same, (w + k, bad, r[0])
                    elif not same:
                        planted_differs = True
        if g is golden_atoms:
            assert planted_differs, "the lockstep harness must catch the unmasked-refill variant"


def test_fsm_decoder_level_fields(kc, tabs):
    """Level fields that reach every bin state (Exp-Golomb magnitudes up to the
    int16 limit, EOB at band edges and at 32768) come back from the state
    machine, with and without the per-block edge test, as the straight-line
    decoder gives them and as they were encoded."""
    cm, gs, scan, bs = tabs
    rng = np.random.default_rng(11)
    for lv, eob in _level_fields(rng):
        n, ovf, rc, _ = _enc(kc, bs, lv, eob, 32768 * 4, False)
        assert not ovf
        ref = loop_decode(kc, tabs, rc)
        assert ref[0] == 0
        pos, lev = entry_levels(ref[1])
        nz = np.nonzero(lv[:eob])[0]
        assert np.array_equal(pos, nz) and np.array_equal(lev, lv[nz].astype(np.int64)), eob
        for check in (0, 1):
            got = fsm_decode(kc, tabs, rc, 3, check)
            assert got[0] == 0 and np.array_equal(got[1], ref[1]), (eob, check)


def _enc(kc, bs, lv, eob, cap, fsm, count=False, budget=None, check=0):
    out = np.zeros(max(cap, 1) + 64, np.uint8)
    ovf, stop = C.c_int(), C.c_int()
    if not fsm:
        n = kc.kc_encode(lv.ctypes.data, eob, bs.ctypes.data, out.ctypes.data, cap, C.byref(ovf))
        return n, ovf.value, out[:min(n, cap)].tobytes(), 0
    kb = np.zeros(1, np.uint32)
    n = kc.kc_encode_fsm(lv.ctypes.data, eob, bs.ctypes.data, None if count else out.ctypes.data, cap,
                         C.byref(ovf), kb.ctypes.data if budget is not None else None,
                         budget or 0, C.byref(stop), check)
    return n, ovf.value, (b"" if count else out[:min(n, cap)].tobytes()), stop.value


def _level_fields(rng):
    """Level fields that reach every bin state: sparse and dense fields, large
    magnitudes (Exp-Golomb up to the int16 limit), EOBs at 0, 1, band edges and
    32768, and long zero and long equal-level runs."""
    out = []
    for eob in (0, 1, 2, 363, 364, 365, 2600, 8376, 17153, 32767, 32768):
        lv = np.zeros(32768, np.int16)
        if eob:
            n = max(1, eob // 7)
            idx = rng.choice(eob, n, replace=False)
            lv[idx] = rng.integers(-40, 41, n)
            lv[eob - 1] = rng.choice([-1, 1, 2, -30000, 32767])
        out.append((lv, eob))
    for scale in (1, 3, 14, 300, 32767):
        lv = rng.integers(-scale, scale + 1, 32768).astype(np.int16)
        nz = np.nonzero(lv)[0]
        out.append((lv, int(nz[-1]) + 1 if nz.size else 0))
    lv = np.zeros(32768, np.int16)
    lv[::5] = 13                       # smallest Exp-Golomb value, repeated
    lv[::7] = -12
    out.append((lv, 32766))
    return out


def test_fsm_encoder_matches(kc, tabs, golden_atoms):
    """The kernels' state-machine encoder writes the same bytes, length and
    overflow flag as the straight-line encoder, and counts the same length,
    on the reference atoms' levels and on adversarial level fields, at full
    and at too-small capacities."""
    cm, gs, scan, bs = tabs
    atoms, qs = golden_atoms["atoms"], golden_atoms["qs"]
    cases = []
    for ai, qi, pay, _ in payloads_of(golden_atoms):
        z, lv, co, eob, dc = k_fdct(kc, tabs, atoms[ai], 0, 0, 0, float(qs[qi]))
        if not z:
            assert bytes([dc]) + _enc(kc, bs, lv, eob, 131136, True)[2] == pay, (ai, qi)
            cases.append((lv, eob))
    cases += _level_fields(np.random.default_rng(7))
    for k, (lv, eob) in enumerate(cases):
        ref = _enc(kc, bs, lv, eob, 131136, False)
        for check in (0, 1):
            got = _enc(kc, bs, lv, eob, 131136, True, check=check)
            assert got[:3] == ref[:3], (k, check)
            assert _enc(kc, bs, lv, eob, 131136, True, count=True, check=check)[:2] == ref[:2], (k, check)
        small = max(1, ref[0] // 2)
        assert _enc(kc, bs, lv, eob, small, True)[:3] == _enc(kc, bs, lv, eob, small, False)[:3], k
        assert _enc(kc, bs, lv, eob, small, True, count=True)[:2] == _enc(kc, bs, lv, eob, small, False)[:2], k


def test_fsm_encoder_budget(kc, tabs):
    """The counting encoder gives up only when the byte total has passed the
    budget, and otherwise returns the exact count."""
    cm, gs, scan, bs = tabs
    for lv, eob in _level_fields(np.random.default_rng(11)):
        full = _enc(kc, bs, lv, eob, 131136, False)[0]
        for budget in (0, 1023, 1024, 2047, full - 1, full, full + 5000):
            if budget < 0:
                continue
            n, ovf, _, stop = _enc(kc, bs, lv, eob, 131136, True, count=True, budget=budget)
            if stop:
                assert full > budget, (eob, budget, full)
            else:
                assert n == full, (eob, budget)
