"""Pins the CPU oracle (oracle/vc_oracle.c) before it is trusted as the
checker: closed forms, matrix-form transforms, exhaustive float quantizer,
entropy round trips, the reference's golden fixtures and (when built) the
reference library itself."""
import hashlib

import numpy as np
import pytest

from conftest import container_volume, payloads_of


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def freq_of(idx):
    idx = np.asarray(idx, np.int64)
    return (idx >> 10) + ((idx >> 5) & 31) + (idx & 31)


# ---------------------------------------------------------------- tables
def test_dct_basis_closed_form(oracle):
    cm, _, _, _ = oracle.tables()
    k = np.arange(32)[:, None]
    n = np.arange(32)[None, :]
    ck = np.where(k == 0, np.sqrt(1 / 32), np.sqrt(2 / 32))
    exact = ck * np.cos(np.pi * (2 * n + 1) * k / 64) * 16384
    assert np.all(np.abs(cm - exact) <= 0.5)
    assert cm[0, 0] == 2896 and np.abs(cm).max() == 4091
    assert cm[0].sum() == 92672 and np.all(cm[1:].sum(axis=1) == 0)
    # symmetry used by the partial butterfly: C[k][31-n] = (-1)^k C[k][n]
    sign = np.where(np.arange(32) % 2 == 0, 1, -1)[:, None]
    assert np.array_equal(cm[:, ::-1], sign * cm)
    g = cm.astype(np.int64) @ cm.T.astype(np.int64)
    assert np.all(np.abs(g - np.diag(np.diag(g))) < 2 ** 28 * 2e-3)
    assert np.all(np.abs(np.diag(g) / 2 ** 28 - 1) < 2e-3)


def test_gstep_scan_band(oracle):
    _, gs, scan, band = oracle.tables()
    f = np.arange(94)
    ref = np.power(1.0 + f, float(np.float32(0.65)))     # powf(1+f, 0.65f)
    assert gs[0] == 1.0 and np.all(np.diff(gs) > 0)
    assert np.all(np.abs(gs.astype(np.float64) - ref) <= np.spacing(gs))   # within 1 ulp
    assert np.array_equal(np.sort(scan), np.arange(32768))
    fr = freq_of(scan)
    assert np.all(np.diff(fr) >= 0)
    same = np.diff(fr) == 0
    assert np.all(np.diff(scan.astype(np.int64))[same] > 0)     # ties by raster index
    assert np.array_equal(band, np.minimum(7, fr * 8 // 96))


def test_dq_tables(oracle):
    _, gs, _, _ = oracle.tables()
    for q in (0.05, 1.0, 3.7, 4096.0):
        st, bi = oracle.dq_tables(q)
        step = np.float32(q) * gs
        assert np.array_equal(st, np.rint(step * np.float32(256)).astype(np.int32))
        c = np.float32(np.float32(0.4) - np.float32(1.0)) + np.float32(0.8)
        assert np.array_equal(bi, np.rint((c * step) * np.float32(256)).astype(np.int32))


# ---------------------------------------------------------------- transforms
def _pass(v, cm, axis, inverse):
    v = np.moveaxis(v.astype(np.int64), axis, -1)
    m = cm.T if inverse else cm          # forward: out[k] = sum_n C[k][n] v[n]
    out = (v @ m.T.astype(np.int64) + 8192) >> 14
    return np.moveaxis(out, -1, axis)


def test_fdct_equals_matrix_form(oracle):
    cm, _, _, _ = oracle.tables()
    rng = np.random.default_rng(0)
    for t in range(6):
        vox = rng.integers(0, 256, (32, 32, 32), dtype=np.uint8) if t < 3 else \
            (rng.integers(0, 2, (32, 32, 32)) * 255).astype(np.uint8)
        dc = int((int(vox.sum()) + 16384) // 32768)
        v = vox.astype(np.int64) - dc
        for ax in (2, 1, 0):               # x, then y, then z
            v = _pass(v, cm, ax, False)
        assert np.array_equal(oracle.fdct(vox, dc).reshape(32, 32, 32), v)


def test_idct_equals_matrix_form(oracle):
    cm, _, _, _ = oracle.tables()
    rng = np.random.default_rng(1)
    for t in range(6):
        c = np.zeros((32, 32, 32), np.int64)
        k = rng.integers(1, 3000)
        idx = rng.integers(0, 32768, k)
        c.ravel()[idx] = rng.integers(-4000, 4000, k)
        v = c
        for ax in (2, 1, 0):
            v = _pass(v, cm, ax, True)
        dc = int(rng.integers(0, 256))
        exp = np.clip(v + dc, 0, 255).astype(np.uint8)
        assert np.array_equal(oracle.idct(c.astype(np.int32), dc).reshape(32, 32, 32), exp)


def test_quantizer_exhaustive_float32(oracle):
    """Every integer coefficient -65535..65535 at every frequency: the oracle's
    quantizer equals strict float32 evaluation in source order."""
    _, gs, _, _ = oracle.tables()
    c = np.arange(-65535, 65536, dtype=np.int64)
    af = np.abs(c).astype(np.float32)
    for q in (0.05, 0.77, 1.0, 13.3, 4096.0):
        for f in range(94):
            got = oracle.quant_levels(q, f, c)
            step = np.float32(q) * gs[f]
            dz = np.float32(0.8) * step
            lev = np.where(af >= dz, ((af - dz) / step + np.float32(1.0)).astype(np.int64), 0)
            lev = np.minimum(lev, 32767) * np.sign(c)
            assert np.array_equal(got, lev), (q, f)


# ---------------------------------------------------------------- entropy coder
def test_entropy_roundtrip(oracle):
    rng = np.random.default_rng(2)
    _, _, scan, _ = oracle.tables()
    cases = []
    for density, mag in [(0.0, 1), (0.001, 3), (0.05, 20), (0.3, 200), (0.9, 4000), (1.0, 32767)]:
        lv = np.zeros(32768, np.int16)
        m = rng.random(32768) < density
        lv[m] = (rng.integers(1, mag + 1, m.sum()) * rng.choice([-1, 1], m.sum())).astype(np.int16)
        cases.append(lv)
    lv = np.zeros(32768, np.int16); lv[scan[32767]] = -32767; cases.append(lv)   # EOB = 32768
    lv = np.zeros(32768, np.int16); lv[scan[:12]] = np.arange(1, 13); cases.append(lv)  # unary edges
    for lv in cases:
        b = oracle.entropy_encode(lv)
        assert b is not None
        st, dec, eob = oracle.entropy_decode(b)
        assert st == 0 and np.array_equal(dec, lv)


class _TinyEncoder:
    """Minimal LZMA-style binary range encoder used only to build adversarial
    test streams (bins the codec's own encoder never emits)."""

    def __init__(self):
        self.low, self.range, self.cache, self.csize, self.out = 0, 0xFFFFFFFF, 0, 1, bytearray()

    def _shift(self):
        if (self.low >> 32) != 0 or self.low < 0xFF000000:
            carry = self.low >> 32
            while True:
                self.out.append((self.cache + carry) & 0xFF)
                self.cache = 0xFF
                self.csize -= 1
                if self.csize == 0:
                    break
            self.cache = (self.low >> 24) & 0xFF
        self.csize += 1
        self.low = (self.low << 8) & 0xFFFFFFFF

    def _norm(self):
        while self.range < (1 << 24):
            self._shift()
            self.range = (self.range << 8) & 0xFFFFFFFF

    def bit(self, ctx, i, b):
        r0 = (self.range * ctx[i]) >> 12
        if not b:
            self.range = r0
            ctx[i] += (4096 - ctx[i]) >> 5
        else:
            self.low += r0
            self.range -= r0
            ctx[i] -= ctx[i] >> 5
        self._norm()

    def bypass(self, b):
        self.range >>= 1
        if b:
            self.low += self.range
        self._norm()

    def flush(self):
        for _ in range(5):
            self._shift()
        return bytes(self.out)


def eg_prefix_stream(nbits):
    """EOB=1, coefficient 0 significant, 12 escape bins, then nbits one-bypasses and a zero."""
    e = _TinyEncoder()
    for i in range(15, -1, -1):
        e.bypass((1 >> i) & 1)
    sig, mag = [2048] * 8, [2048] * 12
    e.bit(sig, 0, 1)
    for k in range(12):
        e.bit(mag, k, 1)
    for _ in range(nbits):
        e.bypass(1)
    e.bypass(0)
    for _ in range(nbits):
        e.bypass(0)
    e.bypass(0)
    return e.flush()


def test_eg_prefix_cap(oracle):
    # an Exp-Golomb prefix longer than 64 bits is rejected (zeros + format error);
    # up to the cap the value wraps modulo 2^32 and the decode completes
    st, q, eob = oracle.entropy_decode(eg_prefix_stream(70))
    assert st == 1 and not q.any()
    st, q, eob = oracle.entropy_decode(eg_prefix_stream(20))
    assert st == 0 and eob == 1 and q[0] == ((11 + (1 << 20)) & 0xFFFF) - 0x10000 * (((11 + (1 << 20)) & 0xFFFF) >= 0x8000)


# ---------------------------------------------------------------- golden fixtures
def test_golden_payloads(oracle, golden_atoms):
    atoms = golden_atoms["atoms"]
    qs = golden_atoms["qs"]
    for ai, qi, pay, rsha in payloads_of(golden_atoms):
        assert oracle.encode_atom(atoms[ai], float(qs[qi])) == pay, (ai, qi)
        if pay:
            st, rec = oracle.decode_payload(pay, float(qs[qi]))
            assert st == 0 and sha(rec) == rsha, (ai, qi)
        else:
            assert rsha == sha(np.zeros(32768, np.uint8))


def test_golden_corrupt(oracle, golden_corrupt):
    g = golden_corrupt
    blob = g["blob"].tobytes()
    for k in range(len(g["lens"])):
        p = blob[int(g["offs"][k]):int(g["offs"][k]) + int(g["lens"][k])]
        st, rec = oracle.decode_payload(p, float(g["qs"][k]))
        assert st == 0 and sha(rec) == str(g["rec_sha"][k])


def test_golden_downsample(oracle):
    import json, os
    from conftest import GOLDEN
    meta = json.load(open(os.path.join(GOLDEN, "downsample.json")))
    rng = np.random.default_rng(3)
    for shp_s in meta:
        shp = tuple(int(x) for x in shp_s.split("x"))
        v = rng.integers(0, 256, shp, dtype=np.uint8)
        assert sha(v) == meta[shp_s]["in_sha"]
        assert sha(oracle.downsample2x(v)) == meta[shp_s]["out_sha"]


def test_golden_containers(oracle, golden_containers):
    for name, (m, b) in golden_containers.items():
        vol = container_volume(oracle, m)
        assert sha(vol) == m["vol_sha"], name
        st, ours = oracle.encode_q(vol, m["target_ratio"], m["qs"])
        assert st == 0 and ours == b["v2"], name
        a1, a2 = oracle.open(b["v1"]), oracle.open(b["v2"])
        nz, ny, nx = m["shape_zyx"]
        for lod in range(oracle.nlod(nx, ny, nz)):
            dx, dy, dz = nx, ny, nz
            for _ in range(lod):
                dx, dy, dz = (dx + 1) // 2, (dy + 1) // 2, (dz + 1) // 2
            s1, v1 = a1.decode_lod(lod, (dx, dy, dz))
            s2, v2 = a2.decode_lod(lod, (dx, dy, dz))
            assert s1 == s2 == 0 and np.array_equal(v1, v2), (name, lod)


def test_region_equals_atoms(oracle, golden_containers):
    m, b = golden_containers["synth_70x80x96"]
    a = oracle.open(b["v2"])
    rng = np.random.default_rng(4)
    for _ in range(10):
        x0, y0, z0 = rng.integers(0, 60), rng.integers(0, 50), rng.integers(0, 40)
        x1, y1, z1 = rng.integers(x0 + 1, 97), rng.integers(y0 + 1, 81), rng.integers(z0 + 1, 71)
        st, reg = a.decode_region(0, (x0, y0, z0, x1, y1, z1))
        assert st == 0
        for z in (z0, z1 - 1):
            for y in (y0, y1 - 1):
                for x in (x0, x1 - 1):
                    _, at = a.decode_atom(0, x // 32, y // 32, z // 32)
                    assert reg[z - z0, y - y0, x - x0] == at[z % 32, y % 32, x % 32]


def test_rate_control_bracket(oracle):
    """q0 meets the target on the calibration sample; (almost) any finer q does not."""
    vol = oracle.synth(11, 128, 128, 128, 20)
    for target in (10.0, 50.0):
        q0 = oracle.rate_control(vol, target)
        assert 0.05 <= q0 <= 4096
        # sample = all 64 LOD0 atoms for 128^3
        atoms = [vol[z:z + 32, y:y + 32, x:x + 32] for z in range(0, 128, 32)
                 for y in range(0, 128, 32) for x in range(0, 128, 32)]

        def ratio(q):
            b = sum(len(oracle.encode_atom(a, q)) for a in atoms)
            return 64 * 32768 / b

        assert ratio(q0) >= target
        assert ratio(q0 / 1.002) < target * 1.01


# ---------------------------------------------------------------- reference library
def test_oracle_matches_reference_strict_build(oracle, ref):
    """Payload bytes equal the reference compiled with strict IEEE float (the
    pinned quantizer order).  The reference's own -ffast-math flags give
    ISA-dependent levels at exact-integer boundaries (e.g. 14115.0001
    truncating to 14114 in the AVX2/SSE2 builds), so those builds are not the
    oracle."""
    rng = np.random.default_rng(9)
    vol = oracle.synth(21, 128, 96, 64, 12)
    atoms = [vol[z:z + 32, y:y + 32, x:x + 32] for z in range(0, 64, 32) for y in range(0, 96, 32)
             for x in range(0, 128, 32)]
    atoms += [rng.integers(0, 256, (32, 32, 32), dtype=np.uint8) for _ in range(2)]
    for q in (0.11, 0.9, 2.5, 17.0):
        assert ref.payloads(atoms, q) == [oracle.encode_atom(a, q) for a in atoms], q


def test_oracle_reader_matches_reference(oracle, ref, golden_containers):
    for name, (m, b) in golden_containers.items():
        for tag in ("v1", "v2"):
            ra, oa = ref.open(b[tag]), oracle.open(b[tag])
            nz, ny, nx = m["shape_zyx"]
            for az in range((nz + 31) // 32):
                for ay in range((ny + 31) // 32):
                    for ax in range((nx + 31) // 32):
                        c1 = ra.payload_range(0, az, ay, ax)
                        c2 = oa.coverage(0, az, ay, ax)
                        assert c1 == c2
                        s1, v1 = ra.decode_atom(0, ax, ay, az)
                        s2, v2 = oa.decode_atom(0, ax, ay, az)
                        assert s1 == s2 and np.array_equal(v1, v2)


# ---------------------------------------------------------------- container rewrite
def _v2_dir(b):
    """(directory entries [(off, len, flags)], total length field) of a v2 image"""
    import struct
    dir_off, cnt, tot = struct.unpack_from("<QQQ", b, 40)
    return [struct.unpack_from("<QII", b, dir_off + 16 * k) for k in range(cnt)], tot


def test_golden_repack_fixture(oracle, golden_repack):
    """The rewrite fixture is what its metadata says: several regions per LOD,
    a region of zero-marked atoms that becomes a ZERO entry, absent atoms, and
    v1 and v2 that decode alike."""
    for name, (m, b) in golden_repack.items():
        assert hashlib.sha256(b["v1"]).hexdigest() == m["v1_sha"]
        assert hashlib.sha256(b["v2"]).hexdigest() == m["v2_sha"]
        ents, tot = _v2_dir(b["v2"])
        assert tot == len(b["v2"]) == m["v2_len"]
        flags = [f for _, _, f in ents]
        assert flags.count(2) >= 2 and flags.count(1) >= 1, flags
        a1, a2 = oracle.open(b["v1"]), oracle.open(b["v2"])
        nz, ny, nx = m["shape_zyx"]
        seen = {0: 0, 1: 0, 2: 0}
        for az in range((nz + 31) // 32):
            for ay in range((ny + 31) // 32):
                for ax in range((nx + 31) // 32):
                    c1, c2 = a1.coverage(0, az, ay, ax), a2.coverage(0, az, ay, ax)
                    assert c1[0] == c2[0] and c1[2] == c2[2]
                    seen[c1[0]] += 1
                    s1, v1 = a1.decode_atom(0, ax, ay, az)
                    s2, v2 = a2.decode_atom(0, ax, ay, az)
                    assert s1 == s2 == 0 and np.array_equal(v1, v2)
        assert all(seen[k] > 0 for k in seen), seen


def test_oracle_repack_equals_reference_tool(oracle, golden_containers, golden_repack):
    """vco_repack reproduces tools/vc_repack.c's output on every golden v1 (and
    maps each canonical v2 to itself)."""
    cases = dict(golden_containers)
    cases.update(golden_repack)
    for name, (m, b) in cases.items():
        st, out = oracle.repack(b["v1"])
        assert st == 0 and out == b["v2"], name
        st, out = oracle.repack(b["v2"])
        assert st == 0 and out == b["v2"], name


def test_oracle_repack_rejects_payload_outside_archive(oracle, golden_repack):
    """A planted fault: one present slot's length made to run past the end."""
    import struct
    m, b = golden_repack["repack_multi"]
    buf = bytearray(b["v2"])
    ents, _ = _v2_dir(buf)
    off = next(o for o, _, f in ents if f == 2)
    for s in range(32768):
        so = off + 16 * s
        if buf[so + 12] == 1:
            po, = struct.unpack_from("<Q", buf, so)
            struct.pack_into("<I", buf, so + 8, len(buf) - po + 1)
            break
    st, out = oracle.repack(bytes(buf))
    assert st == 1 and out is None


# ---------------------------------------------------------------- exporter downscale
def test_oracle_ds2x_golden(oracle):
    """vco_ds2x against the reference's tools/downscale.h (strict build) on
    every golden case, and against the built reference variants directly."""
    import json
    import os
    import sys
    from conftest import GOLDEN
    sys.path.insert(0, GOLDEN)
    from make_golden import ds_input
    from pyoracle import ds_ref
    cases = json.load(open(os.path.join(GOLDEN, "ds.json")))
    refs = [r for r in (ds_ref("strict"), ds_ref("fast"), ds_ref("fast_v3")) if r]
    for c in cases:
        v = ds_input(tuple(c["shape_zyx"]), c["kind"], c["seed"])
        got = oracle.ds2x(v, c["method"], c["alpha"])
        assert sha(got) == c["sha"], c
        for r in refs:
            assert np.array_equal(got, r(v, c["method"], c["alpha"])), c


def test_fast_math_builds_vs_strict(oracle, golden_containers, capsys):
    """The reference built with its own flags (-ffast-math -ffp-contract=fast,
    CMakeLists.txt:46) at three ISA levels against the strict build that pins
    the quantizer's evaluation order: payloads of 64 atoms of the 128^3
    synthetic volume (20 sheets) at the LOD0..2 base_q its rate control picks
    for ratio 10, and at q = 5.  The strict build equals the oracle on every
    atom; every build decodes the golden payloads to the same voxels (the
    decoder has no float arithmetic); each fast build's agreement with strict
    is recorded (FMA contraction can move a level across a rounding edge)."""
    from pyoracle import Ref, host_has_avx512, ref_variants
    v = ref_variants()
    if "strict" not in v:
        pytest.skip("reference build oracle/_ref not present (make -f oracle/ref.mk)")
    vol = oracle.synth(1, 128, 128, 128, 20)
    q0 = oracle.rate_control(vol, 10.0)
    qs = [float(q) for q in oracle.lod_schedule(q0, 8)[:3]] + [5.0]
    rng = np.random.default_rng(3)
    atoms = []
    for _ in range(64):
        z, y, x = (int(t) for t in rng.integers(0, 4, 3))
        atoms.append(vol[z * 32:z * 32 + 32, y * 32:y * 32 + 32, x * 32:x * 32 + 32].copy())
    strict = Ref("strict")
    base = {q: strict.payloads(atoms, q) for q in qs}
    for q in qs:
        assert base[q] == [oracle.encode_atom(a, q) for a in atoms], q
    names = [n for n in ("fast_v4", "fast_v3", "base") if n in v and (n != "fast_v4" or host_has_avx512())]
    table = {}
    for n in names:
        r = Ref(n)
        table[n] = {f"{q:.4g}": sum(p == b for p, b in zip(r.payloads(atoms, q), base[q])) for q in qs}
        for name, (m, b) in golden_containers.items():
            ra, sa = r.open(b["v1"]), strict.open(b["v1"])
            for lod in range(3):
                st, dims = sa.lod_dims(lod)
                if st:
                    break
                box = (0, 0, 0) + tuple(dims)
                assert np.array_equal(ra.decode_region(lod, box)[1], sa.decode_region(lod, box)[1]), (n, name, lod)
            ra.close()
            sa.close()
        assert all(c >= 0.8 * len(atoms) for c in table[n].values()), (n, table[n])
    with capsys.disabled():
        print(f"\nfast-math payload agreement with strict (of {len(atoms)} atoms): {table}")
