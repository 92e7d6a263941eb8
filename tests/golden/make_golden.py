"""Generate the golden fixtures under tests/golden/ from the reference build.

Runs in the build container only (needs oracle/_ref, made by
`make -f oracle/ref.mk` from /root/reference).  The fixtures are data:
inputs plus the reference's outputs for them.

  atoms.npz       test atoms x base_q: payload bytes of the reference's
                  encode_atom (extracted from a v1 archive its writer made) and
                  SHA-256 of its reconstruction (vc_test_atom_rt)
  corrupt.npz     damaged payloads the reference decodes (terminating cases
                  only) and SHA-256 of what it decodes them to
  downsample.json SHA-256 of vc_test_downsample2x outputs on odd shapes
  containers.json + *.vca.gz
                  whole archives: v1 from the reference writer (single thread,
                  LOD-major raster append) and v2 from tools/vc_repack.c on it
  repack.json + repack_*.vca.gz
                  container-rewrite cases: a v1 archive of several regions per
                  LOD from the reference writer, atoms appended in a shuffled
                  order with some left absent, some marked zero and one region
                  all zero, and v2 from tools/vc_repack.c on it

  ds.json         SHA-256 of the exporter's within-cell downscale
                  (tools/downscale.h, box and cbox at several alphas) on
                  seeded inputs of odd and even shapes

usage: python tests/golden/make_golden.py [--only repack|ds]
"""
import gzip
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from pyoracle import Oracle, Ref  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
QS = [0.05, 0.3, 1.0, 3.7, 40.0, 4096.0]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_atoms(o):
    """Deterministic atom set (also rebuilt by the tests from this file's data)."""
    atoms = []
    vol = o.synth(5, 96, 64, 64, 8)
    for az in range(2):
        for ay in range(2):
            for ax in range(3):
                atoms.append(vol[az * 32:az * 32 + 32, ay * 32:ay * 32 + 32, ax * 32:ax * 32 + 32])
    rng = np.random.default_rng(1)
    atoms.append(rng.integers(0, 256, (32, 32, 32), dtype=np.uint8))               # noise
    atoms.append(np.full((32, 32, 32), 255, np.uint8))                             # flat 255
    a = np.zeros((32, 32, 32), np.uint8); a[5, 5, 5] = 200; atoms.append(a)        # single voxel
    a = np.zeros((32, 32, 32), np.uint8); a[5, 5, 5:13] = 200; atoms.append(a)     # 8-voxel line
    atoms.append(np.zeros((32, 32, 32), np.uint8))                                 # zero
    z, y, x = np.meshgrid(np.arange(32), np.arange(32), np.arange(32), indexing="ij")
    atoms.append(((x + y + z) * 255 // 93).astype(np.uint8))                       # ramp
    atoms.append((128 + 60 * np.sin(x * 0.06) + 40 * np.cos(y * 0.05) + 20 * np.sin(z * 0.07)).astype(np.uint8))
    atoms.append((((x + y + z) & 1) * 255).astype(np.uint8))                       # checker
    edge = np.zeros((32, 32, 32), np.uint8); edge[:7, :19, :3] = vol[:7, :19, :3]; atoms.append(edge)  # partial
    return np.stack(atoms)


def make_atoms(o, ref, ref_fast):
    atoms = test_atoms(o)
    blob, offs, lens, rec_sha, qidx, aidx = [], [], [], [], [], []
    pos = 0
    for qi, q in enumerate(QS):
        ours = [o.encode_atom(a, q) for a in atoms]
        theirs = ref.payloads(list(atoms), q)
        theirs_fast = ref_fast.payloads(list(atoms), q)
        for i, a in enumerate(atoms):
            assert ours[i] == theirs[i] == theirs_fast[i], (q, i)
            n, rec = ref.atom_rt(a, q)
            assert n == len(theirs[i])
            blob.append(theirs[i]); offs.append(pos); lens.append(len(theirs[i])); pos += len(theirs[i])
            rec_sha.append(sha(rec)); qidx.append(qi); aidx.append(i)
    np.savez_compressed(os.path.join(OUT, "atoms.npz"), atoms=atoms, qs=np.array(QS, np.float32),
                        blob=np.frombuffer(b"".join(blob), np.uint8), offs=np.array(offs, np.uint64),
                        lens=np.array(lens, np.uint32), qidx=np.array(qidx, np.int32),
                        aidx=np.array(aidx, np.int32), rec_sha=np.array(rec_sha))
    print("atoms.npz:", len(lens), "payloads,", pos, "bytes")


def make_corrupt(o, ref):
    """Damage valid payloads; keep the cases the reference terminates on."""
    atoms = test_atoms(o)
    rng = np.random.default_rng(7)
    pays, qs, shas = [], [], []
    base = [(atoms[i], q) for i in (0, 5, 12, 15) for q in (0.3, 3.7)]
    for vox, q in base:
        p = bytearray(o.encode_atom(vox, q))
        for trial in range(6):
            d = bytearray(p)
            kind = trial % 3
            if kind == 0:                                   # flip random bytes
                for _ in range(3):
                    d[1 + rng.integers(0, len(d) - 1)] ^= int(rng.integers(1, 256))
            elif kind == 1:                                 # truncate (past-end bytes read as 0)
                d = d[:max(2, int(len(d) * rng.uniform(0.2, 0.9)))]
            else:                                           # random tail
                cut = max(2, len(d) // 2)
                d = d[:cut] + bytearray(rng.integers(0, 256, len(d) - cut, dtype=np.uint8).tobytes())
            st, rec = o.decode_payload(bytes(d), q)
            if st != 0:
                continue                                    # the reference would not return
            # the reference decodes the same bytes through a v1 slot
            got = ref_decode_payload(ref, bytes(d), q)
            assert np.array_equal(got, rec), "oracle/reference disagree on a damaged payload"
            pays.append(bytes(d)); qs.append(q); shas.append(sha(rec))
    # EOB above 32768 is clamped (vc.c:549): all-ones payload start
    d = bytes([7]) + b"\xff\xff\xff\xff" + bytes(40)
    st, rec = o.decode_payload(d, 1.0)
    print("eob-clamp payload oracle status", st)
    lens = np.array([len(p) for p in pays], np.uint32)
    offs = np.concatenate([[0], np.cumsum(lens[:-1])]).astype(np.uint64) if pays else np.zeros(0, np.uint64)
    np.savez_compressed(os.path.join(OUT, "corrupt.npz"), blob=np.frombuffer(b"".join(pays), np.uint8),
                        offs=offs, lens=lens, qs=np.array(qs, np.float32), rec_sha=np.array(shas))
    print("corrupt.npz:", len(pays), "payloads")


def ref_decode_payload(ref, pay, q):
    """Decode raw payload bytes with the reference by planting them in a v1 archive."""
    import tempfile
    fd, path = tempfile.mkstemp(suffix=".vca")
    os.close(fd)
    try:
        one = np.zeros((32, 32, 32), np.uint8); one[0, 0, 0] = 1
        ref.write_v1(path, (32, 32, 32), 10.0, [q] * 8, [(0, 0, 0, 0, one)])
        buf = bytearray(open(path, "rb").read())
    finally:
        os.unlink(path)
    arc = ref.open(bytes(buf))
    c, off, ln = arc.payload_range(0, 0, 0, 0)
    arc.close()
    # append the damaged payload at the end and repoint the slot
    slot = None
    for so in range(0, len(buf) - 16):
        if int.from_bytes(buf[so:so + 8], "little") == off and int.from_bytes(buf[so + 8:so + 12], "little") == ln \
                and buf[so + 12] == 1:
            slot = so
            break
    assert slot is not None
    newoff = len(buf)
    buf += pay
    buf[slot:slot + 8] = newoff.to_bytes(8, "little")
    buf[slot + 8:slot + 12] = len(pay).to_bytes(4, "little")
    arc = ref.open(bytes(buf))
    st, out = arc.decode_atom(0, 0, 0, 0)
    arc.close()
    assert st == 0
    return out.ravel()


def make_downsample(o, ref):
    res = {}
    rng = np.random.default_rng(3)
    for shp in [(1, 1, 1), (2, 3, 5), (33, 17, 5), (64, 64, 64), (97, 131, 65), (70, 80, 96)]:
        v = rng.integers(0, 256, shp, dtype=np.uint8)
        r = ref.downsample2x(v)
        assert np.array_equal(r, o.downsample2x(v))
        res["x".join(map(str, shp))] = {"seed": 3, "out_sha": sha(r), "in_sha": sha(v)}
    json.dump(res, open(os.path.join(OUT, "downsample.json"), "w"), indent=1)
    print("downsample.json:", len(res))


CONTAINERS = [
    # name, (nz, ny, nx), synth seed, sheets, target_ratio, q0, zero-slab
    ("synth_70x80x96", (70, 80, 96), 3, 5, 10.0, 1.3, False),
    ("zeros_40x33x17", (40, 33, 17), 3, 3, 20.0, 0.7, True),
    ("one_1x1x1", (1, 1, 1), 4, 0, 10.0, 1.0, False),
    ("synth_64x64x64", (64, 64, 64), 9, 6, 50.0, 5.0, False),
]


def ref_v1_v2(o, ref, vol, ratio, qs):
    nz, ny, nx = vol.shape
    nlod = o.nlod(nx, ny, nz)
    lods = [vol]
    for _ in range(1, nlod):
        lods.append(o.downsample2x(lods[-1]))
    atoms = []
    for lod, v in enumerate(lods):
        lz, ly, lx = v.shape
        for az in range((lz + 31) // 32):
            for ay in range((ly + 31) // 32):
                for ax in range((lx + 31) // 32):
                    a = np.zeros((32, 32, 32), np.uint8)
                    s = v[az * 32:az * 32 + 32, ay * 32:ay * 32 + 32, ax * 32:ax * 32 + 32]
                    a[:s.shape[0], :s.shape[1], :s.shape[2]] = s
                    atoms.append((lod, az, ay, ax, a))
    p1, p2 = "/tmp/golden_v1.vca", "/tmp/golden_v2.vca"
    ref.write_v1(p1, (nx, ny, nz), ratio, list(qs[:nlod]), atoms)
    subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "vc_repack"), p1, p2], stdout=subprocess.DEVNULL)
    b1, b2 = open(p1, "rb").read(), open(p2, "rb").read()
    os.unlink(p1); os.unlink(p2)
    return b1, b2


def make_containers(o, ref):
    meta = {}
    for name, shp, seed, ns, ratio, q0, zslab in CONTAINERS:
        nz, ny, nx = shp
        vol = o.synth(seed, nx, ny, nz, ns)
        if zslab:
            vol[:32, :32, :] = 0
        qs = o.lod_schedule(q0, o.nlod(nx, ny, nz))
        b1, b2 = ref_v1_v2(o, ref, vol, ratio, qs)
        st, ours = o.encode_q(vol, ratio, qs)
        assert st == 0 and ours == b2, name
        for tag, b in (("v1", b1), ("v2", b2)):
            with gzip.open(os.path.join(OUT, f"{name}.{tag}.vca.gz"), "wb", 9) as f:
                f.write(b)
        meta[name] = {"shape_zyx": list(shp), "seed": seed, "sheets": ns, "target_ratio": ratio, "q0": q0,
                      "zero_slab": zslab, "qs": [float(x) for x in qs], "vol_sha": sha(vol),
                      "v1_sha": hashlib.sha256(b1).hexdigest(), "v2_sha": hashlib.sha256(b2).hexdigest(),
                      "v1_len": len(b1), "v2_len": len(b2)}
        print(name, len(b1), len(b2))
    json.dump(meta, open(os.path.join(OUT, "containers.json"), "w"), indent=1)


DS_SHAPES = [(2, 64, 96), (7, 9, 11), (1, 1, 1), (33, 17, 5), (2, 130, 66), (40, 41, 43)]
DS_ALPHAS = [0.0, 0.3, 0.5, 0.77, 1.0]


def ds_input(shape, kind, seed):
    """seeded downscale input: uniform bytes, 80% zeros, or sparse bright sheets on a dim background"""
    rng = np.random.default_rng(seed)
    v = rng.integers(0, 256, shape).astype(np.uint8)
    if kind == "sparse":
        v[rng.random(shape) < 0.8] = 0
    elif kind == "sheet":
        v = np.where(rng.random(shape) < 0.1, 200, 12).astype(np.uint8)
    return v


def make_ds():
    from pyoracle import ds_ref
    run = ds_ref("strict")
    cases = []
    for si, shp in enumerate(DS_SHAPES):
        for ki, kind in enumerate(("rand", "sparse", "sheet")):
            v = ds_input(shp, kind, 100 * si + ki)
            for m in (0, 1):
                for a in DS_ALPHAS:
                    cases.append({"shape_zyx": list(shp), "kind": kind, "seed": 100 * si + ki, "method": m,
                                  "alpha": a, "sha": sha(run(v, m, a))})
    json.dump(cases, open(os.path.join(OUT, "ds.json"), "w"), indent=0)
    print("ds cases", len(cases))


def make_repack(o, ref):
    meta = {}
    name, (nz, ny, nx), seed = "repack_multi", (40, 72, 1100), 31
    vol = o.synth(seed, nx, ny, nz, 6)
    nlod = o.nlod(nx, ny, nz)
    qs = list(o.lod_schedule(8.0, nlod))
    lods = [vol]
    for _ in range(1, nlod):
        lods.append(o.downsample2x(lods[-1]))
    rng = np.random.default_rng(seed)
    atoms, absent, zero = [], 0, 0
    for lod, v in enumerate(lods):
        lz, ly, lx = v.shape
        for az in range((lz + 31) // 32):
            for ay in range((ly + 31) // 32):
                for ax in range((lx + 31) // 32):
                    if lod == 0 and ax >= 32:          # LOD0's second x region: every atom marked zero
                        atoms.append((lod, az, ay, ax, None)); zero += 1
                        continue
                    u = rng.random()
                    if u < 0.1:
                        absent += 1
                        continue
                    if u < 0.2:
                        atoms.append((lod, az, ay, ax, None)); zero += 1
                        continue
                    a = np.zeros((32, 32, 32), np.uint8)
                    sl = v[az * 32:az * 32 + 32, ay * 32:ay * 32 + 32, ax * 32:ax * 32 + 32]
                    a[:sl.shape[0], :sl.shape[1], :sl.shape[2]] = sl
                    atoms.append((lod, az, ay, ax, a))
    order = rng.permutation(len(atoms))
    p1, p2 = "/tmp/golden_rp_v1.vca", "/tmp/golden_rp_v2.vca"
    ref.write_v1(p1, (nx, ny, nz), 10.0, qs, [atoms[i] for i in order])
    subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "vc_repack"), p1, p2], stdout=subprocess.DEVNULL)
    b1, b2 = open(p1, "rb").read(), open(p2, "rb").read()
    os.unlink(p1); os.unlink(p2)
    for tag, b in (("v1", b1), ("v2", b2)):
        with gzip.open(os.path.join(OUT, f"{name}.{tag}.vca.gz"), "wb", 9) as f:
            f.write(b)
    meta[name] = {"shape_zyx": [nz, ny, nx], "seed": seed, "nlod": nlod, "qs": [float(x) for x in qs],
                  "atoms_written": len(atoms), "absent": absent, "zero_marked": zero,
                  "v1_sha": hashlib.sha256(b1).hexdigest(), "v2_sha": hashlib.sha256(b2).hexdigest(),
                  "v1_len": len(b1), "v2_len": len(b2)}
    print(name, len(b1), len(b2), "atoms", len(atoms), "absent", absent, "zero", zero)
    json.dump(meta, open(os.path.join(OUT, "repack.json"), "w"), indent=1)


def main():
    o = Oracle()
    ref = Ref("strict")
    if sys.argv[1:] == ["--only", "repack"]:
        make_repack(o, ref)
        return
    if sys.argv[1:] == ["--only", "ds"]:
        make_ds()
        return
    ref_fast = Ref("fast_v4" if "fast_v4" in __import__("pyoracle").ref_variants() else "fast_v3")
    make_atoms(o, ref, ref_fast)
    make_corrupt(o, ref)
    make_downsample(o, ref)
    make_containers(o, ref)
    make_repack(o, ref)
    make_ds()


if __name__ == "__main__":
    main()
