"""CPU tests of libvc.so: it loads without a GPU, exports every function the
include/*.h headers declare, and its host-only reader functions (open,
validation, coverage, payload ranges, blob ranges, LOD dims) agree with the
reference and the oracle on the golden archives.  No codec call is made."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def declared_functions():
    names = set()
    for h in ("vc.h", "vc_gpu.h", "vc_testhooks.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(vc[x]?_[a-z0-9_]+)\s*\((?!\*)", txt):
            if not m.group(1).endswith("_fn"):
                names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol(pyvc):
    L = pyvc.lib()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_region_atoms(pyvc):
    assert pyvc.lib().vc_region_atoms() == 32


def test_open_and_index_queries_match_reference(pyvc, oracle, golden_containers):
    for name, (m, b) in golden_containers.items():
        nz, ny, nx = m["shape_zyx"]
        for tag in ("v1", "v2"):
            a = pyvc.Archive(b[tag])
            o = oracle.open(b[tag])
            assert a.version() == (1 if tag == "v1" else 2)
            nlod = oracle.nlod(nx, ny, nz)
            dx, dy, dz = nx, ny, nz
            for lod in range(nlod):
                st, d = a.lod_dims(lod)
                assert st == 0 and d == (dx, dy, dz)
                for az in range((dz + 31) // 32 + 1):
                    for ay in range((dy + 31) // 32 + 1):
                        for ax in range((dx + 31) // 32 + 1):
                            c, off, ln = a.payload_range(lod, az, ay, ax)
                            assert (c, off, ln) == o.coverage(lod, az, ay, ax)
                            assert a.coverage(lod, az, ay, ax) == c
                dx, dy, dz = (dx + 1) // 2, (dy + 1) // 2, (dz + 1) // 2
            assert a.lod_dims(nlod)[0] == 2 and a.lod_dims(-1)[0] == 2
            # v2 blob ranges: every DATA region's blob holds its slots + payloads
            if tag == "v2":
                c, off, ln = a.blob_range(0, 0, 0, 0)
                assert c == 2 and ln >= 32768 * 16 and off + ln <= len(b[tag])
                c, off, ln = a.blob_range(nlod - 1, 0, 0, 0)       # coarsest LOD first
                assert c == 2 and off % 4096 == 0 and off < 4096 + 128 + 16 * 64
            else:
                assert a.blob_range(0, 0, 0, 0)[0] == 0
            a.close()


def test_reference_coverage_on_golden(pyvc, ref, golden_containers):
    for name, (m, b) in golden_containers.items():
        nz, ny, nx = m["shape_zyx"]
        for tag in ("v1", "v2"):
            a, r = pyvc.Archive(b[tag]), ref.open(b[tag])
            for az in range((nz + 31) // 32):
                for ay in range((ny + 31) // 32):
                    for ax in range((nx + 31) // 32):
                        assert a.payload_range(0, az, ay, ax) == r.payload_range(0, az, ay, ax)


def test_header_rejection(pyvc, golden_containers):
    """src/vc/vc.c:1285-1314: bad magic, version, atom size, region size, LOD
    count, truncated buffers and a directory that does not cover every region."""
    m, b = golden_containers["synth_64x64x64"]
    for tag in ("v1", "v2"):
        good = bytearray(b[tag])
        pyvc.Archive(bytes(good)).close()
        for off, val in ((0, 0x00), (4, 7), (8, 16), (16, 31), (20, 0), (20, 9)):
            bad = bytearray(good)
            bad[off] = val if off else bad[0] ^ 0xFF
            with pytest.raises(ValueError):
                pyvc.Archive(bytes(bad))
        with pytest.raises(ValueError):
            pyvc.Archive(bytes(good[:100]))
    bad = bytearray(b["v2"])
    bad[48] ^= 1                                     # dir_count != regions
    with pytest.raises(ValueError):
        pyvc.Archive(bytes(bad))
    bad = bytearray(b["v1"])
    bad[40] = 3                                      # l1_cap not a power of two
    with pytest.raises(ValueError):
        pyvc.Archive(bytes(bad))


def test_streaming_open_index_queries(pyvc, oracle, golden_containers):
    m, b = golden_containers["synth_70x80x96"]
    for tag in ("v1", "v2"):
        s = pyvc.StreamingArchive(b[tag])
        o = oracle.open(b[tag])
        for az in range(3):
            for ay in range(3):
                for ax in range(4):
                    assert s.payload_range(0, az, ay, ax) == o.coverage(0, az, ay, ax)
        assert s.reads > 0
