"""GPU: many host threads decoding from one archive at once (the reference's
reentrant reader, src/vc/vc.h:46).  16 threads each decode whole LODs and
unaligned boxes into host memory; every result equals the oracle's, and the
device memory the library keeps afterwards stays within the shared decoder
scratch pool plus each thread's host staging, not 16 private scratch sets."""
import ctypes as C
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NT = 16


def _free_bytes():
    h = C.CDLL("libamdhip64.so")
    f, t = C.c_size_t(), C.c_size_t()
    assert h.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
    return f.value


def test_16_threads_full_lods_within_budget(pyvc, oracle):
    assert pyvc.device_init() == 0
    dims = (512, 512, 384)
    vol = pyvc.synth(8, dims, 60)
    st, arc = pyvc.encode(vol, 10.0)
    assert st == 0
    a = pyvc.Archive(arc)
    oa = oracle.open(arc)
    refs = {lod: oa.decode_lod(lod, a.lod_dims(lod)[1])[1] for lod in range(3)}
    assert a.decode_atom(0, 0, 0, 0)[0] == 0          # the device copy of the archive exists
    free0 = _free_bytes()
    errs = []

    def work(k):
        try:
            rng = np.random.default_rng(100 + k)
            for lod in (k % 3, (k + 1) % 3):
                st, got = a.decode_lod(lod)
                if st != 0 or not np.array_equal(got, refs[lod]):
                    errs.append((k, "lod", lod))
            for _ in range(3):
                lo = [int(rng.integers(0, d - 1)) for d in dims]
                hi = [int(rng.integers(l + 1, min(d, l + 200) + 1)) for l, d in zip(lo, dims)]
                st, got = a.decode_region(0, (lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]))
                ref = refs[0][lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
                if st != 0 or not np.array_equal(got, ref):
                    errs.append((k, "box", lo, hi))
        except Exception as e:                       # noqa: BLE001  (reported below)
            errs.append((k, repr(e)))

    ts = [threading.Thread(target=work, args=(k,)) for k in range(NT)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs, errs[:4]
    used = free0 - _free_bytes()
    atoms0 = 16 * 16 * 12                             # LOD0 atoms, the largest batch
    pool = 3 * atoms0 * (32832 * 4 + 16)              # shared decoder scratch
    staging = NT * (512 * 512 * 384 + 4 * 32768 * 16) # per-thread host staging and quads
    assert used <= pool + staging + (1 << 30), (used, pool, staging)
    a.close()
