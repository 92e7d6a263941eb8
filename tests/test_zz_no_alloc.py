"""The allocation-free batched atom decode (vcx_decode_atoms_noalloc) against
vcx_decode_atoms, the entry point it varies, and the CPU oracle: two calls in
a row on one caller-created stream, the second smaller, plus device-side
range checks."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so")
    h.hipStreamCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


def _quads(a, rng, n):
    nlod = 0
    while nlod < 8 and a.lod_dims(nlod)[0] == 0:
        nlod += 1
    q = []
    for _ in range(n):
        lod = 0 if rng.random() < 0.6 else int(rng.integers(1, nlod))
        _, (nx, ny, nz) = a.lod_dims(lod)
        q.append((lod, int(rng.integers(0, (nx + 31) // 32)), int(rng.integers(0, (ny + 31) // 32)),
                  int(rng.integers(0, (nz + 31) // 32))))
    return np.array(q, np.int32)


def test_noalloc_equals_decode_atoms(pyvc, oracle):
    assert pyvc.device_init() == 0
    L = pyvc.lib()
    dims = (160, 128, 96)
    vol = pyvc.synth(31, dims, 8)
    st, arc = pyvc.encode(vol, 10.0)
    assert st == 0
    a = pyvc.Archive(arc)
    oa = oracle.open(arc)
    n0 = np.zeros(1, np.int32)
    # no device copy yet: refused (creating it would allocate)
    assert L.vcx_decode_atoms_noalloc(a.h, 1, n0.ctypes.data, n0.ctypes.data, None, None) == pyvc.VC_ERR_RANGE
    assert a.decode_atom(0, 0, 0, 0)[0] == 0          # makes the device copy
    hip = _hip()
    stream = C.c_void_p()
    assert hip.hipStreamCreateWithFlags(C.byref(stream), 1) == 0
    rng = np.random.default_rng(5)
    cap = L.vcx_decode_atoms_noalloc
    dq = pyvc.dev_alloc(64 * 16)
    dout = pyvc.dev_alloc(64 * 32768)
    dst = pyvc.dev_alloc(64 * 4)
    try:
        assert cap(a.h, 65, dq, dout, dst, stream) == pyvc.VC_ERR_RANGE       # above the limit
        for n in (64, 23):                                                  # second call smaller
            q = _quads(a, rng, n)
            pyvc.h2d(dq, q.view(np.uint8))
            assert cap(a.h, n, dq, dout, dst, stream) == 0
            assert hip.hipStreamSynchronize(stream) == 0
            got = pyvc.d2h(dout, n * 32768).reshape(n, 32768)
            sts = pyvc.d2h(dst, n * 4).view(np.int32)
            rst, ref, per = a.decode_atoms(q)
            assert rst == 0 and not per.any() and not sts.any(), (rst, per, sts)
            assert np.array_equal(got, ref), n
            for k in range(0, n, 7):
                ost, oat = oa.decode_atom(*(int(v) for v in q[k]))
                assert ost == 0 and np.array_equal(got[k].reshape(32, 32, 32), oat), k
        # quads outside the archive: zeros and status 2, the others decoded
        q = _quads(a, rng, 6)
        q[1] = (8, 0, 0, 0)                 # lod beyond nlod
        q[3] = (0, 5, 0, 0)                 # ax beyond the LOD0 atom grid (5 atoms)
        q[4] = (1 | (3 << 8), 0, 0, 0)      # a placement tag is not a lod
        q[5] = (-1, 0, 0, 0)
        pyvc.h2d(dq, q.view(np.uint8))
        assert cap(a.h, 6, dq, dout, dst, stream) == 0
        assert hip.hipStreamSynchronize(stream) == 0
        got = pyvc.d2h(dout, 6 * 32768).reshape(6, 32768)
        sts = pyvc.d2h(dst, 6 * 4).view(np.int32)
        assert list(sts) == [0, 2, 0, 2, 2, 2], sts
        for k in (1, 3, 4, 5):
            assert not got[k].any()
        for k in (0, 2):
            rst, ref, _ = a.decode_atoms(q[k:k + 1])
            assert rst == 0 and np.array_equal(got[k], ref[0])
    finally:
        pyvc.dev_free(dq)
        pyvc.dev_free(dout)
        pyvc.dev_free(dst)
        hip.hipStreamDestroy(stream)
        a.close()
