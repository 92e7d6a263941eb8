"""GPU: vcx_encode_batch, the batched multi-volume encode service.  Every
archive of a batch of volumes of uneven dims and mixed target ratios is
byte-identical to the oracle's vc_encode of that volume alone; invalid
entries get the status vc_encode gives them and leave the rest unaffected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RATIOS = (10.0, 20.0, 50.0, 100.0)


def say(*a):
    print(*a, flush=True)


def _volumes(oracle, seed, n, lo, hi):
    rng = np.random.default_rng(seed)
    vols, targets = [], []
    for k in range(n):
        nx, ny, nz = (int(v) for v in rng.integers(lo, hi + 1, 3))
        ns = max(1, (nx + ny + nz) // 20)
        vols.append(oracle.synth(1000 + k, nx, ny, nz, ns))
        targets.append(RATIOS[int(rng.integers(0, 4))])
    return vols, targets


def test_encode_batch_matches_oracle(pyvc, oracle):
    """18 volumes, one all-zero and one of a single voxel, dims not multiples
    of 32."""
    vols, targets = _volumes(oracle, 5, 16, 40, 110)
    vols += [np.zeros((45, 38, 70), np.uint8), oracle.synth(77, 1, 1, 1, 0)]
    targets += [10.0, 50.0]
    say("batch", len(vols), [v.shape for v in vols])
    rc, res = pyvc.encode_batch(vols, targets)
    assert rc == 0
    for k, (v, t) in enumerate(zip(vols, targets)):
        ost, oarc = oracle.encode(v, t)
        st, arc = res[k]
        assert st == ost == 0, k
        assert arc == oarc, (k, v.shape, t)


def test_encode_batch_small_group_and_errors(pyvc, oracle):
    """A group below the shared-round width, with invalid entries between
    valid ones: those get VC_ERR_RANGE, the call returns it, the rest encode."""
    vols, targets = _volumes(oracle, 9, 4, 33, 97)
    vols.insert(1, np.zeros((4, 4, 4), np.uint8)); targets.insert(1, 0.5)      # target below 1
    vols.insert(3, np.zeros((0, 4, 4), np.uint8)); targets.insert(3, 10.0)     # empty volume
    rc, res = pyvc.encode_batch(vols, targets)
    assert rc == 2
    assert res[1] == (2, None) and res[3] == (2, None)
    for k, (v, t) in enumerate(zip(vols, targets)):
        if k in (1, 3):
            continue
        st, arc = pyvc.encode(v, t)
        assert st == 0 and res[k] == (0, arc), k
    ost, oarc = oracle.encode(vols[0], targets[0])
    assert res[0][1] == oarc


def test_encode_batch_launch_capacities(pyvc, oracle):
    """The number of bisection levels evaluated per launch (and the split of
    a round over launches) follows the launch capacity; the archives do not.
    Capacities from 512 rows (one level per launch, rounds over several
    launches) up to the default."""
    vols, targets = _volumes(oracle, 21, 7, 60, 230)
    ref = []
    for v, t in zip(vols, targets):
        st, arc = pyvc.encode(v, t)
        assert st == 0
        ref.append(arc)
    try:
        for rows in (512, 1500, 4000, 9000, 0):
            pyvc.test_rc_rows(rows)
            say("rc rows", rows)
            rc, res = pyvc.encode_batch(vols, targets)
            assert rc == 0
            for k in range(len(vols)):
                assert res[k] == (0, ref[k]), (rows, k)
    finally:
        pyvc.test_rc_rows(0)
    ost, oarc = oracle.encode(vols[0], targets[0])
    assert oarc == ref[0]


def test_encode_batch_device_resident(pyvc):
    """Volumes already in device memory: each archive equals vcx_encode_device
    of that volume alone."""
    dims = [(130, 70, 90), (64, 200, 41), (97, 97, 97)]
    ptrs = []
    try:
        for k, d in enumerate(dims):
            p = pyvc.dev_alloc(d[0] * d[1] * d[2])
            ptrs.append(p)
            assert pyvc.synth_device(300 + k, d, 6, p) == 0
        rc, res = pyvc.encode_batch(ptrs, [10.0, 100.0, 20.0], on_device=True, dims=dims)
        assert rc == 0
        for k, d in enumerate(dims):
            st, arc = pyvc.encode_device(ptrs[k], d, [10.0, 100.0, 20.0][k])
            assert st == 0 and res[k] == (0, arc), k
    finally:
        for p in ptrs:
            pyvc.dev_free(p)
