"""Benchmark: batched vc_decode_lod of LOD0 of a 1024^3 u8 synthetic volume.

One step = one device-resident decode of every LOD0 atom (32768 atoms, one
entropy launch + one transform launch) from the archive already resident in
HBM into a device output buffer.  Steps are issued back to back through the
asynchronous device API, which overlaps the transform of one call with the
entropy pass of the next; the timed region brackets all K steps with HIP
events on the library's own stream.  The latency of one synchronous call is
reported beside it (single_call_ms).  Also reported: per-kernel durations
(stage events), the dominant kernel's roofline, plain per-call vc_decode_atom
latency p50/p99, a bit-exactness check against the CPU oracle on sampled
atoms, and the CPU baseline (the reference's vc_decode_atom under OpenMP on
the host cores).

Further sections, each beside its plain per-call counterpart:
  other_ratios   : the same volume at target ratios 50 and 100 (decode GB/s,
                   encode time, sampled atoms against the oracle);
  serving        : a 2048^3 ratio-50 archive (320 sheets) served through the
                   continuous-batching request loop under open-loop Poisson
                   traces of atom requests (1-4096 atoms, log-uniform, LOD 0
                   with p = 0.7) and of region requests (log-uniform edges
                   17-600 voxels, LOD 0-7), p50/p99 latency; and the same
                   archive rewritten into the canonical v2 layout in device
                   memory (vcx_repack_device) beside the reference's own
                   repack tool on the host; and opened with vc_open_streaming
                   (LOD0 from the region-blob cache, cold and warm, and LOD1
                   by per-atom reads);
  downscale      : the exporter's within-cell 2x downscale (cbox) of the
                   1024^3 volume on the device (HBM-bound);
  batched_encode : 256 device-resident volumes of uneven dims (axes uniform
                   in [40, 512]) at ratios {10, 20, 50, 100} in one
                   vcx_encode_batch call, against per-volume encodes.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 1024] [--ratio 10]
                       [--no-serve] [--no-ratios] [--no-encode-batch] [--no-cpu]
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "LOD0 decode GB/s + p50/p99 atom-request latency, batched vs plain, bit-exact"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_pyvc():
    spec = importlib.util.spec_from_file_location("pyvc", os.path.join(ROOT, "volume-compressor_amd", "pyvc.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


class Hip:
    """hipEvent helpers over libamdhip64 (events recorded on the library's stream)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.L.hipEventSynchronize.argtypes = [C.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.L.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, s):
        assert self.L.hipEventRecord(e, s) == 0

    def ms(self, a, b):
        self.L.hipEventSynchronize(b)
        t = C.c_float()
        assert self.L.hipEventElapsedTime(C.byref(t), a, b) == 0
        return t.value


def cpu_baseline(arc, atoms_zyx, budget_s=12.0):
    """The reference's own vc_decode_atom (oracle/_ref, dlopen) under OpenMP,
    or the oracle port when the reference build is absent, on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    threads = min(threads, os.cpu_count() or threads)
    arr = np.frombuffer(arc, np.uint8)
    v = pyoracle.ref_variants()
    pick = None
    if v:
        pick = v.get("fast_v4") if pyoracle.host_has_avx512() and "fast_v4" in v else v.get("fast_v3", v.get("strict"))
    rng = np.random.default_rng(17)
    if pick:
        L = C.CDLL(os.path.join(ROOT, "oracle", "libcpubench.so"))
        L.cpu_decode_atoms.restype = C.c_double
        L.cpu_decode_atoms.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_size_t,
                                       C.c_int, C.POINTER(C.c_long), C.c_void_p]

        def run(idx):
            z = np.ascontiguousarray(atoms_zyx[idx], np.int32)
            bad = C.c_long()
            s = L.cpu_decode_atoms(pick.encode(), arr.ctypes.data, arr.size, 0, z.ctypes.data, len(idx),
                                   threads, C.byref(bad), None)
            assert s > 0 and bad.value == 0, (s, bad.value)
            return s
        kind, label = "reference", os.path.basename(pick)
    else:
        o = pyoracle.Oracle()
        oa = o.open(arc)

        def run(idx):
            return oa.decode_atoms_omp(0, np.ascontiguousarray(atoms_zyx[idx], np.uint32), threads)[0]
        kind, label = "port", "oracle/liboracle.so"
    # probe with 256 atoms, then decode a random sample (whole LOD0 at most) as
    # many times as fills the budget
    n = 256
    t = run(rng.choice(len(atoms_zyx), n, replace=False))
    n = int(min(len(atoms_zyx), max(n, n * budget_s / max(t, 1e-3))))
    idx = rng.choice(len(atoms_zyx), n, replace=False)
    t1 = run(idx)
    reps = max(1, min(64, int(budget_s / max(t1, 1e-3))))
    t = t1 + sum(run(idx) for _ in range(reps - 1))
    gbs = reps * n * 32768 / t / 1e9
    return {"value": round(gbs, 4), "unit": "GB/s", "cores": threads, "kind": kind,
            "sample": f"{n} random LOD0 atoms of the same archive x {reps} passes "
                      f"({reps * n * 32768 / 2**20:.0f} MiB out), {t:.1f} s, {label}, OpenMP schedule(dynamic)"}


def _ref_lib():
    """(path of the reference build used as the CPU baseline, kind) or (None, None)"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    v = pyoracle.ref_variants()
    if not v:
        return None
    return v.get("fast_v4") if pyoracle.host_has_avx512() and "fast_v4" in v else v.get("fast_v3", v.get("strict"))


def _threads():
    t = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    return min(t, os.cpu_count() or t)


def cpu_serve_bench(arc, reqs, qps, seed=11):
    """The same open-loop trace (same requests, same Poisson gaps as
    serve_bench) served on the host cores by the reference's own
    vc_decode_atom / vc_decode_region (oracle/_ref): a pool of one worker per
    core takes whole requests in arrival order."""
    lib = _ref_lib()
    if not lib:
        return None
    L = C.CDLL(os.path.join(ROOT, "oracle", "libcpubench.so"))
    L.cpu_serve.restype = C.c_double
    L.cpu_serve.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.c_size_t] + [C.c_void_p] * 6 + \
        [C.c_int, C.c_void_p, C.POINTER(C.c_long)]
    n = len(reqs)
    gaps = np.random.default_rng(seed).exponential(1.0 / qps, n)
    t_arr = np.ascontiguousarray(2e-3 + np.cumsum(gaps))
    kind = np.array([0 if k == "atoms" else 1 for k, _ in reqs], np.int32)
    nat = np.array([len(x) if k == "atoms" else 0 for k, x in reqs], np.int32)
    qoff = np.concatenate([[0], np.cumsum(nat)[:-1]]).astype(np.int64)
    quads = np.ascontiguousarray(np.concatenate([x for k, x in reqs if k == "atoms"] or [np.zeros((1, 4), np.int32)]),
                                 np.int32)
    boxes = np.zeros((n, 7), np.int32)
    for r, (k, x) in enumerate(reqs):
        if k == "region":
            boxes[r] = (x[0],) + tuple(x[1])
    arr = np.frombuffer(arc, np.uint8)
    t_done = np.zeros(n)
    bad = C.c_long()
    th = _threads()
    span = L.cpu_serve(lib.encode(), arr.ctypes.data, arr.size, n, kind.ctypes.data, qoff.ctypes.data,
                       nat.ctypes.data, quads.ctypes.data, boxes.ctypes.data, t_arr.ctypes.data, th,
                       t_done.ctypes.data, C.byref(bad))
    assert span > 0 and bad.value == 0, (span, bad.value)
    lat = (t_done - t_arr) * 1e3
    out = {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
           "requests": n, "offered_qps": qps, "achieved_rps": round(float(n / (span - t_arr[0])), 1), "cores": th,
           "kind": "reference", "lib": os.path.basename(lib)}
    if kind[0] == 0:
        out["atoms_per_s"] = round(float(nat.sum() / (span - t_arr[0])), 1)
    return out


def cpu_encode_bench(vols, targets, budget_s=12.0):
    """vc_encode on the host cores: the oracle's OpenMP encoder (the reference
    ships no whole-volume encoder; this restates it from the reference's
    primitives), four volumes at a time, on as many of the volumes as fit the
    budget (at least one).  GB/s of LOD0 input."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from concurrent.futures import ThreadPoolExecutor
    from pyoracle import Oracle
    o = Oracle()
    t0 = time.perf_counter()
    done, nbytes = 0, 0
    with ThreadPoolExecutor(min(4, len(vols))) as ex:
        while done < len(vols) and (done == 0 or time.perf_counter() - t0 < budget_s):
            k = min(len(vols), done + 4)
            for st, _ in ex.map(lambda i: o.encode(vols[i], targets[i]), range(done, k)):
                assert st == 0
            nbytes += sum(int(v.size) for v in vols[done:k])
            done = k
    t = time.perf_counter() - t0
    return {"value": round(nbytes / t / 1e9, 4), "unit": "GB/s", "cores": _threads(), "kind": "port",
            "sample": f"{done} of {len(vols)} volumes, {nbytes / 1e9:.3f} GB, {t:.1f} s, oracle/liboracle.so "
                      "vco_encode (OpenMP), 4 volumes at a time"}


def lod_count(a):
    nlod = 0
    while nlod < 8 and a.lod_dims(nlod)[0] == 0:
        nlod += 1
    return nlod


def atom_requests(a, nreq, rng):
    """Requests of round(exp(U(ln 1, ln 4096))) atoms, LOD 0 with p = 0.7 else
    uniform over the coarser LODs, coordinates uniform over that LOD's atom grid."""
    nlod = lod_count(a)
    reqs = []
    for _ in range(nreq):
        n = int(round(np.exp(rng.uniform(0.0, np.log(4096.0)))))
        lod = 0 if rng.random() < 0.7 or nlod == 1 else int(rng.integers(1, nlod))
        _, (nx, ny, nz) = a.lod_dims(lod)
        q = np.stack([np.full(n, lod), rng.integers(0, (nx + 31) // 32, n), rng.integers(0, (ny + 31) // 32, n),
                      rng.integers(0, (nz + 31) // 32, n)], 1).astype(np.int32)
        reqs.append(("atoms", q))
    return reqs


def region_requests(a, nreq, rng):
    """Boxes with log-uniform edges in [17, 600] voxels per axis (independent,
    clipped to the LOD), offsets uniform (not 32-aligned), LOD uniform over
    every LOD of the archive."""
    nlod = lod_count(a)
    reqs = []
    for _ in range(nreq):
        lod = int(rng.integers(0, nlod))
        _, dims = a.lod_dims(lod)
        lo, hi = [], []
        for d in dims:
            e = min(d, int(round(np.exp(rng.uniform(np.log(17.0), np.log(600.0))))))
            o = int(rng.integers(0, d - e + 1))
            lo.append(o)
            hi.append(o + e)
        reqs.append(("region", (lod, (lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]))))
    return reqs


def req_bytes(r):
    kind, x = r
    if kind == "atoms":
        return len(x) * 32768
    _, (x0, y0, z0, x1, y1, z1) = x
    return (x1 - x0) * (y1 - y0) * (z1 - z0)


def tune_section(pyvc, a, device_bytes=16 << 30):
    """vcx_server_autotune on the serving archive: the launch-shape table
    (per LOD class and batch-size bucket, default and tuned ms of one
    captured batch of that size), then the tuning time."""
    srv = pyvc.Server(a, device_bytes)
    try:
        t0 = time.perf_counter()
        assert srv.autotune() == 0
        t = time.perf_counter() - t0
        rows = srv.tuning()
    finally:
        srv.close()
    table = [{"class": "lod0" if c == 0 else "coarse", "bucket": b, "ent_waves": w, "ent_groups": ge,
              "idct_groups": gt, "default_ms": round(md, 4), "tuned_ms": round(mt, 4),
              "gain": round(md / mt, 3) if mt > 0 else None} for c, b, w, ge, gt, md, mt in rows]
    return {"tune_s": round(t, 2), "table": table}


def serve_bench(pyvc, a, reqs, qps, seed=11, device_bytes=16 << 30, tune=True):
    """Open-loop Poisson trace of `reqs` at `qps` through the continuous-batching
    server.  Latency = completion (server clock) - scheduled arrival.  Three
    served requests are checked against the plain per-call path."""
    rng = np.random.default_rng(seed)
    nreq = len(reqs)
    gaps = rng.exponential(1.0 / qps, nreq)
    sizes = [req_bytes(r) for r in reqs]
    offs = np.cumsum([0] + sizes)
    arena = pyvc.PinnedBuffer(int(offs[-1]))
    srv = pyvc.Server(a, device_bytes)
    if tune:
        assert srv.autotune() == 0
    ids, arrive = {}, np.zeros(nreq)
    done_t, bad = {}, 0
    rec = {}                      # k -> (t_submit, t_admit, t_done, batch class, batch atoms)

    def take(got):
        nonlocal bad
        for g, s_, cls, ts, ta, td, ba in got:
            k = ids[int(g)]
            done_t[k] = int(td)
            rec[k] = (int(ts), int(ta), int(td), int(cls), int(ba))
            bad += int(s_ != 0)
    try:
        t0 = pyvc.now_ns() + 2_000_000
        sched = t0 + np.cumsum(gaps) * 1e9
        for k, (kind, x) in enumerate(reqs):
            while pyvc.now_ns() < sched[k]:
                pass
            if kind == "atoms":
                i = srv.submit_atoms(x, arena.p + int(offs[k]))
            else:
                i = srv.submit_region(x[0], x[1], arena.p + int(offs[k]))
            assert i >= 0, (kind, x if kind == "region" else len(x))
            ids[int(i)] = k
            arrive[k] = sched[k]
            take(srv.poll_served(wait_ms=0))
        while len(done_t) < nreq:
            take(srv.poll_served(wait_ms=100))
    finally:
        srv.close()
    lat = np.array([(done_t[k] - arrive[k]) / 1e6 for k in range(nreq)])
    span = (max(done_t.values()) - arrive[0]) / 1e9
    exact = True
    for k in (0, nreq // 2, nreq - 1):
        kind, x = reqs[k]
        base = int(offs[k])
        if kind == "atoms":
            for m in range(min(4, len(x))):
                lod, ax, ay, az = (int(v) for v in x[m])
                st, ref = a.decode_atom(lod, ax, ay, az)
                b = base + m * 32768
                exact &= st == 0 and np.array_equal(arena.a[b:b + 32768], ref.ravel())
        else:
            st, ref = a.decode_region(x[0], x[1])
            exact &= st == 0 and np.array_equal(arena.a[base:base + sizes[k]], ref.ravel())
    arena.free()
    out = {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
           "requests": nreq, "offered_qps": qps, "achieved_rps": round(nreq / span, 1),
           "out_GBps": round(sum(sizes) / span / 1e9, 3), "errors": bad, "spot_check_exact": bool(exact)}
    # latency by class: requests at LOD0 and at the coarser LODs
    lod = np.array([int(x[0, 0]) if kind == "atoms" else x[0] for kind, x in reqs])
    for name, m in (("lod0", lod == 0), ("coarse", lod != 0)):
        if m.any():
            out[f"{name}_p50"] = round(float(np.percentile(lat[m], 50)), 3)
            out[f"{name}_p99"] = round(float(np.percentile(lat[m], 99)), 3)
    # where the slowest 1% of requests spent their time: submission lag
    # (scheduled arrival to submit), queueing (submit to admission into a
    # batch) and service (admission to completion: the batch's launches and
    # result copy), with the class and size of the batch that completed them
    slow = np.argsort(lat)[-max(1, nreq // 100):]
    out["p99_breakdown_ms"] = {
        "requests": int(slow.size),
        "submit_lag": round(float(np.mean([(rec[k][0] - arrive[k]) / 1e6 for k in slow])), 3),
        "queue": round(float(np.mean([(rec[k][1] - rec[k][0]) / 1e6 for k in slow])), 3),
        "service": round(float(np.mean([(rec[k][2] - rec[k][1]) / 1e6 for k in slow])), 3),
        "coarse_batches": int(sum(rec[k][3] for k in slow)),
        "batch_atoms": int(np.mean([rec[k][4] for k in slow])),
    }
    out["queue_ms_p50"] = round(float(np.percentile([(rec[k][1] - rec[k][0]) / 1e6 for k in range(nreq)], 50)), 3)
    out["service_ms_p50"] = round(float(np.percentile([(rec[k][2] - rec[k][1]) / 1e6 for k in range(nreq)], 50)), 3)
    if reqs[0][0] == "atoms":
        atoms = sum(len(x) for _, x in reqs)
        out["atoms_per_s"] = round(atoms / span, 1)
        out["mean_atoms_per_request"] = round(atoms / nreq, 1)
    else:
        out["mean_MB_per_request"] = round(sum(sizes) / nreq / 1e6, 3)
    return out


def qps_sweep(pyvc, a, reqs, rates):
    """The same trace offered at increasing rates: p50/p99 and achieved rps
    per rate (the knee is the highest rate still served at about the offered
    rate)."""
    out = []
    for q in rates:
        r = serve_bench(pyvc, a, reqs, float(q))
        out.append({k: r[k] for k in ("offered_qps", "achieved_rps", "p50", "p99", "errors", "p99_breakdown_ms")})
        log(f"  sweep {q}: {out[-1]}")
    return out


def dest_bench(pyvc, a, n=4096, reps=5, seed=19):
    """One request of n random LOD0 atoms at a time through the tuned server,
    its output in pinned host memory vs device memory (no PCIe copy): the
    median latency of each."""
    rng = np.random.default_rng(seed)
    _, (nx, ny, nz) = a.lod_dims(0)
    q = np.stack([np.zeros(n, np.int64), rng.integers(0, (nx + 31) // 32, n), rng.integers(0, (ny + 31) // 32, n),
                  rng.integers(0, (nz + 31) // 32, n)], 1).astype(np.int32)
    pin = pyvc.PinnedBuffer(n * 32768)
    dev = pyvc.dev_alloc(n * 32768)
    srv = pyvc.Server(a)
    out = {"atoms": n}
    try:
        assert srv.autotune() == 0
        for name, ptr in (("pinned_host", pin.p), ("device", dev)):
            lat = []
            for _ in range(reps + 1):
                t0 = pyvc.now_ns()
                i = srv.submit_atoms(q, ptr)
                assert i >= 0
                while True:
                    got, st, t = srv.poll(wait_ms=100)
                    if len(got):
                        assert int(st[0]) == 0
                        lat.append((int(t[0]) - t0) / 1e6)
                        break
            out[f"{name}_ms"] = round(float(np.median(lat[1:])), 3)
    finally:
        srv.close()
        pin.free()
        pyvc.dev_free(dev)
    return out


def atom_floor(pyvc, a, sample=3000, seed=13):
    """Per LOD, the heaviest atom of a random sample (largest payload, i.e.
    most serial range-coder bins) decoded alone through vc_decode_atom: the
    latency no request containing that atom can beat, since one lane decodes
    one atom serially."""
    L = pyvc.lib()
    rng = np.random.default_rng(seed)
    out = np.zeros(32768, np.uint8)
    res = []
    for lod in range(lod_count(a)):
        _, (nx, ny, nz) = a.lod_dims(lod)
        ax, ay, az = (nx + 31) // 32, (ny + 31) // 32, (nz + 31) // 32
        best = (-1, None)
        for _ in range(min(sample, ax * ay * az)):
            c = (int(rng.integers(0, az)), int(rng.integers(0, ay)), int(rng.integers(0, ax)))
            ln = a.payload_range(lod, *c)[2]
            if ln > best[0]:
                best = (ln, c)
        z, y, x = best[1]
        ts = []
        for _ in range(6):
            t0 = time.perf_counter()
            L.vc_decode_atom(a.h, lod, x, y, z, out.ctypes.data)
            ts.append((time.perf_counter() - t0) * 1e3)
        res.append({"lod": lod, "max_payload_bytes": best[0], "decode_ms": round(min(ts[1:]), 3)})
    return res


def plain_latency(pyvc, a, reqs, limit_s=4.0, min_n=5):
    """The plain per-call path: vc_decode_atom per atom / vc_decode_region per
    box, host wall clock per request, on the first requests of the trace until
    `limit_s` is spent and at least min_n requests are done."""
    L = pyvc.lib()
    out = np.zeros(32768, np.uint8)
    lat, t_all = [], time.perf_counter()
    for kind, x in reqs:
        t0 = time.perf_counter()
        if kind == "atoms":
            for lod, ax, ay, az in x:
                L.vc_decode_atom(a.h, int(lod), int(ax), int(ay), int(az), out.ctypes.data)
        else:
            a.decode_region(x[0], x[1])
        lat.append((time.perf_counter() - t0) * 1e3)
        if time.perf_counter() - t_all > limit_s and len(lat) >= min_n:
            break
    lat = np.array(lat)
    return {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
            "mean": round(float(lat.mean()), 3), "n": int(lat.size),
            "rps": round(1e3 / float(lat.mean()), 3)}


def serving_section(pyvc, args):
    """Serving configurations: the 2048^3 ratio-50 archive (320 sheets), an
    atom-request trace and a region-request trace through the server, each
    beside the plain per-call path on the same requests."""
    n = args.serve_size
    dims = (n, n, n)
    dvol = pyvc.dev_alloc(n ** 3)
    try:
        assert pyvc.synth_device(3, dims, args.serve_sheets, dvol) == 0
        st, arc, t_enc = pyvc.encode_device(dvol, dims, 50.0, timed=True)
        assert st == 0, st
    finally:
        pyvc.dev_free(dvol)
    log(f"serving archive {n}^3 ratio 50: {len(arc) / 1e6:.1f} MB, encode {t_enc:.2f} s")
    a = pyvc.Archive(arc)
    rng = np.random.default_rng(23)
    areqs = atom_requests(a, 300, rng)
    rreqs = region_requests(a, 200, rng)
    # warm: one request of each kind through the per-call path uploads the archive
    plain_latency(pyvc, a, areqs[:1], 0.0)
    res = {"workload": f"{n}^3 u8 synthetic, {args.serve_sheets} sheets, seed 3, target_ratio 50",
           "archive_bytes": len(arc), "encode_s": round(t_enc, 3)}
    res["autotune"] = tune_section(pyvc, a)
    log(f"autotune: {res['autotune']}")
    sa = serve_bench(pyvc, a, areqs, args.qps)
    sa["untuned"] = {k: v for k, v in serve_bench(pyvc, a, areqs, args.qps, tune=False).items()
                     if k in ("p50", "p99", "achieved_rps")}
    pa = plain_latency(pyvc, a, areqs, limit_s=15.0, min_n=12)
    sa["rps_vs_plain"] = round(sa["achieved_rps"] / pa["rps"], 1)
    ca = None if args.no_cpu else cpu_serve_bench(arc, areqs, args.qps)
    res["atom_requests"] = {"batched": sa, "plain": pa, "cpu": ca,
                            "sweep": qps_sweep(pyvc, a, areqs, (200, 400, 800, 1600, 3200))}
    log(f"atom serving: {sa}; plain {pa}; cpu {ca}")
    sr = serve_bench(pyvc, a, rreqs, args.region_qps)
    pr = plain_latency(pyvc, a, rreqs, limit_s=8.0, min_n=50)
    sr["rps_vs_plain"] = round(sr["achieved_rps"] / pr["rps"], 1)
    cr = None if args.no_cpu else cpu_serve_bench(arc, rreqs, args.region_qps)
    res["region_requests"] = {"batched": sr, "plain": pr, "cpu": cr,
                              "sweep": qps_sweep(pyvc, a, rreqs, (200, 400, 800, 1600, 3200))}
    log(f"region serving: {sr}; plain {pr}; cpu {cr}")
    res["destination_4096_atoms"] = dest_bench(pyvc, a)
    log(f"4096-atom request by destination: {res['destination_4096_atoms']}")
    res["atom_floor"] = atom_floor(pyvc, a)
    log(f"atom floor: {res['atom_floor']}")
    res["repack"] = repack_bench(pyvc, a, arc, no_cpu=args.no_cpu)
    log(f"repack: {res['repack']}")
    res["streaming"] = stream_bench(pyvc, a, arc)
    log(f"streaming: {res['streaming']}")
    a.close()
    return res


def stream_bench(pyvc, a, arc):
    """The serving archive opened with vc_open_streaming over a Python read
    function on host memory: LOD0 decoded into device memory from the
    region-blob cache, cold (every blob fetched by one read) and warm, beside
    the in-memory archive; and LOD1 decoded by per-atom reads (the path the
    cache replaces) with a cache too small for any blob."""
    n = a.lod_dims(0)[1]
    nb = n[0] * n[1] * n[2]
    dev = pyvc.dev_alloc(nb)
    out = {"workload": "2048^3 ratio-50 archive, vc_open_streaming over host memory", "archive_bytes": len(arc)}
    try:
        s = pyvc.StreamingArchive(arc)
        t0 = time.perf_counter()
        assert s.decode_lod_device(0, dev, 1) == 0
        t_cold = time.perf_counter() - t0
        reads, _ = s.stats()
        t0 = time.perf_counter()
        assert s.decode_lod_device(0, dev, 1) == 0
        t_warm = time.perf_counter() - t0
        got = pyvc.d2h(dev, nb)
        t0 = time.perf_counter()
        assert a.decode_lod_device(0, dev, 1) == 0
        t_flat = time.perf_counter() - t0
        same = np.array_equal(got, pyvc.d2h(dev, nb))
        out.update({"lod0_cold_s": round(t_cold, 4), "lod0_warm_s": round(t_warm, 4), "lod0_in_memory_s": round(t_flat, 4),
                    "lod0_cold_GBps": round(nb / t_cold / 1e9, 2), "blob_reads": reads, "read_calls": s.reads,
                    "identical": bool(same)})
        s.close()
        s = pyvc.StreamingArchive(arc)
        s.set_cache(1 << 20)                      # no blob fits: per-atom reads
        d1 = s.lod_dims(1)[1]
        t0 = time.perf_counter()
        st, v1 = s.decode_lod(1)
        t_atoms = time.perf_counter() - t0
        st2, f1 = a.decode_lod(1)
        out.update({"lod1_per_atom_reads_s": round(t_atoms, 3), "lod1_per_atom_read_calls": s.reads,
                    "lod1_GBps_per_atom_reads": round(d1[0] * d1[1] * d1[2] / t_atoms / 1e9, 3),
                    "lod1_identical": bool(st == st2 and np.array_equal(v1, f1))})
        s.close()
    finally:
        pyvc.dev_free(dev)
    return out


def repack_bench(pyvc, a, arc, reps=3, no_cpu=False):
    """Container rewrite of the serving archive into device memory
    (vcx_repack_device, a synchronous call, host wall clock).  The archive is
    the encoder's canonical v2, so the image must equal it byte for byte.
    Beside it, the reference's own tool (tools/vc_repack.c, built from source
    into oracle/_ref) on the same file with 16 threads, when built."""
    st, n = a.repack_device(None, 0)
    assert st == 2 and n == len(arc), (st, n)
    dev = pyvc.dev_alloc(n)
    try:
        ts = []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            st, _ = a.repack_device(dev, n)
            ts.append(time.perf_counter() - t0)
            assert st == 0, st
        same = pyvc.d2h(dev, n).tobytes() == arc
    finally:
        pyvc.dev_free(dev)
    t = min(ts[1:])
    # payload bytes: data blobs less their 512 KiB slot arrays (v2 directory)
    import struct
    dir_off, cnt = struct.unpack_from("<QQ", arc, 40)
    ents = np.frombuffer(arc, np.uint32, 4 * cnt, dir_off).reshape(cnt, 4)
    data = ents[:, 3] == 2
    payload = int(ents[data, 2].astype(np.int64).sum()) - int(data.sum()) * 32768 * 16
    out = {"workload": "2048^3 ratio-50 archive (v2) rewritten into device memory", "bytes": n,
           "payload_bytes": payload, "data_regions": int(data.sum()),
           "device_s": round(t, 5), "GBps": round(n / t / 1e9, 2), "identical": bool(same)}
    tool = os.path.join(ROOT, "oracle", "_ref", "vc_repack")
    if not no_cpu and os.path.exists(tool):
        import subprocess
        import tempfile
        d = tempfile.mkdtemp()
        p_in, p_out = os.path.join(d, "in.vca"), os.path.join(d, "out.vca")
        try:
            open(p_in, "wb").write(arc)
            t0 = time.perf_counter()
            subprocess.run([tool, p_in, p_out, "--threads", "16"], check=True, stdout=subprocess.DEVNULL, timeout=120)
            out["reference_tool_s"] = round(time.perf_counter() - t0, 3)
            out["reference_tool_identical"] = open(p_out, "rb").read() == arc
        finally:
            for f in (p_in, p_out):
                if os.path.exists(f):
                    os.unlink(f)
            os.rmdir(d)
    return out


def ds_section(pyvc, hip, dvol, dout, n, reps=10):
    """vcx_downscale2x_device (cbox, alpha 0.5, the exporter's default) of the
    device-resident n^3 volume into dout, back to back on the library stream:
    n^3 bytes read and n^3/8 written per call (HBM-bound)."""
    L = pyvc.lib()
    stream = L.vcx_stream()
    d = pyvc.Dims(n, n, n)
    assert L.vcx_downscale2x_device(dvol, d, 1, C.c_float(0.5), dout, 1) == 0
    e0, e1 = hip.event(), hip.event()
    hip.L.hipStreamSynchronize(stream)
    hip.record(e0, stream)
    for _ in range(reps):
        L.vcx_downscale2x_device(dvol, d, 1, C.c_float(0.5), dout, 0)
    hip.record(e1, stream)
    ms = hip.ms(e0, e1) / reps
    assert L.vcx_sync() == 0
    moved = n ** 3 + (n // 2) ** 3
    return {"workload": f"{n}^3 u8 -> {n // 2}^3, cbox alpha 0.5", "ms": round(ms, 4),
            "GBps_moved": round(moved / ms / 1e6, 1), "frac_of_hbm": round(moved / ms / 1e6 / 8000, 3)}


def transfer_section(pyvc, a, dout, n, reps=3):
    """Host<->device copies of the LOD0 volume's size (pageable numpy memory
    and pinned memory), and vc_decode_lod LOD0 into host memory (the drop-in
    call: device decode plus the copy out), against the device-resident rate."""
    L = pyvc.lib()
    nb = n ** 3
    host = np.ones(nb, np.uint8)
    pin = pyvc.PinnedBuffer(nb)
    pin.a[:] = 1
    out = {"bytes": nb}

    def best(f):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            f()
            ts.append(time.perf_counter() - t0)
        return min(ts)
    try:
        out["h2d_pageable_GBps"] = round(nb / best(lambda: L.vcx_h2d(dout, host.ctypes.data, nb)) / 1e9, 2)
        out["h2d_pinned_GBps"] = round(nb / best(lambda: L.vcx_h2d(dout, pin.p, nb)) / 1e9, 2)
        out["d2h_pageable_GBps"] = round(nb / best(lambda: L.vcx_d2h(host.ctypes.data, dout, nb)) / 1e9, 2)
        out["d2h_pinned_GBps"] = round(nb / best(lambda: L.vcx_d2h(pin.p, dout, nb)) / 1e9, 2)
        d = pyvc.Dims()
        t = best(lambda: L.vc_decode_lod(a.h, 0, host.ctypes.data, C.byref(d)))
        out["decode_lod_host_pageable_ms"] = round(t * 1e3, 3)
        out["decode_lod_host_pageable_GBps"] = round(nb / t / 1e9, 2)
        t = best(lambda: L.vc_decode_lod(a.h, 0, pin.p, C.byref(d)))
        out["decode_lod_host_pinned_ms"] = round(t * 1e3, 3)
        out["decode_lod_host_pinned_GBps"] = round(nb / t / 1e9, 2)
    finally:
        pin.free()
    return out


def export_section(pyvc, dvol, n, no_cpu=False):
    """vcx_export_volume of the device-resident n^3 volume (the exporter's
    cascade: cbox 0.5, q0 1.0, falloff 0.5) into a v1 file, wall clock of the
    C call; beside it the reference writer appending the same volume's LOD0
    atoms (the exporter's LOD0 share) on the host cores."""
    import tempfile
    d = tempfile.mkdtemp()
    path = os.path.join(d, "export.vca")
    try:
        ts = []
        for _ in range(2):
            t0 = time.perf_counter()
            assert pyvc.export_volume(dvol, path, dims=(n, n, n)) == 0
            ts.append(time.perf_counter() - t0)
        size = os.path.getsize(path)
        out = {"workload": f"{n}^3 device-resident volume -> v1 archive (exporter cascade, cbox 0.5, q0 1.0, falloff 0.5)",
               "s": round(min(ts), 3), "GBps": round(n ** 3 / min(ts) / 1e9, 3), "archive_bytes": size}
        lib = None if no_cpu else _ref_lib()
        if lib:
            L = C.CDLL(os.path.join(ROOT, "oracle", "libcpubench.so"))
            L.cpu_append_lod0.restype = C.c_double
            L.cpu_append_lod0.argtypes = [C.c_char_p, C.c_char_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                          C.c_float, C.c_int, C.POINTER(C.c_long)]
            hv = pyvc.d2h(dvol, n ** 3)
            bad = C.c_long()
            t = L.cpu_append_lod0(lib.encode(), path.encode(), hv.ctypes.data, n, n, n, C.c_float(1.0), _threads(),
                                  C.byref(bad))
            assert t > 0 and bad.value == 0, (t, bad.value)
            out["cpu_lod0_appends"] = {"s": round(t, 3), "GBps": round(n ** 3 / t / 1e9, 3), "cores": _threads(),
                                       "kind": "reference", "lib": os.path.basename(lib),
                                       "sample": f"every LOD0 atom of the {n}^3 volume at q 1.0 (the LOD0 share of "
                                                 "the export), reference vc_append_atom under OpenMP"}
    finally:
        if os.path.exists(path):
            os.unlink(path)
        os.rmdir(d)
    return out


def ratio_section(pyvc, hip, o, dvol, dout, n, ratios, steps=30):
    """The same volume encoded at other target ratios: encode time, then
    back-to-back device LOD0 decodes (GB/s) and 16 sampled atoms against the
    oracle."""
    L = pyvc.lib()
    stream = L.vcx_stream()
    nax = (n + 31) // 32
    rng = np.random.default_rng(9)
    out = {}
    for r in ratios:
        st, arc, t_enc = pyvc.encode_device(dvol, (n, n, n), r, timed=True)
        assert st == 0, st
        a = pyvc.Archive(arc)
        for _ in range(2):
            assert a.decode_lod_device(0, dout, 0) == 0
        assert L.vcx_sync() == 0
        e0, e1 = hip.event(), hip.event()
        hip.L.hipStreamSynchronize(stream)
        hip.record(e0, stream)
        for _ in range(steps):
            a.decode_lod_device(0, dout, 0)
        hip.record(e1, stream)
        ms = hip.ms(e0, e1) / steps
        assert L.vcx_sync() == 0
        dec = pyvc.d2h(dout, n ** 3).reshape(n, n, n)
        oa = o.open(arc)
        exact = True
        for _ in range(16):
            az, ay, ax = (int(x) for x in rng.integers(0, nax, 3))
            ost, oat = oa.decode_atom(0, ax, ay, az)
            exact &= ost == 0 and np.array_equal(
                dec[az * 32:az * 32 + 32, ay * 32:ay * 32 + 32, ax * 32:ax * 32 + 32], oat)
        del dec
        a.close()
        out[f"ratio_{r:g}"] = {"decode_GBps": round(n ** 3 / (ms / 1e3) / 1e9, 3), "ms_per_step": round(ms, 4),
                               "encode_s": round(t_enc, 3), "encode_GBps": round(n ** 3 / t_enc / 1e9, 3),
                               "archive_bytes": len(arc), "bit_exact_sampled": bool(exact)}
        log(f"ratio {r:g}: {out[f'ratio_{r:g}']}")
    return out


def encode_batch_section(pyvc, nvol=256, nseq=16, seed=2024, with_cpu=False):
    """Batched encode service: nvol device-resident volumes with each axis
    uniform in [40, 512], sheets scaled with the volume's edge (20 per 128
    voxels), target ratios uniform over {10, 20, 50, 100}; one
    vcx_encode_batch call, beside vcx_encode_device called per volume on the
    first nseq volumes (their archives must match byte for byte)."""
    rng = np.random.default_rng(seed)
    dims = [tuple(int(v) for v in rng.integers(40, 513, 3)) for _ in range(nvol)]
    targets = [float((10, 20, 50, 100)[int(rng.integers(0, 4))]) for _ in range(nvol)]
    ptrs, tot = [], 0
    try:
        for k, d in enumerate(dims):
            nb = d[0] * d[1] * d[2]
            tot += nb
            p = pyvc.dev_alloc(nb)
            ptrs.append(p)
            assert pyvc.synth_device(5000 + k, d, max(1, round(20 * nb ** (1 / 3) / 128)), p) == 0
        pyvc.encode_batch(ptrs[:2], targets[:2], on_device=True, dims=dims[:2])   # warm
        # times are those of the C calls: the archives' copies into Python
        # bytes objects (for the comparison below) are left out
        rc, res, t_batch = pyvc.encode_batch(ptrs, targets, on_device=True, dims=dims, timed=True)
        assert rc == 0, rc
        sub = sum(d[0] * d[1] * d[2] for d in dims[:nseq])
        t_seq = 0.0
        same = True
        for k in range(nseq):
            st, arc, tk = pyvc.encode_device(ptrs[k], dims[k], targets[k], timed=True)
            t_seq += tk
            same &= st == 0 and arc == res[k][1]
        cpu = None
        if with_cpu:
            hv = [pyvc.d2h(ptrs[k], dims[k][0] * dims[k][1] * dims[k][2]).reshape(dims[k][2], dims[k][1], dims[k][0])
                  for k in range(64)]
            cpu = cpu_encode_bench(hv, targets[:64])
            del hv
    finally:
        for p in ptrs:
            pyvc.dev_free(p)
    out = {"workload": f"{nvol} volumes, axes uniform in [40, 512], ratios in {{10, 20, 50, 100}}, device-resident",
           "input_GB": round(tot / 1e9, 3), "archive_GB": round(sum(len(x) for _, x in res) / 1e9, 3),
           "batched_s": round(t_batch, 3), "batched_GBps": round(tot / t_batch / 1e9, 3),
           "per_call_GBps": round(sub / t_seq / 1e9, 3) if nseq else None, "per_call_volumes": nseq,
           "batched_vs_per_call": round((tot / t_batch) / (sub / t_seq), 2) if nseq else None,
           "identical_to_per_call": bool(same), "cpu_baseline": cpu}
    log(f"batched encode: {out}")
    return out


def pmc_traffic(kernel, size, ratio):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (calibrated against a plain copy), or None."""
    f = os.path.join(ROOT, "profiles", f"decode_lod0_{size}_ratio{ratio:g}_pmc_bytes.md")
    if not os.path.exists(f):
        return None
    for line in open(f):
        w = line.split()
        if len(w) == 3 and w[0] == "traffic_bytes" and w[1] == kernel:
            return int(w[2])
    return None


def sq_counters(size, ratio):
    """{kernel: VALU wave-instructions per launch} from the committed rocprofv3
    SQ-counter summary of this workload, or {}."""
    f = os.path.join(ROOT, "profiles", f"decode_lod0_{size}_ratio{ratio:g}_sq_counters.md")
    out = {}
    if os.path.exists(f):
        for line in open(f):
            w = line.split()
            if len(w) == 3 and w[0] == "valu_insts":
                out[w[1]] = int(w[2])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--ratio", type=float, default=10.0)
    ap.add_argument("--sheets", type=int, default=160)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--qps", type=float, default=400.0, help="offered rate of the atom-request trace")
    ap.add_argument("--region-qps", type=float, default=400.0, help="offered rate of the region-request trace")
    ap.add_argument("--serve-size", type=int, default=2048)
    ap.add_argument("--serve-sheets", type=int, default=320)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-ratios", action="store_true")
    ap.add_argument("--no-encode-batch", action="store_true")
    ap.add_argument("--downscale-only", action="store_true",
                    help="profiler run: only the within-cell downscale of the 1024^3 volume, its JSON on stdout")
    ap.add_argument("--repack-only", action="store_true",
                    help="profiler run: only the container rewrite of the 2048^3 ratio-50 archive, its JSON on stdout")
    ap.add_argument("--encode-batch-only", action="store_true",
                    help="profiler run: only the batched-encode section (256 volumes), its JSON on stdout")
    ap.add_argument("--decode-only", action="store_true",
                    help="profiler run: only the LOD0 decode section, the timed calls back to back as in a "
                         "full run (no per-atom calls, no other sections, no CPU leg)")
    ap.add_argument("--profile", action="store_true",
                    help="profiler run: only the LOD0 decode launches, each call synchronous (no per-atom "
                         "calls, no serving trace, no CPU leg), then one 1 GiB calibration copy kernel")
    args = ap.parse_args()
    # sections beyond the LOD0 decode run only in a full run
    args.quick = args.profile or args.decode_only
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the decode shards by atom "
              "with no inter-GPU traffic", file=sys.stderr)
        sys.exit(2)

    pyvc = load_pyvc()
    hip = Hip()
    assert pyvc.device_init() == 0, "no HIP device"
    if args.encode_batch_only:
        print(json.dumps(encode_batch_section(pyvc, nseq=0)), flush=True)
        return
    if args.downscale_only:
        n = args.size
        dvol, dout = pyvc.dev_alloc(n ** 3), pyvc.dev_alloc((n // 2) ** 3)
        try:
            assert pyvc.synth_device(2, (n, n, n), args.sheets, dvol) == 0
            print(json.dumps(ds_section(pyvc, hip, dvol, dout, n)), flush=True)
        finally:
            pyvc.dev_free(dvol)
            pyvc.dev_free(dout)
        return
    if args.repack_only:
        n = args.serve_size
        dvol = pyvc.dev_alloc(n ** 3)
        try:
            assert pyvc.synth_device(3, (n, n, n), args.serve_sheets, dvol) == 0
            st, arc = pyvc.encode_device(dvol, (n, n, n), 50.0)
            assert st == 0, st
        finally:
            pyvc.dev_free(dvol)
        a = pyvc.Archive(arc)
        print(json.dumps(repack_bench(pyvc, a, arc, reps=5, no_cpu=True)), flush=True)
        a.close()
        return
    L = pyvc.lib()
    n = args.size
    dims = (n, n, n)
    nvox = n ** 3

    # ---- synthetic volume + encode (setup, untimed for the metric) ----
    t0 = time.time()
    dvol = pyvc.dev_alloc(nvox)
    assert pyvc.synth_device(2, dims, args.sheets, dvol) == 0
    log(f"synth {n}^3 ({args.sheets} sheets): {time.time() - t0:.2f} s")
    # the C call, timed on its second run (the first also grows the encoder's
    # device scratch and pinned staging, once per process); the archives match
    st, arc0, t_enc_first = pyvc.encode_device(dvol, dims, args.ratio, timed=True)
    assert st == 0, st
    st, arc, t_enc = pyvc.encode_device(dvol, dims, args.ratio, timed=True)
    assert st == 0 and arc == arc0, st
    del arc0
    a = pyvc.Archive(arc)
    nax = (n + 31) // 32
    natoms = nax ** 3
    pay = np.zeros(natoms, np.int64)
    for az in range(nax):
        for ay in range(nax):
            for ax in range(nax):
                pay[(az * nax + ay) * nax + ax] = a.payload_range(0, az, ay, ax)[2]
    pay_bytes = int(pay.sum())
    # encode's algorithmic bytes (input voxels, each pyramid level read and
    # the next written, every payload byte, the archive bytes)
    lvl = [a.lod_dims(l)[1] for l in range(lod_count(a))]
    vox = [int(x) * int(y) * int(z) for x, y, z in lvl]
    all_pay = pay_bytes
    for l in range(1, len(lvl)):
        gx, gy, gz = ((d + 31) // 32 for d in lvl[l])
        all_pay += sum(a.payload_range(l, az, ay, ax)[2] for az in range(gz) for ay in range(gy) for ax in range(gx))
    enc_bytes = vox[0] + sum(vox[l - 1] + vox[l] for l in range(1, len(vox))) + all_pay + len(arc)
    log(f"encode (ratio {args.ratio}): {t_enc:.2f} s, archive {len(arc) / 1e6:.1f} MB, "
        f"LOD0 payload {pay_bytes / 1e6:.1f} MB (ratio {nvox / pay_bytes:.2f})")

    # ---- device decode: warm up (also uploads the archive) ----
    dout = pyvc.dev_alloc(nvox)
    stream = L.vcx_stream()
    for _ in range(args.warmup):
        assert a.decode_lod_device(0, dout, 0) == 0
    assert L.vcx_sync() == 0

    # ---- bit-exactness: sampled atoms vs the CPU oracle ----
    if args.warmup == 0:                      # the check reads a decode's output
        assert a.decode_lod_device(0, dout, 0) == 0
        assert L.vcx_sync() == 0
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle
    o = Oracle()
    oa = o.open(arc)
    dec = pyvc.d2h(dout, nvox).reshape(n, n, n)
    rng = np.random.default_rng(5)
    checks = 64
    exact = True
    for _ in range(checks):
        az, ay, ax = (int(x) for x in rng.integers(0, nax, 3))
        ost, oat = oa.decode_atom(0, ax, ay, az)
        if ost != 0 or not np.array_equal(dec[az * 32:az * 32 + 32, ay * 32:ay * 32 + 32, ax * 32:ax * 32 + 32], oat):
            exact = False
    del dec
    log(f"bit-exact vs oracle on {checks} sampled atoms: {exact}")

    # ---- timed region: K steps bracketed by events on the library stream ----
    e0, e1 = hip.event(), hip.event()
    hip.L.hipStreamSynchronize(stream)
    hip.record(e0, stream)
    for _ in range(args.steps):
        # profile runs keep every call synchronous so per-kernel durations are
        # those of isolated launches (the throughput loop overlaps them)
        a.decode_lod_device(0, dout, 1 if args.profile else 0)
    hip.record(e1, stream)
    ms_total = hip.ms(e0, e1)
    assert L.vcx_sync() == 0
    ms_step = ms_total / args.steps
    gbs = nvox / (ms_step / 1e3) / 1e9

    # ---- per-kernel durations via stage events ----
    ev = [hip.event() for _ in range(3)]
    evarr = (C.c_void_p * 3)(*[e.value for e in ev])
    L.vcx_set_stage_events(evarr)
    k1, k2, k12 = [], [], []
    for _ in range(args.steps):
        a.decode_lod_device(0, dout, 1)
        k1.append(hip.ms(ev[0], ev[1]))
        k2.append(hip.ms(ev[1], ev[2]))
        k12.append(hip.ms(ev[0], ev[2]))
    L.vcx_set_stage_events(None)
    k1_ms, k2_ms = float(np.mean(k1)), float(np.mean(k2))
    # one call at a time, as a caller sees it (events on the library stream
    # around each call)
    single = []
    for _ in range(args.steps):
        hip.L.hipStreamSynchronize(stream)
        hip.record(e0, stream)
        a.decode_lod_device(0, dout, 0)
        hip.record(e1, stream)
        single.append(hip.ms(e0, e1))
    assert L.vcx_sync() == 0
    slot_bytes = 16 * natoms
    region_bytes = 16 * ((nax + 31) // 32) ** 3
    # algorithmic bytes of one decode of the LOD0 (32768 atoms: one launch of
    # each kernel): every touched atom's payload + its 16-byte slot, 16 bytes
    # per touched index region, and the voxels written (1 byte each)
    alg_step = pay_bytes + slot_bytes + region_bytes + nvox
    if k1_ms >= k2_ms:
        dom, dom_ms = "k_entropy", k1_ms
    else:
        dom, dom_ms = "k_idct_m", k2_ms
    dom_bytes = alg_step
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    log(f"step {ms_step:.3f} ms ({gbs:.1f} GB/s); k_entropy {k1_ms:.3f} ms, transform {k2_ms:.3f} ms")

    # ---- plain per-call path: vc_decode_atom latency (host wall clock) ----
    lat = []
    out = np.zeros(32768, np.uint8)
    for i in range(0 if args.quick else 200):
        az, ay, ax = (int(x) for x in rng.integers(0, nax, 3))
        t0 = time.perf_counter()
        L.vc_decode_atom(a.h, 0, ax, ay, az, out.ctypes.data)
        lat.append((time.perf_counter() - t0) * 1e3)
    lat = np.array(lat[10:])
    p50, p99 = (float(np.percentile(lat, 50)), float(np.percentile(lat, 99))) if lat.size else (None, None)

    # ---- host transfers and the host-destination drop-in call ----
    xfer = None if args.quick else transfer_section(pyvc, a, dout, n)
    if xfer:
        log(f"transfers: {xfer}")
    # ---- vc_encode on the host cores (the oracle's OpenMP encoder) ----
    cpu_enc = None
    if not (args.no_cpu or args.quick):
        hv = pyvc.d2h(dvol, nvox).reshape(n, n, n)
        cpu_enc = cpu_encode_bench([hv], [args.ratio])
        cpu_enc["gpu_GBps"] = round(nvox / t_enc / 1e9, 3)
        del hv
        log(f"cpu encode {n}^3: {cpu_enc}")

    # ---- the same volume at the other target ratios ----
    ratios = None
    if not (args.no_ratios or args.quick):
        ratios = ratio_section(pyvc, hip, o, dvol, dout, n, [r for r in (10.0, 50.0, 100.0) if r != args.ratio])
    # ---- the exporter's within-cell downscale of the same volume (HBM-bound) ----
    ds = None
    if not args.quick:
        ds = ds_section(pyvc, hip, dvol, dout, n)
        log(f"downscale: {ds}")
    exp = None
    if not args.quick:
        exp = export_section(pyvc, dvol, n, no_cpu=args.no_cpu)
        log(f"export: {exp}")
    pyvc.dev_free(dout)
    pyvc.dev_free(dvol)
    dout = dvol = None

    # ---- serving: continuous-batching request loop under Poisson traces ----
    served = None
    if not (args.no_serve or args.quick):
        served = serving_section(pyvc, args)

    # ---- batched encode service ----
    enc_batch = None
    if not (args.no_encode_batch or args.quick):
        enc_batch = encode_batch_section(pyvc, with_cpu=not args.no_cpu)

    # ---- CPU baseline ----
    cpu = None
    if not (args.no_cpu or args.quick):
        zyx = np.array([(z, y, x) for z in range(nax) for y in range(nax) for x in range(nax)], np.int32)
        cpu = cpu_baseline(arc, zyx)
        log(f"cpu baseline: {cpu}")

    vi = sq_counters(n, args.ratio)
    res = {
        "metric": METRIC,
        "value": round(gbs, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"vc_decode_lod LOD0 (device-resident, batched), {n}^3 u8 synthetic, "
                               f"{args.sheets} sheets, seed 2, target_ratio {args.ratio}",
                   "volume": f"{n}^3", "target_ratio": args.ratio, "atoms": natoms,
                   "lod0_payload_bytes": pay_bytes, "achieved_lod0_ratio": round(nvox / pay_bytes, 3),
                   "archive_bytes": len(arc), "parallelism": "single GPU"},
        "bit_exact": exact,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
                     "traffic": pmc_traffic(dom, n, args.ratio),
                     "algorithmic_bytes": dom_bytes, "avg_ms": round(dom_ms, 4),
                     # the kernel's real limit: share of the chip's VALU issue
                     # slots it uses (profiled instruction count x 4 cycles over
                     # 1024 SIMDs at 2.4 GHz for this run's launch time)
                     "valu_issue_frac": (round(vi[dom] * 4 / (1024 * 2.4e9 * dom_ms / 1e3), 4)
                                         if dom in vi else None)},
        "step_roofline": {"achieved": round(alg_step / (ms_step / 1e3) / 1e9, 3), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(alg_step / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 6),
                          "algorithmic_bytes": alg_step},
        # the transform stage: k_idct_m, then k_idct_ovf (a scan that ends at
        # once when no atom is past the matrix-core bound)
        "kernels_ms": {"k_entropy": round(k1_ms, 4), "k_idct_m+k_idct_ovf": round(k2_ms, 4)},
        # one call at a time (back-to-back calls overlap the transform of call
        # k with the entropy pass of k+1)
        "single_call_ms": round(float(np.mean(single)), 4),
        "plain_atom_latency_ms": ({"p50": round(p50, 4), "p99": round(p99, 4), "n": int(lat.size)}
                                  if lat.size else None),
        "serving": served,
        "encode_s": round(t_enc, 3), "encode_first_call_s": round(t_enc_first, 3),
        # the whole vc_encode call (rate control, pyramid, transforms, range
        # coding, assembly) against HBM peak; its time is set by the serial
        # range coding of the heaviest coarse-LOD atoms
        "encode_roofline": {"bound": "hbm", "achieved": round(enc_bytes / t_enc / 1e9, 3), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(enc_bytes / t_enc / 1e9 / HBM_PEAK_GBS, 6),
                            "algorithmic_bytes": enc_bytes, "s": round(t_enc, 4)},
        "other_ratios": ratios,
        "downscale": ds,
        "export": exp,
        "batched_encode": enc_batch,
        "transfers": xfer,
        "cpu_baseline": cpu,
        "cpu_encode": cpu_enc,
    }
    if args.profile:
        # 1 GiB device copy by a plain kernel (k_copy16): byte-counter reference
        cb = 1 << 30
        src, dst = pyvc.dev_alloc(cb), pyvc.dev_alloc(cb)
        assert L.vcx_copy_kernel(dst, src, cb) == 0
        pyvc.dev_free(src)
        pyvc.dev_free(dst)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
