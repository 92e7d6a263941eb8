"""Benchmark: batched vc_decode_lod of LOD0 of a 1024^3 u8 synthetic volume.

One step = one device-resident decode of every LOD0 atom (32768 atoms, one
entropy launch + one transform launch) from the archive already resident in
HBM into a device output buffer.  Steps are issued back to back through the
asynchronous device API, which overlaps the transform of one call with the
entropy pass of the next; the timed region brackets all K steps with HIP
events on the library's own stream.  The latency of one synchronous call is
reported beside it (single_call_ms).  Also reported: per-kernel durations (stage events), the dominant
kernel's roofline, plain per-call vc_decode_atom latency p50/p99, a
bit-exactness check against the CPU oracle on sampled atoms, and the CPU
baseline (the reference's vc_decode_atom under OpenMP on the host cores).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 1024] [--ratio 10]
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


def serve_bench(pyvc, a, nreq=300, qps=400.0, seed=11):
    """Open-loop Poisson trace through the continuous-batching server:
    requests of round(exp(U(ln 1, ln 4096))) atoms, LOD 0 with p = 0.7 else
    uniform over the coarser LODs, coordinates uniform over that LOD's atom
    grid.  Latency = completion (server clock) - scheduled arrival."""
    rng = np.random.default_rng(seed)
    nlod = 0
    while a.lod_dims(nlod)[0] == 0:
        nlod += 1
    reqs = []
    for _ in range(nreq):
        n = int(round(np.exp(rng.uniform(0.0, np.log(4096.0)))))
        lod = 0 if rng.random() < 0.7 or nlod == 1 else int(rng.integers(1, nlod))
        _, (nx, ny, nz) = a.lod_dims(lod)
        q = np.stack([np.full(n, lod), rng.integers(0, (nx + 31) // 32, n), rng.integers(0, (ny + 31) // 32, n),
                      rng.integers(0, (nz + 31) // 32, n)], 1).astype(np.int32)
        reqs.append(q)
    gaps = rng.exponential(1.0 / qps, nreq)
    total = sum(len(q) for q in reqs) * 32768
    arena = pyvc.PinnedBuffer(total)
    srv = pyvc.Server(a, 6 << 30)
    offs = np.cumsum([0] + [len(q) * 32768 for q in reqs])
    ids, arrive = {}, np.zeros(nreq)
    done_t, bad = {}, 0
    try:
        t0 = pyvc.now_ns() + 2_000_000
        sched = t0 + np.cumsum(gaps) * 1e9
        for k, q in enumerate(reqs):
            while pyvc.now_ns() < sched[k]:
                pass
            i = srv.submit_atoms(q, arena.p + int(offs[k]))
            assert i >= 0
            ids[int(i)] = k
            arrive[k] = sched[k]
            got, st, t = srv.poll(wait_ms=0)
            for g, s_, tt in zip(got, st, t):
                done_t[ids[int(g)]] = int(tt)
                bad += int(s_ != 0)
        while len(done_t) < nreq:
            got, st, t = srv.poll(wait_ms=100)
            for g, s_, tt in zip(got, st, t):
                done_t[ids[int(g)]] = int(tt)
                bad += int(s_ != 0)
    finally:
        srv.close()
    lat = np.array([(done_t[k] - arrive[k]) / 1e6 for k in range(nreq)])
    span = (max(done_t.values()) - arrive[0]) / 1e9
    atoms = sum(len(q) for q in reqs)
    # spot check: three served requests against the plain per-atom path
    exact = True
    for k in (0, nreq // 2, nreq - 1):
        for m in range(min(4, len(reqs[k]))):
            lod, ax, ay, az = (int(x) for x in reqs[k][m])
            st, ref = a.decode_atom(lod, ax, ay, az)
            base = int(offs[k]) + m * 32768
            exact &= st == 0 and np.array_equal(arena.a[base:base + 32768], ref.ravel())
    arena.free()
    return {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
            "requests": nreq, "offered_qps": qps, "achieved_rps": round(nreq / span, 1),
            "atoms_per_s": round(atoms / span, 1), "mean_atoms_per_request": round(atoms / nreq, 1),
            "errors": bad, "spot_check_exact": bool(exact)}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--ratio", type=float, default=10.0)
    ap.add_argument("--sheets", type=int, default=160)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--qps", type=float, default=400.0, help="offered request rate of the serving trace")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="profiler run: only the LOD0 decode launches, each call synchronous (no per-atom "
                         "calls, no serving trace, no CPU leg), then one 1 GiB calibration copy kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the decode shards by atom "
              "with no inter-GPU traffic, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    pyvc = load_pyvc()
    hip = Hip()
    assert pyvc.device_init() == 0, "no HIP device"
    L = pyvc.lib()
    n = args.size
    dims = (n, n, n)
    nvox = n ** 3

    # ---- synthetic volume + encode (setup, untimed for the metric) ----
    t0 = time.time()
    dvol = pyvc.dev_alloc(nvox)
    assert pyvc.synth_device(2, dims, args.sheets, dvol) == 0
    log(f"synth {n}^3 ({args.sheets} sheets): {time.time() - t0:.2f} s")
    t0 = time.time()
    st, arc = pyvc.encode_device(dvol, dims, args.ratio)
    t_enc = time.time() - t0
    assert st == 0, st
    a = pyvc.Archive(arc)
    nax = (n + 31) // 32
    natoms = nax ** 3
    pay = np.zeros(natoms, np.int64)
    for az in range(nax):
        for ay in range(nax):
            for ax in range(nax):
                pay[(az * nax + ay) * nax + ax] = a.payload_range(0, az, ay, ax)[2]
    pay_bytes = int(pay.sum())
    log(f"encode (ratio {args.ratio}): {t_enc:.2f} s, archive {len(arc) / 1e6:.1f} MB, "
        f"LOD0 payload {pay_bytes / 1e6:.1f} MB (ratio {nvox / pay_bytes:.2f})")

    # ---- device decode: warm up (also uploads the archive) ----
    dout = pyvc.dev_alloc(nvox)
    stream = L.vcx_stream()
    for _ in range(args.warmup):
        assert a.decode_lod_device(0, dout, 0) == 0
    assert L.vcx_sync() == 0

    # ---- bit-exactness: sampled atoms vs the CPU oracle ----
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
    slot_bytes = 16 * natoms
    region_bytes = 16 * ((nax + 31) // 32) ** 3
    alg_k1 = pay_bytes + slot_bytes + region_bytes           # entropy: payload + index reads
    alg_k2 = nvox                                            # transform: voxel writes
    alg_step = pay_bytes + slot_bytes + region_bytes + nvox  # whole decode (BASELINE.md formula)
    if k1_ms >= k2_ms:
        dom, dom_ms, dom_bytes = "k_entropy", k1_ms, alg_k1
    else:
        dom, dom_ms, dom_bytes = "k_idct", k2_ms, alg_k2
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    log(f"step {ms_step:.3f} ms ({gbs:.1f} GB/s); k_entropy {k1_ms:.3f} ms, k_idct {k2_ms:.3f} ms")

    # ---- plain per-call path: vc_decode_atom latency (host wall clock) ----
    lat = []
    out = np.zeros(32768, np.uint8)
    for i in range(0 if args.profile else 200):
        az, ay, ax = (int(x) for x in rng.integers(0, nax, 3))
        t0 = time.perf_counter()
        L.vc_decode_atom(a.h, 0, ax, ay, az, out.ctypes.data)
        lat.append((time.perf_counter() - t0) * 1e3)
    lat = np.array(lat[10:])
    p50, p99 = (float(np.percentile(lat, 50)), float(np.percentile(lat, 99))) if lat.size else (None, None)

    # ---- batched serving: continuous-batching request loop under a Poisson trace ----
    served = None
    if not (args.no_serve or args.profile):
        served = serve_bench(pyvc, a, qps=args.qps)
        log(f"serving: {served}")
        if p50:
            # the plain path serves a request with one vc_decode_atom call per atom
            served["plain_rps_estimate"] = round(1e3 / (p50 * served["mean_atoms_per_request"]), 3)
            served["rps_vs_plain"] = round(served["achieved_rps"] / served["plain_rps_estimate"], 1)

    # ---- CPU baseline ----
    cpu = None
    if not (args.no_cpu or args.profile):
        zyx = np.array([(z, y, x) for z in range(nax) for y in range(nax) for x in range(nax)], np.int32)
        cpu = cpu_baseline(arc, zyx)
        log(f"cpu baseline: {cpu}")

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
                     "algorithmic_bytes": dom_bytes, "avg_ms": round(dom_ms, 4)},
        "step_roofline": {"achieved": round(alg_step / (ms_step / 1e3) / 1e9, 3), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(alg_step / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 6),
                          "algorithmic_bytes": alg_step},
        "kernels_ms": {"k_entropy": round(k1_ms, 4), "k_idct": round(k2_ms, 4)},
        # one synchronous call (entropy start to transform end); back-to-back
        # calls overlap the transform of call k with the entropy pass of k+1
        "single_call_ms": round(float(np.mean(k12)), 4),
        "plain_atom_latency_ms": ({"p50": round(p50, 4), "p99": round(p99, 4), "n": int(lat.size)}
                                  if lat.size else None),
        "batched_atom_latency_ms": served,
        "encode_s": round(t_enc, 3),
        "cpu_baseline": cpu,
    }
    if args.profile:
        # 1 GiB device copy by a plain kernel (k_copy16): byte-counter reference
        cb = 1 << 30
        src, dst = pyvc.dev_alloc(cb), pyvc.dev_alloc(cb)
        assert L.vcx_copy_kernel(dst, src, cb) == 0
        pyvc.dev_free(src)
        pyvc.dev_free(dst)
    pyvc.dev_free(dout)
    pyvc.dev_free(dvol)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
