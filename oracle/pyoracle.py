"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and for the
reference builds under oracle/_ref/ (made by oracle/ref.mk).

TEST INFRASTRUCTURE ONLY: imported by tests/, tests/golden/, __graft_entry__
.smoke() and bench.py's cpu_baseline leg, as the checker.  Never imported by
the product library.
"""
import ctypes as C
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
A3 = 32768

_u8p = C.POINTER(C.c_uint8)


def _ptr(a):
    return a.ctypes.data_as(_u8p)


class Oracle:
    def __init__(self, path=None):
        path = path or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = self.L = C.CDLL(path)
        vp, sz = C.c_void_p, C.c_size_t
        L.vco_tables.argtypes = [vp, vp, vp, vp]
        L.vco_dq_tables.argtypes = [C.c_float, vp, vp]
        L.vco_fdct.argtypes = [vp, C.c_int32, vp]
        L.vco_quantize.argtypes = [vp, C.c_float, vp]
        L.vco_dequantize.argtypes = [vp, C.c_float, vp]
        L.vco_quant_levels.argtypes = [C.c_float, C.c_int, vp, sz, vp]
        L.vco_idct.argtypes = [vp, C.c_int32, vp]
        L.vco_entropy_encode.argtypes = [vp, vp, sz]
        L.vco_entropy_encode.restype = C.c_long
        L.vco_entropy_decode.argtypes = [vp, sz, vp, C.POINTER(C.c_uint32)]
        L.vco_encode_atom.argtypes = [vp, C.c_float, vp, sz, C.POINTER(C.c_int)]
        L.vco_encode_atom.restype = C.c_long
        L.vco_decode_payload.argtypes = [vp, C.c_uint32, C.c_float, vp]
        L.vco_downsample2x.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, vp]
        L.vco_nlod.argtypes = [C.c_uint32] * 3
        L.vco_gather_atom.argtypes = [vp] + [C.c_uint32] * 6 + [vp]
        L.vco_encode_q.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, vp,
                                   C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]
        L.vco_encode.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
                                 C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]
        L.vco_rate_control.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float]
        L.vco_rate_control.restype = C.c_float
        L.vco_lod_schedule.argtypes = [C.c_float, C.c_int, vp]
        L.vco_open.argtypes = [vp, sz]
        L.vco_open.restype = vp
        L.vco_close.argtypes = [vp]
        L.vco_decode_atom.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]
        L.vco_coverage.argtypes = [vp, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]
        L.vco_decode_region.argtypes = [vp, C.c_int, vp, vp]
        L.vco_decode_lod.argtypes = [vp, C.c_int, vp, vp]
        L.vco_decode_atoms_omp.argtypes = [vp, C.c_int, vp, sz, C.c_int, vp]
        L.vco_decode_atoms_omp.restype = C.c_double
        L.vco_synth.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, vp]
        L.vco_repack.argtypes = [vp, sz, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]
        L.vco_ds2x.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_float, vp]
        self.libc = C.CDLL("libc.so.6")
        self.libc.free.argtypes = [C.c_void_p]

    # ---- tables ----
    def tables(self):
        cm = np.zeros((32, 32), np.int32)
        gs = np.zeros(94, np.float32)
        scan = np.zeros(A3, np.uint16)
        band = np.zeros(A3, np.uint8)
        self.L.vco_tables(cm.ctypes.data, gs.ctypes.data, scan.ctypes.data, band.ctypes.data)
        return cm, gs, scan, band

    def dq_tables(self, q):
        st = np.zeros(94, np.int32)
        bi = np.zeros(94, np.int32)
        self.L.vco_dq_tables(C.c_float(q), st.ctypes.data, bi.ctypes.data)
        return st, bi

    # ---- stages ----
    def fdct(self, vox, dc):
        vox = np.ascontiguousarray(vox, np.uint8).ravel()
        out = np.zeros(A3, np.int32)
        self.L.vco_fdct(vox.ctypes.data, int(dc), out.ctypes.data)
        return out

    def quantize(self, coef, q):
        coef = np.ascontiguousarray(coef, np.int32).ravel()
        out = np.zeros(A3, np.int16)
        self.L.vco_quantize(coef.ctypes.data, C.c_float(q), out.ctypes.data)
        return out

    def quant_levels(self, q, freq, coef):
        coef = np.ascontiguousarray(coef, np.int32)
        out = np.zeros(coef.size, np.int32)
        self.L.vco_quant_levels(C.c_float(q), freq, coef.ctypes.data, coef.size, out.ctypes.data)
        return out

    def dequantize(self, lv, q):
        lv = np.ascontiguousarray(lv, np.int16).ravel()
        out = np.zeros(A3, np.int32)
        self.L.vco_dequantize(lv.ctypes.data, C.c_float(q), out.ctypes.data)
        return out

    def idct(self, coef, dc):
        coef = np.ascontiguousarray(coef, np.int32).ravel()
        out = np.zeros(A3, np.uint8)
        self.L.vco_idct(coef.ctypes.data, int(dc), out.ctypes.data)
        return out

    def entropy_encode(self, lv):
        lv = np.ascontiguousarray(lv, np.int16).ravel()
        buf = np.zeros(A3 * 4 + 64, np.uint8)
        n = self.L.vco_entropy_encode(lv.ctypes.data, buf.ctypes.data, buf.size)
        return None if n < 0 else buf[:n].tobytes()

    def entropy_decode(self, rc):
        b = np.frombuffer(rc, np.uint8) if len(rc) else np.zeros(1, np.uint8)
        out = np.zeros(A3, np.int16)
        eob = C.c_uint32(0)
        st = self.L.vco_entropy_decode(b.ctypes.data, len(rc), out.ctypes.data, C.byref(eob))
        return st, out, eob.value

    def encode_atom(self, vox, q):
        vox = np.ascontiguousarray(vox, np.uint8).ravel()
        buf = np.zeros(A3 * 4 + 65, np.uint8)
        z = C.c_int(0)
        n = self.L.vco_encode_atom(vox.ctypes.data, C.c_float(q), buf.ctypes.data, buf.size, C.byref(z))
        if z.value:
            return b""
        if n < 0:
            raise OverflowError("range coder overflow")
        return buf[:n].tobytes()

    def decode_payload(self, pay, q):
        b = np.frombuffer(pay, np.uint8)
        out = np.zeros(A3, np.uint8)
        st = self.L.vco_decode_payload(b.ctypes.data, len(pay), C.c_float(q), out.ctypes.data)
        return st, out

    # ---- volumes ----
    def downsample2x(self, vol):
        nz, ny, nx = vol.shape
        vol = np.ascontiguousarray(vol, np.uint8)
        out = np.zeros(((nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2), np.uint8)
        self.L.vco_downsample2x(vol.ctypes.data, nx, ny, nz, out.ctypes.data)
        return out

    def ds2x(self, vol, method, alpha):
        """the exporter's within-cell 2x downscale (tools/downscale.h); method 0 box, 1 cbox"""
        nz, ny, nx = vol.shape
        vol = np.ascontiguousarray(vol, np.uint8)
        out = np.zeros(((nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2), np.uint8)
        self.L.vco_ds2x(vol.ctypes.data, nx, ny, nz, method, C.c_float(alpha), out.ctypes.data)
        return out

    def nlod(self, nx, ny, nz):
        return self.L.vco_nlod(nx, ny, nz)

    def synth(self, seed, nx, ny, nz, nsheets):
        out = np.zeros((nz, ny, nx), np.uint8)
        self.L.vco_synth(C.c_uint64(seed), nx, ny, nz, nsheets, out.ctypes.data)
        return out

    def _take(self, p, n):
        b = C.string_at(p, n)
        self.libc.free(p)
        return b

    def encode_q(self, vol, target_ratio, qs):
        nz, ny, nx = vol.shape
        vol = np.ascontiguousarray(vol, np.uint8)
        q8 = np.zeros(8, np.float32)
        q8[:len(qs)] = qs
        p, n = C.c_void_p(), C.c_size_t()
        st = self.L.vco_encode_q(vol.ctypes.data, nx, ny, nz, C.c_float(target_ratio),
                                 q8.ctypes.data, C.byref(p), C.byref(n))
        return st, (self._take(p, n.value) if st == 0 else None)

    def encode(self, vol, target_ratio):
        nz, ny, nx = vol.shape
        vol = np.ascontiguousarray(vol, np.uint8)
        p, n = C.c_void_p(), C.c_size_t()
        st = self.L.vco_encode(vol.ctypes.data, nx, ny, nz, C.c_float(target_ratio),
                               C.byref(p), C.byref(n))
        return st, (self._take(p, n.value) if st == 0 else None)

    def rate_control(self, vol, target_ratio):
        nz, ny, nx = vol.shape
        vol = np.ascontiguousarray(vol, np.uint8)
        return self.L.vco_rate_control(vol.ctypes.data, nx, ny, nz, C.c_float(target_ratio))

    def lod_schedule(self, q0, nlod):
        qs = np.zeros(8, np.float32)
        self.L.vco_lod_schedule(C.c_float(q0), nlod, qs.ctypes.data)
        return qs

    def repack(self, buf):
        """canonical v2 image of an archive (tools/vc_repack.c): (status, bytes or None)"""
        b = np.frombuffer(bytes(buf), np.uint8).copy()
        p, n = C.c_void_p(), C.c_size_t()
        st = self.L.vco_repack(b.ctypes.data, b.size, C.byref(p), C.byref(n))
        return st, (self._take(p, n.value) if st == 0 else None)

    def open(self, buf):
        return OracleArchive(self, buf)


class OracleArchive:
    def __init__(self, o, buf):
        self.o = o
        self.arr = np.frombuffer(bytes(buf), np.uint8).copy()
        self.h = o.L.vco_open(self.arr.ctypes.data, self.arr.size)
        if not self.h:
            raise ValueError("vco_open rejected the archive")

    def close(self):
        if self.h:
            self.o.L.vco_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def decode_atom(self, lod, ax, ay, az):
        out = np.zeros(A3, np.uint8)
        st = self.o.L.vco_decode_atom(self.h, lod, ax, ay, az, out.ctypes.data)
        return st, out.reshape(32, 32, 32)

    def coverage(self, lod, az, ay, ax):
        off, ln = C.c_uint64(), C.c_uint32()
        c = self.o.L.vco_coverage(self.h, lod, az, ay, ax, C.byref(off), C.byref(ln))
        return c, off.value, ln.value

    def decode_region(self, lod, box):
        x0, y0, z0, x1, y1, z1 = box
        out = np.zeros((z1 - z0, y1 - y0, x1 - x0), np.uint8)
        b = np.array(box, np.uint32)
        st = self.o.L.vco_decode_region(self.h, lod, b.ctypes.data, out.ctypes.data)
        return st, out

    def decode_lod(self, lod, dims):
        nx, ny, nz = dims
        out = np.zeros((nz, ny, nx), np.uint8)
        d = np.zeros(3, np.uint32)
        st = self.o.L.vco_decode_lod(self.h, lod, out.ctypes.data, d.ctypes.data)
        return st, out

    def decode_atoms_omp(self, lod, atoms_zyx, nthreads, keep=False):
        atoms = np.ascontiguousarray(atoms_zyx, np.uint32)
        n = atoms.shape[0]
        out = np.zeros((n, A3), np.uint8) if keep else None
        secs = self.o.L.vco_decode_atoms_omp(self.h, lod, atoms.ctypes.data, n, nthreads,
                                             out.ctypes.data if keep else None)
        return secs, out


# ---------------------------------------------------------------------------
# Reference build (oracle/_ref): the reference's own vc.c, compiled in place.
# ---------------------------------------------------------------------------
def ds_ref(variant="strict"):
    """tools/downscale.h from the reference tree, built by oracle/ref.mk (None if absent):
    a function (vol, method, alpha) -> downscaled volume"""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", f"libdsref_{variant}.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.ref_ds2x.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_void_p]

    def run(vol, method, alpha):
        nz, ny, nx = vol.shape
        vol = np.ascontiguousarray(vol, np.uint8)
        out = np.zeros(((nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2), np.uint8)
        L.ref_ds2x(vol.ctypes.data, nx, ny, nz, method, C.c_float(alpha), out.ctypes.data)
        return out
    return run


def ref_variants():
    if not os.path.isdir(REF_DIR):
        return {}
    return {f[len("libvcref_"):-3]: os.path.join(REF_DIR, f)
            for f in sorted(os.listdir(REF_DIR)) if f.startswith("libvcref_") and f.endswith(".so")}


def host_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            return " avx512f " in f.read()
    except OSError:
        return False


class _Dims(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32)]


class _Box(C.Structure):
    _fields_ = [("x0", C.c_uint32), ("y0", C.c_uint32), ("z0", C.c_uint32),
                ("x1", C.c_uint32), ("y1", C.c_uint32), ("z1", C.c_uint32)]


class Ref:
    """The reference library (one build variant)."""

    def __init__(self, variant="strict"):
        v = ref_variants()
        if variant not in v:
            raise FileNotFoundError(f"reference variant {variant!r} not built (make -f oracle/ref.mk)")
        self.variant = variant
        L = self.L = C.CDLL(v[variant])
        vp = C.c_void_p
        L.vc_create.argtypes = [C.c_char_p, _Dims, C.c_float]
        L.vc_create.restype = vp
        L.vc_set_base_q.argtypes = [vp, C.c_int, C.c_float]
        L.vc_append_atom.argtypes = [vp, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, vp]
        L.vc_mark_zero_atom.argtypes = [vp, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32]
        L.vc_mark_zero_region.argtypes = [vp, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32]
        L.vc_writer_decode_2x2x2.argtypes = [vp, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, vp,
                                             C.POINTER(C.c_int)]
        L.vc_writer_close.argtypes = [vp]
        L.vc_open.argtypes = [vp, C.c_size_t]
        L.vc_open.restype = vp
        L.vc_close.argtypes = [vp]
        L.vc_decode_atom.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]
        L.vc_decode_region.argtypes = [vp, C.c_int, _Box, vp]
        L.vc_atom_payload_range.argtypes = [vp, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                            C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]
        L.vc_atom_coverage.argtypes = [vp, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32]
        L.vc_lod_dims.argtypes = [vp, C.c_int, C.POINTER(_Dims)]
        L.vc_archive_version.argtypes = [vp]
        L.vc_region_blob_range.argtypes = [vp, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                           C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.vc_test_atom_rt.argtypes = [vp, C.c_float, vp]
        L.vc_test_atom_rt.restype = C.c_uint32
        L.vc_test_downsample2x.argtypes = [vp, _Dims, C.POINTER(_Dims)]
        L.vc_test_downsample2x.restype = vp
        self.libc = C.CDLL("libc.so.6")
        self.libc.free.argtypes = [vp]

    def atom_rt(self, vox, q):
        vox = np.ascontiguousarray(vox, np.uint8).ravel()
        rec = np.zeros(A3, np.uint8)
        n = self.L.vc_test_atom_rt(vox.ctypes.data, C.c_float(q), rec.ctypes.data)
        return n, rec

    def downsample2x(self, vol):
        nz, ny, nx = vol.shape
        vol = np.ascontiguousarray(vol, np.uint8)
        od = _Dims()
        p = self.L.vc_test_downsample2x(vol.ctypes.data, _Dims(nx, ny, nz), C.byref(od))
        out = np.frombuffer(C.string_at(p, od.nx * od.ny * od.nz), np.uint8).reshape(od.nz, od.ny, od.nx).copy()
        self.libc.free(p)
        return out

    def write_v1(self, path, dims, target_ratio, qs, atoms):
        """atoms: iterable of (lod, az, ay, ax, vox or None for mark-zero)."""
        nx, ny, nz = dims
        w = self.L.vc_create(path.encode(), _Dims(nx, ny, nz), C.c_float(target_ratio))
        if not w:
            raise RuntimeError("vc_create failed")
        for l, q in enumerate(qs):
            self.L.vc_set_base_q(w, l, C.c_float(q))
        for lod, az, ay, ax, vox in atoms:
            if vox is None:
                st = self.L.vc_mark_zero_atom(w, lod, az, ay, ax)
            else:
                v = np.ascontiguousarray(vox, np.uint8).ravel()
                st = self.L.vc_append_atom(w, lod, az, ay, ax, v.ctypes.data)
            if st != 0:
                self.L.vc_writer_close(w)
                raise RuntimeError(f"append failed {st}")
        self.L.vc_writer_close(w)

    def export_volume(self, oracle, vol, path, q0=1.0, q_falloff=0.5, method=1, alpha=0.5, tile=0,
                      chunk=(128, 128, 128)):
        """The exporter's single-threaded order (tools/vc_export_stream.c) on a
        local volume whose source chunks are all present, driving this
        reference writer: padded dims (:688-694), base_q schedule (:711-719),
        prepass of regions past the source (:778-790, :338-352), units in
        claim order (:360-367) with the empty-band stamp (:505-522) or the
        plane cascade (push_plane / flush_slab / finish_cube, :395-466), then
        the coarse tail from the writer's own decoded parents (:603-626).
        The downscale is the oracle's restatement of tools/downscale.h."""
        A = 32
        REG = 1024
        nz, ny, nx = vol.shape
        px, py, pz = [(d + REG - 1) // REG * REG for d in (nx, ny, nz)]
        w = self.L.vc_create(path.encode(), _Dims(px, py, pz), C.c_float(1.0))
        if not w:
            raise RuntimeError("vc_create failed")
        nlod = 1
        d = (px, py, pz)
        for l in range(1, 8):
            d = tuple((v + 1) // 2 for v in d)
            nlod = l + 1
            if max(d) <= 1:
                break
        q = np.float32(q0)
        for l in range(nlod):
            if q < np.float32(0.05):
                q = np.float32(0.05)
            self.L.vc_set_base_q(w, l, C.c_float(float(q)))
            q = np.float32(q * np.float32(q_falloff))
        lx, ly, lz = px, py, pz
        for L in range(nlod):
            nrz = ((lz + A - 1) // A + 31) // 32
            nry = ((ly + A - 1) // A + 31) // 32
            nrx = ((lx + A - 1) // A + 31) // 32
            for rz in range(nrz):
                for ry in range(nry):
                    for rx in range(nrx):
                        if (rx * REG << L) >= nx or (ry * REG << L) >= ny or (rz * REG << L) >= nz:
                            self.L.vc_mark_zero_region(w, L, rz, ry, rx)
            lx, ly, lz = (lx + 1) // 2, (ly + 1) // 2, (lz + 1) // 2
        fine = min(nlod - 1, 2)
        band = A << fine
        sb = tile if tile else 1024
        sb = min(sb, A << (nlod - 1))
        sb = (sb + band - 1) // band * band
        # the padded pyramid LOD0..fine
        lv = [np.zeros((pz, py, px), np.uint8)]
        lv[0][:nz, :ny, :nx] = vol
        for l in range(1, fine + 1):
            lv.append(oracle.ds2x(lv[-1], method, alpha))

        def put(L, az, ay, ax, atom):
            a = np.ascontiguousarray(atom)
            if not a.any():
                st = self.L.vc_mark_zero_atom(w, L, az, ay, ax)
            else:
                st = self.L.vc_append_atom(w, L, az, ay, ax, a.ctypes.data)
            if st != 0:
                raise RuntimeError(f"append {L} {az} {ay} {ax}: {st}")

        ntx, nty, nbz = (px + sb - 1) // sb, (py + sb - 1) // sb, (pz + band - 1) // band
        ncx, ncy, ncz = [(d + c - 1) // c for d, c in zip((nx, ny, nz), chunk)]
        for ti in range(ntx * nty * nbz):
            bz, txy = ti % nbz, ti // nbz
            ty, tx = txy // ntx, txy % ntx
            vx0, vy0, vz0 = tx * sb, ty * sb, bz * band
            vx1, vy1, vz1 = min(vx0 + sb, px), min(vy0 + sb, py), min(vz0 + band, pz)
            if vx0 // chunk[0] >= ncx or vy0 // chunk[1] >= ncy or vz0 // chunk[2] >= ncz:
                for L in range(fine + 1):
                    az0, ay0, ax0 = (vz0 >> L) // A, (vy0 >> L) // A, (vx0 >> L) // A
                    for az in range(((vz1 - vz0) >> L) // A):
                        for ay in range(((vy1 - vy0) >> L) // A):
                            for ax in range(((vx1 - vx0) >> L) // A):
                                self.L.vc_mark_zero_atom(w, L, az0 + az, ay0 + ay, ax0 + ax)
                continue
            # the plane cascade: a level's slab flushes after its 32nd plane;
            # every second plane of a level feeds one plane of the next
            slabfill, pairfill, zatom = [0] * 8, [0] * 8, [(vz0 >> L) // A for L in range(8)]

            def push(L):
                slabfill[L] += 1
                if slabfill[L] == A:
                    v = lv[L]
                    z0 = zatom[L] * A
                    for ay in range(((vy1 - vy0) >> L) // A):
                        for ax in range(((vx1 - vx0) >> L) // A):
                            y0, x0 = (vy0 >> L) + ay * A, (vx0 >> L) + ax * A
                            put(L, zatom[L], y0 // A, x0 // A, v[z0:z0 + A, y0:y0 + A, x0:x0 + A])
                    zatom[L] += 1
                    slabfill[L] = 0
                if L + 1 <= fine:
                    pairfill[L] += 1
                    if pairfill[L] == 2:
                        push(L + 1)
                        pairfill[L] = 0
            for _ in range(vz1 - vz0):
                push(0)
        # coarse tail
        out8 = np.zeros((8, A3), np.uint8)
        anyp = C.c_int(0)
        for L in range(fine + 1, nlod):
            dl = [px, py, pz]
            for _ in range(L):
                dl = [(v + 1) // 2 for v in dl]
            dcx, dcy, dcz = [(v + A - 1) // A for v in dl]
            for daz in range(dcz):
                for day in range(dcy):
                    for dax in range(dcx):
                        self.L.vc_writer_decode_2x2x2(w, L - 1, 2 * daz, 2 * day, 2 * dax, out8.ctypes.data,
                                                      C.byref(anyp))
                        if not anyp.value:
                            self.L.vc_mark_zero_atom(w, L, daz, day, dax)
                            continue
                        blk = np.zeros((64, 64, 64), np.uint8)
                        for c in range(8):
                            dz, dy, dx = c >> 2, (c >> 1) & 1, c & 1
                            blk[dz * A:dz * A + A, dy * A:dy * A + A, dx * A:dx * A + A] = out8[c].reshape(A, A, A)
                        put(L, daz, day, dax, oracle.ds2x(blk, method, alpha))
        self.L.vc_writer_close(w)

    def payloads(self, vox_list, q):
        """Exact payload bytes of encode_atom via a one-LOD v1 archive in /tmp."""
        n = len(vox_list)
        fd, path = tempfile.mkstemp(suffix=".vca")
        os.close(fd)
        try:
            self.write_v1(path, (32 * n, 32, 32), 10.0, [q] * 8,
                          [(0, 0, 0, i, v) for i, v in enumerate(vox_list)])
            buf = open(path, "rb").read()
        finally:
            os.unlink(path)
        arc = RefArchive(self, buf)
        out = []
        for i in range(n):
            c, off, ln = arc.payload_range(0, 0, 0, i)
            out.append(buf[off:off + ln] if c == 2 else b"")
        arc.close()
        return out

    def open(self, buf):
        return RefArchive(self, buf)


class RefArchive:
    def __init__(self, ref, buf):
        self.ref = ref
        self.arr = np.frombuffer(bytes(buf), np.uint8).copy()
        self.h = ref.L.vc_open(self.arr.ctypes.data, self.arr.size)
        if not self.h:
            raise ValueError("reference vc_open rejected the archive")

    def close(self):
        if self.h:
            self.ref.L.vc_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def version(self):
        return self.ref.L.vc_archive_version(self.h)

    def lod_dims(self, lod):
        d = _Dims()
        st = self.ref.L.vc_lod_dims(self.h, lod, C.byref(d))
        return st, (d.nx, d.ny, d.nz)

    def decode_atom(self, lod, ax, ay, az):
        out = np.zeros(A3, np.uint8)
        st = self.ref.L.vc_decode_atom(self.h, lod, ax, ay, az, out.ctypes.data)
        return st, out.reshape(32, 32, 32)

    def decode_region(self, lod, box):
        x0, y0, z0, x1, y1, z1 = box
        out = np.zeros((z1 - z0, y1 - y0, x1 - x0), np.uint8)
        st = self.ref.L.vc_decode_region(self.h, lod, _Box(*box), out.ctypes.data)
        return st, out

    def payload_range(self, lod, az, ay, ax):
        off, ln = C.c_uint64(), C.c_uint32()
        c = self.ref.L.vc_atom_payload_range(self.h, lod, az, ay, ax, C.byref(off), C.byref(ln))
        return c, off.value, ln.value

    def coverage(self, lod, az, ay, ax):
        return self.ref.L.vc_atom_coverage(self.h, lod, az, ay, ax)
