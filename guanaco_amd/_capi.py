"""ctypes binding of the C ABI in include/guanaco_hip.h (libguanaco_hip.so).

The library must be built in-tree (``python -m guanaco_amd._build``); there
is no fallback: loading fails loudly when it is missing.
"""
import ctypes
import os

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libguanaco_hip.so")

GX_OK = 0
GX_EINVAL = 1
GX_ENOMEM = 2
GX_EDEVICE = 3
GX_EUNSUPPORTED = 4

STORE_F32 = 0
STORE_F16 = 1
STORE_BF16 = 2
STORAGE = {"f32": STORE_F32, "f16": STORE_F16, "bf16": STORE_BF16}
BP_DIRECT = 0
BP_STAGED = 1
BACKPROJECTOR = {"direct": BP_DIRECT, "staged": BP_STAGED}

INPUT_PROJECTIONS = 0
INPUT_CORRECTED = 1
INPUT_SINOGRAMS = 2

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_int = ctypes.c_int
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


class GxError(RuntimeError):
    """A non-zero status from the C ABI (the reference raises RuntimeError
    from guanaco::Error)."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class ReconCfg(ctypes.Structure):
    _fields_ = [
        ("num_pixels", _sz), ("num_angles", _sz), ("num_defocus", _sz),
        ("grid_height", _sz), ("grid_width", _sz),
        ("pixel_size", ctypes.c_float), ("centre", ctypes.c_float),
        ("min_defocus", ctypes.c_float), ("max_defocus", ctypes.c_float),
        ("angles", _f32p), ("gpu_index", _int),
    ]


class VolumeDesc(ctypes.Structure):
    _fields_ = [
        ("num_pixels", _sz), ("num_rows", _sz), ("num_angles", _sz),
        ("num_defocus", _sz), ("depth", _sz), ("width", _sz),
        ("pixel_size", ctypes.c_float), ("centre", ctypes.c_float),
        ("min_defocus", ctypes.c_float), ("max_defocus", ctypes.c_float),
        ("angles", _f32p), ("weights", _f32p),
        ("storage", _int), ("input_layout", _int),
        ("defocus", _f64p),
        ("wavelength", ctypes.c_double), ("Cs", ctypes.c_double),
        ("Ca", ctypes.c_double), ("Pa", ctypes.c_double),
        ("phi", ctypes.c_double),
        ("tilt_chunk", _sz), ("gpu_index", _int),
        ("centres", _f32p), ("band_rows", _sz),
    ]


class SynthDesc(ctypes.Structure):
    _fields_ = [
        ("num_pixels", _sz), ("num_rows", _sz), ("num_angles", _sz),
        ("depth", _sz), ("width", _sz), ("centre", ctypes.c_float),
        ("angles", _f32p), ("num_ellipsoids", _sz),
        ("seed", ctypes.c_uint64),
        ("semi_min", ctypes.c_float), ("semi_max", ctypes.c_float),
        ("density_min", ctypes.c_float), ("density_max", ctypes.c_float),
        ("noise", _int), ("noise_level", ctypes.c_float),
        ("contrast", ctypes.c_float), ("ctf", _int),
        ("wavelength", ctypes.c_double), ("Cs", ctypes.c_double),
        ("defocus", ctypes.c_double), ("phi", ctypes.c_double),
        ("pixel_size", ctypes.c_double),
    ]


class ServerCfg(ctypes.Structure):
    _fields_ = [("budget_bytes", _sz), ("max_inflight", _int),
                ("backfill", _int), ("gpu_index", _int)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("state", _int), ("submit_s", ctypes.c_double),
                ("start_s", ctypes.c_double), ("end_s", ctypes.c_double),
                ("device_ms", ctypes.c_double), ("updates", ctypes.c_double),
                ("footprint_bytes", _sz), ("volume", _vp),
                ("error", ctypes.c_char * 256)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", _sz), ("completed", _sz), ("failed", _sz),
                ("cancelled", _sz), ("queued", _sz), ("running", _sz),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("updates_per_s", ctypes.c_double),
                ("span_s", ctypes.c_double), ("budget_bytes", _sz),
                ("used_bytes", _sz), ("peak_bytes", _sz)]


class CtfDesc(ctypes.Structure):
    _fields_ = [("width", _sz), ("height", _sz), ("num_defocus", _sz),
                ("pixel_size", ctypes.c_double),
                ("defocus", ctypes.POINTER(ctypes.c_double)),
                ("wavelength", ctypes.c_double), ("Cs", ctypes.c_double),
                ("Ca", ctypes.c_double), ("Pa", ctypes.c_double),
                ("dd", ctypes.c_double), ("theta_c", ctypes.c_double),
                ("phi", ctypes.c_double), ("envelope", _int),
                ("recentre", _int), ("gpu_index", _int)]


class TuneEntry(ctypes.Structure):
    _fields_ = [("kind", _int), ("rows", _int), ("filter_ms", ctypes.c_float),
                ("bp_ms", ctypes.c_float)]


TUNE_MAX = 8


class TuneResult(ctypes.Structure):
    _fields_ = [("kind", _int), ("rows", _int), ("n", _int),
                ("cand", TuneEntry * TUNE_MAX)]


# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "gx_last_error": [],
    "gx_version": [],
    "gx_device_count": [ctypes.POINTER(_int)],
    "gx_set_device": [_int],
    "gx_get_device": [ctypes.POINTER(_int)],
    "gx_device_info": [_int, ctypes.c_char_p, _sz, ctypes.POINTER(_sz),
                       ctypes.POINTER(_int)],
    "gx_mem_get_info": [ctypes.POINTER(_sz), ctypes.POINTER(_sz)],
    "gx_malloc": [ctypes.POINTER(_vp), _sz],
    "gx_free": [_vp],
    "gx_memcpy_htod": [_vp, _vp, _sz, _vp],
    "gx_memcpy_dtoh": [_vp, _vp, _sz, _vp],
    "gx_memcpy_dtod": [_vp, _vp, _sz, _vp],
    "gx_fill_u32": [_vp, ctypes.c_uint32, _sz, _vp],
    "gx_copy_u32": [_vp, _vp, _sz, _vp],
    "gx_stream_create": [ctypes.POINTER(_vp)],
    "gx_stream_destroy": [_vp],
    "gx_stream_synchronize": [_vp],
    "gx_device_synchronize": [],
    "gx_event_create": [ctypes.POINTER(_vp)],
    "gx_event_destroy": [_vp],
    "gx_event_record": [_vp, _vp],
    "gx_event_synchronize": [_vp],
    "gx_event_elapsed_ms": [_vp, _vp, ctypes.POINTER(ctypes.c_float)],
    "gx_graph_capture_begin": [_vp],
    "gx_graph_capture_end": [_vp, ctypes.POINTER(_vp)],
    "gx_graph_launch": [_vp, _vp],
    "gx_graph_destroy": [_vp],
    "gx_tilt_geometry": [ctypes.POINTER(ReconCfg), _f32p],
    "gx_recon": [ctypes.POINTER(ReconCfg), _f32p, _f32p],
    "gx_recon_noalloc": [ctypes.POINTER(ReconCfg), _vp, _vp, _vp],
    "gx_project": [ctypes.POINTER(ReconCfg), _f32p, _f32p],
    "gx_filter": [_f32p, _f32p, _sz, _sz, _int],
    "gx_corr": [_f32p, _f32p, _f32p, _sz, _sz, _sz, _int],
    "gx_volume_plan_bytes": [ctypes.POINTER(VolumeDesc), ctypes.POINTER(_sz)],
    "gx_volume_plan_create": [ctypes.POINTER(VolumeDesc), ctypes.POINTER(_vp)],
    "gx_volume_plan_destroy": [_vp],
    "gx_volume_run": [_vp, _vp, _vp, _vp],
    "gx_volume_filter": [_vp, _vp, _vp],
    "gx_volume_backproject": [_vp, _vp, _vp],
    "gx_volume_backproject_rows": [_vp, _sz, _sz, _vp, _vp],
    "gx_volume_filter_rows": [_vp, _vp, _sz, _sz, _vp],
    "gx_volume_plan_band": [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_sz),
                            ctypes.POINTER(_sz)],
    "gx_volume_rows_per_block": [_vp, ctypes.POINTER(_sz)],
    "gx_volume_plan_set_backprojector": [_vp, _int],
    "gx_volume_plan_get_backprojector": [_vp, ctypes.POINTER(_int),
                                         ctypes.POINTER(_int)],
    "gx_volume_plan_slab_range": [_vp, ctypes.POINTER(_int),
                                  ctypes.POINTER(_int)],
    "gx_volume_slabs": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_sz),
                        ctypes.POINTER(_int)],
    "gx_volume_plan_autotune": [_vp, _vp, _vp, _vp,
                                ctypes.POINTER(TuneResult)],
    "gx_volume_plan_get_tile": [_vp, ctypes.POINTER(_int),
                                ctypes.POINTER(_int)],
    "gx_volume_plan_set_tile": [_vp, _int, _int],
    "gx_tune_save": [ctypes.c_char_p],
    "gx_tune_load": [ctypes.c_char_p],
    "gx_tune_clear": [],
    "gx_synth_projections": [ctypes.POINTER(SynthDesc), _vp, _vp],
    "gx_ctf_grid": [ctypes.POINTER(CtfDesc), _vp],
    "gx_server_create": [ctypes.POINTER(ServerCfg), ctypes.POINTER(_vp)],
    "gx_server_destroy": [_vp],
    "gx_server_request_bytes": [ctypes.POINTER(VolumeDesc), _int,
                                ctypes.POINTER(_sz)],
    "gx_server_submit": [_vp, ctypes.POINTER(VolumeDesc), _vp, _int,
                         ctypes.POINTER(ctypes.c_uint64)],
    "gx_server_wait": [_vp, ctypes.c_uint64, ctypes.c_double,
                       ctypes.POINTER(RequestInfo)],
    "gx_server_info": [_vp, ctypes.c_uint64, ctypes.POINTER(RequestInfo)],
    "gx_server_cancel": [_vp, ctypes.c_uint64],
    "gx_server_release": [_vp, ctypes.c_uint64],
    "gx_server_get_stats": [_vp, ctypes.POINTER(ServerStats)],
}
_RESTYPE = {"gx_last_error": ctypes.c_char_p}

_lib = None


def lib():
    """Load libguanaco_hip.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libguanaco_hip.so is not built (run python -m "
                "guanaco_amd._build); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, _int)
        _lib = L
    return _lib


def check(rc):
    if rc != GX_OK:
        msg = lib().gx_last_error().decode(errors="replace")
        raise GxError(rc, msg or "guanaco_hip error %d" % rc)
    return rc


def call(name, *args):
    return check(getattr(lib(), name)(*args))


def device_count():
    n = _int(0)
    call("gx_device_count", ctypes.byref(n))
    return n.value


def device_info(dev=0):
    name = ctypes.create_string_buffer(256)
    mem = _sz(0)
    cus = _int(0)
    call("gx_device_info", dev, name, 256, ctypes.byref(mem), ctypes.byref(cus))
    return {"name": name.value.decode(), "hbm_bytes": mem.value,
            "cus": cus.value}


def f32p(a):
    assert a.dtype == numpy.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_f32p)


def f64p(a):
    assert a.dtype == numpy.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_f64p)


def recon_cfg(N, T, ndf, H, W, pixel_size, centre, min_defocus, max_defocus,
              angles, gpu_index=-1):
    angles = numpy.ascontiguousarray(angles, dtype=numpy.float32)
    cfg = ReconCfg(N, T, ndf, H, W, float(pixel_size), float(centre),
                   float(min_defocus), float(max_defocus), f32p(angles),
                   gpu_index)
    cfg._keep = angles
    return cfg


# ---- device buffers -----------------------------------------------------


class DeviceArray:
    """Owning handle of a device allocation with numpy-side shape/dtype."""

    def __init__(self, shape, dtype=numpy.float32):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = numpy.dtype(dtype)
        self.nbytes = int(numpy.prod(self.shape, dtype=numpy.int64)) * \
            self.dtype.itemsize
        p = _vp()
        call("gx_malloc", ctypes.byref(p), self.nbytes)
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, stream=None):
        a = numpy.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        call("gx_memcpy_htod", d.ptr, a.ctypes.data, a.nbytes, stream)
        if stream is not None:
            call("gx_stream_synchronize", stream)
        return d

    def to_host(self, stream=None):
        out = numpy.empty(self.shape, self.dtype)
        call("gx_memcpy_dtoh", out.ctypes.data, self.ptr, self.nbytes, stream)
        if stream is not None:
            call("gx_stream_synchronize", stream)
        return out

    def free(self):
        if self.ptr:
            call("gx_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def mem_free():
    """Free device memory in bytes."""
    f, tot = _sz(), _sz()
    call("gx_mem_get_info", ctypes.byref(f), ctypes.byref(tot))
    return f.value


class DeviceView:
    """Non-owning view of device memory (e.g. a slice of a DeviceArray)."""

    def __init__(self, ptr, shape, dtype=numpy.float32):
        self.ptr = int(ptr)
        self.shape = tuple(int(s) for s in shape)
        self.dtype = numpy.dtype(dtype)
        self.nbytes = int(numpy.prod(self.shape, dtype=numpy.int64)) * \
            self.dtype.itemsize

    def to_host(self, stream=None):
        out = numpy.empty(self.shape, self.dtype)
        call("gx_memcpy_dtoh", out.ctypes.data, self.ptr, self.nbytes, stream)
        if stream is not None:
            call("gx_stream_synchronize", stream)
        return out


class Stream:
    def __init__(self):
        p = _vp()
        call("gx_stream_create", ctypes.byref(p))
        self.handle = p.value

    def synchronize(self):
        call("gx_stream_synchronize", self.handle)

    def destroy(self):
        if self.handle:
            call("gx_stream_destroy", self.handle)
            self.handle = None


class Event:
    def __init__(self):
        p = _vp()
        call("gx_event_create", ctypes.byref(p))
        self.handle = p.value

    def record(self, stream=None):
        call("gx_event_record", self.handle,
             stream.handle if isinstance(stream, Stream) else stream)

    def synchronize(self):
        call("gx_event_synchronize", self.handle)

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        call("gx_event_elapsed_ms", self.handle, end.handle, ctypes.byref(ms))
        return ms.value

    def destroy(self):
        if self.handle:
            call("gx_event_destroy", self.handle)
            self.handle = None


# ---- whole-volume plans -------------------------------------------------


def volume_desc(num_pixels, num_rows, angles, depth, width=None,
                num_defocus=1, pixel_size=1.0, centre=None, min_defocus=0.0,
                max_defocus=0.0, weights=None, storage="f32",
                input_layout=INPUT_PROJECTIONS, defocus=None, wavelength=0.0,
                Cs=0.0, Ca=0.0, Pa=0.0, phi=0.0, tilt_chunk=0, gpu_index=-1,
                band_rows=0):
    """A gx_volume_desc and the arrays it points to (keep them alive).
    band_rows > 0: the plan holds the filtered slabs of that many detector
    rows at a time (rounded up to a multiple of 32) and runs band by band."""
    angles = numpy.ascontiguousarray(angles, dtype=numpy.float32)
    T = angles.size
    weights = None if weights is None else numpy.ascontiguousarray(
        weights, dtype=numpy.float32)
    defocus = None if defocus is None else numpy.ascontiguousarray(
        defocus, dtype=numpy.float64)
    width = num_pixels if width is None else width
    centre = num_pixels / 2.0 if centre is None else centre
    centres = None
    if numpy.size(centre) > 1:   # one centre per row
        centres = numpy.ascontiguousarray(centre, dtype=numpy.float32)
        assert centres.size == num_rows, "one centre per row"
        centre = centres[0]
    desc = VolumeDesc(
        num_pixels, num_rows, T, num_defocus, depth, width,
        float(pixel_size), float(centre), float(min_defocus),
        float(max_defocus), f32p(angles),
        None if weights is None else f32p(weights),
        STORAGE[storage] if isinstance(storage, str) else storage,
        input_layout, None if defocus is None else f64p(defocus),
        float(wavelength), float(Cs), float(Ca), float(Pa), float(phi),
        tilt_chunk, gpu_index, None if centres is None else f32p(centres),
        int(band_rows))
    return desc, (angles, weights, defocus, centres)


class VolumePlan:
    """A device-resident reconstruction plan (gx_volume_plan)."""

    def __init__(self, num_pixels, num_rows, angles, depth, width=None,
                 num_defocus=1, pixel_size=1.0, centre=None, min_defocus=0.0,
                 max_defocus=0.0, weights=None, storage="f32",
                 input_layout=INPUT_PROJECTIONS, defocus=None,
                 wavelength=0.0, Cs=0.0, Ca=0.0, Pa=0.0, phi=0.0,
                 tilt_chunk=0, gpu_index=-1, band_rows=0):
        self.desc, self._keep = volume_desc(
            num_pixels, num_rows, angles, depth, width, num_defocus,
            pixel_size, centre, min_defocus, max_defocus, weights, storage,
            input_layout, defocus, wavelength, Cs, Ca, Pa, phi, tilt_chunk,
            gpu_index, band_rows)
        width = num_pixels if width is None else width
        self.shape = (num_rows, depth, width)
        self.storage = storage
        self.num_angles = len(angles)
        p = _vp()
        call("gx_volume_plan_create", ctypes.byref(self.desc), ctypes.byref(p))
        self.handle = p.value

    def footprint(self):
        n = _sz()
        call("gx_volume_plan_bytes", ctypes.byref(self.desc), ctypes.byref(n))
        return n.value

    def run(self, d_in, d_vol, stream=None):
        call("gx_volume_run", self.handle, d_in.ptr, d_vol.ptr,
             stream.handle if isinstance(stream, Stream) else stream)

    def filter(self, d_in, stream=None):
        call("gx_volume_filter", self.handle, d_in.ptr,
             stream.handle if isinstance(stream, Stream) else stream)

    def backproject(self, d_vol, stream=None):
        call("gx_volume_backproject", self.handle, d_vol.ptr,
             stream.handle if isinstance(stream, Stream) else stream)

    def filter_rows(self, d_in, y0, y1, stream=None):
        """Filter detector rows [y0, y1) into the slabs (one band)."""
        call("gx_volume_filter_rows", self.handle, d_in.ptr, int(y0), int(y1),
             stream.handle if isinstance(stream, Stream) else stream)

    def backproject_rows(self, y0, y1, d_vol, stream=None):
        """Backproject rows [y0, y1) of the filtered band into d_vol,
        which holds those rows."""
        call("gx_volume_backproject_rows", self.handle, int(y0), int(y1),
             d_vol.ptr, stream.handle if isinstance(stream, Stream)
             else stream)

    def band(self):
        """(rows held per band, first and end row of the band held)."""
        b, y0, y1 = _sz(), _sz(), _sz()
        call("gx_volume_plan_band", self.handle, ctypes.byref(b),
             ctypes.byref(y0), ctypes.byref(y1))
        return b.value, y0.value, y1.value

    def set_backprojector(self, kind):
        """BP_DIRECT or BP_STAGED (the default where a tile spans <= 4
        defocus slabs)."""
        call("gx_volume_plan_set_backprojector", self.handle,
             BACKPROJECTOR[kind] if isinstance(kind, str) else kind)

    def backprojector(self):
        """(kind, slabs held per tilt by the staged kernel; 0 = none)."""
        k = _int()
        n = _int()
        call("gx_volume_plan_get_backprojector", self.handle, ctypes.byref(k),
             ctypes.byref(n))
        return k.value, n.value

    def tile(self):
        """(rows per backprojection tile, backprojector kind)."""
        r, k = _int(), _int()
        call("gx_volume_plan_get_tile", self.handle, ctypes.byref(r),
             ctypes.byref(k))
        return r.value, k.value

    def set_tile(self, rows, kind):
        """One of the autotuner's candidates, by hand (takes effect with the
        next filter pass)."""
        call("gx_volume_plan_set_tile", self.handle, int(rows),
             BACKPROJECTOR[kind] if isinstance(kind, str) else kind)

    def autotune(self, d_in, d_vol, stream=None):
        """Times every tile candidate on this plan's data and keeps the
        fastest (remembered for later plans of this shape); d_vol holds the
        reconstruction afterwards. Returns the choice and the timings."""
        r = TuneResult()
        call("gx_volume_plan_autotune", self.handle, d_in.ptr, d_vol.ptr,
             stream.handle if isinstance(stream, Stream) else stream,
             ctypes.byref(r))
        names = {BP_DIRECT: "direct", BP_STAGED: "staged"}
        return {"backprojector": names[r.kind], "rows": r.rows,
                "candidates": [
                    {"backprojector": names[c.kind], "rows": c.rows,
                     "filter_ms": round(c.filter_ms, 4),
                     "bp_ms": round(c.bp_ms, 4)} for c in r.cand[:r.n]]}

    def slab_range(self):
        """(first, last): per tilt, the defocus slabs the filter computes
        (those some voxel of the volume reads)."""
        import numpy
        T = self.num_angles
        first = numpy.zeros(T, numpy.int32)
        last = numpy.zeros(T, numpy.int32)
        call("gx_volume_plan_slab_range", self.handle,
             first.ctypes.data_as(ctypes.POINTER(_int)),
             last.ctypes.data_as(ctypes.POINTER(_int)))
        return first, last

    def slabs(self):
        p = _vp()
        n = _sz()
        yb = _int()
        call("gx_volume_slabs", self.handle, ctypes.byref(p), ctypes.byref(n),
             ctypes.byref(yb))
        return p.value, n.value, yb.value

    def destroy(self):
        if getattr(self, "handle", None):
            call("gx_volume_plan_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def ctf_grid(width, height, pixel_size, defocus, wavelength, Cs=0.0, Ca=0.0,
             Pa=0.0, dd=0.0, theta_c=0.0, phi=0.0, envelope=True,
             recentre=False, out=None, gpu_index=-1):
    """complex64 (ndf, height, width) CTF images computed on the device
    (gx_ctf_grid); units as the reference's CTF: A and rad."""
    df = numpy.ascontiguousarray(numpy.atleast_1d(defocus), numpy.float64)
    shape = (df.size, int(height), int(width))
    if out is None:
        out = numpy.empty(shape, numpy.complex64)
    assert out.shape == shape and out.dtype == numpy.complex64 \
        and out.flags.c_contiguous and out.flags.writeable
    d = CtfDesc(int(width), int(height), df.size, float(pixel_size),
                df.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                float(wavelength), float(Cs), float(Ca), float(Pa), float(dd),
                float(theta_c), float(phi), 1 if envelope else 0,
                1 if recentre else 0, int(gpu_index))
    call("gx_ctf_grid", ctypes.byref(d), out.ctypes.data)
    return out


def tune_save(path):
    """Writes the remembered tile choices to `path` (text)."""
    call("gx_tune_save", os.fsencode(path))


def tune_load(path):
    """Merges the tile choices in `path` into the remembered ones."""
    call("gx_tune_load", os.fsencode(path))


def tune_clear():
    call("gx_tune_clear")


def plan_bytes(**kw):
    """Device footprint of a plan with these arguments (admission)."""
    angles = numpy.ascontiguousarray(kw.pop("angles"), dtype=numpy.float32)
    defocus = kw.pop("defocus", None)
    if defocus is not None:
        defocus = numpy.ascontiguousarray(defocus, dtype=numpy.float64)
    storage = kw.pop("storage", "f32")
    d = VolumeDesc(
        kw["num_pixels"], kw["num_rows"], angles.size,
        kw.get("num_defocus", 1), kw["depth"],
        kw.get("width", kw["num_pixels"]), float(kw.get("pixel_size", 1.0)),
        float(kw.get("centre", kw["num_pixels"] / 2.0)),
        float(kw.get("min_defocus", 0.0)), float(kw.get("max_defocus", 0.0)),
        f32p(angles), None,
        STORAGE[storage] if isinstance(storage, str) else storage,
        kw.get("input_layout", INPUT_PROJECTIONS),
        None if defocus is None else f64p(defocus),
        0.0, 0.0, 0.0, 0.0, 0.0, kw.get("tilt_chunk", 0), -1, None,
        int(kw.get("band_rows", 0)))
    n = _sz()
    call("gx_volume_plan_bytes", ctypes.byref(d), ctypes.byref(n))
    return n.value


def synth_projections(d_out, num_pixels, num_rows, angles, depth=0, width=0,
                      centre=None, num_ellipsoids=20, seed=1000,
                      semi=(5.0, 40.0), density=(0.1, 1.0), noise=None,
                      noise_level=0.0, contrast=0.2, ctf=None, stream=None):
    """Fill d_out (T, Y, X) with a seeded synthetic tilt series.

    noise: None, "gaussian" (sigma = noise_level * max) or "poisson" (mean
    noise_level electrons per pixel). ctf: None or a dict with wavelength,
    Cs, defocus (A), phi (rad) and pixel_size (A).
    """
    angles = numpy.ascontiguousarray(angles, dtype=numpy.float32)
    mode = {None: 0, "gaussian": 1, "poisson": 2}[noise]
    c = ctf or {}
    d = SynthDesc(num_pixels, num_rows, angles.size, depth, width,
                  float(num_pixels / 2.0 if centre is None else centre),
                  f32p(angles), num_ellipsoids, seed, float(semi[0]),
                  float(semi[1]), float(density[0]), float(density[1]), mode,
                  float(noise_level), float(contrast), 1 if ctf else 0,
                  float(c.get("wavelength", 0.0)), float(c.get("Cs", 0.0)),
                  float(c.get("defocus", 0.0)), float(c.get("phi", 0.0)),
                  float(c.get("pixel_size", 1.0)))
    call("gx_synth_projections", ctypes.byref(d), d_out.ptr,
         stream.handle if isinstance(stream, Stream) else stream)
