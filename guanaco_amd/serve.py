"""Serving loop for whole-volume reconstructions (no reference counterpart).

A Server owns one device allocation (its budget) and runs requests of
different sizes concurrently, one stream each: a request is admitted, in
arrival order, as soon as its footprint (plan, input copy, volume) fits in
the budget and a stream is free; with backfill, later requests that fit may
start while an earlier one waits. Each request's whole path (CTF correction,
ramp filter, backprojection of all defocus slabs) runs on the device with no
allocation or device-wide synchronisation, and its volume stays in server
memory until released. The scheduling loop is native (libguanaco_hip.so,
gx_server_*); this module is its Python face.

    srv = Server(max_inflight=4)
    req = srv.submit(projections, angles, depth=256, pixel_size=2.0,
                     defocus=[20000.0], wavelength=l, Cs=2.7e7, phi=phi,
                     storage="f16")
    vol = req.result()          # numpy (Y, Z, W) float32
    req.release()
    srv.stats()                 # p50/p99 latency, throughput, memory
"""
import ctypes

import numpy

from . import _capi as C

STATES = {0: "queued", 1: "running", 2: "done", 3: "failed", 4: "cancelled"}


class Request:
    """Handle of a submitted request."""

    def __init__(self, server, rid, shape, keep):
        self.server = server
        self.id = rid
        self.shape = shape
        self._keep = keep  # host input and descriptor arrays

    def info(self, wait=False, timeout=None):
        i = C.RequestInfo()
        if wait:
            C.call("gx_server_wait", self.server.handle, self.id,
                   -1.0 if timeout is None else float(timeout),
                   ctypes.byref(i))
        else:
            C.call("gx_server_info", self.server.handle, self.id,
                   ctypes.byref(i))
        return {"state": STATES[i.state], "submit_s": i.submit_s,
                "start_s": i.start_s, "end_s": i.end_s,
                "latency_ms": 1e3 * (i.end_s - i.submit_s)
                if i.state >= 2 else None,
                "device_ms": i.device_ms, "updates": i.updates,
                "footprint_bytes": i.footprint_bytes, "volume_ptr": i.volume,
                "error": i.error.decode(errors="replace")}

    def wait(self, timeout=None):
        """Blocks until the request has finished; returns info()."""
        inf = self.info(wait=True, timeout=timeout)
        if inf["state"] in ("done", "failed", "cancelled"):
            self._keep = None  # the input is no longer read
        return inf

    def result(self, out=None):
        """The volume as a host array (waits for the request)."""
        inf = self.wait()
        if inf["state"] != "done":
            raise RuntimeError("request %d %s: %s" % (
                self.id, inf["state"], inf["error"]))
        if out is None:
            out = numpy.empty(self.shape, numpy.float32)
        assert out.shape == self.shape and out.dtype == numpy.float32 \
            and out.flags.c_contiguous
        C.call("gx_memcpy_dtoh", out.ctypes.data, inf["volume_ptr"],
               out.nbytes, None)
        return out

    def cancel(self):
        C.call("gx_server_cancel", self.server.handle, self.id)

    def release(self):
        """Frees the volume in server memory; the request is then gone."""
        C.call("gx_server_release", self.server.handle, self.id)


class Server:
    def __init__(self, budget_bytes=0, max_inflight=4, backfill=False,
                 gpu_index=-1):
        cfg = C.ServerCfg(int(budget_bytes), int(max_inflight),
                          1 if backfill else 0, int(gpu_index))
        p = C._vp()
        C.call("gx_server_create", ctypes.byref(cfg), ctypes.byref(p))
        self.handle = p.value

    @staticmethod
    def request_bytes(num_pixels, num_rows, angles, depth, input_on_host=True,
                      **kw):
        desc, _ = C.volume_desc(num_pixels, num_rows, angles, depth, **kw)
        n = C._sz()
        C.call("gx_server_request_bytes", ctypes.byref(desc),
               1 if input_on_host else 0, ctypes.byref(n))
        return n.value

    def submit(self, projections, angles, depth, **kw):
        """projections: numpy (host), DeviceArray or DeviceView, laid out per
        input_layout ((T, Y, X) raw projections by default); keyword
        arguments as VolumePlan's (width, num_defocus, pixel_size, centre,
        min/max_defocus, weights, storage, input_layout, defocus,
        wavelength, Cs, Ca, Pa, phi). A host input is copied by the server
        and held here until the request ends."""
        on_host = not isinstance(projections, (C.DeviceArray, C.DeviceView))
        if on_host:
            projections = numpy.ascontiguousarray(projections,
                                                  dtype=numpy.float32)
            ptr = projections.ctypes.data
            shape = projections.shape
        else:
            ptr = projections.ptr
            shape = projections.shape
        layout = kw.get("input_layout", C.INPUT_PROJECTIONS)
        if layout == C.INPUT_SINOGRAMS:
            Y, X = shape[0], shape[-1]
        else:
            Y, X = shape[-2], shape[-1]
        desc, arrays = C.volume_desc(X, Y, angles, depth, **kw)
        rid = ctypes.c_uint64()
        C.call("gx_server_submit", self.handle, ctypes.byref(desc), ptr,
               1 if on_host else 0, ctypes.byref(rid))
        width = kw.get("width") or X
        return Request(self, rid.value, (Y, depth, width),
                       (projections, desc, arrays))

    def stats(self):
        st = C.ServerStats()
        C.call("gx_server_get_stats", self.handle, ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in C.ServerStats._fields_}

    def close(self):
        if getattr(self, "handle", None):
            C.call("gx_server_destroy", self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
