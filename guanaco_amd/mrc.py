"""Memory-mapped MRC2014 files, as guanaco's file drivers use them.

The reference reads and writes tilt series, CTF-corrected stacks and volumes
through the third-party ``mrcfile`` package (detail/reconstruct.py:214-402,
detail/correct.py:321-473, plot.py:248); that package is not a dependency
here, so this module implements the part of the format those drivers need,
with the same attribute names (``data``, ``voxel_size``,
``extended_header``, ``indexed_extended_header``, ``header``) and the same
data-shape rules, so the drivers read like the reference's.

Format (MRC2014, Cheng et al., J. Struct. Biol. 192 (2015) 146-150): a
1024-byte header, ``nsymbt`` bytes of extended header, then the data,
fastest axis x. Byte order from the machine stamp (0x44 0x44 little, 0x11
0x11 big); files are written little-endian. Modes: 0 int8, 1 int16,
2 float32, 4 complex64, 6 uint16, 12 float16. Data shape: (nz // mz, mz, ny,
nx) for a volume stack (space group 401-630), (ny, nx) for a single image
(space group 0, nz 1), else (nz, ny, nx). New 3-D files are volumes (space
group 1, mz = nz), new 4-D files volume stacks (space group 401).

FEI extended header (exttyp "FEI1" / "FEI2", Thermo Fisher's per-image
metadata blocks): one record per image of "Metadata size" bytes; the fields
below are at their published byte offsets. Records written here are 768
bytes (the FEI1 block size) with only "Metadata size" and the fields the
drivers set filled in, as the reference writes them.
"""
import os
import time

import numpy

__all__ = ["MrcFile", "open", "mmap", "new_mmap", "new", "read", "write",
           "HEADER_DTYPE", "FEI1_EXTENDED_HEADER_DTYPE", "MODES",
           "mode_from_dtype"]

_builtin_open = open

HEADER_BYTES = 1024

_HEADER_FIELDS = [
    ("nx", "i4"), ("ny", "i4"), ("nz", "i4"), ("mode", "i4"),
    ("nxstart", "i4"), ("nystart", "i4"), ("nzstart", "i4"),
    ("mx", "i4"), ("my", "i4"), ("mz", "i4"),
    ("cella", [("x", "f4"), ("y", "f4"), ("z", "f4")]),
    ("cellb", [("alpha", "f4"), ("beta", "f4"), ("gamma", "f4")]),
    ("mapc", "i4"), ("mapr", "i4"), ("maps", "i4"),
    ("dmin", "f4"), ("dmax", "f4"), ("dmean", "f4"),
    ("ispg", "i4"), ("nsymbt", "i4"),
    ("extra1", "V8"), ("exttyp", "S4"), ("nversion", "i4"), ("extra2", "V84"),
    ("origin", [("x", "f4"), ("y", "f4"), ("z", "f4")]),
    ("map", "S4"), ("machst", "u1", (4,)), ("rms", "f4"), ("nlabl", "i4"),
    ("label", "S80", (10,)),
]


def _header_dtype(order):
    def conv(fmt):
        if isinstance(fmt, list):
            return [(n, conv(f)) for n, f in fmt]
        if fmt[0] in "if":
            return order + fmt
        return fmt
    return numpy.dtype([(f[0], conv(f[1])) + tuple(f[2:])
                        for f in _HEADER_FIELDS])


HEADER_DTYPE = _header_dtype("<")
assert HEADER_DTYPE.itemsize == HEADER_BYTES

# mode -> little-endian data type
MODES = {0: numpy.dtype("i1"), 1: numpy.dtype("<i2"), 2: numpy.dtype("<f4"),
         4: numpy.dtype("<c8"), 6: numpy.dtype("<u2"), 12: numpy.dtype("<f2")}

# (name, format, byte offset) of the FEI metadata block fields
_FEI_FIELDS = [
    ("Metadata size", "i4", 0), ("Metadata version", "i4", 4),
    ("Bitmask 1", "u4", 8), ("Timestamp", "f8", 12),
    ("Microscope type", "S16", 20), ("D-Number", "S16", 36),
    ("Application", "S16", 52), ("Application version", "S16", 68),
    ("HT", "f8", 84), ("Dose", "f8", 92), ("Alpha tilt", "f8", 100),
    ("Beta tilt", "f8", 108), ("X-Stage", "f8", 116), ("Y-Stage", "f8", 124),
    ("Z-Stage", "f8", 132), ("Tilt axis angle", "f8", 140),
    ("Dual axis rotation", "f8", 148), ("Pixel size X", "f8", 156),
    ("Pixel size Y", "f8", 164), ("Defocus", "f8", 220),
    ("STEM Defocus", "f8", 228), ("Applied defocus", "f8", 236),
    ("Instrument mode", "i4", 244), ("Projection mode", "i4", 248),
    ("Objective lens mode", "S16", 252),
    ("High magnification mode", "S16", 268), ("Probe mode", "i4", 284),
    ("EFTEM On", "?", 288), ("Magnification", "f8", 289),
]
FEI1_BLOCK_BYTES = 768


def fei_dtype(itemsize=FEI1_BLOCK_BYTES, order="<"):
    """Structured type of one FEI metadata block of `itemsize` bytes."""
    return numpy.dtype({
        "names": [f[0] for f in _FEI_FIELDS],
        "formats": [(order + f[1]) if f[1][0] in "iuf" else f[1]
                    for f in _FEI_FIELDS],
        "offsets": [f[2] for f in _FEI_FIELDS],
        "itemsize": itemsize})


FEI1_EXTENDED_HEADER_DTYPE = fei_dtype()

_VOXEL_DTYPE = numpy.dtype([("x", "f4"), ("y", "f4"), ("z", "f4")])


def mode_from_dtype(dtype):
    """The MRC mode storing `dtype` (mrcfile.utils.mode_from_dtype)."""
    dt = numpy.dtype(dtype).newbyteorder("<") if numpy.dtype(
        dtype).itemsize > 1 else numpy.dtype(dtype)
    for mode, d in MODES.items():
        if d == dt:
            return mode
    raise ValueError("data type %s has no MRC mode" % numpy.dtype(dtype))


def _is_volume_stack(ispg):
    return 401 <= ispg <= 630


def data_shape(header):
    """Array shape of the data a header describes (MRC2014 space groups)."""
    nx, ny, nz = int(header["nx"]), int(header["ny"]), int(header["nz"])
    mz = int(header["mz"])
    ispg = int(header["ispg"])
    if _is_volume_stack(ispg):
        if mz <= 0 or nz % mz:
            raise ValueError("volume stack with nz %d not a multiple of mz %d"
                             % (nz, mz))
        return (nz // mz, mz, ny, nx)
    if ispg == 0 and nz == 1:
        return (ny, nx)
    return (nz, ny, nx)


class MrcFile:
    """An MRC file whose data is a numpy memory map.

    mode "r" (read-only), "r+" (data and header writable) or "w+" (a new
    file; use new_mmap). Closing a writable file writes the header and the
    extended header back and flushes the data.
    """

    def __init__(self, name, mode="r", permissive=False):
        if mode not in ("r", "r+", "w+"):
            raise ValueError("mode must be 'r', 'r+' or 'w+'")
        self._name = os.fspath(name)
        self._mode = mode
        self._closed = False
        self._indexed = None
        self.data = None
        if mode != "w+":
            self._read(permissive)

    # -- reading --------------------------------------------------------
    def _read(self, permissive):
        with _builtin_open(self._name, "rb") as f:
            raw = f.read(HEADER_BYTES)
            if len(raw) < HEADER_BYTES:
                raise ValueError("%s: shorter than an MRC header" % self._name)
            order = self._byte_order(raw)
            self.header = numpy.frombuffer(
                raw, dtype=_header_dtype(order)).astype(HEADER_DTYPE).reshape(())
            if self.header["map"].item() != b"MAP " and not permissive:
                raise ValueError("%s: no MRC map identifier" % self._name)
            nsym = int(self.header["nsymbt"])
            if nsym < 0:
                raise ValueError("%s: negative extended header size"
                                 % self._name)
            self.extended_header = numpy.frombuffer(
                f.read(nsym), dtype="V1") if nsym else numpy.zeros(0, "V1")
            if self.extended_header.size != nsym:
                raise ValueError("%s: truncated extended header" % self._name)
        self.extended_header = numpy.array(self.extended_header)
        self._order = order
        mode = int(self.header["mode"])
        if mode not in MODES:
            raise ValueError("%s: MRC mode %d is not supported"
                             % (self._name, mode))
        dtype = MODES[mode].newbyteorder(order) if MODES[mode].itemsize > 1 \
            else MODES[mode]
        shape = data_shape(self.header)
        offset = HEADER_BYTES + int(self.header["nsymbt"])
        need = offset + int(numpy.prod(shape)) * dtype.itemsize
        size = os.path.getsize(self._name)
        if size < need:
            raise ValueError("%s: file holds %d bytes, the header describes %d"
                             % (self._name, size, need))
        self.data = numpy.memmap(self._name, dtype=dtype,
                                 mode="r" if self._mode == "r" else "r+",
                                 offset=offset, shape=shape)

    @staticmethod
    def _byte_order(raw):
        st = raw[212]
        if st == 0x11:
            return ">"
        if st == 0x44:
            return "<"
        # no usable machine stamp: take the order in which the mode is valid
        le = int.from_bytes(raw[12:16], "little")
        return "<" if le in MODES else ">"

    # -- creating -------------------------------------------------------
    def _create(self, shape, dtype, extended_header=None, exttyp=None,
                overwrite=False):
        if os.path.exists(self._name) and not overwrite:
            raise ValueError("%s already exists; pass overwrite=True"
                             % self._name)
        dtype = numpy.dtype(dtype)
        mode = mode_from_dtype(dtype)
        shape = tuple(int(s) for s in shape)
        if len(shape) not in (2, 3, 4):
            raise ValueError("MRC data must be 2, 3 or 4 dimensional")
        h = numpy.zeros((), HEADER_DTYPE)
        h["mode"] = mode
        h["cellb"] = (90.0, 90.0, 90.0)
        h["mapc"], h["mapr"], h["maps"] = 1, 2, 3
        h["map"] = b"MAP "
        h["machst"] = (0x44, 0x44, 0, 0)
        h["nversion"] = 20140
        ny, nx = shape[-2:]
        if len(shape) == 2:
            nz, mz, ispg = 1, 1, 0
        elif len(shape) == 3:
            nz = mz = shape[0]
            ispg = 1
        else:
            nz, mz, ispg = shape[0] * shape[1], shape[1], 401
        h["nx"], h["ny"], h["nz"] = nx, ny, nz
        h["mx"], h["my"], h["mz"] = nx, ny, mz
        h["ispg"] = ispg
        self.header = h
        self.reset_header_stats()
        self.add_label("Created by guanaco_amd: %s"
                       % time.strftime("%Y-%m-%d %H:%M:%S"))
        if extended_header is None:
            ext = numpy.zeros(0, "V1")
        else:
            ext = numpy.ascontiguousarray(extended_header).reshape(-1).view(
                "V1")
        self.extended_header = numpy.array(ext)
        h["nsymbt"] = ext.size
        if exttyp is not None:
            h["exttyp"] = exttyp
        self._order = "<"
        offset = HEADER_BYTES + ext.size
        nbytes = offset + int(numpy.prod(shape)) * dtype.itemsize
        with _builtin_open(self._name, "wb") as f:
            f.truncate(nbytes)
        self._write_header()
        self.data = numpy.memmap(self._name, dtype=MODES[mode], mode="r+",
                                 offset=offset, shape=shape)
        self._mode = "r+"
        return self

    # -- header helpers -------------------------------------------------
    @property
    def voxel_size(self):
        """Voxel size in Angstrom: cella / (mx, my, mz), as a record with
        fields x, y, z."""
        h = self.header
        v = numpy.zeros((), _VOXEL_DTYPE)
        for ax, m in (("x", "mx"), ("y", "my"), ("z", "mz")):
            n = int(h[m])
            v[ax] = float(h["cella"][ax]) / n if n else 0.0
        return v.view(numpy.recarray)

    @voxel_size.setter
    def voxel_size(self, value):
        if isinstance(value, numpy.ndarray) and value.dtype.names:
            xyz = (float(value["x"]), float(value["y"]), float(value["z"]))
        elif numpy.ndim(value) == 0:
            xyz = (float(value),) * 3
        else:
            xyz = tuple(float(v) for v in value)
            if len(xyz) != 3:
                raise ValueError("voxel size: one value or (x, y, z)")
        self._check_writable()
        h = self.header
        h["cella"] = (xyz[0] * int(h["mx"]), xyz[1] * int(h["my"]),
                      xyz[2] * int(h["mz"]))

    @property
    def indexed_extended_header(self):
        """The FEI metadata blocks as a structured array, one record per
        image (None when the extended header is not FEI1/FEI2). The array
        views the extended header, so assignments to it are written back
        when the file is closed."""
        if self._indexed is not None:
            return self._indexed
        if self.header["exttyp"].item() not in (b"FEI1", b"FEI2"):
            return None
        raw = self.extended_header.view(numpy.uint8)
        if raw.size < 4:
            return None
        size = int(numpy.frombuffer(raw[:4].tobytes(), self._order + "i4")[0])
        if size < 297 or raw.size % size:
            raise ValueError("%s: FEI metadata size %d does not divide the "
                             "%d-byte extended header" % (self._name, size,
                                                          raw.size))
        self._indexed = raw.view(fei_dtype(size, self._order))
        return self._indexed

    def reset_header_stats(self):
        """Marks dmin/dmax/dmean/rms as not determined (dmax < dmin,
        dmean < min(dmin, dmax), rms < 0)."""
        self._check_writable(creating=True)
        self.header["dmin"], self.header["dmax"] = 0.0, -1.0
        self.header["dmean"], self.header["rms"] = -2.0, -1.0

    def update_header_stats(self, chunk_bytes=1 << 28):
        """Sets dmin/dmax/dmean/rms from the data, in chunks of about
        `chunk_bytes` (one pass; the data may be larger than memory)."""
        self._check_writable()
        d = self.data.reshape(-1) if self.data.size else self.data
        n = d.size
        if n == 0:
            self.reset_header_stats()
            return
        step = max(1, chunk_bytes // d.dtype.itemsize)
        lo, hi, s1 = numpy.inf, -numpy.inf, 0.0
        for o in range(0, n, step):
            c = numpy.asarray(d[o:o + step])
            if numpy.iscomplexobj(c):
                c = numpy.abs(c)
            c = c.astype(numpy.float64)
            lo, hi, s1 = min(lo, c.min()), max(hi, c.max()), s1 + c.sum()
        mean = s1 / n
        s2 = 0.0
        for o in range(0, n, step):
            c = numpy.asarray(d[o:o + step])
            if numpy.iscomplexobj(c):
                c = numpy.abs(c)
            s2 += float(numpy.sum((c.astype(numpy.float64) - mean) ** 2))
        h = self.header
        h["dmin"], h["dmax"], h["dmean"] = lo, hi, mean
        h["rms"] = numpy.sqrt(s2 / n)

    def add_label(self, text):
        self._check_writable(creating=True)
        n = int(self.header["nlabl"])
        if n >= 10:
            raise ValueError("an MRC header holds at most 10 labels")
        self.header["label"][n] = text.encode("ascii", "replace")[:80]
        self.header["nlabl"] = n + 1

    def is_image_stack(self):
        return int(self.header["ispg"]) == 0

    def is_volume(self):
        return 1 <= int(self.header["ispg"]) <= 230

    def is_volume_stack(self):
        return _is_volume_stack(int(self.header["ispg"]))

    def set_image_stack(self):
        """Marks a 3-D file as a stack of images (space group 0, mz 1)."""
        self._check_writable()
        if self.data.ndim != 3:
            raise ValueError("only 3-D data can be an image stack")
        self.header["ispg"] = 0
        self.header["mz"] = 1

    def set_volume(self):
        self._check_writable()
        if self.data.ndim != 3:
            raise ValueError("only 3-D data can be a volume")
        self.header["ispg"] = 1
        self.header["mz"] = self.header["nz"]

    # -- writing back ---------------------------------------------------
    def _check_writable(self, creating=False):
        if self._mode == "r" and not creating:
            raise ValueError("%s is open read-only" % self._name)

    def _write_header(self):
        with _builtin_open(self._name, "r+b") as f:
            # in the file's own byte order (the data is not rewritten)
            f.write(self.header.astype(_header_dtype(self._order)).tobytes())
            f.write(self.extended_header.tobytes())

    def flush(self):
        if self._mode == "r" or self._closed:
            return
        if self.header["nsymbt"] != self.extended_header.size:
            raise ValueError("the extended header size cannot change")
        self.data.flush()
        self._write_header()

    def close(self):
        if self._closed:
            return
        self.flush()
        self._indexed = None
        self.data = None
        self._closed = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __repr__(self):
        return "MrcFile(%r, mode=%r)" % (self._name, self._mode)


def open(name, mode="r", permissive=False):  # noqa: A001
    """Opens an existing MRC file (data memory-mapped)."""
    return MrcFile(name, mode, permissive)


mmap = open


def new_mmap(name, shape, mrc_mode=2, overwrite=False, extended_header=None,
             exttyp=None):
    """Creates an MRC file for data of `shape` in `mrc_mode` and returns it
    open for writing; the data is zero-filled (a sparse file) and the header
    statistics are marked not determined (mrcfile.new_mmap's contract)."""
    if mrc_mode not in MODES:
        raise ValueError("MRC mode %d is not supported" % mrc_mode)
    f = MrcFile(name, "w+")
    return f._create(shape, MODES[mrc_mode], extended_header, exttyp,
                     overwrite)


def new(name, data, voxel_size=None, overwrite=False, extended_header=None,
        exttyp=None):
    """Writes `data` to a new MRC file (with header statistics)."""
    data = numpy.asarray(data)
    f = new_mmap(name, data.shape, mode_from_dtype(data.dtype), overwrite,
                 extended_header, exttyp)
    f.data[...] = data
    if voxel_size is not None:
        f.voxel_size = voxel_size
    f.update_header_stats()
    return f


def read(name):
    """The data of an MRC file as an in-memory array."""
    with open(name) as f:
        return numpy.array(f.data)


def write(name, data, voxel_size=None, overwrite=False, extended_header=None,
          exttyp=None):
    new(name, data, voxel_size, overwrite, extended_header, exttyp).close()
