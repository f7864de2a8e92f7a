"""Data and geometry files for the projectors: elsa's MHD and EDF formats and a geometry file.

  MHD.read(path) / MHD.write(data, meta_path, raw_path, spacing=None)
      elsa/io/MHDHandler.cpp (MetaImage: key = value header, raw little-endian data file)
  EDF.read(path) / EDF.write(data, path, spacing=None)
      elsa/io/EDFHandler.cpp (ESRF data format: {key = value; ...} header padded to a multiple
      of 1024 bytes, then the raw data)
  read(path) / write(data, path)
      elsa/io/IO.cpp: dispatch on the extension (.mhd, .edf)
  write_geometry(path, vol_dims, det_dims, poses) / read_geometry(path)
      the per-pose matrices P, Pinv, C, ext, sdd that ehp_geometry_create takes (FP32,
      column-major, elsa's Geometry::buildMatrices, Geometry.cpp:225-270), as a file: one
      JSON header line behind a magic line, then the five arrays raw and little-endian.

Arrays are numpy in C order of elsa's x-fastest layout: shape ([z,] y, x) for a volume,
(poses, [v,] u) for a sinogram; file dimension 1 is the last numpy axis. A read returns
(data, spacing) with spacing per file dimension. Like elsa, reading supports unsigned 16-bit,
float and double data; writing supports the six element types elsa names.
"""
from __future__ import annotations

import json
import os

import numpy as np

_MHD_TYPES = {"MET_CHAR": np.int8, "MET_UCHAR": np.uint8, "MET_SHORT": np.int16, "MET_USHORT": np.uint16,
              "MET_FLOAT": np.float32, "MET_DOUBLE": np.float64}
_EDF_TYPES = {"signedbyte": np.int8, "unsignedbyte": np.uint8, "signedshort": np.int16, "unsignedshort": np.uint16,
              "float": np.float32, "floatvalue": np.float32, "real": np.float32, "double": np.float64,
              "doublevalue": np.float64}
_EDF_NAMES = {np.dtype(np.int8): "SignedByte", np.dtype(np.uint8): "UnsignedByte", np.dtype(np.int16): "SignedShort",
              np.dtype(np.uint16): "UnsignedShort", np.dtype(np.float32): "FloatValue",
              np.dtype(np.float64): "DoubleValue"}
# the element types elsa's readers parse (DataUtils::parseRawData instantiations)
_READABLE = (np.dtype(np.uint16), np.dtype(np.float32), np.dtype(np.float64))


def _fmt(v):
    """A number as elsa's streams print it (integers without a decimal point)."""
    f = float(v)
    return str(int(f)) if f == int(f) else repr(f)


def _raw(fobj, dt, dims, path):
    n = int(np.prod(dims))
    data = np.fromfile(fobj, dtype=np.dtype(dt).newbyteorder("<"), count=n)
    if data.size != n:
        raise IOError(f"{path}: expected {n} elements, found {data.size}")
    return data.reshape(tuple(reversed(dims)))


class MHD:
    @staticmethod
    def read(path):
        """elsa MHD::read (MHDHandler.cpp): returns (data, spacing)."""
        if not os.path.isfile(path):
            raise IOError(f"MHD::read: cannot read from '{path}'")
        props = {}
        with open(path) as f:
            for line in f:
                line = line.strip()
                if not line:
                    continue
                if "=" not in line:
                    raise ValueError("MHD::readHeader: found non-empty line without name/value pair")
                name, value = line.split("=", 1)
                props[name.strip().lower()] = value.strip()
        if "ndims" not in props:
            raise ValueError("MHD::parseHeader: tag 'ndims' not found")
        ndims = int(props["ndims"])
        if props.get("elementbyteordermsb", "false").lower() not in ("false", "no"):
            raise ValueError("MHD::parseHeader: only supporting little endian byte order")
        if "elementtype" not in props:
            raise ValueError("MHD::parseHeader: tag 'element type' not found")
        et = props["elementtype"].upper()
        if et not in _MHD_TYPES:
            raise ValueError("MHD::parseHeader: tag 'element type' of unsupported value")
        if "elementdatafile" not in props:
            raise ValueError("MHD::parseHeader: tag 'element data file' not found")
        if "dimsize" not in props:
            raise ValueError("MHD::parseHeader: tag 'dim size' not found")
        dims = [int(v) for v in props["dimsize"].split()]
        if len(dims) != ndims:
            raise ValueError("MHD::parseHeader: dimension size mismatch")
        spacing = [1.0] * ndims
        if "elementspacing" in props:
            spacing = [float(v) for v in props["elementspacing"].split()]
            if len(spacing) != ndims:
                raise ValueError("MHD::parseHeader: spacing size mismatch")
        dt = np.dtype(_MHD_TYPES[et])
        if dt not in _READABLE:
            raise ValueError("MHD::read: invalid/unsupported data type")
        raw = props["elementdatafile"]
        raw = raw if os.path.isabs(raw) else os.path.join(os.path.dirname(os.path.abspath(path)), raw)
        if not os.path.isfile(raw):
            raise IOError(f"MHD::read: can not read from '{props['elementdatafile']}'")
        with open(raw, "rb") as f:
            return _raw(f, dt, dims, raw), np.asarray(spacing)

    @staticmethod
    def write(data, meta_path, raw_path, spacing=None):
        """elsa MHD::write: the header names raw_path as given."""
        data = np.ascontiguousarray(data)
        names = {np.dtype(v): k for k, v in _MHD_TYPES.items()}
        if data.dtype not in names:
            raise ValueError("MHD::getDataTypeName: invalid/unsupported data type")
        dims = list(reversed(data.shape))
        spacing = [1.0] * len(dims) if spacing is None else list(spacing)
        with open(meta_path, "w") as f:
            f.write(f"NDims = {len(dims)}\n")
            f.write("DimSize = " + " ".join(str(d) for d in dims) + "\n")
            f.write("ElementSpacing = " + " ".join(_fmt(s) for s in spacing) + "\n")
            f.write(f"ElementType = {names[data.dtype]}\n")
            f.write("ElementByteOrderMSB = False\n")
            f.write(f"ElementDataFile = {raw_path}\n")
        raw = raw_path if os.path.isabs(raw_path) else os.path.join(os.path.dirname(os.path.abspath(meta_path)),
                                                                    raw_path)
        data.astype(data.dtype.newbyteorder("<"), copy=False).tofile(raw)


class EDF:
    @staticmethod
    def _header(f):
        if f.read(1) != b"{":
            raise ValueError("EDF::readHeader: no header opening marker")
        text = bytearray()
        while True:
            c = f.read(1)
            if not c or c == b"}":
                break
            text += c
        f.read(1)  # the newline after the closing brace (the header ends with "}\n")
        props, sq, dq, cur = {}, False, False, ""
        for ch in text.decode("latin-1"):
            if ch == ";" and not (sq or dq):
                if cur.strip():
                    if "=" not in cur:
                        raise ValueError("failed reading name/value delimiter")
                    name, value = cur.split("=", 1)
                    value = value.strip()
                    if len(value) >= 2 and value[0] == value[-1] and value[0] in "'\"":
                        value = value[1:-1]
                    props[name.strip().lower()] = value
                cur = ""
                continue
            if ch == "'":
                sq = not sq
            if ch == '"':
                dq = not dq
            cur += ch
        return props

    @staticmethod
    def read(path):
        """elsa EDF::read (EDFHandler.cpp): returns (data, spacing)."""
        if not os.path.isfile(path):
            raise IOError(f"EDF::read: cannot read from '{path}'")
        with open(path, "rb") as f:
            p = EDF._header(f)
            dims = []
            while f"dim_{len(dims) + 1}" in p:
                dims.append(int(p[f"dim_{len(dims) + 1}"]))
            if not dims:
                raise ValueError("EDF::parseHeader: dimension information not found")
            spacing = [float(v) for v in p["spacing"].split()] if "spacing" in p else []
            if p.get("byteorder", "lowbytefirst").lower() != "lowbytefirst":
                raise ValueError("EDF::parseHeader: unsupported byte order value")
            if "datatype" not in p:
                raise ValueError("EDF::parseHeader: data type not found")
            tname = p["datatype"].lower()
            if tname not in _EDF_TYPES:
                raise ValueError("EDF::parseHeader: invalid/unsupported data type")
            if "compression" in p:
                raise ValueError("EDF::parseHeader: compression not supported")
            if "image" in p and int(p["image"]) != 1:
                raise ValueError("EDF::parseHeader: image not set to 1")
            dt = np.dtype(_EDF_TYPES[tname])
            if int(np.prod(dims)) * dt.itemsize != int(p.get("size", 0)):
                raise ValueError("EDF::parseHeader: size inconsistency")
            if spacing and len(spacing) != len(dims):
                raise ValueError("EDF::parseHeader: spacing inconsistency")
            if dt not in _READABLE:
                raise ValueError("EDF::read: invalid/unsupported data type")
            return _raw(f, dt, dims, path), np.asarray(spacing or [1.0] * len(dims))

    @staticmethod
    def header_text(data, spacing=None):
        """The header EDF::writeHeader writes (padded so that it ends at a multiple of 1024)."""
        data = np.asarray(data)
        if data.dtype not in _EDF_NAMES:
            raise ValueError("EDF::getDataTypeName: invalid/unsupported data type")
        dims = list(reversed(data.shape))
        spacing = [1.0] * len(dims) if spacing is None else list(spacing)
        h = "{\n"
        h += "HeaderID = EH:000001:000000:000000;\n"
        h += "Image = 1;\n"
        h += "ByteOrder = LowByteFirst;\n"
        h += f"DataType = {_EDF_NAMES[data.dtype]};\n"
        for i, d in enumerate(dims):
            h += f"Dim_{i + 1} = {d};\n"
        h += f"Size = {data.size * data.dtype.itemsize};\n"
        h += "Spacing =" + "".join(" " + _fmt(s) for s in spacing) + ";\n"
        n = 1024
        while n < len(h) + 3:
            n += 1024
        return h + " " * (n - len(h) - 3) + "\n}\n"

    @staticmethod
    def write(data, path, spacing=None):
        data = np.ascontiguousarray(data)
        with open(path, "wb") as f:
            f.write(EDF.header_text(data, spacing).encode("latin-1"))
            data.astype(data.dtype.newbyteorder("<"), copy=False).tofile(f)


def read(path):
    """elsa io::read: by extension (.edf, .mhd)."""
    ext = os.path.splitext(path)[1].lower()
    if ext == ".edf":
        return EDF.read(path)
    if ext == ".mhd":
        return MHD.read(path)
    raise ValueError(f"io::read: file extension '{ext}' not supported")


def write(data, path, spacing=None):
    """elsa io::write: .edf, or .mhd with the raw data beside it (same name, .raw)."""
    ext = os.path.splitext(path)[1].lower()
    if ext == ".edf":
        return EDF.write(data, path, spacing)
    if ext == ".mhd":
        return MHD.write(data, path, os.path.splitext(os.path.basename(path))[0] + ".raw", spacing)
    raise ValueError(f"io::write: file extension '{ext}' not supported")


_GEOM_MAGIC = b"ELSA_AMD_GEOMETRY 1\n"


def write_geometry(path, vol_dims, det_dims, poses):
    """poses: (P, Pinv, C, ext, sdd) arrays of n poses (e.g. a PlanarDetectorDescriptor's
    pose_arrays()), FP32, each pose's matrices column-major."""
    P, Pinv, C, ext, sdd = (np.ascontiguousarray(a, np.float32) for a in poses)
    dim = len(vol_dims)
    n = sdd.shape[0]
    nm = dim * (dim + 1)
    for a, w in ((P, nm), (Pinv, nm), (C, dim), (ext, nm)):
        if a.size != n * w:
            raise ValueError("write_geometry: array sizes do not match the number of poses")
    head = {"dim": dim, "poses": n, "volume": [int(v) for v in vol_dims], "detector": [int(v) for v in det_dims],
            "arrays": ["P", "Pinv", "C", "ext", "sdd"], "dtype": "float32<", "order": "column-major per pose"}
    with open(path, "wb") as f:
        f.write(_GEOM_MAGIC)
        f.write((json.dumps(head) + "\n").encode())
        for a in (P, Pinv, C, ext, sdd):
            a.astype("<f4", copy=False).tofile(f)


def read_geometry(path):
    """Returns (vol_dims, det_dims, (P, Pinv, C, ext, sdd)) as written by write_geometry."""
    with open(path, "rb") as f:
        if f.readline() != _GEOM_MAGIC:
            raise ValueError(f"{path}: not a geometry file")
        head = json.loads(f.readline().decode())
        dim, n = int(head["dim"]), int(head["poses"])
        nm = dim * (dim + 1)
        arrays = []
        for w in (nm, nm, dim, nm, 1):
            a = np.fromfile(f, dtype="<f4", count=n * w)
            if a.size != n * w:
                raise ValueError(f"{path}: truncated")
            arrays.append(a.reshape(n, w) if w > 1 else a)
    return head["volume"], head["detector"], tuple(arrays)
