"""Reader for the VTK XML files the host writers produce (binary DataArrays
with UInt32 headers, raw or vtkZLibDataCompressor), for the format tests."""
import base64
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np

_TYPES = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8, "Int32": np.int32, "UInt32": np.uint32}


def _decode(text, compressed):
    text = text.strip()
    if not compressed:
        raw = base64.b64decode(text)
        (n,) = struct.unpack("<I", raw[:4])
        assert len(raw) == 4 + n, "raw header size"
        return raw[4:]
    hdr = base64.b64decode(text[:24])
    nblocks, usize, last, csize = struct.unpack("<4I", hdr)
    assert nblocks == 1 and usize == last
    data = zlib.decompress(base64.b64decode(text[24:]))
    assert len(data) == usize, "uncompressed size"
    assert len(base64.b64decode(text[24:])) == csize, "compressed size"
    return data


def read(path):
    """{'type': ..., 'attrs': root attributes, 'piece': piece attributes,
    'arrays': {name: ndarray (n, ncomp) or (n,)}}"""
    root = ET.parse(path).getroot()
    compressed = root.get("compressor") == "vtkZLibDataCompressor"
    assert root.get("header_type") == "UInt32" and root.get("byte_order") == "LittleEndian"
    out = {"type": root.get("type"), "attrs": dict(root.attrib), "arrays": {}}
    piece = root.find(".//Piece")
    out["piece"] = dict(piece.attrib) if piece is not None else {}
    for da in root.iter("DataArray"):
        assert da.get("format") == "binary"
        a = np.frombuffer(_decode(da.text, compressed), _TYPES[da.get("type")])
        nc = int(da.get("NumberOfComponents", "1"))
        out["arrays"][da.get("Name")] = a.reshape(-1, nc) if nc > 1 else a
    ext = root.find(".//RectilinearGrid")
    if ext is not None:
        out["extent"] = ext.get("WholeExtent")
    return out


def read_pvd(path):
    """[(timestep, file)] of a ParaView collection"""
    root = ET.parse(path).getroot()
    assert root.get("type") == "Collection"
    return [(float(d.get("timestep")), d.get("file")) for d in root.iter("DataSet")]
