"""A reader of the VTK files the solver writes, written from the VTK XML file-format rules (not
from either writer): ImageData with one Piece, Float32 point arrays in an <AppendedData
encoding="raw"> block that starts after a '_' marker, each array at its `offset` from the marker
and prefixed by its byte count as the file's header_type (UInt64 here), little-endian."""
import xml.etree.ElementTree as ET

import numpy as np

MARKER = b'<AppendedData encoding="raw">\n_'
TRAILER = b"\n  </AppendedData>\n</VTKFile>\n"


def read_vti(path):
    """Return (attributes, {name: array (nz, ny, nx) float32}, trailer bytes)."""
    data = open(path, "rb").read()
    k = data.index(MARKER)
    start = k + len(MARKER)
    head = data[:k].decode("ascii")
    root = ET.fromstring(head[head.index("<VTKFile"):] + "</VTKFile>")
    assert root.tag == "VTKFile"
    attrs = dict(root.attrib)
    img = root.find("ImageData")
    x0, x1, y0, y1, z0, z1 = (int(v) for v in img.attrib["WholeExtent"].split())
    piece = img.find("Piece")
    assert piece.attrib["Extent"] == img.attrib["WholeExtent"]
    shape = (z1 - z0 + 1, y1 - y0 + 1, x1 - x0 + 1)
    n = shape[0] * shape[1] * shape[2]
    attrs.update(origin=img.attrib["Origin"], spacing=img.attrib["Spacing"], shape=shape)
    assert attrs["header_type"] == "UInt64" and attrs["byte_order"] == "LittleEndian"
    arrays = {}
    end = start
    for da in piece.find("PointData").findall("DataArray"):
        assert da.attrib["type"] == "Float32" and da.attrib["format"] == "appended"
        pos = start + int(da.attrib["offset"])
        nbytes = int(np.frombuffer(data, "<u8", 1, pos)[0])
        assert nbytes == 4 * n, (da.attrib["Name"], nbytes, n)
        arrays[da.attrib["Name"]] = np.frombuffer(data, "<f4", n, pos + 8).reshape(shape)
        end = max(end, pos + 8 + nbytes)
    return attrs, arrays, data[end:]


def read_pvd(path):
    """Return [(timestep, file)] of a .pvd collection."""
    root = ET.parse(path).getroot()
    assert root.tag == "VTKFile" and root.attrib["type"] == "Collection"
    return [(int(ds.attrib["timestep"]), ds.attrib["file"]) for ds in root.find("Collection").findall("DataSet")]
