"""Reader for the VTK XML ImageData files the exporter writes (test helper).

Handles format="ascii" arrays and format="appended" raw arrays with UInt64
block sizes; returns {name: ndarray (npoints, ncomp)} and the extent.
"""
import re

import numpy as np

_TYPES = {"Float32": np.float32, "UInt8": np.uint8}


def read_vti(path):
    raw = open(path, "rb").read()
    cut = raw.find(b"<AppendedData")
    head = (raw if cut < 0 else raw[:cut]).decode()
    ext = [int(v) for v in re.search(r'WholeExtent="([^"]+)"', head).group(1).split()]
    npts = (ext[1] - ext[0] + 1) * (ext[3] - ext[2] + 1) * (ext[5] - ext[4] + 1)
    arrays = {}
    if cut >= 0:
        base = raw.index(b"_", cut) + 1
        for m in re.finditer(r'<DataArray type="(\w+)" Name="(\w+)" NumberOfComponents="(\d+)" '
                             r'format="appended" offset="(\d+)"/>', head):
            t, name, nc, off = m.group(1), m.group(2), int(m.group(3)), int(m.group(4))
            p = base + off
            nbytes = int(np.frombuffer(raw[p:p + 8], np.uint64)[0])
            a = np.frombuffer(raw[p + 8:p + 8 + nbytes], _TYPES[t])
            assert a.size == npts * nc, (name, a.size, npts * nc)
            arrays[name] = a.reshape(npts, nc)
    for m in re.finditer(r'<DataArray type="(\w+)" Name="(\w+)" NumberOfComponents="(\d+)" format="ascii">\n'
                         r'(.*?)        </DataArray>', head, re.S):
        t, name, nc = m.group(1), m.group(2), int(m.group(3))
        a = np.array(m.group(4).split(), dtype=np.float64).astype(_TYPES[t])
        assert a.size == npts * nc, (name, a.size, npts * nc)
        arrays[name] = a.reshape(npts, nc)
    return arrays, ext, head
