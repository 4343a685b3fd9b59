"""Numpy restatement of VLSV files and of vlasiator's velocity block arrays (test
infrastructure only: the checker of libvacc's VLSV I/O, never part of the product).

Format (fmihpc/vlsv, the external library vlasiator links; it is not in the reference
tree and the reference pins no version of it -- its Makefile only records the commit it
was built with, Makefile:268):
  bytes 0-7   endianness word, byte 0 = 0 for little-endian
  bytes 8-15  uint64 byte offset of the XML footer
  arrays      back to back, each arraysize * vectorsize * datasize bytes
  footer      <VLSV> root; one child tag per array with the mandatory attributes
              arraysize, vectorsize, datatype (int | uint | float), datasize plus the
              writer's own; the tag text is the array's byte offset in ascii
              (doc/vlsv2/documentation.lyx:85-276 describes the tags and their mandatory
              attributes; that document's older layout starts the file with the footer
              offset itself, which read() accepts too).
The header word order is restated from the public library; no VLSV file in the reference
tree pins it ("parity unpinned" for the header against files written by fmihpc/vlsv).

Block arrays of one population (iowrite.cpp:141-318, ioread.cpp:371-700):
  CELLSWITHBLOCKS  uint64 cell IDs            {mesh: SpatialGrid, name: pop}
  BLOCKSPERCELL    uint32 blocks per cell     {mesh: SpatialGrid, name: pop}
  MESH_BBOX        uint64 [Lx Ly Lz 4 4 4]    {mesh: pop, type: amr_ucd, max_velocity_ref_level: 0}
  MESH_NODE_CRDS_X/Y/Z  float64 vmin + i*dv, i = 0..L*4  {same as MESH_BBOX}
  BLOCKIDS         uint32 GIDs                {mesh: SpatialGrid, name: pop}
  BLOCKVARIABLE    float32 x 64 per block     {mesh: SpatialGrid, name: pop}
"""
from __future__ import annotations

import struct
import xml.etree.ElementTree as ET

import numpy as np

WID = 4
SPATIAL_MESH = "SpatialGrid"
_KIND = {"int": "i", "uint": "u", "float": "f"}


def _type_of(dt: np.dtype) -> str:
    return {"i": "int", "u": "uint", "f": "float"}[np.dtype(dt).kind]


def write(path, arrays, legacy_header: bool = False):
    """arrays: list of (tag, attrs dict, ndarray of shape (n,) or (n, vectorsize)).
    legacy_header: start the file with the footer offset (the older layout)."""
    head = 8 if legacy_header else 16
    body = bytearray()
    tags = []
    for tag, attrs, a in arrays:
        a = np.ascontiguousarray(a)
        a = a.astype(a.dtype.newbyteorder("<"), copy=False)
        n = a.shape[0]
        vs = 1 if a.ndim == 1 else int(np.prod(a.shape[1:]))
        full = dict(attrs)
        full.update(arraysize=str(n), vectorsize=str(vs), datatype=_type_of(a.dtype), datasize=str(a.dtype.itemsize))
        tags.append((tag, full, head + len(body)))
        body += a.tobytes()
    foot = "<VLSV>\n" + "".join(
        "   <%s %s>%d</%s>\n" % (t, " ".join('%s="%s"' % kv for kv in sorted(at.items())), off, t)
        for t, at, off in tags) + "</VLSV>\n"
    footer_offset = head + len(body)
    with open(path, "wb") as fh:
        if legacy_header:
            fh.write(struct.pack("<Q", footer_offset))
        else:
            fh.write(struct.pack("<QQ", 0, footer_offset))
        fh.write(bytes(body))
        fh.write(foot.encode())


def _footer(raw: bytes):
    for pos in (8, 0):
        (off,) = struct.unpack_from("<Q", raw, pos)
        if 8 <= off < len(raw) and raw[off:].lstrip().startswith(b"<VLSV>"):
            return off, ET.fromstring(raw[off:].decode())
    raise ValueError("not a VLSV file")


def read(path):
    """Returns (footer offset, [(tag, attrs, array)]) -- arrays as stored, 2-D when
    vectorsize > 1."""
    raw = open(path, "rb").read()
    off, root = _footer(raw)
    out = []
    for el in root:
        at = dict(el.attrib)
        n, vs, ds = int(at["arraysize"]), int(at["vectorsize"]), int(at["datasize"])
        dt = np.dtype("<%s%d" % (_KIND[at["datatype"]], ds))
        start = int(el.text.strip())
        a = np.frombuffer(raw, dt, n * vs, start).copy()
        out.append((el.tag, at, a.reshape(n, vs) if vs > 1 else a))
    return off, out


def find(entries, tag, **attrs):
    for t, at, a in entries:
        if t == tag and all(at.get(k) == v for k, v in attrs.items()):
            return a
    raise KeyError((tag, attrs))


def cell_size(vmin, vmax, length):
    """velocity_mesh_parameters.cpp:153-163: (vmax - vmin) / L, then / WID."""
    return (float(vmax) - float(vmin)) / int(length) / WID


def velocity_distribution_arrays(grid_length, vmin, vmax, pop, cell_ids, cells):
    """The arrays writeVelocityDistributionData writes for one population, in its order
    (iowrite.cpp:141-318).  cells: [(gids, f)] in file order."""
    cattr = {"mesh": SPATIAL_MESH, "name": pop}
    mattr = {"mesh": pop, "type": "amr_ucd", "max_velocity_ref_level": "0"}
    arrays = [("CELLSWITHBLOCKS", cattr, np.asarray(cell_ids, np.uint64)),
              ("BLOCKSPERCELL", cattr, np.array([len(g) for g, _ in cells], np.uint32)),
              ("MESH_BBOX", mattr, np.array(list(grid_length) + [WID] * 3, np.uint64))]
    for d, t in enumerate(("MESH_NODE_CRDS_X", "MESH_NODE_CRDS_Y", "MESH_NODE_CRDS_Z")):
        dv = cell_size(vmin[d], vmax[d], grid_length[d])
        arrays.append((t, mattr, float(vmin[d]) + np.arange(grid_length[d] * WID + 1, dtype=np.float64) * dv))
    g = np.concatenate([np.asarray(g, np.uint32) for g, _ in cells]) if cells else np.zeros(0, np.uint32)
    f = (np.concatenate([np.asarray(f).reshape(-1, 64) for _, f in cells]) if cells
         else np.zeros((0, 64), np.float32))
    arrays.append(("BLOCKIDS", cattr, g))
    arrays.append(("BLOCKVARIABLE", cattr, f))
    return arrays


def read_block_data(path, grid_length, vmin, vmax, pop, mesh_name=SPATIAL_MESH, cell_offset=0, ncells=None):
    """readBlockData + _readBlockData (ioread.cpp:371-700) for one process reading cells
    [cell_offset, cell_offset + ncells): returns (cell IDs, [(gids, f float32)]).
    A smaller file velocity mesh is placed into ours by renumbering (ioread.cpp:520-631);
    errors raise ValueError."""
    _, E = read(path)
    bbox = [int(x) for x in find(E, "MESH_BBOX", mesh=pop).reshape(-1)[:6]]
    L = [int(x) for x in grid_length]
    remap = None
    if bbox[:3] != L:
        if any(L[d] < bbox[d] for d in range(3)):
            raise ValueError("trying to shrink velocity space")
        shift = []
        for d, t in enumerate(("MESH_NODE_CRDS_X", "MESH_NODE_CRDS_Y", "MESH_NODE_CRDS_Z")):
            crds = find(E, t, mesh=pop).astype(np.float64)[: bbox[d] * bbox[d + 3] + 1]
            dv = cell_size(vmin[d], vmax[d], L[d])
            ci = (crds - float(vmin[d])) / dv
            if np.any(np.abs(np.rint(ci) - ci) > 1e-4):
                raise ValueError("cell coordinates do not match")
            c0 = int(np.rint(ci[0]))
            if c0 % WID:
                raise ValueError("not at a block boundary")
            shift.append(c0 // WID)

        def remap(g):
            g = g.astype(np.int64)
            x, y, z = g % bbox[0], (g // bbox[0]) % bbox[1], g // (bbox[0] * bbox[1])
            return ((x + shift[0]) + (y + shift[1]) * L[0] + (z + shift[2]) * L[0] * L[1]).astype(np.uint32)
    ids = find(E, "CELLSWITHBLOCKS", mesh=mesh_name, name=pop).astype(np.uint64)
    per = find(E, "BLOCKSPERCELL", mesh=mesh_name, name=pop).astype(np.int64)
    if ncells is None:
        ncells = len(ids) - cell_offset
    gids = find(E, "BLOCKIDS", mesh=mesh_name, name=pop)
    f = find(E, "BLOCKVARIABLE", mesh=mesh_name, name=pop)
    start = np.concatenate([[0], np.cumsum(per)])
    cells = []
    for c in range(cell_offset, cell_offset + ncells):
        g = gids[start[c]:start[c + 1]].astype(np.uint32)
        if remap is not None:
            g = remap(g)
        cells.append((g, f[start[c]:start[c + 1]].astype(np.float32)))
    return ids[cell_offset:cell_offset + ncells], cells
