"""Native section files (.snp) of the time loop's sections.

Layout of src/swpc_3d/m_snap.f90, snp_format = 'native' (stream access,
machine endian):
  * header (write_snp_header, :846-892): binary type 'STREAMIO', code type
    'SWPC_3D ', header version 6, title (80 characters), exedate, coordinate
    ('xy', 'yz', 'xz', 'fs', 'ob'), data type ('ps', 'v3', 'u3'), ns1, ns2,
    the first coordinate of each axis, the two spacings, the frame interval
    dt*ntdec_s, the absorber widths in section points na1, na2, nmed, nsnp,
    clon, clat, phi and three words 1.0;
  * medium (newfile_yz/xz/xy, :324-473): rho, lambda, mu at the section's
    points, (ns1, ns2) each, first axis fastest; xy, fs and ob sections add
    the topography -bddep(i, j, 0) * 1000 (metres, positive up);
  * frames: nsnp arrays (ns1, ns2) per output step (wbuf_*, e.g.
    :1587-1653), div, rot_x, rot_y, rot_z for ps, x, y, z for v3 and u3.
The sections' points, coordinates and sizes are m_snap.f90:125-160; the
absorber widths :335-336,380-381,425-426; file names :255-272.
The maxima maps of fs and ob sections are not written in this format
(output__put_maxval, :2306-2310).
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

BINARY_TYPE = b"STREAMIO"
CODE_TYPE = b"SWPC_3D "
HEADER_VERSION = 6

KINDS = {"ps": ("ps", 4), "v": ("v3", 3), "u": ("u3", 3)}   # section kind -> data type, nsnp
PLANES = ("xy", "yz", "xz", "fs", "ob")


def _chars(s, n):
    b = s.encode() if isinstance(s, str) else bytes(s)
    return b[:n].ljust(n, b" ")


def coords(n: int, dec: int, beg: float, d: float) -> np.ndarray:
    """section coordinates: i2x (src/shared/m_fdtool.f90:639-649) at grid
    cell i*dec - dec/2 for i = 1..n, in single precision"""
    f32 = np.float32
    ii = np.arange(1, n + 1) * dec - dec // 2
    return (f32(beg) + (ii.astype(f32) - f32(0.5)) * f32(d)).astype(f32)


@dataclass
class SnapGeometry:
    """What the sections of one run share (snap__setup)"""
    nx: int
    ny: int
    nz: int
    na: int
    dx: float
    dy: float
    dz: float
    dt: float
    xbeg: float
    ybeg: float
    zbeg: float
    idec: int = 1
    jdec: int = 1
    kdec: int = 1
    ntdec_s: int = 10
    k0_xy: int = 1
    i0_yz: int = 1
    j0_xz: int = 1

    @property
    def nxs(self):
        return (self.nx + self.idec // 2) // self.idec

    @property
    def nys(self):
        return (self.ny + self.jdec // 2) // self.jdec

    @property
    def nzs(self):
        return (self.nz + self.kdec // 2) // self.kdec

    def axes(self, plane):
        """(ns1, ns2, axis-1 coordinates, axis-2 coordinates, na1, na2)"""
        xs = coords(self.nxs, self.idec, self.xbeg, self.dx)
        ys = coords(self.nys, self.jdec, self.ybeg, self.dy)
        zs = coords(self.nzs, self.kdec, self.zbeg, self.dz)
        if plane == "yz":
            return self.nys, self.nzs, ys, zs, self.na // self.jdec, self.na // self.kdec
        if plane == "xz":
            return self.nxs, self.nzs, xs, zs, self.na // self.idec, self.na // self.kdec
        return self.nxs, self.nys, xs, ys, self.na // self.idec, self.na // self.jdec


def _spacing(c: np.ndarray, dec: int, beg: float, d: float) -> np.float32:
    """xs(2) - xs(1) of the header; with a single point, the spacing the
    second point would have"""
    c2 = c if c.size >= 2 else coords(2, dec, beg, d)
    return np.float32(c2[1] - c2[0])


def header_bytes(geo: SnapGeometry, plane: str, kind: str, title: str, exedate: int, clon: float = 139.7604,
                 clat: float = 35.7182, phi: float = 0.0) -> bytes:
    """write_snp_header (m_snap.f90:846-892)"""
    snaptype, nsnp = KINDS[kind]
    ns1, ns2, c1, c2, na1, na2 = geo.axes(plane)
    dec1, beg1, d1 = {"yz": (geo.jdec, geo.ybeg, geo.dy)}.get(plane, (geo.idec, geo.xbeg, geo.dx))
    dec2, beg2, d2 = (geo.kdec, geo.zbeg, geo.dz) if plane in ("yz", "xz") else (geo.jdec, geo.ybeg, geo.dy)
    nmed = 3 if plane in ("yz", "xz") else 6
    i4 = lambda v: np.int32(v).tobytes()
    f4 = lambda v: np.float32(v).tobytes()
    out = [BINARY_TYPE, CODE_TYPE, i4(HEADER_VERSION), _chars(title, 80), i4(exedate),
           _chars(plane, 2), _chars(snaptype, 2), i4(ns1), i4(ns2), f4(c1[0]), f4(c2[0]),
           f4(_spacing(c1, dec1, beg1, d1)), f4(_spacing(c2, dec2, beg2, d2)),
           f4(np.float32(geo.dt) * np.float32(geo.ntdec_s)), i4(na1), i4(na2), i4(nmed), i4(nsnp),
           f4(clon), f4(clat), f4(phi), f4(1.0), f4(1.0), f4(1.0)]
    return b"".join(out)


def medium_block(geo: SnapGeometry, plane: str, rho, lam, mu, kfs=None, kob=None, topo_depth: float = 0.0) -> bytes:
    """rho, lambda, mu at the section points (newfile_yz/xz/xy,
    m_snap.f90:324-473); arrays in the (k,i,j) halo layout as numpy
    (ny+6, nx+6, nz+6); kfs/kob (ny+6, nx+6) for fs/ob sections.  xy, fs and
    ob add the topography -topo_depth * 1000 (bddep(:,:,0) of the layered
    models is the first layer's depth, m_vmodel_lhm.f90:104)."""
    ns1, ns2, _, _, _, _ = geo.axes(plane)
    i1 = np.arange(1, ns1 + 1)
    i2 = np.arange(1, ns2 + 1)
    if plane == "yz":
        jj = i1 * geo.jdec - geo.jdec // 2
        kk = i2 * geo.kdec - geo.kdec // 2
        J, K = np.meshgrid(jj, kk)                       # (ns2, ns1): first axis fastest
        idx = (J + 2, np.full_like(J, geo.i0_yz + 2), K + 2)
    elif plane == "xz":
        ii = i1 * geo.idec - geo.idec // 2
        kk = i2 * geo.kdec - geo.kdec // 2
        I, K = np.meshgrid(ii, kk)
        idx = (np.full_like(I, geo.j0_xz + 2), I + 2, K + 2)
    else:
        ii = i1 * geo.idec - geo.idec // 2
        jj = i2 * geo.jdec - geo.jdec // 2
        I, J = np.meshgrid(ii, jj)
        if plane == "fs":
            K = np.asarray(kfs)[J + 2, I + 2] + 1
        elif plane == "ob":
            K = np.asarray(kob)[J + 2, I + 2] + 1
        else:
            K = np.full_like(I, geo.k0_xy)
        idx = (J + 2, I + 2, K + 2)
    out = [np.ascontiguousarray(np.asarray(a, np.float32)[idx]).tobytes() for a in (rho, lam, mu)]
    if plane not in ("yz", "xz"):
        t = np.float32(-np.float32(topo_depth) * np.float32(1000))
        out.append(np.full((ns2, ns1), t, np.float32).tobytes())
    return b"".join(out)


def file_name(odir: str, title: str, plane: str, kind: str) -> str:
    """m_snap.f90:255-272"""
    return os.path.join(odir, f"{title}.3d.{plane}.{kind}.snp")


class SnpFile:
    """One native section file: header and medium on creation, then one
    frame (nsnp, ns2, ns1) -- the layout swpc_get_snapshot returns -- per
    output step."""

    def __init__(self, path: str, geo: SnapGeometry, plane: str, kind: str, medium: bytes, title: str,
                 exedate: int, clon: float = 139.7604, clat: float = 35.7182, phi: float = 0.0):
        self.plane, self.kind = plane, kind
        self.ns1, self.ns2, _, _, _, _ = geo.axes(plane)
        self.nsnp = KINDS[kind][1]
        self.f = open(path, "wb")
        self.f.write(header_bytes(geo, plane, kind, title, exedate, clon, clat, phi))
        self.f.write(medium)

    def write_frame(self, frame) -> None:
        a = np.ascontiguousarray(frame, np.float32)
        if a.shape != (self.nsnp, self.ns2, self.ns1):
            raise ValueError(f"frame of shape {a.shape}, expected {(self.nsnp, self.ns2, self.ns1)}")
        self.f.write(a.tobytes())

    def close(self) -> None:
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def write_sections(sim, geo: SnapGeometry, sections, nframes: int, odir: str, title: str, exedate: int, rho, lam,
                   mu, kfs=None, kob=None, topo_depth: float = 0.0, clon: float = 139.7604, clat: float = 35.7182,
                   phi: float = 0.0) -> list:
    """Write the first nframes frames that a simulation holds on the device
    for each section name ('xy_v', 'fs_u', 'yz_ps', ...) to its .snp file;
    returns the paths."""
    paths = []
    for name in sections:
        plane, kind = name.split("_")
        path = file_name(odir, title, plane, kind)
        med = medium_block(geo, plane, rho, lam, mu, kfs, kob, topo_depth)
        with SnpFile(path, geo, plane, kind, med, title, exedate, clon, clat, phi) as f:
            for fr in range(nframes):
                f.write_frame(sim.snapshot(name, fr))
        paths.append(path)
    return paths


def read_snp(path: str) -> dict:
    """Read a native section file back: header fields in the order of
    write_snp_header, the medium arrays and the frames, as (ns2, ns1)
    arrays."""
    b = open(path, "rb").read()
    o = [0]

    def take(n):
        v = b[o[0]:o[0] + n]
        o[0] += n
        return v

    i4 = lambda: int(np.frombuffer(take(4), np.int32)[0])
    f4 = lambda: float(np.frombuffer(take(4), np.float32)[0])
    h = dict(bintype=take(8), codetype=take(8), hdrver=i4(), title=take(80).decode().rstrip(), exedate=i4(),
             coordinate=take(2).decode(), datatype=take(2).decode(), ns1=i4(), ns2=i4(), beg1=f4(), beg2=f4(),
             ds1=f4(), ds2=f4(), dt=f4(), na1=i4(), na2=i4(), nmed=i4(), nsnp=i4(), clon=f4(), clat=f4(), phi=f4(),
             dum=(f4(), f4(), f4()))
    n = h["ns1"] * h["ns2"]
    nmed_written = 3 if h["coordinate"] in ("yz", "xz") else 4
    med = [np.frombuffer(take(4 * n), np.float32).reshape(h["ns2"], h["ns1"]) for _ in range(nmed_written)]
    rest = np.frombuffer(b[o[0]:], np.float32)
    per = h["nsnp"] * n
    if rest.size % per:
        raise ValueError(f"{path}: {rest.size} trailing words are not whole frames of {per}")
    frames = rest.reshape(-1, h["nsnp"], h["ns2"], h["ns1"])
    return dict(header=h, medium=med, frames=frames)
