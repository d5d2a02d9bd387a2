"""SAC, CSF and tar output of station traces and Green's functions.

Byte layouts follow the reference writers:
  * SAC: a 158-word header (70 floats, 35 integers, 5 logicals, 48 four-byte
    character words) then the samples, machine endian
    (src/shared/m_sac.f90:267-449, defaults of sac__init :452-562);
  * CSF: 'CSFD', the trace count and length, then header + samples per trace
    (m_sac.f90:585-648);
  * tar: a ustar header per SAC member, zero padding to the next 512-byte
    block, two zero blocks at the end (src/shared/m_tar.f90:48-73,165-234,
    m_sac.f90:812-829).
Which header fields a station or Green's-function trace carries follows
src/swpc_3d/m_wav.f90:272-396 and src/swpc_3d/m_green.f90:298-401; the file
names follow m_wav.f90:658-830 and m_green.f90:302-312,553-612.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field, fields

import numpy as np

UNDEF_F = -12345.0
UNDEF_I = -12345

# header word (1-based) of each field sac__whdr writes (m_sac.f90:342-434)
_FWORD = dict(delta=1, depmin=2, depmax=3, b=6, e=7, o=8, a=9,
              t0=11, t1=12, t2=13, t3=14, t4=15, t5=16, t6=17, t7=18, t8=19, t9=20,
              stla=32, stlo=33, stel=34, stdp=35, evla=36, evlo=37, evel=38, evdp=39, mag=40,
              user0=41, user1=42, user2=43, user3=44, user4=45, user5=46, user6=47, user7=48, user8=49,
              user9=50, dist=51, az=52, baz=53, gcarc=54, depmen=57, cmpaz=58, cmpinc=59,
              user10=64, user11=65, user12=66, user13=67, user14=68, user15=69, user16=70)
_IWORD = dict(nzyear=71, nzjday=72, nzhour=73, nzmin=74, nzsec=75, nzmsec=76, nvhdr=77, npts=80,
              iftype=86, idep=87, ievtyp=93, iuser0=98, iuser1=99, iuser2=100, iuser3=101, iuser4=102,
              iuser5=103, iuser6=104, iuser7=105)
_LWORD = dict(leven=106, lpspol=107, lovrok=108, lcalda=109, luser0=110)
_KFIELD = (("kstnm", 8), ("kevnm", 16), ("khole", 8), ("ko", 8), ("ka", 8), ("kt0", 8), ("kt1", 8), ("kt2", 8),
           ("kt3", 8), ("kt4", 8), ("kt5", 8), ("kt6", 8), ("kt7", 8), ("kt8", 8), ("kt9", 8), ("kf", 8),
           ("kuser0", 8), ("kuser1", 8), ("kuser2", 8), ("kcmpnm", 8), ("knetwk", 8), ("kdatrd", 8),
           ("kinst", 8))


def _f(default=UNDEF_F):
    return field(default=default)


@dataclass
class SacHeader:
    """The fields of sac__hdr that reach the file, with sac__init's defaults
    (m_sac.f90:452-562).  Floats are double here and rounded to single
    precision on writing, as sac__whdr does; tim (seconds since 1970) is the
    tar member time."""
    delta: float = UNDEF_F
    depmin: float = UNDEF_F
    depmax: float = UNDEF_F
    b: float = UNDEF_F
    e: float = UNDEF_F
    o: float = UNDEF_F
    a: float = UNDEF_F
    t0: float = UNDEF_F
    t1: float = UNDEF_F
    t2: float = UNDEF_F
    t3: float = UNDEF_F
    t4: float = UNDEF_F
    t5: float = UNDEF_F
    t6: float = UNDEF_F
    t7: float = UNDEF_F
    t8: float = UNDEF_F
    t9: float = UNDEF_F
    stla: float = UNDEF_F
    stlo: float = UNDEF_F
    stel: float = UNDEF_F
    stdp: float = UNDEF_F
    evla: float = UNDEF_F
    evlo: float = UNDEF_F
    evel: float = UNDEF_F
    evdp: float = UNDEF_F
    mag: float = UNDEF_F
    user0: float = UNDEF_F
    user1: float = UNDEF_F
    user2: float = UNDEF_F
    user3: float = UNDEF_F
    user4: float = UNDEF_F
    user5: float = UNDEF_F
    user6: float = UNDEF_F
    user7: float = UNDEF_F
    user8: float = UNDEF_F
    user9: float = UNDEF_F
    dist: float = UNDEF_F
    az: float = UNDEF_F
    baz: float = UNDEF_F
    gcarc: float = UNDEF_F
    depmen: float = UNDEF_F
    cmpaz: float = UNDEF_F
    cmpinc: float = UNDEF_F
    user10: float = UNDEF_F
    user11: float = UNDEF_F
    user12: float = UNDEF_F
    user13: float = UNDEF_F
    user14: float = UNDEF_F
    user15: float = UNDEF_F
    user16: float = UNDEF_F
    nzyear: int = UNDEF_I
    nzjday: int = UNDEF_I
    nzhour: int = UNDEF_I
    nzmin: int = UNDEF_I
    nzsec: int = UNDEF_I
    nzmsec: int = UNDEF_I
    nvhdr: int = 6
    npts: int = UNDEF_I
    iftype: int = 1
    idep: int = UNDEF_I
    ievtyp: int = UNDEF_I
    iuser0: int = UNDEF_I
    iuser1: int = UNDEF_I
    iuser2: int = UNDEF_I
    iuser3: int = UNDEF_I
    iuser4: int = UNDEF_I
    iuser5: int = UNDEF_I
    iuser6: int = UNDEF_I
    iuser7: int = UNDEF_I
    leven: bool = True
    lpspol: bool = False
    lovrok: bool = True
    lcalda: bool = False
    luser0: bool = False
    kstnm: str = "-12345"
    kevnm: str = "-12345"
    khole: str = "-12345"
    ko: str = "-12345"
    ka: str = "-12345"
    kt0: str = "-12345"
    kt1: str = "-12345"
    kt2: str = "-12345"
    kt3: str = "-12345"
    kt4: str = "-12345"
    kt5: str = "-12345"
    kt6: str = "-12345"
    kt7: str = "-12345"
    kt8: str = "-12345"
    kt9: str = "-12345"
    kf: str = "-12345"
    kuser0: str = "-12345"
    kuser1: str = "-12345"
    kuser2: str = "-12345"
    kcmpnm: str = "-12345"
    knetwk: str = "-12345"
    kdatrd: str = "-12345"
    kinst: str = "-12345"
    tim: int = 0

    def copy(self, **kw) -> "SacHeader":
        d = {f.name: getattr(self, f.name) for f in fields(self)}
        d.update(kw)
        return SacHeader(**d)


def _chars(s: str, n: int) -> bytes:
    """Fortran character assignment: truncate or blank-pad to n"""
    b = s.encode("ascii")[:n]
    return b + b" " * (n - len(b))


def header_bytes(h: SacHeader) -> bytes:
    """The 632-byte header of sac__whdr (m_sac.f90:314-449)."""
    fh = np.full(70, UNDEF_F, np.float32)
    ih = np.full(35, UNDEF_I, np.int32)
    lh = np.zeros(5, np.int32)
    for name, w in _FWORD.items():
        fh[w - 1] = np.float32(getattr(h, name))
    # delta is kept to 1e-7 s: real(int(delta*1d7)) / 1e7 in single precision
    fh[0] = np.float32(np.float32(int(h.delta * 1e7)) / np.float32(1e7))
    for name, w in _IWORD.items():
        ih[w - 71] = getattr(h, name)
    for name, w in _LWORD.items():
        lh[w - 106] = 1 if getattr(h, name) else 0
    ah = b"".join(_chars(getattr(h, name), n) for name, n in _KFIELD)
    assert len(ah) == 192
    return fh.tobytes() + ih.tobytes() + lh.tobytes() + ah


def sac_bytes(h: SacHeader, data) -> bytes:
    d = np.ascontiguousarray(data, np.float32).reshape(-1)
    if h.npts != d.size:
        raise ValueError(f"SAC header npts {h.npts} but {d.size} samples")
    return header_bytes(h) + d.tobytes()


def write_sac(path: str, h: SacHeader, data) -> None:
    """sac__write with overwrite (m_sac.f90:267-311)"""
    with open(path, "wb") as fh:
        fh.write(sac_bytes(h, data))


def write_csf(path: str, headers, data) -> None:
    """csf__write (m_sac.f90:585-648): data is (ntrace, npts).  The file is
    written from its start without truncation, as the reference's
    status='unknown' stream open does."""
    d = np.ascontiguousarray(data, np.float32)
    ntrace, npts = d.shape
    if len(headers) != ntrace:
        raise ValueError("one header per trace")
    for h in headers:
        if h.npts != npts:
            raise ValueError("csf: npts mismatch")
    buf = [b"CSFD", np.int32(ntrace).tobytes(), np.int32(npts).tobytes()]
    for h, row in zip(headers, d):
        buf.append(header_bytes(h))
        buf.append(row.tobytes())
    _write_in_place(path, b"".join(buf))


def _write_in_place(path: str, payload: bytes) -> None:
    mode = "r+b" if os.path.exists(path) else "wb"
    with open(path, mode) as fh:
        fh.seek(0)
        fh.write(payload)


# ---- tar (src/shared/m_tar.f90) ----
def tar_header(name: str, fsize: int, mtime: int) -> bytes:
    """tar__inithdr + tar__whdr (m_tar.f90:48-73,165-201).  The block starts
    as one NUL and 511 blanks (a one-character NUL assigned to a 512-character
    variable); each string field is written with a NUL after it, octal fields
    as 8 or 12 digits without terminator, and the checksum is summed with its
    own field blank.  The version field is never written."""
    nb = name.encode("ascii")
    if len(nb) >= 100:
        raise ValueError("tar member name must be shorter than 100 bytes")
    blk = bytearray(b"\0" + b" " * 511)
    blk[0:len(nb) + 1] = nb + b"\0"
    blk[100:108] = b"%08o" % 420
    blk[108:116] = b"%08o" % 0
    blk[116:124] = b"%08o" % 0
    blk[124:136] = b"%012o" % fsize
    blk[136:148] = b"%012o" % mtime
    blk[148:156] = b" " * 8
    blk[156:157] = b"0"
    for pos, s in ((157, b"\0"), (257, b"ustar"), (265, b"root"), (297, b"root"), (329, b"\0"), (337, b"\0"),
                   (345, b"\0")):
        blk[pos:pos + len(s) + 1] = s + b"\0"   # linkname, magic, uname, gname, devmajor, devminor, prefix
    blk[148:156] = b"%08o" % sum(blk)
    return bytes(blk)


def tar_member(name: str, h: SacHeader, data) -> bytes:
    """sac__wtar (m_sac.f90:812-829): header, SAC bytes, then zeros up to the
    next 512-byte boundary -- a whole zero block when the size is already a
    multiple of 512 (tar__wpad, m_tar.f90:214-225)."""
    body = sac_bytes(h, data)
    fsize = (158 + h.npts) * 4
    return tar_header(name, fsize, h.tim) + body + bytes(512 - fsize % 512)


TAR_END = bytes(1024)   # tar__wend (m_tar.f90:227-234)


# ---- trace headers of a run ----
def moment_magnitude(m0: float) -> float:
    """moment_magnitude (src/shared/m_fdtool.f90:281-293), single precision"""
    m = np.float32(m0)
    if m < np.finfo(np.float32).eps:
        return UNDEF_F
    return float((np.float32(np.log10(m)) - np.float32(9.1)) * np.float32(2.0) / np.float32(3.0))


def _localtime(tim: int):
    """daytim__localtime + daytim__ymd2jul: the host's local time of tim"""
    import time
    t = time.localtime(tim)
    return dict(nzyear=t.tm_year, nzjday=t.tm_yday, nzhour=t.tm_hour, nzmin=t.tm_min, nzsec=t.tm_sec, nzmsec=0)


@dataclass
class RunInfo:
    """Run-level values every trace header carries (m_wav.f90:353-386)."""
    title: str
    exedate: int                   # seconds since 1970 (the run's start)
    tbeg: float
    dt: float
    ntdec_w: int
    ntw: int
    M0: float
    evlo: float = UNDEF_F
    evla: float = UNDEF_F
    evdp: float = UNDEF_F
    sx0: float = 0.0               # source x, y for distance and azimuths
    sy0: float = 0.0
    mij0: tuple | None = None      # mxx myy mzz myz mxz mxy of the source
    f0: tuple | None = None        # fx fy fz (body-force runs)
    clon: float = UNDEF_F
    clat: float = UNDEF_F
    phi: float = 0.0
    otim: float = UNDEF_F


WAV_COMPONENTS = {"vel": (("Vx", "Vy", "Vz"), 7), "disp": (("Ux", "Uy", "Uz"), 6),
                  "stress": (("Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy"), 5),
                  "strain": (("Exx", "Eyy", "Ezz", "Eyz", "Exz", "Exy"), 5)}


def station_headers(run: RunInfo, kind: str, stnm: str, xst: float, yst: float, zst: float,
                    stlo: float = UNDEF_F, stla: float = UNDEF_F):
    """Headers of one station's traces of one kind: initialize_sac_header and
    set_sac_header (m_wav.f90:272-396)."""
    f32 = np.float32
    base = SacHeader(evlo=run.evlo, evla=run.evla, evdp=run.evdp, tim=run.exedate, b=run.tbeg,
                     delta=float(f32(run.ntdec_w) * f32(run.dt)), npts=run.ntw, mag=moment_magnitude(run.M0),
                     user6=run.clon, user7=run.clat, user8=run.phi, o=run.otim, **_localtime(run.exedate))
    if run.f0 is not None:
        base = base.copy(user0=run.f0[0], user1=run.f0[1], user2=run.f0[2])
    elif run.mij0 is not None:
        base = base.copy(**{f"user{q}": run.mij0[q] for q in range(6)})
    xs, ys, x0, y0 = f32(xst), f32(yst), f32(run.sx0), f32(run.sy0)
    r2d = f32(180.0 / math.pi)
    base = base.copy(kevnm=run.title[:16].strip(), kstnm=stnm.strip(), stlo=stlo, stla=stla,
                     stdp=float(f32(zst) * f32(1000)), lcalda=False,
                     dist=float(np.sqrt((x0 - xs) ** 2 + (y0 - ys) ** 2)),
                     az=float(np.arctan2(ys - y0, xs - x0) * r2d), baz=float(np.arctan2(y0 - ys, x0 - xs) * r2d))
    names, idep = WAV_COMPONENTS[kind]
    out = []
    for q, c in enumerate(names):
        h = base.copy(kcmpnm=c, idep=idep)
        if kind in ("vel", "disp"):
            h.cmpinc = 90.0
            h.cmpaz = float(f32(0.0) + f32(run.phi)) if q == 0 else (
                float(f32(90.0) + f32(run.phi)) if q == 1 else 0.0)
        out.append(h)
    return out


def write_wav(odir: str, run: RunInfo, stations, traces: dict, fmt: str = "sac") -> list:
    """wav__write (m_wav.f90:658-765): stations is a list of
    (stnm, x, y, z[, stlo, stla]); traces maps a kind to (nst, ncmp, ntw)
    arrays as swpc_get_seismograms returns them.  fmt: 'sac', 'csf',
    'tar_st' or 'tar_node'.  Returns the files written."""
    os.makedirs(os.path.join(odir, "wav"), exist_ok=True)
    kinds = [k for k in ("vel", "disp", "stress", "strain") if k in traces]
    heads = {k: [station_headers(run, k, *st) for st in stations] for k in kinds}
    files = []
    if fmt == "sac":
        for n, st in enumerate(stations):
            for k in kinds:
                for q, h in enumerate(heads[k][n]):
                    fn = os.path.join(odir, "wav", f"{run.title}.3d.{h.kstnm}.{h.kcmpnm}.sac")
                    write_sac(fn, h, traces[k][n, q])
                    files.append(fn)
    elif fmt == "csf":
        fn = os.path.join(odir, "wav", f"{run.title}__00000__.csf")
        for k in kinds:
            hs = [h for per in heads[k] for h in per]
            write_csf(fn, hs, traces[k].reshape(-1, run.ntw))
        files.append(fn)
    elif fmt in ("tar_st", "tar_node"):
        members = {}
        for n, st in enumerate(stations):
            buf = []
            for k in kinds:
                for q, h in enumerate(heads[k][n]):
                    buf.append(tar_member(f"{run.title}.{h.kstnm}.3d.{h.kcmpnm}.sac", h, traces[k][n, q]))
            members[st[0].strip()] = b"".join(buf)
        if fmt == "tar_st":
            for stnm, body in members.items():
                fn = os.path.join(odir, "wav", f"{run.title}.3d.{stnm}.sac.tar")
                _write_in_place(fn, body + TAR_END)
                files.append(fn)
        else:
            fn = os.path.join(odir, "wav", f"{run.title}.3d.000000.sac.tar")
            _write_in_place(fn, b"".join(members.values()) + TAR_END)
            files.append(fn)
    else:
        raise ValueError(f"wav_format {fmt!r}")
    return files


GREEN_CMP = ("mxx", "myy", "mzz", "myz", "mxz", "mxy", "fx_", "fy_", "fz_")


def green_headers(run: RunInfo, green_stnm: str, green_cmp: str, ncmp: int, gid: int, point, cell_xyz,
                  evlo0: float = UNDEF_F, evla0: float = UNDEF_F, zsrc: float = 0.0, long: float = UNDEF_F,
                  latg: float = UNDEF_F):
    """Headers of one GF point's traces (m_green.f90:298-401): the virtual
    source (station) as the SAC station, the GF point as the event; point =
    (x, y, z) as listed, cell_xyz = the cell centre it maps to."""
    cid8 = "%08d" % gid
    base = SacHeader(stlo=evlo0, stla=evla0, stdp=float(np.float32(zsrc) * np.float32(1000)), kevnm=cid8,
                     kstnm=green_stnm.strip(), evlo=long, evla=latg, evdp=point[2], tim=run.exedate, nzmsec=0,
                     b=run.tbeg, delta=float(np.float32(run.ntdec_w) * np.float32(run.dt)), npts=run.ntw,
                     user0=point[0], user1=point[1], user2=point[2], user3=cell_xyz[0], user4=cell_xyz[1],
                     user5=cell_xyz[2], user6=run.clon, user7=run.clat, user8=run.phi,
                     **{k: v for k, v in _localtime(run.exedate).items() if k != "nzmsec"})
    out = []
    for j in range(ncmp):
        h = base.copy(idep=7 if j < 6 else 6, kcmpnm=f"G_V{green_cmp}_{GREEN_CMP[j]}")
        if green_cmp == "x":
            h.cmpinc, h.cmpaz = 90.0, float(np.float32(0.0) + np.float32(run.phi))
        elif green_cmp == "y":
            h.cmpinc, h.cmpaz = 90.0, float(np.float32(90.0) + np.float32(run.phi))
        else:
            h.cmpinc, h.cmpaz = 0.0, 0.0
        out.append(h)
    return out


def write_green(odir: str, run: RunInfo, green_stnm: str, green_cmp: str, gf, gids, points, cells,
                fmt: str = "sac", **station) -> list:
    """green__export (m_green.f90:553-612) for the 'sac' and 'csf' formats:
    gf is (ng*ncmp, ntw) as swpc_get_green returns it; the sign is flipped
    for green_cmp = 'z'."""
    gf = np.asarray(gf, np.float32)
    ng = len(gids)
    ncmp = gf.shape[0] // max(ng, 1)
    if green_cmp == "z":
        gf = -gf
    d = os.path.join(odir, "green", green_stnm.strip())
    os.makedirs(d, exist_ok=True)
    heads = [green_headers(run, green_stnm, green_cmp, ncmp, gids[i], points[i], cells[i], **station)
             for i in range(ng)]
    files = []
    if fmt == "sac":
        for i in range(ng):
            for j in range(ncmp):
                fn = os.path.join(d, f"{run.title}__{'%08d' % gids[i]}__{green_stnm.strip()}__{green_cmp}__"
                                     f"{GREEN_CMP[j].strip()}__.sac")
                write_sac(fn, heads[i][j], gf[i * ncmp + j])
                files.append(fn)
    elif fmt == "csf":
        if ng:
            fn = os.path.join(d, f"{run.title}__{green_stnm.strip()}__{green_cmp}__000000__.csf")
            write_csf(fn, [h for hs in heads for h in hs], gf)
            files.append(fn)
    else:
        raise ValueError(f"green wav_format {fmt!r} (sac or csf)")
    return files
