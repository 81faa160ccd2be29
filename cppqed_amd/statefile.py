"""Trajectory state files (``<trajectory file>.state``): checkpoint and resume.

The reference appends one self-contained archive per saved instant to the state
file (``writeViaSStream`` / ``readViaSStream``, CPPQEDutils/Trajectory.h:353-410):
the archive's byte length in decimal, then a Boost.Serialization binary archive,
and the whole file is a bzip2 stream (``openStateFileWriting``,
CPPQEDutils/Trajectory.cc:40-72; every appending run adds a stream, so a file
may hold several concatenated bzip2 streams).  Reading keeps the last archive
(``run`` reads until end of file, Trajectory.h:492-500).

Archive content (each class's one-time header is written where the class first
occurs in the archive):

* header: ``serialization::archive``, library version 18, the sizes of int,
  long, float, double (4, 8, 4, 8) and an int 1;
* ``SerializationMetadata`` (Trajectory.h:135-160): protocol version, rank,
  typeID (``"CArray"``), trajectoryID (``"MCWF_Trajectory"``,
  ``"Ensemble of MCWF_Trajectory"``, ``"Master"`` or ``"ArrayOnly"``);
* then, per trajectory (an ensemble writes all of them in order,
  StochasticTrajectory.h:238), ``MCWF_Trajectory::stateIO``
  (MCWF_Trajectory.h:134-147): state array, time, ODE logger (derivative calls,
  steps, failed steps, ODE.h:99-122), dtDid, dtTry, random engine (pcg64's
  text form, Random.h:208-240), MCWF logger (steps, dpLimit overshoots,
  tolerance overshoots, the three maxima and the jump list,
  MCWF_TrajectoryLogger.h:59-101); ``Master::stateIO`` (Master.h:89-102) stops
  after dtTry;
* the state array is a Blitz++ ``Array<complex<double>,R>``: its memory block
  (through a tracked pointer: class id, object id, element count, data), the
  offset of the view into the block, extents, storage (padding policy,
  ascending flags, ordering, bases), strides and the zero offset.  Only the reference's own
  layout (row-major, ascending, zero-based, contiguous) is accepted.

The device ensembles draw from a Philox4x64-10 counter generator instead of
pcg64; they store ``"philox4x64-10 <seed> <key> <ordinal>"`` in the random-engine
record, which the reference's pcg64 reader rejects with its own error rather
than continuing from a wrong stream.
"""
import bz2
import re
import struct
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

LIBRARY_VERSION = 18
ARRAY_ONLY = "ArrayOnly"
MCWF_ID = "MCWF_Trajectory"
ENSEMBLE_ID = "Ensemble of MCWF_Trajectory"
MASTER_ID = "Master"


class StateFileError(RuntimeError):
    pass


@dataclass
class TrajRecord:
    """One ``MCWF_Trajectory`` archive body (MCWF_Trajectory.h:134-147)."""
    psi: np.ndarray                       # complex128, shape = subsystem dims
    t: float = 0.0
    dtDid: float = 0.0
    dtTry: float = 0.0
    ode_derivs: int = 0                   # ODE logger (ODE.h:117)
    ode_steps: int = 0
    ode_failed: int = 0
    rng: str = ""                         # random-engine record (text)
    mcwf_steps: int = 0                   # MCWF logger (MCWF_TrajectoryLogger.h:86-88)
    overshot: int = 0
    tolerance_overshot: int = 0
    dp_max_overshoot: float = 0.0
    dp_tolerance_max_overshoot: float = 0.0
    norm_max_deviation: float = 0.0
    jumps: List[Tuple[float, int]] = field(default_factory=list)


@dataclass
class MasterRecord:
    """One ``Master`` archive body (Master.h:89-102); rho has shape dims + dims."""
    rho: np.ndarray
    t: float = 0.0
    dtDid: float = 0.0
    dtTry: float = 0.0
    ode_derivs: int = 0
    ode_steps: int = 0
    ode_failed: int = 0

    def matrix(self):
        d = int(np.prod(self.rho.shape[: self.rho.ndim // 2]))
        return self.rho.reshape(d, d)


@dataclass
class Frame:
    """One archive of a state file."""
    trajectory_id: str
    type_id: str = "CArray"
    rank: int = 0
    protocol: int = 0
    trajs: List[TrajRecord] = field(default_factory=list)    # MCWF_Trajectory / Ensemble
    master: Optional[MasterRecord] = None                      # Master
    array: Optional[np.ndarray] = None                         # ArrayOnly


# ------------------------------------------------------------------ binary archive

class _Writer:
    def __init__(self):
        self.b = bytearray()
        self.seen = set()
        self.objects = 0

    def raw(self, fmt, *v):
        self.b += struct.pack("<" + fmt, *v)

    def string(self, s):
        e = s.encode()
        self.raw("Q", len(e))
        self.b += e

    def cls(self, key):
        # first occurrence of a class: tracking flag and class version (both 0)
        if key not in self.seen:
            self.seen.add(key)
            self.b += b"\0" * 5

    def header(self):
        self.string("serialization::archive")
        self.raw("H", LIBRARY_VERSION)
        self.b += bytes([4, 8, 4, 8])
        self.raw("i", 1)

    def metadata(self, rank, type_id, traj_id, protocol=0):
        self.cls("meta")
        self.raw("ii", protocol, rank)
        self.string(type_id)
        self.string(traj_id)

    def array(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        r = a.ndim
        self.cls("array")
        self.cls("blockref")
        # tracked pointer to the memory block: class id 3, one-time (tracking 1, version 0), object id
        self.raw("h", 3)
        if "block" not in self.seen:
            self.seen.add("block")
            self.raw("BI", 1, 0)
        self.raw("I", self.objects)
        self.objects += 1
        self.raw("Q", a.size)
        self.b += a.tobytes()
        self.raw("Q", 0)                                   # view offset into the block
        self.cls("tv_int")
        self.raw("Q", r)
        self.raw("%di" % r, *a.shape)                     # extents
        self.cls("storage")
        self.raw("i", 0)                                   # padding policy: contiguous
        self.cls("tv_bool")
        self.raw("Q", r)
        self.b += b"\1" * r                                # ascending
        self.raw("Q", r)
        self.raw("%di" % r, *range(r - 1, -1, -1))        # ordering: row major
        self.raw("Q", r)
        self.raw("%di" % r, *([0] * r))                   # bases
        self.cls("tv_long")
        self.raw("Q", r)
        strides = [int(np.prod(a.shape[k + 1:])) for k in range(r)]
        self.raw("%dq" % r, *strides)
        self.raw("q", 0)                                   # zero offset

    def ode(self, rec):
        self.cls("odelog")
        self.raw("QQQ", rec.ode_derivs, rec.ode_steps, rec.ode_failed)
        self.raw("dd", rec.dtDid, rec.dtTry)

    def traj(self, rec):
        self.array(rec.psi)
        self.raw("d", rec.t)
        self.ode(rec)
        self.cls("rng")
        self.string(rec.rng)
        self.cls("mcwflog")
        self.raw("QQQ", rec.mcwf_steps, rec.overshot, rec.tolerance_overshot)
        self.raw("ddd", rec.dp_max_overshoot, rec.dp_tolerance_max_overshoot, rec.norm_max_deviation)
        self.cls("list")
        self.raw("QI", len(rec.jumps), 0)
        for (tj, ch) in rec.jumps:
            self.cls("pair")
            self.raw("dQ", tj, ch)


class _Reader:
    def __init__(self, b):
        self.b = b
        self.p = 0
        self.seen = set()

    def raw(self, fmt):
        fmt = "<" + fmt
        n = struct.calcsize(fmt)
        if self.p + n > len(self.b):
            raise StateFileError("truncated archive")
        v = struct.unpack_from(fmt, self.b, self.p)
        self.p += n
        return v if len(v) > 1 else v[0]

    def string(self):
        n = self.raw("Q")
        s = self.b[self.p:self.p + n]
        self.p += n
        return s.decode()

    def cls(self, key):
        if key not in self.seen:
            self.seen.add(key)
            if self.b[self.p:self.p + 5] != b"\0" * 5:
                raise StateFileError("unexpected class header for %s" % key)
            self.p += 5

    def header(self):
        if self.string() != "serialization::archive":
            raise StateFileError("not a binary archive")
        lv = self.raw("H")
        sizes = bytes(self.b[self.p:self.p + 4])
        self.p += 4
        one = self.raw("i")
        if sizes != bytes([4, 8, 4, 8]) or one != 1:
            raise StateFileError("archive from a platform with other type sizes or byte order")
        return lv

    def metadata(self):
        self.cls("meta")
        protocol, rank = self.raw("ii")
        return protocol, rank, self.string(), self.string()

    def array(self):
        self.cls("array")
        self.cls("blockref")
        if self.raw("h") != 3:
            raise StateFileError("unexpected memory-block class id")
        if "block" not in self.seen:
            self.seen.add("block")
            self.raw("BI")
        self.raw("I")                                       # object id
        n = self.raw("Q")
        data = np.frombuffer(bytes(self.b[self.p:self.p + 16 * n]), dtype=np.complex128).copy()
        self.p += 16 * n
        off = self.raw("Q")
        self.cls("tv_int")
        r = self.raw("Q")
        ext = self.raw("%di" % r)
        ext = (ext,) if r == 1 else tuple(ext)
        self.cls("storage")
        padding = self.raw("i")
        self.cls("tv_bool")
        rb = self.raw("Q")
        asc = bytes(self.b[self.p:self.p + rb])
        self.p += rb
        self.raw("Q")
        order = self.raw("%di" % r)
        order = (order,) if r == 1 else tuple(order)
        self.raw("Q")
        base = self.raw("%di" % r)
        base = (base,) if r == 1 else tuple(base)
        self.cls("tv_long")
        self.raw("Q")
        strides = self.raw("%dq" % r)
        strides = (strides,) if r == 1 else tuple(strides)
        zero = self.raw("q")
        want = tuple(int(np.prod(ext[k + 1:])) for k in range(r))
        if (off != 0 or padding != 0 or asc != b"\1" * r or order != tuple(range(r - 1, -1, -1)) or any(base)
                or strides != want or zero != 0 or n != int(np.prod(ext))):
            raise StateFileError("only contiguous row-major zero-based arrays are supported")
        return data.reshape(ext)

    def ode(self, rec):
        self.cls("odelog")
        rec.ode_derivs, rec.ode_steps, rec.ode_failed = self.raw("QQQ")
        rec.dtDid, rec.dtTry = self.raw("dd")

    def traj(self):
        rec = TrajRecord(psi=self.array())
        rec.t = self.raw("d")
        self.ode(rec)
        self.cls("rng")
        rec.rng = self.string()
        self.cls("mcwflog")
        rec.mcwf_steps, rec.overshot, rec.tolerance_overshot = self.raw("QQQ")
        rec.dp_max_overshoot, rec.dp_tolerance_max_overshoot, rec.norm_max_deviation = self.raw("ddd")
        self.cls("list")
        n, _item_version = self.raw("QI")
        for _ in range(n):
            self.cls("pair")
            tj, ch = self.raw("dQ")
            rec.jumps.append((tj, int(ch)))
        return rec


def encode_frame(fr: Frame) -> bytes:
    """One archive (without the decimal length prefix)."""
    w = _Writer()
    w.header()
    if fr.trajectory_id == ARRAY_ONLY:
        w.metadata(fr.array.ndim, fr.type_id, ARRAY_ONLY, fr.protocol)
        w.array(fr.array)
    elif fr.trajectory_id == MASTER_ID:
        m = fr.master
        w.metadata(m.rho.ndim, fr.type_id, MASTER_ID, fr.protocol)
        w.array(m.rho)
        w.raw("d", m.t)
        w.ode(m)
    elif fr.trajectory_id in (MCWF_ID, ENSEMBLE_ID):
        if fr.trajectory_id == MCWF_ID and len(fr.trajs) != 1:
            raise StateFileError("an MCWF_Trajectory archive holds one trajectory")
        w.metadata(fr.trajs[0].psi.ndim, fr.type_id, fr.trajectory_id, fr.protocol)
        for rec in fr.trajs:
            w.traj(rec)
    else:
        raise StateFileError("unsupported trajectory type %r" % fr.trajectory_id)
    return bytes(w.b)


def decode_frame(b: bytes) -> Frame:
    r = _Reader(b)
    r.header()
    protocol, rank, type_id, traj_id = r.metadata()
    fr = Frame(trajectory_id=traj_id, type_id=type_id, rank=rank, protocol=protocol)
    if type_id != "CArray":
        raise StateFileError("unsupported state type %r" % type_id)
    if traj_id == ARRAY_ONLY:
        fr.array = r.array()
    elif traj_id == MASTER_ID:
        m = MasterRecord(rho=r.array())
        m.t = r.raw("d")
        r.ode(m)
        fr.master = m
    elif traj_id == MCWF_ID:
        fr.trajs.append(r.traj())
    elif traj_id == ENSEMBLE_ID:
        while r.p < len(b):
            fr.trajs.append(r.traj())
    else:
        raise StateFileError("unsupported trajectory type %r" % traj_id)
    if r.p != len(b):
        raise StateFileError("trailing bytes in archive")
    return fr


def split_frames(data: bytes) -> List[bytes]:
    """The archives of a decompressed state stream (readViaSStream, Trajectory.h:392-405)."""
    out, p = [], 0
    while p < len(data):
        m = re.match(rb"\s*(\d+)", data[p:])
        if not m:
            raise StateFileError("bad archive length at byte %d" % p)
        n = int(m.group(1))
        s = p + m.end()
        if s + n > len(data):
            raise StateFileError("truncated state file")
        out.append(data[s:s + n])
        p = s + n
    return out


def join_frames(archives) -> bytes:
    return b"".join(str(len(a)).encode() + a for a in archives)


def read_stream(path) -> bytes:
    raw = open(path, "rb").read()
    return bz2.decompress(raw) if raw[:3] == b"BZh" else raw


def read(path) -> List[Frame]:
    """Every archive of a state file, oldest first."""
    return [decode_frame(a) for a in split_frames(read_stream(path))]


def read_last(path) -> Frame:
    """The archive a continuing run starts from (the last one, Trajectory.h:492-500)."""
    frames = split_frames(read_stream(path))
    if not frames:
        raise StateFileError("empty state file")
    return decode_frame(frames[-1])


def write(path, frames, append=True, compress=True):
    """Appends archives to a state file, bzip2-compressed as the reference writes them
    (a new or empty file is compressed; an existing uncompressed file stays so, Trajectory.cc:18-30)."""
    blob = join_frames([encode_frame(f) for f in frames])
    mode = "ab" if append else "wb"
    if append:
        try:
            with open(path, "rb") as f:
                head = f.read(3)
            if head and head != b"BZh":
                compress = False
        except FileNotFoundError:
            pass
    with open(path, mode) as f:
        f.write(bz2.compress(blob, 9) if compress else blob)


# ------------------------------------------------------------------ random-engine record

def parse_pcg64(text):
    """pcg64's stream form (multiplier, increment, state as decimal 128-bit integers) ->
    (state, increment)."""
    parts = text.split()
    if len(parts) != 3:
        raise StateFileError("not a pcg64 state: %r" % text)
    mult, inc, state = (int(x) for x in parts)
    return state, inc


def format_pcg64(state, inc, mult=0x2360ED051FC65DA44385DF649FCCF645):
    return "%d %d %d " % (mult, inc, state)


PHILOX_TAG = "philox4x64-10"


def format_philox(seed, key, ordinal):
    return "%s %d %d %d" % (PHILOX_TAG, seed, key, ordinal)


def parse_philox(text):
    parts = text.split()
    if len(parts) != 4 or parts[0] != PHILOX_TAG:
        return None
    return tuple(int(x) for x in parts[1:])
