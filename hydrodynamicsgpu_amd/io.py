"""On-disk formats: per-channel FITS images as Solver::save writes them, a
raw checkpoint that resumes a run exactly, and the solid-cell image a config
names with `solidFilename` (PNG).

FITS (src/Solver/Solver.cpp:547-589): one file per state channel, named
`<channel name><index>.fits` with the channel names of the equation
(Euler: DENSITY, MOMENTUM_X[, _Y, _Z], ENERGY_TOTAL, src/Equation/Euler.cpp:36-40;
SRHD: REST_MASS_DENSITY, MOMENTUM_DENSITY_X.., TOTAL_ENERGY_DENSITY,
src/Equation/SRHD.cpp:36-40; Maxwell: ELECTRIC_X.., MAGNETIC_X..,
src/Equation/Maxwell.cpp:39-46) and the first index whose file for the first
channel does not exist yet (Solver::getSaveIndex).  The image is the full
grid, ghost cells included, as 32-bit floats, x fastest.  The reference
hands the image to its own image library, which is not part of the
reference sources; the files here are single primary-HDU FITS images
(BITPIX -32, NAXIS1 = size x, NAXIS2 = size y, NAXIS3 = size z when the grid
is 3D), which any FITS reader opens.

Checkpoint (this build's own format, little-endian):
    magic   8 bytes  b"HGPUCKP1"
    u32     version (1)
    u32     number of states N
    u64     volume V (cells, ghosts included)
    f64     simulated time
    i64     completed steps
    u32     flags: bit 0 SRHD primitives follow, bit 1 FP32 low words follow
    u32     size of the descriptor record that follows
    bytes   the hydro_sim_desc record (include/hydro_capi.h)
    f64[N*V] state, reference AoS layout state[s + N*cell] (compensated FP32
             storage: its high words)
    f32[N*V] low words of compensated FP32 storage (flag bit 1)
    f64[N*V] primitives (flag bit 0)
The C++ host layer writes and reads the same layout
(HydroGPU::Solver::Solver::saveCheckpoint / loadCheckpoint).
"""
import ctypes
import os
import struct

import numpy as np

from . import config

_MAGIC = b"HGPUCKP1"
_HEAD = struct.Struct("<8sIIQdqII")

EULER_NAMES = ["DENSITY", "MOMENTUM_X", "MOMENTUM_Y", "MOMENTUM_Z", "ENERGY_TOTAL"]
SRHD_NAMES = ["REST_MASS_DENSITY", "MOMENTUM_DENSITY_X", "MOMENTUM_DENSITY_Y", "MOMENTUM_DENSITY_Z",
              "TOTAL_ENERGY_DENSITY"]
MAXWELL_NAMES = ["ELECTRIC_X", "ELECTRIC_Y", "ELECTRIC_Z", "MAGNETIC_X", "MAGNETIC_Y", "MAGNETIC_Z"]


def channel_names(desc):
    """Equation::states of the descriptor's equation, in state order."""
    if desc.equation == config.EQUATIONS["Maxwell"]:
        return list(MAXWELL_NAMES)
    names = SRHD_NAMES if desc.equation == config.EQUATIONS["SRHD"] else EULER_NAMES
    return names[:1 + desc.dim] + names[-1:]


# ---------------------------------------------------------------- FITS
def _card(key, value, comment=None):
    if isinstance(value, bool):
        v = "%20s" % ("T" if value else "F")
    elif isinstance(value, int):
        v = "%20d" % value
    else:
        v = "%20s" % value
    s = "%-8s= %s" % (key, v)
    if comment:
        s += " / " + comment
    return s[:80].ljust(80).encode("ascii")


def fits_bytes(image):
    """A primary-HDU FITS file of a float32 image shaped (z, y, x) or (y, x)."""
    a = np.asarray(image, dtype=np.float32)
    shape = a.shape[::-1]  # NAXIS1 = x, fastest
    cards = [_card("SIMPLE", True, "conforms to FITS standard"), _card("BITPIX", -32, "IEEE single precision"),
             _card("NAXIS", len(shape))]
    for i, n in enumerate(shape):
        cards.append(_card("NAXIS%d" % (i + 1), int(n)))
    cards.append(_card("EXTEND", True))
    cards.append(b"END".ljust(80))
    head = b"".join(cards)
    head += b" " * (-len(head) % 2880)
    data = a.astype(">f4").tobytes()
    data += b"\0" * (-len(data) % 2880)
    return head + data


def read_fits(path):
    """(header dict, float array shaped (..., NAXIS2, NAXIS1)) of a primary
    HDU image with BITPIX -32 or -64."""
    with open(path, "rb") as f:
        raw = f.read()
    hdr = {}
    pos = 0
    while True:
        card = raw[pos:pos + 80].decode("ascii")
        pos += 80
        key = card[:8].strip()
        if key == "END":
            break
        if card[8:10] == "= ":
            val = card[10:].split("/")[0].strip()
            if val in ("T", "F"):
                hdr[key] = val == "T"
            else:
                try:
                    hdr[key] = int(val)
                except ValueError:
                    hdr[key] = val
    pos += -pos % 2880
    shape = [hdr["NAXIS%d" % (i + 1)] for i in range(hdr["NAXIS"])][::-1]
    dt = {-32: ">f4", -64: ">f8"}[hdr["BITPIX"]]
    n = int(np.prod(shape))
    data = np.frombuffer(raw, dtype=dt, count=n, offset=pos).astype(np.float64).reshape(shape)
    return hdr, data


def save_index(desc, directory="."):
    """Solver::getSaveIndex: the first i with no `<first channel><i>.fits`."""
    first = channel_names(desc)[0]
    for i in range(1000000):
        if not os.path.exists(os.path.join(directory, "%s%d.fits" % (first, i))):
            return i
    raise RuntimeError("failed to find available save filename")


def save_fits(desc, state, directory=".", index=None):
    """Solver::save: one FITS image per state channel; returns the paths."""
    N = config.num_states(desc)
    sx, sy, sz = desc.size[0], desc.size[1], desc.size[2]
    q = np.asarray(state, dtype=np.float64).reshape(sz, sy, sx, N)
    if index is None:
        index = save_index(desc, directory)
    paths = []
    for c, name in enumerate(channel_names(desc)):
        img = q[..., c] if sz > 1 else q[0, ..., c]
        path = os.path.join(directory, "%s%d.fits" % (name, index))
        with open(path, "wb") as f:
            f.write(fits_bytes(img))
        paths.append(path)
    return paths


# ---------------------------------------------------------------- checkpoint
def checkpoint_bytes(desc, state, t, steps, prims=None, low=None):
    N = config.num_states(desc)
    V = config.volume(desc)
    q = np.ascontiguousarray(state, dtype="<f8")
    if q.size != N * V:
        raise ValueError("state size must be N * volume")
    rec = bytes(desc)
    flags = (1 if prims is not None else 0) | (2 if low is not None else 0)
    out = [_HEAD.pack(_MAGIC, 1, N, V, float(t), int(steps), flags, len(rec)), rec, q.tobytes()]
    if low is not None:
        lo = np.ascontiguousarray(low, dtype="<f4")
        if lo.size != N * V:
            raise ValueError("low words must have the state's size")
        out.append(lo.tobytes())
    if prims is not None:
        p = np.ascontiguousarray(prims, dtype="<f8")
        if p.size != N * V:
            raise ValueError("primitives must have the state's size")
        out.append(p.tobytes())
    return b"".join(out)


def parse_checkpoint(raw):
    """dict(desc, state, t, steps, prims, low) from checkpoint bytes (prims
    and low None when absent)."""
    if len(raw) < _HEAD.size:
        raise ValueError("not a checkpoint (too short)")
    magic, ver, N, V, t, steps, flags, nrec = _HEAD.unpack_from(raw, 0)
    if magic != _MAGIC or ver != 1:
        raise ValueError("not a checkpoint of this format")
    if nrec != ctypes.sizeof(config.HydroSimDesc):
        raise ValueError("descriptor record of %d bytes, expected %d" % (nrec, ctypes.sizeof(config.HydroSimDesc)))
    pos = _HEAD.size
    desc = config.HydroSimDesc.from_buffer_copy(raw[pos:pos + nrec])
    pos += nrec
    if config.num_states(desc) != N or config.volume(desc) != V:
        raise ValueError("checkpoint header does not match its descriptor")
    need = pos + N * V * (8 + (4 if flags & 2 else 0) + (8 if flags & 1 else 0))
    if len(raw) != need:
        raise ValueError("checkpoint is %d bytes, expected %d" % (len(raw), need))
    out = {"desc": desc, "t": t, "steps": steps, "prims": None, "low": None}
    out["state"] = np.frombuffer(raw, dtype="<f8", count=N * V, offset=pos).astype(np.float64)
    pos += 8 * N * V
    if flags & 2:
        out["low"] = np.frombuffer(raw, dtype="<f4", count=N * V, offset=pos).astype(np.float32)
        pos += 4 * N * V
    if flags & 1:
        out["prims"] = np.frombuffer(raw, dtype="<f8", count=N * V, offset=pos).astype(np.float64)
    return out


def _compensated(desc):
    return desc.precision == config.PRECISION["fp32_mixed"] and bool(desc.flags & config.FLAG_KAHAN)


def save_checkpoint(sim, path):
    """Write a Sim's state in its stored form (high and low words for
    compensated FP32 storage), primitives (SRHD) and clock."""
    t, steps, _ = sim.time()
    prims = sim.get_primitives() if sim.desc.equation == config.EQUATIONS["SRHD"] else None
    if _compensated(sim.desc):
        hi, lo = sim.get_storage()
    else:
        hi, lo = sim.get_state(), None
    with open(path, "wb") as f:
        f.write(checkpoint_bytes(sim.desc, hi, t, steps, prims, lo))


def load_checkpoint(path, device=0, desc=None):
    """A Sim resumed from a checkpoint (the saved descriptor unless `desc`
    is given, e.g. to continue in another precision)."""
    from .capi import Sim

    with open(path, "rb") as f:
        c = parse_checkpoint(f.read())
    d = desc if desc is not None else c["desc"]
    sim = Sim(d, device)
    state, low = c["state"], c["low"]
    if low is not None and not _compensated(d):
        state, low = state + low.astype(np.float64), None
    sim.restore(state, c["t"], c["steps"], prims=c["prims"], low=low)
    return sim


# ---------------------------------------------------------------- solid image
def read_png(path_or_bytes):
    """A PNG image as a uint8 array (height, width, channels): 8-bit
    greyscale, grey+alpha, RGB, RGBA or palette (expanded to RGB),
    non-interlaced -- the forms an image library reads a mask from.  Rows
    are un-filtered as the PNG specification defines (None, Sub, Up,
    Average, Paeth)."""
    import struct
    import zlib

    raw = path_or_bytes if isinstance(path_or_bytes, (bytes, bytearray)) else open(path_or_bytes, "rb").read()
    if raw[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos, idat, plte, hdr = 8, [], None, None
    while pos < len(raw):
        (n,) = struct.unpack(">I", raw[pos:pos + 4])
        kind, body = raw[pos + 4:pos + 8], raw[pos + 8:pos + 8 + n]
        pos += 12 + n
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"PLTE":
            plte = np.frombuffer(body, dtype=np.uint8).reshape(-1, 3)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
    if hdr is None:
        raise ValueError("PNG without IHDR")
    w, h, depth, ctype, _, _, interlace = hdr
    chans = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}.get(ctype)
    if depth != 8 or chans is None or interlace != 0:
        raise ValueError("PNG form not read here (bit depth %d, colour type %d, interlace %d)" % (depth, ctype,
                                                                                                  interlace))
    data = zlib.decompress(b"".join(idat))
    stride = w * chans
    out = np.zeros((h, stride), dtype=np.uint8)
    prev = np.zeros(stride, dtype=np.int32)
    for y in range(h):
        f = data[y * (stride + 1)]
        line = np.frombuffer(data, dtype=np.uint8, count=stride, offset=y * (stride + 1) + 1).astype(np.int32)
        cur = np.zeros(stride, dtype=np.int32)
        if f == 0:
            cur = line
        elif f == 2:
            cur = (line + prev) & 255
        elif f in (1, 3, 4):
            for i in range(stride):
                a = cur[i - chans] if i >= chans else 0
                b = prev[i]
                if f == 1:
                    p = a
                elif f == 3:
                    p = (a + b) >> 1
                else:
                    c = prev[i - chans] if i >= chans else 0
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
                cur[i] = (line[i] + p) & 255
        else:
            raise ValueError("PNG row filter %d" % f)
        out[y] = cur
        prev = cur
    img = out.reshape(h, w, chans)
    if ctype == 3:
        if plte is None:
            raise ValueError("palette PNG without PLTE")
        img = plte[img[:, :, 0]]
    return img


def solid_from_image(desc, img):
    """SelfGravitation::resetState's solid mask from an image
    (src/Solver/SelfGravitation.cpp:56-74): every cell (ghosts included)
    samples pixel (x * W / size_x, H - 1 - y * H / size_y) of plane
    z * planes / size_z, channel 0, solid where it exceeds 127.  `img` is
    (H, W, channels) or (planes, H, W, channels)."""
    a = np.asarray(img)
    if a.ndim == 3:
        a = a[None]
    planes, H, W = a.shape[0], a.shape[1], a.shape[2]
    sx, sy, sz = desc.size[0], desc.size[1], desc.size[2]
    xs = (np.arange(sx) * W) // sx
    ys = H - 1 - (np.arange(sy) * H) // sy
    zs = (np.arange(sz) * planes) // sz
    v = a[zs[:, None, None], ys[None, :, None], xs[None, None, :], 0]
    return (v > 127).astype(np.uint8).reshape(-1)
