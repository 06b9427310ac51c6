"""On-disk formats (CPU): FITS images as Solver::save writes them and the
checkpoint record, in Python and in the C++ host layer, byte for byte."""
import os
import subprocess

import numpy as np
import pytest

from hydrodynamicsgpu_amd import config, io

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "host_api_test")


def _binary():
    if not os.path.exists(BIN):
        subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "cpp")])
    return BIN


def test_channel_names_follow_the_equations():
    assert io.channel_names(config.make_desc([16])) == ["DENSITY", "MOMENTUM_X", "ENERGY_TOTAL"]
    assert io.channel_names(config.make_desc([16, 16, 16])) == [
        "DENSITY", "MOMENTUM_X", "MOMENTUM_Y", "MOMENTUM_Z", "ENERGY_TOTAL"]
    assert io.channel_names(config.make_desc([16, 16], solverName="SRHDRoe")) == [
        "REST_MASS_DENSITY", "MOMENTUM_DENSITY_X", "MOMENTUM_DENSITY_Y", "TOTAL_ENERGY_DENSITY"]
    assert io.channel_names(config.make_desc([16], solverName="MaxwellRoe"))[3] == "MAGNETIC_X"


@pytest.mark.parametrize("shape", [(5, 7), (3, 4, 6)])
def test_fits_round_trip(tmp_path, shape):
    a = np.random.default_rng(1).standard_normal(shape).astype(np.float32)
    raw = io.fits_bytes(a)
    assert len(raw) % 2880 == 0
    cards = [raw[i:i + 80].decode() for i in range(0, 2880, 80)]
    assert cards[0].startswith("SIMPLE  =                    T")
    assert cards[1].startswith("BITPIX  =                  -32")
    assert cards[2].startswith("NAXIS   =                    %d" % len(shape))
    assert cards[3].startswith("NAXIS1  = %20d" % shape[-1])
    p = tmp_path / "x.fits"
    p.write_bytes(raw)
    hdr, back = io.read_fits(str(p))
    assert hdr["NAXIS"] == len(shape)
    np.testing.assert_array_equal(back, a.astype(np.float64))
    # big-endian payload after the header block
    assert np.frombuffer(raw, dtype=">f4", count=a.size, offset=2880)[0] == a.reshape(-1)[0]


def test_save_fits_names_and_indices(tmp_path):
    d = config.make_desc([6, 5])
    q = np.arange(4 * 30, dtype=np.float64)
    p0 = io.save_fits(d, q, str(tmp_path))
    p1 = io.save_fits(d, q, str(tmp_path))
    assert [os.path.basename(p) for p in p0] == ["DENSITY0.fits", "MOMENTUM_X0.fits", "MOMENTUM_Y0.fits",
                                                 "ENERGY_TOTAL0.fits"]
    assert os.path.basename(p1[0]) == "DENSITY1.fits"
    _, img = io.read_fits(p0[3])
    np.testing.assert_array_equal(img, q.reshape(5, 6, 4)[..., 3])


@pytest.mark.parametrize("dims", [(7, 5, 1), (4, 3, 5)])
def test_fits_matches_cpp_writer(tmp_path, dims):
    sx, sy, sz = dims
    p = tmp_path / "c.fits"
    subprocess.check_call([_binary(), "fits", str(p), str(sx), str(sy), str(sz)])
    img = (.5 * np.arange(sx * sy * sz, dtype=np.float32) - 3).reshape((sz, sy, sx) if sz > 1 else (sy, sx))
    assert p.read_bytes() == io.fits_bytes(img)


def test_checkpoint_round_trip_and_validation():
    d = config.make_desc([12, 10], solverName="SRHDRoe")
    q = 1.0 + 1e-3 * np.arange(480)
    pr = -2.0 + 1e-2 * np.arange(480)
    lo = (1e-9 * np.arange(480)).astype(np.float32)
    raw = io.checkpoint_bytes(d, q, 0.125, 7, pr, lo)
    c = io.parse_checkpoint(raw)
    assert bytes(c["desc"]) == bytes(d) and c["t"] == 0.125 and c["steps"] == 7
    np.testing.assert_array_equal(c["state"], q)
    np.testing.assert_array_equal(c["prims"], pr)
    np.testing.assert_array_equal(c["low"], lo)
    c = io.parse_checkpoint(io.checkpoint_bytes(d, q, 0.0, 0))
    assert c["prims"] is None and c["low"] is None
    with pytest.raises(ValueError):
        io.parse_checkpoint(raw[:-8])
    with pytest.raises(ValueError):
        io.parse_checkpoint(b"NOTACKPT" + raw[8:])
    with pytest.raises(ValueError):
        io.checkpoint_bytes(d, q[:-1], 0.0, 0)


def test_checkpoint_shared_with_cpp(tmp_path):
    # C++ writes, Python reads
    p = tmp_path / "c.ckpt"
    subprocess.check_call([_binary(), "ckpt-write", str(p)])
    c = io.parse_checkpoint(p.read_bytes())
    d = c["desc"]
    assert d.dim == 2 and list(d.size)[:2] == [12, 10] and d.equation == config.EQUATIONS["SRHD"]
    assert c["t"] == 0.125 and c["steps"] == 7
    np.testing.assert_array_equal(c["state"], 1.0 + 1e-3 * np.arange(480))
    np.testing.assert_array_equal(c["prims"], -2.0 + 1e-2 * np.arange(480))
    np.testing.assert_array_equal(c["low"], np.float32(1e-9) * np.arange(480, dtype=np.float32))
    # Python writes, C++ reads
    d = config.make_desc([9, 8, 7], solverName="EulerHLLC")
    q = np.linspace(-1, 1, 5 * 504)
    p2 = tmp_path / "py.ckpt"
    p2.write_bytes(io.checkpoint_bytes(d, q, 2.5, 11))
    out = subprocess.run([_binary(), "ckpt-read", str(p2)], capture_output=True, text=True, check=True).stdout
    f = out.split()
    assert f[f.index("dim") + 1] == "3" and f[f.index("size") + 1:f.index("size") + 4] == ["9", "8", "7"]
    assert f[f.index("states") + 1] == "5" and f[f.index("volume") + 1] == "504"
    assert float(f[f.index("t") + 1]) == 2.5 and f[f.index("steps") + 1] == "11"
    assert float(f[f.index("sum") + 1]) == pytest.approx(q.sum(), abs=1e-12)
    assert f[f.index("prims") + 1] == "0" and f[f.index("low") + 1] == "0"


# ---------------------------------------------------------------- solid image
def _png(img, filters=(0,)):
    """A PNG of `img` (H, W, C uint8, C = 1, 2, 3 or 4), row r filtered with
    filters[r % len(filters)] as the PNG specification defines them."""
    import struct
    import zlib

    h, w, ch = img.shape
    ctype = {1: 0, 2: 4, 3: 2, 4: 6}[ch]
    rows = img.reshape(h, w * ch).astype(np.int32)
    out = bytearray()
    prev = np.zeros(w * ch, np.int32)
    for y in range(h):
        f = filters[y % len(filters)]
        cur = rows[y]
        left = np.concatenate([np.zeros(ch, np.int32), cur[:-ch]])
        upleft = np.concatenate([np.zeros(ch, np.int32), prev[:-ch]])
        if f == 0:
            pred = np.zeros_like(cur)
        elif f == 1:
            pred = left
        elif f == 2:
            pred = prev
        elif f == 3:
            pred = (left + prev) >> 1
        else:
            pa, pb, pc = np.abs(prev - upleft), np.abs(left - upleft), np.abs(left + prev - 2 * upleft)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, upleft))
        out.append(f)
        out += ((cur - pred) & 255).astype(np.uint8).tobytes()
        prev = cur

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0))
            + chunk(b"IDAT", zlib.compress(bytes(out))) + chunk(b"IEND", b""))


@pytest.mark.parametrize("ch", [1, 2, 3, 4])
def test_png_reader_undoes_every_row_filter(ch):
    rng = np.random.default_rng(ch)
    img = rng.integers(0, 256, (23, 17, ch), dtype=np.uint8)
    np.testing.assert_array_equal(io.read_png(_png(img, filters=(0, 1, 2, 3, 4))), img)


def test_solid_mask_from_image_follows_the_reference_sampling():
    # src/Solver/SelfGravitation.cpp:60-71: pixel (x W / sx, H - 1 - y H / sy),
    # channel 0 above 127; a 2-pixel-wide bar and a single bright pixel
    H, W = 10, 16
    img = np.zeros((H, W, 3), np.uint8)
    img[:, 4:6, 0] = 200          # bar in columns 4..5
    img[0, 15, 0] = 128           # top-right pixel (the grid's top row after the flip)
    img[:, :, 1] = 255            # other channels do not count
    d = config.make_desc([32, 20])
    m = io.solid_from_image(d, io.read_png(_png(img))).reshape(20, 32)
    for y in range(20):
        for x in range(32):
            px, py = x * W // 32, H - 1 - y * H // 20
            assert m[y, x] == (1 if img[py, px, 0] > 127 else 0), (x, y)
    # the bar covers grid columns 8..11; the pixel grid rows 18..19 x columns 30..31
    assert m[:, 8:12].all() and m[18:, 30:].all() and m.sum() == 20 * 4 + 4


def test_config_solid_filename_replaces_the_solid_channel(tmp_path):
    from hydrodynamicsgpu_amd import luaconfig

    img = np.zeros((8, 8, 3), np.uint8)
    img[2:6, 2:6, 0] = 255
    (tmp_path / "mask.png").write_bytes(_png(img))
    cfg = tmp_path / "config.lua"
    cfg.write_text("size = {24, 16}\nsolverName = 'EulerRoe'\ninitCondName = 'Sod'\n"
                   "initConds[initCondName].setup()\nsolidFilename = 'mask.png'\n")
    desc, info = luaconfig.load(str(cfg))
    q, pot, solid = info["initChannels"](desc)
    np.testing.assert_array_equal(solid, io.solid_from_image(desc, img))
    assert solid.sum() == (24 // 2) * (16 // 2)


REF_SOLID = "/root/reference/res/test-solid.png"


@pytest.mark.skipif(not os.path.exists(REF_SOLID), reason="reference sources not present")
def test_reference_solid_image_decodes():
    # the image res/config.lua names in its commented solidFilename line: a
    # 256x256 RGB PNG; the mask is its bright pixels
    img = io.read_png(REF_SOLID)
    assert img.shape == (256, 256, 3)
    frac = (img[:, :, 0] > 127).mean()
    assert 0.0 < frac < 0.5
    m = io.solid_from_image(config.make_desc([256, 256]), img).reshape(256, 256)
    np.testing.assert_array_equal(m, (img[::-1, :, 0] > 127).astype(np.uint8))
