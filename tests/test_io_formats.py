"""Input formats of the CLI (CPU only; host code of libfasttau): TIFF folders, binarize, `.sph`
sphere lists, subvolumes.

TIFF slices are written by Pillow (an independent writer) and by a minimal big-endian writer
here; the loader must return exactly the written bytes in sorted file-name order and refuse
what it cannot read as 8-bit samples.  The sphere rasterizer is checked against a float32
restatement of rasterizeSpheres (VolumeGenerator.cpp:210-251) and against geometry."""
import struct

import numpy as np
import pytest
from PIL import Image

import fast_amd as F
from fast_amd._lib import FastError


def _write_tiff_be(path, img):
    """Big-endian baseline TIFF, one strip per 3 rows, SHORT strip counts."""
    h, w = img.shape
    rows = 3
    strips = [img[r:r + rows].tobytes() for r in range(0, h, rows)]
    n = len(strips)
    entries = []
    data_off = 8
    blob = b"".join(strips)
    offs = []
    o = data_off
    for sdata in strips:
        offs.append(o)
        o += len(sdata)
    arrays = b""
    ifd_off = o + (o & 1)
    ne = 9
    after_ifd = ifd_off + 2 + 12 * ne + 4
    off_arr = after_ifd
    cnt_arr = off_arr + 4 * n
    arrays = b"".join(struct.pack(">I", v) for v in offs) + b"".join(struct.pack(">H", len(sd)) for sd in strips)

    def ent(tag, typ, count, value):
        if typ == 3 and count == 1:
            return struct.pack(">HHIHH", tag, typ, count, value, 0)
        return struct.pack(">HHII", tag, typ, count, value)

    entries = [ent(256, 4, 1, w), ent(257, 4, 1, h), ent(258, 3, 1, 8), ent(259, 3, 1, 1), ent(262, 3, 1, 1),
               ent(273, 4, n, off_arr) if n > 1 else ent(273, 4, 1, offs[0]), ent(277, 3, 1, 1), ent(278, 3, 1, rows),
               ent(279, 3, n, cnt_arr) if n > 2 else struct.pack(">HHIHH", 279, 3, n, len(strips[0]),
                                                                  len(strips[1]) if n > 1 else 0)]
    with open(path, "wb") as f:
        f.write(b"MM" + struct.pack(">HI", 42, ifd_off))
        f.write(blob)
        f.write(b"\0" * (ifd_off - o))
        f.write(struct.pack(">H", ne) + b"".join(entries) + struct.pack(">I", 0))
        f.write(arrays)


def _folder(tmp_path, vol, names):
    d = tmp_path / "slices"
    d.mkdir()
    for name, sl in zip(names, vol):
        Image.fromarray(sl, mode="L").save(d / name, tiffinfo={278: 16})   # 16 rows per strip
    return d


def test_tiff_folder_sorted_slices_and_ignored_files(tmp_path):
    rng = np.random.default_rng(1)
    vol = rng.integers(0, 256, size=(5, 257, 300), dtype=np.uint8)
    names = ["c.tif", "a.tiff", "e.tif", "b.tif", "d.tiff"]
    d = _folder(tmp_path, vol, names)
    (d / "notes.txt").write_text("not a slice")
    Image.fromarray(vol[0], mode="L").save(d / "UPPER.TIF")      # extension test is case-sensitive
    (d / "dir.tif").mkdir()                                     # directories are skipped
    got = F.loadTiffFolder(str(d))
    order = np.argsort(names)
    assert got.shape == vol.shape
    assert np.array_equal(got, vol[order])


def test_tiff_big_endian_multi_strip(tmp_path):
    rng = np.random.default_rng(2)
    vol = rng.integers(0, 256, size=(3, 11, 7), dtype=np.uint8)
    d = tmp_path / "be"
    d.mkdir()
    for k, sl in enumerate(vol):
        _write_tiff_be(d / f"s{k}.tif", sl)
    assert np.array_equal(F.loadTiffFolder(str(d)), vol)


@pytest.mark.parametrize("kind", ["lzw", "16bit", "rgb", "size_mismatch", "empty"])
def test_tiff_refused(tmp_path, kind):
    d = tmp_path / kind
    d.mkdir()
    a = np.zeros((8, 9), np.uint8)
    if kind == "lzw":
        Image.fromarray(a, mode="L").save(d / "a.tif", compression="tiff_lzw")
    elif kind == "16bit":
        Image.fromarray(a.astype(np.uint16)).save(d / "a.tif")
    elif kind == "rgb":
        Image.fromarray(np.zeros((8, 9, 3), np.uint8), mode="RGB").save(d / "a.tif")
    elif kind == "size_mismatch":
        Image.fromarray(a, mode="L").save(d / "a.tif")
        Image.fromarray(np.zeros((8, 10), np.uint8), mode="L").save(d / "b.tif")
    with pytest.raises(FastError):
        F.loadTiffFolder(str(d))


def test_binarize_reference_threshold():
    v = np.arange(256, dtype=np.uint8)
    assert np.array_equal(F.binarize(v, 1.0), np.where(v < 255, 0, 255).astype(np.uint8))
    assert np.array_equal(F.binarize(v, 0.5), np.where(v < 127, 0, 255).astype(np.uint8))   # uchar(127.5) = 127


def _raster_f32(spheres, res):
    """rasterizeSpheres in float32: int((dim-1)(c -+ r)) clamped bounds, float(i)/dim positions,
    length2 summed x, y, z, strictly inside."""
    nx, ny, nz = res
    m = np.zeros((nz, ny, nx), np.uint8)
    f = np.float32
    for x0, y0, z0, r in spheres:
        c = [f(x0), f(y0), f(z0)]
        r = f(r)
        r2 = f(r * r)
        lo = [min(max(int(f(d - 1) * f(c[a] - r)), 0), d - 1) for a, d in enumerate(res)]
        hi = [min(max(int(f(d - 1) * f(c[a] + r)), 0), d - 1) for a, d in enumerate(res)]
        for z in range(lo[2], hi[2] + 1):
            for y in range(lo[1], hi[1] + 1):
                for x in range(lo[0], hi[0] + 1):
                    dx = f(f(x) / f(nx)) - c[0]
                    dy = f(f(y) / f(ny)) - c[1]
                    dz = f(f(z) / f(nz)) - c[2]
                    d2 = f(f(f(dx * dx) + f(dy * dy)) + f(dz * dz))
                    if d2 < r2:
                        m[z, y, x] = 255
    return m


def test_rasterize_sph(tmp_path):
    spheres = [(0.5, 0.5, 0.5, 0.25), (0.1, 0.9, 0.3, 0.15), (0.95, 0.05, 0.6, 0.2), (1.5, 1.5, 1.5, 0.1)]
    p = tmp_path / "s.sph"
    p.write_text(f"{len(spheres)}\n" + "\n".join(" ".join(repr(v) for v in s) for s in spheres) + "\n")
    res = (24, 20, 28)
    got = F.rasterizeSpheres(str(p), res)
    assert np.array_equal(got, _raster_f32(spheres, res))
    # geometry: the centred sphere covers about 4/3 pi r^3 of the unit cube
    one = tmp_path / "one.sph"
    one.write_text("1\n0.5 0.5 0.5 0.25\n")
    m = F.rasterizeSpheres(str(one), 64)
    assert abs((m > 0).mean() - 4 / 3 * np.pi * 0.25 ** 3) < 0.01
    empty = tmp_path / "far.sph"
    empty.write_text("1\n3.0 3.0 3.0 0.1\n")
    assert not F.rasterizeSpheres(str(empty), 16).any()


def test_subvolume():
    rng = np.random.default_rng(3)
    v = rng.integers(0, 256, size=(7, 9, 11), dtype=np.uint8)
    assert np.array_equal(F.getSubvolume(v, (2, 1, 3), (5, 6, 4)), v[3:7, 1:7, 2:7])
    with pytest.raises(FastError):
        F.getSubvolume(v, (8, 0, 0), (5, 1, 1))
