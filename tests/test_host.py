"""Host-side boundary of the PIV path, no GPU needed: the C ABI library loads
and exports what include/piv_xcorr2d.h declares, parameter validation mirrors
the reference's init checks, and the file formats at the boundary
(Parameter.dat, pair index, output names, TIFF, Tecplot) match the
reference's fixtures."""
import json
import os
import re
import struct

import numpy as np
import pytest

from conftest import GOLDEN, REPO

HEADER = os.path.join(REPO, "include", "piv_xcorr2d.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(pivx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(pivx):
    L = pivx.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"
    # and the Python binding covers every declared function
    assert sorted(set(names) - set(pivx.exported_symbols())) == []


@pytest.mark.parametrize("field,value,msg", [
    ("ws_x1", 2, "The interrogation window size (1st grid) is too small!"),
    ("ws_y1", 1028, "The interrogation window size (1st grid) is too big!"),
    ("ws_y2", 0, "The interrogation window size (2nd grid) is too small!"),
    ("ws_x2", 2048, "The interrogation window size (2nd grid) is too big!"),
    ("ws_x1", 30, "The window x-dimension is not divisible by 4!"),
    ("ws_y2", 34, "The window y-dimension is not divisible by 4!"),
    ("nsec_cc", 0, "Nsec_cc must be >= 1"),
    ("spacing_x2", 0, "vector spacing must be >= 1"),
    ("npeak_save", 0, "Npeak_save must be >= 1"),
])
def test_init_rejects_what_the_reference_rejects(pivx, field, value, msg):
    """cu:896-920 checks (same messages), plus the divide-by-zero guards."""
    p = pivx.init_params(**{field: value})
    with pytest.raises(pivx.PivError) as e:
        pivx.Xcorr2dMulti(p)
    assert e.value.code == 1
    assert msg in str(e.value)


def test_init_accepts_boundary_window_sizes_before_the_device_check(pivx):
    """ws = 4 and 1024 pass validation; any failure is then about the device
    (or none on a GPU box)."""
    p = pivx.init_params(nx=4096, ny=4096, ws_x1=4, ws_y1=1024, ws_x2=1024, ws_y2=4, spacing_x1=1024,
                         spacing_y1=1024, spacing_x2=1024, spacing_y2=1024)
    try:
        x = pivx.Xcorr2dMulti(p)
        x.freemem()
    except pivx.PivError as e:
        assert e.code == 2, str(e)


# ---------------------------------------------------------------- Parameter.dat
GEOMETRY_CASES = [
    # (nx, ny, ws_x, ws_y, sp_x, sp_y, pct_x, pct_y, nsec)
    (512, 512, 32, 32, 16, 16, 0.5, 0.5, 1),        # config 1
    (6600, 4400, 64, 64, 64, 64, 0.4, 0.4, 1),      # config 2, grid 1
    (6600, 4400, 64, 64, 32, 32, 0.4, 0.4, 1),      # sample Parameter.dat grid 1
    (6600, 4400, 32, 32, 16, 16, 0.4, 0.4, 1),      # config 2, grid 2
    (600, 400, 32, 32, 16, 16, 0.4, 0.4, 1),        # odd vector-column count (37)
    (588, 400, 24, 24, 12, 12, 0.4, 0.4, 1),
    (600, 400, 16, 32, 8, 16, 0.4, 0.4, 1),
    (584, 300, 16, 16, 8, 8, 0.3, 0.45, 1),
    (2048, 2048, 32, 32, 8, 8, 0.4, 0.4, 1),        # 75% overlap
    (4096, 2048, 64, 32, 32, 16, 0.4, 0.4, 1),      # non-square windows
    (1000, 1000, 32, 32, 16, 16, 0.4, 0.4, 3),      # sections
    (301, 177, 20, 12, 50, 9, 0.25, 0.1, 2),        # spacing beyond the window
    (97, 61, 4, 8, 3, 5, 0.4, 0.4, 7),
]


@pytest.mark.parametrize("case", GEOMETRY_CASES)
def test_engine_geometry_equals_oracle(pivx, orc, case):
    """The geometry the engine derives at init (pads, aug sizes, borders,
    slot cap, sections) equals the oracle's, which the reference result
    files pin at 6600x4400 (test_oracle.py)."""
    nx, ny, wx, wy, sx, sy, px, py, nsec = case
    ip = pivx.init_params(nx=nx, ny=ny, nsec_cc=nsec, ws_x1=wx, ws_y1=wy, spacing_x1=sx, spacing_y1=sy,
                          max_dispx_percent1=px, max_dispy_percent1=py, ws_x2=wx, ws_y2=wy, spacing_x2=sx,
                          spacing_y2=sy, max_dispx_percent2=px, max_dispy_percent2=py)
    want = orc.geometry(nx, ny, nsec, wx, wy, sx, sy, px, py)
    for grid in (1, 2):
        g = pivx.compute_geometry(ip, grid)
        got = {k: getattr(g, k) for k in want}
        assert got == want, (case, grid)
        assert (g.ws_x, g.ws_y, g.spacing_x, g.spacing_y, g.nt_vec) == (wx, wy, sx, sy, g.nx_vec * g.ny_vec)


def test_compute_geometry_applies_init_checks(pivx):
    with pytest.raises(pivx.PivError, match="not divisible by 4"):
        pivx.compute_geometry(pivx.init_params(ws_x1=30), 1)
    with pytest.raises(pivx.PivError):
        pivx.compute_geometry(pivx.init_params(), 3)


def test_parameter_file_sample_parses(pivx):
    """The reference's 50-line CRLF Parameter.dat (no trailing newline)."""
    pf = pivx.parse_parameter_file(os.path.join(GOLDEN, "Parameter.dat"))
    assert pf.prefix_f1.decode() == r"E:\JHUME\c_coding\xcorr2d_V1.0\data\data_organized\1_"
    assert pf.suffix_f2.decode() == ".tif"
    assert pf.folder_result.decode().endswith("result\\")
    assert (pf.device_id, pf.ifmhe, pf.mhe_ws) == (0, 1, 40)
    assert abs(pf.mhe_thresh - 0.91) < 1e-6
    i = pf.init
    assert (i.nx, i.ny, i.ws_x1, i.ws_y1, i.spacing_x1, i.spacing_y1, i.npass1) == (6600, 4400, 64, 64, 32, 32, 2)
    assert (i.ws_x2, i.ws_y2, i.spacing_x2, i.spacing_y2, i.npass2, i.nsec_cc) == (32, 32, 16, 16, 2, 1)
    assert (i.wx1_median, i.wy1_median, pf.npass1_median, pf.min_nneigh_p1) == (5, 5, 2, 5)
    assert (pf.residual_threshold_p1, pf.insert_threshold_p1) == (2.0, 3.0)
    assert (i.wx1_smooth, i.wy1_smooth, pf.npass1_smooth) == (3, 3, 1)
    assert (i.wx2_median, i.wy2_median, pf.npass2_median, pf.min_nneigh_p2) == (5, 5, 2, 5)
    assert (i.wx2_smooth, i.wy2_smooth, pf.npass2_smooth) == (3, 3, 1)
    assert abs(i.max_dispx_percent1 - 0.4) < 1e-7 and abs(i.max_dispy_percent2 - 0.4) < 1e-7
    assert i.npeak_save == 4
    assert (i.dispx_lim1, i.dispx_lim2, i.dispy_lim1, i.dispy_lim2) == (-100, 100, -100, 100)


def test_parameter_file_41_lines_is_missing_parameters(pivx):
    with pytest.raises(pivx.PivError) as e:
        pivx.parse_parameter_file(os.path.join(GOLDEN, "Parameter41.dat"))
    assert "Parameter missing" in str(e.value)


def test_parameter_file_unopenable(pivx, tmp_path):
    with pytest.raises(pivx.PivError) as e:
        pivx.parse_parameter_file(str(tmp_path / "nope.dat"))
    assert "Can't open parameter file" in str(e.value)


def test_numeric_lines_parse_like_wtoi_wtof(pivx, tmp_path):
    lines = open(os.path.join(GOLDEN, "Parameter.dat"), "rb").read().split(b"\r\n")
    lines[10] = b"  6600abc"   # Nx: leading number
    lines[12] = b"garbage"     # ws_x1 -> 0
    lines[41] = b"-0.25e0x"    # max_dispx_percent1
    p = tmp_path / "p.dat"
    p.write_bytes(b"\n".join(lines))  # LF line ends are accepted too
    pf = pivx.parse_parameter_file(str(p))
    assert pf.init.nx == 6600 and pf.init.ws_x1 == 0 and pf.init.max_dispx_percent1 == -0.25


def test_pair_index(pivx, tmp_path):
    assert pivx.read_pair_index(os.path.join(GOLDEN, "pair_index.txt")) == [1]
    p = tmp_path / "idx.txt"
    p.write_text("3\r\n17\n 42\n")
    assert pivx.read_pair_index(str(p)) == [3, 17, 42]
    e = tmp_path / "empty.txt"
    e.write_text("")
    assert pivx.read_pair_index(str(e)) == []
    with pytest.raises(pivx.PivError):
        pivx.read_pair_index(str(tmp_path / "missing.txt"))


def test_output_names_match_reference_fixtures(pivx):
    """main.cpp:327-376 naming, checked against the committed result files."""
    pf = pivx.parse_parameter_file(os.path.join(GOLDEN, "Parameter.dat"))
    pf.folder_result = b""
    fixtures = json.load(open(os.path.join(GOLDEN, "result_fixtures.json")))
    names = [pivx.output_name(pf, 1, k) for k in range(5)]
    assert names[0] == "0001_vector_GPU_Grid1_4-peak-candidate_3DCC-64-64-2pass_4peaks_raw.dat"
    assert names[4] == "0001_vector_GPU_Grid2_4-peak-candidate_3DCC-32-32-2pass_median-5-5-3.0-2pass.dat"
    for n in names:
        assert n in fixtures, n
    pf.folder_result = b"out/"
    assert pivx.output_name(pf, 12, 2).startswith("out/0012_vector_GPU_Grid2_")


# ---------------------------------------------------------------- Tecplot writer
def test_tecplot_format_matches_fixture(pivx, tmp_path):
    """Header lines and row layout of cu:1875-1886 (fixed, 2 decimals, tabs),
    CRLF as in the committed files, on the 6600x4400 grid-2 geometry."""
    fixtures = json.load(open(os.path.join(GOLDEN, "result_fixtures.json")))
    raw = fixtures["0001_vector_GPU_Grid2_4-peak-candidate_3DCC-32-32-2pass_4peaks_raw.dat"]
    filt = fixtures["0001_vector_GPU_Grid2_4-peak-candidate_3DCC-32-32-2pass_median-5-5-3.0-2pass_smooth-3-3-1pass.dat"]
    nxv, nyv = 412, 275
    k, j = np.meshgrid(np.arange(nxv), np.arange(nyv))
    gx = (8 + 16 * k).ravel().astype(np.float32)
    gy = (8 + 16 * j).ravel().astype(np.float32)
    z = np.zeros((nxv * nyv, 4), np.float32)
    p = tmp_path / "raw.dat"
    pivx.write_tecplot(str(p), nxv, nyv, gx, gy, z, z, raw=True, crlf=True)
    lines = p.read_bytes().decode().split("\n")
    got = [l + "\n" for l in lines[:-1]]
    assert got[:3] == raw["header"]
    assert got[3:8] == raw["first_rows"]
    assert got[-1] == raw["last_row"]
    assert len(got) == raw["n_lines"]
    # filtered format, value formatting of the fixture's first rows
    vals = [[float(t) for t in r.split("\t")[2:]] for r in filt["first_rows"]]
    dx = np.zeros(nxv * nyv, np.float32)
    dy = np.zeros(nxv * nyv, np.float32)
    for i, (a, b) in enumerate(vals):
        dx[i], dy[i] = a, b
    q = tmp_path / "filt.dat"
    pivx.write_tecplot(str(q), nxv, nyv, gx, gy, dx, dy, raw=False, crlf=True)
    got = [l + "\n" for l in q.read_bytes().decode().split("\n")[:-1]]
    assert got[:3] == filt["header"]
    assert got[3:8] == filt["first_rows"]


def test_tecplot_rounding_and_missing_candidates(pivx, tmp_path):
    gx = np.array([16.0], np.float32)
    gy = np.array([16.0], np.float32)
    dx = np.array([[-0.004, 1.005]], np.float32)
    dy = np.array([[2.5, -3.999]], np.float32)
    p = tmp_path / "r.dat"
    pivx.write_tecplot(str(p), 1, 1, gx, gy, dx, dy, raw=True)
    row = p.read_text().split("\n")[3]
    f = row.split("\t")
    assert f[:2] == ["16.00", "16.00"]
    assert f[2] == "-0.00"            # fixed/setprecision(2) keeps the sign
    assert f[3] == "2.50"
    assert f[5] == "-4.00"
    assert f[6:] == ["0.00"] * 4      # K = 2 < 4: missing candidates are written as zeros


# ---------------------------------------------------------------- TIFF
def test_tiff_roundtrip_and_info(pivx, tmp_path):
    img = (np.arange(37 * 23) % 251).astype(np.uint8).reshape(23, 37)
    p = tmp_path / "a.tif"
    pivx.tiff_write8(str(p), img)
    assert pivx.tiff_info(str(p)) == (37, 23, 8)
    np.testing.assert_array_equal(pivx.tiff_read8(str(p), 37, 23), img)


def _tiff_bytes(img, big=False, strips=1, short_offsets=False):
    """Hand-built baseline TIFF (independent of the library's writer)."""
    h, w = img.shape
    e = ">" if big else "<"
    rows = -(-h // strips)
    data = [img[r:r + rows].tobytes() for r in range(0, h, rows)]
    n = len(data)
    entries = 8
    ifd = 8
    tables = ifd + 2 + 12 * entries + 4
    typ = 3 if short_offsets else 4
    sz = 2 if short_offsets else 4
    off_tab = tables
    cnt_tab = off_tab + n * sz
    pix = cnt_tab + n * sz
    offs, o = [], pix
    for d in data:
        offs.append(o)
        o += len(d)
    out = bytearray(b"MM" if big else b"II")
    out += struct.pack(e + "HI", 42, ifd)
    out += struct.pack(e + "H", entries)

    def ent(tag, t, count, val):
        if t == 3 and count == 1:
            return struct.pack(e + "HHIHH", tag, t, count, val, 0)
        return struct.pack(e + "HHII", tag, t, count, val)

    out += ent(256, 4, 1, w) + ent(257, 4, 1, h) + ent(258, 3, 1, 8) + ent(259, 3, 1, 1) + ent(262, 3, 1, 1)
    if n == 1:
        out += ent(273, typ, 1, offs[0]) + ent(278, 4, 1, rows) + ent(279, typ, 1, len(data[0]))
    elif n * sz <= 4:  # two SHORT values fit in the entry itself (TIFF 6.0, IFD entry rule)
        out += struct.pack(e + "HHI", 273, typ, n) + struct.pack(e + "HH", *offs)
        out += ent(278, 4, 1, rows)
        out += struct.pack(e + "HHI", 279, typ, n) + struct.pack(e + "HH", *[len(d) for d in data])
    else:
        out += ent(273, typ, n, off_tab) + ent(278, 4, 1, rows) + ent(279, typ, n, cnt_tab)
    out += struct.pack(e + "I", 0)
    fmt = e + ("H" if short_offsets else "I")
    for v in offs:
        out += struct.pack(fmt, v)
    for d in data:
        out += struct.pack(fmt, len(d))
    for d in data:
        out += d
    return bytes(out)


@pytest.mark.parametrize("big,strips,short", [(False, 1, False), (False, 5, False), (True, 3, False),
                                              (False, 4, True), (True, 2, True)])
def test_tiff_reader_strips_and_byte_order(pivx, tmp_path, big, strips, short):
    img = np.random.default_rng(1).integers(0, 256, (29, 41), dtype=np.uint8)
    p = tmp_path / "x.tif"
    p.write_bytes(_tiff_bytes(img, big, strips, short))
    assert pivx.tiff_info(str(p)) == (41, 29, 8)
    np.testing.assert_array_equal(pivx.tiff_read8(str(p), 41, 29), img)


def test_tiff_rejects_corrupt_directories_and_short_strips(pivx, tmp_path):
    """Counts and offsets taken from the file are bounded by the file size:
    a strip table claiming 2^32-1 entries, an IFD beyond the end of the file
    and an entry count running past it are format errors (not a huge
    allocation or a long seek loop), and strips holding fewer bytes than the
    frame are an error rather than a partly stale frame."""
    img = np.random.default_rng(2).integers(0, 256, (29, 41), dtype=np.uint8)
    good = bytearray(_tiff_bytes(img, False, 5, False))
    n_ent = struct.unpack("<H", good[8:10])[0]
    cases = {}
    for k in range(n_ent):  # strip-offset entry: count -> 0xffffffff
        o = 10 + 12 * k
        if struct.unpack("<H", good[o:o + 2])[0] == 273:
            b = bytearray(good)
            b[o + 4:o + 8] = struct.pack("<I", 0xFFFFFFFF)
            cases["huge strip count"] = b
    b = bytearray(good)
    b[4:8] = struct.pack("<I", len(good) + 100)
    cases["IFD beyond the file"] = b
    b = bytearray(good)
    b[8:10] = struct.pack("<H", 0xFFFF)
    cases["entry count beyond the file"] = b
    assert len(cases) == 3
    for name, data in cases.items():
        p = tmp_path / "bad.tif"
        p.write_bytes(bytes(data))
        with pytest.raises(pivx.PivError) as ei:
            pivx.tiff_read8(str(p), 41, 29)
        assert ei.value.code == 5, name
    p = tmp_path / "short.tif"
    p.write_bytes(_tiff_bytes(img[:20], False, 2, False))  # 20 rows for a 29-row frame
    with pytest.raises(pivx.PivError, match="fewer bytes"):
        pivx.tiff_read8(str(p), 41, 29)


def test_tiff_rejects_non_tiff(pivx, tmp_path):
    p = tmp_path / "bad.tif"
    p.write_bytes(b"II\x2b\x00\x08\x00\x00\x00")
    with pytest.raises(pivx.PivError):
        pivx.tiff_info(str(p))


# ---------------------------------------------------------------- synthetic inputs
def test_synthetic_pairs_are_seeded(pivx):
    sp = pivx.synth_params(96, 80, seed=5, flow=0, a0=1.5, a1=-0.5)
    a1, b1 = pivx.synth_pair(sp)
    a2, b2 = pivx.synth_pair(sp)
    np.testing.assert_array_equal(a1, a2)
    np.testing.assert_array_equal(b1, b2)
    sp.seed = 6
    a3, _ = pivx.synth_pair(sp)
    assert (a3 != a1).mean() > 0.3
    assert 10 < a1.mean() < 80 and a1.max() == 255


def test_synthetic_flows(pivx):
    sp = pivx.synth_params(6600, 4400, flow=1, a0=300.0, a1=8.0, a2=0.0)
    xc, yc = (6600 - 1) / 2, (4400 - 1) / 2
    r = np.linspace(1, 2000, 4000)
    u, v = pivx.synth_flow(sp, xc + r, np.full_like(r, yc))
    assert abs(np.abs(v).max() - 8.0) < 1e-3                # peak swirl
    assert abs(r[np.argmax(np.abs(v))] / 300.0 - 1.1209) < 0.01
    assert np.abs(u).max() < 1e-6
    sp.a2 = 0.002
    u, v = pivx.synth_flow(sp, [xc, xc], [yc - 1000, yc + 1000])
    # Lamb-Oseen swirl at r = 1000 (counter-clockwise) plus 0.002*(y - yc) shear
    s_max = max((1 - np.exp(-s * s)) / s for s in np.arange(1, 40000) * 1e-4)
    ut = 300.0 * 8.0 / s_max / 1000.0 * (1 - np.exp(-(1000.0 / 300.0) ** 2))
    assert np.allclose(u, [ut - 2.0, -ut + 2.0], atol=1e-3)
    assert np.allclose(v, [0.0, 0.0], atol=1e-6)
    sp2 = pivx.synth_params(2048, 2048, seed=3, flow=2, a0=12.0)
    g = np.arange(0, 2048, 8, dtype=np.float32)
    X, Y = np.meshgrid(g, g)
    u, v = pivx.synth_flow(sp2, X, Y)
    assert abs(np.sqrt(u ** 2 + v ** 2).max() - 12.0) < 1e-3


def test_tecplot_numbers_match_printf(pivx, tmp_path):
    """The writer's number formatting equals C's "%.2f" (iostream fixed +
    setprecision(2), cu:1879-1880) byte for byte: ties to even on the exact
    binary value, "-0.00" for small negatives, wide values."""
    rng = np.random.default_rng(5)
    nx, ny = 97, 41
    nt = nx * ny
    v = rng.normal(0, 30, nt * 8).astype(np.float32)
    edge = np.array([0.125, 0.375, -0.125, 0.005, -0.005, -0.004, 0.0, -0.0, 2.675, 1.005, 12345.675, -12345.675,
                     0.045, 0.015, 1e6 + 0.125, -1e-30, 99.995, 0.5, -0.5, 3e9, 1e20, -2.5e-3], np.float32)
    v[:edge.size] = edge
    dx, dy = v[:nt * 4].reshape(nt, 4), v[nt * 4:].reshape(nt, 4)
    gx = (rng.random(nt) * 6600).astype(np.float32)
    gy = (rng.random(nt) * 4400).astype(np.float32)
    p = str(tmp_path / "t.dat")
    pivx.write_tecplot(p, nx, ny, gx, gy, dx, dy, raw=True, crlf=False)
    rows = open(p, "rb").read().split(b"\n")[3:-1]
    assert len(rows) == nt
    for i in range(nt):
        exp = "\t".join("%.2f" % float(x) for x in [gx[i], gy[i]] + [c for k in range(4) for c in (dx[i, k], dy[i, k])])
        assert rows[i] == exp.encode(), (i, rows[i], exp)
