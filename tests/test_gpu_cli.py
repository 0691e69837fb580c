"""The drop-in entry point: piv_xcorr2d on a 50-line Parameter.dat with 8-bit
TIFF frames writes the reference's five Tecplot files per pair (main.cpp:
327-376), with values identical to the engine's own fields."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _write_params(path, lines):
    with open(path, "wb") as f:
        f.write(b"\r\n".join(l.encode() for l in lines))


def _param_lines(tmp, nx, ny):
    base = open(os.path.join(GOLDEN, "Parameter.dat"), "rb").read().decode().split("\r\n")
    lines = list(base)
    lines[0] = str(tmp / "1_")
    lines[1] = ".tif"
    lines[2] = str(tmp / "2_")
    lines[3] = ".tif"
    lines[4] = str(tmp / "pair index.txt")
    lines[5] = str(tmp / "result") + "/"
    lines[7] = "0"          # MHE off
    lines[10] = str(nx)
    lines[11] = str(ny)
    return lines


def _read_tecplot(path):
    rows = open(path).read().split("\n")
    head = rows[:3]
    data = np.array([[float(t) for t in r.split("\t")] for r in rows[3:] if r])
    return head, data


def test_cli_runs_parameter_file(pivx, tmp_path):
    nx, ny = 640, 448
    pairs = [1, 7]
    for k in pairs:
        sp = pivx.synth_params(nx, ny, seed=100 + k, flow=1, a0=90.0, a1=5.0, a2=0.002)
        a, b = pivx.synth_pair(sp)
        pivx.tiff_write8(str(tmp_path / f"1_{k:04d}.tif"), a)
        pivx.tiff_write8(str(tmp_path / f"2_{k:04d}.tif"), b)
    (tmp_path / "pair index.txt").write_text("\n".join(str(k) for k in pairs))
    lines = _param_lines(tmp_path, nx, ny)
    par = tmp_path / "Parameter.dat"
    _write_params(par, lines)
    r = subprocess.run([pivx.CLI_PATH, str(par), "--crlf"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr
    assert "NO ERROR!" in r.stdout
    pf = pivx.parse_parameter_file(str(par))
    for k in pairs:
        names = [pivx.output_name(pf, k, w) for w in range(5)]
        for n in names:
            assert os.path.exists(n), n
            assert open(n, "rb").read().count(b"\r\n") > 3
        # same values as the library run on the same frames
        a = pivx.tiff_read8(str(tmp_path / f"1_{k:04d}.tif"), nx, ny)
        b = pivx.tiff_read8(str(tmp_path / f"2_{k:04d}.tif"), nx, ny)
        fp1 = pivx.FilterParams(pf.npass1_median, pf.min_nneigh_p1, pf.residual_threshold_p1,
                                pf.insert_threshold_p1, pf.npass1_smooth)
        fp2 = pivx.FilterParams(pf.npass2_median, pf.min_nneigh_p2, pf.residual_threshold_p2,
                                pf.insert_threshold_p2, pf.npass2_smooth)
        with pivx.Xcorr2dMulti(pf.init) as x:
            x.prep_img_data(a, b)
            x.run_pair(fp1, fp2)
            f2 = x.filtered(2)
            mo = x.filtered(3)
            r1 = x.raw(1)
            nx2, ny2 = x.dims(2)
        head, d = _read_tecplot(names[3])
        assert head[2].strip() == f'ZONE T="Frame 0", I={nx2}, J={ny2}'
        assert np.abs(d[:, 2] - f2[0]).max() <= 0.005 + 1e-6 and np.abs(d[:, 3] - f2[1]).max() <= 0.005 + 1e-6
        _, d = _read_tecplot(names[4])
        assert np.abs(d[:, 2] - mo[0]).max() <= 0.005 + 1e-6
        _, d = _read_tecplot(names[0])
        assert d.shape[1] == 10 and np.abs(d[:, 2] - r1[0][:, 0]).max() <= 0.005 + 1e-6


def test_cli_errors(pivx, tmp_path):
    r = subprocess.run([pivx.CLI_PATH, str(tmp_path / "none.dat")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Can't open parameter file!" in r.stderr
    r = subprocess.run([pivx.CLI_PATH, os.path.join(GOLDEN, "Parameter41.dat")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "Parameter missing" in r.stderr
    lines = _param_lines(tmp_path, 256, 256)
    par = tmp_path / "p.dat"
    (tmp_path / "pair index.txt").write_text("")
    _write_params(par, lines)
    r = subprocess.run([pivx.CLI_PATH, str(par)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "No image pair index detected!" in r.stderr
    (tmp_path / "pair index.txt").write_text("3")
    r = subprocess.run([pivx.CLI_PATH, str(par)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "missing" in r.stderr
    # MHE requested with a window the reference cannot tile (line 9 = 0)
    img = np.zeros((256, 256), np.uint8)
    pivx.tiff_write8(str(tmp_path / "1_0003.tif"), img)
    pivx.tiff_write8(str(tmp_path / "2_0003.tif"), img)
    lines[7] = "1"
    lines[8] = "0"
    _write_params(par, lines)
    r = subprocess.run([pivx.CLI_PATH, str(par)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "mhe_gpu" in r.stderr


def test_cli_mhe_matches_enhanced_frames(pivx, orc, tmp_path):
    """Parameter.dat line 8 = 1: the CLI enhances both frames (window 40,
    threshold 0.91, lines 9-10) before the pair runs; the vectors equal a
    library run on frames enhanced by the oracle's MHE."""
    nx, ny = 640, 448
    sp = pivx.synth_params(nx, ny, seed=77, flow=0, a0=3.3, a1=-1.7)
    a, b = pivx.synth_pair(sp)
    pivx.tiff_write8(str(tmp_path / "1_0005.tif"), a)
    pivx.tiff_write8(str(tmp_path / "2_0005.tif"), b)
    (tmp_path / "pair index.txt").write_text("5")
    lines = _param_lines(tmp_path, nx, ny)
    lines[7] = "1"
    par = tmp_path / "Parameter.dat"
    _write_params(par, lines)
    r = subprocess.run([pivx.CLI_PATH, str(par)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    pf = pivx.parse_parameter_file(str(par))
    assert (pf.ifmhe, pf.mhe_ws) == (1, 40)
    ea, eb = orc.mhe(a, 40, pf.mhe_thresh), orc.mhe(b, 40, pf.mhe_thresh)
    fp1 = pivx.FilterParams(pf.npass1_median, pf.min_nneigh_p1, pf.residual_threshold_p1, pf.insert_threshold_p1,
                            pf.npass1_smooth)
    fp2 = pivx.FilterParams(pf.npass2_median, pf.min_nneigh_p2, pf.residual_threshold_p2, pf.insert_threshold_p2,
                            pf.npass2_smooth)
    with pivx.Xcorr2dMulti(pf.init) as x:
        x.prep_img_data(ea, eb)
        x.run_pair(fp1, fp2)
        f2 = x.filtered(2)
    head, d = _read_tecplot(pivx.output_name(pf, 5, 3))
    assert np.abs(d[:, 2] - f2[0]).max() <= 0.005 + 1e-6 and np.abs(d[:, 3] - f2[1]).max() <= 0.005 + 1e-6


def test_cli_pipelined_equals_serial(pivx, tmp_path):
    """The pipelined driver (next pair read and previous pair written on host
    threads while the GPU runs) writes the same bytes as the serial order."""
    nx, ny = 512, 384
    pairs = [2, 3, 4]
    for k in pairs:
        sp = pivx.synth_params(nx, ny, seed=300 + k, flow=1, a0=80.0, a1=4.0, a2=0.002)
        a, b = pivx.synth_pair(sp)
        pivx.tiff_write8(str(tmp_path / f"1_{k:04d}.tif"), a)
        pivx.tiff_write8(str(tmp_path / f"2_{k:04d}.tif"), b)
    (tmp_path / "pair index.txt").write_text("\n".join(str(k) for k in pairs))
    outs = {}
    for mode in ("pipelined", "serial"):
        lines = _param_lines(tmp_path, nx, ny)
        lines[5] = str(tmp_path / f"res_{mode}") + "/"
        par = tmp_path / f"P_{mode}.dat"
        _write_params(par, lines)
        r = subprocess.run([pivx.CLI_PATH, str(par)] + (["--serial"] if mode == "serial" else []),
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert "Total: 3 pairs" in r.stdout
        pf = pivx.parse_parameter_file(str(par))
        outs[mode] = [open(pivx.output_name(pf, k, w), "rb").read() for k in pairs for w in range(5)]
    assert outs["pipelined"] == outs["serial"]
