"""The reference's file formats (fd2_io.cpp), checked against the example
files the reference ships (example/hybridMANTIS_input.in and
example/output_files/*.dat, copied under tests/golden/io): the input
parser, the output-file names, byte-identical PRF and PHS writers and the
console statistics block (1e5xrays_console.out:275-299).  The
EDE reader is checked on records written in the tally's format
(main/hybridMANTIS_tallyEnergyDepositionEvents.f:173-174)."""
import os

import numpy as np
import pytest

from hybridmantis_amd import fd2

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "io")
STEM = "150.0_5.1_0.1_1e-4_0.2_280.0_0.25_0.055_9_1"
PRF_FILE = os.path.join(HERE, f"myimage_{STEM}.dat")
PHS_FILE = os.path.join(HERE, f"detected_{STEM}.dat")
INPUT = os.path.join(HERE, "hybridMANTIS_input.in")
N_HIST = 100000   # example/output_files/1e5xrays_console.out


def test_read_example_input():
    p, tokens = fd2.read_input(INPUT)
    assert p["num_hist"] == 100000 and p["min_ph"] == 0 and p["max_ph"] == 1400
    assert p["num_bins"] == 140
    assert p["xdet"] == 909.0 and p["ydet"] == 909.0 and p["H"] == 150.0 and p["R"] == 5.1
    assert (p["n_C"], p["n_IC"], p["top_absfrac"], p["bulk"], p["beta"]) == (1.8, 1.0, 0.1, 1e-4, 0.2)
    assert (p["d_min"], p["d_max"]) == (1.0, 280.0)
    assert (p["lbx"], p["lby"], p["ubx"], p["uby"]) == (0.0, 0.0, 909.0, 909.0)
    assert p["yield_per_ev"] == 0.055 and p["pixel"] == 9 and p["sensor_refl"] == 0.25
    assert p["rungpu"] == 1 and p["machine"] == 1
    assert tokens == ["150.0", "5.1", "0.1", "1e-4", "0.2", "280.0", "0.25", "0.055", "9", "1"]


def test_output_names_match_example_files():
    _, tokens = fd2.read_input(INPUT)
    prf, phs = fd2.output_names(tokens)
    assert prf == os.path.basename(PRF_FILE)
    assert phs == os.path.basename(PHS_FILE)


def test_read_input_rejects_short_or_bad_files(tmp_path):
    lines = open(INPUT).read().splitlines()
    short = tmp_path / "short.in"
    short.write_text("\n".join(lines[:20]) + "\n")
    with pytest.raises(fd2.Fd2Error):
        fd2.read_input(str(short))
    bad = tmp_path / "bad.in"
    body = [l for l in lines if l.strip() and not l.lstrip().startswith("#")]
    body[4] = "abc   # x-dimension"
    bad.write_text("\n".join(body) + "\n")
    with pytest.raises(fd2.Fd2Error):
        fd2.read_input(str(bad))
    with pytest.raises(fd2.Fd2Error):
        fd2.read_input(str(tmp_path / "missing.in"))


def read_prf_values(path):
    """Values of a PRF file in file order (y-pixel blocks of x pixels)."""
    blocks, cur = [], []
    for line in open(path):
        if line.strip():
            cur.append(float(line))
        else:
            blocks.append(cur)
            cur = []
    return np.array(blocks)  # [y][x]


def test_prf_writer_reproduces_example_file(tmp_path):
    """Counts whose count / 1e5 prints as the example's values give the
    example file byte for byte (10302 records: 101 blocks of 101 values and
    a blank record, pixelX = int(909 / 9) - 1 = 100)."""
    p, _ = fd2.read_input(INPUT)
    v = read_prf_values(PRF_FILE)
    assert v.shape == (101, 101)
    counts_yx = np.rint(v * N_HIST).astype(np.uint64)
    prf = np.ascontiguousarray(counts_yx.T)  # [x][y]
    out = tmp_path / "myimage.dat"
    fd2.write_prf(str(out), p, prf, N_HIST)
    assert out.read_bytes() == open(PRF_FILE, "rb").read()
    # STATUS='NEW': an existing file is not overwritten
    with pytest.raises(fd2.Fd2Error):
        fd2.write_prf(str(out), p, prf, N_HIST)


def test_prf_writer_orientation_and_edges(tmp_path):
    """Outer blocks run over y, values inside a block over x; pixels of the
    written range beyond the tallied image print as zero."""
    p = dict(lbx=0.0, lby=0.0, ubx=4.0, uby=4.0, pixel=1)
    prf = np.zeros((4, 4), np.uint64)
    prf[1, 3] = 2    # x = 1, y = 3
    out = tmp_path / "o.dat"
    fd2.write_prf(str(out), p, prf, 2.0)
    v = read_prf_values(str(out))
    assert v.shape == (4, 4) and v[3, 1] == 1.0 and v.sum() == 1.0
    out2 = tmp_path / "o2.dat"
    fd2.write_prf(str(out2), p, prf[:2, :2], 2.0)  # smaller tally than the file
    assert read_prf_values(str(out2)).shape == (4, 4)


def test_phs_writer_reproduces_example_file(tmp_path):
    phs = np.array([int(l) for l in open(PHS_FILE)], np.int32)
    assert len(phs) == 140 and phs.sum() > 0
    out = tmp_path / "detected.dat"
    fd2.write_phs(str(out), phs)
    assert out.read_bytes() == open(PHS_FILE, "rb").read()
    with pytest.raises(fd2.Fd2Error):
        fd2.write_phs(str(out), phs)


def ede_line(x, y, z, de, h):
    """One record as the tally writes it: (3(1x,1pe16.9),1x,1pe12.4,i8)."""
    return ("".join(" " + f"{v:.9E}".rjust(16) for v in (x, y, z)) + " " +
            f"{de:.4E}".rjust(12) + f"{h:8d}")


def test_ede_reader(tmp_path):
    rows = [(0.04545, 0.04545, -0.0074991, 25000.0, 1),
            (0.0455, 0.0454, 0.00123, 812.5, 1),
            (0.1, 0.2, -0.0001, 3.5e3, 12345678)]   # 8-digit history touches de
    lines = ["# x y z de hist", ""] + [ede_line(*r) for r in rows]
    assert all(len(l) == 72 for l in lines[2:])
    assert "E+0312345678" in lines[-1]
    path = tmp_path / "ede.dat"
    path.write_text("\n".join(lines) + "\n")
    ev = fd2.read_ede(str(path), 10)
    assert len(ev) == 3
    for e, (x, y, z, de, h) in zip(ev, rows):
        # the drivers' conversion: double(cm) * 1e4 rounded to float
        assert e["x"] == np.float32(float(f"{x:.9E}") * 10000.0)
        assert e["y"] == np.float32(float(f"{y:.9E}") * 10000.0)
        assert e["z"] == np.float32(float(f"{z:.9E}") * 10000.0)
        assert e["E"] == np.float32(float(f"{de:.4E}"))
        assert e["hist"] == h and e["N"] == 0
    # free-format records are read too
    path2 = tmp_path / "ede2.dat"
    path2.write_text("0.01 0.02 -0.003 1500.0 7\n")
    ev2 = fd2.read_ede(str(path2), 1)
    assert ev2["hist"][0] == 7 and ev2["x"][0] == np.float32(0.01 * 10000.0)
    with pytest.raises(fd2.Fd2Error):   # more records than room
        fd2.read_ede(str(path), 2)
    bad = tmp_path / "bad.dat"
    bad.write_text("0.01 0.02 x 1500.0 7\n")
    with pytest.raises(fd2.Fd2Error):
        fd2.read_ede(str(bad), 4)


def test_console_stats_block_reproduces_example(tmp_path):
    """fd2_write_stats against the reference example run's console block
    (example/output_files/1e5xrays_console.out:275-299, copied as
    tests/golden/io/console_stats_1e5xrays.txt): the counters of that run,
    its 0.21 s of GPU time and 15.39 s of optical wall time."""
    from hybridmantis_amd import fd2

    want = open(os.path.join(HERE, "console_stats_1e5xrays.txt"), "rb").read()
    stats = [95222286.0, 51718347.0, 8184844.0, 3175019.0, 20111907.0, 29389.0, 0.0, 210.0]
    out = tmp_path / "stats.txt"
    fd2.write_stats(str(out), stats, rungpu=1, penoptical_s=15.39)
    assert out.read_bytes() == want
    fd2.write_stats(str(out), stats, rungpu=0, penoptical_s=15.39)
    got = out.read_bytes().decode()
    assert "   !! Simulation in the CPU !!\n" in got
    # F12.2 overflow prints asterisks, as gfortran does
    fd2.write_stats(str(out), [1e13] + stats[1:], rungpu=1, penoptical_s=0.0)
    assert out.read_text().splitlines()[8] == " ************"
