"""The rocprofv3 summary of pllcart3d_amd/profiling.py on synthetic trace and
counter databases (no GPU): only the profiled handle's dispatches count
(the stream of the last dispatch, from its first copy kernel on), the
counter passes keep each kernel's last dispatches, and the copy kernel
calibrates the byte counts."""
import argparse
import os
import sqlite3
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "pllcart3d_amd"))
import profiling  # noqa: E402


def _trace_db(path, rows):
    """rows: (name, dispatch_id, stream_id, start_ns, end_ns)"""
    con = sqlite3.connect(path)
    con.execute("create table kernels (name text, dispatch_id int, stream_id int, start int, end int)")
    con.executemany("insert into kernels values (?, ?, ?, ?, ?)", rows)
    con.commit()
    con.close()


def _counter_db(path, cname, rows):
    """rows: (kernel_name, dispatch_id, value)"""
    con = sqlite3.connect(path)
    con.execute("create table counters_collection (kernel_name text, dispatch_id int, counter_name text, value real)")
    con.executemany("insert into counters_collection values (?, ?, ?, ?)", [(n, d, cname, v) for n, d, v in rows])
    con.commit()
    con.close()


def test_summary_counts_the_profiled_handle_only(tmp_path, capsys):
    cp = "void pllc::k_copy<double>(double const*, double*, long)"
    sw = "void pllc::k_sweep_part<double, 0>(pllc::Dims)"
    tr = []
    # a tuning handle on stream 1 (its own copies and sweeps), then the
    # profiled handle on stream 2: a set-up sweep, two copies, two sweeps
    tr += [(cp, 1, 1, 0, 1000), (sw, 2, 1, 1000, 9000), (sw, 3, 1, 9000, 19000)]
    tr += [(sw, 4, 2, 20000, 20500), (cp, 5, 2, 21000, 22000), (cp, 6, 2, 22000, 23000)]
    tr += [(sw, 7, 2, 23000, 27000), (sw, 8, 2, 27000, 33000)]
    _trace_db(str(tmp_path / "t.db"), tr)
    rows = {r[0]: r for r in profiling._trace_rows(str(tmp_path / "t.db"))}
    # the sweeps after the profiled handle's first copy: 4 us and 6 us
    assert rows[sw][1] == 2 and abs(rows[sw][2] - 10.0) < 1e-9 and abs(rows[sw][3] - 5.0) < 1e-9
    assert rows[cp][1] == 2 and abs(rows[cp][3] - 1.0) < 1e-9

    # counter passes (KiB): one field of 1 Mi points x 8 B = 8192 KiB; the
    # copy reads half of it by FETCH_SIZE and all of it by WRITE_SIZE
    _counter_db(str(tmp_path / "f.db"), "FETCH_SIZE",
                [("k_copy<double>", 1, 4096.0), ("k_sweep_part<double, 0>", 2, 99.0),
                 ("k_sweep_part<double, 0>", 3, 99.0), ("k_sweep_part<double, 0>", 4, 99.0),
                 ("k_copy<double>", 5, 4096.0), ("k_copy<double>", 6, 4096.0),
                 ("k_sweep_part<double, 0>", 7, 6144.0), ("k_sweep_part<double, 0>", 8, 10240.0)])
    _counter_db(str(tmp_path / "w.db"), "WRITE_SIZE",
                [("k_copy<double>", 1, 8192.0), ("k_copy<double>", 5, 8192.0), ("k_copy<double>", 6, 8192.0),
                 ("k_sweep_part<double, 0>", 7, 2048.0), ("k_sweep_part<double, 0>", 8, 2048.0)])
    args = argparse.Namespace(trace=str(tmp_path / "t.db"), fetch=str(tmp_path / "f.db"),
                              write=str(tmp_path / "w.db"), points=1024 * 1024, bytes=8)
    profiling.summary(args)
    out = capsys.readouterr().out
    assert "factors are 2.000 (fetch) and 1.000 (write)" in out
    line = [l for l in out.splitlines() if l.startswith("| `k_sweep_part<double, 0>` | 2 |")]
    # the sweep's last two dispatches (not the tuning handle's 99s): fetch
    # (6144 + 10240)/2 x 2 = 16384 KiB, write 2048 KiB, in MB; 2.25 fields
    assert line, out
    cells = [c.strip() for c in line[-1].split("|")]
    assert abs(float(cells[3]) - 16384 * 1024 / 1e6) < 0.05
    assert abs(float(cells[4]) - 2048 * 1024 / 1e6) < 0.05
    assert abs(float(cells[5]) - 2.25) < 1e-6


def test_ensemble_summary_takes_every_stream_after_the_last_copy(tmp_path, capsys):
    cp = "void pllc::k_copy<float>(float const*, float*, long)"
    cv = "void pllc::k_corr_pg<float, true>(pllc::Dims)"
    con = sqlite3.connect(str(tmp_path / "e.db"))
    con.execute("create table kernels (name text, dispatch_id int, start int, end int, grid_x int, grid_y int, "
                "grid_z int)")
    rows = [(cv, 1, 0, 5000, 64, 64, 4224),                        # warm-up step
            (cp, 2, 6000, 7000, 512, 1, 1), (cp, 3, 7000, 8000, 512, 1, 1),   # the marker
            (cv, 4, 10000, 14000, 64, 64, 4224), (cv, 5, 11000, 13000, 64, 48, 3328),   # overlapping batches
            (cv, 6, 20000, 24000, 64, 64, 4224), (cv, 7, 21000, 22000, 64, 48, 3328)]
    con.executemany("insert into kernels values (?, ?, ?, ?, ?, ?, ?)", rows)
    con.commit()
    con.close()
    args = argparse.Namespace(trace=str(tmp_path / "e.db"), ensemble=True, steps=2)
    profiling.summary(args)
    out = capsys.readouterr().out
    # span 10 us .. 24 us = 0.014 ms over 2 steps
    assert "4 dispatches over 0.01 ms: 0.01 ms per step (2 steps" in out, out
    line = [l for l in out.splitlines() if l.startswith("| `k_corr_pg<float, true>` | 64 x 64 x 4224 |")]
    assert line and line[0].split("|")[3].strip() == "2" and abs(float(line[0].split("|")[5]) - 4.0) < 1e-9, out
    line = [l for l in out.splitlines() if l.startswith("| `k_corr_pg<float, true>` | 64 x 48 x 3328 |")]
    assert line and abs(float(line[0].split("|")[5]) - 1.5) < 1e-9, out


def test_issue_summary_shares(tmp_path, capsys):
    con = sqlite3.connect(str(tmp_path / "p.db"))
    con.execute("create table kernels (name text, start int, end int)")
    con.execute("create table counters_collection (kernel_name text, counter_name text, value real)")
    k = "void pllc::k_rhs_c<float, true, false>(pllc::Dims)"
    # two dispatches of 50 us; 3.84e7 instructions each: 3.84e7 x 4 / (1024 x 50e-6 x 2.4e9) = 1.25
    con.executemany("insert into kernels values (?, ?, ?)", [(k, 0, 50000), (k, 60000, 110000)])
    con.executemany("insert into counters_collection values (?, ?, ?)",
                    [(k, "SQ_INSTS_VALU", 3.84e7), (k, "SQ_INSTS_VALU", 3.84e7), (k, "SQ_WAVES", 65536.0),
                     (k, "SQ_WAVES", 65536.0)])
    con.commit()
    con.close()
    profiling.summary(argparse.Namespace(trace=str(tmp_path / "p.db"), issue=True, cus=256, clock_ghz=2.4))
    out = capsys.readouterr().out
    line = [l for l in out.splitlines() if l.startswith("| `k_rhs_c<float, true, false>`")]
    assert line, out
    cells = [c.strip() for c in line[0].strip("|").split("|")]
    assert cells[1] == "50.0" and cells[3] == "65536" and cells[4] == "586" and cells[5] == "125%", cells
