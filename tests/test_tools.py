"""The rocprofv3 database summaries used for profiles/ (proto_amd.rocpd_summary),
on small synthetic databases with the tables and columns rocprofv3 writes."""
import sqlite3

import pytest

from proto_amd import rocpd_summary as rs


def _kernels_db(path, rows):
    c = sqlite3.connect(path)
    c.execute("create table kernels (dispatch_id integer, name text, duration integer, grid_x integer, "
              "grid_y integer, grid_z integer, workgroup_x integer, vgpr_count integer, "
              "accum_vgpr_count integer, lds_size integer)")
    for i, row in enumerate(rows):
        name, dur, grid = row[:3]
        wg = row[3] if len(row) > 3 else 256
        c.execute("insert into kernels values (?,?,?,?,?,?,?,?,?,?)", (i, name, dur, grid[0], grid[1], grid[2], wg,
                                                                       40, 0, 1024))
    c.commit()
    c.close()


def _counters_db(path, rows):
    c = sqlite3.connect(path)
    c.execute("create table counters_collection (kernel_name text, value real, grid_size integer)")
    c.executemany("insert into counters_collection values (?,?,?)", rows)
    c.commit()
    c.close()


def test_trace_and_calls(tmp_path):
    db = str(tmp_path / "t.db")
    # durations in ns: kernel A 1,2,3,4,5,6 us (last five mean 4 us), kernel B 10 us
    _kernels_db(db, [("A(int)", 1000 * k, (256, 1, 1)) for k in range(1, 7)] + [("B(float)", 10000, (512, 2, 1))])
    out = rs.trace(db).splitlines()
    assert out[0].split()[:2] == ["kernel", "calls"]
    a = next(l for l in out if l.startswith("A(int)")).split()
    assert a[1] == "6" and float(a[2]) == pytest.approx(3.5) and float(a[5]) == pytest.approx(4.0)
    assert out[1].startswith("A(int)")  # sorted by total time (21 us > 10 us)
    calls = rs.calls(db, "B").splitlines()
    assert len(calls) == 2 and "512x2x1" in calls[1] and "10.00" in calls[1]


def test_groups_by_kernel_and_grid(tmp_path):
    db = str(tmp_path / "g.db")
    _kernels_db(db, [("void k<double>(int)", 2000, (1024, 512, 1)), ("void k<double>(int)", 4000, (1024, 512, 1)),
                     ("void k<double>(int)", 500, (1024, 64, 1)), ("void j(float)", 1500, (64, 1, 1)),
                     ("void j(float)", 700, (64, 1, 1), 128)])
    out = rs.groups(db).splitlines()
    assert out[0] == "total 0.009 ms over 5 dispatches"
    first = out[2].split()
    assert first[:2] == ["void", "k<double>"] and first[2] == "1024x512x1/256" and first[3] == "2"
    assert float(first[4]) == pytest.approx(0.006) and float(first[5]) == pytest.approx(3.0)
    # the same grid with another workgroup size is another level: its own line
    assert any(l.split()[:3] == ["void", "j", "64x1x1/128"] for l in out[2:])
    assert len(out) == 6


def test_pmc_calibrates_against_the_copy(tmp_path):
    f, w = str(tmp_path / "f.db"), str(tmp_path / "w.db")
    copy = 1 << 30  # a 1 GiB copy reads and writes 512 MiB each
    # counters read FETCH at half the true bytes, WRITE at the true bytes (KiB)
    _counters_db(f, [("__amd_rocclr_copyBuffer", 262144.0, 4096), ("kern", 100000.0, 8192)])
    _counters_db(w, [("__amd_rocclr_copyBuffer", 524288.0, 4096), ("kern", 50000.0, 8192)])
    out = rs.pmc(f, w, copy).splitlines()
    assert "fetch x2.0000, write x1.0000" in out[0]
    row = next(l for l in out if l.startswith("kern ("))
    fg, wg, tg = (float(v) for v in row.split()[-3:])
    assert fg == pytest.approx(2 * 100000 * 1024 / 1e9, abs=1e-4)
    assert wg == pytest.approx(50000 * 1024 / 1e9, abs=1e-4)
    assert tg == pytest.approx(fg + wg, abs=2e-4)
