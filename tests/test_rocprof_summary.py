"""The profile summariser (cgdycore.jl_amd/rocprof_summary.py) on small
SQLite files with the tables rocprofv3 writes and that it reads: kernel
table, HBM traffic from FETCH_SIZE / WRITE_SIZE scaled by a copy's
counters, and SQ wave-time shares, each against values worked out by hand."""
import importlib.util
import json
import os
import sqlite3

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def rs():
    spec = importlib.util.spec_from_file_location("rocprof_summary",
                                                  os.path.join(ROOT, "cgdycore.jl_amd", "rocprof_summary.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _db(path, kernels=(), counters=()):
    con = sqlite3.connect(path)
    con.execute("create table top_kernels (name text, total_calls int, total_duration real, average real, "
                "percentage real)")
    con.execute("create table counters_collection (kernel_name text, counter_name text, value real)")
    con.executemany("insert into top_kernels values (?,?,?,?,?)", kernels)
    con.executemany("insert into counters_collection values (?,?,?)", counters)
    con.commit()
    con.close()
    return str(path)


def test_kernels_table(rs, tmp_path):
    db = _db(tmp_path / "k.db", kernels=[("void k_fcn2<5, 8, true, double>(KArgs, int)", 8, 1600.0, 200.0, 80.0),
                                          ("k_prepare(KArgs)", 4, 400.0, 100.0, 20.0)])
    out = tmp_path / "k.txt"
    tab = rs.write_kernels(db, "cmd", str(out))
    assert [t[1] for t in tab] == [8, 4]
    text = out.read_text()
    assert "# command: cmd" in text and "k_fcn2<5, 8, true, double>" in text and "200.00" in text


def test_traffic(rs, tmp_path):
    # the copy of 1 MiB reads 512 KiB on FETCH_SIZE (factor 0.5) and 1024 KiB on WRITE_SIZE (factor 1)
    cal_f = _db(tmp_path / "cf.db", counters=[("__amd_rocclr_copyBuffer", "FETCH_SIZE", 512.0)] * 3)
    cal_w = _db(tmp_path / "cw.db", counters=[("__amd_rocclr_copyBuffer", "WRITE_SIZE", 1024.0)] * 3)
    # 8 dispatches = 2 evaluations of 4: 100 KiB fetched, 50 KiB written per dispatch
    f = _db(tmp_path / "f.db", counters=[("k_fcn2<5, 8, true, double>", "FETCH_SIZE", 100.0)] * 8)
    w = _db(tmp_path / "w.db", counters=[("k_fcn2<5, 8, true, double>", "WRITE_SIZE", 50.0)] * 8)

    class A:
        pass

    a = A()
    a.fetch, a.write, a.cal_fetch, a.cal_write = f, w, cal_f, cal_w
    a.kernel, a.workload, a.out = "k_fcn2", "wl", str(tmp_path / "t.json")
    a.command_fetch = a.command_write = "cmd"
    a.copy_bytes, a.per_eval, a.algorithmic = 1 << 20, 4, 4 * 200 * 1024
    res = rs.write_traffic(a)
    assert res["fetch_factor"] == 0.5 and res["write_factor"] == 1.0 and res["evals_profiled"] == 2.0
    # per evaluation: 4 x 100 KiB / 0.5 fetched, 4 x 50 KiB / 1 written
    assert res["fetch_bytes_per_eval"] == 800 * 1024
    assert res["write_bytes_per_eval"] == 200 * 1024
    assert res["traffic_over_algorithmic"] == pytest.approx(1000 / 800)
    assert json.loads(open(a.out).read())["traffic_bytes_per_eval"] == 1000 * 1024


def test_sq_shares(rs, tmp_path):
    rows = []
    for _ in range(2):  # two dispatches of the kernel, one of another
        rows += [("k_fcn2<5, 8, true, double>", "SQ_WAVE_CYCLES", 100.0), ("k_fcn2<5, 8, true, double>",
                                                                            "SQ_WAIT_ANY", 50.0),
                 ("k_fcn2<5, 8, true, double>", "SQ_ACTIVE_INST_ANY", 25.0)]
    rows += [("k_fcn1", "SQ_WAVE_CYCLES", 1000.0)]
    db = _db(tmp_path / "s.db", counters=rows)
    lines = rs.write_sq(db, "k_fcn2", "cmd", str(tmp_path / "s.txt"))
    text = "\n".join(lines)
    assert "2 dispatches" in text
    assert "SQ_WAVE_CYCLES         2.0000e+02" in text
    assert "share SQ_WAIT_ANY          0.500" in text
    assert "share SQ_ACTIVE_INST_ANY   0.250" in text
