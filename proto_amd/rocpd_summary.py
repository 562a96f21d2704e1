"""Summarise rocprofv3 rocpd databases (kernel durations; FETCH_SIZE /
WRITE_SIZE per dispatch, calibrated against a copy kernel of known size) into
the short text tables kept under profiles/.

    python -m proto_amd.rocpd_summary trace  DB [NAME_SUBSTRING]
    python -m proto_amd.rocpd_summary calls  DB NAME_SUBSTRING [LAST_N]
    python -m proto_amd.rocpd_summary groups DB [TOP_N]
    python -m proto_amd.rocpd_summary pmc    FETCH_DB WRITE_DB COPY_BYTES
"""
import sqlite3
import sys


def trace(db, pattern=""):
    """Per kernel: dispatches, mean / median / min duration, the mean of the
    last 5 dispatches (a benchmark's timed launches come last), registers and
    LDS.  Durations in microseconds."""
    c = sqlite3.connect(db)
    rows = {}
    for name, dur, vgpr, agpr, lds in c.execute(
            "select name, duration, vgpr_count, accum_vgpr_count, lds_size from kernels order by dispatch_id"):
        if pattern and pattern not in name:
            continue
        r = rows.setdefault(name, {"d": [], "vgpr": vgpr, "agpr": agpr, "lds": lds})
        r["d"].append(dur / 1000.0)
    out = ["%-84s %6s %10s %10s %10s %10s %5s %6s" % ("kernel", "calls", "avg_us", "median_us", "min_us",
                                                     "last5_us", "vgpr", "lds")]
    for name, r in sorted(rows.items(), key=lambda kv: -sum(kv[1]["d"])):
        d = r["d"]
        sd = sorted(d)
        last = d[-5:]
        out.append("%-84s %6d %10.2f %10.2f %10.2f %10.2f %5d %6d" % (
            name[:84], len(d), sum(d) / len(d), sd[len(sd) // 2], sd[0], sum(last) / len(last),
            r["vgpr"] + (r["agpr"] or 0), r["lds"]))
    return "\n".join(out)


def calls(db, pattern, last=40):
    """The last `last` dispatches of the kernels whose name contains `pattern`:
    duration, grid and workgroup size, in dispatch order."""
    c = sqlite3.connect(db)
    rows = [r for r in c.execute("select name, duration, grid_x, grid_y, grid_z, workgroup_x from kernels "
                                 "order by dispatch_id") if pattern in r[0]]
    out = ["%-70s %10s %20s %6s" % ("kernel", "dur_us", "grid (threads)", "wg")]
    for name, dur, gx, gy, gz, wg in rows[-last:]:
        out.append("%-70s %10.2f %20s %6s" % (name[:70], dur / 1000.0, "%dx%dx%d" % (gx, gy, gz), wg))
    return "\n".join(out)


def groups(db, top=30):
    """Total time per (kernel, grid, workgroup) over the whole database -- one
    solve's breakdown by kernel and level size when the profiled program runs
    one (kernels that split z to fill the GPU launch the same thread count on
    several levels; their workgroup of 64 x boxes-along-x tells those apart) --
    with calls and the mean per call, largest first."""
    c = sqlite3.connect(db)
    acc = {}
    for name, dur, gx, gy, gz, wx in c.execute("select name, duration, grid_x, grid_y, grid_z, workgroup_x "
                                               "from kernels"):
        k = (name.split("(")[0][:64], "%dx%dx%d/%d" % (gx, gy, gz, wx))
        a = acc.setdefault(k, [0, 0.0])
        a[0] += 1
        a[1] += dur / 1e6
    tot = sum(v[1] for v in acc.values())
    out = ["total %.3f ms over %d dispatches" % (tot, sum(v[0] for v in acc.values())),
           "%-64s %-22s %6s %10s %10s" % ("kernel", "grid (threads)/wg", "calls", "total_ms", "us/call")]
    for (name, grid), (n, t) in sorted(acc.items(), key=lambda kv: -kv[1][1])[:top]:
        out.append("%-64s %-22s %6d %10.3f %10.1f" % (name, grid, n, t, 1e3 * t / n))
    return "\n".join(out)


def _per_kernel(db):
    c = sqlite3.connect(db)
    acc = {}
    for name, value, grid in c.execute("select kernel_name, value, grid_size from counters_collection"):
        acc.setdefault((name, grid), []).append(value)
    return {k: sum(v) / len(v) for k, v in acc.items()}


def pmc(fetch_db, write_db, copy_bytes):
    f, w = _per_kernel(fetch_db), _per_kernel(write_db)
    copy_keys = [k for k in f if "copyBuffer" in k[0] and f[k] > 1e4]
    ck = max(copy_keys, key=lambda k: f[k])
    half = copy_bytes / 2 / 1024.0  # a copy reads and writes copy_bytes/2 each
    ff, wf = half / f[ck], half / w[ck]
    out = ["calibration: plain copy of %d B reads and writes %.0f KiB each; counters read FETCH %.0f KiB, "
           "WRITE %.0f KiB -> factors fetch x%.4f, write x%.4f" % (copy_bytes // 2, half, f[ck], w[ck], ff, wf),
           "%-60s %14s %14s %14s" % ("kernel (grid)", "fetch_GB", "write_GB", "traffic_GB")]
    for k in sorted(f, key=lambda k: -f[k]):
        if f[k] < 1e3 or k not in w:
            continue
        fg, wg = f[k] * ff * 1024 / 1e9, w[k] * wf * 1024 / 1e9
        out.append("%-60s %14.4f %14.4f %14.4f" % ((k[0][:48] + " (%d)" % k[1]), fg, wg, fg + wg))
    return "\n".join(out)


if __name__ == "__main__":
    if sys.argv[1] == "trace":
        print(trace(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else ""))
    elif sys.argv[1] == "calls":
        print(calls(sys.argv[2], sys.argv[3], int(sys.argv[4]) if len(sys.argv) > 4 else 40))
    elif sys.argv[1] == "groups":
        print(groups(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 30))
    else:
        print(pmc(sys.argv[2], sys.argv[3], int(sys.argv[4])))
