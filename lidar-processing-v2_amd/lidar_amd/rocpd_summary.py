"""Per-kernel summaries of rocprofv3 result databases (rocpd SQLite).

  python -m lidar_amd.rocpd_summary trace <run_results.db> [frames_per_launch]
      kernel durations (calls, total, average, min, max), optionally only the
      launches whose grid covers that many frames (grid y / block y)
  python -m lidar_amd.rocpd_summary trace_wg <run_results.db> frames
      the same for launches of one workgroup per frame (grid y = 1 and grid
      x / block x = frames: the row solve, the per-frame scans and
      union-find, ...)
  python -m lidar_amd.rocpd_summary pmc <run_results.db>
      per-kernel counter values averaged per dispatch
  python -m lidar_amd.rocpd_summary traffic <fetch.db> <write.db> <frames> <copy_bytes>
      HBM bytes per frame of every kernel as JSON: FETCH_SIZE and WRITE_SIZE
      (KiB per dispatch, one run each) scaled so that the device-to-device
      copy of copy_bytes in the same runs reads and writes exactly its size
"""
import json
import re
import sqlite3
import sys
from collections import defaultdict


def _short(name):
    m = re.search(r"(k_\w+)", name)
    if m:
        return m.group(1)
    return name.replace("void ", "")[:40]


def trace(db, frames=None, per_frame_grid=False):
    con = sqlite3.connect(db)
    agg = {}
    rows = con.execute("select name, duration, grid_x, workgroup_x, grid_y, workgroup_y from kernels")
    for name, dur, gx, wx, gy, wy in rows:
        nb = gy // max(wy, 1) if gy else 1
        if per_frame_grid:
            if nb != 1 or gx // max(wx, 1) != frames:
                continue
        elif frames is not None and nb != frames:
            continue
        a = agg.setdefault(_short(name), [0, 0.0, float("inf"), 0.0])
        a[0] += 1
        a[1] += dur
        a[2] = min(a[2], dur)
        a[3] = max(a[3], dur)
    tot = sum(v[1] for v in agg.values()) or 1.0
    lines = [f"{'kernel':32s} {'calls':>6s} {'total_ms':>9s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'pct':>6s}"]
    for k, (c, t, mn, mx) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"{k[:32]:32s} {c:6d} {t/1e6:9.3f} {t/c/1e3:9.2f} {mn/1e3:9.2f} {mx/1e3:9.2f} {100*t/tot:6.2f}")
    return "\n".join(lines)


def _per_dispatch(db, counter):
    con = sqlite3.connect(db)
    per = defaultdict(float)
    disp = defaultdict(set)
    rows = con.execute("select dispatch_id, kernel_name, counter_name, value, grid_size from counters_collection")
    for did, name, cn, v, gs in rows:
        if cn != counter:
            continue
        k = _short(name) if "k_" in name else f"copy {gs}" if "copyBuffer" in name else None
        if k is None:
            continue
        per[k] += v
        disp[k].add(did)
    return {k: (per[k] / len(disp[k]), len(disp[k])) for k in per}


def traffic(fetch_db, write_db, frames, copy_bytes):
    fe = _per_dispatch(fetch_db, "FETCH_SIZE")
    wr = _per_dispatch(write_db, "WRITE_SIZE")
    # The calibration copy is the copy launch with the largest grid.
    big = max((k for k in fe if k.startswith("copy ")), key=lambda k: int(k.split()[1]))
    fe["copy"], wr["copy"] = fe[big], wr[big]
    fs = copy_bytes / 1024.0 / fe["copy"][0]
    ws = copy_bytes / 1024.0 / wr["copy"][0]
    out = {"calibration": {"copy_bytes": copy_bytes, "fetch_kib": round(fe["copy"][0], 3),
                           "write_kib": round(wr["copy"][0], 3), "fetch_scale": round(fs, 4),
                           "write_scale": round(ws, 4)}}
    for k in sorted(set(fe) & set(wr)):
        if k.startswith("copy"):
            continue
        f_kib, nd = fe[k]
        w_kib = wr[k][0]
        out[k] = {"fetch_kib_per_dispatch": round(f_kib, 1), "write_kib_per_dispatch": round(w_kib, 1),
                  "dispatches": nd, "frames_per_dispatch": frames,
                  "hbm_bytes_per_frame": round((f_kib * 1024 * fs + w_kib * 1024 * ws) / frames)}
    return out


def pmc(db):
    con = sqlite3.connect(db)
    per = defaultdict(lambda: defaultdict(float))
    disp = defaultdict(set)
    dur = defaultdict(dict)
    rows = con.execute("select dispatch_id, kernel_name, counter_name, value, duration, grid_size from counters_collection")
    for did, name, cn, v, du, gs in rows:
        k = _short(name) + ("" if "k_" in name else f"[grid {gs}]")
        per[k][cn] += v
        disp[k].add(did)
        dur[k][did] = du
    names = sorted({cn for k in per for cn in per[k]})
    lines = [f"{'kernel':40s} {'disp':>5s} {'avg_us':>9s} " + " ".join(f"{n:>14s}" for n in names)]
    for k in sorted(per, key=lambda k: -sum(dur[k].values())):
        nd = len(disp[k])
        lines.append(f"{k[:40]:40s} {nd:5d} {sum(dur[k].values())/nd/1e3:9.1f} " +
                     " ".join(f"{per[k][n]/nd:14.1f}" for n in names))
    return "\n".join(lines)


if __name__ == "__main__":
    mode, db = sys.argv[1], sys.argv[2]
    if mode == "trace":
        print(trace(db, int(sys.argv[3]) if len(sys.argv) > 3 else None))
    elif mode == "trace_wg":
        print(trace(db, int(sys.argv[3]), per_frame_grid=True))
    elif mode == "traffic":
        print(json.dumps(traffic(db, sys.argv[3], int(sys.argv[4]), int(sys.argv[5])), indent=1))
    else:
        print(pmc(db))
