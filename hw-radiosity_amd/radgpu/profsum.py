"""Summaries of rocprofv3 results databases.

    python -m radgpu.profsum trace <results.db>            per-kernel durations
    python -m radgpu.profsum pmc <fetch.db> <write.db> <kernel-substring> <copy_bytes>
    python -m radgpu.profsum sq <results.db> <kernel-substring>    SQ counters, summed

The pmc form reports FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the
kernels whose name contains the substring, each scaled by the factor that
makes the same counter read `copy_bytes` for the plain device copy
(__amd_rocclr_copyBuffer) profiled in the same run.  The sq form sums every
counter of the run over the kernel's dispatches and derives the VALU lane
use, SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU): the share of a wave's
64 lanes active over the cycles it spends in VALU instructions.
"""
import json
import sqlite3
import sys


def _short(name):
    return name.replace("rgpu::(anonymous namespace)::", "").replace("void ", "").split("(")[0]


def trace(db_path, min_us=0.0):
    con = sqlite3.connect(db_path)
    rows = con.execute(
        "select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
        "from kernels group by name order by sum(duration) desc").fetchall()
    total = sum(r[2] for r in rows) or 1
    out = ["| kernel | calls | total us | avg us | min us | max us | share |", "|---|---|---|---|---|---|---|"]
    for name, n, tot, avg, mn, mx in rows:
        out.append("| `%s` | %d | %.1f | %.2f | %.2f | %.2f | %.1f%% |" % (
            _short(name), n, tot / 1e3, avg / 1e3, mn / 1e3, mx / 1e3, 100.0 * tot / total))
    return "\n".join(out)


def _counter(db_path, counter):
    con = sqlite3.connect(db_path)
    return con.execute("select kernel_name, dispatch_id, sum(value) from counters_collection "
                       "where counter_name = ? group by kernel_name, dispatch_id", (counter,)).fetchall()


def pmc(fetch_db, write_db, kernel, copy_bytes):
    res = {}
    for counter, db in (("FETCH_SIZE", fetch_db), ("WRITE_SIZE", write_db)):
        rows = _counter(db, counter)
        copy_kib = sum(v for n, _, v in rows if "copyBuffer" in n)
        scale = (copy_bytes / 1024.0) / copy_kib if copy_kib else None
        ker = [v for n, _, v in rows if kernel in n]
        busy = [v for v in ker if v > 0]
        res[counter] = {"copy_kib": copy_kib, "scale": scale, "dispatches": len(ker),
                        "kib_per_dispatch": sum(ker) / max(len(ker), 1),
                        "bytes_per_dispatch_scaled": (sum(ker) / max(len(ker), 1)) * 1024.0 * (scale or 0.0),
                        "nonzero_dispatches": len(busy)}
    res["traffic_bytes_per_dispatch"] = (res["FETCH_SIZE"]["bytes_per_dispatch_scaled"]
                                         + res["WRITE_SIZE"]["bytes_per_dispatch_scaled"])
    return res


def sq(db_path, kernel):
    con = sqlite3.connect(db_path)
    rows = con.execute("select kernel_name, counter_name, dispatch_id, sum(value) from counters_collection "
                       "group by kernel_name, counter_name, dispatch_id").fetchall()
    tot, disp = {}, set()
    for name, counter, d, v in rows:
        if kernel in name:
            tot[counter] = tot.get(counter, 0.0) + v
            disp.add(d)
    res = {"kernel_substring": kernel, "dispatches": len(disp), "counters": tot}
    nd = max(len(disp), 1)
    if tot.get("SQ_ACTIVE_INST_VALU"):
        res["valu_lane_use"] = tot.get("SQ_THREAD_CYCLES_VALU", 0.0) / (64.0 * tot["SQ_ACTIVE_INST_VALU"])
    if tot.get("SQ_THREAD_CYCLES_VALU"):
        # lanes active over the VALU instructions issued: the kernel's VALU work
        res["thread_valu_per_dispatch"] = tot["SQ_THREAD_CYCLES_VALU"] / nd
    if tot.get("SQ_ACTIVE_INST_VALU") and tot.get("GRBM_GUI_ACTIVE"):
        # GRBM_GUI_ACTIVE is summed over the 8 XCDs; a wave64 VALU
        # instruction holds its SIMD (16 lanes) for 4 clocks; 1024 SIMDs
        cycles = tot["GRBM_GUI_ACTIVE"] / 8.0
        res["gui_cycles_per_dispatch"] = cycles / nd
        res["valu_issue_frac"] = 4.0 * tot["SQ_ACTIVE_INST_VALU"] / (cycles * 1024.0)
    if tot.get("SQ_INSTS_VALU") and tot.get("SQ_WAVES"):
        res["valu_insts_per_wave"] = tot["SQ_INSTS_VALU"] / tot["SQ_WAVES"]
    return res


if __name__ == "__main__":
    if sys.argv[1] == "trace":
        print(trace(sys.argv[2]))
    elif sys.argv[1] == "sq":
        print(json.dumps(sq(sys.argv[2], sys.argv[3]), indent=1))
    else:
        print(json.dumps(pmc(sys.argv[2], sys.argv[3], sys.argv[4], float(sys.argv[5])), indent=1))
