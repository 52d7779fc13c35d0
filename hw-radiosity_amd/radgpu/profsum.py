"""Per-kernel summary of a rocprofv3 --kernel-trace results database.

    python -m radgpu.profsum gpurun_out/prof/hall_results.db
"""
import sqlite3
import sys


def summarize(db_path):
    con = sqlite3.connect(db_path)
    rows = con.execute(
        "select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
        "from kernels group by name order by sum(duration) desc").fetchall()
    total = sum(r[2] for r in rows) or 1
    out = ["| kernel | calls | total us | avg us | min us | max us | share |", "|---|---|---|---|---|---|---|"]
    for name, n, tot, avg, mn, mx in rows:
        short = name.replace("rgpu::(anonymous namespace)::", "").replace("void ", "").split("(")[0]
        out.append("| `%s` | %d | %.1f | %.2f | %.2f | %.2f | %.1f%% |" % (
            short, n, tot / 1e3, avg / 1e3, mn / 1e3, mx / 1e3, 100.0 * tot / total))
    return "\n".join(out)


if __name__ == "__main__":
    print(summarize(sys.argv[1]))
