"""Summaries of rocprofv3 output for the profiles kept with the sources.

    python -m specmc_amd.profsum stats  RUN_kernel_stats.csv  "title" "command"
    python -m specmc_amd.profsum traffic FETCH.csv WRITE.csv KERNEL PROBE_MIB "command"

``stats`` turns a ``--kernel-trace --stats`` CSV into a markdown table.
``traffic`` reads two ``--pmc`` counter CSVs (FETCH_SIZE, WRITE_SIZE; KiB per
dispatch) and prints JSON with the mean HBM bytes per dispatch of ``KERNEL``:
each counter is scaled by the factor that makes it read the known size of
the device-to-device copy of PROBE_MIB MiB in the same run (its first
``__amd_rocclr_copyBuffer`` dispatch of that size).
"""
import csv
import json
import sys


def stats_md(path, title, command, top=25):
    rows = list(csv.DictReader(open(path)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    out = ["# " + title, "", "Command (repository root, one MI355X):", "", "```", command,
           "```", "", "Total kernel time: %.1f ms over %d kernel names." % (tot / 1e6, len(rows)),
           "", "| % time | total ms | calls | avg us | kernel |", "|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:top]:
        name = r["Name"].replace("|", "/")
        if len(name) > 90:
            name = name[:90] + "..."
        out.append("| %.2f | %.2f | %s | %.1f | `%s` |" % (
            float(r["Percentage"]), float(r["TotalDurationNs"]) / 1e6, r["Calls"],
            float(r["AverageNs"]) / 1e3, name))
    return "\n".join(out)


def _per_kernel(path, kernel):
    rows = list(csv.DictReader(open(path)))
    vals = [float(r["Counter_Value"]) for r in rows if r["Kernel_Name"].startswith(kernel)]
    probe = [r for r in rows if r["Kernel_Name"].startswith("__amd_rocclr_copyBuffer")]
    return vals, probe


def traffic(fetch_csv, write_csv, kernel, probe_mib, command=""):
    res = dict(kernel=kernel, command=command, probe_mib=probe_mib)
    total = 0.0
    for name, path in (("fetch", fetch_csv), ("write", write_csv)):
        vals, probe = _per_kernel(path, kernel)
        if not vals:
            raise SystemExit("no dispatches of %s in %s" % (kernel, path))
        big = max(probe, key=lambda r: float(r["Counter_Value"]))
        factor = probe_mib * 1024.0 / float(big["Counter_Value"])
        mean = sum(vals) / len(vals) * 1024.0 * factor
        res[name + "_kib_probe"] = float(big["Counter_Value"])
        res[name + "_scale"] = factor
        res[name + "_bytes_per_launch"] = mean
        res[name + "_dispatches"] = len(vals)
        total += mean
    res["hbm_bytes_per_launch"] = total
    return res


def main(argv):
    if argv[:1] == ["stats"]:
        print(stats_md(argv[1], argv[2], argv[3]))
    elif argv[:1] == ["traffic"]:
        print(json.dumps(traffic(argv[1], argv[2], argv[3], float(argv[4]),
                                 argv[5] if len(argv) > 5 else ""), indent=1))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
