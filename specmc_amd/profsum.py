"""Summaries of rocprofv3 output for the profiles kept with the sources.

    python -m specmc_amd.profsum stats  RUN_kernel_stats.csv  "title" "command"
    python -m specmc_amd.profsum traffic FETCH.csv WRITE.csv KERNEL PROBE_MIB "command" [UNITS]
    python -m specmc_amd.profsum letters STATS.csv FETCH.csv WRITE.csv PROBE_MIB "title" \
        "command" [MFMA.csv]

``stats`` turns a ``--kernel-trace --stats`` CSV into a markdown table.
``traffic`` reads two ``--pmc`` counter CSVs (FETCH_SIZE, WRITE_SIZE; KiB per
dispatch) and prints JSON with the mean HBM bytes per dispatch of ``KERNEL``:
each counter is scaled by the factor that makes it read the known size of
the device-to-device copy of PROBE_MIB MiB in the same run (its first
``__amd_rocclr_copyBuffer`` dispatch of that size).  With UNITS (the
algorithmic units every dispatch of the kernel processes) it also gives the
bytes per unit.
``letters`` groups the kernels by the part of the time step they serve --
(a) x-transforms, (b) z-transforms, (c) viscosity law and pseudo-spectral
products, (d) Stokes / Krylov / preconditioner, (e) theta BDF2 pieces,
(f) CFL, rescue -- and prints, per group and per kernel, the HBM bytes per
dispatch (the counter passes, scaled by the copy probe), the mean duration
(the kernel-trace pass) and the achieved GB/s against 8 TB/s; with an MFMA
counter CSV (SQ_INSTS_VALU_MFMA_MOPS_F64 / _F32, 512 flops per unit) also
the matrix-core FLOP/s against the FP64 (78.6 TF) or FP32 (157.3 TF) peak.
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


def traffic(fetch_csv, write_csv, kernel, probe_mib, command="", units=None):
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
    if units:
        res["units_per_launch"] = units
        res["hbm_bytes_per_unit"] = total / units
    return res


PEAK_HBM = 8.0e12
PEAK_F64, PEAK_F32 = 78.6e12, 157.3e12
GROUPS = (
    ("a", "x-transforms (half-grid DFTs)", lambda k: _role(k) == 1),
    ("b", "z-transforms and z-derivatives", lambda k: _role(k) == 0),
    ("c", "viscosity law and pseudo-spectral products",
     lambda k: k.startswith(("k_eta_law", "k_eta_guard", "k_xprod", "k_advect",
                             "k_phi_grid"))),
    ("d", "Stokes operator, Krylov and preconditioner",
     lambda k: _role(k) == 2 or k.startswith(("k_fg_", "k_inv_apply", "k_lu_solve",
                                              "k_lu_factor", "k_gj_invert",
                                              "k_stokes_combine", "k_schur_combine",
                                              "k_tau_assemble", "k_rowscale",
                                              "k_cheb_mult"))),
    ("e", "theta BDF2 pieces", lambda k: k.startswith(("k_theta_", "k_axpby"))),
    ("f", "CFL and diagnostics", lambda k: k.startswith(("k_maxabs", "k_cfl", "k_diag"))),
    ("rescue", "dense rescue (assembly, LU, solves)",
     lambda k: _role(k) == 3 or k.startswith(("k_lud_", "k_ludm_", "k_schur_blocks",
                                               "k_schur_tau"))),
)


def _short(name):
    """Kernel name without 'void ' and the argument list."""
    n = name[5:] if name.startswith("void ") else name
    depth = 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return n[:i]
    return n


def _role(k):
    """Role template argument of the GEMM kernels (last template argument)."""
    if not k.startswith(("k_gemm<", "k_gemm_mfma<")):
        return -1
    return int(k[k.rindex(",") + 1:k.rindex(">")].strip())


def _counter_means(path, probe_mib=None):
    """-> ({short kernel name: mean counter value per dispatch}, probe scale)."""
    rows = list(csv.DictReader(open(path)))
    acc, scale = {}, 1.0
    if probe_mib:
        probe = [r for r in rows if r["Kernel_Name"].startswith("__amd_rocclr_copyBuffer")]
        big = max(probe, key=lambda r: float(r["Counter_Value"]))
        scale = probe_mib * 1024.0 / float(big["Counter_Value"])
    # one row per (dispatch, counter): the counters of a pass are summed per
    # dispatch (an FP64 and an FP32 MFMA counter of one kernel add up)
    for r in rows:
        k = _short(r["Kernel_Name"])
        a = acc.setdefault(k, [0.0, set()])
        a[0] += float(r["Counter_Value"])
        a[1].add(r.get("Dispatch_Id", len(a[1])))
    return {k: v[0] / len(v[1]) for k, v in acc.items()}, scale


def letters_md(stats_csv, fetch_csv, write_csv, probe_mib, title, command, mfma_csv=None):
    st = {}
    for r in csv.DictReader(open(stats_csv)):
        k = _short(r["Name"])
        a = st.setdefault(k, [0, 0.0])
        a[0] += int(r["Calls"])
        a[1] += float(r["TotalDurationNs"])
    fetch, fs = _counter_means(fetch_csv, probe_mib)
    write, ws = _counter_means(write_csv, probe_mib)
    mf = _counter_means(mfma_csv)[0] if mfma_csv else {}
    out = ["# " + title, "", "Command (repository root, one MI355X):", "", "```", command, "```",
           "", "HBM bytes per dispatch: FETCH_SIZE x %.4f + WRITE_SIZE x %.4f (KiB; the "
           "factors make each counter read the %g MiB device-to-device copy of the same "
           "run as its size).  Duration: mean of the kernel-trace pass." % (fs, ws, probe_mib),
           "", "| part | kernel | calls | mean us | HBM MB / dispatch | GB/s | of 8 TB/s |"
           + (" MFMA TFLOP/s | of peak |" if mfma_csv else ""),
           "|---|---|---|---|---|---|---|" + ("---|---|" if mfma_csv else "")]
    summary = []
    for key, label, pred in GROUPS:
        ks = sorted((k for k in st if pred(k)), key=lambda k: -st[k][1])
        if not ks:
            continue
        tb = tt = tf = 0.0
        for k in ks:
            calls, ns = st[k]
            mean_ns = ns / calls
            byt = 1024.0 * (fetch.get(k, 0.0) * fs + write.get(k, 0.0) * ws)
            tb += byt * calls
            tt += ns
            row = "| %s | `%s` | %d | %.1f | %.2f | %.0f | %.3f |" % (
                key, k if len(k) < 60 else k[:60] + "...", calls, mean_ns / 1e3, byt / 1e6,
                byt / mean_ns, byt / mean_ns * 1e9 / PEAK_HBM)
            if mfma_csv:
                fl = 512.0 * mf.get(k, 0.0)
                peak = PEAK_F32 if "<float" in k else PEAK_F64
                tf += fl * calls
                row += " %.2f | %.3f |" % (fl / mean_ns / 1e3, fl / mean_ns * 1e9 / peak)
            out.append(row)
        summary.append((key, label, tt / 1e6, tb / tt if tt else 0.0, tf / tt / 1e3 if tt else 0.0))
    out += ["", "Per part (time-weighted over its kernels):", "",
            "| part | what | kernel time ms | achieved GB/s | of 8 TB/s |"
            + (" MFMA TFLOP/s |" if mfma_csv else ""),
            "|---|---|---|---|---|" + ("---|" if mfma_csv else "")]
    for key, label, ms, gbs, tfl in summary:
        out.append("| %s | %s | %.1f | %.0f | %.3f |" % (key, label, ms, gbs,
                                                        gbs * 1e9 / PEAK_HBM)
                   + (" %.2f |" % tfl if mfma_csv else ""))
    return "\n".join(out)


def main(argv):
    if argv[:1] == ["stats"]:
        print(stats_md(argv[1], argv[2], argv[3]))
    elif argv[:1] == ["letters"]:
        print(letters_md(argv[1], argv[2], argv[3], float(argv[4]), argv[5], argv[6],
                         argv[7] if len(argv) > 7 else None))
    elif argv[:1] == ["traffic"]:
        print(json.dumps(traffic(argv[1], argv[2], argv[3], float(argv[4]),
                                 argv[5] if len(argv) > 5 else "",
                                 float(argv[6]) if len(argv) > 6 else None), indent=1))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
