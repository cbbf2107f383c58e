"""Summaries of rocprofv3 output databases (SQLite, ``*_results.db``).

    python -m dem_mph_amd.profiling --trace T.db [--fetch F.db --write W.db --copy-mib 256]
                                    [--title ...] [--command ...] --out profiles/<name>.md [--json ...]

- kernel statistics from a ``--kernel-trace --stats`` run (calls, average and
  total duration per kernel);
- HBM traffic per dispatch from two ``--pmc`` runs, FETCH_SIZE and WRITE_SIZE
  (KiB per dispatch), each scaled by what the same counter reads for a plain
  device-to-device copy of known size (``bench.py --copy-probe MIB``, or
  ``python -m dem_mph_amd.surrogate --copy-probe MIB``): the copy
  kernel is the ``__amd_rocclr_copyBuffer`` dispatch with the largest count.
"""
from __future__ import annotations

import argparse
import collections
import json
import sqlite3
from typing import Dict, List, Tuple


def copy_probe(mib: int) -> None:
    """One device-to-device copy of `mib` MiB through the HIP runtime: the copy
    kernel's FETCH_SIZE / WRITE_SIZE under rocprof calibrate the counters
    against a known byte count."""
    import ctypes as C
    from dem_mph_amd import share_hip_runtime
    share_hip_runtime()
    hip = C.CDLL("libamdhip64.so.7")   # by soname: the runtime already loaded, if any
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def kernel_stats(db: str) -> List[Tuple[str, int, float, float]]:
    """(name, calls, average us, total us) per kernel, by total time."""
    c = sqlite3.connect(db)
    rows = []
    for name, n, tot in c.execute("select name, count(*), sum(duration) from kernels group by name"):
        rows.append((name, int(n), tot / n / 1e3, tot / 1e3))
    rows.sort(key=lambda r: -r[3])
    return rows


def pmc_per_dispatch(db: str, counter: str) -> Dict[str, List[float]]:
    """Counter value per dispatch (summed over hardware instances), grouped by kernel name."""
    c = sqlite3.connect(db)
    per = collections.defaultdict(float)
    names = {}
    for name, v, did in c.execute("select name, counter_value, dispatch_id from pmc_events where counter_name = ?",
                                  (counter,)):
        per[did] += v
        names[did] = name
    out = collections.defaultdict(list)
    for did, v in per.items():
        out[names[did]].append(v)
    return out


def pmc_means(db: str) -> Dict[str, Dict[str, float]]:
    """Mean per dispatch of every counter in a --pmc database, per kernel, with
    the mean dispatch duration of that run (ns) under "duration_ns"."""
    c = sqlite3.connect(db)
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    names = {}
    for name, cn, v, did, du in c.execute(
            "select name, counter_name, counter_value, dispatch_id, duration from pmc_events"):
        per[did][cn] += v
        per[did]["duration_ns"] = float(du)
        names[did] = name
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for did, cs in per.items():
        for cn, v in cs.items():
            acc[names[did]][cn].append(v)
    return {k: {cn: sum(v) / len(v) for cn, v in cs.items()} for k, cs in acc.items()}


def traffic(fetch_db: str, write_db: str, copy_mib: int) -> Tuple[Dict[str, Dict[str, float]], Dict[str, float]]:
    """Calibrated HBM bytes per dispatch per kernel, and the calibration factors."""
    f = pmc_per_dispatch(fetch_db, "FETCH_SIZE")
    w = pmc_per_dispatch(write_db, "WRITE_SIZE")
    known_kib = copy_mib * 1024.0
    cf = known_kib / max(f.get("__amd_rocclr_copyBuffer", [known_kib]))
    cw = known_kib / max(w.get("__amd_rocclr_copyBuffer", [known_kib]))
    out = {}
    for name in set(f) | set(w):
        fr = sum(f.get(name, [0.0])) / max(len(f.get(name, [1])), 1)
        wr = sum(w.get(name, [0.0])) / max(len(w.get(name, [1])), 1)
        out[name] = {"fetch_kib_raw": fr, "write_kib_raw": wr, "read_bytes": fr * 1024 * cf,
                     "write_bytes": wr * 1024 * cw, "bytes": fr * 1024 * cf + wr * 1024 * cw}
    return out, {"fetch": cf, "write": cw}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--title", default="rocprofv3 summary")
    ap.add_argument("--command", default="")
    ap.add_argument("--out", required=True)
    ap.add_argument("--json")
    ap.add_argument("--top", type=int, default=14)
    ap.add_argument("--counters", help="a --pmc database with SQ_ counters (VALU utilisation)")
    ap.add_argument("--l1", help="a --pmc database with TCP_TOTAL_CACHE_ACCESSES_sum and TCP_TCC_READ_REQ_sum")
    ap.add_argument("--cus", type=int, default=256, help="CUs on the device (one L1 each)")
    ap.add_argument("--sclk-mhz", type=float, default=2400.0, help="shader clock for the utilisation estimate")
    ap.add_argument("--simds", type=int, default=1024, help="SIMDs on the device (256 CUs x 4)")
    a = ap.parse_args(argv)
    st = kernel_stats(a.trace)
    lines = ["# " + a.title, ""]
    if a.command:
        lines += ["Command:", "", "```", a.command, "```", ""]
    lines += ["| kernel | calls | avg us | total us |", "|---|---|---|---|"]
    for name, n, avg, tot in st[: a.top]:
        lines.append("| `%s` | %d | %.2f | %.1f |" % (name, n, avg, tot))
    js = {"kernels": {n: {"calls": c, "avg_us": av, "total_us": t} for n, c, av, t in st}}
    if a.fetch and a.write:
        tr, fac = traffic(a.fetch, a.write, a.copy_mib)
        lines += ["", "HBM traffic per dispatch (FETCH_SIZE x %.3f, WRITE_SIZE x %.3f; factors from a %d MiB "
                      "device-to-device copy):" % (fac["fetch"], fac["write"], a.copy_mib), "",
                  "| kernel | read MB | write MB | total MB |", "|---|---|---|---|"]
        for name, v in sorted(tr.items(), key=lambda kv: -kv[1]["bytes"]):
            if v["bytes"] < 1e6:
                continue
            lines.append("| `%s` | %.1f | %.1f | %.1f |" % (name, v["read_bytes"] / 1e6, v["write_bytes"] / 1e6,
                                                          v["bytes"] / 1e6))
        js["traffic"] = tr
        js["calibration"] = fac
    if a.counters:
        pm = pmc_means(a.counters)
        lines += ["", "Counters per dispatch (one `--pmc` pass).  VALU issue utilisation = SQ_INSTS_VALU x 4 cycles "
                      "(a wave64 VALU instruction occupies a 16-lane SIMD for 4 cycles) / (%d SIMDs x the dispatch's "
                      "cycles at %.0f MHz), durations from the same pass:" % (a.simds, a.sclk_mhz), "",
                  "| kernel | waves | VALU instr / wave | VMEM reads / wave | LDS instr / wave | VALU issue util |",
                  "|---|---|---|---|---|---|"]
        js["counters"] = {}
        for name, n, avg, tot in st[: a.top]:
            c = pm.get(name)
            if not c or "SQ_WAVES" not in c or c["SQ_WAVES"] <= 0:
                continue
            w = c["SQ_WAVES"]
            cyc = c.get("duration_ns", 0.0) * a.sclk_mhz / 1e3
            util = c.get("SQ_INSTS_VALU", 0.0) * 4.0 / (a.simds * cyc) if cyc > 0 else 0.0
            lines.append("| `%s` | %d | %.0f | %.1f | %.1f | %.2f |" % (
                name, w, c.get("SQ_INSTS_VALU", 0) / w, c.get("SQ_INSTS_VMEM_RD", 0) / w, c.get("SQ_INSTS_LDS", 0) / w,
                util))
            js["counters"][name] = dict(c, valu_issue_util=util)
    if a.l1:
        pl = pmc_means(a.l1)
        pq = pmc_means(a.counters) if a.counters else {}
        lines += ["", "Vector L1 (TCP) per dispatch (one `--pmc` pass): tag lookups, lookups per vector-memory "
                      "instruction (SQ_INSTS_VMEM_RD from the counter pass above), misses sent to L2, and the lookup "
                      "rate against one lookup per clock per CU (%d CUs at %.0f MHz):" % (a.cus, a.sclk_mhz), "",
                  "| kernel | L1 lookups (M) | lookups / VMEM instr | to L2 (M) | L1 lookup util |", "|---|---|---|---|---|"]
        js["l1"] = {}
        for name, n, avg, tot in st[: a.top]:
            c = pl.get(name)
            if not c or "TCP_TOTAL_CACHE_ACCESSES_sum" not in c:
                continue
            acc = c["TCP_TOTAL_CACHE_ACCESSES_sum"]
            cyc = c.get("duration_ns", 0.0) * a.sclk_mhz / 1e3
            util = acc / (a.cus * cyc) if cyc > 0 else 0.0
            q = pq.get(name, {})
            vm = q.get("SQ_INSTS_VMEM_RD", 0.0)
            per = acc / vm if vm > 0 else float("nan")
            if acc < 1e5:
                continue
            lines.append("| `%s` | %.1f | %.1f | %.2f | %.2f |" % (name, acc / 1e6, per, c.get("TCP_TCC_READ_REQ_sum", 0.0) / 1e6,
                                                                 util))
            js["l1"][name] = dict(c, lookups_per_vmem=per, l1_lookup_util=util)
    open(a.out, "w").write("\n".join(lines) + "\n")
    if a.json:
        json.dump(js, open(a.json, "w"), indent=1, sort_keys=True)
    print("\n".join(lines))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
