"""Summaries of rocprofv3 output for the step kernels.

* ``kernel_stats(dir)``: the per-kernel duration table of a ``rocprofv3 --kernel-trace --stats --output-format csv``
  run (``*_kernel_stats.csv``);
* ``counter_per_dispatch(dir, counter, match)``: mean value per dispatch of one counter of a ``rocprofv3 --pmc
  <counter> --output-format csv`` run (``*_counter_collection.csv``), over the dispatches whose kernel name contains
  ``match``;
* ``traffic(...)``: HBM bytes per launch of a kernel from a FETCH_SIZE run and a WRITE_SIZE run, each scaled by what
  the same counter reads for a streaming copy of known size in the same run (the calibration kernel);
* ``valu(...)``: VALU busy share and lane utilisation of a kernel from one ``--pmc`` run of VALU_COUNTERS;
* ``mfma(...)``: matrix-core FP32 operations and busy share of a kernel from one ``--pmc`` run of MFMA_COUNTERS;
* ``source_digest()``: digest of the kernel sources and build flags, recorded with a profile so that a report can tell
  whether the profile belongs to the kernels it is reporting on.

Command line: ``python -m genesis_amd.profiling --trace DIR --fetch DIR --write DIR --kernel k_fused_wave
--copy-bytes N --out profiles/NAME`` writes ``NAME.md`` (tables and figures) and ``NAME.json`` (the figures).
"""

from __future__ import annotations

import argparse
import csv
import glob
import hashlib
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
CALIBRATION_MATCH = "elementwise"  # torch's elementwise kernel that bench.py's copy calibration launches


def source_digest() -> str:
    from .build_lib import FLAGS

    h = hashlib.sha256(" ".join(FLAGS).encode())
    csrc = os.path.join(HERE, "csrc")
    for name in sorted(os.listdir(csrc)):
        with open(os.path.join(csrc, name), "rb") as f:
            h.update(name.encode())
            h.update(f.read())
    return h.hexdigest()[:16]


def _rows(directory: str, suffix: str) -> list[dict]:
    files = sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True))
    if not files:
        raise FileNotFoundError(f"no *{suffix} under {directory}")
    out = []
    for path in files:
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                out.append({k.strip().lower(): v for k, v in r.items() if k is not None})
    return out


def kernel_stats(directory: str) -> list[dict]:
    """[{name, calls, avg_us, total_ms, share}] sorted by total time."""
    rows = _rows(directory, "kernel_stats.csv")
    out = []
    for r in rows:
        out.append({"name": r["name"], "calls": int(r["calls"]), "avg_us": float(r["averagens"]) * 1e-3,
                    "total_ms": float(r["totaldurationns"]) * 1e-6, "share": float(r.get("percentage", 0.0))})
    out.sort(key=lambda x: -x["total_ms"])
    return out


def dispatch_avg_us(directory: str, kernel: str, skip: int, count: int) -> tuple[float, int]:
    """Mean duration (us) of dispatches skip .. skip + count - 1, in start order, of the kernels whose name contains
    ``kernel``, from the per-dispatch table (``*_kernel_trace.csv``) of the same run; and how many were averaged."""
    rows = [r for r in _rows(directory, "kernel_trace.csv") if kernel in r["kernel_name"]]
    rows.sort(key=lambda r: int(r["start_timestamp"]))
    sel = rows[skip: skip + count]
    if not sel:
        raise ValueError(f"no dispatches {skip}..{skip + count - 1} of {kernel} under {directory}")
    us = [(int(r["end_timestamp"]) - int(r["start_timestamp"])) * 1e-3 for r in sel]
    return sum(us) / len(us), len(us)


def counter_per_dispatch(directory: str, counter: str, match: str, min_value: float = 0.0,
                         skip: int = 0) -> tuple[float, int]:
    """Mean of ``counter`` over the dispatches of kernels whose name contains ``match`` (and whose value is at least
    ``min_value``), leaving out the first ``skip`` of them in dispatch order (a run's warm-up); returns (mean, number
    of dispatches)."""
    vals: dict[str, float] = {}
    for r in _rows(directory, "counter_collection.csv"):
        if r.get("counter_name") != counter or match not in r.get("kernel_name", ""):
            continue
        key = r.get("dispatch_id") or r.get("correlation_id")
        vals[key] = vals.get(key, 0.0) + float(r["counter_value"])
    ordered = [vals[k] for k in sorted(vals, key=lambda k: int(float(k)))][skip:]
    sel = [v for v in ordered if v >= min_value]
    if not sel:
        raise ValueError(f"no dispatch of *{match}* with {counter}")
    return sum(sel) / len(sel), len(sel)


def traffic(fetch_dir: str, write_dir: str, kernel: str, copy_bytes: int, skip: int = 0) -> dict:
    """HBM bytes per launch of ``kernel``. FETCH_SIZE and WRITE_SIZE are in KiB per dispatch; each is scaled by
    copy_bytes / (what it reads for the calibration copy, which reads and writes copy_bytes)."""
    out = {"kernel": kernel, "copy_bytes": copy_bytes}
    total = 0.0
    for name, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        k_kib, n_k = counter_per_dispatch(d, name, kernel, skip=skip)
        c_kib, n_c = counter_per_dispatch(d, name, CALIBRATION_MATCH, min_value=0.5 * copy_bytes / 1024)
        factor = copy_bytes / (c_kib * 1024.0)
        out[name] = {"kernel_kib_per_dispatch": k_kib, "kernel_dispatches": n_k, "copy_kib_per_dispatch": c_kib,
                     "copy_dispatches": n_c, "factor": factor, "bytes_per_launch": k_kib * 1024.0 * factor}
        total += k_kib * 1024.0 * factor
    out["bytes_per_launch"] = total
    return out


VALU_COUNTERS = ("SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_WAVE_CYCLES", "GRBM_GUI_ACTIVE")


def valu(directory: str, kernel: str, simds: int, xcds: int = 8, skip: int = 0) -> dict:
    """VALU figures of ``kernel`` from one ``--pmc`` run of VALU_COUNTERS (means per dispatch). GRBM_GUI_ACTIVE is
    summed over the device's XCDs (it reads about xcds times the kernel's duration in cycles), so the device's busy
    cycles are GRBM_GUI_ACTIVE / xcds:
    valu_busy = SQ_ACTIVE_INST_VALU * 4 / (simds * GRBM_GUI_ACTIVE / xcds), the share of SIMD cycles issuing VALU work;
    valu_lane_util = SQ_THREAD_CYCLES_VALU / (64 * SQ_ACTIVE_INST_VALU), the share of lanes active in them."""
    out = {"simds": simds, "xcds": xcds}
    for c in VALU_COUNTERS:
        out[c], out["dispatches"] = counter_per_dispatch(directory, c, kernel, skip=skip)
    out["valu_busy"] = out["SQ_ACTIVE_INST_VALU"] * 4.0 / (simds * out["GRBM_GUI_ACTIVE"] / xcds)
    out["valu_lane_util"] = out["SQ_THREAD_CYCLES_VALU"] / (64.0 * out["SQ_ACTIVE_INST_VALU"])
    return out


MFMA_COUNTERS = ("SQ_INSTS_VALU_MFMA_MOPS_F32", "SQ_VALU_MFMA_BUSY_CYCLES", "SQ_INSTS_MFMA", "GRBM_GUI_ACTIVE")
MFMA_FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix peak


def mfma(directory: str, kernel: str, simds: int, xcds: int = 8, skip: int = 0) -> dict:
    """Matrix-core figures of ``kernel`` from one ``--pmc`` run of MFMA_COUNTERS (means per dispatch):
    flops = SQ_INSTS_VALU_MFMA_MOPS_F32 * 512 (the counter counts FP32 matrix multiply-adds in units of 512 FLOPs);
    mfma_busy = SQ_VALU_MFMA_BUSY_CYCLES / (simds * GRBM_GUI_ACTIVE / xcds), the share of SIMD cycles the matrix
    cores were busy (GRBM_GUI_ACTIVE summed over the XCDs, as in ``valu``)."""
    out = {"simds": simds, "xcds": xcds}
    for c in MFMA_COUNTERS:
        out[c], out["dispatches"] = counter_per_dispatch(directory, c, kernel, skip=skip)
    out["flops"] = out["SQ_INSTS_VALU_MFMA_MOPS_F32"] * 512.0
    out["mfma_busy"] = out["SQ_VALU_MFMA_BUSY_CYCLES"] / (simds * out["GRBM_GUI_ACTIVE"] / xcds)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--kernel", default="k_fused_wave")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--bench-json", default=None, help="bench.py output of the traced run")
    ap.add_argument("--title", default="")
    ap.add_argument("--commands", default="", help="text of the commands that produced the runs")
    ap.add_argument("--valu", default=None, help="directory of a --pmc run of VALU_COUNTERS")
    ap.add_argument("--mfma", default=None, help="directory of a --pmc run of MFMA_COUNTERS")
    ap.add_argument("--pmc-skip", type=int, default=0,
                    help="leave out each counter run's first N dispatches of the kernel (its warm-up substeps)")
    ap.add_argument("--simds", type=int, default=1024, help="SIMDs of the device (MI355X: 256 CUs x 4)")
    ap.add_argument("--xcds", type=int, default=8, help="XCDs of the device (MI355X: 8)")
    ap.add_argument("--out", required=True, help="output path without extension")
    a = ap.parse_args()
    stats = kernel_stats(a.trace)
    tr = traffic(a.fetch, a.write, a.kernel, a.copy_bytes, a.pmc_skip)
    vu = valu(a.valu, a.kernel, a.simds, a.xcds, a.pmc_skip) if a.valu else None
    mf = mfma(a.mfma, a.kernel, a.simds, a.xcds, a.pmc_skip) if a.mfma else None
    k = next(s for s in stats if a.kernel in s["name"])
    bench = None
    if a.bench_json:
        with open(a.bench_json) as f:
            lines = [ln for ln in f if ln.startswith("{")]
        bench = json.loads(lines[-1]) if lines else None
    fig = {"kernel": a.kernel, "avg_us": k["avg_us"], "calls": k["calls"], "traffic": tr, "valu": vu, "mfma": mf,
           "source_digest": source_digest(), "title": a.title}
    if bench:
        rf = bench["roofline"]
        fig["bench"] = {"value": bench["value"], "kernel_ms": rf["kernel_ms"],
                        "bytes_per_env_step": rf["bytes_per_env_step"], "n_envs": bench["config"]["n_envs"]}
        fig["algorithmic_bytes_per_launch"] = rf["bytes_per_env_step"] * bench["config"]["n_envs"]
        # bench.py runs its warm-up substeps, then the timed ones: the same dispatches of the trace
        fig["timed_avg_us"], fig["timed_dispatches"] = dispatch_avg_us(a.trace, a.kernel, int(bench["warmup"]),
                                                                       int(bench["steps"]))
    with open(a.out + ".json", "w") as f:
        json.dump(fig, f, indent=1)
    lines = [a.title or f"{a.kernel}: rocprofv3 summary", "=" * 100, "", "Commands (repo root, one MI355X):", ""]
    lines += ["  " + ln for ln in a.commands.strip().split("\n")] + [""]
    lines += ["Kernel trace (--kernel-trace --stats):", "",
              f"{'kernel':60s} {'calls':>7s} {'avg us':>10s} {'total ms':>10s} {'share':>7s}"]
    for s in stats[:12]:
        nm = s["name"] if len(s["name"]) <= 60 else s["name"][:57] + "..."
        lines.append(f"{nm:60s} {s['calls']:7d} {s['avg_us']:10.1f} {s['total_ms']:10.2f} {s['share']:6.1f}%")
    skipped = f", after the first {a.pmc_skip} (warm-up) dispatches" if a.pmc_skip else ""
    lines += ["", f"HBM traffic of {a.kernel} (--pmc FETCH_SIZE, then --pmc WRITE_SIZE; KiB per dispatch{skipped}):", ""]
    for name in ("FETCH_SIZE", "WRITE_SIZE"):
        t = tr[name]
        lines.append(f"  {name}: kernel {t['kernel_kib_per_dispatch']:.1f} KiB/dispatch ({t['kernel_dispatches']} "
                     f"dispatches); calibration copy of {a.copy_bytes} B reads {t['copy_kib_per_dispatch']:.1f} KiB "
                     f"({t['copy_dispatches']} dispatches) -> factor {t['factor']:.4f} -> "
                     f"{t['bytes_per_launch'] / 1e6:.3f} MB/launch")
    lines.append(f"  traffic: {tr['bytes_per_launch'] / 1e6:.3f} MB per launch")
    if vu:
        lines += ["", f"VALU of {a.kernel} (--pmc {' '.join(VALU_COUNTERS)}; means per dispatch, {vu['simds']} SIMDs, "
                      f"{vu['xcds']} XCDs):",
                  "", "  " + ", ".join(f"{c} {vu[c]:.4g}" for c in VALU_COUNTERS),
                  f"  VALU busy (SQ_ACTIVE_INST_VALU * 4 / (SIMDs * GRBM_GUI_ACTIVE / XCDs)): {100 * vu['valu_busy']:.1f}%",
                  f"  VALU lane utilisation (SQ_THREAD_CYCLES_VALU / (64 * SQ_ACTIVE_INST_VALU)): "
                  f"{100 * vu['valu_lane_util']:.1f}%"]
    if mf:
        t_ms = fig.get("timed_avg_us", k["avg_us"]) / 1e3
        lines += ["", f"Matrix cores of {a.kernel} (--pmc {' '.join(MFMA_COUNTERS)}; means per dispatch over "
                      f"{mf['dispatches']} dispatches):",
                  "", "  " + ", ".join(f"{c} {mf[c]:.4g}" for c in MFMA_COUNTERS),
                  f"  FP32 matrix FLOPs per dispatch (SQ_INSTS_VALU_MFMA_MOPS_F32 * 512): {mf['flops']:.4g}; over the "
                  f"kernel's mean duration (timed region) {t_ms:.4f} ms: {mf['flops'] / (t_ms * 1e-3) / 1e12:.3f} TFLOP/s of "
                  f"{MFMA_FP32_PEAK_TFLOPS} ({100 * mf['flops'] / (t_ms * 1e-3) / 1e12 / MFMA_FP32_PEAK_TFLOPS:.3f}%)",
                  f"  matrix cores busy (SQ_VALU_MFMA_BUSY_CYCLES / (SIMDs * GRBM_GUI_ACTIVE / XCDs)): "
                  f"{100 * mf['mfma_busy']:.2f}%"]
    if bench:
        alg = fig["algorithmic_bytes_per_launch"]
        lines += ["", f"Algorithmic bytes per launch: {rf['bytes_per_env_step']:.1f} B/env-step x "
                      f"{bench['config']['n_envs']} envs = {alg / 1e6:.3f} MB; traffic / algorithmic = "
                      f"{tr['bytes_per_launch'] / alg:.2f}",
                  f"bench.py on the traced run: kernel_ms {rf['kernel_ms']:.4f} (HIP events on the launch stream over "
                  f"its timed region) vs rocprof {fig['timed_avg_us'] / 1e3:.4f} ms (mean of the same "
                  f"{fig['timed_dispatches']} dispatches, after the {int(bench['warmup'])} warm-up ones, from the "
                  f"per-dispatch trace) and {k['avg_us'] / 1e3:.4f} ms (all {k['calls']} dispatches, --stats)"]
    lines += ["", f"source digest (kernel sources + build flags): {fig['source_digest']}"]
    with open(a.out + ".md", "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


if __name__ == "__main__":
    main()
