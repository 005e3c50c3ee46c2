"""Summaries of rocprofv3 output for the k-NN build (the text kept under profiles/).

    python -m rinnd_amd.profiling stats <run_kernel_stats.csv>
    python -m rinnd_amd.profiling traffic <fetch run_counter_collection.csv> <write run_counter_collection.csv>
    python -m rinnd_amd.profiling counters <run_counter_collection.csv>

`counters` prints, per timed kernel family, the mean of every collected counter per launch, and for the MFMA
counters the MFMA flops per launch (SQ_INSTS_VALU_MFMA_MOPS_* count units of 512 flops); divided by the mean
launch duration of `stats` they give the MFMA rate against the dense peak.

`stats` prints the kernels with at least 0.1 % of the kernel time.  `traffic` prints, per timed kernel
family, the mean HBM bytes per launch: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what the
same counters read for the largest plain device-to-device copy in the run (bench.py --copy-calibration
copies 1 GiB), i.e. bytes = counter x 1024 x (copy bytes / (counter on the copy x 1024)).
"""
from __future__ import annotations

import csv
import re
import sys

# timed kernel family -> (name used by bench.py, regex over the dispatched kernel names)
FAMILIES = {
    "margins": ("k_lvl_margins", r"k_lvl_margins"),
    "leaf_dist_sq": ("k_leaf_dist_sq", r"k_leaf_dist_sq"),
    "leaf_dist_mfma": ("k_leaf_dist_mfma/k_leaf_surv_mfma", r"k_leaf_dist_mfma|k_leaf_surv_mfma"),
    "leaf_push": ("k_leaf_push", r"k_leaf_push"),  # k_leaf_push<16|32> and k_leaf_push_surv
    "join_fp32": ("k_join_fp32", r"k_join_fp32"),
    "join_bf16": ("k_join_bf16", r"k_join_bf16"),
}
COPY = r"__amd_rocclr_copyBuffer"
CALIBRATION_BYTES = 1 << 30  # the copy bench.py --copy-calibration makes


def short_name(name: str) -> str:
    return re.sub(r"^void ", "", re.sub(r"\(.*", "", name))


def kernel_stats(path: str, min_pct: float = 0.1) -> list[str]:
    rows = list(csv.DictReader(open(path)))
    out = [f"{'kernel':60s} {'calls':>6s} {'total_ms':>10s} {'avg_us':>10s} {'pct':>6s}"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        pct = float(r["Percentage"])
        if pct < min_pct:
            continue
        calls = int(r["Calls"])
        tot = float(r["TotalDurationNs"]) / 1e6
        out.append(f"{short_name(r['Name'])[:60]:60s} {calls:6d} {tot:10.3f} {1e3 * tot / calls:10.1f} {pct:6.2f}")
    return out


def _per_dispatch(path: str, counter: str) -> list[tuple[str, float]]:
    acc: dict[int, list] = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        d = acc.setdefault(int(r["Dispatch_Id"]), [r["Kernel_Name"], 0.0])
        d[1] += float(r["Counter_Value"])
    return [tuple(v) for _, v in sorted(acc.items())]


def traffic(fetch_csv: str, write_csv: str) -> list[str]:
    fetch = _per_dispatch(fetch_csv, "FETCH_SIZE")
    write = _per_dispatch(write_csv, "WRITE_SIZE")
    cf = max((v for n, v in fetch if re.search(COPY, n)), default=0.0)
    cw = max((v for n, v in write if re.search(COPY, n)), default=0.0)
    if cf <= 0 or cw <= 0:
        raise SystemExit("no copy dispatch to calibrate against (run bench.py with --copy-calibration)")
    sf, sw = CALIBRATION_BYTES / (cf * 1024), CALIBRATION_BYTES / (cw * 1024)
    out = [f"# calibration: the {CALIBRATION_BYTES >> 30} GiB copy reads FETCH_SIZE {cf:.0f} KiB and WRITE_SIZE "
           f"{cw:.0f} KiB, so",
           f"#   fetch bytes = FETCH_SIZE x 1024 x {sf:.3f} and write bytes = WRITE_SIZE x 1024 x {sw:.3f}"]
    for fam, (name, pat) in FAMILIES.items():
        fv = [v for n, v in fetch if re.search(pat, n)]
        wv = [v for n, v in write if re.search(pat, n)]
        if not fv or not wv:
            continue
        fb = sum(fv) / len(fv) * 1024 * sf
        wb = sum(wv) / len(wv) * 1024 * sw
        out.append(f"traffic kernel={name} launches={len(fv)} fetch_bytes={fb:.4e} write_bytes={wb:.4e} "
                   f"traffic_bytes={fb + wb:.4e}")
    return out


def counters(path: str) -> list[str]:
    per: dict[int, dict] = {}
    for r in csv.DictReader(open(path)):
        d = per.setdefault(int(r["Dispatch_Id"]), {"name": r["Kernel_Name"]})
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    fams = dict(FAMILIES)
    fams["search"] = ("k_search", r"k_search")
    out = []
    for fam, (name, pat) in fams.items():
        ds = [d for d in per.values() if re.search(pat, d["name"])]
        if not ds:
            continue
        keys = sorted({k for d in ds for k in d if k != "name"})
        mean = {k: sum(d.get(k, 0.0) for d in ds) / len(ds) for k in keys}
        line = f"counters kernel={name} launches={len(ds)} " + " ".join(f"{k}={v:.4e}" for k, v in mean.items())
        for kind in ("BF16", "F8"):
            mops = mean.get(f"SQ_INSTS_VALU_MFMA_MOPS_{kind}")
            if mops:
                line += f" mfma_{kind.lower()}_flops={mops * 512:.4e}"
        out.append(line)
    return out


def main(argv: list[str]) -> int:
    if len(argv) == 2 and argv[0] == "counters":
        print("\n".join(counters(argv[1])))
        return 0
    if len(argv) == 2 and argv[0] == "stats":
        print("\n".join(kernel_stats(argv[1])))
        return 0
    if len(argv) == 3 and argv[0] == "traffic":
        print("\n".join(traffic(argv[1], argv[2])))
        return 0
    print(__doc__)
    return 2


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
