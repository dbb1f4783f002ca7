"""Benchmark of the ray-acceleration and radiance path on one MI355X.

Workload (one step): the 65,536-triangle soup + emissive quad (65,538 shapes, seed 1), rebuilt as an
LBVH from device-resident shapes (bounds + Morton, onesweep sort, Karras topology, refit) and
rendered at 512x512, 8 spp, MAX_DEPTH 5 (6 path segments) with BVH closest-hit and any-hit shadow
queries.  `value` is Mrays/s = (closest-hit + any-hit queries) / wall time of the timed steps,
build included.  The line also carries `served`: a Poisson request trace served with continuous
batching against the plain one-scene-per-launch path (frames/s, p50/p99 latency, Mrays/s), and
`batched`: 64 scenes built in one batched launch sequence and rendered in one launch against the
plain per-scene build + render.
Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-serve]
  python bench.py --serve            # only the served comparison
  python bench.py --batch            # only the batched (64-scene) comparison
  python bench.py --c3               # only the 327K-triangle icosphere, LBVH vs KD trees
  python bench.py --c1               # only the Cornell spheres frame (no acceleration)
  python bench.py --build-sweep 16,20,23
"""
from __future__ import annotations

import argparse
import json
import re
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
W = H = 512
SPP = 8
DEPTH = 5
N_TRI = 65536


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(shapes, lc, cam, btn, btl, budget_s=12.0, w=None, h=None, depth=None, what=None):
    """The reference's own host path (geomfunc.h RadiancePathTracing, BVH mode, brute-force shadow
    rays), built from the reference sources into oracle/_ref/, timed on this host's cores on a
    pixel subset; its query count comes from the oracle in host-twin semantics, which follows the
    same paths bit for bit.  Falls back to the oracle itself when oracle/_ref was not built."""
    rays, elapsed, total_pix, kind, threads = cpu_render_rate(shapes, lc, cam, btn, btl, budget_s, w, h, depth)
    W, H = (w or globals()["W"]), (h or globals()["H"])
    what = what or f"the same {W}x{H} frame"
    return {"value": rays / elapsed / 1e6, "unit": "Mrays/s", "cores": threads, "kind": kind,
            "sample": f"{total_pix} pixels x 1 spp of {what} (BVH closest hit, brute-force shadow rays over "
                      f"{len(shapes):,} shapes), {elapsed:.1f} s"}


def cpu_render_rate(shapes, lc, cam, btn, btl, budget_s, w=None, h=None, depth=None):
    """(queries, seconds, pixels, kind, threads) of the host render path on pixel samples of one frame."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O
    from ptgpu_pc_amd.layout import default_seeds
    W, H, DEPTH = (w or globals()["W"]), (h or globals()["H"]), (depth or globals()["DEPTH"])
    threads = min(os.cpu_count() or 1, 16)
    ref = O.ref_fn("ref_render", 1, DEPTH)
    kind = "reference" if ref is not None else "port"
    rng = np.random.default_rng(7)
    npix = 64
    elapsed = 0.0
    total_pix = 0
    rays = 0
    t_all = time.perf_counter()
    while elapsed < budget_s and time.perf_counter() - t_all < 3 * budget_s:
        pix = np.sort(rng.choice(W * H, npix, replace=False)).astype(np.int32)
        colors = np.zeros(W * H * 3, np.float32)
        seeds = default_seeds(W, H)
        t0 = time.perf_counter()
        if ref is not None:
            ref(colors.ctypes.data, seeds.ctypes.data, cam.ctypes.data, shapes.ctypes.data, len(shapes), lc,
                btn.ctypes.data, btl.ctypes.data, 0, 0, W, H, 0, 1, pix.ctypes.data, len(pix), threads)
        else:
            O.radiance_frame(colors, seeds, cam, shapes, lc, O.accel_bvh(btn, btl, quirks=1), W, H, 0, 1, None,
                             DEPTH, sem=1, pix_list=pix, threads=threads)
        elapsed += time.perf_counter() - t0
        c2 = np.zeros(W * H * 3, np.float32)
        s2 = default_seeds(W, H)
        q = O.radiance_frame(c2, s2, cam, shapes, lc, O.accel_bvh(btn, btl, quirks=1), W, H, 0, 1, None, DEPTH,
                             sem=1, pix_list=pix, threads=threads)
        rays += q[0] + q[1]
        total_pix += npix
        npix = min(npix * 2, 4096)
    return rays, elapsed, total_pix, kind, threads


RENDER_KERNEL = "k_render_lanes"
# the timed instance: single frame, no statistics (<stack, STATS, WIDE, MULTI, HALF>)
RENDER_INSTANCE = "k_render_lanes<%, false, %, false, %>"
WORKLOAD = ("LBVH build + BVH traversal: 65,536-triangle soup (seed 1) + emissive quad, 512x512, 8 spp, depth 5, "
            "30-bit Morton, one leaf per primitive")
TRAFFIC_FILE = ROOT / "profiles" / "k_render_lanes_c2_traffic.json"
CALIB_BYTES = 256 << 20      # the calibration copy reads and writes this many bytes


def _db_rows(db, sql):
    import sqlite3
    con = sqlite3.connect(str(db))
    try:
        return list(con.execute(sql))
    finally:
        con.close()


def _counter(db, name, kernel_like):
    import sqlite3
    con = sqlite3.connect(str(db))
    try:
        r = con.execute("select avg(value), count(*) from counters_collection where counter_name = ? and kernel_name "
                        "like ?", (name, f"%{kernel_like}%")).fetchone()
    finally:
        con.close()
    return r


def profile_summary(d: Path, out_md: Path) -> None:
    """Summarise rocprofv3 results of the default workload into profiles/: kernel-trace statistics,
    FETCH_SIZE / WRITE_SIZE calibrated against the known-size copy kernel, SQ/TCC ratios."""
    d = Path(d)
    kt = next(d.glob("kt/**/*results.db"))
    rows = _db_rows(kt, "select name, total_calls, average, percentage from top_kernels order by total_duration desc "
                        "limit 12")
    fetch_db, write_db = next(d.glob("fetch/**/*results.db")), next(d.glob("write/**/*results.db"))
    f_k, f_n = _counter(fetch_db, "FETCH_SIZE", RENDER_INSTANCE)
    w_k, w_n = _counter(write_db, "WRITE_SIZE", RENDER_INSTANCE)
    f_c = _counter(fetch_db, "FETCH_SIZE", "k_copy_calib")[0]
    w_c = _counter(write_db, "WRITE_SIZE", "k_copy_calib")[0]
    f_scale = CALIB_BYTES / (f_c * 1024.0)     # bytes per reported KiB, from the known copy
    w_scale = CALIB_BYTES / (w_c * 1024.0)
    traffic = f_k * 1024.0 * f_scale + w_k * 1024.0 * w_scale
    sq = {}
    for dbp in d.glob("sq*/**/*results.db"):
        for name, val in _db_rows(dbp, "select counter_name, avg(value) from counters_collection where kernel_name "
                                       f"like '%{RENDER_INSTANCE}%' group by counter_name"):
            sq[name] = val
    timed = re.compile(r"k_render_lanes<\d+, false, (true|false), false, (true|false)>")
    render_avg = next((r[2] / 1e3 for r in rows if timed.search(r[0])), None)
    cmds = json.loads((d / "commands.json").read_text())
    lines = [f"# {RENDER_KERNEL} — C2 (65,538-shape soup, 512x512, 8 spp, depth 5), MI355X", "",
             "Commands (run from the repository root; rocprofv3 output directories as named):", ""]
    lines += [f"    {c}" for c in cmds]
    lines += ["", "## Kernel trace (rocprofv3 --kernel-trace --stats)", "",
              "| kernel | calls | average ms | % of GPU time |", "|---|---|---|---|"]
    lines += [f"| `{r[0][:90]}` | {r[1]} | {r[2] / 1e3:.4f} | {r[3]:.2f} |" for r in rows]   # rocpd: µs
    lines += ["", "The timed frames run three in flight: each step renders one frame (its tree built during "
              "the previous step, on a third stream) while the previous frame's render drains on the other "
              "render stream. Kernels of consecutive frames overlap, so a render launch's duration spans more "
              "than one step, build kernels' durations are spans beside a render, and summed kernel time "
              "exceeds the wall time."]
    lines += ["", "## HBM traffic per render launch", "",
              f"* calibration: `k_copy_calib` reads and writes {CALIB_BYTES >> 20} MiB; FETCH_SIZE reports {f_c:.0f} KiB "
              f"(true/reported = {f_scale:.4f}), WRITE_SIZE reports {w_c:.0f} KiB (true/reported = {w_scale:.4f})",
              f"* {RENDER_KERNEL}: FETCH_SIZE {f_k:.0f} KiB ({f_n} dispatches), WRITE_SIZE {w_k:.0f} KiB",
              f"* calibrated traffic: {traffic / 1e9:.3f} GB per launch "
              f"({traffic / 1e9 / render_avg * 1e3 if render_avg else 0:.0f} GB/s at the traced average duration)"]
    if sq:
        lines += ["", "## SQ / TCC ratios", ""]
        if "SQ_THREAD_CYCLES_VALU" in sq and "SQ_ACTIVE_INST_VALU" in sq:
            lines.append(f"* VALU lane utilisation: {sq['SQ_THREAD_CYCLES_VALU'] / (sq['SQ_ACTIVE_INST_VALU'] * 64):.3f}")
        if "SQ_ACTIVE_INST_VALU" in sq and "GRBM_GUI_ACTIVE" in sq:
            simds, xcds = 1024, 8
            busy = sq["SQ_ACTIVE_INST_VALU"] * 4 / (simds * sq["GRBM_GUI_ACTIVE"] / xcds)
            lines.append(f"* VALU busy (quad-cycles x4 / (1024 SIMDs x GPU cycles)): {busy:.3f}")
        if "TCC_HIT_sum" in sq:
            lines.append(f"* L2 hit rate: {sq['TCC_HIT_sum'] / (sq['TCC_HIT_sum'] + sq['TCC_MISS_sum']):.4f}")
        for k in sorted(sq):
            lines.append(f"* {k}: {sq[k]:.4g}")
    out_md.write_text("\n".join(lines) + "\n")
    ktb = next(d.glob("ktb/**/*results.db"), None)
    if ktb is not None:
        batch_build_summary(ktb, cmds, ROOT / "profiles" / "lbvh_batch_build_64scenes.md")
    ktc = next(d.glob("ktc/**/*results.db"), None)
    if ktc is not None and (d / "c3.json").exists():
        c3_summary(ktc, json.loads((d / "c3.json").read_text().strip().splitlines()[-1]), cmds,
                   ROOT / "profiles" / "c3_icosphere_1024x768x16spp.md")
    ratios = {}
    if "SQ_THREAD_CYCLES_VALU" in sq and "SQ_ACTIVE_INST_VALU" in sq:
        ratios["valu_lane_util"] = sq["SQ_THREAD_CYCLES_VALU"] / (sq["SQ_ACTIVE_INST_VALU"] * 64)
    if "SQ_ACTIVE_INST_VALU" in sq and "GRBM_GUI_ACTIVE" in sq:
        ratios["valu_busy"] = sq["SQ_ACTIVE_INST_VALU"] * 4 / (1024 * sq["GRBM_GUI_ACTIVE"] / 8)
    if "TCC_HIT_sum" in sq:
        ratios["l2_hit"] = sq["TCC_HIT_sum"] / (sq["TCC_HIT_sum"] + sq["TCC_MISS_sum"])
    TRAFFIC_FILE.write_text(json.dumps({"kernel": RENDER_KERNEL, "workload": WORKLOAD, "traffic_bytes": traffic,
                                        "fetch_kib": f_k, "write_kib": w_k, "fetch_factor": f_scale,
                                        "write_factor": w_scale, "avg_ms_rocprof": render_avg,
                                        **{k: round(v, 4) for k, v in ratios.items()},
                                        "summary": out_md.name}, indent=1) + "\n")
    print(out_md.read_text())


# stage of each batched-build kernel, for the per-stage algorithmic GB/s of the 64-scene build
BATCH_STAGE = {"kb_fill": "sort", "kb_bounds": "morton", "kb_bounds_final": "morton", "kb_morton": "morton",
               "kb_karras": "refit", "kb_sort_hist": "sort", "kb_sort_scan": "sort", "kb_sort_pass": "sort",
               "kb_refit": "refit", "kb_karras_morton": "refit", "kb_mrefit_chunk": "refit", "kb_mrefit_mid": "refit",
               "kb_mrefit_top": "refit",
               "kb_gather": "refit"}


def batch_build_summary(db: Path, cmds: list, out_md: Path) -> None:
    """Per-kernel and per-stage times of the batched 64-scene build (bench.py --batch under rocprofv3)."""
    rows = _db_rows(db, "select name, total_calls, average from top_kernels where name like '%kb_%'")
    n_total = sum(len(w["shapes"]) for w in batch_workload())
    steps = None
    per_stage, table = {}, []
    for name, calls, avg_us in rows:
        m = re.search(r"(kb_\w+)(<[^>(]*>)?", name)
        short = m.group(1) + (m.group(2) or "")
        if steps is None:
            steps = _db_rows(db, "select total_calls from top_kernels where name like '%kb_gather%'")[0][0]
        ms = avg_us / 1e3 * calls / steps          # per batched build
        stage = BATCH_STAGE[m.group(1)]
        per_stage[stage] = per_stage.get(stage, 0.0) + ms
        table.append((short, calls // steps, avg_us / 1e3, ms))
    total = sum(per_stage.values())
    lines = ["# Batched LBVH build — 64 scenes, one launch per stage, MI355X", "",
             f"Workload: the 64-scene set of `bench.py --batch` ({n_total:,} shapes; triangles log-uniform in "
             "[1e3, 1e6], soup or displaced icosphere by scene seed), every stage of every scene's build in one "
             "launch.", "", "Command (run from the repository root):", "",
             "    " + next(c for c in cmds if "--batch" in c), "",
             "| kernel | launches per build | average ms | ms per build |", "|---|---|---|---|"]
    for short, k, avg, ms in sorted(table, key=lambda r: -r[3]):
        lines.append(f"| `{short}` | {k} | {avg:.4f} | {ms:.4f} |")
    lines += ["", f"Total kernel time per batched build: {total:.3f} ms = {n_total / total / 1e3:.0f} Mtri/s.", "",
              "| stage | ms | algorithmic B/prim | achieved GB/s | fraction of 8 TB/s |", "|---|---|---|---|---|"]
    for st in STAGE_BYTES:
        ms = per_stage.get(st, 0.0)
        gbs = STAGE_BYTES[st] * n_total / (ms * 1e-3) / 1e9 if ms else 0.0
        lines.append(f"| {st} | {ms:.3f} | {STAGE_BYTES[st]} | {gbs:.0f} | {gbs / HBM_PEAK_GBS:.3f} |")
    out_md.write_text("\n".join(lines) + "\n")
    print(out_md.read_text())


def c3_summary(db: Path, c3: dict, cmds: list, out_md: Path) -> None:
    """The 327K-triangle icosphere comparison (bench.py --c3) with its kernel trace."""
    rows = _db_rows(db, "select name, total_calls, average, percentage from top_kernels order by total_duration desc "
                        "limit 10")
    lines = ["# C3 — 327,680-triangle icosphere in a box, 1024x768, 16 spp, depth 6, MI355X", "",
             "Commands (run from the repository root):", ""]
    lines += ["    " + c for c in cmds if "--c3" in c]
    lines += ["", "| tree | build | render ms | Mrays/s | same pixels |", "|---|---|---|---|---|"]
    for name, t in c3["trees"].items():
        build = f"{t['build_ms']} ms (GPU)" if "build_ms" in t else f"{t['build_ms_host']} ms (host)"
        lines.append(f"| {name} | {build} | {t['render_ms']} | {t['mrays_s']:,} | {t['pixels_identical']} |")
    lines += ["", f"Autotuner choice for one frame per build: `{json.dumps(c3.get('tuning'))}`", "",
              "## Kernel trace (rocprofv3 --kernel-trace --stats)", "", "| kernel | calls | average ms | % of GPU time |",
              "|---|---|---|---|"]
    for name, calls, avg, pct in rows:
        lines.append(f"| `{name[:90]}` | {calls} | {avg / 1e3:.4f} | {pct:.2f} |")
    out_md.write_text("\n".join(lines) + "\n")
    print(out_md.read_text())


def measured_counters():
    """Counter figures of the render kernel from the committed rocprofv3 summary of this workload (or
    None): HBM bytes per launch (calibrated FETCH_SIZE + WRITE_SIZE), the profile's average duration, L2
    hit rate, VALU busy fraction and VALU lane utilisation."""
    try:
        t = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    if t.get("kernel") != RENDER_KERNEL or t.get("workload") != WORKLOAD:
        return None
    return t


def measured_traffic():
    """HBM bytes per render launch from the committed rocprofv3 summary of this workload (or None)."""
    t = measured_counters()
    return int(t["traffic_bytes"]) if t else None


# algorithmic bytes per primitive of each build stage (each logical load/store counted once at its
# natural width, atomics 8 B): Morton 36+36+4, 8-bit onesweep 4 + 16*4, refit 156 + the topology's 28
# (the refit writes whole reference records, Karras fields included; the topology is a closed form
# of N, so it has no pass of its own)
STAGE_BYTES = {"morton": 76, "sort": 68, "refit": 156 + 28}
BUILD_BYTES_PER_PRIM = sum(STAGE_BYTES.values())   # 328 B per primitive with 8-bit digits
# the Morton-keyed traversal tree, per primitive: Karras over sorted keys (reads 4 B per key,
# writes a 16-B node and 8 B of parents) + refit (reads a 24-B leaf bound, the 16-B node and a 4-B
# parent, writes a 64-B traversal node)
TRAV_TREE_BYTES = 4 + 16 + 8 + 24 + 16 + 4 + 64


def render_roofline(alg_bytes, avg_ms, box_tests, prim_tests, box_bytes, measured=True, kernel=None, step_ms=None):
    """Roofline object of a render launch.  The path-tracing kernel is bound neither by HBM nor by VALU:
    its time is the longest per-pixel chain of dependent node fetches (samples of a pixel run in order,
    each query a chain of node visits), so `bound` says "latency"; `achieved` / `frac` price the
    algorithmic bytes against the HBM peak as the contract asks, and `measured` gives what the counters
    say the kernel actually moves (calibrated HBM traffic, L2 hit rate, VALU busy and lane use)."""
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    out = {"bound": "latency", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
           "kernel": kernel or f"{RENDER_KERNEL} (path-tracing megakernel, per-lane states)",
           "limiter": "per-pixel chains of dependent node fetches (the samples of a pixel run in order)",
           "per_launch": {"box_tests": box_tests, "prim_tests": prim_tests, "bytes_per_box_test": box_bytes,
                          "alg_bytes": alg_bytes, "avg_ms": round(avg_ms, 4)}}
    if step_ms:
        # launches overlap (a frame's render starts while the previous one drains), so a launch's span
        # is longer than the step: the same bytes over the step time give the device-level rate
        agg = alg_bytes / (step_ms * 1e-3) / 1e9
        out["per_step"] = {"achieved": round(agg, 1), "frac": round(agg / HBM_PEAK_GBS, 4), "step_ms": round(step_ms, 4),
                           "note": "algorithmic bytes of one launch over the step time; launches of consecutive "
                                   "frames overlap, so avg_ms spans more than one step"}
    t = measured_counters() if measured else None
    if t:
        out["traffic"] = int(t["traffic_bytes"])
        hbm = t["traffic_bytes"] / (t["avg_ms_rocprof"] * 1e-3) / 1e9
        out["measured"] = {"hbm_gbs": round(hbm, 1), "hbm_frac": round(hbm / HBM_PEAK_GBS, 4),
                           **{k: t[k] for k in ("l2_hit", "valu_busy", "valu_lane_util") if k in t},
                           "source": f"profiles/{t['summary']}"}
    return out


def cpu_build_baseline(shapes):
    """The reference's host-side LBVH construction as the oracle restates it, single-threaded: host
    Morton codes (CLBVH.cpp:90-164), the literal stable descending bubble sort (CLBVH.cpp:31-47),
    Karras topology over leaf indices (BVH.cl:141-208) and the bottom-up refit (BVH.cl:214-258)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O
    t0 = time.perf_counter()
    O.build_lbvh(shapes, literal_sort=True)
    sec = time.perf_counter() - t0
    return {"value": round(len(shapes) / sec / 1e6, 4), "unit": "Mtri/s", "cores": 1, "kind": "port",
            "sample": f"one whole LBVH build of the {len(shapes):,} shapes with the literal O(N^2) bubble sort, "
                      f"{sec:.2f} s"}


def c1_compare(steps: int = 10, warmup: int = 2, cpu: bool = True):
    """Cornell spheres (9 shapes, no acceleration), 256x256, 4 spp, MAX_DEPTH 5: the GPU frame (one
    fused 4-spp launch) against the reference's own host path (geomfunc.h, oracle/_ref) on the whole
    frame on this host's cores."""
    import ptgpu_pc_amd as pt
    from ptgpu_pc_amd import scenes
    from ptgpu_pc_amd.layout import default_seeds
    w = h = 256
    spp, depth = 4, 5
    shapes, lc, cam = scenes.cornell(w, h)
    pt.update_camera(cam)
    ctx = pt.Context(0)
    d_sh, d_cam = ctx.to_device(shapes), ctx.to_device(cam)
    seeds0 = ctx.to_device(default_seeds(w, h))
    d_seeds = ctx.alloc(w * h * 8, np.uint32)
    d_col = ctx.alloc(w * h * 12, np.float32)
    d_pix = ctx.alloc(w * h * 4, np.uint32)
    d_rays = ctx.alloc(16, np.uint64).zero()
    d_stats = ctx.alloc(16, np.uint64).zero()

    def frame(rays=None, stats=None):
        d_seeds.copy_from(seeds0)
        pt.render(ctx, None, None, colors=d_col, seeds=d_seeds, camera=d_cam, shapes=d_sh, shapeCnt=len(shapes),
                  lightCnt=lc, width=w, height=h, currentSample=0, spp=spp, max_depth=depth, pixels=d_pix,
                  ray_counter=rays, trav_stats=stats)
    for _ in range(warmup):
        frame()
    frame(stats=d_stats)
    ctx.synchronize()
    _, sphere_tests = (int(x) for x in d_stats.download(np.uint64)[:2])
    ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.record(0)
        frame(rays=d_rays)
        ctx.record(1)
        ms.append(ctx.elapsed_ms(0, 1))
    ctx.synchronize()
    wall = time.perf_counter() - t0
    rays = int(d_rays.download(np.uint64)[:2].sum()) // steps
    avg = float(np.mean(ms))
    # algorithmic bytes: 16 B per sphere test (radius and centre) and the frame buffers
    alg = 16 * sphere_tests + w * h * (24 + 16 + 4)
    out = {"workload": "Cornell spheres (9 shapes, no acceleration), 256x256, 4 spp, MAX_DEPTH 5",
           "mrays_s": round(rays * steps / wall / 1e6, 1), "frame_ms": round(wall * 1e3 / steps, 3),
           "render_ms": round(avg, 4), "rays_per_frame": rays,
           "roofline": {"bound": "latency", "achieved": round(alg / (avg * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(alg / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                        "kernel": "k_render (lockstep megakernel; brute force over the 9 spheres)",
                        "limiter": "per-pixel sample chains: 65,536 pixels fill a quarter of the chip's lanes",
                        "per_launch": {"sphere_tests": sphere_tests, "alg_bytes": alg}}}
    ctx.close()
    if cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle as O
        threads = min(os.cpu_count() or 1, 16)
        ref = O.ref_fn("ref_render", 0, depth)
        colors = np.zeros(w * h * 3, np.float32)
        seeds = default_seeds(w, h)
        t0 = time.perf_counter()
        if ref is not None:
            ref(colors.ctypes.data, seeds.ctypes.data, cam.ctypes.data, shapes.ctypes.data, len(shapes), lc, None,
                None, 0, 0, w, h, 0, spp, None, 0, threads)
        else:
            O.radiance_frame(colors, seeds, cam, shapes, lc, O.accel_none(), w, h, 0, spp, None, depth, sem=1,
                             threads=threads)
        sec = time.perf_counter() - t0
        q = O.radiance_frame(np.zeros(w * h * 3, np.float32), default_seeds(w, h), cam, shapes, lc, O.accel_none(),
                             w, h, 0, spp, None, depth, sem=1, threads=threads)
        out["cpu_baseline"] = {"value": round((q[0] + q[1]) / sec / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                               "kind": "reference" if ref is not None else "port",
                               "sample": f"the whole 256x256 x 4 spp frame, {sec:.2f} s"}
    return out


def build_sweep(sizes, reps=5):
    """LBVH build throughput and achieved algorithmic GB/s per stage on large soups (working sets
    beyond the 256 MB MALL from about 2^21 primitives)."""
    import ptgpu_pc_amd as pt
    rng = np.random.default_rng(11)
    ctx = pt.Context(0)
    # one copy of known size, so that counter runs of this sweep can be calibrated
    c_src, c_dst = ctx.alloc(CALIB_BYTES, np.uint8), ctx.alloc(CALIB_BYTES, np.uint8)
    ctx.copy_calibration(c_dst, c_src, CALIB_BYTES)
    ctx.synchronize()
    c_src.free()
    c_dst.free()
    for n in sizes:
        cen = rng.random((n, 3), dtype=np.float32)
        from ptgpu_pc_amd.layout import SHAPE_DTYPE, TRIANGLE
        sh = np.zeros(n, dtype=SHAPE_DTYPE)
        sh["type"] = TRIANGLE
        for k in ("p1", "p2", "p3"):
            sh[k] = cen + rng.uniform(-0.01, 0.01, size=(n, 3)).astype(np.float32)
        d_sh = ctx.to_device(sh)
        del sh
        bvh = pt.CLBVH(ctx, d_sh, n)
        best = None
        for _ in range(reps):
            bvh.reset()
            bvh.buildRadixTree()
            bvh.buildBVHTree()
            st = bvh.stage_times6()
            tot = st["morton"] + st["sort"] + st["topology"] + st["refit"] + st["traversal_tree"]   # topology ~0
            if best is None or tot < best[0]:
                best = (tot, st)
        tot, st = best
        row = {"n": n, "build_ms": round(tot, 4), "mtri_s": round(n / tot / 1e3, 1),
               "working_set_mb": round(n * (104 + 120 + 64 + 48 + 16) / 2 ** 20, 1)}
        for k, b in list(STAGE_BYTES.items()) + [("traversal_tree", TRAV_TREE_BYTES)]:
            row[f"{k}_ms"] = round(st[k], 4)
            row[f"{k}_gbs"] = round(b * n / (st[k] * 1e-3) / 1e9, 1) if st[k] > 0 else None
        print(json.dumps(row), flush=True)
        bvh.close()
        d_sh.free()
    ctx.close()


def serve_trace(n_req: int, rate: float, pool: int, seed: int = 2024):
    from ptgpu_pc_amd import scenes
    return scenes.serve_trace(n_req, rate, pool, seed)


def run_plain(ctx, pool_scenes, reqs):
    """One scene per launch, in arrival order: upload + LBVH build + one fused spp render + pixel readback."""
    import ptgpu_pc_amd as pt
    from ptgpu_pc_amd import scenes
    from ptgpu_pc_amd.layout import default_seeds
    d_rays = ctx.alloc(16, np.uint64).zero()
    lat, pix_out = [], []
    t0 = time.perf_counter()
    for r in reqs:
        now = time.perf_counter() - t0
        if now < r["arrival"]:
            time.sleep(r["arrival"] - now)
        _, shapes, lc, _ = pool_scenes[r["scene"]]
        w, h = r["w"], r["h"]
        cam = scenes.orbit_camera(r["yaw"], w, h)
        pt.update_camera(cam)
        s = shapes.copy()
        d_sh = ctx.to_device(s)
        bvh = pt.CLBVH(ctx, d_sh, len(s))
        bvh.buildRadixTree()
        bvh.buildBVHTree()
        d_col = ctx.alloc(w * h * 12, np.float32)
        d_seed = ctx.to_device(default_seeds(w, h))
        d_pix = ctx.alloc(w * h * 4, np.uint32)
        pt.render(ctx, bvh, None, colors=d_col, seeds=d_seed, camera=ctx.to_device(cam), shapes=d_sh, shapeCnt=len(s),
                  lightCnt=lc, width=w, height=h, currentSample=0, spp=r["spp"], max_depth=5, pixels=d_pix,
                  ray_counter=d_rays)
        px = d_pix.download(np.uint32)
        lat.append(time.perf_counter() - t0 - r["arrival"])
        pix_out.append(px)
        bvh.close()
        for a in (d_sh, d_col, d_seed, d_pix):
            a.free()
    ctx.synchronize()
    end = time.perf_counter() - t0
    rays = int(d_rays.download(np.uint64)[:2].sum())
    d_rays.free()
    return lat, end, rays, pix_out


def run_served(ctx, pool_scenes, reqs, budget, spp_per_tick):
    """Continuous batching: submit on arrival, tick, take finished frames (pixel readback)."""
    import ptgpu_pc_amd as pt
    from ptgpu_pc_amd import scenes
    srv = pt.Server(ctx, budget, spp_per_tick)
    cams = []
    for r in reqs:
        cam = scenes.orbit_camera(r["yaw"], r["w"], r["h"])
        pt.update_camera(cam)
        cams.append(cam)
    lat = [None] * len(reqs)
    pix_out = [None] * len(reqs)
    ids = {}
    nxt, done = 0, 0
    t0 = time.perf_counter()
    while done < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt]["arrival"] <= now:
            r = reqs[nxt]
            ss, shapes, lc, _ = pool_scenes[r["scene"]]
            ids[srv.submit(ss, shapes, lc, cams[nxt], r["w"], r["h"], r["spp"], 5)] = nxt
            nxt += 1
        active, queued = srv.tick()
        for rid in srv.finished():
            k = ids[rid]
            _, px, _ = srv.result(rid, colors=False, pixels=True)
            lat[k] = time.perf_counter() - t0 - reqs[k]["arrival"]
            pix_out[k] = px
            done += 1
        if active == 0 and queued == 0 and nxt < len(reqs):
            wait = reqs[nxt]["arrival"] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
    end = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    return lat, end, st["closest"] + st["anyhit"], pix_out, st


def serve_compare(n_req, rate, pool, spp_per_tick, budget_gb, blocks_per_cu=None, regroup=None):
    import ptgpu_pc_amd as pt
    pool_scenes, reqs = serve_trace(n_req, rate, pool)
    ctx = pt.Context(0)
    if blocks_per_cu:
        ctx.set_tuning("batch_blocks_per_cu", blocks_per_cu)
    if regroup:
        ctx.set_tuning("batch_regroup", regroup)
    # warm the code paths once (first-launch costs are not part of either path)
    warm = [dict(reqs[0], arrival=0.0)]
    run_plain(ctx, pool_scenes, warm)
    run_served(ctx, pool_scenes, warm, budget_gb << 30, spp_per_tick)
    lat_p, end_p, rays_p, pix_p = run_plain(ctx, pool_scenes, reqs)
    lat_s, end_s, rays_s, pix_s, st = run_served(ctx, pool_scenes, reqs, budget_gb << 30, spp_per_tick)
    ctx.close()
    first = reqs[0]["arrival"]
    same = all(a.tobytes() == b.tobytes() for a, b in zip(pix_p, pix_s))

    def summary(lat, end, rays):
        lat = np.array(lat)
        return {"frames_s": round(len(lat) / (end - first), 2), "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 1),
                "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 1), "mrays_s": round(rays / (end - first) / 1e6, 1),
                "makespan_s": round(end - first, 3)}
    plain, served = summary(lat_p, end_p, rays_p), summary(lat_s, end_s, rays_s)
    return {"requests": n_req, "rate_per_s": rate, "scene_pool": pool, "spp_per_tick": spp_per_tick,
            "hbm_budget_gb": budget_gb, "plain": plain, "served": served,
            "frames_speedup": round(served["frames_s"] / plain["frames_s"], 2), "pixels_identical": same,
            "served_builds": st["builds"], "workload": "Poisson trace; scenes log-uniform [1e3, 2e6] triangles "
            "(soup or displaced icosphere by seed), 256^2/512^2/1024^2, spp 1/4/16, depth 5"}


def batch_workload(n_scenes: int = 64, seed: int = 4000):
    from ptgpu_pc_amd import scenes
    return scenes.batch_set(n_scenes, seed)


def batch_cpu_baseline(W, render_scenes: int = 6, budget_s: float = 1.5):
    """The many-scene workload on the host: every scene's LBVH as the oracle restates the reference's
    host construction (CLBVH.cpp:90-164 Morton codes, BVH.cl:141-258 topology and refit) with
    std::stable_sort in place of the O(N^2) bubble sort (CLBVH.cpp:31-47, about 10^12 steps at this
    size), single-threaded, summed over the scenes; and the reference's host render path on a pixel
    sample of every len(W) // render_scenes-th scene, its rays over its seconds summed over them."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O
    build_s, n_total, picked = 0.0, 0, []
    step = max(1, len(W) // render_scenes)
    for k, w in enumerate(W):
        t0 = time.perf_counter()
        s_h, _, btn, btl = O.build_lbvh(w["shapes"])
        build_s += time.perf_counter() - t0
        n_total += len(w["shapes"])
        if k % step == 0 and len(picked) < render_scenes:
            picked.append((s_h, w, btn, btl))
    rays = secs = 0.0
    pix = 0
    kind, cores = "port", 1
    for s_h, w, btn, btl in picked:
        q, sec, npx, kind, cores = cpu_render_rate(s_h, w["lc"], w["cam"], btn, btl, budget_s, w["w"], w["h"], 5)
        rays += q
        secs += sec
        pix += npx
    sizes = ", ".join(f"{len(s_h):,}" for s_h, _, _, _ in picked)
    return {"build": {"value": round(n_total / build_s / 1e6, 3), "unit": "Mtri/s", "cores": 1, "kind": "port",
                      "sample": f"the LBVH of each of the {len(W)} scenes ({n_total:,} shapes) with std::stable_sort "
                                f"for the bubble sort, {build_s:.2f} s"},
            "render": {"value": round(rays / secs / 1e6, 4) if secs else None, "unit": "Mrays/s", "cores": cores,
                       "kind": kind,
                       "sample": f"{pix} pixels x 1 spp over {len(picked)} of the scenes ({sizes} shapes; BVH "
                                 f"closest hit, brute-force shadow rays), {secs:.1f} s"}}


def batch_compare(steps: int = 3, warmup: int = 1, n_scenes: int = 64, tune: bool = True, record: bool = True,
                  cpu: bool = False):
    """Plain path (per scene: LBVH build, then one fused 4-spp render launch) against the batched
    path (every scene's build in one batched sequence of launches, every frame in one render
    launch) on the same resident scenes.  Both rebuild every tree and re-render every frame per step."""
    import ptgpu_pc_amd as pt
    from ptgpu_pc_amd.layout import default_seeds
    W = batch_workload(n_scenes)
    ctx = pt.Context(0)
    total_n = sum(len(w["shapes"]) for w in W)
    d_sh = [ctx.to_device(w["shapes"]) for w in W]
    d_cam = [ctx.to_device(w["cam"]) for w in W]
    seeds0 = [ctx.to_device(default_seeds(w["w"], w["h"])) for w in W]
    frames = [{"colors": ctx.alloc(w["w"] * w["h"] * 12, np.float32), "seeds": ctx.alloc(w["w"] * w["h"] * 8, np.uint32),
               "pixels": ctx.alloc(w["w"] * w["h"] * 4, np.uint32)} for w in W]
    d_rays = ctx.alloc(16, np.uint64).zero()
    plain_h = [pt.CLBVH(ctx, d, len(w["shapes"])) for d, w in zip(d_sh, W)]
    batch_h = pt.CLBVH.create_batch(ctx, d_sh)

    def job(k, h):
        w = W[k]
        return dict(bvh=h, colors=frames[k]["colors"], seeds=frames[k]["seeds"], pixels=frames[k]["pixels"],
                    camera=d_cam[k], shapes=d_sh[k], shapeCnt=len(w["shapes"]), lightCnt=w["lc"], width=w["w"],
                    height=w["h"], currentSample=0, spp=4, max_depth=5)

    def reseed():
        for k in range(len(W)):
            frames[k]["seeds"].copy_from(seeds0[k])

    def plain_step(counter=None):
        for k, h in enumerate(plain_h):
            h.reset()
            h.buildRadixTree()
            h.buildBVHTree()
            pt.render(ctx, h, None, ray_counter=counter, **{a: b for a, b in job(k, h).items() if a != "bvh"})

    def batch_step(counter=None):
        ctx.record(0)
        pt.CLBVH.build_batch(ctx, batch_h)
        ctx.record(1)
        pt.render_batch(ctx, [job(k, h) for k, h in enumerate(batch_h)], ray_counter=counter)
        ctx.record(2)

    # same frames from both paths (from the same seeds)
    reseed()
    plain_step()
    ctx.synchronize()
    pix_plain = [f["pixels"].download(np.uint32) for f in frames]
    reseed()
    batch_step()
    ctx.synchronize()
    same = all(a.tobytes() == f["pixels"].download(np.uint32).tobytes() for a, f in zip(pix_plain, frames))
    tuning = pt.autotune_batch(ctx, [job(k, h) for k, h in enumerate(batch_h)], trials=1) if tune else None

    def timed(step_fn, events=False):
        for _ in range(warmup):
            step_fn()
        ctx.synchronize()
        d_rays.zero()
        ctx.synchronize()
        b_ms, r_ms = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            step_fn(d_rays)
            if events:
                b_ms.append(ctx.elapsed_ms(0, 1))
                r_ms.append(ctx.elapsed_ms(1, 2))
        ctx.synchronize()
        wall = time.perf_counter() - t0
        rays = int(d_rays.download(np.uint64)[:2].sum())
        return wall / steps, rays / steps, (float(np.mean(b_ms)) if b_ms else None), (float(np.mean(r_ms)) if r_ms else None)

    t_plain, rays_plain, _, _ = timed(plain_step)
    t_batch, rays_batch, b_ms, r_ms = timed(batch_step, events=True)
    # the plain path's builds alone
    ctx.synchronize()
    t0 = time.perf_counter()
    for h in plain_h:
        h.reset()
        h.buildRadixTree()
        h.buildBVHTree()
    ctx.synchronize()
    plain_build_ms = (time.perf_counter() - t0) * 1e3
    # the batched step recorded once (batched rebuild + batched render, nothing read back) and replayed
    # (record=False leaves it out: rocprofv3's kernel trace stopped with a host segmentation fault inside
    # hipGraphLaunch of this step, so the profiled runs skip it)
    recorded = None
    if record:
        reseed()
        t_rec, same_rec = record_batch_step(ctx, batch_h, [job(k, h) for k, h in enumerate(batch_h)], frames,
                                            pix_plain, steps, warmup)
        recorded = {"frames_s": round(len(W) / t_rec, 2), "ms_per_step": round(t_rec * 1e3, 2),
                    "pixels_identical": same_rec}
    for h in plain_h + batch_h:
        h.close()
    ctx.close()
    cpu_res = batch_cpu_baseline(W) if cpu else None
    alg = BUILD_BYTES_PER_PRIM * total_n
    nfr = len(W)
    return {"scenes": nfr, "shapes": total_n, "soup_scenes": sum(1 for w in W if w["kind"] == 0),
            "pixels": sum(w["w"] * w["h"] for w in W), "spp": 4, "max_depth": 5,
            "plain": {"frames_s": round(nfr / t_plain, 2), "ms_per_step": round(t_plain * 1e3, 2),
                      "mrays_s": round(rays_plain / t_plain / 1e6, 1), "build_ms": round(plain_build_ms, 2)},
            "batched": {"frames_s": round(nfr / t_batch, 2), "ms_per_step": round(t_batch * 1e3, 2),
                        "mrays_s": round(rays_batch / t_batch / 1e6, 1), "build_ms": round(b_ms, 2),
                        "render_ms": round(r_ms, 2), "build_mtri_s": round(total_n / b_ms / 1e3, 1),
                        "build_alg_gbs": round(alg / (b_ms * 1e-3) / 1e9, 1)},
            "batched_recorded": recorded,
            "frames_speedup": round(t_plain / t_batch, 2), "build_speedup": round(plain_build_ms / b_ms, 2),
            "pixels_identical": same, "tuning": tuning, "cpu_baseline": cpu_res,
            "workload": f"{nfr} scenes, triangles log-uniform [1e3, 1e6] (soup or displaced icosphere by seed), "
                        "orbit cameras, 256^2/512^2/1024^2 by seed, 4 spp, depth 5; every tree rebuilt and every "
                        "frame rendered per step"}


def record_batch_step(ctx, batch_h, jobs, frames, pix_plain, steps, warmup):
    """The batched step (batched rebuild + batched render) recorded into a graph and replayed, first from
    the reset seeds: (seconds per step, the first replay's pixels equal the plain path's)."""
    import ptgpu_pc_amd as pt
    ctx.graph_begin()
    try:
        pt.CLBVH.build_batch(ctx, batch_h)
        pt.render_batch(ctx, jobs)
    finally:
        g = ctx.graph_end()
    g.launch()
    ctx.synchronize()
    same_rec = all(a.tobytes() == f["pixels"].download(np.uint32).tobytes() for a, f in zip(pix_plain, frames))
    for _ in range(warmup):
        g.launch()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.launch()
    ctx.synchronize()
    t_rec = (time.perf_counter() - t0) / steps
    g.close()
    return t_rec, same_rec


def frames_compare(tuning: dict | None, counts=(1, 2, 4), steps: int = 5, warmup: int = 1):
    """The C2 frame rendered one per launch against 2 and 4 frames of it per launch (ptgpu_render_batch).
    Every frame has its own tree, rebuilt each step (LBVH + the tuned treelet pass, one after another
    on one stream), its own buffers and the same seeds, so every frame's pixels must equal the
    one-per-launch frame's.  A single frame is bound by its longest per-pixel chain of samples (the
    RNG state runs through a pixel's samples in order); more frames per launch fill the device
    while those chains run."""
    import ptgpu_pc_amd as pt
    from ptgpu_pc_amd import scenes
    from ptgpu_pc_amd.layout import default_seeds
    shapes, lc, cam = scenes.soup(N_TRI, seed=1, width=W, height=H)
    pt.update_camera(cam)
    n = len(shapes)
    ctx = pt.Context(0)
    tun = tuning or {}
    for k in ("blocks_per_cu", "regroup", "wide", "batch_blocks_per_cu", "batch_regroup"):
        if k in tun:
            ctx.set_tuning(k, int(tun[k]))
    treelet = int(tun.get("treelet") or 0)
    fmax = max(counts)
    d_sh = [ctx.to_device(shapes) for _ in range(fmax)]
    hs = [pt.CLBVH(ctx, d_sh[i], n, digit_bits=int(tun.get("digit_bits", 8))) for i in range(fmax)]
    d_cam = ctx.to_device(cam)
    seeds0 = ctx.to_device(default_seeds(W, H))
    bufs = [dict(colors=ctx.alloc(W * H * 12, np.float32).zero(), seeds=ctx.alloc(W * H * 8, np.uint32),
                 pixels=ctx.alloc(W * H * 4, np.uint32)) for _ in range(fmax)]
    d_rays = ctx.alloc(16, np.uint64).zero()

    def job(i):
        return dict(bvh=hs[i], camera=d_cam, shapes=d_sh[i], shapeCnt=n, lightCnt=lc, width=W, height=H,
                    currentSample=0, spp=SPP, max_depth=DEPTH, **bufs[i])

    def step(f, counter=None):
        for i in range(f):
            hs[i].reset()
            hs[i].buildRadixTree()
            hs[i].buildBVHTree()
            if treelet:
                hs[i].optimize(treelet)
            bufs[i]["seeds"].copy_from(seeds0)
        ctx.record(0)
        if f == 1:
            pt.render(ctx, hs[0], None, ray_counter=counter, **{a: b for a, b in job(0).items() if a != "bvh"})
        else:
            pt.render_batch(ctx, [job(i) for i in range(f)], ray_counter=counter)
        ctx.record(1)

    rows, ref, same = [], None, True
    for f in counts:
        for _ in range(warmup):
            step(f)
        ctx.synchronize()
        pix = [bufs[i]["pixels"].download(np.uint32).tobytes() for i in range(f)]
        if ref is None:
            ref = pix[0]
        same = same and all(p == ref for p in pix)
        d_rays.zero()
        ctx.synchronize()
        r_ms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            step(f, d_rays)
            r_ms.append(ctx.elapsed_ms(0, 1))
        ctx.synchronize()
        wall = time.perf_counter() - t0
        rays = int(d_rays.download(np.uint64)[:2].sum())
        rows.append({"frames_per_launch": f, "render_ms_per_frame": round(float(np.mean(r_ms)) / f, 3),
                     "ms_per_frame": round(wall * 1e3 / (steps * f), 3), "mrays_s": round(rays / wall / 1e6, 1)})
    for h in hs:
        h.close()
    ctx.close()
    return {"workload": WORKLOAD, "treelet": treelet, "rows": rows, "pixels_identical": same,
            "note": "each step rebuilds every frame's tree, then renders the frames in one launch"}


def graph_compare(frames: int = 200, sizes=(128, 256), spp: int = 1):
    """Per-frame sequence of an interactive viewer on the C2 scene (camera upload from pinned memory,
    seed reset, colour clear, render, pixel readback, wait), issued call by call against replayed as
    one recorded HIP graph (ptgpu_graph_*)."""
    import ptgpu_pc_amd as pt
    from ptgpu_pc_amd import scenes
    from ptgpu_pc_amd.layout import CAMERA_DTYPE, default_seeds
    shapes, lc, _ = scenes.soup(N_TRI, seed=1)
    ctx = pt.Context(0)
    d_sh = ctx.to_device(shapes)
    bvh = pt.CLBVH(ctx, d_sh, len(shapes))
    bvh.buildRadixTree()
    bvh.buildBVHTree()
    out = {}
    for res in sizes:
        cams = []
        for k in range(frames):
            c = scenes.orbit_camera(2 * np.pi * k / frames, res, res)
            pt.update_camera(c)
            cams.append(c)
        h_cam = ctx.pinned(CAMERA_DTYPE.itemsize, CAMERA_DTYPE)
        h_pix = ctx.pinned(res * res * 4, np.uint32)
        d_cam = ctx.alloc(CAMERA_DTYPE.itemsize, CAMERA_DTYPE)
        d_seeds0 = ctx.to_device(default_seeds(res, res))
        d_seeds = ctx.alloc(res * res * 8, np.uint32)
        d_col = ctx.alloc(res * res * 12, np.float32)
        d_pix = ctx.alloc(res * res * 4, np.uint32)

        def frame():
            ctx.memcpy_async(d_cam, h_cam, CAMERA_DTYPE.itemsize)
            ctx.memcpy_async(d_seeds, d_seeds0, res * res * 8)
            d_col.zero()
            pt.render(ctx, bvh, None, colors=d_col, seeds=d_seeds, camera=d_cam, shapes=d_sh, shapeCnt=len(shapes),
                      lightCnt=lc, width=res, height=res, currentSample=0, spp=spp, max_depth=DEPTH, pixels=d_pix)
            ctx.memcpy_async(h_pix, d_pix, res * res * 4)

        h_cam.array[:] = cams[0]
        frame()                      # prepares the handle (height, four-wide nodes, light list)
        ctx.synchronize()
        ctx.graph_begin()
        try:
            frame()
        finally:
            g = ctx.graph_end()
        times = {}
        last = {}
        for mode in ("eager", "graph", "eager", "graph"):
            t0 = time.perf_counter()
            for k in range(frames):
                h_cam.array[:] = cams[k]
                if mode == "graph":
                    g.launch()
                else:
                    frame()
                ctx.synchronize()
            times[mode] = min(times.get(mode, 1e9), time.perf_counter() - t0)
            last[mode] = h_pix.array.copy()
        g.close()
        h_cam.close()
        h_pix.close()
        out[f"{res}x{res}"] = {"eager_frames_s": round(frames / times["eager"], 1),
                               "graph_frames_s": round(frames / times["graph"], 1),
                               "speedup": round(times["eager"] / times["graph"], 3),
                               "pixels_identical": last["eager"].tobytes() == last["graph"].tobytes()}
    bvh.close()
    ctx.close()
    return {"frames": frames, "spp": spp, "sizes": out, "c2_step": graph_c2_step(),
            "workload": "C2 scene, orbiting camera, per frame: camera upload, seed reset, colour clear, render, "
                        "pixel readback into pinned memory, wait"}


def graph_c2_step(steps: int = 20, treelet: int = 5):
    """The whole C2 step -- LBVH rebuild from the device-resident shapes, treelet pass, 8-spp 512x512
    render, pixel readback, wait -- issued call by call against replayed as one recorded graph.  Issued
    eagerly, the first traversal after each rebuild reads the tree height back (a wait in mid-step);
    recorded, the render sizes its stack from the last height read plus a margin and checks the
    built height on the device."""
    import ptgpu_pc_amd as pt
    from ptgpu_pc_amd import scenes
    from ptgpu_pc_amd.layout import default_seeds
    shapes, lc, cam = scenes.soup(N_TRI, seed=1)
    pt.update_camera(cam)
    n = len(shapes)
    ctx = pt.Context(0)
    d_sh = ctx.to_device(shapes)
    d_cam = ctx.to_device(cam)
    d_seeds0 = ctx.to_device(default_seeds(W, H))
    d_seeds = ctx.alloc(W * H * 8, np.uint32)
    d_col = ctx.alloc(W * H * 12, np.float32)
    d_pix = ctx.alloc(W * H * 4, np.uint32)
    h_pix = ctx.pinned(W * H * 4, np.uint32)
    bvh = pt.CLBVH(ctx, d_sh, n)

    def step():
        bvh.reset()
        bvh.buildRadixTree()
        bvh.buildBVHTree()
        if treelet:
            bvh.optimize(treelet)
        ctx.memcpy_async(d_seeds, d_seeds0, W * H * 8)
        d_col.zero()
        pt.render(ctx, bvh, None, colors=d_col, seeds=d_seeds, camera=d_cam, shapes=d_sh, shapeCnt=n, lightCnt=lc,
                  width=W, height=H, currentSample=0, spp=SPP, max_depth=DEPTH, pixels=d_pix)
        ctx.memcpy_async(h_pix, d_pix, W * H * 4)

    step()                           # heights read back, buffers sized, light list built
    ctx.synchronize()
    ctx.graph_begin()
    try:
        step()
    finally:
        g = ctx.graph_end()
    times, last = {}, {}
    for mode in ("eager", "graph", "eager", "graph"):
        t0 = time.perf_counter()
        for _ in range(steps):
            if mode == "graph":
                g.launch()
            else:
                step()
            ctx.synchronize()
        times[mode] = min(times.get(mode, 1e9), time.perf_counter() - t0)
        last[mode] = h_pix.array.copy()
    g.close()
    h_pix.close()
    bvh.close()
    ctx.close()
    return {"steps": steps, "treelet": treelet,
            "eager_ms_per_step": round(1e3 * times["eager"] / steps, 3),
            "graph_ms_per_step": round(1e3 * times["graph"] / steps, 3),
            "speedup": round(times["eager"] / times["graph"], 3),
            "pixels_identical": last["eager"].tobytes() == last["graph"].tobytes(),
            "workload": f"C2 step: LBVH rebuild + treelet {treelet} + {SPP}-spp {W}x{H} render + pixel readback, "
                        "one stream, one step in flight"}


def c3_compare(steps: int = 3, tune: bool = True, cpu: bool = False):
    """The 327,680-triangle displaced icosphere (seed 3) in a 12-triangle box with an emissive top,
    1024x768, 16 spp, MAX_DEPTH 6: the LBVH (plain and after treelet-7 restructuring) against the
    reference's KD tree (mean split) and the SAH KD tree (depth 24).  Every tree renders the same
    image; per tree: build time and Mrays/s of one fused 16-spp render."""
    import ptgpu_pc_amd as pt
    from ptgpu_pc_amd import scenes
    from ptgpu_pc_amd.layout import default_seeds
    W3, H3, SPP3, D3 = 1024, 768, 16, 6
    shapes, lc, cam = scenes.icosphere_in_box(7, seed=3, width=W3, height=H3)
    pt.update_camera(cam)
    n = len(shapes)
    ctx = pt.Context(0)
    d_sh = ctx.to_device(shapes)
    d_cam = ctx.to_device(cam)
    seeds0 = ctx.to_device(default_seeds(W3, H3))
    d_seeds = ctx.alloc(W3 * H3 * 8, np.uint32)
    d_col = ctx.alloc(W3 * H3 * 12, np.float32)
    d_pix = ctx.alloc(W3 * H3 * 4, np.uint32)
    d_rays = ctx.alloc(16, np.uint64).zero()

    d_stats = ctx.alloc(16, np.uint64)

    def roofline_of(bvh, kd, ms, box_bytes):
        """one untimed counting launch of the same frame: box and triangle tests"""
        d_seeds.copy_from(seeds0)
        d_stats.zero()
        pt.render(ctx, bvh, kd, colors=d_col, seeds=d_seeds, camera=d_cam, shapes=d_sh, shapeCnt=n, lightCnt=lc,
                  width=W3, height=H3, currentSample=0, spp=SPP3, max_depth=D3, pixels=d_pix, trav_stats=d_stats)
        box, prim = (int(x) for x in d_stats.download(np.uint64)[:2])
        alg = box_bytes * box + 36 * prim + W3 * H3 * (24 + 16 + 4)
        return render_roofline(alg, ms, box, prim, box_bytes, measured=False)

    def render(bvh, kd):
        best, rays = 1e9, 0
        for k in range(steps + 1):
            d_seeds.copy_from(seeds0)
            d_rays.zero()
            ctx.record(0)
            pt.render(ctx, bvh, kd, colors=d_col, seeds=d_seeds, camera=d_cam, shapes=d_sh, shapeCnt=n, lightCnt=lc,
                      width=W3, height=H3, currentSample=0, spp=SPP3, max_depth=D3, pixels=d_pix, ray_counter=d_rays)
            ctx.record(1)
            ms = ctx.elapsed_ms(0, 1)
            if k:   # the first render also prepares the tree (height, four-wide nodes, light list)
                best = min(best, ms)
            rays = int(d_rays.download(np.uint64)[:2].sum())
        return best, rays, d_pix.download(np.uint32)

    out = {}
    bvh = pt.CLBVH(ctx, d_sh, n)
    for name, treelet in (("lbvh", 0), ("lbvh_treelet7", 7)):
        ctx.synchronize()
        t0 = time.perf_counter()
        bvh.reset()
        bvh.buildRadixTree()
        bvh.buildBVHTree()
        if treelet:
            bvh.optimize(treelet)
        ctx.synchronize()
        build_ms = (time.perf_counter() - t0) * 1e3
        ms, rays, pix = render(bvh, None)
        out[name] = {"build_ms": round(build_ms, 2), "render_ms": round(ms, 2), "mrays_s": round(rays / ms / 1e3, 1),
                     "rays": rays, "roofline": roofline_of(bvh, None, ms, 16)}
        ref_pix = out.setdefault("_pix", pix)
        out[name]["pixels_identical"] = pix.tobytes() == ref_pix.tobytes()
    bvh.close()
    for name, sah in (("kd_mean_split", False), ("kd_sah", True)):
        t0 = time.perf_counter()
        kd = pt.KDTree(ctx, shapes, sah=sah, max_depth=24, max_leaf=4) if sah else pt.KDTree(ctx, shapes)
        build_ms = (time.perf_counter() - t0) * 1e3
        ms, rays, pix = render(None, kd)
        out[name] = {"build_ms_host": round(build_ms, 1), "render_ms": round(ms, 2),
                     "mrays_s": round(rays / ms / 1e3, 1), "rays": rays,
                     "pixels_identical": pix.tobytes() == out["_pix"].tobytes(), **kd.stats(),
                     "roofline": roofline_of(None, kd, ms, 32)}   # KD leaves share primitives: fp32 four-wide nodes
        kd.close()
    out.pop("_pix")
    tuning = None
    if tune:
        # the autotuner on this scene and frame (one frame per build): digit width, launch shape, treelet size
        bvh = pt.CLBVH(ctx, d_sh, n)
        bvh.buildRadixTree()
        bvh.buildBVHTree()
        d_seeds.copy_from(seeds0)
        tuning = pt.autotune(ctx, bvh, frames_per_build=1, trials=1, colors=d_col, seeds=d_seeds, camera=d_cam,
                             shapes=d_sh, shapeCnt=n, lightCnt=lc, width=W3, height=H3, currentSample=0, spp=SPP3,
                             max_depth=D3, pixels=d_pix)
        bvh.close()
    ctx.close()
    res = {"workload": "327,680-triangle displaced icosphere (seed 3) + 12-triangle box, 2 emissive, 1024x768, "
                       "16 spp, MAX_DEPTH 6", "shapes": n, "trees": out, "tuning": tuning}
    if cpu:
        # the reference's host builds as the oracle restates them, single-threaded (the KD tree of
        # KDTree.cpp:89-270; the LBVH with a stable sort in place of the O(N^2) bubble sort, which would
        # take minutes at this size), and the reference's own host render path on a pixel subset
        sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle as O
        t0 = time.perf_counter()
        s_h, _, btn, btl = O.build_lbvh(shapes)
        lbvh_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.KD(shapes)
        kd_s = time.perf_counter() - t0
        res["cpu_build_baseline"] = {
            "lbvh": {"value": round(n / lbvh_s / 1e6, 3), "unit": "Mtri/s", "cores": 1, "kind": "port",
                     "sample": f"one LBVH build of the {n:,} shapes (std::stable_sort for the bubble sort), "
                               f"{lbvh_s:.2f} s"},
            "kd_mean_split": {"value": round(n / kd_s / 1e6, 3), "unit": "Mtri/s", "cores": 1, "kind": "port",
                              "sample": f"one KDTree::build of the {n:,} shapes, {kd_s:.2f} s"}}
        res["cpu_baseline"] = cpu_baseline(s_h, lc, cam, btn, btl, budget_s=6.0, w=W3, h=H3, depth=D3,
                                           what="the same 1024x768 frame")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--engine", default="auto", choices=["auto", "megakernel", "wavefront", "lanes"])
    ap.add_argument("--regroup", type=int, default=None)
    ap.add_argument("--blocks-per-cu", type=int, default=None)
    ap.add_argument("--wide", type=int, default=None)
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="a launch tuning value (ptgpu_ctx_set_tuning) for every render context, e.g. split=0")
    ap.add_argument("--treelet", type=int, default=0, help="treelet size for CLBVH::optimize (0: no optimize)")
    ap.add_argument("--treelet-passes", type=int, default=1)
    ap.add_argument("--serve", action="store_true", help="only the served vs plain comparison (a JSON line)")
    ap.add_argument("--no-serve", action="store_true", help="skip the served comparison in the default line")
    ap.add_argument("--batch", action="store_true", help="only the 64-scene batched vs plain comparison (a JSON line)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched comparison in the default line")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--no-tune", action="store_true", help="keep the built-in defaults instead of autotuning")
    ap.add_argument("--graph", action="store_true", help="only the recorded-graph vs eager per-frame comparison")
    ap.add_argument("--c3", action="store_true", help="only the 327K-triangle icosphere: LBVH (+treelet 7) vs KD trees")
    ap.add_argument("--no-c3", action="store_true", help="skip the 327K-triangle icosphere block in the default line")
    ap.add_argument("--c1", action="store_true", help="only the Cornell spheres (no acceleration) frame")
    ap.add_argument("--no-c1", action="store_true", help="skip the Cornell spheres block in the default line")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the graph comparison in the default line (with --batch: the recorded batched step)")
    ap.add_argument("--frames", action="store_true",
                    help="only the C2 frame rendered 1, 2 and 4 per launch (settings of --tuning FILE, else the defaults)")
    ap.add_argument("--no-frames", action="store_true", help="skip the frames-per-launch comparison in the default line")
    ap.add_argument("--no-pipeline", action="store_true",
                    help="build and render each frame in turn on one stream (the same as --frames-in-flight 1)")
    ap.add_argument("--frames-in-flight", type=int, default=3, choices=(1, 2, 3, 4, 5),
                    help="1: build and render each frame in turn on one stream; 2: the next frame's tree is built "
                         "on a second stream while the current frame renders; 3 (default): as 2, and frames "
                         "alternate between two render streams with their own frame buffers, so a frame's render "
                         "starts while the previous frame's render drains; 4, 5: as 3 with 3 or 4 render streams")
    ap.add_argument("--tuning", default=None, help="JSON file with the tuning object of an earlier run (no autotune)")
    ap.add_argument("--requests", type=int, default=256)
    ap.add_argument("--rate", type=float, default=500.0)
    ap.add_argument("--pool", type=int, default=16)
    ap.add_argument("--spp-per-tick", type=int, default=16)
    ap.add_argument("--budget-gb", type=int, default=64)
    ap.add_argument("--profile-summary", type=str, default=None,
                    help="directory of rocprofv3 results (kt/, fetch/, write/, sq*/, commands.json) to summarise")
    ap.add_argument("--build-sweep", type=str, default=None,
                    help="comma-separated log2 sizes, e.g. 16,18,20,22,24: LBVH build GB/s per stage")
    args = ap.parse_args()
    if args.graph:
        print(json.dumps(graph_compare()), flush=True)
        return
    if args.c3:
        print(json.dumps(c3_compare(cpu=not args.no_cpu)), flush=True)
        return
    if args.c1:
        print(json.dumps(c1_compare(cpu=not args.no_cpu)), flush=True)
        return
    if args.frames:
        tun = json.loads(Path(args.tuning).read_text()) if args.tuning else None
        print(json.dumps(frames_compare(tun)), flush=True)
        return
    if args.batch:
        print(json.dumps(batch_compare(n_scenes=args.batch_scenes, tune=not args.no_tune, record=not args.no_graph)),
              flush=True)
        return
    if args.serve:
        print(json.dumps(serve_compare(args.requests, args.rate, args.pool, args.spp_per_tick, args.budget_gb,
                                       args.blocks_per_cu, args.regroup)), flush=True)
        return
    if args.profile_summary:
        profile_summary(Path(args.profile_summary), ROOT / "profiles" / "k_render_lanes_c2_soup65538_512x512x8spp.md")
        return
    if args.build_sweep:
        build_sweep([1 << int(x) for x in args.build_sweep.split(",")])
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path shards by scene or "
              f"pixel tile, see the design notes)", file=sys.stderr)
        sys.exit(2)

    import ptgpu_pc_amd as pt
    from ptgpu_pc_amd import scenes
    from ptgpu_pc_amd.layout import default_seeds

    shapes, lc, cam = scenes.soup(N_TRI, seed=1, width=W, height=H)
    pt.update_camera(cam)
    n = len(shapes)
    ctx = pt.Context(0)
    ctx.set_engine(args.engine)
    if args.regroup:
        ctx.set_tuning("regroup", args.regroup)
    if args.blocks_per_cu:
        ctx.set_tuning("blocks_per_cu", args.blocks_per_cu)
    if args.wide is not None:
        ctx.set_tuning("wide", args.wide)
    extra = [(kv.split("=")[0], int(kv.split("=")[1])) for kv in args.set]
    for k, v in extra:
        ctx.set_tuning(k, v)
    d_shapes = ctx.to_device(shapes)
    d_cam = ctx.to_device(cam)
    d_col = ctx.alloc(W * H * 12, np.float32).zero()
    seeds0 = ctx.to_device(default_seeds(W, H))
    d_seeds = ctx.alloc(W * H * 8, np.uint32)
    d_pix = ctx.alloc(W * H * 4, np.uint32)
    d_rays = ctx.alloc(16, np.uint64).zero()
    d_stats = ctx.alloc(16, np.uint64).zero()
    bvh = pt.CLBVH(ctx, d_shapes, n)
    treelet, passes = args.treelet, args.treelet_passes
    tuning = None
    if args.tuning:
        # settings from an earlier autotuned run (its line's "tuning" object), e.g. for profiled runs
        tuning = json.loads(Path(args.tuning).read_text())
        for k in ("blocks_per_cu", "regroup", "wide"):
            ctx.set_tuning(k, int(tuning[k]))
        bvh.set_digit_bits(int(tuning["digit_bits"]))
        if not args.treelet and tuning["treelet"]:
            treelet, passes = int(tuning["treelet"]), 1
    elif not args.no_tune:
        # autotuner: digit width, launch shape, treelet size for this scene and frame (one frame per
        # build; frames_per_build 0: each build runs beside the previous frame's render)
        bvh.buildRadixTree()
        bvh.buildBVHTree()
        d_seeds.copy_from(seeds0)
        tuning = pt.autotune(ctx, bvh, frames_per_build=1 if (args.no_pipeline or args.frames_in_flight == 1) else 0, trials=2, colors=d_col, seeds=d_seeds, camera=d_cam,
                             shapes=d_shapes, shapeCnt=n, lightCnt=lc, width=W, height=H, currentSample=0, spp=SPP,
                             max_depth=DEPTH, pixels=d_pix)
        if not args.treelet and tuning["treelet"]:
            treelet, passes = tuning["treelet"], 1
    # --set values win over the replayed or autotuned ones (the line's tuning object shows them)
    for k, v in extra:
        ctx.set_tuning(k, v)
        if tuning is not None and k in tuning:
            tuning[k] = v

    def step(ray_counter=None, trav_stats=None):
        ctx.record(0)
        bvh.reset()
        bvh.buildRadixTree()
        bvh.buildBVHTree()
        for _ in range(passes if treelet else 0):
            bvh.optimize(treelet)
        d_seeds.copy_from(seeds0)
        ctx.record(1)
        pt.render(ctx, bvh, None, colors=d_col, seeds=d_seeds, camera=d_cam, shapes=d_shapes, shapeCnt=n,
                  lightCnt=lc, width=W, height=H, currentSample=0, spp=SPP, max_depth=DEPTH, pixels=d_pix,
                  ray_counter=ray_counter, trav_stats=trav_stats)
        ctx.record(2)

    # Two frames in flight: frame k's tree is built on a second context's stream (its own handle and
    # shape array, one of two used in turn) while frame k-1 renders.  A build waits for the render
    # that last used its handle; the render of frame k waits for its build (the library inserts
    # that wait when the handle's context differs from the render's).  Event slots: build stream
    # 2i/2i+1 = build begin/end of parity i; render stream 6+2i/7+2i = render begin/end.
    # Frames in flight 2: frame k's tree is built on a second context's stream while frame k-1 renders
    # (two handles, each over its own copy of the shapes, used in turn).  A build waits for the
    # render that last used its handle; the render of frame k waits for its build (the library
    # inserts that wait when the handle's context differs from the render's).
    # Frames in flight 3: three handles, and each step enqueues frame k's render (its tree was built
    # during the previous step) and then frame k+1's build; frames alternate between two render
    # contexts (streams) with their own frame buffers.  A render then starts while the previous
    # frame's render drains (its longest pixel chains finish on a few waves while the rest of the
    # device would idle), and a build runs in whatever a draining render leaves free.  Every frame
    # still rebuilds its tree and renders all its samples.
    # Frames in flight 4 and 5: the same with 3 or 4 render contexts and as many handles as frames in
    # flight (measured no faster than 3 at C2: two renders already fill what a draining one leaves).
    # Event slots: build context 2s/2s+1 = build begin/end of handle s; render contexts 6+2s/7+2s =
    # render begin/end of the frame on handle s.
    fif = 1 if args.no_pipeline else args.frames_in_flight
    pipeline = fif >= 2
    if pipeline:
        cb = pt.Context(0)
        bits = int(tuning["digit_bits"]) if tuning else 8
        nslot = fif                                  # handles (each over its own copy of the shapes)
        nr = 1 if fif == 2 else fif - 1              # render contexts, each with its own frame buffers
        shape_bufs = [d_shapes] + [ctx.to_device(shapes) for _ in range(nslot - 1)]
        slots = [(pt.CLBVH(cb, sb, n, digit_bits=bits), sb) for sb in shape_bufs]
        rctx = [ctx]
        fbuf = [(d_col, d_seeds, d_pix)]
        for _ in range(nr - 1):
            r1 = pt.Context(0)
            r1.set_engine(args.engine)
            for k in ("blocks_per_cu", "regroup", "wide"):
                v = tuning[k] if tuning else {"blocks_per_cu": args.blocks_per_cu, "regroup": args.regroup,
                                              "wide": args.wide}[k]
                if v:
                    r1.set_tuning(k, int(v))
            for k, v in extra:
                r1.set_tuning(k, v)
            rctx.append(r1)
            fbuf.append((r1.alloc(W * H * 12, np.float32).zero(), r1.alloc(W * H * 8, np.uint32),
                         r1.alloc(W * H * 4, np.uint32)))
        for sl in range(nslot):
            for c in rctx:
                c.record(7 + 2 * sl)
        frame = [0]
        last = [None]          # (slot, render context) of the last frame enqueued
        built = [False]        # fif 3: the next frame's tree has been enqueued

        def build(sl, wait_ctx):
            b = slots[sl][0]
            cb.wait_event(wait_ctx, 7 + 2 * sl)    # the render that last used this handle
            cb.record(2 * sl)
            b.reset()
            b.buildRadixTree()
            b.buildBVHTree()
            for _ in range(passes if treelet else 0):
                b.optimize(treelet)
            cb.record(2 * sl + 1)

        def render(k, sl, ray_counter, trav_stats):
            rc = rctx[k % nr]
            col_i, seeds_i, pix_i = fbuf[k % nr]
            seeds_i.copy_from(seeds0)
            rc.wait_event(cb, 2 * sl + 1)
            rc.record(6 + 2 * sl)
            pt.render(rc, slots[sl][0], None, colors=col_i, seeds=seeds_i, camera=d_cam, shapes=slots[sl][1],
                      shapeCnt=n, lightCnt=lc, width=W, height=H, currentSample=0, spp=SPP, max_depth=DEPTH,
                      pixels=pix_i, ray_counter=ray_counter, trav_stats=trav_stats)
            rc.record(7 + 2 * sl)

        def pstep(ray_counter=None, trav_stats=None):
            """Enqueue frame k; returns (build_ms, render_ms) of frame k-1 (None for the first), read
            after frame k's render is queued so the host's wait never idles the device."""
            k = frame[0]
            frame[0] += 1
            sl = k % nslot
            if fif == 2:
                build(sl, ctx)
                render(k, sl, ray_counter, trav_stats)
            else:
                if not built[0]:
                    build(sl, rctx[(k - nslot) % nr])
                render(k, sl, ray_counter, trav_stats)
                # the next frame's tree; its handle was last rendered by frame k+1-nslot
                build((k + 1) % nslot, rctx[(k + 1 - nslot) % nr])
                built[0] = True
            prev, last[0] = last[0], (sl, rctx[k % nr])
            if prev is None:
                return None
            return cb.elapsed_ms(2 * prev[0], 2 * prev[0] + 1), prev[1].elapsed_ms(6 + 2 * prev[0], 7 + 2 * prev[0])

        def pdrain():
            """Timings of the last frame enqueued; leaves every stream idle (a build enqueued for a
            next frame is finished and used by the next step)."""
            prev, last[0] = last[0], None
            for c in (*rctx, cb):
                c.synchronize()
            return cb.elapsed_ms(2 * prev[0], 2 * prev[0] + 1), prev[1].elapsed_ms(6 + 2 * prev[0], 7 + 2 * prev[0])

    # counter calibration: one plain copy of a known 256 MiB (beyond the 256 MB MALL)
    cal_src = ctx.alloc(256 << 20, np.uint8)
    cal_dst = ctx.alloc(256 << 20, np.uint8)
    ctx.copy_calibration(cal_dst, cal_src, 256 << 20)
    ctx.synchronize()
    cal_src.free()
    cal_dst.free()

    if pipeline:
        for _ in range(args.warmup):
            pstep()
        pdrain()
        # one counting frame (untimed) for the roofline's algorithmic bytes
        pstep(trav_stats=d_stats)
        pdrain()
    else:
        for _ in range(args.warmup):
            step()
        ctx.synchronize()
        step(trav_stats=d_stats)
        ctx.synchronize()
    box_tests, prim_tests = (int(x) for x in d_stats.download(np.uint64)[:2])

    d_rays.zero()
    ctx.synchronize()
    render_ms, build_ms = [], []
    t0 = time.perf_counter()
    if pipeline:
        for _ in range(args.steps):
            r = pstep(ray_counter=d_rays)
            if r is not None:
                build_ms.append(r[0])
                render_ms.append(r[1])
        r = pdrain()
        build_ms.append(r[0])
        render_ms.append(r[1])
    else:
        for _ in range(args.steps):
            step(ray_counter=d_rays)
            build_ms.append(ctx.elapsed_ms(0, 1))
            render_ms.append(ctx.elapsed_ms(1, 2))
        ctx.synchronize()
    wall = time.perf_counter() - t0
    closest, anyhit = (int(x) for x in d_rays.download(np.uint64)[:2])
    rays = closest + anyhit
    overlapped_ms = None
    if pipeline:
        # the builds above ran beside a render, on what it left of the device; the build rate is
        # measured on builds alone (after the timed region), as the one-stream step runs them
        overlapped_ms = float(np.mean(build_ms))
        b = slots[0][0]
        build_ms = []
        for _ in range(5):
            cb.record(0)
            b.reset()
            b.buildRadixTree()
            b.buildBVHTree()
            for _ in range(passes if treelet else 0):
                b.optimize(treelet)
            cb.record(1)
            build_ms.append(cb.elapsed_ms(0, 1))
        stage = b.stage_times()
    else:
        stage = bvh.stage_times()
    mrays = rays / wall / 1e6
    render_avg = float(np.mean(render_ms))
    build_avg = float(np.mean(build_ms))

    # algorithmic bytes of one render launch: the node bytes per child box tested (16 B with the
    # default four-wide binary16 nodes, 32 B with fp32 four-wide or binary nodes), 36 B of vertices per
    # triangle test, and the frame buffers (colors 12 B read+write, seeds 8 B read+write, pixels 4 B
    # write per pixel)
    wide = int(tuning["wide"]) if tuning else (args.wide if args.wide is not None else 2)
    box_bytes = 16 if wide == 2 else 32
    alg_bytes = box_bytes * box_tests + 36 * prim_tests + W * H * (24 + 16 + 4)
    roofline = render_roofline(alg_bytes, render_avg, box_tests, prim_tests, box_bytes,
                               measured=args.engine in ("auto", "lanes") and not args.treelet,
                               step_ms=wall * 1e3 / args.steps if fif >= 3 else None)

    served = None
    if not args.no_serve:
        served = serve_compare(args.requests, args.rate, args.pool, args.spp_per_tick, args.budget_gb,
                               args.blocks_per_cu, None)

    graphed = None
    if not args.no_graph:
        graphed = graph_compare()

    per_launch = None
    if not args.no_frames:
        per_launch = frames_compare(tuning)

    batched = None
    if not args.no_batch:
        batched = batch_compare(n_scenes=args.batch_scenes, tune=not args.no_tune, cpu=not args.no_cpu)

    c1 = None
    if not args.no_c1:
        c1 = c1_compare(cpu=not args.no_cpu)

    c3 = None
    if not args.no_c3:
        c3 = c3_compare(steps=2, tune=False, cpu=not args.no_cpu)

    cpu = cpu_build = None
    if not args.no_cpu:
        btn, btl = bvh.getTrees()
        s_host = d_shapes.download(dtype=shapes.dtype, shape=(n,))
        cpu = cpu_baseline(s_host, lc, cam, btn, btl)
        cpu_build = cpu_build_baseline(shapes)

    out = {
        "metric": "Mrays/s + LBVH build Mtri/s; served p50/p99 frame latency vs plain GPU path",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": WORKLOAD,
                   "shapes": n, "width": W, "height": H, "spp": SPP, "max_depth": DEPTH,
                   "frames_in_flight": fif},
        "rays_per_step": rays // args.steps,
        "closest_per_step": closest // args.steps,
        "anyhit_per_step": anyhit // args.steps,
        "render_ms": round(render_avg, 3),
        # build_ms: the step's whole build (LBVH + the tuner's treelet pass, if any); build_mtri_s: the
        # LBVH construction alone (Morton, sort, topology, refit stages of the last step);
        # build_with_optimize_mtri_s: primitives over build_ms
        "build_ms": round(build_avg, 4),
        # with two frames in flight: the mean span of a build on its own stream beside the previous
        # frame's render (hidden behind it)
        "build_overlapped_ms": round(overlapped_ms, 4) if overlapped_ms is not None else None,
        "build_mtri_s": round(n / ((stage["morton"] + stage["sort"] + stage["topology"] + stage["refit"]) * 1e-3)
                              / 1e6, 1),
        "build_with_optimize_mtri_s": round(n / (build_avg * 1e-3) / 1e6, 1),
        "build_stage_ms": {k: round(v, 4) for k, v in stage.items()},
        "msamples_s": round(W * H * SPP * args.steps / wall / 1e6, 2),
        "roofline": roofline,
        "tuning": tuning,
        "served": served,
        "batched": batched,
        "frames_per_launch": per_launch,
        "graph": graphed,
        "c1": c1,
        "c3": c3,
        "cpu_baseline": cpu,
        "cpu_build_baseline": cpu_build,
    }
    print(json.dumps(out), flush=True)
    if pipeline:
        for b, _ in slots:
            b.close()
        cb.close()
        for c in rctx[1:]:
            c.close()
    bvh.close()
    ctx.close()


if __name__ == "__main__":
    main()
