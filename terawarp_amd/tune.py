"""Autotuner of the library's per-context tunables (tw_set_tune).

The reference hard-codes its counterparts: threads per block (ele.cuh:4), the
1000^3 block size (q_imgwarp_tps_quicksmallmemory.cpp:1264) and the read halo
(:1279).  Here the settings that do not change a result are measured on the
workloads they serve and persisted per device:

  tile   TW_TUNE_TILE: the fused resample's gather tile (64x4xTZ) or LDS
         column tile (16x16x64), bit-identical outputs
  tz     TW_TUNE_TZ / TW_TUNE_TZ_BATCH: the gather tile's depth along z in
         single-volume and in persistent (brick, batch) launches (0: 4 grid
         cells)
  brick  the brick edge of Warper.warp_stream (outputs identical: positions
         are global)

Stored settings come from the package's tuning file
terawarp_amd/tuning/<arch>.json (produced by `python -m terawarp_amd.tune` on
that GPU and shipped with the package), overridden by a user file
($TERAWARP_TUNE_FILE, default ~/.cache/terawarp/tune.json).  Every Warper
applies them when it creates its context (Warper(tuned=False) keeps the
library defaults).

Not tuned here: the FP32-factor threshold (TW_TUNE_FP32_FACTOR_MIN; its
library default comes from `bench.py --fit-sweep`), the read halo (computed
exactly from the field's bounds) and the grid spacing (it changes the output).
"""
import argparse
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PACKAGE_DIR = os.path.join(HERE, "tuning")
DEFAULTS = {"tile": 0, "tz": 0, "tz_batch": 0, "brick": 512}
TILES = {0: "auto", 1: "gather", 2: "lds"}


def user_file():
    return os.environ.get("TERAWARP_TUNE_FILE") or os.path.join(os.path.expanduser("~"), ".cache", "terawarp",
                                                                "tune.json")


def device_arch(device=0):
    """gfx name of the device ("gfx950"), the key of the stored settings."""
    import torch
    return torch.cuda.get_device_properties(device).gcnArchName.split(":")[0]


def _read(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def load_settings(arch):
    """The stored settings for `arch`: package file, then user file."""
    out = dict(DEFAULTS)
    out.update(_read(os.path.join(PACKAGE_DIR, arch + ".json")).get("settings", {}))
    out.update(_read(user_file()).get(arch, {}).get("settings", {}))
    return out


def apply(warper, settings):
    """Set the context tunables of a Warper from stored settings."""
    warper.set_tune("tile", int(settings.get("tile", 0)))
    warper.set_tune("tz", int(settings.get("tz", 0)))
    warper.set_tune("tz_batch", int(settings.get("tz_batch", 0)))
    warper.stream_brick = int(settings.get("brick", 512))


# ---- measurement ---------------------------------------------------------------

def _median_ms(fn, warper, stage, reps):
    fn()
    ms = []
    for _ in range(reps):
        fn()
        ms.append(warper.timings()[stage])
    return float(np.median(ms))


def bricks_workload(warper, nb=4, threads=16):
    """A brick layer of the streaming configuration (8192x8192x2048 uint16,
    K = 2048 landmarks, |d| <= 40, grid 8, 512^3 bricks), nb x nb bricks of
    its middle layer with their exact source windows in device memory."""
    import torch
    from . import synth, warp
    from .warp import Synth
    size, brick, g, zl = (8192, 8192, 2048), 512, 8, 1024
    src = Synth(size, np.uint16, seed=4, cell=256, blobs_per_cell=16)
    sub, tar = synth.landmarks(size, 2048, 40.0, seed=4)
    c, d = warper.tps_fit(tar, sub, precision=warp.MIXED)
    x0 = y0 = (size[0] // brick // 2 - nb // 2) * brick
    gsh = (warp.grid_dim(nb * brick, g), warp.grid_dim(nb * brick, g), warp.grid_dim(brick, g))
    grid = warper.new_grid(gsh, g, origin=(x0, y0, zl), storage=warp.F16, device=True)
    warper.grid_eval(tar, c, d, grid, precision=warp.MIXED)
    g32 = warper.new_grid(gsh, g, origin=(x0, y0, zl), device=True)
    warper.grid_eval(tar, c, d, g32, precision=warp.PLAIN)
    srcs, dsts, so, do = [], [], [], []
    nbn = brick // g
    for by in range(nb):
        for bx in range(nb):
            blk = g32.data[:, :, by * nbn:by * nbn + nbn + 3, bx * nbn:bx * nbn + nbn + 3].reshape(3, -1)
            dmin, dmax = blk.amin(dim=1).tolist(), blk.amax(dim=1).tolist()
            o = (x0 + bx * brick, y0 + by * brick, zl)
            lo = [max(0, int(np.floor(o[a] + dmin[a])) - 2) for a in range(3)]
            hi = [min(size[a], int(np.ceil(o[a] + brick - 1 + dmax[a])) + 3) for a in range(3)]
            lo[0] -= lo[0] % 8
            hi[0] = min(size[0], hi[0] + (-(hi[0] - lo[0])) % 8)
            n = [hi[a] - lo[a] for a in range(3)]
            host = np.empty((n[2], n[1], n[0]), np.uint16)
            src.fill_into(host, lo[0], lo[1], lo[2], n[0], n[1], n[2], threads)
            srcs.append(torch.from_numpy(host.view(np.int16)).cuda())
            dsts.append(torch.empty((brick, brick, brick), dtype=torch.int16, device="cuda"))
            so.append(tuple(lo))
            do.append(o)
    del g32
    prepared = warper.prepare_bricks(srcs, dsts, so, do, size)
    state = {"grid": grid, "srcs": srcs, "dsts": dsts, "prepared": prepared}

    def run():
        warper.resample_bricks(None, None, grid, None, None, None, precision=warp.MIXED, prepared=prepared)

    return run, state


def batch_workload(warper):
    """The batched configuration: 256 uneven instances (TPS, affine, rigid)
    in one tw_warp_batch call."""
    import torch
    from . import synth, warp
    specs = synth.batch_instances(seed=5, n=256, dmax=8.0)
    items = []
    for sp in specs:
        s = synth.blob_volume_device(sp["size"], sp["dtype"], sp["seed"])
        it = {"src": s, "dst": torch.zeros_like(s)}
        if sp["kind"] == "tps":
            it.update(centres=sp["centres"], targets=sp["targets"])
        else:
            it["M"] = sp["M"]
        items.append(it)

    def run():
        warper.warp_batch(items, f=8, lam=0.2, precision=warp.MIXED)

    return run, items


def warp_workload(warper):
    """The plain-GPU configuration's chain: 1024x1024x256 uint16, K = 512
    landmarks, |d| <= 20, grid 8, one tw_warp_tps call on device buffers."""
    import torch
    from . import synth, warp
    size = (1024, 1024, 256)
    src = synth.blob_volume_device(size, np.uint16, 2)
    dst = torch.empty_like(src)
    sub, tar = synth.landmarks(size, 512, 20.0, seed=2)

    def run():
        warper.warp_tps(src, tar, sub, f=8, lam=0.2, precision=warp.MIXED, dst=dst)

    return run, (src, dst)


def tune(warper, reps=3, tz_candidates=(16, 24, 32, 48, 64, 96, 128), bricks=4, stream=True, log=print):
    """Measure every (tile, tile depth) candidate on the brick-layer and batch
    workloads (persistent launches) and on the plain-GPU chain (a
    single-volume launch) by resample-stage device time, and the streaming
    brick edge on a 1024x1024x512 volume by wall time.  The depths minimise
    the summed time relative to the library defaults over their family's
    workloads; the tile (shared by both families) is the LDS column tile only
    if it beats the best gather depths summed over all three."""
    import torch
    from . import synth, warp
    from .warp import Synth
    warper.set_timing(True)
    keys = ["auto/0"] + ["gather/%d" % tz for tz in tz_candidates] + ["lds/0"]
    evidence = {}
    kinds = (("bricks", "tz_batch", lambda: bricks_workload(warper, nb=bricks)),
             ("batch", "tz_batch", lambda: batch_workload(warper)),
             ("warp", "tz", lambda: warp_workload(warper)))
    for kind, knob, make in kinds:
        log("tuning on", kind)
        run, keep = make()
        times = {}
        for key in keys:
            tile, tz = key.split("/")
            warper.set_tune("tile", {v: k for k, v in TILES.items()}[tile])
            warper.set_tune(knob, int(tz))
            times[key] = _median_ms(run, warper, "resample_ms", reps)
            log("  ", kind, key, "%.3f ms" % times[key])
        warper.set_tune(knob, 0)
        warper.set_tune("tile", 0)
        evidence[kind] = {"resample_ms": times, "fixed": "auto/0", "knob": knob}
        del keep, run
        torch.cuda.empty_cache()

    def rel(kind, key):
        return evidence[kind]["resample_ms"][key] / evidence[kind]["resample_ms"]["auto/0"]

    gathers = ["gather/%d" % tz for tz in tz_candidates]
    best_b = min(gathers, key=lambda k: rel("bricks", k) + rel("batch", k))
    best_w = min(gathers, key=lambda k: rel("warp", k))
    score = {"gather": rel("bricks", best_b) + rel("batch", best_b) + rel("warp", best_w),
             "lds": sum(rel(k, "lds/0") for k in ("bricks", "batch", "warp"))}
    lds = score["lds"] < score["gather"]
    settings = {"tile": 2 if lds else 1, "tz": int(best_w.split("/")[1]), "tz_batch": int(best_b.split("/")[1]),
                "brick": DEFAULTS["brick"]}
    evidence["bricks"]["chosen"] = evidence["batch"]["chosen"] = "lds/0" if lds else best_b
    evidence["warp"]["chosen"] = "lds/0" if lds else best_w
    apply(warper, settings)
    if stream:
        log("tuning the streaming brick edge")
        size = (1024, 1024, 512)
        src = Synth(size, np.uint16, seed=4, cell=256, blobs_per_cell=16)
        sub, tar = synth.landmarks(size, 512, 20.0, seed=4)
        walls = {}
        for brick in (128, 256, 384, 512):
            warper.warp_stream(size, np.uint16, tar, sub, src, f=8, brick=brick, precision=warp.MIXED)
            t = []
            for _ in range(reps):
                t0 = time.perf_counter()
                warper.warp_stream(size, np.uint16, tar, sub, src, f=8, brick=brick, precision=warp.MIXED)
                t.append(time.perf_counter() - t0)
            walls[str(brick)] = float(np.median(t)) * 1e3
            log("   stream brick", brick, "%.1f ms" % walls[str(brick)])
        evidence["stream"] = {"wall_ms": walls, "fixed": str(DEFAULTS["brick"]), "size_xyz": list(size)}
        settings["brick"] = int(min(walls, key=walls.get))
        warper.stream_brick = settings["brick"]
    return {"settings": settings, "score": score, "evidence": evidence}


def report_md(arch, result):
    """The evidence as a markdown table (tuned vs fixed)."""
    lines = ["| workload | setting | time ms | vs fixed |", "|---|---|---:|---:|"]
    s = result["settings"]
    for kind, ev in result["evidence"].items():
        if kind == "stream":
            fixed = ev["wall_ms"][ev["fixed"]]
            for b, t in ev["wall_ms"].items():
                mark = " (chosen)" if int(b) == s["brick"] else (" (fixed)" if b == ev["fixed"] else "")
                lines.append("| stream %s | brick %s%s | %.1f | %.3f |" % ("x".join(map(str, ev["size_xyz"])), b, mark, t,
                                                                        t / fixed))
            continue
        fixed = ev["resample_ms"][ev["fixed"]]
        for key, t in ev["resample_ms"].items():
            mark = " (chosen)" if key == ev.get("chosen") else (" (fixed)" if key == ev["fixed"] else "")
            lines.append("| %s | %s%s | %.3f | %.3f |" % (kind, key, mark, t, t / fixed))
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=None, help="write the settings here (default: the user file)")
    ap.add_argument("--package", action="store_true", help="write the package's tuning file for this device")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--no-stream", action="store_true")
    ap.add_argument("--md", default=None, help="also write the evidence as markdown to this path")
    args = ap.parse_args()
    from .warp import Warper
    arch = device_arch()
    w = Warper(0, timing=True, tuned=False)
    res = tune(w, reps=args.reps, stream=not args.no_stream)
    w.close()
    res["arch"] = arch
    if args.package:
        os.makedirs(PACKAGE_DIR, exist_ok=True)
        path = os.path.join(PACKAGE_DIR, arch + ".json")
        with open(path, "w") as f:
            json.dump(res, f, indent=1, sort_keys=True)
    else:
        path = args.out or user_file()
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        data = _read(path)
        data[arch] = res
        with open(path, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
    md = report_md(arch, res)
    if args.md:
        with open(args.md, "w") as f:
            f.write(md + "\n")
    print(json.dumps({"arch": arch, "settings": res["settings"], "file": path}))
    print(md)


if __name__ == "__main__":
    main()
