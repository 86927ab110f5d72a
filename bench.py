"""Benchmark: MPTC per-frame encode throughput (Mblocks/s) at 1080p on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload: the 1080p plain-encode sequence (64 frames of 1920x1080 RGB8 from
the seeded generator: gradient, 32 rectangles with velocities in [-4, 4]
px/frame, 1 % scene cuts, noise in [-8, 8]), search area S = 8 (16-block
window), error threshold 0, an intra frame every 8 frames.  All frames are
uploaded to HBM before timing.  A step is the device-resident encode of one
frame (DXT1 fit + inter search + intra search + re-fit + commit, entropy
stage excluded); the K timed steps walk the sequence from frame 0.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "encode Mblocks/s at 1080p (plain vs learned-proposal) at stated PSNR delta (dB)"
WORKLOAD = "c2_t0"
# Algorithmic VALU work per candidate evaluation (lane-ops): 288 fp32 for the
# least-squares sums (16 px x 18 mul/add, no FMA), ~28 solve, ~96 quantise+snap,
# ~24 palette, ~10 validity, ~154 integer SSE.
W_CAND = 600
# MI355X vector ALU: 256 CU x 4 SIMD x 16 lanes x 2.4 GHz, one lane-op per lane-cycle
VALU_PEAK_TLOPS = 256 * 4 * 16 * 2.4e9 / 1e12
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=2)
    ap.add_argument("--workload", default=WORKLOAD)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device-to-device copy of known size (rocprof counter scale)")
    return ap.parse_args()


def cpu_baseline(spec, frames, n_frames):
    """The reference encoder compiled from its sources (oracle/_ref), or the C
    oracle where that was not built, on the first n_frames frames."""
    import oracle
    nb = (spec["width"] // 4) * (spec["height"] // 4)
    sample = frames[:n_frames]
    if oracle.reference_available():
        R = oracle.Reference()
        res = R.encode_sequence(sample, spec["S"], spec["thr"], spec["intra_interval"])
        t = sum(r["t_fit"] + r["t_reencode"] for r in res)
        kind = "reference"
    else:
        O = oracle.Oracle()
        t0 = time.perf_counter()
        O.encode_sequence(sample, spec["S"], spec["thr"], spec["intra_interval"])
        t = time.perf_counter() - t0
        kind = "port"
    return {"value": n_frames * nb / t / 1e6, "unit": "Mblocks/s", "cores": 1, "kind": kind,
            "sample": f"first {n_frames} frames of the same 1080p sequence (frame 0 intra, then inter), "
                      f"{n_frames * nb} blocks, {t:.1f} s single-threaded; fit (PNG decode excluded) + Reencode"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    from mptc_amd import hip, lib
    from mptc_amd.workloads import SEQUENCES, synth_cfg

    hip.set_device(0)
    if a.copy_calibration:
        n = 256 << 20
        src, dst = hip.Buffer(n), hip.Buffer(n)
        for _ in range(5):
            hip.check(hip.rt().hipMemcpy(hip.C.c_void_p(dst.ptr), hip.C.c_void_p(src.ptr), hip.C.c_size_t(n),
                                         hip.hipMemcpyDeviceToDevice))
        hip.synchronize()
        print(json.dumps({"copy_bytes": n, "copies": 5}))
        return

    spec = SEQUENCES[a.workload]
    W, H = spec["width"], spec["height"]
    nb = (W // 4) * (H // 4)
    n_seq = spec["frames"]
    frames = lib.synth_frames(synth_cfg(spec), 0, n_seq)
    enc = lib.Encoder(W, H, spec["S"], spec["thr"])
    stream = hip.Stream()
    d_frames = hip.Buffer.from_array(frames)
    fbytes = W * H * 3
    n_out = max(a.steps, a.warmup, 1)
    d_phys = [hip.Buffer(8 * nb) for _ in range(n_out)]
    d_motion = [hip.Buffer(2 * nb) for _ in range(n_out)]
    d_unique = [hip.Buffer(4 * nb) for _ in range(n_out)]
    d_nuniq = hip.Buffer(4 * n_out)
    d_eps = hip.Buffer(6 * nb)

    def run(n_steps, collect=False):
        set_intra, num = True, 1
        stats = []
        for f in range(n_steps):
            fi = f % n_seq
            if fi == 0:
                set_intra, num = True, 1
            enc.encode_frame_device(d_frames.offset(fi * fbytes), W, H, set_intra,
                                    None if set_intra else d_phys[(f - 1) % n_out], d_phys[f % n_out],
                                    d_motion[f % n_out], d_unique[f % n_out], d_nuniq.offset(4 * (f % n_out)),
                                    d_eps if collect else None, None, stream=stream.handle)
            if collect:
                stats.append(enc.stats())
            set_intra, num = lib.next_intra(set_intra, num, spec["intra_interval"])
        return stats

    # warmup
    run(a.warmup)
    stream.synchronize()
    # timed region: K steps bracketed by device synchronisation
    hip.synchronize()
    t0 = time.perf_counter()
    run(a.steps)
    stream.synchronize()
    hip.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = a.steps * nb / elapsed / 1e6
    # the same K steps again with a HIP event pair around every kernel launch
    # (recorded on the launch stream) for the per-kernel durations
    enc.profile_enable(True)
    hip.synchronize()
    tp0 = time.perf_counter()
    run(a.steps)
    stream.synchronize()
    hip.synchronize()
    tp1 = time.perf_counter()
    kernels = {}
    for name in ("fit", "table", "inter", "inter_combine", "intra_prep", "intra_dense", "intra_combine",
                 "intra_apply", "intra_reeval", "intra_select", "commit", "unique", "zero"):
        n, ms = enc.profile_read(name)
        if n:
            kernels[name] = {"launches": n, "total_ms": round(ms, 4), "avg_ms": round(ms / n, 5)}
    enc.profile_enable(False)

    # work counters (one untimed pass over the same frames)
    stats = run(a.steps, collect=True)
    inter_c = sum(s["inter_candidates"] for s in stats)
    intra_c = sum(s["intra_candidates"] for s in stats)
    reev_c = sum(s["reeval_candidates"] for s in stats)
    kinds = {k: int(sum(s[k] for s in stats)) for k in ("n_inter", "n_intra", "n_unique")}
    iters = [s["intra_iterations"] for s in stats]

    # roofline of the dominant kernel (VALU-bound: no contraction maps to MFMA)
    cand_of = {"inter": inter_c, "intra_dense": intra_c, "intra_reeval": reev_c}
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    roof = None
    if dom in cand_of and kernels[dom]["launches"]:
        per_launch_ops = cand_of[dom] * W_CAND / kernels[dom]["launches"]
        achieved = per_launch_ops / (kernels[dom]["avg_ms"] * 1e-3) / 1e12
        roof = {"kernel": dom, "bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_PEAK_TLOPS, 2),
                "unit": "Tlane-op/s", "frac": round(achieved / VALU_PEAK_TLOPS, 4), "traffic": None,
                "work_per_launch": f"{cand_of[dom] / kernels[dom]['launches']:.4g} candidate evaluations x "
                                   f"{W_CAND} lane-ops"}

    # quality of the last timed pass (decoder-side PSNR over physical blocks)
    psnr = None
    try:
        import oracle
        O = oracle.Oracle()
        last = (a.steps - 1) % n_out
        fi = (a.steps - 1) % n_seq
        phys = d_phys[last].download(np.uint64, nb)
        psnr = round(O.psnr_physical(frames[fi], phys), 4)
    except Exception:
        pass

    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "Mblocks/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(elapsed / a.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+i32",
        "data": "synthetic (seeded generator, mptc_amd/csrc/framegen.h)",
        "config": {"workload": f"1080p plain encode ({a.workload}): {n_seq} frames {W}x{H} RGB8, "
                               f"S={spec['S']} ({2 * spec['S']}-block window), threshold {spec['thr']}, "
                               f"intra every {spec['intra_interval']} frames; bit-exact to the reference",
                   "blocks_per_frame": nb, "device": hip.device_name(0)},
        "roofline": roof,
        "kernels": kernels,
        "ms_per_step_with_kernel_events": round((tp1 - tp0) / a.steps * 1e3, 4),
        "work": {"inter_candidates": inter_c, "intra_candidates": intra_c, "reeval_candidates": reev_c,
                 **kinds, "max_intra_iterations": max(iters) if iters else 0},
        "psnr_db_last_frame": psnr,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(spec, frames, a.cpu_frames)
        out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)
    enc.close()


if __name__ == "__main__":
    main()
