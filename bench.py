"""Benchmark: MPTC per-frame encode throughput (Mblocks/s) at 1080p on one MI355X,
plain (bit-exact) search vs learned-proposal search, with the PSNR delta.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-learned]

Workload: the 1080p sequence (64 frames of 1920x1080 RGB8 from the seeded
generator: gradient, 32 rectangles with velocities in [-4, 4] px/frame, 1 %
scene cuts, noise in [-8, 8]), search area S = 8 (16-block window), error
threshold 0, an intra frame every 8 frames.  All frames are uploaded to HBM
before timing.  A step is the device-resident encode of one frame (DXT1 fit +
inter search + intra search + re-fit + commit + unique palette, entropy stage
excluded); the K timed steps walk the sequence from frame 0.  `value` is the
plain path, bit-exact to the reference encoder.

The learned path is trained inside the run: the plain search records which
offsets pass its accept test on 1080p training sequences of other seeds, both
proposal networks are trained on the GPU, and the same K steps are encoded
with top-8 proposals (exact acceptance).  Prints one JSON line.
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
# Algorithmic work of one candidate evaluation (lane-ops, the unit of the
# search kernels): 288 fp32 for the least-squares sums (16 px x 18 mul/add,
# no FMA), ~28 solve, ~96 quantise + snap, ~24 palette, ~10 validity, ~154
# integer SSE.
W_CAND = 600
# Work of one candidate bound of the pruned search: 36 v_dot4 group sums, 9
# subtractions for the last group, 20 integer multiply/adds for |S_g|^2, 11
# fp32 convert/multiply/adds for sum |S_g|^2 / n_g, subtract and compare.
W_BOUND = 78
# MI355X vector ALU: 157.3 TFLOP/s fp32 = 256 CU x 128 lanes (packed / dual
# issue) x 2 (FMA) x 2.4 GHz, i.e. 78.6 T lane-ops/s of non-FMA fp32 work.
VALU_PEAK = 78.6
MFMA_BF16_PEAK = 2500.0  # TFLOP/s, dense
NET_FLOP_L12 = 2 * (48 * 256 + 256 * 256)  # per block
NET_FLOP_L3 = 2 * 256 * 262
TRAIN_SEEDS = list(range(1000, 1007))  # 7 x 32 frames of 1080p


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--cpu-frames", type=int, default=2)
    ap.add_argument("--top-k", type=int, default=8)
    ap.add_argument("--train-steps", type=int, default=3000)
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


KERNELS = ("fit", "inter", "inter_combine", "intra_prep", "intra_dense", "intra_combine", "intra_apply",
           "intra_reeval", "intra_select", "commit", "unique", "net_input", "net_l12", "net_l3",
           "propose_inter", "top_intra", "lintra_pass", "pruned_inter", "pruned_intra", "fold_intra")


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "encode_1080p_plain_hbm_traffic.md")
KERNEL_SYMBOL = {"inter": "k_inter", "intra_dense": "k_intra_dense", "intra_reeval": "k_intra_reeval",
                 "pruned_inter": "k_pruned_search<false>", "pruned_intra": "k_pruned_search<true>"}


def profiled_traffic(kernel):
    """HBM bytes per dispatch of `kernel` (read + write), from the rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated against a copy
    of known size); None when the summary has no row for it."""
    sym = KERNEL_SYMBOL.get(kernel)
    if sym is None or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0] == sym:
            return round((float(cells[3]) + float(cells[4])) * 1e6)
    return None


class Runner:
    """Device-resident encodes of the workload sequence."""

    def __init__(self, hip, lib, enc, spec, frames, n_out):
        self.hip, self.lib, self.enc, self.spec = hip, lib, enc, spec
        self.W, self.H = spec["width"], spec["height"]
        self.nb = (self.W // 4) * (self.H // 4)
        self.n_seq = frames.shape[0]
        self.frames = frames
        self.d_frames = hip.Buffer.from_array(frames)
        self.n_out = n_out
        self.d_phys = [hip.Buffer(8 * self.nb) for _ in range(n_out)]
        self.d_motion = [hip.Buffer(2 * self.nb) for _ in range(n_out)]
        self.d_unique = [hip.Buffer(4 * self.nb) for _ in range(n_out)]
        self.d_nuniq = hip.Buffer(4 * n_out)
        self.stream = hip.Stream()

    def run(self, n_steps, collect=False):
        set_intra, num = True, 1
        stats = []
        fb = self.W * self.H * 3
        for f in range(n_steps):
            fi = f % self.n_seq
            if fi == 0:
                set_intra, num = True, 1
            o = f % self.n_out
            self.enc.encode_frame_device(self.d_frames.offset(fi * fb), self.W, self.H, set_intra,
                                         None if set_intra else self.d_phys[(f - 1) % self.n_out], self.d_phys[o],
                                         self.d_motion[o], self.d_unique[o], self.d_nuniq.offset(4 * o), None, None,
                                         stream=self.stream.handle)
            if collect:
                stats.append(self.enc.stats())
            set_intra, num = self.lib.next_intra(set_intra, num, self.spec["intra_interval"])
        return stats

    def autotune(self, steps=3):
        """Picks the launch shape (mptc_autotune) on frame 1 of the sequence, an
        inter frame encoded against frame 0; before any timed work."""
        self.run(1)
        fb = self.W * self.H * 3
        io = dict(d_rgb=self.d_frames.offset(fb), width=self.W, height=self.H, is_intra=False,
                  d_ref=self.d_phys[0], d_phys=self.d_phys[1 % self.n_out], d_motion=self.d_motion[0],
                  d_unique=self.d_unique[0], d_n_unique=self.d_nuniq)
        best, ms = self.enc.autotune([io], steps=steps)
        return best, ms

    def timed(self, steps, warmup):
        self.run(warmup)
        self.stream.synchronize()
        self.hip.synchronize()
        t0 = time.perf_counter()
        self.run(steps)
        self.stream.synchronize()
        self.hip.synchronize()
        return time.perf_counter() - t0

    def kernels(self, steps):
        enc = self.enc
        enc.profile_enable(True)
        self.hip.synchronize()
        t0 = time.perf_counter()
        self.run(steps)
        self.stream.synchronize()
        t1 = time.perf_counter()
        out = {}
        for name in KERNELS:
            n, ms = enc.profile_read(name)
            if n:
                out[name] = {"launches": n, "total_ms": round(ms, 4), "avg_ms": round(ms / n, 5)}
        enc.profile_enable(False)
        return out, (t1 - t0) / steps * 1e3

    def quality(self, steps):
        """Mean decoder-side PSNR of the frames the last run wrote (the last
        min(steps, n_out) of them), and their unique-block count."""
        import oracle
        O = oracle.Oracle()
        ps = []
        for f in range(max(0, steps - self.n_out), steps):
            phys = self.d_phys[f % self.n_out].download(np.uint64, self.nb)
            ps.append(O.psnr_physical(self.frames[f % self.n_seq], phys))
        return float(np.mean(ps))


def work_of(stats):
    return {"inter_candidates": sum(s["inter_candidates"] for s in stats),
            "intra_candidates": sum(s["intra_candidates"] for s in stats),
            "reeval_candidates": sum(s["reeval_candidates"] for s in stats),
            "exact_inter_candidates": sum(s["exact_inter_candidates"] for s in stats),
            "exact_intra_candidates": sum(s["exact_intra_candidates"] for s in stats),
            "n_inter": sum(s["n_inter"] for s in stats), "n_intra": sum(s["n_intra"] for s in stats),
            "n_unique": sum(s["n_unique"] for s in stats),
            "max_intra_iterations": max(s["intra_iterations"] for s in stats) if stats else 0}


def train_networks(hip, lib, enc, spec, train_steps):
    """Acceptance sets recorded by the exhaustive search on 1080p training
    sequences (seeds other than the benchmark's), then GPU training of both
    networks."""
    from mptc_amd.workloads import synth_cfg
    W, H = spec["width"], spec["height"]
    nb = (W // 4) * (H // 4)
    cap = 1 << 22
    bufs = {net: dict(px=hip.Buffer(48 * cap), tgt=hip.Buffer(32 * cap), eps=hip.Buffer(6 * cap),
                      count=hip.Buffer.from_array(np.zeros(1, np.uint32))) for net in (0, 1)}
    d_phys = [hip.Buffer(8 * nb) for _ in range(2)]
    d_motion, d_unique, d_nu, d_eps = hip.Buffer(2 * nb), hip.Buffer(4 * nb), hip.Buffer(4), hip.Buffer(6 * nb)
    t0 = time.perf_counter()
    enc.set_search_mode(lib.SEARCH_EXHAUSTIVE, 8)
    enc.record_pairs(True)
    n_frames = 0
    for seed in TRAIN_SEEDS:
        cfg = dict(spec)
        cfg["seed"] = seed
        frames = lib.synth_frames(synth_cfg(cfg), 0, 32)
        d_fr = hip.Buffer.from_array(frames)
        set_intra, num = True, 1
        for f in range(32):
            rgb = d_fr.offset(f * W * H * 3)
            enc.encode_frame_device(rgb, W, H, set_intra, None if set_intra else d_phys[(f + 1) % 2], d_phys[f % 2],
                                    d_motion, d_unique, d_nu, d_eps, None)
            for net in (0, 1):
                b = bufs[net]
                enc.collect_pairs(net, b["px"], b["tgt"], b["eps"], b["count"], cap)
            set_intra, num = lib.next_intra(set_intra, num, spec["intra_interval"])
        n_frames += 32
        d_fr.free()
    enc.record_pairs(False)
    t_collect = time.perf_counter() - t0
    n_pairs = {net: min(cap, int(bufs[net]["count"].download(np.uint32, 1)[0])) for net in (0, 1)}
    info = {"training_frames": n_frames, "training_seeds": TRAIN_SEEDS, "pairs_inter": n_pairs[0],
            "pairs_intra": n_pairs[1], "collect_s": round(t_collect, 2)}
    for net in (0, 1):
        enc.net_init(net, 11 + net)
        b = bufs[net]
        t1 = time.perf_counter()
        losses = enc.net_train(net, b["px"], b["tgt"], b["eps"], n_pairs[net], batch=4096, steps=train_steps,
                               lr=2e-3, seed_weight=1.0, seed=7, lr_final=1e-4)
        key = "inter" if net == 0 else "intra"
        info[f"train_s_{key}"] = round(time.perf_counter() - t1, 2)
        info[f"loss_{key}"] = [round(float(losses[:20].mean()), 4), round(float(losses[-50:].mean()), 4)]
    for b in bufs.values():
        for v in b.values():
            if hasattr(v, "free"):
                v.free()
    return info


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
    frames = lib.synth_frames(synth_cfg(spec), 0, spec["frames"])
    enc = lib.Encoder(W, H, spec["S"], spec["thr"])
    R = Runner(hip, lib, enc, spec, frames, max(a.steps, a.warmup, 2))

    # plain (exhaustive, bit-exact) path, in the launch shape the autotuner picks
    tuning, tune_ms = R.autotune() if not a.no_autotune else (enc.get_tuning(), None)
    elapsed = R.timed(a.steps, a.warmup)
    value = a.steps * nb / elapsed / 1e6
    kernels, ms_ev = R.kernels(a.steps)
    work = work_of(R.run(a.steps, collect=True))
    psnr_plain = R.quality(a.steps)

    # roofline of the dominant search kernel (vector-ALU bound: no contraction maps to MFMA).
    # Algorithmic lane-ops of each search kernel over the run: exact evaluations
    # x W_CAND; the pruned kernels bound every candidate and evaluate the survivors.
    ev, bd = W_CAND, W_BOUND
    work_of_kernel = {
        "inter": (work["inter_candidates"] * ev, [(work["inter_candidates"], "evaluations", ev)]),
        "intra_dense": (work["intra_candidates"] * ev, [(work["intra_candidates"], "evaluations", ev)]),
        "intra_reeval": (work["reeval_candidates"] * ev, [(work["reeval_candidates"], "evaluations", ev)]),
        "pruned_inter": (work["inter_candidates"] * bd + work["exact_inter_candidates"] * ev,
                         [(work["inter_candidates"], "bounds", bd), (work["exact_inter_candidates"], "evaluations", ev)]),
        "pruned_intra": (work["intra_candidates"] * bd + work["exact_intra_candidates"] * ev,
                         [(work["intra_candidates"], "bounds", bd), (work["exact_intra_candidates"], "evaluations", ev)]),
    }
    dom = max((k for k in kernels if k in work_of_kernel), key=lambda k: kernels[k]["total_ms"])
    lane_ops, parts = work_of_kernel[dom]
    n_launch = kernels[dom]["launches"]
    achieved = lane_ops / n_launch / (kernels[dom]["avg_ms"] * 1e-3) / 1e12
    roof = {"kernel": dom, "bound": "valu", "achieved": round(achieved, 3), "peak": VALU_PEAK,
            "unit": "T lane-op/s", "frac": round(achieved / VALU_PEAK, 4), "traffic": profiled_traffic(dom),
            "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT),
            "work_per_launch": " + ".join(f"{n / n_launch:.4g} candidate {what} x {w} lane-ops"
                                          for n, what, w in parts)}

    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "Mblocks/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(elapsed / a.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+i32",
        "data": "synthetic (seeded generator, mptc_amd/csrc/framegen.h)",
        "config": {"workload": f"1080p encode ({a.workload}): {spec['frames']} frames {W}x{H} RGB8, "
                               f"S={spec['S']} ({2 * spec['S']}-block window), threshold {spec['thr']}, "
                               f"intra every {spec['intra_interval']} frames; value = plain search, "
                               f"bit-exact to the reference", "blocks_per_frame": nb,
                   "tuning": tuning},
        "roofline": roof,
        "kernels": kernels,
        "ms_per_step_with_kernel_events": round(ms_ev, 4),
        "work": work,
        "psnr_db": round(psnr_plain, 4),
    }

    if not a.no_learned:
        info = train_networks(hip, lib, enc, spec, a.train_steps)
        learned = {"top_k": a.top_k, "training": info}
        enc.set_search_mode(lib.SEARCH_LEARNED, a.top_k)
        el = R.timed(a.steps, a.warmup)
        lk, lms_ev = R.kernels(a.steps)
        lw = work_of(R.run(a.steps, collect=True))
        lpsnr = R.quality(a.steps)
        lval = a.steps * nb / el / 1e6
        reused_plain = work["n_inter"] + work["n_intra"]
        reused = lw["n_inter"] + lw["n_intra"]
        learned.update({
            "value": round(lval, 4), "ms_per_step": round(el / a.steps * 1e3, 4),
            "speedup_vs_plain": round(lval / value, 3),
            "psnr_db": round(lpsnr, 4), "psnr_delta_db": round(lpsnr - psnr_plain, 4),
            "reused_blocks": reused, "reused_blocks_plain": reused_plain,
            "reuse_recall": round(reused / max(1, reused_plain), 4),
            "unique_blocks_delta_pct": round(100.0 * (lw["n_unique"] - work["n_unique"]) / max(1, work["n_unique"]), 3),
            "kernels": lk, "ms_per_step_with_kernel_events": round(lms_ev, 4), "work": lw,
        })
        # MFMA roofline of the network's layer-1+2 kernel (bf16 dense)
        if "net_l12" in lk:
            rows = nb  # each launch covers every block of the frame
            ach = rows * NET_FLOP_L12 / (lk["net_l12"]["avg_ms"] * 1e-3) / 1e12
            learned["roofline_net_l12"] = {"bound": "mfma", "achieved": round(ach, 2), "peak": MFMA_BF16_PEAK,
                                           "unit": "TFLOP/s", "frac": round(ach / MFMA_BF16_PEAK, 4),
                                           "work_per_launch": f"{rows} blocks x {NET_FLOP_L12} flop"}
        # ablation: the same search with fixed nearest-first scores, and with untrained networks
        Ls = np.arange(256)
        abl = {}
        for name in ("nearest_prior", "untrained"):
            if name == "nearest_prior":
                for net in (0, 1):
                    dx = Ls % 16 - 8
                    dy = Ls // 16 - 8 if net == 0 else Ls // 16
                    enc.set_prior(net, (-(dx * dx + dy * dy) - 1e-3 * Ls).astype(np.float32))
            else:
                for net in (0, 1):
                    enc.set_prior(net, None)
                    enc.net_init(net, 99 + net)
            e2 = R.timed(a.steps, a.warmup)
            w2 = work_of(R.run(a.steps, collect=True))
            abl[name] = {"value": round(a.steps * nb / e2 / 1e6, 2),
                         "reused_blocks": w2["n_inter"] + w2["n_intra"],
                         "n_inter": w2["n_inter"], "n_intra": w2["n_intra"],
                         "psnr_db": round(R.quality(a.steps), 4)}
        learned["ablation"] = abl
        out["learned"] = learned
        enc.set_search_mode(lib.SEARCH_EXHAUSTIVE, 8)

    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(spec, frames, a.cpu_frames)
        out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)
    enc.close()


if __name__ == "__main__":
    main()
