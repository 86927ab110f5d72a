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
offsets pass its accept test on 1080p training sequences of other seeds (and
how often an offset passes by the number of pixel indices its word differs
from the block's own), both proposal networks are trained on the GPU, and the
same K steps are encoded with top-4 proposals (the inter ranking adding the
index-similarity scores; exact acceptance).  Prints one JSON line.
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


_T0 = time.perf_counter()


def note(msg: str):
    """A progress line on stderr (the JSON result goes to stdout)."""
    print(f"[bench {time.perf_counter() - _T0:6.1f}s] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the config 3 (4K) and config 4 (12 streams) runs")
    ap.add_argument("--config-steps", type=int, default=4, help="timed frames per config 3 / config 4 run")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--top-k", type=int, default=4)
    ap.add_argument("--no-threshold8", action="store_true", help="skip the learned frontier at threshold 8")
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--workload", default=WORKLOAD)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device-to-device copy of known size (rocprof counter scale)")
    return ap.parse_args()


def _cpu_sample(args):
    """One CPU worker: the reference encoder (oracle/_ref, compiled from its
    sources), or the C oracle where that was not built, on two frames of the
    workload starting at `first` -- an intra frame, then an inter frame
    against it, as the reference's own frame-chunk threading restarts each
    chunk with an intra frame (codec/codec.cpp:1775-1856).  Returns the
    seconds of the intra and of the inter frame (fit + Reencode; PNG decode
    excluded for the reference) and the kind."""
    spec, first = args
    import oracle
    from oracle import SynthCfg
    O = oracle.Oracle()
    cfg = SynthCfg(spec["width"], spec["height"], spec["seed"], spec["n_rects"], spec["vel_mode"], spec["vel"],
                   8, spec["cut_permille"])
    frames = np.stack([O.synth_frame(cfg, first + t) for t in range(2)])
    if oracle.reference_available():
        res = oracle.Reference().encode_sequence(frames, spec["S"], spec["thr"], 2)
        return res[0]["t_fit"] + res[0]["t_reencode"], res[1]["t_fit"] + res[1]["t_reencode"], "reference"
    ts = []
    prev = None
    for f in range(2):
        t0 = time.perf_counter()
        fit = O.fit_frame(frames[f])
        r = O.reencode(frames[f], fit, spec["S"], spec["thr"], f == 0, prev)
        ts.append(time.perf_counter() - t0)
        prev = r["phys"]
    return ts[0], ts[1], "port"


def _cpu_info():
    model, mhz = "unknown", None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model == "unknown":
                model = line.split(":", 1)[1].strip()
            if line.startswith("cpu MHz") and mhz is None:
                mhz = float(line.split(":", 1)[1])
    except OSError:
        pass
    return model, mhz


def cpu_baseline(spec, steps):
    """The reference CPU encoder on this host, on one core and on every core
    this process may use (at most 16, the host share of one GPU), in worker
    processes started before the GPU is touched.  Each worker times one
    two-frame chunk (intra, inter) of the workload; a worker's rate is taken
    at the intra/inter mix of the GPU's timed frames (the first `steps`
    frames of the sequence, intra every intra_interval), and the all-core
    rate is the sum over concurrently running workers."""
    import multiprocessing as mp
    nb = (spec["width"] // 4) * (spec["height"] // 4)
    n_intra = sum(1 for f in range(steps) if (f % spec["frames"]) % spec["intra_interval"] == 0)
    n_inter = steps - n_intra

    def rate(t_i, t_p):
        return steps * nb / (n_intra * t_i + n_inter * t_p) / 1e6

    ctx = mp.get_context("fork")
    with ctx.Pool(1) as pool:
        t_i, t_p, kind = pool.map(_cpu_sample, [(spec, 0)])[0]
    try:
        n_cores = len(os.sched_getaffinity(0))
    except AttributeError:
        n_cores = os.cpu_count() or 1
    workers = max(1, min(16, n_cores))
    t0 = time.perf_counter()
    firsts = [(2 * w) % (spec["frames"] - 1) for w in range(workers)]
    with ctx.Pool(workers) as pool:
        res = pool.map(_cpu_sample, [(spec, f) for f in firsts])
    wall = time.perf_counter() - t0
    model, mhz = _cpu_info()
    one = {"value": rate(t_i, t_p), "unit": "Mblocks/s", "cores": 1, "kind": kind,
           "sample": f"frames 0 (intra) and 1 (inter) of the same 1080p sequence: {t_i:.2f} s + {t_p:.2f} s "
                     f"single-threaded (fit, PNG decode excluded, + Reencode), rate taken at the timed frames' "
                     f"mix of {n_intra} intra and {n_inter} inter frames"}
    allc = {"value": sum(rate(a, b) for a, b, _ in res), "unit": "Mblocks/s", "cores": workers, "kind": kind,
            "sample": f"{workers} worker processes at once, each a two-frame chunk (intra, inter) starting at "
                      f"frames {firsts}; {wall:.1f} s wall; per-worker rates at the timed mix, summed"}
    host = {"cpu_model": model, "cpu_mhz": mhz, "nproc": os.cpu_count(), "usable_cores": n_cores}
    return one, allc, host


KERNELS = ("fit", "inter", "intra_prep", "intra_dense", "intra_combine", "intra_apply",
           "intra_reeval", "intra_tile", "commit", "unique", "net_input", "net_l12", "net_l3",
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
        if len(cells) >= 5 and cells[0] == sym:  # the last two columns: HBM read and write MB
            return round((float(cells[-2]) + float(cells[-1])) * 1e6)
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

    def verify(self, steps, fixture):
        """Compares the frames of the last run (steps frames from frame 0) with
        the reference's per-frame SHA-256 of blocks, motion bytes and unique
        palette; outside the timed region.  Returns (all equal, frames checked)."""
        if fixture is None:
            return None, 0
        n = min(steps, self.n_seq, self.n_out, len(fixture["frames"]))
        ok = True
        for f in range(n):
            o = f % self.n_out
            nu = int(self.d_nuniq.download(np.uint32, 1, offset=4 * o)[0])
            want = fixture["frames"][f]
            ok = ok and _sha(self.d_phys[o].download(np.uint64, self.nb)) == want["phys_sha"]
            ok = ok and _sha(self.d_motion[o].download(np.uint8, 2 * self.nb)) == want["motion_sha"]
            ok = ok and _sha(self.d_unique[o].download(np.uint32, nu)) == want["unique_sha"]
        return ok, n

    def end_to_end(self, steps):
        """The same steps fed from pinned host memory: per frame the RGB8
        frame is copied to the device, encoded, and the blocks, motion bytes,
        unique palette (full capacity) and its count are copied back, all on
        the step's stream.  Returns seconds for `steps` frames."""
        hip = self.hip
        fb = self.W * self.H * 3
        n_in = min(steps, self.n_seq)
        h_in = hip.HostBuffer(n_in * fb)
        h_in.array()[:] = self.frames[:n_in].reshape(-1)
        h_out = hip.HostBuffer(self.nb * (8 + 2 + 4) + 4)
        d_rgb = [hip.Buffer(fb) for _ in range(2)]
        out_bytes = self.nb * 14 + 4

        def step(f, set_intra):
            fi = f % n_in
            slot = f % 2
            o = f % self.n_out
            hip.memcpy_async(d_rgb[slot].ptr, h_in.ptr + fi * fb, fb, hip.hipMemcpyHostToDevice, self.stream)
            self.enc.encode_frame_device(d_rgb[slot], self.W, self.H, set_intra,
                                         None if set_intra else self.d_phys[(f - 1) % self.n_out], self.d_phys[o],
                                         self.d_motion[o], self.d_unique[o], self.d_nuniq.offset(4 * o), None, None,
                                         stream=self.stream.handle)
            base = h_out.ptr
            hip.memcpy_async(base, self.d_phys[o].ptr, 8 * self.nb, hip.hipMemcpyDeviceToHost, self.stream)
            hip.memcpy_async(base + 8 * self.nb, self.d_motion[o].ptr, 2 * self.nb, hip.hipMemcpyDeviceToHost,
                             self.stream)
            hip.memcpy_async(base + 10 * self.nb, self.d_unique[o].ptr, 4 * self.nb, hip.hipMemcpyDeviceToHost,
                             self.stream)
            hip.memcpy_async(base + 14 * self.nb, self.d_nuniq.ptr + 4 * o, 4, hip.hipMemcpyDeviceToHost,
                             self.stream)

        def run(n):
            set_intra, num = True, 1
            for f in range(n):
                if f % n_in == 0:
                    set_intra, num = True, 1
                step(f, set_intra)
                set_intra, num = self.lib.next_intra(set_intra, num, self.spec["intra_interval"])

        run(2)
        self.stream.synchronize()
        t0 = time.perf_counter()
        run(steps)
        self.stream.synchronize()
        t = time.perf_counter() - t0
        h_in.free()
        h_out.free()
        return t, fb + out_bytes

    def end_to_end_overlapped(self, steps):
        """The same host-fed steps with the copies on a second stream: frame
        f + 1's RGB8 upload runs during frame f's encode and frame f's
        downloads during frame f + 1's (input triple-buffered on the device,
        output double-buffered on the host; events order the streams).  The
        upload is queued before the encode call, which polls the fixpoint
        on the host.  Returns seconds for `steps` frames."""
        hip = self.hip
        fb = self.W * self.H * 3
        n_in = min(steps, self.n_seq)
        h_in = hip.HostBuffer(n_in * fb)
        h_in.array()[:] = self.frames[:n_in].reshape(-1)
        out_bytes = self.nb * 14 + 4
        h_out = [hip.HostBuffer(out_bytes) for _ in range(2)]
        d_rgb = [hip.Buffer(fb) for _ in range(3)]
        cs = hip.Stream()
        up = [hip.Event() for _ in range(3)]        # upload into device slot s done
        enc_done = [hip.Event() for _ in range(3)]  # the encode that read slot s done

        def upload(f):
            s = f % 3
            if f >= 3:  # the encode of frame f - 3 read this slot
                hip.stream_wait(cs, enc_done[s])
            hip.memcpy_async(d_rgb[s].ptr, h_in.ptr + (f % n_in) * fb, fb, hip.hipMemcpyHostToDevice, cs)
            up[s].record(cs)

        def run(n):
            set_intra, num = True, 1
            upload(0)
            for f in range(n):
                if f % n_in == 0:
                    set_intra, num = True, 1
                s, o = f % 3, f % self.n_out
                if f + 1 < n:
                    upload(f + 1)
                hip.stream_wait(self.stream, up[s])
                self.enc.encode_frame_device(d_rgb[s], self.W, self.H, set_intra,
                                             None if set_intra else self.d_phys[(f - 1) % self.n_out],
                                             self.d_phys[o], self.d_motion[o], self.d_unique[o],
                                             self.d_nuniq.offset(4 * o), None, None, stream=self.stream.handle)
                enc_done[s].record(self.stream)
                hip.stream_wait(cs, enc_done[s])
                base = h_out[f % 2].ptr
                hip.memcpy_async(base, self.d_phys[o].ptr, 8 * self.nb, hip.hipMemcpyDeviceToHost, cs)
                hip.memcpy_async(base + 8 * self.nb, self.d_motion[o].ptr, 2 * self.nb, hip.hipMemcpyDeviceToHost,
                                 cs)
                hip.memcpy_async(base + 10 * self.nb, self.d_unique[o].ptr, 4 * self.nb,
                                 hip.hipMemcpyDeviceToHost, cs)
                hip.memcpy_async(base + 14 * self.nb, self.d_nuniq.ptr + 4 * o, 4, hip.hipMemcpyDeviceToHost, cs)
                set_intra, num = self.lib.next_intra(set_intra, num, self.spec["intra_interval"])

        run(2)
        cs.synchronize()
        self.stream.synchronize()
        t0 = time.perf_counter()
        run(steps)
        cs.synchronize()
        self.stream.synchronize()
        t = time.perf_counter() - t0
        for e in up + enc_done:
            e.destroy()
        cs.destroy()
        for b in [h_in] + h_out:
            b.free()
        return t

    def quality(self, steps):
        """Mean decoder-side PSNR of the frames the last run wrote (the last
        min(steps, n_out) of them), from mptc_frame_sse on the device."""
        lib = self.lib
        fb = self.W * self.H * 3
        d_sse = self.hip.Buffer(8)
        ps = []
        for f in range(max(0, steps - self.n_out), steps):
            lib.frame_sse(self.d_frames.offset((f % self.n_seq) * fb), self.W, self.H, self.d_phys[f % self.n_out],
                          d_sse, self.stream.handle)
            self.stream.synchronize()
            ps.append(lib.psnr_from_sse(int(d_sse.download(np.uint64, 1)[0]), self.W, self.H))
        d_sse.free()
        return float(np.mean(ps))


def _sha(a) -> str:
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load_fixture(workload):
    p = os.path.join(ROOT, "tests", "golden", f"seq_{workload}.json")
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        return json.load(fh)


def work_of(stats):
    return {"inter_candidates": sum(s["inter_candidates"] for s in stats),
            "intra_candidates": sum(s["intra_candidates"] for s in stats),
            "reeval_candidates": sum(s["reeval_candidates"] for s in stats),
            "exact_inter_candidates": sum(s["exact_inter_candidates"] for s in stats),
            "exact_intra_candidates": sum(s["exact_intra_candidates"] for s in stats),
            "n_inter": sum(s["n_inter"] for s in stats), "n_intra": sum(s["n_intra"] for s in stats),
            "n_unique": sum(s["n_unique"] for s in stats),
            "max_intra_iterations": max(s["intra_iterations"] for s in stats) if stats else 0}


def nearest_scores(net: int) -> np.ndarray:
    """Fixed nearest-first scores over the 16 x 16 offsets of a network's
    window (inter: dx, dy in -8..7; intra: dx in -8..7, dy in 0..15)."""
    L = np.arange(256)
    dx = L % 16 - 8
    dy = L // 16 - 8 if net == 0 else L // 16
    return (-(dx * dx + dy * dy) - 1e-3 * L).astype(np.float32)


def collect_pairs(hip, lib, enc, spec, seeds, n_frames=32):
    """Acceptance sets recorded by the exhaustive search (at the context's
    threshold) on 1080p sequences of the given seeds."""
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
    for seed in seeds:
        cfg = dict(spec)
        cfg["seed"] = seed
        frames = lib.synth_frames(synth_cfg(cfg), 0, n_frames)
        d_fr = hip.Buffer.from_array(frames)
        set_intra, num = True, 1
        for f in range(n_frames):
            rgb = d_fr.offset(f * W * H * 3)
            enc.encode_frame_device(rgb, W, H, set_intra, None if set_intra else d_phys[(f + 1) % 2], d_phys[f % 2],
                                    d_motion, d_unique, d_nu, d_eps, None)
            for net in (0, 1):
                b = bufs[net]
                enc.collect_pairs(net, b["px"], b["tgt"], b["eps"], b["count"], cap)
            set_intra, num = lib.next_intra(set_intra, num, spec["intra_interval"])
        d_fr.free()
    enc.record_pairs(False)
    hip.synchronize()
    for buf in d_phys + [d_motion, d_unique, d_nu, d_eps]:
        buf.free()
    for net in (0, 1):
        bufs[net]["n"] = min(cap, int(bufs[net]["count"].download(np.uint32, 1)[0]))
    info = {"training_frames": n_frames * len(seeds), "training_seeds": list(seeds),
            "pairs_inter": bufs[0]["n"], "pairs_intra": bufs[1]["n"],
            "collect_s": round(time.perf_counter() - t0, 2)}
    return bufs, info


def free_pairs(bufs):
    for b in bufs.values():
        for v in b.values():
            if hasattr(v, "free"):
                v.free()


def train_networks(lib, enc, bufs, train_steps, prior_alpha):
    """GPU training of both networks on recorded pairs.  With prior_alpha > 0
    the logits are learned as corrections on top of fixed base scores
    prior_alpha x the nearest-first ranking: the logit columns of W3 and b3
    start at zero, so training starts from that ranking."""
    info = {"prior_alpha": prior_alpha}
    for net in (0, 1):
        enc.net_init(net, 11 + net)
        if prior_alpha > 0:
            enc.net_set_params(net, lib.zero_logit_layer(enc.net_params(net)))
            enc.net_set_base(net, prior_alpha * nearest_scores(net))
        else:
            enc.net_set_base(net, None)
        b = bufs[net]
        t1 = time.perf_counter()
        losses = enc.net_train(net, b["px"], b["tgt"], b["eps"], b["n"], batch=4096, steps=train_steps,
                               lr=2e-3, seed_weight=1.0, seed=7, lr_final=1e-4)
        key = "inter" if net == 0 else "intra"
        info[f"train_s_{key}"] = round(time.perf_counter() - t1, 2)
        info[f"loss_{key}"] = [round(float(losses[:20].mean()), 4), round(float(losses[-50:].mean()), 4)]
    return info


def dominant_roofline(kernels, work):
    """Roofline of the search kernel with the most time (work-done convention:
    exact evaluations x W_CAND, bounds x W_BOUND)."""
    ev, bd = W_CAND, W_BOUND
    table = {
        "inter": work["inter_candidates"] * ev,
        "intra_dense": work["intra_candidates"] * ev,
        "intra_reeval": work["reeval_candidates"] * ev,
        "pruned_inter": work["inter_candidates"] * bd + work["exact_inter_candidates"] * ev,
        "pruned_intra": work["intra_candidates"] * bd + work["exact_intra_candidates"] * ev,
    }
    names = [k for k in kernels if k in table]
    if not names:
        return None
    dom = max(names, key=lambda k: kernels[k]["total_ms"])
    ach = table[dom] / (kernels[dom]["total_ms"] * 1e-3) / 1e12
    return {"kernel": dom, "bound": "valu", "achieved": round(ach, 3), "peak": VALU_PEAK, "unit": "T lane-op/s",
            "frac": round(ach / VALU_PEAK, 4), "share_of_step": None}


def measure_config3(hip, lib, steps, warmup):
    """3840x2160, S = 32 (64x64-block window), thresholds 0, 8, 32, 128 (the
    reference's floor(SSE/48) units): the dense search kernels.  Per threshold
    the Mblocks/s of `steps` frames from frame 0, the dominant kernel's
    roofline, and the first frames against the reference's fixture."""
    from mptc_amd.workloads import SEQUENCES, synth_cfg
    names = ("c3_t0", "c3_t8", "c3_t32", "c3_t128")
    spec0 = SEQUENCES[names[0]]
    W, H = spec0["width"], spec0["height"]
    nb = (W // 4) * (H // 4)
    frames = lib.synth_frames(synth_cfg(spec0), 0, max(steps, warmup, 2))
    enc = lib.Encoder(W, H, spec0["S"], spec0["thr"])
    R = Runner(hip, lib, enc, spec0, frames, max(steps, warmup, 2))
    out = {}
    # intra fixpoint passes per threshold, the faster of itile_w 0/8/16/32/64
    # measured on MI355X at this size: tile-resolved 64 x 1 passes at
    # threshold 8, plain re-fold passes elsewhere
    itile = {8: 64}
    for name in names:
        spec = SEQUENCES[name]
        enc.set_params(spec["S"], spec["thr"])
        enc.set_tuning(**dict(enc.get_tuning(), itile_w=itile.get(spec["thr"], 0)))
        R.spec = spec
        el = R.timed(steps, warmup)
        ok, n = R.verify(steps, load_fixture(name))
        k, _ = R.kernels(steps)
        w = work_of(R.run(steps, collect=True))
        roof = dominant_roofline(k, w)
        if roof:
            roof["share_of_step"] = round(k[roof["kernel"]]["total_ms"] / sum(v["total_ms"] for v in k.values()), 3)
        out[name] = {"value": round(steps * nb / el / 1e6, 3), "ms_per_step": round(el / steps * 1e3, 3),
                     "threshold": spec["thr"], "bit_exact": ok, "bit_exact_frames_checked": n, "roofline": roof,
                     "kernels": {x: k[x] for x in k if k[x]["total_ms"] > 0.01 * steps},
                     "candidates_per_frame": round((w["inter_candidates"] + w["intra_candidates"]) / steps),
                     "max_intra_iterations": w["max_intra_iterations"], "itile_w": enc.get_tuning()["itile_w"]}
    enc.close()
    return {"workload": f"config 3: {W}x{H} RGB8, S={spec0['S']} (64x64-block window), seeded generator, "
                        f"{steps} timed frames from frame 0 per threshold", "blocks_per_frame": nb, "results": out}


def measure_config4(hip, lib, steps, warmup):
    """Twelve streams of uneven size (3 each of 640x360, 1280x720, 1920x1080,
    3840x2160; S = 8, threshold 8) encoded together: one step encodes the
    next frame of every stream, every kernel of the step covering all twelve
    frames.  The launch shape (graph replay included) is autotuned on the
    batch first.  Reports Mblocks/s summed over the streams and checks the
    first frames of every stream against the reference's fixtures."""
    from mptc_amd.workloads import BATCH_STREAMS, SEQUENCES, synth_cfg
    specs = [SEQUENCES[n] for n in BATCH_STREAMS]
    nbs = [(s["width"] // 4) * (s["height"] // 4) for s in specs]
    total = sum(nbs)
    n_fr = max(steps, warmup, 3)
    enc = lib.Encoder(3840, 2160, specs[0]["S"], specs[0]["thr"], max_batch_frames=len(specs),
                      max_batch_blocks=total)
    st = []
    for s, nb in zip(specs, nbs):
        fr = lib.synth_frames(synth_cfg(s), 0, n_fr)
        st.append(dict(rgb=hip.Buffer.from_array(fr), w=s["width"], h=s["height"], nb=nb,
                       phys=[hip.Buffer(8 * nb) for _ in range(n_fr)], motion=[hip.Buffer(2 * nb) for _ in range(n_fr)],
                       unique=[hip.Buffer(4 * nb) for _ in range(n_fr)], nu=hip.Buffer(4 * n_fr)))
    stream = hip.Stream()

    def ios(f, set_intra):
        return [dict(d_rgb=x["rgb"].offset(f * x["w"] * x["h"] * 3), width=x["w"], height=x["h"],
                     is_intra=set_intra, d_ref=None if set_intra else x["phys"][f - 1], d_phys=x["phys"][f],
                     d_motion=x["motion"][f], d_unique=x["unique"][f], d_n_unique=x["nu"].offset(4 * f))
                for x in st]

    def run(n):
        set_intra, num = True, 1
        for f in range(n):
            enc.encode_batch_device(ios(f, set_intra), stream=stream.handle)
            set_intra, num = lib.next_intra(set_intra, num, specs[0]["intra_interval"])

    run(1)
    stream.synchronize()
    tuning, _ = enc.autotune(ios(1, False), steps=2)
    # the autotuner times one inter step; the fixpoint tile is then chosen on
    # the timed mix (an intra step first), where the intra frames' passes weigh
    best = None
    for itw in (0, 8, 64):
        enc.set_tuning(**dict(tuning, itile_w=itw))
        run(1)
        stream.synchronize()
        t0 = time.perf_counter()
        run(steps)
        stream.synchronize()
        dt = time.perf_counter() - t0
        if best is None or dt < best[0]:
            best = (dt, itw)
    tuning = dict(tuning, itile_w=best[1])
    enc.set_tuning(**tuning)
    run(warmup)
    stream.synchronize()
    hip.synchronize()
    t0 = time.perf_counter()
    run(steps)
    stream.synchronize()
    hip.synchronize()
    el = time.perf_counter() - t0
    ok, checked = True, 0
    for name, x in zip(BATCH_STREAMS, st):
        fx = load_fixture(name)
        for f in range(min(steps, len(fx["frames"]))):
            nu = int(x["nu"].download(np.uint32, 1, offset=4 * f)[0])
            want = fx["frames"][f]
            ok = ok and _sha(x["phys"][f].download(np.uint64, x["nb"])) == want["phys_sha"]
            ok = ok and _sha(x["motion"][f].download(np.uint8, 2 * x["nb"])) == want["motion_sha"]
            ok = ok and _sha(x["unique"][f].download(np.uint32, nu)) == want["unique_sha"]
            checked += 1
    enc.profile_enable(True)
    run(steps)
    stream.synchronize()
    k = {}
    for name in KERNELS:
        n, ms = enc.profile_read(name)
        if n:
            k[name] = {"launches": n, "total_ms": round(ms, 4), "avg_ms": round(ms / n, 5)}
    enc.profile_enable(False)
    stream.destroy()
    enc.close()
    return {"workload": "config 4: 12 streams (3 each of 640x360, 1280x720, 1920x1080, 3840x2160), S=8, "
                        f"threshold 8, one step = the next frame of every stream, {steps} timed steps",
            "blocks_per_step": total, "value": round(steps * total / el / 1e6, 3), "unit": "Mblocks/s",
            "ms_per_step": round(el / steps * 1e3, 3), "tuning": tuning, "bit_exact": ok,
            "bit_exact_frames_checked": checked, "kernels": k}


def learned_point(lib, enc, R, a, k, plain, with_quality=True):
    """Timed learned-mode encode of the workload at top-k = k: throughput and
    what it keeps of the plain search's reuse (plain: that search's value and
    work at the same threshold)."""
    nb = R.nb
    enc.set_search_mode(lib.SEARCH_LEARNED, k)
    t = R.timed(a.steps, a.warmup)
    w = work_of(R.run(a.steps, collect=True))
    v = a.steps * nb / t / 1e6
    reused_plain = plain["work"]["n_inter"] + plain["work"]["n_intra"]
    reused = w["n_inter"] + w["n_intra"]
    pt = {"top_k": k, "value": round(v, 4), "ms_per_step": round(t / a.steps * 1e3, 4),
          "speedup_vs_plain": round(v / plain["value"], 3), "reused_blocks": reused,
          "n_inter": w["n_inter"], "n_intra": w["n_intra"], "n_unique": w["n_unique"],
          "reuse_recall": round(reused / max(1, reused_plain), 4),
          "unique_blocks_delta_pct": round(100.0 * (w["n_unique"] - plain["work"]["n_unique"])
                                           / max(1, plain["work"]["n_unique"]), 3)}
    if with_quality:
        q = R.quality(a.steps)
        pt["psnr_db"] = round(q, 4)
        pt["psnr_delta_db"] = round(q - plain["psnr_db"], 4)
    return pt, w


def learned_frontier(hip, lib, enc, R, a, plain, prior_alpha, sims):
    """k in {4, 8, 16} for: the trained networks with the inter network's
    index-similarity scores (the headline configuration, already set);
    untrained networks with the same scores (the base ranking prior_alpha x
    nearest-first + similarity: the method without a trained network); the
    trained networks with similarity scores on both networks and on neither;
    and the plain nearest-first ranking."""
    fr = {}
    for method in ("trained+inter_similarity", "untrained+inter_similarity", "trained+similarity", "trained",
                   "nearest_prior"):
        if method == "untrained+inter_similarity":
            saved = [enc.net_params(net).copy() for net in (0, 1)]
            for net in (0, 1):
                enc.net_init(net, 99 + net)
                enc.net_set_params(net, lib.zero_logit_layer(enc.net_params(net)))
                enc.net_set_base(net, prior_alpha * nearest_scores(net))
        elif method == "trained+similarity":
            for net in (0, 1):
                enc.net_set_params(net, saved[net])
                enc.net_set_similarity(net, sims[net])
        elif method == "trained":
            for net in (0, 1):
                enc.net_set_similarity(net, None)
        elif method == "nearest_prior":
            for net in (0, 1):
                enc.set_prior(net, nearest_scores(net))
        fr[method] = {str(k): learned_point(lib, enc, R, a, k, plain)[0] for k in (4, 8, 16)}
    for net in (0, 1):
        enc.set_prior(net, None)
    return fr


def similarity_of(lib, enc):
    """Index-similarity scores from the statistics of the last recording."""
    out, sims = {}, {}
    for net in (0, 1):
        tries, acc = enc.similarity_counts(net)
        sims[net] = lib.similarity_scores(tries, acc)
        out["inter" if net == 0 else "intra"] = {"tries": tries.tolist(), "accepts": acc.tolist(),
                                                 "scores": [round(float(v), 3) for v in sims[net]]}
    return sims, out


PRIOR_ALPHA = 0.25  # base scores = 0.25 x nearest-first ranking (corrections learned on top)


def measure_learned(hip, lib, enc, R, spec, a, value, work, psnr_plain, tuning):
    """The learned-proposal path at the workload's threshold (trained on pairs
    of other seeds, launch shape autotuned in learned mode), its kernels and
    MFMA roofline, and the k frontier against the nearest-first ranking and
    untrained networks; then the same at threshold 8 (the c2_t8 sequence:
    the same frames), where reuse costs PSNR."""
    nb = R.nb
    plain = {"value": value, "work": work, "psnr_db": psnr_plain}
    note("learned: collecting pairs and training")
    bufs, info = collect_pairs(hip, lib, enc, spec, TRAIN_SEEDS)
    sims, info["similarity"] = similarity_of(lib, enc)
    info.update(train_networks(lib, enc, bufs, a.train_steps, PRIOR_ALPHA))
    free_pairs(bufs)
    enc.net_set_similarity(0, sims[0])  # the inter network's ranking gains the similarity scores
    learned = {"top_k": a.top_k, "training": info,
               "proposals": "trained networks; inter ranking + index-similarity scores"}
    enc.set_search_mode(lib.SEARCH_LEARNED, a.top_k)
    if not a.no_autotune:
        ltune, _ = R.autotune()
        learned["tuning"] = ltune
    pt, lw = learned_point(lib, enc, R, a, a.top_k, plain)
    learned.update(pt)
    lk, lms_ev = R.kernels(a.steps)
    learned.update({"kernels": lk, "ms_per_step_with_kernel_events": round(lms_ev, 4), "work": lw})
    # MFMA roofline of the network's layer-1+2 kernel (bf16 dense)
    if "net_l12" in lk:
        ach = nb * NET_FLOP_L12 / (lk["net_l12"]["avg_ms"] * 1e-3) / 1e12
        learned["roofline_net_l12"] = {"bound": "mfma", "achieved": round(ach, 2), "peak": MFMA_BF16_PEAK,
                                       "unit": "TFLOP/s", "frac": round(ach / MFMA_BF16_PEAK, 4),
                                       "work_per_launch": f"{nb} blocks x {NET_FLOP_L12} flop"}
    if "net_l3" in lk:
        ach = nb * NET_FLOP_L3 / (lk["net_l3"]["avg_ms"] * 1e-3) / 1e12
        learned["roofline_net_l3"] = {"bound": "mfma", "achieved": round(ach, 2), "peak": MFMA_BF16_PEAK,
                                      "unit": "TFLOP/s", "frac": round(ach / MFMA_BF16_PEAK, 4),
                                      "work_per_launch": f"{nb} blocks x {NET_FLOP_L3} flop (logits) + top-k"}
    note("learned: k frontier")
    learned["frontier"] = learned_frontier(hip, lib, enc, R, a, plain, PRIOR_ALPHA, sims)
    # against the method without the network and the plain nearest-first ranking, at the same k
    for ref in ("untrained+inter_similarity", "nearest_prior"):
        p = learned["frontier"][ref][str(a.top_k)]
        learned[f"vs_{ref}"] = {"speedup": round(learned["value"] / p["value"], 3),
                                "reused_blocks_delta_pct": round(100.0 * (learned["reused_blocks"] - p["reused_blocks"])
                                                                 / max(1, p["reused_blocks"]), 3),
                                "unique_blocks_delta_pct": round(100.0 * (lw["n_unique"] - p["n_unique"])
                                                                 / max(1, p["n_unique"]), 3)}
    if a.no_threshold8:
        return learned
    note("learned: threshold 8")
    spec8 = dict(spec)
    spec8["thr"] = 8
    enc.set_search_mode(lib.SEARCH_EXHAUSTIVE, 8)
    enc.set_params(spec["S"], 8)
    enc.set_tuning(**tuning)
    t8 = R.timed(a.steps, a.warmup)
    w8 = work_of(R.run(a.steps, collect=True))
    plain8 = {"value": a.steps * nb / t8 / 1e6, "work": w8, "psnr_db": R.quality(a.steps)}
    bufs, info8 = collect_pairs(hip, lib, enc, spec8, TRAIN_SEEDS[:4])
    sims8, info8["similarity"] = similarity_of(lib, enc)
    info8.update(train_networks(lib, enc, bufs, a.train_steps, PRIOR_ALPHA))
    free_pairs(bufs)
    enc.net_set_similarity(0, sims8[0])
    if "tuning" in learned:
        enc.set_tuning(**learned["tuning"])
    learned["threshold8"] = {
        "plain": {"value": round(plain8["value"], 4), "reused_blocks": w8["n_inter"] + w8["n_intra"],
                  "n_unique": w8["n_unique"], "psnr_db": round(plain8["psnr_db"], 4)},
        "training": info8,
        "frontier": learned_frontier(hip, lib, enc, R, a, plain8, PRIOR_ALPHA, sims8)}
    enc.set_params(spec["S"], spec["thr"])
    return learned


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    from mptc_amd.workloads import SEQUENCES, synth_cfg

    # the CPU baseline first, in worker processes forked before any GPU work
    cpu = None
    if not a.no_cpu_baseline and not a.copy_calibration:
        note("CPU baseline")
        cpu = cpu_baseline(SEQUENCES[a.workload], a.steps)

    from mptc_amd import hip, lib
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
    note("plain path: autotune, timed steps, checks")
    tuning, tune_ms = R.autotune() if not a.no_autotune else (enc.get_tuning(), None)
    elapsed = R.timed(a.steps, a.warmup)
    value = a.steps * nb / elapsed / 1e6
    # outside the timed region: the timed frames against the reference's fixture
    bit_exact, n_checked = R.verify(a.steps, load_fixture(a.workload))
    # the same frames fed from and returned to pinned host memory
    e2e_s, e2e_bytes = R.end_to_end(a.steps)
    # the same with the copies overlapping the encodes on a second stream, and
    # its frames checked against the fixture too
    e2e_ov_s = R.end_to_end_overlapped(a.steps)
    e2e_ov_exact, _ = R.verify(a.steps, load_fixture(a.workload))
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
                                          for n, what, w in parts),
            "frac_is": "work done: every candidate the kernel bounds x W_BOUND + every exact evaluation x W_CAND"}
    # the same launch time against two other work conventions: every window
    # candidate at W_CAND (the full-window count, which pruning skips most of)
    # and the exact evaluations alone
    cand_key = {"inter": "inter_candidates", "pruned_inter": "inter_candidates", "intra_dense": "intra_candidates",
                "pruned_intra": "intra_candidates", "intra_reeval": "reeval_candidates"}[dom]
    exact_key = {"pruned_inter": "exact_inter_candidates", "pruned_intra": "exact_intra_candidates"}.get(dom, cand_key)
    per_s = n_launch * kernels[dom]["avg_ms"] * 1e-3 * 1e12
    roof["frac_full_window"] = round(work[cand_key] * ev / per_s / VALU_PEAK, 4)
    roof["frac_exact_evaluations"] = round(work[exact_key] * ev / per_s / VALU_PEAK, 4)

    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "Mblocks/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(elapsed / a.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+i32",
        "data": "synthetic (seeded generator, mptc_amd/csrc/framegen.h)",
        "config": {"workload": f"encode ({a.workload}): {spec['frames']} frames {W}x{H} RGB8, "
                               f"S={spec['S']} ({2 * spec['S']}-block window), threshold {spec['thr']}, "
                               f"intra every {spec['intra_interval']} frames; value = plain search, "
                               f"bit-exact to the reference", "blocks_per_frame": nb,
                   "tuning": tuning},
        "roofline": roof,
        "kernels": kernels,
        "ms_per_step_with_kernel_events": round(ms_ev, 4),
        "work": work,
        "psnr_db": round(psnr_plain, 4),
        "bit_exact": bit_exact, "bit_exact_frames_checked": n_checked,
        "end_to_end": {"value": round(a.steps * nb / e2e_s / 1e6, 4), "unit": "Mblocks/s",
                       "ms_per_step": round(e2e_s / a.steps * 1e3, 4),
                       "note": f"same steps with the RGB8 frame copied in from pinned host memory and blocks, "
                               f"motion, unique palette and count copied back per step "
                               f"({e2e_bytes} bytes over PCIe per frame)"},
        "end_to_end_overlapped": {"value": round(a.steps * nb / e2e_ov_s / 1e6, 4), "unit": "Mblocks/s",
                                  "ms_per_step": round(e2e_ov_s / a.steps * 1e3, 4), "bit_exact": e2e_ov_exact,
                                  "note": "the same host-fed steps with the copies on a second stream: frame f+1's "
                                          "upload and frame f's downloads overlap the encodes"},
    }

    if not a.no_learned:
        note("learned path")
        out["learned"] = measure_learned(hip, lib, enc, R, spec, a, value, work, psnr_plain, tuning)
        enc.set_search_mode(lib.SEARCH_EXHAUSTIVE, 8)
        enc.set_tuning(**tuning)

    if not a.no_configs:
        note("config 3 (4K, S=32, threshold sweep)")
        out["config3"] = measure_config3(hip, lib, steps=a.config_steps, warmup=1)
        note("config 4 (12 uneven streams)")
        out["config4"] = measure_config4(hip, lib, steps=a.config_steps, warmup=1)

    if cpu is not None:
        one, allc, host = cpu
        out["cpu_baseline"] = allc
        out["cpu_baseline_1core"] = one
        out["cpu_host"] = host
        out["gpu_over_cpu"] = round(value / allc["value"], 1)
        out["gpu_over_cpu_1core"] = round(value / one["value"], 1)
    note("done")
    print(json.dumps(out), flush=True)
    enc.close()


if __name__ == "__main__":
    main()
