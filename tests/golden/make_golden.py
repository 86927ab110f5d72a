"""Writes the golden fixtures under tests/golden/ from the reference encoder
compiled from its own sources (oracle/build_ref.sh -> oracle/_ref/libmptc_ref.so).

Run in the build container (where the reference sources are), e.g.
    python tests/golden/make_golden.py units
    python tests/golden/make_golden.py seq c1_s8_t0
    python tests/golden/make_golden.py all-seq      # every sequence, one process each
    python tests/golden/make_golden.py psnr         # the reference's PSNR of leading frames
    python tests/golden/make_golden.py bitstream    # the reference's .mptc streams
The fixtures are data only: inputs are regenerated from the seeded frame
generator (mptc_amd/csrc/framegen.h), outputs are the reference's.
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402
from tests.golden.workloads import SEQUENCES, synth_cfg  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def unit_vectors(R: "oracle.Reference", O: "oracle.Oracle") -> None:
    rng = np.random.default_rng(20261202)
    blocks = []
    # random, near-constant, two-colour, constant, gradient and extreme blocks
    for t in range(3000):
        k = t % 6
        if k == 0:
            px = rng.integers(0, 256, 48)
        elif k == 1:
            px = np.tile(rng.integers(0, 256, 3), 16) + rng.integers(-4, 5, 48)
        elif k == 2:
            a, b = rng.integers(0, 256, 3), rng.integers(0, 256, 3)
            px = np.array([a if rng.random() < 0.5 else b for _ in range(16)]).reshape(48)
        elif k == 3:
            px = np.tile(rng.integers(0, 256, 3), 16)
        elif k == 4:
            base = rng.integers(0, 256, 3)
            g = rng.integers(-40, 41, (2, 3))
            px = np.array([base + g[0] * (i % 4) + g[1] * (i // 4) for i in range(16)]).reshape(48)
        else:
            px = rng.choice(np.array([0, 1, 127, 128, 254, 255]), 48)
        blocks.append(np.clip(px, 0, 255).astype(np.uint8))
    blocks = np.stack(blocks)
    fit = np.array([R.fit_block(b) for b in blocks], dtype=np.uint64)

    # candidates: random words, the own word, single-index words, words of neighbours
    cand_px, cand_own, cand_word = [], [], []
    singles = [0x00000000, 0x55555555, 0xAAAAAAAA, 0xFFFFFFFF]
    for t in range(6000):
        b = blocks[rng.integers(0, len(blocks))]
        own = int(fit[rng.integers(0, len(blocks))]) if t % 11 == 0 else R.fit_block(b)
        k = t % 5
        if k == 0:
            c = own >> 32
        elif k == 1:
            c = singles[t % 4]
        elif k == 2:
            c = int(fit[rng.integers(0, len(blocks))]) >> 32
        else:
            c = int(rng.integers(0, 2**32))
        cand_px.append(b)
        cand_own.append(own)
        cand_word.append(c)
    cand_px = np.stack(cand_px)
    cand_own = np.array(cand_own, dtype=np.uint64)
    cand_word = np.array(cand_word, dtype=np.uint32)
    valid = np.zeros(len(cand_word), np.uint8)
    err = np.zeros(len(cand_word), np.int32)
    rea = np.zeros(len(cand_word), np.uint8)
    eps = np.zeros((len(cand_word), 6), np.uint8)
    for i in range(len(cand_word)):
        ok, e, r, ep = R.eval_candidate(int(cand_own[i]), cand_px[i], int(cand_word[i]))
        valid[i] = ok
        if ok:
            err[i], rea[i], eps[i] = e, r, ep
    five = np.array([R.L.ref_to_five_bits(x) for x in range(256)], np.uint8)
    six = np.array([R.L.ref_to_six_bits(x) for x in range(256)], np.uint8)
    np.savez_compressed(os.path.join(HERE, "unit_vectors.npz"), fit_px=blocks, fit_out=fit,
                        cand_px=cand_px, cand_own=cand_own, cand_word=cand_word,
                        cand_valid=valid, cand_err=err, cand_reassigned=rea, cand_eps=eps,
                        to_five=five, to_six=six)
    print("unit vectors:", len(blocks), "fit,", len(cand_word), "candidates,",
          int(valid.sum()), "valid")


def sequence(name: str) -> None:
    spec = SEQUENCES[name]
    O = oracle.Oracle()
    R = oracle.Reference()
    cfg = synth_cfg(spec)
    n = spec["golden_frames"]
    frames = O.synth_frames(cfg, n)
    t0 = time.time()
    res = R.encode_sequence(frames, spec["S"], spec["thr"], spec["intra_interval"])
    wall = time.time() - t0
    out = dict(name=name, spec=spec, frames=[], wall_s=wall)
    for f, r in enumerate(res):
        kinds = np.zeros(3, int)
        mot = r["motion"].reshape(-1, 2)
        uniq = (mot[:, 0] == 255) & (mot[:, 1] == 255)
        inter = (mot[:, 0] & 0x80).astype(bool) & ~uniq
        kinds[0] = int(uniq.sum())
        kinds[2] = int(inter.sum())
        kinds[1] = len(mot) - kinds[0] - kinds[2]
        out["frames"].append(dict(
            frame=f, input_sha=sha(frames[f]), phys_sha=sha(r["phys"]),
            motion_sha=sha(r["motion"]), unique_sha=sha(r["unique"]), eps_sha=sha(r["eps"]),
            n_unique=int(len(r["unique"])), kinds=[int(x) for x in kinds],
            psnr_logical=O.psnr_logical(frames[f], r["phys"], r["eps"]),
            psnr_physical=O.psnr_physical(frames[f], r["phys"]),
            t_fit=r["t_fit"], t_reencode=r["t_reencode"]))
    if spec.get("full_arrays"):
        np.savez_compressed(os.path.join(HERE, f"seq_{name}.npz"),
                            phys=np.stack([r["phys"] for r in res]),
                            motion=np.stack([r["motion"] for r in res]),
                            eps=np.stack([r["eps"] for r in res]),
                            unique=np.concatenate([r["unique"] for r in res]),
                            n_unique=np.array([len(r["unique"]) for r in res]))
    with open(os.path.join(HERE, f"seq_{name}.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(name, "done in %.1fs" % wall)


# Sequences whose leading frames get the reference's own PSNR values, and how
# many frames of each (the 1080p ones take ~10 s per frame single-threaded).
PSNR_FRAMES = {"c1_s8_t0": 8, "c1_s8_t8": 8, "c1_s16_t0": 8, "c2_t0": 3, "c2_t8": 3}


def psnr_fixture() -> None:
    """psnr_ref.json: DXTImage::PSNR (codec/dxt_image.cpp:363-383) of the
    reference's own encodes, encoder side (its logical blocks right after
    Reencode) and decoder side (after SetLogicalBlocks, codec/dxt_image.cpp:452),
    for the leading frames of PSNR_FRAMES, with the SHA of each frame's blocks
    so that a test can tie the values to the sequence fixtures."""
    O = oracle.Oracle()
    R = oracle.Reference()
    out = {}
    for name, n in PSNR_FRAMES.items():
        spec = SEQUENCES[name]
        frames = O.synth_frames(synth_cfg(spec), n)
        res = R.encode_sequence(frames, spec["S"], spec["thr"], spec["intra_interval"], psnr=True)
        out[name] = [dict(frame=f, phys_sha=sha(r["phys"]), psnr_enc=r["psnr_enc"], psnr_dec=r["psnr_dec"])
                     for f, r in enumerate(res)]
        print(name, [round(r["psnr_dec"], 4) for r in res], flush=True)
    with open(os.path.join(HERE, "psnr_ref.json"), "w") as fh:
        json.dump(out, fh, indent=1)


# Bitstream cases: (name, sequence, frames, unique_interval).  The 512x256
# case has a 128x64 block grid (two wavelet blocks side by side); the last
# case leaves two frames after its last whole group.
BITSTREAMS = [("bs_c1_s8_t0_u8", "c1_s8_t0", 8, 8), ("bs_c1_s8_t8_u4", "c1_s8_t8", 8, 4),
              ("bs_c1_s16_t0_u3", "c1_s16_t0", 8, 3), ("bs_w512_t8_u2", "w512_t8", 4, 2)]


def _bitstream_one(case: str) -> dict:
    """Runs the reference's CompressMultiUnique (codec/codec.cpp:1307-1532) on
    the PNG frames of one case in a scratch directory; returns the stream's
    SHA-256, size and header.  One process per case: the reference keeps the
    header's maxima in process globals."""
    import ctypes
    import tempfile
    name, seq, n, ui = next(c for c in BITSTREAMS if c[0] == case)
    spec = SEQUENCES[seq]
    O = oracle.Oracle()
    frames = O.synth_frames(synth_cfg(spec), n)
    L = ctypes.CDLL(oracle.REF_CODEC_SO)
    with tempfile.TemporaryDirectory() as tmp:
        fdir = os.path.join(tmp, "frames")
        os.makedirs(fdir)
        for i, f in enumerate(frames):
            a = np.ascontiguousarray(f)
            rc = L.ref_write_png(os.path.join(fdir, f"frame{i:03d}.png").encode(), a.ctypes.data_as(ctypes.c_void_p),
                                 spec["width"], spec["height"])
            assert rc == 0
        out = os.path.join(tmp, "out.mptc")
        cwd = os.getcwd()
        os.chdir(tmp)  # the reference also writes out.txt into the working directory
        try:
            L.ref_compress_dir(fdir.encode(), out.encode(), spec["S"], spec["thr"], spec["intra_interval"], ui)
        finally:
            os.chdir(cwd)
        data = open(out, "rb").read()
    return dict(name=name, sequence=seq, frames=n, unique_interval=ui, size=len(data), sha256=sha(np.frombuffer(data, np.uint8)),
                header=data[:34].hex())


def bitstream_fixture() -> None:
    """bitstreams.json: the reference encoder's .mptc streams of the cases in
    BITSTREAMS (SHA-256, size, header)."""
    out = []
    for case in BITSTREAMS:
        r = subprocess.run([sys.executable, __file__, "bitstream-one", case[0]], capture_output=True, text=True,
                           check=True)
        out.append(json.loads(r.stdout.strip().splitlines()[-1]))
        print(out[-1]["name"], out[-1]["size"], flush=True)
    with open(os.path.join(HERE, "bitstreams.json"), "w") as fh:
        json.dump(out, fh, indent=1)


def main(argv):
    oracle.build_oracle()
    oracle.build_ref()
    what = argv[1] if len(argv) > 1 else "units"
    if what == "units":
        unit_vectors(oracle.Reference(), oracle.Oracle())
    elif what == "psnr":
        psnr_fixture()
    elif what == "bitstream":
        bitstream_fixture()
    elif what == "bitstream-one":
        import contextlib
        import io
        buf = io.StringIO()
        with contextlib.redirect_stdout(buf):
            r = _bitstream_one(argv[2])
        print(json.dumps(r))
    elif what == "seq":
        for name in argv[2:]:
            sequence(name)
    elif what == "all-seq":
        procs = [subprocess.Popen([sys.executable, __file__, "seq", name]) for name in SEQUENCES]
        sys.exit(max(p.wait() for p in procs))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv)
