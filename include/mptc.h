/*
 * mptc.h -- C-ABI of libmptc_hip.so, the MI355X (gfx950) implementation of
 * MPTC's per-frame encode hot path: the DXT1 endpoint fit of every 4x4 block,
 * the index-block reuse search over the inter and intra windows with its
 * error threshold, and the least-squares endpoint re-fit of reused indices.
 *
 * Plain C types only: pointers, sizes, status codes.  Every call returns 0
 * (MPTC_OK) or a negative status; mptc_last_error() gives the message for the
 * calling thread.  Nothing throws across the boundary.  A context owns one HIP
 * stream and the device work buffers for frames up to its configured maximum
 * size; use one context per host thread.  Output is bit-identical to the
 * reference encoder for any launch configuration.
 *
 * Block layout (codec/dxt_image.h:16-23): one uint64_t per 4x4 block in raster
 * order, bits 0-15 ep1 (RGB565), 16-31 ep2, 32-63 the 2-bit index word with
 * pixel i = row*4 + col at bits 2i..2i+1.  Motion: two bytes (x, y) per block;
 * inter reuse sets bit 7 of both, unique blocks are (255, 255)
 * (codec/dxt_image.cpp:877, 892-894, 917, 953).
 */
#ifndef MPTC_H
#define MPTC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MPTC_OK = 0,
  MPTC_EINVAL = -1, /* bad argument: null pointer, size, search area, threshold */
  MPTC_EHIP = -2,   /* HIP runtime error */
  MPTC_ENOMEM = -3, /* device allocation failed */
  MPTC_ESTATE = -4  /* call not valid in the context's current state */
};

typedef struct mptc_ctx mptc_ctx;

typedef struct {
  int32_t device;           /* HIP device ordinal */
  int32_t max_width;        /* largest frame width in pixels (multiple of 4) */
  int32_t max_height;       /* largest frame height in pixels (multiple of 4) */
  int32_t search_area;      /* S in 1..63: window side 2S blocks */
  int32_t err_threshold;    /* reuse accepted iff error increase <= threshold,
                               in floor(SSE/48) units, -32768..30000 */
  int32_t max_batch_frames; /* frames per mptc_encode_batch_device call, 1..64 (0 = 1) */
  int32_t max_batch_blocks; /* 4x4 blocks summed over a batch
                               (0 = max_batch_frames * (max_width/4) * (max_height/4)) */
} mptc_config;

/* Launch shape of the step, chosen by mptc_autotune or set directly.  Every
 * setting gives bit-identical output. */
typedef struct {
  int32_t wg_size;    /* workgroup size of the per-block kernels: 64, 128, 256 */
  int32_t tile_w;     /* blocks per wave row in the search kernels: 64, 32, 16, 8
                         (the wave covers tile_w x 64/tile_w blocks) */
  int32_t pass_batch; /* intra fixpoint passes queued per convergence check, 1..16 */
  int32_t use_graph;  /* replay the step as captured HIP graphs */
  int32_t win_chunks; /* window tile: the 2S window rows of a block are split into
                         this many runs scanned by separate waves, 1..8 */
  int32_t prune;      /* 1: for S <= 8 and thresholds 0..4, evaluate exactly only the
                         candidates an error lower bound cannot rule out (same output);
                         0: evaluate every candidate */
  int32_t itile_w;    /* intra fixpoint passes: 8, 16, 32 or 64: a wave settles the reuse
                         chains inside a tile of itile_w x 64/itile_w blocks per pass;
                         0: exhaustive search re-folds flagged blocks once per pass, learned
                         search uses 8 x 8 tiles (same output for every value) */
} mptc_tuning;

/* One frame of a batched step: device pointers, layouts as mptc_reencode. */
typedef struct {
  const uint8_t* d_rgb;
  int32_t width, height;
  int32_t is_intra;
  const uint64_t* d_ref_phys; /* this stream's previous output (ignored if is_intra) */
  uint64_t* d_phys_out;
  uint8_t* d_motion_out;
  uint32_t* d_unique_out;
  uint32_t* d_n_unique_out;
  uint8_t* d_eps_out;         /* optional */
  uint8_t* d_kind_out;        /* optional */
} mptc_frame_io;

/* Counters of the last encoded frame. */
typedef struct {
  int64_t n_blocks;
  int64_t n_inter;          /* blocks reusing a previous-frame index word */
  int64_t n_intra;          /* blocks reusing an index word of this frame */
  int64_t n_unique;         /* blocks keeping their own index word */
  int64_t inter_candidates; /* in-frame inter window positions searched */
  int64_t intra_candidates; /* in-frame intra window positions of searched blocks */
  int64_t reeval_candidates;/* intra positions re-evaluated after a word changed */
  int32_t intra_iterations; /* passes until the intra words reached their fixpoint */
  int32_t is_intra;
  int64_t exact_inter_candidates; /* with pruning: candidates evaluated exactly (the rest */
  int64_t exact_intra_candidates; /* were ruled out by the error bound); 0 without pruning */
} mptc_frame_stats;

/* Context lifetime.  Replaces the process-global state the reference keeps in
 * DXTImage's static members and vErrThreshold (codec/dxt_image.cpp:25,
 * 357-360, 393-394). */
int32_t mptc_create(const mptc_config* cfg, mptc_ctx** out);
void mptc_destroy(mptc_ctx* ctx);
const char* mptc_last_error(void);
int32_t mptc_set_params(mptc_ctx* ctx, int32_t search_area, int32_t err_threshold);

/* The DXTImage constructor's fit loop (codec/dxt_image.cpp:419-431,
 * CompressRGB :123-149 -> stb_compress_dxt_block HIGHQUAL): host buffers.
 * rgb: width*height*3 bytes; phys_out: (width/4)*(height/4) blocks. */
int32_t mptc_fit(mptc_ctx* ctx, const uint8_t* rgb, int32_t width, int32_t height,
                 uint64_t* phys_out);

/* DXTImage::Reencode (codec/dxt_image.cpp:868-957): host buffers.
 *  phys_inout  in: this frame's fitted blocks (as from mptc_fit); out: re-encoded
 *  ref_phys    previous frame's re-encoded blocks; ignored (may be NULL) if is_intra
 *  motion_out  2 bytes per block
 *  unique_out  the unique palette (own index words of blocks not reused), raster order;
 *              capacity one word per block; *n_unique_out receives the count
 *  eps_out     optional (NULL): 6 bytes per block, logical ep1 rgb then ep2 rgb after
 *              re-encode (the re-fit's unquantised-grid values, as _logical_blocks)
 *  kind_out    optional (NULL): 0 unique, 1 intra reuse, 2 inter reuse */
int32_t mptc_reencode(mptc_ctx* ctx, const uint8_t* rgb, int32_t width, int32_t height,
                      int32_t is_intra, const uint64_t* ref_phys, uint64_t* phys_inout,
                      uint8_t* motion_out, uint32_t* unique_out, uint32_t* n_unique_out,
                      uint8_t* eps_out, uint8_t* kind_out);

/* Fused, device-resident per-frame step: fit + Reencode of one frame whose
 * RGB8 pixels are already in device memory.  All pointers are device
 * pointers; d_ref_phys is the previous call's d_phys_out (ignored if is_intra).
 * d_eps_out and d_kind_out may be NULL.  Work is queued on `hip_stream`
 * (a hipStream_t; NULL = the context's stream); the call returns once the
 * frame's work is queued and its intra fixpoint is reached (it synchronises
 * with the stream to test convergence).  Same layouts as mptc_reencode.
 * Successive calls on one context may use different streams: a call on
 * another stream than the previous call's first waits for that call's work
 * (the context's buffers are shared), so outputs never race. */
int32_t mptc_encode_frame_device(mptc_ctx* ctx, const uint8_t* d_rgb, int32_t width,
                                 int32_t height, int32_t is_intra, const uint64_t* d_ref_phys,
                                 uint64_t* d_phys_out, uint8_t* d_motion_out,
                                 uint32_t* d_unique_out, uint32_t* d_n_unique_out,
                                 uint8_t* d_eps_out, uint8_t* d_kind_out, void* hip_stream);

/* One step of several independent streams (uneven sizes allowed): frame i of
 * the batch is encoded against frames[i].d_ref_phys exactly as
 * mptc_encode_frame_device would, but all frames share every kernel launch
 * of the step.  No reference counterpart (the reference encodes one stream per
 * process); each frame's output equals DXTImage::Reencode on that frame. */
int32_t mptc_encode_batch_device(mptc_ctx* ctx, int32_t n_frames, const mptc_frame_io* frames,
                                 void* hip_stream);

/* Launch-shape tuning.  mptc_autotune times every candidate shape (wg_size x
 * tile_w x win_chunks x use_graph, prune 0/1 where the pruned search applies;
 * then itile_w 0/8/16/32/64 on the winner) on
 * `steps` encodes of the given batch (outputs are overwritten), keeps the
 * fastest and reports it. */
int32_t mptc_get_tuning(mptc_ctx* ctx, mptc_tuning* out);
int32_t mptc_set_tuning(mptc_ctx* ctx, const mptc_tuning* t);
int32_t mptc_autotune(mptc_ctx* ctx, int32_t n_frames, const mptc_frame_io* frames, int32_t steps,
                      mptc_tuning* best, double* best_ms_per_step);

/* Waits for all work queued on the context's own stream (the stream used when
 * a device entry point is given hip_stream = NULL) and for the last call's
 * work on whatever stream it used. */
int32_t mptc_synchronize(mptc_ctx* ctx);

/* Counters of the last step encoded with this context, summed over its
 * frames (synchronises). */
int32_t mptc_get_stats(mptc_ctx* ctx, mptc_frame_stats* out);

/* Per-kernel timing with HIP events recorded on the launch stream.
 * mptc_profile_enable(ctx, 1) starts recording; mptc_profile_read returns the
 * accumulated launch count and milliseconds of kernel `name` ("fit": the step's
 * first kernel: counters cleared, fit, reference tables, network input),
 * "inter", "intra_prep" (with the inter decisions of an exhaustive inter
 * search), "intra_dense", "intra_combine",
 * "intra_apply", "intra_reeval", "intra_tile", "commit", "unique";
 * learned search: "net_input", "net_l12", "net_l3" (inference), "propose_inter",
 * "top_intra", "lintra_pass"; "net_forward": training steps (the fused
 * forward/loss/backward, weight-gradient and reduction/Adam launches); pruned search:
 * "pruned_inter", "pruned_intra", "fold_intra" (only when the window is split
 * into a number of chunks other than 1, 2 or 4; otherwise the intra search
 * takes those decisions)) since it was enabled
 * (synchronises).  While enabled, steps run as individual launches, not graphs. */
int32_t mptc_profile_enable(mptc_ctx* ctx, int32_t enable);
int32_t mptc_profile_read(mptc_ctx* ctx, const char* name, int64_t* launches, double* total_ms);

/* ---------------------------------------------------------------------------
 * Learned proposals (no counterpart in the reference).  Two networks of one
 * shape, MPTC_NET_INTER for the inter window and MPTC_NET_INTRA for the causal
 * intra window: 48 inputs (a block's pixels, pixel-major RGB, / 255) -> 256
 * ReLU -> 256 ReLU -> 262 outputs (256 logits over the 16x16 offsets of the
 * window, then 6 endpoint seeds: ep1 rgb, ep2 rgb / 255).  BF16 weights and
 * activations, FP32 accumulation, FP32 biases and master weights.
 *
 * Logit index of a window offset (dx = i - bx, in -8..7):
 *   inter: (dy + 8) * 16 + dx + 8 with dy = j - by in -8..7
 *   intra: dy * 16 + dx + 8        with dy = by - j in 0..15 (dy = 0: dx < 0 only)
 *
 * Parameters are one fp32 vector of MPTC_NET_PARAMS values: W1 [48][256],
 * b1 [256], W2 [256][256], b2 [256], W3 [256][262], b3 [262] (W[k][n]: input
 * k, output n).  Training loss: -log of the probability the softmax of the
 * logits gives a target set of offsets (the offsets the exhaustive search
 * found passing; one offset = cross-entropy) + seed_weight * mean squared
 * error of the seeds against the endpoints / 255, averaged over the batch;
 * Adam.
 * Proposals: the top_k admitted offsets by score (scores compared as fp32
 * truncated to 24 bits, ties to the lower logit index).
 * ------------------------------------------------------------------------- */
#define MPTC_NET_INTER 0
#define MPTC_NET_INTRA 1
#define MPTC_NET_IN 48
#define MPTC_NET_HIDDEN 256
#define MPTC_NET_OUT 262
#define MPTC_NET_LOGITS 256
#define MPTC_NET_PARAMS 145670

/* Search modes (mptc_set_search_mode). */
#define MPTC_SEARCH_EXHAUSTIVE 0 /* the reference's search: bit-exact */
#define MPTC_SEARCH_LEARNED_INTER 1 /* inter: top-k proposals checked exactly;
                                       intra: exhaustive for every block the
                                       proposals leave */
#define MPTC_SEARCH_LEARNED 2    /* inter and intra: top-k proposals checked
                                    exactly; a block none passes stays unique */

typedef struct {
  int32_t batch;       /* rows per step */
  int32_t steps;       /* optimizer steps */
  float lr;            /* Adam step size */
  float seed_weight;   /* weight of the endpoint-seed loss */
  uint64_t seed;       /* shuffling seed */
  float beta1, beta2, eps; /* Adam constants (0 = 0.9, 0.999, 1e-8) */
  float lr_final;      /* step size of the last step, cosine decay (0 = constant lr) */
} mptc_train_cfg;

/* He-normal weights from `seed`, zero biases, Adam state cleared. */
int32_t mptc_net_init(mptc_ctx* ctx, int32_t net, uint64_t seed);
/* Host copies of the fp32 parameter vector; set clears the Adam state. */
int32_t mptc_net_get_params(mptc_ctx* ctx, int32_t net, float* params_out);
int32_t mptc_net_set_params(mptc_ctx* ctx, int32_t net, const float* params);
/* Forward pass over n blocks: d_px n x 48 bytes, d_out n x 262 fp32 (device). */
int32_t mptc_net_forward(mptc_ctx* ctx, int32_t net, const uint8_t* d_px, int32_t n, float* d_out);
/* Loss and its gradient over n pairs (device inputs: pixels n x 48 bytes,
 * target sets n x 8 uint32 -- bit L of word L / 32 marks logit L as a
 * target -- endpoints n x 6 bytes); grad_out: host, MPTC_NET_PARAMS floats. */
int32_t mptc_net_gradients(mptc_ctx* ctx, int32_t net, const uint8_t* d_px, const uint32_t* d_target,
                           const uint8_t* d_eps, int32_t n, float seed_weight, float* grad_out,
                           float* loss_out);
/* cfg->steps Adam steps on batches drawn without replacement (reshuffled
 * each epoch) from n_pairs device pairs (layouts as mptc_net_gradients);
 * loss_out (host, optional) receives each step's batch loss. */
int32_t mptc_net_train(mptc_ctx* ctx, int32_t net, const uint8_t* d_px, const uint32_t* d_target,
                       const uint8_t* d_eps, int32_t n_pairs, const mptc_train_cfg* cfg,
                       float* loss_out);
/* Training pairs from the exhaustive search.  With recording on, every
 * exhaustive step also records, per block, which offsets of its inter and
 * intra windows pass the accept test (inter: against the previous frame;
 * intra: against the fixpoint words).  mptc_collect_pairs then appends, for
 * network `net`, one pair per block of the last step whose set is not empty:
 * 48 pixels to d_px, the set (8 words) to d_target, the block's endpoints
 * after the encode (the frame's eps output, else its own fit's) to
 * d_eps_out, at index *d_count (device counter, advanced; pairs beyond
 * capacity are dropped). */
int32_t mptc_record_pairs(mptc_ctx* ctx, int32_t enable);
int32_t mptc_collect_pairs(mptc_ctx* ctx, int32_t net, uint8_t* d_px, uint32_t* d_target,
                           uint8_t* d_eps_out, uint32_t* d_count, int32_t capacity);
/* Search mode and proposals per block (top_k 1..16) for later encodes. */
int32_t mptc_set_search_mode(mptc_ctx* ctx, int32_t mode, int32_t top_k);
/* The offsets (logit indices, -1 = none) the last learned step proposed:
 * top_k per block of the step, raster order over its frames; out: host,
 * n_blocks * top_k int16.  Blocks that did not search hold stale values. */
int32_t mptc_get_proposals(mptc_ctx* ctx, int32_t net, int16_t* out, int64_t n_blocks);
/* Replaces a network's scores by one fixed score per logit index (host,
 * MPTC_NET_LOGITS floats; NULL returns to the network): the same offsets are
 * proposed for every block. */
int32_t mptc_set_prior(mptc_ctx* ctx, int32_t net, const float* scores);
/* Fixed base scores under a network's logits (host, MPTC_NET_LOGITS floats;
 * NULL removes them): logit L = network output L + base[L] wherever the
 * network is used -- training loss, mptc_net_forward and the learned search.
 * With the logit columns of W3 and b3 at zero the network then starts as the
 * fixed ranking the base gives and learns corrections to it. */
int32_t mptc_net_set_base(mptc_ctx* ctx, int32_t net, const float* scores);
/* Index similarity in the learned search's ranking.  h(L) = the number of
 * pixels (0..16) whose 2-bit index differs between a block's own index word
 * and the word of the candidate at offset L (inter: the previous frame's
 * word; intra: the word the intra search starts from).  With scores set
 * (host, MPTC_SIM_BINS floats; NULL removes them) the proposals rank offset L
 * by its network logit (or prior score) + scores[h(L)].  Training is not
 * affected. */
#define MPTC_SIM_BINS 17
int32_t mptc_net_set_similarity(mptc_ctx* ctx, int32_t net, const float* scores);
/* Statistics for those scores, gathered by mptc_collect_pairs since recording
 * was last enabled: out[h] = admitted offsets with h differing indices over
 * the blocks that searched the window, out[MPTC_SIM_BINS + h] = those that
 * passed the accept test (2 x MPTC_SIM_BINS values). */
int32_t mptc_get_similarity_counts(mptc_ctx* ctx, int32_t net, uint64_t* out);

/* Allocation-free variant of mptc_encode_frame_device (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code inside
 * a serving loop that must not call hipMalloc, harnesses that forbid it.  It
 * takes no context: the work buffers live in a device array inside this
 * library (one copy per device, about 53 MiB, held until the process ends),
 * and the call makes no device or host allocation after its first call on a
 * device (which looks up that array's address).  Limits: frames of at most
 * 65536 4x4 blocks (1280x720 = 57,600 fits, 1920x1080 does not) and
 * search_area 1..8; larger inputs return MPTC_EINVAL.  Same inputs, outputs
 * (device pointers) and bytes as mptc_encode_frame_device with that search
 * area and threshold.  Work is queued on `hip_stream` (NULL = the null
 * stream) and the call returns once the frame is done.  Calls on one device
 * share its array: they are serialised within the process (a call waits for
 * the previous one), and no state carries over from one call to the next. */
int32_t mptc_encode_frame_noalloc(int32_t device, const uint8_t* d_rgb, int32_t width, int32_t height,
                                  int32_t search_area, int32_t err_threshold, int32_t is_intra,
                                  const uint64_t* d_ref_phys, uint64_t* d_phys_out, uint8_t* d_motion_out,
                                  uint32_t* d_unique_out, uint32_t* d_n_unique_out, uint8_t* d_eps_out,
                                  uint8_t* d_kind_out, void* hip_stream);

/* Quality of an encoded frame: *d_sse_out (device, one uint64) receives the
 * sum over all pixels and channels of the squared difference between the
 * RGB8 frame and the decoding of its physical blocks d_phys (device, raster
 * order).  PSNR = 10 log10(3 * 255^2 * width * height / sse), the formula of
 * DXTImage::PSNR (codec/dxt_image.cpp:363-383) over the blocks a decoder
 * reads.  Queued on hip_stream (NULL: the null stream); no context needed. */
int32_t mptc_frame_sse(const uint8_t* d_rgb, int32_t width, int32_t height, const uint64_t* d_phys,
                       uint64_t* d_sse_out, void* hip_stream);

/* Seeded synthetic RGB8 frames (see mptc_amd/csrc/framegen.h): writes
 * n_frames frames starting at first_frame, each width*height*3 bytes.
 * The layout of cfg is mptc_synth_cfg. */
int32_t mptc_synth_frames(const void* cfg, int32_t first_frame, int32_t n_frames, uint8_t* rgb_out);

#ifdef __cplusplus
}
#endif
#endif
