/*
 * leela_amd.h — C ABI of the MI355X leaf-evaluation backend.
 *
 * This library replaces Leela's OpenCL backend (OpenCL.h / OpenCL.cpp) behind
 * the unchanged Network API. Each entry point below names the reference
 * interface it stands in for. Plain pointers and sizes only; no exception or
 * C++ type crosses this boundary. Every call returns LZ_OK (0) or a negative
 * status; lz_last_error() gives the message of the last failure on a context.
 *
 * Data formats at the boundary
 *   planes  uint8 [n][32][S*S], values 0/1, plane order as the reference's
 *           gather_features_policy / gather_features_value (Network.cpp:
 *           883-1201), index y*S+x, already rotated by the caller exactly as
 *           get_scored_moves_internal does (Network.cpp:765-773).
 *   policy  raw = ELU'd output of the last 1-channel conv, float [n][S*S]
 *           (what the reference hands to softmax, Network.cpp:777,808);
 *           out = softmax(raw, T) float [n][S*S] (Network.cpp:450-469).
 *   value   raw = output of the last inner product, float [n];
 *           out = (1 + tanh(raw)) / 2, float [n] (Network.cpp:704,736).
 */
#ifndef LEELA_AMD_H
#define LEELA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LZ_OK 0
#define LZ_ERR_ARG (-1)         /* bad argument or shape */
#define LZ_ERR_HIP (-2)         /* HIP runtime error */
#define LZ_ERR_STATE (-3)       /* call out of order (e.g. before finalize) */
#define LZ_ERR_NOMEM (-4)       /* device or host allocation failed */
#define LZ_ERR_UNSUPPORTED (-5) /* layer table the kernels do not implement */

#define LZ_NET_POLICY 0
#define LZ_NET_VALUE 1

#define LZ_PREC_FP32 0 /* plain form: FP32 everywhere, FP32 MFMA */
#define LZ_PREC_BF16 1 /* BF16 trunk on MFMA, FP32 accumulate/epilogue */
#define LZ_PREC_FP8 2  /* FP8 e4m3 operands after the 5x5 stem (FP16):
                          per-output-channel weight scales, per-board
                          activation scale per layer; 19x19 only */
#define LZ_PREC_FP16 3 /* FP16 operands on the BF16 kernels' MFMA path:
                          same rate, 3 more mantissa bits */

typedef struct lz_ctx lz_ctx;

/* Opens device `device` (HIP ordinal) and creates an empty context.
 * Replaces OpenCL::initialize (OpenCL.cpp:764-908), which picks a device
 * and builds the kernels; here the kernels are precompiled for gfx950. */
int lz_init(lz_ctx **ctx, int device);

/* Appends a same-padded conv layer (ELU applied to its output) to net
 * LZ_NET_POLICY or LZ_NET_VALUE. w is OIHW [cout][cin][filter][filter],
 * b is [cout]; both are copied. Replaces OpenCL_Network::push_convolve
 * (OpenCL.h:67-77), whose channel count is derived as W/(B*k*k). */
int lz_push_convolve(lz_ctx *ctx, int net, int filter, int cin, int cout,
                     const float *w, const float *b);

/* Appends an inner-product layer y = W x + b, ELU iff outputs > 1.
 * w is [outputs][inputs]. Replaces OpenCL_Network::push_innerproduct
 * (OpenCL.h:79-87) and the kernel at OpenCL.cpp:407-438. */
int lz_push_innerproduct(lz_ctx *ctx, int net, int inputs, int outputs,
                         const float *w, const float *b);

/* Activation of a conv (lz_push_activation). LZ_ACT_ELU is the reference's
 * (every conv kernel applies it: Network.cpp:382-392, OpenCL.cpp convolve
 * kernels); ReLU and none are the other forms a Caffe net's layer list
 * takes (the shipped model_4258 / model_4750 nets: Convolution, ReLU,
 * BatchNorm per layer, and a last Convolution with no activation). */
#define LZ_ACT_ELU 0
#define LZ_ACT_RELU 1
#define LZ_ACT_NONE 2

/* Sets the activation of the last conv pushed to `net` (LZ_ACT_ELU unless
 * changed): a Caffe activation layer applied in place to that conv's
 * output. Nets with any conv not ELU, or with batch normalisation, run in
 * LZ_PREC_FP32 only (the other forms return LZ_ERR_UNSUPPORTED). No
 * OpenCL counterpart: the reference's kernels hard-code ELU. */
int lz_push_activation(lz_ctx *ctx, int net, int kind);

/* Appends batch normalisation to the last conv pushed to `net`, applied to
 * its output after the activation: out = (x - means[c] / scale) /
 * sqrt(variances[c] / scale + 1e-5) per channel (Caffe's BatchNorm with
 * stored statistics; channels must equal that conv's outputs). Replaces
 * OpenCL_Network::push_batchnorm (OpenCL.h:52-61) and its kernel
 * (OpenCL.cpp:379-405; Network.cpp:427-448). Not on the last conv. */
int lz_push_batchnorm(lz_ctx *ctx, int net, int channels, const float *means,
                      const float *variances, float scale);

/* Validates both layer tables and uploads the weights in the packed
 * layouts every precision uses. Must follow the last push and precede any
 * forward. The reference copies weights at push time (OpenCL.cpp:474-488);
 * packing for MFMA needs the whole table, hence the extra step. */
int lz_finalize(lz_ctx *ctx);

/* Batches of at most max_boards 19x19 positions (0..64; default 64) run
 * the 16-bit forms layered: one launch per conv, each conv of a board split
 * over several workgroups, instead of one workgroup per board running the
 * whole chain. Same outputs bit for bit; lower latency for small batches
 * (the leaf queue's smallest bucket, synchronous single-position calls).
 * 0 keeps every batch on the persistent kernel. Launch-shape setting with
 * no reference counterpart; applies to forwards and queue graphs captured
 * after the call. */
int lz_set_layered_max(lz_ctx *ctx, int max_boards);

/* Softmax temperature of the policy head (cfg_softmax_temp, GTP.cpp:89-93;
 * default 0.75). */
int lz_set_softmax_temp(lz_ctx *ctx, float temperature);

/* Batched forward of `net` over n positions of 19x19, host buffers in and
 * out, synchronous. planes is uint8 [n][C][361] with C the first conv's
 * input channels (the reference's 32 planes; a Caffe net's own count, e.g.
 * 22, in LZ_PREC_FP32). Replaces OpenCL_Network::forward (OpenCL.cpp:490-569),
 * which evaluates one position per call and blocks on queue.finish().
 * raw or out may be NULL when not wanted. */
int lz_forward(lz_ctx *ctx, int net, int precision, int n,
               const uint8_t *planes, float *raw, float *out);

/* Same as lz_forward with device-resident buffers, enqueued on `stream`
 * (a hipStream_t; NULL is the HIP default stream, as in every HIP call);
 * returns after enqueueing. Buffers must stay valid until the stream
 * reaches the work. Device-buffer calls on one context share its
 * workspace: calls on different streams must be ordered by the caller. */
int lz_forward_device(lz_ctx *ctx, int net, int precision, int n,
                      const uint8_t *d_planes, float *d_raw, float *d_out,
                      void *stream);

/* Caller-owned workspaces: lz_forward_device_ws is lz_forward_device with
 * the activation buffers, value-head planes and layered tiles of `ws`
 * instead of the context's, so forwards on different streams with
 * different workspaces run concurrently (one workspace still serves one
 * stream at a time: calls sharing it must be ordered by the caller). A
 * workspace grows to the largest batch it is given (it is created for
 * max_boards). No OpenCL counterpart: the reference keeps one set of
 * buffers per thread (OpenCL.cpp:446-472); this is that, for device-buffer
 * callers. lz_workspace_destroy waits for the device to be idle. */
typedef struct lz_workspace lz_workspace;
int lz_workspace_create(lz_ctx *ctx, int max_boards, lz_workspace **ws);
void lz_workspace_destroy(lz_workspace *ws);
int lz_forward_device_ws(lz_ctx *ctx, lz_workspace *ws, int net,
                         int precision, int n, const uint8_t *d_planes,
                         float *d_raw, float *d_out, void *stream);

/* Ragged batch: n boards of sides[i] in 5..19 (e.g. mixed 9x9 / 13x13 /
 * 19x19) in one persistent launch of the 16-bit fused kernels
 * (LZ_PREC_BF16 / LZ_PREC_FP16; FP32 returns LZ_ERR_UNSUPPORTED). Planes
 * are packed board after board, uint8 [32][S_i*S_i] each; policy raw/out
 * likewise (S_i*S_i floats per board), value one float per board. The
 * reference nets are 19x19 only (Network.cpp:541-543); for smaller boards
 * the convs are the same same-padded layers and the value head's 361-input
 * inner product sees the S x S plane at y*19+x with zeros elsewhere.
 * Same role as lz_forward: replaces OpenCL_Network::forward. */
int lz_forward_ragged(lz_ctx *ctx, int net, int precision, int n,
                      const int32_t *sides, const uint8_t *planes, float *raw,
                      float *out);
/* Device-buffer form (sides stays a host array), enqueued on `stream`. */
int lz_forward_ragged_device(lz_ctx *ctx, int net, int precision, int n,
                             const int32_t *sides, const uint8_t *d_planes,
                             float *d_raw, float *d_out, void *stream);

/* As lz_forward_device, and also times every kernel launch of the forward
 * with HIP events recorded on `stream` between launches. Writes up to
 * max_launches durations (ms, launch order) to launch_ms and the launch
 * count to *n_launches; synchronises the stream before returning.
 * Measurement hook; no counterpart in the reference. */
int lz_forward_timed(lz_ctx *ctx, int net, int precision, int n,
                     const uint8_t *d_planes, float *d_raw, float *d_out,
                     void *stream, float *launch_ms, int max_launches,
                     int *n_launches);

/* As lz_forward_device for LZ_PREC_BF16 / LZ_PREC_FP16 / LZ_PREC_FP8 and
 * 0 <= n <= LZ_NOALLOC_MAX_BOARDS boards of 19x19, with no allocation: no
 * hipMalloc, hipFree or other allocating call, so it can be made where
 * allocation is not allowed (code under HIP stream capture, tight serving
 * loops, harnesses that forbid it). No counterpart in the reference; it is
 * for those callers. The value net's head plane (n x 361 floats) goes to a
 * static device array of LZ_NOALLOC_MAX_BOARDS x 361 floats (2.96 MB,
 * held by each device the library runs on until the process ends) instead
 * of the context's workspace. Every call on one device shares that array:
 * calls that overlap on one GPU (other streams, other contexts) race, so
 * the caller orders them. Larger n returns LZ_ERR_ARG; LZ_PREC_FP32 returns
 * LZ_ERR_UNSUPPORTED (its activation maps need a workspace). */
#define LZ_NOALLOC_MAX_BOARDS 2048
int lz_forward_device_noalloc(lz_ctx *ctx, int net, int precision, int n,
                              const uint8_t *d_planes, float *d_raw,
                              float *d_out, void *stream);

/* ---- Autotuner ----------------------------------------------------------
 * Times every launch shape of the fused kernels (wave grid 1x4 or 2x4;
 * persistent grid of CUs x occupancy x {1, 2} or one workgroup per slot)
 * for batches of n on the device and keeps the fastest per net for n's
 * bucket (<=64, <=256, <=1024, more). Replaces the reference's fixed
 * OpenCL work sizes (cfg_rowtiles, OpenCL.cpp:605-619). report (optional)
 * receives one line per candidate. Tuning is per context; save/load keep
 * it in a text file. */
int lz_autotune(lz_ctx *ctx, int precision, int n, char *report,
                size_t report_len);
/* The same for a ragged batch of n boards of sides[0..n) (5..19): times
 * lz_forward_ragged's launches of every shape and keeps the fastest for
 * the bucket of the batch's slot count (config: 1024 x 19x19 + 512 x
 * 13x13 + 512 x 9x9 packs into 1408 slots). */
int lz_autotune_ragged(lz_ctx *ctx, int precision, int n,
                       const int32_t *sides, char *report,
                       size_t report_len);
int lz_autotune_save(lz_ctx *ctx, const char *path);
int lz_autotune_load(lz_ctx *ctx, const char *path);

/* ---- Batched leaf queue -------------------------------------------------
 * Replaces the per-thread asynchronous path: Network::async_scored_moves
 * -> OpenCL_Network::forward with forward_cb (Network.cpp:471-587,
 * OpenCL.cpp:490-569, per-thread queues OpenCL.cpp:446-472), which
 * evaluates one position per forward per search thread. Any number of
 * threads submit single positions; they are written straight into one of
 * four batch buffers. A dispatcher thread closes a batch at max_batch rows
 * of one net (or when its oldest row has waited max_wait_us, or on flush)
 * and replays the hipGraph captured for the smallest bucket of 64 / 256 /
 * 1024 / 2048 boards that holds it: the dense graph when every board is
 * 19x19, the ragged graph (mixed sides packed into slots, LZ_PREC_BF16 /
 * LZ_PREC_FP16) otherwise. The 16-bit and FP8 graphs evaluate exactly the
 * batch's rows; the FP32 graphs the whole bucket. While the device works
 * the dispatcher fills and launches the next batch; a completion thread
 * runs the callbacks. The softmax temperature in effect at lz_queue_start
 * is baked into the graphs. */

/* Completion callback, run on the queue's completion thread (never inside
 * the HIP runtime). raw/out point at the position's results (policy: S*S
 * floats for a board of side S; value: 1 float) and are valid only during
 * the call; status != LZ_OK means the batch failed and raw/out are NULL. A
 * callback may submit, but when every batch buffer is busy such a submit
 * returns LZ_ERR_STATE instead of waiting for one. */
typedef void (*lz_done_fn)(void *user, int status, const float *raw,
                           const float *out);

int lz_queue_start(lz_ctx *ctx, int precision, int max_batch,
                   int max_wait_us);

/* Copies planes (uint8 [32][361]) into the queue; `done` runs once with
 * the result. Thread-safe; waits while every batch buffer is busy. */
int lz_submit(lz_ctx *ctx, int net, const uint8_t *planes, lz_done_fn done,
              void *user);

/* As lz_submit for a board of side 5..19 (planes uint8 [32][side*side],
 * as lz_encode_planes writes them). Boards of different sides share
 * ragged launches; a queue at LZ_PREC_FP32 or LZ_PREC_FP8 takes 19x19 only
 * (LZ_ERR_UNSUPPORTED). The reference evaluates 19x19 only
 * (Network.cpp:541-543); smaller boards follow lz_forward_ragged. */
int lz_submit_side(lz_ctx *ctx, int net, int side, const uint8_t *planes,
                   lz_done_fn done, void *user);

/* Measurement hook (no counterpart in the reference): `threads` host
 * threads submit the n boards (sides[i] in 5..19; planes packed board after
 * board, uint8 [32][S_i*S_i] each, as lz_forward_ragged takes them) to the
 * running queue, thread t boards t, t + threads, ..., each board's policy
 * row then its value row, `rounds` times over; then lz_join. Writes
 * rounds * n / elapsed seconds: positions evaluated by both nets per
 * second, from the first submit to the last callback. */
int lz_queue_benchmark(lz_ctx *ctx, int n, const int32_t *sides,
                       const uint8_t *policy_planes,
                       const uint8_t *value_planes, int threads, int rounds,
                       double *positions_per_s);

/* Launches what is queued without waiting for a full batch. */
int lz_flush(lz_ctx *ctx);

/* Flushes and waits until every submitted position's callback has run
 * (OpenCL::join_outstanding_cb, OpenCL.cpp:571-577). */
int lz_join(lz_ctx *ctx);

/* Positions submitted whose callbacks have not run yet; the search's
 * backpressure test (OpenCL::thread_can_issue, OpenCL.cpp:446-454). */
int lz_queue_depth(lz_ctx *ctx);

/* stats[6] = {batches launched, positions, batches per bucket 64 / 256 /
 * 1024 / 2048}. */
int lz_queue_stats(lz_ctx *ctx, int64_t *stats);

/* Drains and stops the queue (also done by lz_destroy). */
int lz_queue_stop(lz_ctx *ctx);

/* Fills out[0..n) with N(0, stddev^2) draws: Box-Muller over the
 * reference's xoroshiro128+ (Random.cpp:27-57) seeded with `seed`. The
 * synthetic He-normal weight sets are made with it because the shipped
 * weight blobs (NN128.cpp, NNValue.cpp) are not available. */
int lz_normal_fill(uint32_t seed, size_t n, float stddev, float *out);

/* ---- Host-side input path (no GPU) ------------------------------------
 * Positions, the synthetic-position generator and the plane encoders that
 * feed lz_forward. They restate the reference's board engine and encoders
 * (FastBoard / FullBoard / FastState, Network::gather_features_*), which
 * stay on the caller's side of the OpenCL seam in the reference; they are
 * exported so a batching front end and the tests can build inputs. */
typedef struct lz_position lz_position;

/* New empty side x side game (5..19), black to move. FastState::init_game
 * (FastState.cpp:20-34). NULL on bad arguments. */
lz_position *lz_position_new(int side, float komi);
void lz_position_free(lz_position *pos);

/* Plays `vertex` for the side to move (FastState::play_move, FastState.cpp:
 * 360-389). Vertices use the reference numbering (y+1)*(side+2)+(x+1)
 * (FastBoard::get_vertex, FastBoard.cpp:38-49); -1 passes. */
int lz_position_play(lz_position *pos, int vertex);

/* Resets and plays the synthetic-position generator: move count uniform in
 * 0..max_moves from Random(seed) (Random.cpp:51-57), each move uniform over
 * empty points that are not the ko square, not an own eye and not suicide
 * (the filter of UCTNode::create_children, UCTNode.cpp:149-158). */
int lz_position_random(lz_position *pos, uint32_t seed, int max_moves);

/* info[7] = {side, to_move, komove, last_move, prev_last_move, movenum,
 * passes} (FastState getters, FastState.cpp:391-576). */
int lz_position_info(const lz_position *pos, int32_t *info);

/* squares[y*side+x] = 0 black, 1 white, 2 empty (FastBoard::square_t). */
int lz_position_squares(const lz_position *pos, uint8_t *squares);

/* The 32 input planes of `net`, rotated by `symmetry` (0..7) as
 * get_scored_moves_internal builds them (Network.cpp:765-773):
 * planes uint8 [32][side*side]. Network::gather_features_policy /
 * gather_features_value (Network.cpp:883-1201). */
int lz_encode_planes(const lz_position *pos, int net, int symmetry,
                     uint8_t *planes);

/* Per-point board queries for the side to move, LZ_QUERY_FIELDS int32 per
 * point (y*side+x), -1 where not applicable: {square, string liberties,
 * string size, after_liberties own, after_liberties opponent, empty
 * neighbours, saving_size, check_losing_ladder, check_winning_ladder,
 * is_suicide, no_eye_fill, self_atari}. FastBoard.cpp:191-2837. */
#define LZ_QUERY_FIELDS 12
int lz_board_queries(const lz_position *pos, int32_t *out);

/* The encoder over a batch: planes of n positions of one board side,
 * uint8 [n][32][S*S], each as lz_encode_planes(pos[i], net, symmetry), on
 * `threads` host threads (<= 0: every hardware thread). The per-leaf
 * gather_features_policy / gather_features_value calls of
 * async_scored_moves, get_value and get_scored_moves (Network.cpp:557,
 * 597, 633), batched. */
int lz_encode_batch(const lz_position *const *pos, int n, int net,
                    int symmetry, int threads, uint8_t *planes);

/* ---- Device feature encoder -------------------------------------------
 * The same planes computed on the GPU, one wavefront per board, written
 * straight into device memory where the nets read them: the per-leaf
 * gather_features_policy / gather_features_value calls (Network.cpp:557,
 * 597, 633; encoders at :883-1201) moved off the host. The host side only
 * copies each board's squares and string representatives into a record
 * (1408 bytes) and sends the batch in one copy. */
typedef struct lz_encoder lz_encoder;

/* Pinned staging and device records for batches up to max_batch boards on
 * `device`. */
int lz_encoder_create(lz_encoder **out, int device, int max_batch);
void lz_encoder_destroy(lz_encoder *enc);
const char *lz_encoder_last_error(const lz_encoder *enc);

/* Planes of n positions (sides may differ) into device memory d_planes,
 * packed board after board, uint8 [32][S_i*S_i] each (the layout of
 * lz_forward_device for one side, of lz_forward_ragged_device for mixed
 * sides), each equal to lz_encode_planes(pos[i], net, symmetry).
 * Enqueued on `stream` (NULL: the HIP default stream); returns once the
 * positions have been read, so they may change afterwards. Calls on one
 * encoder on different streams must be ordered by the caller. */
int lz_encode_device(lz_encoder *enc, const lz_position *const *pos, int n,
                     int net, int symmetry, uint8_t *d_planes, void *stream);

/* Both nets' planes from one pass over the boards (the string, liberty
 * and ladder reading is shared): d_policy as lz_encode_device with
 * LZ_NET_POLICY, d_value as with LZ_NET_VALUE. get_scored_moves and
 * get_value on the same leaf (Network.cpp:557, 597) need both. */
int lz_encode_device_both(lz_encoder *enc, const lz_position *const *pos,
                          int n, int symmetry, uint8_t *d_policy,
                          uint8_t *d_value, void *stream);

/* lz_playout_winrate for n positions on the GPU, playouts of position i
 * seeded seed0 + i (as lz_make_value_labels): d_winrates[i] (device
 * memory) equals the host's value bit for bit. One group of 8 lanes per
 * playout; the boards go through the encoder's record slots, so the same
 * ordering rules apply. Playout::mc_owner's playouts (Playout.cpp:51-113,
 * 174-217) for the surrogate's labels, moved onto the device. */
int lz_playout_winrates_device(lz_encoder *enc, const lz_position *const *pos,
                               int n, int playouts, uint32_t seed0,
                               float *d_winrates, void *stream);

/* Generator + encoder for n positions with seeds seed0..seed0+n-1:
 * planes uint8 [n][32][side*side] of `net`, rotation 0. threads <= 0 uses
 * every hardware thread. */
int lz_make_batch(int n, int side, uint32_t seed0, int max_moves, int net,
                  int threads, uint8_t *planes);

/* ---- Network facade ----------------------------------------------------
 * The evaluation API UCTSearch calls, with the reference's Netresult
 * semantics, computed by this backend (C++ class lz::Network in
 * leela_amd/csrc/network.h; these are its C entry points). */

/* Network::get_scored_moves (Network.cpp:624-674): ensemble 0 DIRECT (with
 * rotation 0..7), 1 RANDOM_ROTATION, 2 AVERAGE_ALL (eight symmetries in one
 * batch). Writes the empty points' probabilities and vertices in row-major
 * order (at most 361), losing-ladder escapes zeroed, not renormalised;
 * *count = 0 for boards other than 19x19. */
int lz_network_scored_moves(lz_ctx *ctx, int precision,
                            const lz_position *pos, int ensemble,
                            int rotation, float *probs, int32_t *vertices,
                            int *count);

/* Network::get_value (Network.cpp:590-622): win probability for the side
 * to move (0.5 for boards other than 19x19). */
int lz_network_value(lz_ctx *ctx, int precision, const lz_position *pos,
                     int ensemble, float *value);

/* Network::initialize (Network.cpp:201-233): pushes the reference's two
 * layer tables (policy: 5x5 32->96, 3x3 96->128, 10x 3x3 128->128, 3x3
 * 128->1; value: 5x5 32->64, 10x 3x3 64->64, 3x3 64->1, FC 361->256, FC
 * 256->1) into a fresh context and finalizes it. The reference compiles
 * its weights in and does not ship them; these are the synthetic
 * He-normal set (N(0, 2 / fan_in) from the reference RNG, stream 7000 +
 * 100 * net + layer, biases zero: leela_amd.nets set "A"). */
int lz_network_initialize(lz_ctx *ctx);

/* Network::get_Network (Network.cpp:139-145): the process-wide context,
 * created and initialised (lz_network_initialize) on device 0 by the first
 * call; NULL if that fails. It lives until the process ends. */
lz_ctx *lz_network_get(void);

/* Network::show_heatmap (Network.cpp:834-881) of get_scored_moves's
 * result, as text into text[0..text_len): 19 lines of "%3d " per point
 * (probability x 1000, top row first, 0 on occupied points), then with
 * topmoves the leading moves as "%1.3f (D4)" lines until they sum to 0.85
 * or one is below 0.01. LZ_ERR_ARG if text_len is too small (4 KB is
 * enough). */
int lz_network_heatmap(lz_ctx *ctx, int precision, const lz_position *pos,
                       int ensemble, int rotation, int topmoves, char *text,
                       size_t text_len);

/* Network::benchmark (Network.cpp:147-199), the netbench command: threads
 * threads, each on a copy of pos (19x19), run ceil(2000/threads)
 * lz_network_scored_moves and then ceil(10000/threads) lz_network_value
 * calls, RANDOM_ROTATION; *policy_ps = 2000 and *value_ps = 10000 over the
 * elapsed seconds. The synchronous facade calls go through the leaf queue
 * when one runs at `precision` (lz_queue_start), so the threads' calls are
 * batched together; otherwise each is one lz_forward. */
int lz_network_benchmark(lz_ctx *ctx, int precision, const lz_position *pos,
                         int threads, float *policy_ps, float *value_ps);

/* Network::async_scored_moves (Network.cpp:536-587) with DIRECT rotation,
 * through the leaf queue (lz_queue_start first; the queue's precision
 * applies). `done` gets the Netresult on the dispatcher thread, as
 * forward_cb does (Network.cpp:484-534). */
typedef void (*lz_moves_fn)(void *user, int count, const float *probs,
                            const int32_t *vertices);
int lz_network_async_scored_moves(lz_ctx *ctx, const lz_position *pos,
                                  int rotation, lz_moves_fn done, void *user);

/* ---- Search-side batching ------------------------------------------------
 * UCTNode::netscore_children / UCTNode::run_value_net feeding the batched
 * backend (UCTNode.cpp:74-121, 316-344; UCTSearch.cpp:79-149): `threads`
 * search threads grow one tree from `root` for `simulations` simulations,
 * with virtual loss on the path of each simulation in flight. With
 * LZ_EVAL_NETS every new leaf is expanded with the policy net (priors over
 * the Netresult's points that pass UCTNode::create_children's filter,
 * UCTNode.cpp:149-158, the 35 best as UCTNode.cpp:551-552 considers; at
 * the root the AVERAGE_ALL ensemble, eight rows in one batch, as
 * netscore_children does there, UCTNode.cpp:103-106) and valued with the
 * value net. With the leaf queue running at `precision` (lz_queue_start)
 * the policy row is submitted asynchronously (as async_scored_moves) while
 * the thread runs the value row through the synchronous facade call, and
 * the threads' rows share launches; without it each call is its own
 * forward. This replaces the
 * reference's one blocking forward per search thread and its
 * thread_can_issue backpressure (OpenCL.cpp:446-454). LZ_EVAL_PLAYOUTS
 * (host only, ctx may be NULL) uses uniform priors and the mean of
 * `playouts` light playouts (lz_playout_winrate) instead. Selection is
 * Q + c_puct * P * sqrt(N) / (1 + n), not the reference's UCT-RAVE tree
 * policy (UCTNode.cpp:543-700), which stays on the caller's side.
 * simulations <= 2^22; the tree holds up to 36 nodes per simulation
 * (about 3 KB) with the nets, one per candidate point with playouts. */
#define LZ_EVAL_NETS 0
#define LZ_EVAL_PLAYOUTS 1
typedef struct {
    double seconds;
    int64_t simulations;      /* completed simulations */
    int64_t policy_evals;     /* leaf expansions (one policy row each) */
    int64_t value_evals;      /* leaf evaluations (one value row each) */
    int64_t collisions;       /* simulations retried: leaf being expanded */
    int64_t nodes;            /* tree nodes */
    int64_t queue_batches;    /* leaf-queue launches during the search */
    int64_t queue_positions;  /* rows of those launches */
    int64_t root_visits;
    int64_t virtual_loss_left;    /* after the search; 0 when consistent */
    int64_t visit_sum_mismatches; /* expanded nodes whose children's visits
                                     do not sum to their own minus one */
    int32_t best_move;        /* most visited root child (vertex, -1 pass) */
    int32_t best_visits;
    int32_t root_children;
    float root_value;         /* mean backed-up value for the side to move */
} lz_search_stats;
int lz_search(lz_ctx *ctx, int precision, const lz_position *root,
              int evaluator, int playouts, int threads, int simulations,
              float c_puct, uint32_t seed, lz_search_stats *stats);

/* lz_search with LZ_EVAL_NETS semantics (same expansion, priors, value and
 * selection) but the nets evaluated by the caller's `forward`: it fills out
 * with the softmax probabilities [n][361] (LZ_NET_POLICY) or win rates [n]
 * (LZ_NET_VALUE) of n rotated plane stacks uint8 [n][32][361] and returns
 * LZ_OK. For checking a search against another implementation of the same
 * nets (the reference CPU path restated); no counterpart in the
 * reference. */
typedef int (*lz_host_forward_fn)(void *user, int net, int n,
                                  const uint8_t *planes, float *out);
int lz_search_host(lz_host_forward_fn forward, void *user,
                   const lz_position *root, int threads, int simulations,
                   float c_puct, uint32_t seed, lz_search_stats *stats);

/* ---- Monte Carlo playouts (teacher and fallback of the surrogate) -----
 * Playout::mc_owner (Playout.cpp:174-217): the mean outcome of n playouts
 * from pos for the side to move (win 1, tie 0.5). Each playout follows
 * Playout::run (Playout.cpp:51-113: stop at four passes or move number
 * 2*side^2, score by FastBoard::final_mc_score) with a light policy —
 * uniform over empty points that are not the ko square, an own eye or
 * suicide — in place of the reference's pattern policy. Playout j draws
 * from its own stream, Random(fmix32(seed ^ j * 0x9E3779B9)) (MurmurHash3's
 * finalizer) with two outputs discarded. Host; lz_playout_winrates_device
 * is the same on the GPU. */
int lz_playout_winrate(const lz_position *pos, int n, uint32_t seed,
                       float *winrate);

/* The reference's playout policy: FastState::play_random_move
 * (FastState.cpp:168-255) — captures of strings left in atari, saving
 * moves for the strings the last move put in atari (escape or capture),
 * semeai moves, nakade, and the 3x3 patterns around the last move, each
 * scored by the product of its feature weights and its pattern's weight
 * (Matcher / PolicyWeights, Matcher.cpp:39-112); when no candidate scores
 * 32 (cfg_bound), four random moves from the empty list and a pass join
 * them; one move is drawn in proportion to the scores. Its tables are read
 * from data/mc_policy.bin next to the library (or $LZ_MC_POLICY, or the
 * file lz_mc_policy_load names), written by
 * leela_amd/data/make_mc_policy.py from the reference's tables. */
#define LZ_PLAYOUT_LIGHT 0   /* uniform over legal non-eye points */
#define LZ_PLAYOUT_PATTERN 1 /* the reference's pattern policy */

/* Loads the pattern policy's tables from `path` (err receives a message on
 * failure). Optional: the first pattern playout loads the default file. */
int lz_mc_policy_load(const char *path, char *err, size_t err_len);

/* lz_playout_winrate with the chosen playout policy (LZ_ERR_STATE when the
 * pattern tables cannot be loaded). */
int lz_playout_winrate_policy(const lz_position *pos, int policy, int n,
                              uint32_t seed, float *winrate);

/* The n playouts behind lz_playout_winrate_policy, one by one: scores[j] is
 * playout j's FastBoard::final_mc_score (black's view, komi included), the
 * playout seeded as the win rate's playout j (Playout::run with
 * postpassout, Playout.cpp:51-113). */
int lz_playout_scores(const lz_position *pos, int policy, int n,
                      uint32_t seed, float *scores);

/* lz_playout_winrates_device with the chosen playout policy. With
 * LZ_PLAYOUT_PATTERN each playout runs on one GPU thread, the same source
 * as the host's (its win rates equal lz_playout_winrate_policy's for the
 * same position and seed); a position whose candidate list outgrows the
 * device's 2,048 entries (none of millions of generator positions' move
 * generations needs more than ~300) gets NaN instead of a win rate. The
 * first pattern call uploads the tables and allocates ~4.5 GB of playout
 * states (one per thread of a persistent grid of 196,608 threads), held
 * until lz_encoder_destroy. Pattern calls on one encoder are ordered by
 * the library, whatever their streams. */
int lz_playout_winrates_device_policy(lz_encoder *enc,
                                      const lz_position *const *pos, int n,
                                      int playouts, uint32_t seed0,
                                      int policy, float *d_winrates,
                                      void *stream);

/* lz_make_value_labels with the chosen playout policy. */
int lz_make_value_labels_policy(int policy, int n, int side, uint32_t seed0,
                                int max_moves, int playouts,
                                uint32_t playout_seed, int threads,
                                uint8_t *planes, float *winrates);

/* FastState::set_handicap (FastState.cpp:566-568): the handicap stones'
 * compensation, added to komi wherever a position is scored (playouts,
 * lz_position_score; FastState.cpp:395-401). 0 after lz_position_new. */
int lz_position_set_handicap(lz_position *pos, int handicap);

/* FastBoard::final_mc_score (FastBoard.cpp:420-443) with the position's
 * komi plus handicap (FastState::calculate_mc_score): black stones plus
 * empty points whose four neighbours are black or edge, minus the same for
 * white, minus komi + handicap. */
int lz_position_score(const lz_position *pos, float *score);

/* Training data of the surrogate: generator positions from seeds
 * seed0..seed0+n-1 (as lz_make_batch), their value planes (if planes is not
 * NULL) and their n-playout win rates for the side to move, playouts of
 * position i seeded playout_seed + i. */
int lz_make_value_labels(int n, int side, uint32_t seed0, int max_moves,
                         int playouts, uint32_t playout_seed, int threads,
                         uint8_t *planes, float *winrates);

/* ---- Learned playout surrogate -------------------------------------------
 * A leaf win-rate estimator trained on playout outcomes, standing in for
 * Playout::mc_owner (Playout.cpp:174-217) where a caller wants a playout
 * estimate; the reference has no learned counterpart. Network: 3x3 conv
 * 32->64, six residual blocks of two 3x3 64->64 convs, a 1x1 64->1 conv,
 * FC 361->64 (ELU after each), FC 64->1, tanh; win rate (1 + t) / 2
 * for the side to move, from the value planes (lz_encode_planes with
 * LZ_NET_VALUE). BF16 MFMA compute, FP32 master weights, Adam.
 * Parameters are a flat FP32 vector: per conv W[64][cin][3][3] then
 * b[64]; then wh[64], bh, F1[64][361], c1[64], F2[64], c2. */
typedef struct lz_surrogate lz_surrogate;

/* Allocates for batches up to max_batch on `device` (gfx950) and draws
 * He-normal weights from seed; biases, the second conv of every residual
 * block and the output FC start at zero (each block starts as the
 * identity, the output at win rate 0.5). */
int lz_surrogate_create(lz_surrogate **out, int device, int max_batch,
                        uint32_t seed);
void lz_surrogate_destroy(lz_surrogate *s);
int lz_surrogate_num_params(void);
const char *lz_surrogate_last_error(const lz_surrogate *s);
int lz_surrogate_get_params(lz_surrogate *s, float *params);
/* Replaces the parameters and resets Adam (moments zero, step 0). */
int lz_surrogate_set_params(lz_surrogate *s, const float *params);
int lz_surrogate_set_adam(lz_surrogate *s, float beta1, float beta2,
                          float eps);
long long lz_surrogate_step(const lz_surrogate *s);

/* Win rates [n] of device planes [n][32][361] (u8), on `stream` (NULL: the
 * HIP default stream); asynchronous. */
int lz_surrogate_forward_device(lz_surrogate *s, int n, const uint8_t *planes,
                                float *winrate, void *stream);
/* Host planes in, host win rates out; synchronous. */
int lz_surrogate_predict(lz_surrogate *s, int n, const uint8_t *planes,
                         float *winrate);
/* Loss mean_n (t_n - (2 y_n - 1))^2 and its gradient (flat, num_params)
 * for host planes and targets y (win rates); no update. grad or loss may
 * be NULL. */
int lz_surrogate_gradients(lz_surrogate *s, int n, const uint8_t *planes,
                           const float *targets, float *grad, float *loss);
/* One training step on device planes and targets: forward, backward,
 * Adam at learning rate lr. If loss is not NULL the call waits and stores
 * the step's loss (before the update). */
int lz_surrogate_train_step_device(lz_surrogate *s, int n,
                                   const uint8_t *planes, const float *targets,
                                   float lr, float *loss, void *stream);
/* Inspection of the last forward/backward pass, for layer-by-layer
 * checks. kind 0: the BF16 output map [n][361][64] of conv `layer` (0..12,
 * after ELU); kind 1: the BF16 gradient map at that conv's pre-activation,
 * available once lz_surrogate_keep_maps(s, 1) has been set (it keeps all
 * 13 maps instead of 3 rotating ones). Values are raw BF16 bit patterns.
 * lz_surrogate_read_head copies the head's saved s [n][361], f [n][64] and
 * t [n] (any pointer may be NULL). */
int lz_surrogate_keep_maps(lz_surrogate *s, int on);
int lz_surrogate_read_map(lz_surrogate *s, int kind, int layer, int n,
                          uint16_t *out);
int lz_surrogate_read_head(lz_surrogate *s, int n, float *sv, float *fv,
                           float *tv);
/* Checkpoint: parameters, Adam moments, step and Adam settings. */
int lz_surrogate_save(lz_surrogate *s, const char *path);
int lz_surrogate_load(lz_surrogate *s, const char *path);

/* Device name, as OpenCL::get_device_name (OpenCL.h:121). */
const char *lz_device_name(lz_ctx *ctx);

/* Message of the last failed call on ctx (or of the last failed lz_init
 * when ctx is NULL). */
const char *lz_last_error(lz_ctx *ctx);

/* Frees device memory and the context. */
void lz_destroy(lz_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
