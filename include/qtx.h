/*
 * qtx.h -- C ABI of the MI355X (gfx950) S3TC block codec.
 *
 * Plain C: pointers, sizes and status codes only.  Every entry point below
 * replaces one call that the reference's pybind11 extension `_quicktex` makes
 * into its C++ codecs; the replaced interface is cited beside each one
 * (paths relative to the reference tree).
 *
 * Conventions
 *   - Raw textures are row-major RGBA8, width*height*4 bytes, no row padding
 *     (quicktex/Texture.h:73,87).
 *   - Block textures are row-major blocks, ceil(w/4) x ceil(h/4), 8 bytes per
 *     BC1/BC4 block and 16 bytes per BC3/BC5 block (quicktex/Texture.h:158-162).
 *   - `*_device` entry points take device pointers and a hipStream_t passed as
 *     `void*` (NULL = the null stream); they enqueue work and return without
 *     synchronising.  The host entry points copy host buffers through the
 *     context's staging memory (pageable uploads of 8 MiB and more and downloads
 *     of 2 MiB and more through pinned chunks; pinned buffers directly; calls of
 *     at most 1.5 MiB in and out together run their kernels in place, on the
 *     caller's buffers if both are pinned, else on the pinned chunks if both are
 *     pageable, QTX_HOST_DIRECT_MAX bytes to change it) and return
 *     after the result is back on the host; BC1/BC3 encodes of 16 MiB and more
 *     upload in row bands that overlap the encode.  The results are those of the
 *     device entry points.
 *   - Nothing is thrown across the ABI.  On failure a status is returned and
 *     qtx_last_error() (thread-local) describes it.
 */
#ifndef QTX_H
#define QTX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum qtx_status {
    QTX_OK = 0,
    QTX_EINVAL = 1,  /* bad argument: maps to Python ValueError      */
    QTX_ERANGE = 2,  /* index/size out of range: maps to IndexError  */
    QTX_ENOMEM = 3,  /* host or device allocation failed             */
    QTX_EDEVICE = 4, /* HIP runtime error                            */
    QTX_EBUSY = 5,   /* serving queue full / request not finished    */
    QTX_ENOTSUP = 6, /* request the library cannot serve             */
    QTX_ECANCELED = 7 /* served request cancelled before it ran       */
} qtx_status;

/* Interpolator::Type (quicktex/s3tc/interpolator/Interpolator.h:30) */
enum { QTX_INTERP_IDEAL = 0, QTX_INTERP_IDEAL_ROUND = 1, QTX_INTERP_NVIDIA = 2, QTX_INTERP_AMD = 3 };
/* BC1Encoder::ColorMode (quicktex/s3tc/bc1/BC1Encoder.h:51-65) */
enum { QTX_COLOR_FOUR = 4, QTX_COLOR_THREE = 3, QTX_COLOR_THREE_BLACK = 0x13 };
/* BC1Encoder::ErrorMode (quicktex/s3tc/bc1/BC1Encoder.h:67-79) */
enum { QTX_ERROR_NONE = 0, QTX_ERROR_FASTER = 1, QTX_ERROR_CHECK2 = 2, QTX_ERROR_FULL = 3 };
/* BC1Encoder::EndpointMode (quicktex/s3tc/bc1/BC1Encoder.h:81-94) */
enum { QTX_EP_LEAST_SQUARES = 0, QTX_EP_BOUNDING_BOX = 1, QTX_EP_BOUNDING_BOX_INT = 2, QTX_EP_PCA = 3 };
/* Block formats for the batched and served forms */
enum { QTX_FMT_BC1 = 1, QTX_FMT_BC3 = 3, QTX_FMT_BC4 = 4, QTX_FMT_BC5 = 5 };

/*
 * Complete BC1 encoder state after the reference's clamps: the fields of
 * BC1Encoder that EncodeBlock reads (quicktex/s3tc/bc1/BC1Encoder.h:96-99,
 * 151-168).  Fill it with qtx_bc1_params_init() and then, like the Python
 * advanced API, overwrite single fields.
 */
typedef struct qtx_bc1_params {
    int32_t interpolator;      /* QTX_INTERP_*                                 */
    int32_t color_mode;        /* QTX_COLOR_*                                  */
    int32_t error_mode;        /* QTX_ERROR_*                                  */
    int32_t endpoint_mode;     /* QTX_EP_*                                     */
    int32_t two_ls_passes;     /* bool                                         */
    int32_t two_ep_passes;     /* bool                                         */
    int32_t two_cf_passes;     /* bool                                         */
    int32_t exhaustive;        /* bool                                         */
    uint32_t power_iterations; /* clamped to [4, 10]                           */
    uint32_t search_rounds;    /* any value                                    */
    uint32_t orderings4;       /* clamped to [1, 128]                          */
    uint32_t orderings3;       /* clamped to [1, 32]                           */
} qtx_bc1_params;

/* BC1Encoder(level, color_mode, interpolator) ctor + SetLevel:
 * quicktex/s3tc/bc1/BC1Encoder.cpp:47-249.  EINVAL for level > 19, a colour
 * mode other than 4/3/0x13, or an unknown interpolator. */
qtx_status qtx_bc1_params_init(qtx_bc1_params *out, uint32_t level, int32_t color_mode, int32_t interpolator);
/* Re-applies the clamps of BC1Encoder::SetOrderings4/3 and
 * SetPowerIterations (quicktex/s3tc/bc1/BC1Encoder.cpp:251-258) and validates
 * the enums.  Every encode entry point calls it on a private copy. */
qtx_status qtx_bc1_params_clamp(qtx_bc1_params *p);

/* Thread-local description of the last failure on this thread. */
const char *qtx_last_error(void);

/* ---------------------------------------------------------------------- */
/* Context: one device, its stream, device-resident tables and staging.    */
typedef struct qtx_ctx qtx_ctx;
qtx_status qtx_ctx_create(int32_t device, qtx_ctx **out);
qtx_status qtx_ctx_destroy(qtx_ctx *ctx);
/* Blocks until all work the context enqueued on its own stream is done. */
qtx_status qtx_ctx_synchronize(qtx_ctx *ctx);

/* Which of the two BC1/BC3 encoders serves this context's encodes: QTX_BC1_AUTO
 * (default) picks by setting and amount of work, QTX_BC1_LANE forces the
 * one-block-per-lane encoder (every setting in plain launches; batches keep the
 * exhaustive fit of level 19 on the other), QTX_BC1_GROUP the
 * 16-lanes-per-block encoder.  Both produce identical bytes; this is a speed
 * knob (and lets tests cover both).  Applies to later plain launches and batches
 * and servers created afterwards.  No counterpart in the reference. */
enum { QTX_BC1_AUTO = 0, QTX_BC1_LANE = 1, QTX_BC1_GROUP = 2 };
qtx_status qtx_ctx_set_bc1_encoder(qtx_ctx *ctx, int32_t which);

/* ---------------------------------------------------------------------- */
/* Device-resident entry points (inputs and outputs already in HBM).       */

/* BlockEncoder<BlockTexture<BC1Block>>::Encode with BC1Encoder::EncodeBlock
 * (quicktex/Encoder.h:36-56, quicktex/s3tc/bc1/BC1Encoder.cpp:261-342). */
qtx_status qtx_bc1_encode_device(qtx_ctx *ctx, const qtx_bc1_params *params, const uint8_t *d_rgba, int32_t width,
                                 int32_t height, uint8_t *d_blocks, void *stream);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same encode and bytes as qtx_bc1_encode_device, with
 * the lane kernels' work-queue words and their park records or block lists in a
 * static array of the library (8 MiB + 1 KiB per GPU, held from the library's
 * load until the process ends) instead of the context's lent sets; it calls no
 * allocator.
 * Textures of at most QTX_NOALLOC_MAX_BLOCKS blocks (ceil(w/4) * ceil(h/4)), e.g.
 * 2048 x 2048; larger ones return QTX_ERANGE.  Every call on one GPU uses that
 * GPU's one array, whatever context or stream it comes from: calls whose work
 * overlaps on a GPU race and give wrong blocks, so the caller serialises them
 * (one stream, or an event between them). */
#define QTX_NOALLOC_MAX_BLOCKS 262144
qtx_status qtx_bc1_encode_device_noalloc(qtx_ctx *ctx, const qtx_bc1_params *params, const uint8_t *d_rgba,
                                         int32_t width, int32_t height, uint8_t *d_blocks, void *stream);
/* BC3Encoder::Encode (quicktex/s3tc/bc3/BC3Encoder.cpp:27-32): BC4 of
 * channel 3 then BC1; params->color_mode must be QTX_COLOR_FOUR. */
qtx_status qtx_bc3_encode_device(qtx_ctx *ctx, const qtx_bc1_params *params, const uint8_t *d_rgba, int32_t width,
                                 int32_t height, uint8_t *d_blocks, void *stream);
/* BC4Encoder(channel).Encode (quicktex/s3tc/bc4/BC4Encoder.cpp:30-71). */
qtx_status qtx_bc4_encode_device(qtx_ctx *ctx, uint32_t channel, const uint8_t *d_rgba, int32_t width, int32_t height,
                                 uint8_t *d_blocks, void *stream);
/* BC5Encoder(chan0, chan1).Encode (quicktex/s3tc/bc5/BC5Encoder.cpp:25-30). */
qtx_status qtx_bc5_encode_device(qtx_ctx *ctx, uint32_t chan0, uint32_t chan1, const uint8_t *d_rgba, int32_t width,
                                 int32_t height, uint8_t *d_blocks, void *stream);
/* BC1Decoder(write_alpha, interpolator).Decode
 * (quicktex/Decoder.h:36-54, quicktex/s3tc/bc1/BC1Decoder.cpp:31-52). */
qtx_status qtx_bc1_decode_device(qtx_ctx *ctx, int32_t interpolator, int32_t write_alpha, const uint8_t *d_blocks,
                                 int32_t width, int32_t height, uint8_t *d_rgba, void *stream);
/* BC3Decoder(interpolator).Decode (quicktex/s3tc/bc3/BC3Decoder.cpp:26-32). */
qtx_status qtx_bc3_decode_device(qtx_ctx *ctx, int32_t interpolator, const uint8_t *d_blocks, int32_t width,
                                 int32_t height, uint8_t *d_rgba, void *stream);
/* BC4Decoder(channel).Decode (quicktex/s3tc/bc4/BC4Decoder.cpp:29-49). */
qtx_status qtx_bc4_decode_device(qtx_ctx *ctx, uint32_t channel, const uint8_t *d_blocks, int32_t width,
                                 int32_t height, uint8_t *d_rgba, void *stream);
/* BC5Decoder(chan0, chan1).Decode (quicktex/s3tc/bc5/BC5Decoder.cpp:25-30). */
qtx_status qtx_bc5_decode_device(qtx_ctx *ctx, uint32_t chan0, uint32_t chan1, const uint8_t *d_blocks, int32_t width,
                                 int32_t height, uint8_t *d_rgba, void *stream);

/* ---------------------------------------------------------------------- */
/* Host entry points: what the pybind11 layer calls for encoder.encode()   */
/* and decoder.decode().  `out_bytes` must be at least the texture size.   */

/* BC1Encoder.encode (quicktex/s3tc/bc1/_bindings.cpp:113). */
qtx_status qtx_bc1_encode(qtx_ctx *ctx, const qtx_bc1_params *params, const uint8_t *rgba, int32_t width,
                          int32_t height, uint8_t *out, size_t out_bytes);
/* BC3Encoder.encode (quicktex/s3tc/bc3/_bindings.cpp:83). */
qtx_status qtx_bc3_encode(qtx_ctx *ctx, const qtx_bc1_params *params, const uint8_t *rgba, int32_t width,
                          int32_t height, uint8_t *out, size_t out_bytes);
/* BC4Encoder.encode (quicktex/s3tc/bc4/_bindings.cpp:92). */
qtx_status qtx_bc4_encode(qtx_ctx *ctx, uint32_t channel, const uint8_t *rgba, int32_t width, int32_t height,
                          uint8_t *out, size_t out_bytes);
/* BC5Encoder.encode (quicktex/s3tc/bc5/_bindings.cpp:74). */
qtx_status qtx_bc5_encode(qtx_ctx *ctx, uint32_t chan0, uint32_t chan1, const uint8_t *rgba, int32_t width,
                          int32_t height, uint8_t *out, size_t out_bytes);
/* BC1Decoder.decode (quicktex/s3tc/bc1/_bindings.cpp:180). */
qtx_status qtx_bc1_decode(qtx_ctx *ctx, int32_t interpolator, int32_t write_alpha, const uint8_t *blocks,
                          int32_t width, int32_t height, uint8_t *rgba_out, size_t out_bytes);
/* BC3Decoder.decode (quicktex/s3tc/bc3/_bindings.cpp:108). */
qtx_status qtx_bc3_decode(qtx_ctx *ctx, int32_t interpolator, const uint8_t *blocks, int32_t width, int32_t height,
                          uint8_t *rgba_out, size_t out_bytes);
/* BC4Decoder.decode (quicktex/s3tc/bc4/_bindings.cpp:113). */
qtx_status qtx_bc4_decode(qtx_ctx *ctx, uint32_t channel, const uint8_t *blocks, int32_t width, int32_t height,
                          uint8_t *rgba_out, size_t out_bytes);
/* BC5Decoder.decode (quicktex/s3tc/bc5/_bindings.cpp:98). */
qtx_status qtx_bc5_decode(qtx_ctx *ctx, uint32_t chan0, uint32_t chan1, const uint8_t *blocks, int32_t width,
                          int32_t height, uint8_t *rgba_out, size_t out_bytes);

/* ---------------------------------------------------------------------- */
/* Batched form: many textures (of any sizes, formats and encoder settings)
 * in one launch.  Replaces the per-texture loops of the reference's callers:
 * the mip chain of quicktex.dds.encode (quicktex/dds.py:282-287, one
 * encoder.encode per mip) and DDSFile.decode_all (quicktex/dds.py:209-216,
 * one decoder.decode per mip).  Each job produces the same bytes as the
 * matching plain entry point above. */
enum { QTX_OP_ENCODE = 0, QTX_OP_DECODE = 1 };

typedef struct qtx_job {
    int32_t op;            /* QTX_OP_ENCODE or QTX_OP_DECODE                         */
    int32_t format;        /* QTX_FMT_BC1, _BC3, _BC4 or _BC5                        */
    int32_t width, height; /* texture size in pixels                                 */
    const void *src;       /* device memory: RGBA8 texture (encode) or blocks (decode) */
    void *dst;             /* device memory: blocks (encode) or RGBA8 texture (decode) */
    qtx_bc1_params bc1;    /* BC1/BC3 encode: encoder state; BC1/BC3 decode: only
                              bc1.interpolator is read                              */
    int32_t chan0, chan1;  /* BC4: chan0; BC5: chan0 and chan1                       */
    int32_t write_alpha;   /* BC1 decode                                             */
    int32_t reserved;      /* 0                                                      */
} qtx_job;

/* Launch shape of a batch, chosen by qtx_batch_autotune or set by hand. */
typedef struct qtx_batch_tuning {
    int32_t workgroups_per_cu; /* resident BC1/BC3 workgroups per CU, 1..4 (256 lanes each) */
    uint32_t unit_kops;        /* target work per queue unit, in thousands of ops, 16..16384 */
} qtx_batch_tuning;

typedef struct qtx_batch qtx_batch;
/* Validates the jobs (EINVAL/ERANGE as the plain entry points would), orders
 * them by estimated cost, uploads the job tables and captures the launch as a
 * HIP graph (one persistent BC1/BC3 kernel per interpolator present, on
 * parallel branches, and one kernel for BC4/BC5 encodes and all decodes).
 * Pointers in the jobs are captured: the batch encodes whatever those buffers
 * hold at each launch.  Destination buffers must not overlap. */
qtx_status qtx_batch_create(qtx_ctx *ctx, const qtx_job *jobs, size_t njobs, qtx_batch **out);
/* Enqueues one run of the batch on stream (NULL = null stream).  Runs of one
 * batch are serialised among themselves, whatever streams they are given. */
qtx_status qtx_batch_launch(qtx_batch *batch, void *stream);
/* Times `repeats` launches of each candidate shape on stream and keeps the
 * fastest (the runs write the outputs; results do not depend on the shape).
 * Writes the choice and its mean time per launch (ms) if the pointers are
 * non-NULL.  Synchronises the stream. */
qtx_status qtx_batch_autotune(qtx_batch *batch, void *stream, int32_t repeats, qtx_batch_tuning *best,
                              float *best_ms);
qtx_status qtx_batch_set_tuning(qtx_batch *batch, const qtx_batch_tuning *tuning);
qtx_status qtx_batch_get_tuning(const qtx_batch *batch, qtx_batch_tuning *tuning);
qtx_status qtx_batch_destroy(qtx_batch *batch);
/* Convenience: create, launch on stream, wait, destroy. */
qtx_status qtx_batch_run(qtx_ctx *ctx, const qtx_job *jobs, size_t njobs, void *stream);
/* Host-memory form: src/dst of the jobs are host pointers; inputs are staged
 * into one device arena, the batch runs once, outputs are copied back.
 * dst_bytes[i] is the capacity of jobs[i].dst. */
qtx_status qtx_batch_run_host(qtx_ctx *ctx, const qtx_job *jobs, const size_t *dst_bytes, size_t njobs);

/* ---------------------------------------------------------------------- */
/* Mip chains: the levels quicktex.dds.encode writes (quicktex/dds.py:276-287).
 * Level sizes follow quicktex.image_utils.mip_sizes (image_utils.py:9-38): each
 * level halves both sides of the one before (rounding down, at least 1) until
 * 1x1.  Every level is resampled from the base (level 0) as
 * image_utils.resize_no_premultiply does (image_utils.py:42-52): Pillow's
 * Image.BILINEAR resize of each RGBA8 channel, bit-exact. */

/* Levels in the chain of a width x height texture: the full chain for
 * requested <= 0, else min(requested, full) -- mip_sizes(size, mip_count). */
int32_t qtx_mip_count(int32_t width, int32_t height, int32_t requested);
/* Device workspace qtx_mips_device needs for a chain of `levels` levels (0 when
 * the request is invalid). */
size_t qtx_mips_workspace_bytes(int32_t width, int32_t height, int32_t levels);
/* Resamples levels 1 .. levels-1 of the chain of the RGBA8 texture d_rgba into
 * d_levels[0 .. levels-2] (device memory, RGBA8 of each level's size), on
 * stream.  d_workspace: qtx_mips_workspace_bytes bytes of device memory, not
 * used by anything else until the work completes. */
qtx_status qtx_mips_device(qtx_ctx *ctx, const uint8_t *d_rgba, int32_t width, int32_t height, int32_t levels,
                           uint8_t *const *d_levels, void *d_workspace, size_t workspace_bytes, void *stream);
/* Host-memory form: out_levels[m-1] (capacity out_bytes[m-1]) receives level m. */
qtx_status qtx_mips_host(qtx_ctx *ctx, const uint8_t *rgba, int32_t width, int32_t height, int32_t levels,
                         uint8_t *const *out_levels, const size_t *out_bytes);
/* The encoder side of quicktex.dds.encode: resamples the chain on the device and
 * encodes every level (the base included) as one batch, the levels never
 * leaving device memory.  tmpl: an encode job whose op, format, bc1 params and
 * channels apply to every level (its size and buffers are ignored);
 * out_blocks[m] (capacity out_bytes[m]) receives level m's blocks, byte-identical
 * to encoding each level with the plain entry points. */
qtx_status qtx_mips_encode_host(qtx_ctx *ctx, const qtx_job *tmpl, const uint8_t *rgba, int32_t width, int32_t height,
                                int32_t levels, uint8_t *const *out_blocks, const size_t *out_bytes);

/* ---------------------------------------------------------------------- */
/* The per-image choices of the reference's `quicktex encode` commands
 * (quicktex/cli/encode.py), on the device. */

/* *has_alpha = 1 if any texel's alpha is below 255, else 0: the alpha-histogram
 * test of `encode auto` (cli/encode.py:129-133).  The device form reads a
 * 16-byte aligned texture on stream and waits for it. */
qtx_status qtx_has_alpha_device(qtx_ctx *ctx, const uint8_t *d_rgba, int32_t width, int32_t height, void *stream,
                                int32_t *has_alpha);
qtx_status qtx_has_alpha_host(qtx_ctx *ctx, const uint8_t *rgba, int32_t width, int32_t height, int32_t *has_alpha);
/* `encode auto` for one image (cli/encode.py:103-141): the chain is encoded as
 * BC3 (bc1 params in FourColor, as BC3Encoder(level) builds its colour encoder)
 * when the base has alpha below 255, else as BC1 with bc1 params.  *format
 * receives QTX_FMT_BC3 or QTX_FMT_BC1; out_bytes must hold the BC3 sizes (the
 * larger).  One upload of the base serves the test and the chain. */
qtx_status qtx_mips_encode_auto_host(qtx_ctx *ctx, const qtx_bc1_params *bc1, const uint8_t *rgba, int32_t width,
                                     int32_t height, int32_t levels, uint8_t *const *out_blocks, const size_t *out_bytes,
                                     int32_t *format);
/* The normal-map swizzle of `encode bc3nm` (cli/encode.py:52-55): every texel
 * (R, G, B, A) becomes (255, G, G, R).  Device form in place on stream; host form
 * from rgba into out (width * height * 4 bytes each). */
qtx_status qtx_swizzle_bc3nm_device(qtx_ctx *ctx, uint8_t *d_rgba, int32_t width, int32_t height, void *stream);
qtx_status qtx_swizzle_bc3nm_host(qtx_ctx *ctx, const uint8_t *rgba, int32_t width, int32_t height, uint8_t *out);

/* ---------------------------------------------------------------------- */
/* Serving form: a request loop on its own thread.  Requests (encode or
 * decode, any format, size and settings; host buffers) are admitted in
 * arrival order as device memory frees up, grouped into batches -- every
 * request that is waiting and fits joins the next launch -- and run while
 * earlier batches are still on the device (continuous batching).  Each result
 * is byte-identical to the plain entry point's for the same request.  This
 * replaces the reference's one-texture-at-a-time encode()/decode() calls
 * (quicktex/Encoder.h:36-56, quicktex/Decoder.h:36-54) for callers with
 * many independent requests. */
typedef struct qtx_server_config {
    size_t device_bytes;    /* HBM arena for copies of pageable request inputs, the
                               admission budget (0: 2 GiB); pinned inputs are read in
                               place and outputs are written to host memory directly     */
    size_t staging_bytes;   /* pinned host staging for outputs bound for pageable memory
                               (0: 256 MiB); pinned destinations are written directly   */
    int32_t slots;          /* batches in flight at once, 1..8 (0: 6)                     */
    int32_t max_batch_jobs; /* requests per batch, 1..65536 (0: 1024)                     */
    int32_t max_queue;      /* waiting requests before submit returns EBUSY (0: 65536)    */
    uint32_t unit_kops;     /* colour work-unit size, as in qtx_batch_tuning; 0: 64        */
    size_t batch_bytes;     /* bytes a batch may move (device memory + the PCIe traffic of
                               kernels that read or write host memory); a larger
                               request runs alone (0: min(device_bytes / slots, 256 MiB)) */
} qtx_server_config;

typedef struct qtx_server_result {
    int32_t status;     /* qtx_status of the request                                */
    int32_t batch_jobs; /* requests in the launch that served it                    */
    double submit_s;    /* server clock (seconds since start): submitted            */
    double admit_s;     /* device memory reserved                                   */
    double launch_s;    /* its batch enqueued                                       */
    double done_s;      /* output in the caller's buffer                            */
    char message[160];  /* failure description when status != QTX_OK               */
} qtx_server_result;

typedef struct qtx_server_stats {
    uint64_t submitted, completed, failed, batches;
    uint64_t pixels;           /* width * height summed over completed requests   */
    uint64_t max_waiting;      /* largest number of requests waiting for admission */
    uint64_t max_batch_jobs;   /* largest batch launched                           */
    size_t device_bytes_peak;  /* most request memory reserved at once             */
    double busy_s;             /* time with at least one batch in flight            */
    uint64_t cancelled;        /* requests cancelled before admission               */
} qtx_server_stats;

typedef struct qtx_server qtx_server;
/* Starts the loop on ctx's device (config may be NULL for the defaults).
 * Batches run on the context's two streams (its stream and copy stream),
 * consecutive batches on alternate ones, so that their kernels overlap; within a batch, decodes and
 * BC4/BC5 run before BC1/BC3 work.  A batch whose launch the runtime refuses fails
 * with QTX_EDEVICE and the server goes on; an error raised by work on the device
 * fails everything in flight and waiting, and later submits return QTX_EDEVICE. */
qtx_status qtx_server_start(qtx_ctx *ctx, const qtx_server_config *config, qtx_server **out);
/* Queues one request; job->src/dst are host pointers that must stay valid until
 * the request completes; dst_bytes is the capacity of job->dst.  Validation
 * errors are returned here (EINVAL/ERANGE, as the plain entry points); ENOMEM
 * if the request alone exceeds the device or staging budget; EBUSY if the
 * queue is full.  Thread-safe. */
qtx_status qtx_server_submit(qtx_server *s, const qtx_job *job, size_t dst_bytes, uint64_t *ticket);
/* Waits up to timeout_ms (negative: no limit) for the request; QTX_OK and the
 * result when it is done (the ticket is then released), EBUSY on time-out,
 * EINVAL for an unknown ticket.  Thread-safe. */
qtx_status qtx_server_wait(qtx_server *s, uint64_t ticket, int32_t timeout_ms, qtx_server_result *result);
/* Cancels a request that is still waiting for admission: it completes at once
 * with status QTX_ECANCELED (its buffers are not touched) and its ticket is
 * collected with qtx_server_wait as any other.  EBUSY if it has been admitted
 * (it then completes normally), EINVAL for an unknown or finished ticket.
 * Thread-safe.  No counterpart in the reference, which has no request queue. */
qtx_status qtx_server_cancel(qtx_server *s, uint64_t ticket);
qtx_status qtx_server_stats_get(qtx_server *s, qtx_server_stats *stats);
/* Seconds on the server clock now. */
double qtx_server_now(qtx_server *s);
/* Stops accepting, finishes every queued request, joins the loop, frees the server. */
qtx_status qtx_server_stop(qtx_server *s);

#ifdef __cplusplus
}
#endif

#endif /* QTX_H */
