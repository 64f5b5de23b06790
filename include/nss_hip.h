/*
 * nss_hip: MI355X-native BM3D behind a C ABI.
 *
 * The entry points replace, for the BM3D path only, the reference plugin's
 * interface (AmusementClub/vapoursynth-nssfactory):
 *   - argument resolution of nss.BM3D: parse_bm3d and its helpers
 *     (src/frontend/bm3d_args.cpp:140-170), the numeric/budget checks of
 *     checked_create (src/frontend/validate.hpp:30-50) and the temporal
 *     arguments (src/frontend/temporal_args.cpp:11-59);
 *   - the per-frame work of the filter node: bm3dGetFrame / rollingGetFrame
 *     (src/host/filter_bm3d.cpp:359-494, :985-1072), which the VapourSynth
 *     glue calls once per frame;
 *   - the block matcher spatial_match (include/nss/cpu_api.hpp:146-170) as a
 *     batched device entry point, for match-list parity checks.
 *
 * Only plain C types cross the boundary. Pixel buffers given to the
 * process/match calls are DEVICE pointers (HIP), 32-bit float, with strides
 * in elements; work is queued on the given HIP stream (NULL = the default
 * stream) and is complete when that stream is.
 *
 * Status codes mirror nss::KernelStatus (include/nss/contracts.hpp:18).
 * On failure nss_hip_last_error() returns the message of the calling thread's
 * last failed call (argument errors read "<ns>.BM3D: <message>", as
 * src/frontend/args.hpp:13-15 formats them).
 */
#ifndef NSS_HIP_H
#define NSS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NSS_HIP_OK 0
#define NSS_HIP_BYPASS 1
#define NSS_HIP_INVALID_INPUT 2
#define NSS_HIP_NUMERICAL_FAILURE 3
#define NSS_HIP_UNSUPPORTED 4
#define NSS_HIP_RESOURCE_FAILURE 5

/* Colour families and sample types: VapourSynth API4 values (cfGray = 1,
 * cfRGB = 2, cfYUV = 3; stInteger = 0, stFloat = 1). */
#define NSS_HIP_CF_GRAY 1
#define NSS_HIP_CF_RGB 2
#define NSS_HIP_CF_YUV 3

typedef struct nss_hip_video_info {
    int color_family;
    int sample_type;
    int bits_per_sample;
    int sub_sampling_w;
    int sub_sampling_h;
    int num_planes;
    int width; /* 0 = variable */
    int height;
    int num_frames;
} nss_hip_video_info;

/* One argument of the call, as a VSMap entry holds it. count = number of
 * elements (0 or a missing key = not given). type: 'i' int64 array, 'f'
 * double array, 's' data (utf-8), 'c' clip (the clip itself travels
 * separately; only its video info is read, from clip_info). */
typedef struct nss_hip_arg {
    const char* key;
    char type;
    int count;
    const int64_t* ints;
    const double* floats;
    const char* data;
    const nss_hip_video_info* clip_info;
} nss_hip_arg;

/* Resolved nss.BM3D arguments (include/nss/params/bm3d.hpp:15-31 Bm3dParams,
 * include/nss/params/temporal.hpp RollingParams). sigma and sigma_basic are
 * EFFECTIVE values: (user * (1/255)) * 0.75, 0 = plane copied. */
typedef struct nss_hip_bm3d_params {
    float sigma[3];
    int block_size[3];
    int group_size[3];
    int block_step[3];
    int bm_range[3];
    int ps_num[3];
    int ps_range[3];
    int radius;
    int chroma;
    int final_stage;
    float sigma_basic[3];
    int block_size_basic[3];
    int group_size_basic[3];
    int block_step_basic[3];
    int ps_num_basic[3];
    int has_ref;        /* a ref clip was given (Wiener stage with that ref) */
    int temporal_mode;  /* 0 legacy, 1 rolling (radius > 0 only) */
    int rolling_chunk;
    int rolling_cache_chunks;
    int rolling_cache_limit;
    int rolling_cache_adaptive;
    int memory_limit_mb; /* 0 = not given */
    int device_id;       /* backend tail device_id:int:opt (default 0) */
    int num_streams;     /* backend tail num_streams:int:opt (default 1) */
    int transform;       /* nss_hip extension transform:int:opt, no counterpart
                            in the reference: NSS_HIP_TRANSFORM_* (default FP32) */
} nss_hip_bm3d_params;

/* Arithmetic of the group transforms. FP32 keeps the reference's FP32
 * arithmetic; BF16X3 and FP16 run the 2-D DCT and the Haar as 16x16x16 MFMA
 * products with FP32 accumulation (block 8 with groups of 8 or 16; other
 * shapes stay FP32). Matching, shrinkage and aggregation stay FP32. */
#define NSS_HIP_TRANSFORM_FP32 0
#define NSS_HIP_TRANSFORM_BF16X3 1
#define NSS_HIP_TRANSFORM_FP16 2

typedef struct nss_hip_bm3d nss_hip_bm3d;

/* Message of the calling thread's last failure ("" if none). */
const char* nss_hip_last_error(void);

/* Library and device description (replaces the CUDA plugin's Version /
 * Backend functions, src/cuda/plugin.cpp:35-71): writes a NUL-terminated
 * text of at most len bytes. */
int nss_hip_backend_info(int device, char* text, int len);

/* Creation-time argument resolution of nss.BM3D. ns is the namespace used in
 * error texts ("nss_hip"). ref_info is NULL unless a ref clip was given. On
 * NSS_HIP_INVALID_INPUT, nss_hip_last_error() holds the exact message.
 * Replaces parse_bm3d (src/frontend/bm3d_args.cpp:140-170) with
 * validate_numeric_args and the memory_limit_mb check of checked_create
 * (src/frontend/validate.cpp:13-46, validate.hpp:33-45) and
 * parse_temporal_mode (src/frontend/temporal_args.cpp:11-59). */
int nss_hip_bm3d_parse(const nss_hip_video_info* clip_info, const nss_hip_video_info* ref_info,
                       const nss_hip_arg* args, int nargs, const char* ns, nss_hip_bm3d_params* out);

/* Output format of the node: the clip's, or for legacy temporal output
 * (radius > 0, temporal_mode "legacy") 2r+1 num/den slice pairs stacked
 * vertically (src/host/filter_bm3d.cpp:1155-1157). */
int nss_hip_bm3d_output_info(const nss_hip_bm3d_params* params, const nss_hip_video_info* clip_info,
                             nss_hip_video_info* out);

/* Filter instance for frames of one fixed format and size on one device
 * (the node creation of src/host/filter_bm3d.cpp:1143-1163). Device buffers
 * are sized here, from the parameters. When memory_limit_mb is below what
 * the frame buffers of the num_streams slots need (radius 0), the instance
 * batches the work to fit, as the reference does
 * (docs/api/cuda/README.md:141, src/cuda/common/group_driver.cpp:481-520):
 * every stage adds into frame-sized int64 sums, matching and filtering the
 * reference grid in batches of rows (and running fewer groups at once where
 * group cubes live in global memory); the output bits do not change. The
 * MFMA transform forms have no batched plan. A limit too small for one batch
 * is an INVALID_INPUT error with the reference's message. */
int nss_hip_bm3d_create(const nss_hip_bm3d_params* params, const nss_hip_video_info* info, int device,
                        nss_hip_bm3d** out);

/* Number of temporal slots the process call reads per plane: 2*radius + 1. */
int nss_hip_bm3d_slots(const nss_hip_bm3d* handle);

/* Filter one frame of a spatial instance (radius 0): the work of
 * bm3dGetFrame, both stages when final=1 (the basic estimate stays on the
 * device). src[p] / ref[p] / dst[p]: device pointers of plane p (ref may be
 * NULL when no ref clip was given), strides in floats. frame / frame_count are
 * informational. Asynchronous on stream. Thread-safe: concurrent calls on
 * one handle each lease one of the instance's num_streams frame slots (the
 * reference's SlotPool::acquire, src/cuda/runtime/stream_pool.hpp:1-4),
 * sleeping while all are leased; a call's work is ordered after the slot's
 * previous work, whatever stream that was queued on. */
int nss_hip_bm3d_process(nss_hip_bm3d* handle, const float* const* src, const int* src_stride,
                         const float* const* ref, const int* ref_stride, float* const* dst, const int* dst_stride,
                         int frame, int frame_count, void* stream);

/* No counterpart in the reference: the frame filter of nss_hip_bm3d_process
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It takes
 * the resolved parameters and format instead of an instance and allocates
 * nothing: the frame's working buffers (match lists, tile partials, the basic
 * estimate) live in one fixed device array of 192 MiB per GPU, held from the
 * library's load until the process ends. A frame whose buffers do not fit
 * (a 1920 x 1080 YUV420 two-stage frame at the default arguments needs
 * about 103 MiB) takes the memory_limit_mb plan batched over grid rows, which
 * needs about 20 bytes per pixel of a two-stage frame (a 3840 x 2160 gray
 * frame fits, 3840 x 2160 RGB does not); a frame that fits neither way
 * fails with NSS_HIP_UNSUPPORTED before any work is queued. A
 * memory_limit_mb in the parameters lowers the cap further. Radius 0 only.
 * The MFMA transform forms have no batched plan. Every call on a GPU shares that GPU's
 * array and keeps nothing between calls: calls that overlap on one GPU
 * (different streams, or threads without ordering) race. The first call on
 * a device also loads the DCT tables there (a constant-memory copy), so make
 * it outside a graph capture. Output bits equal nss_hip_bm3d_process's for
 * the same parameters, format and planes. Asynchronous on stream. */
int nss_hip_bm3d_process_noalloc(const nss_hip_bm3d_params* params, const nss_hip_video_info* info, int device,
                                 const float* const* src, const int* src_stride, const float* const* ref,
                                 const int* ref_stride, float* const* dst, const int* dst_stride, void* stream);

/* Temporal instances (radius > 0; rolling, or legacy below): the rolling
 * computation of rollingGetFrame (src/host/filter_bm3d.cpp:730-851,
 * :985-1072) with all state on the device. Push the clip's frames in order
 * (device pointers, copied into the instance's device ring; ref planes when a
 * ref clip was given); each push runs the centres it makes ready. Pop returns
 * finished frames in order: NSS_HIP_OK with *frame set, or NSS_HIP_BYPASS when
 * none is ready. Pop after every push (a push fails with
 * NSS_HIP_RESOURCE_FAILURE while too many finished frames wait). After the
 * last frame (num_frames of the create call) every remaining frame becomes
 * ready. Use one stream for an instance's calls (or order them). A push that
 * fails part-way (a device error while it runs centres) leaves the rolling
 * sums inconsistent: later pushes return NSS_HIP_RESOURCE_FAILURE, frames
 * finished before it still pop; create a new instance to go on. */
int nss_hip_bm3d_push(nss_hip_bm3d* handle, const float* const* src, const int* src_stride, const float* const* ref,
                      const int* ref_stride, void* stream);
int nss_hip_bm3d_pop(nss_hip_bm3d* handle, float* const* dst, const int* dst_stride, int* frame, void* stream);

/* Random access to a temporal clip (the reference's rollingGetFrame over
 * chunks, src/host/filter_bm3d.cpp:985-1072; rolling_chunk_start :548-550).
 * nss_hip_bm3d_chunk: frame n's chunk [start, start + count) of rolling_chunk
 * frames, and the source frames [first, last] its computation reads (every
 * centre whose groups reach the chunk's frames, in every stage: the
 * reference's dependency window). nss_hip_bm3d_seek: makes frame n's chunk
 * the next one pop returns. When it is the next chunk in order nothing is
 * reset (the state carries on, as the reference's carry does); otherwise the
 * instance restarts at the chunk's first source frame. *next_push = the frame
 * to push next: push frames from it through `last` (popping as push asks) and
 * pop until frame start + count - 1 has come out; frames behind the chunk may
 * come out too, in order. The output of a frame does not depend on the order
 * in which chunks are asked for. A seek also ends the failed state of an
 * earlier push. The glue keeps finished chunks (the reference's
 * rolling_cache_chunks / rolling_cache_limit LRU) as output frames. */
int nss_hip_bm3d_chunk(nss_hip_bm3d* handle, int n, int* start, int* count, int* first, int* last);
int nss_hip_bm3d_seek(nss_hip_bm3d* handle, int n, int* next_push, void* stream);

/* Legacy temporal output (radius > 0, temporal_mode "legacy"): push as
 * above; pop returns, per centre frame c in order, its contribution frame
 * (src/host/filter_bm3d.cpp:322-334): per plane 2r+1 slice pairs of plane
 * height, num of slot t at rows 2t*h, den at (2t+1)*h, holding the sums of
 * c's groups' patches taken from frame c-r+t (zeros where that frame is
 * outside the clip; planes with sigma 0 carry the source in the centre slice
 * with weight 1, src/frontend/temporal.hpp:23-36). The VapourSynth glue
 * stamps these frame properties (src/frontend/contribution.hpp:12-27). */
#define NSS_HIP_FAT_VERSION 2        /* _NSSFatVersion, include/nss/contracts.hpp:7 */
#define NSS_HIP_FAT_LAYOUT 1         /* _NSSFatLayout: num/den slices, contracts.hpp:8 */
#define NSS_HIP_MODEL_BM3D 1         /* _NSSModel, contracts.hpp:9 */
#define NSS_HIP_MODEL_VERSION_BM3D 3 /* _NSSModelVersion, contracts.hpp:10-17 */
#define NSS_HIP_NOISE_PROFILE_BM3D 1 /* _NSSNoiseProfile, contracts.hpp:19-20 */

/* nss.VAggregate's reduction of one plane of output frame n
 * (vaggregate_target, src/cpu/bm/agg.cpp:159-201, with the frame logic of
 * src/host/filter_vaggregate.cpp:29-100): nums[k] / dens[k] are device
 * pointers to the num and den rows of contribution k (the slice n-c+r of
 * contribution frame c, for c = max(0, n-r) .. min(N-1, n+r) ascending),
 * strides[k] in floats; dst = sum(num) / sum(den) in that order, or src where
 * sum(den) <= 1e-12. 1 <= count <= 33. Asynchronous on stream. */
int nss_hip_vaggregate(float* dst, int dst_stride, const float* const* nums, const float* const* dens,
                       const int* strides, int count, const float* src, int src_stride, int width, int height,
                       void* stream);

/* Kernel tracing (the role of the reference's NVTX ranges group.match /
 * group.filter / group.aggregate, src/cuda/runtime/nvtx.hpp:15-32): while on,
 * the instance records a HIP event pair around each of its launches on the
 * launch stream. nss_hip_bm3d_kernel_times waits for those events and
 * reports, per kernel kind (0 clear, 1 match, 2 filter + aggregate,
 * 3 finish), the launch count and total milliseconds since the last reset;
 * reset != 0 starts a new window. */
/* Tile-shape autotuner (no counterpart in the reference): times every
 * filter tile shape that fits the LDS budget, one bucket at a time, on a
 * synthetic frame of the instance's format (process on `stream`, repeats
 * frames per shape) and keeps the fastest plan; best_ms = its time per frame.
 * Radius > 0 and ref-only instances have no tile plan (OK, nothing done). */
int nss_hip_bm3d_autotune(nss_hip_bm3d* handle, int repeats, void* stream, double* best_ms);

/* Tile shape of a stage's filter (no counterpart in the reference): every
 * tile-path bucket of stage (0 first) takes tiles of tw x th references;
 * NSS_HIP_UNSUPPORTED when they do not fit its LDS budget. The output does not
 * depend on the shape (tile partials are exact integer sums); the shape only
 * changes the speed. Waits for the instance's queued work. */
int nss_hip_bm3d_set_tiles(nss_hip_bm3d* handle, int stage, int tw, int th);

#define NSS_HIP_KERNEL_KINDS 4
int nss_hip_bm3d_set_tracing(nss_hip_bm3d* handle, int on);
int nss_hip_bm3d_kernel_times(nss_hip_bm3d* handle, int* launches, double* total_ms, int reset);

/* Device memory the instance holds, in bytes, by the kinds of the
 * reference's resource accounting (include/nss/resources.hpp:11
 * ResourceKind: workspace, cached, pinned, inflight, output, framework):
 * bytes[6]; the peak and the memory_limit_mb cap in bytes (0 when none).
 * These are the values a VapourSynth glue stamps on each output frame as
 * _NSSResourceBytes / _NSSResourcePeak / _NSSResourceLimit
 * (src/frontend/ownership.hpp:103-115). Workspace: match lists, partials and
 * sums of one frame; cached: the rolling source, ref and basic rings and the
 * slot sums; output: the rolling output ring. Every buffer is allocated at
 * creation, so the peak is their total; nss_hip_bm3d_create fails when it
 * exceeds memory_limit_mb. */
int nss_hip_bm3d_resources(nss_hip_bm3d* handle, long long* bytes, long long* peak, long long* limit);

/* Frees the instance's device buffers (synchronises its work first). */
int nss_hip_bm3d_destroy(nss_hip_bm3d* handle);

/* Batched block matching over one plane's reference grid (spatial_match for
 * every reference, include/nss/cpu_api.hpp:146-170; grid of
 * src/host/batch_runner.hpp:17-31). plane: device pointer. out_rec: device
 * int32 [nref][group][4] = (x, y, t, bits of the float SSD), reference
 * first, then ascending (distance, y, x); out_count: device int32 [nref].
 * nref = nx * ny with nx = (width - block + step - 1) / step + 1, ny alike. */
int nss_hip_spatial_match(const float* plane, int width, int height, int stride, int block, int step, int bm_range,
                          int group, int* out_rec, int* out_count, void* stream);

/* Batched temporal block matching over one plane's reference grid
 * (predictive_match for every reference: src/cpu/bm/match.cpp:10-33,
 * collect_temporal src/cpu/bm/matcher.hpp:319-373; the batch of
 * predictive_match_batch, src/cpu/batch.cpp:204-241). frames: device pointer
 * to nframes planes of the clip, plane f at frames + f * stride * height.
 * Centre frame `center`, slot t in [0, 2 radius] reads frame
 * center + clamp(t - radius, -center, nframes - 1 - center), and only slots
 * in [max(0, radius - center), radius + min(radius + 1, nframes - center))
 * are searched (src/frontend/temporal.hpp:11-21, src/host/filter_bm3d.cpp:96-97).
 * The spatial search on the centre frame (slot radius) seeds ps_num
 * predictive searches of (2 ps_range + 1)^2 windows per layer. out_rec: device
 * int32 [nref][group][4] = (x, y, slot, bits of the float SSD) in match order
 * (reference first, then ascending distance, slot, y, x); out_count: [nref]. */
int nss_hip_predictive_match(const float* frames, int nframes, int center, int width, int height, int stride,
                             int block, int step, int bm_range, int group, int radius, int ps_num, int ps_range,
                             int* out_rec, int* out_count, void* stream);

/* Matcher choice for this process (instances created afterwards, and
 * nss_hip_spatial_match): 0 automatic, 1 the FP32 matchers only, 2 the MFMA
 * matcher wherever it applies (block 8, group <= 32). Every choice returns
 * the same records; 1 and 2 exist to compare them. The reference picks its
 * tensor-core matcher by window side (src/cuda/common/match.cu:1609-1615
 * spatial_match_tensor) and exposes the forced form to its tests
 * (:1690 spatial_match_tensor_only). */
int nss_hip_set_match_path(int path);

/* ---- WNNM ---------------------------------------------------------------
 * nss.WNNM, the sibling group filter on the same matcher
 * (src/host/filter_wnnm.cpp; the device form src/cuda/wnnm/filter.cpp:16-70):
 * each reference's group is matched as for BM3D, its singular values shrunk
 * by weighted nuclear-norm minimisation, and the groups aggregated. */

/* Resolved arguments (include/nss/params/wnnm.hpp:9-20; defaults
 * include/nss/params.hpp:61-75). sigma is the user value on the 0-255 scale. */
typedef struct nss_hip_wnnm_params {
    float sigma[3];
    int block_size;
    int block_step;
    int group_size;
    int bm_range;
    int radius;
    int ps_num;
    int ps_range;
    int residual;        /* centre each group's rows before the shrinkage */
    int adaptive;        /* adaptive_aggregation: weight 1 / kept values */
    int has_rclip;       /* an rclip was given: matching runs on it */
    int memory_limit_mb; /* 0 = not given */
    int device_id;
    int temporal_mode;   /* radius > 0: 0 legacy, 1 rolling (the default) */
} nss_hip_wnnm_params;

typedef struct nss_hip_wnnm nss_hip_wnnm;

/* Argument resolution of <ns>.WNNM (src/frontend/wnnm_args.cpp:8-29
 * parse_wnnm, then src/cuda/wnnm/filter.cpp:26-33: rclip must match clip,
 * validate_group_planes of src/frontend/validate.hpp:19-25), with the first
 * error's text in nss_hip_last_error(). nss_hip runs radius 0 (spatial);
 * radius > 0 is rejected with a message that says so. */
int nss_hip_wnnm_parse(const nss_hip_video_info* clip_info, const nss_hip_video_info* rclip_info,
                       const nss_hip_arg* args, int nargs, const char* ns, nss_hip_wnnm_params* out);
/* Instance for a clip of `info` (the reference's create,
 * src/cuda/wnnm/filter.cpp:16-70): device buffers for every plane. */
int nss_hip_wnnm_create(const nss_hip_wnnm_params* params, const nss_hip_video_info* info, int device,
                        nss_hip_wnnm** out);
/* One frame (src/host/filter_wnnm.cpp:163-260 wnnmGetFrame at radius 0):
 * device planes src[p] (stride in floats), rclip[p] or NULL (match on src),
 * dst[p]. A plane with sigma 0 is copied. Asynchronous on stream. */
int nss_hip_wnnm_process(nss_hip_wnnm* handle, const float* const* src, const int* src_stride,
                         const float* const* rclip, const int* rclip_stride, float* const* dst,
                         const int* dst_stride, void* stream);
int nss_hip_wnnm_destroy(nss_hip_wnnm* handle);

/* nss.WNNM at radius > 0 (wnnmGetFrame with its fat output,
 * src/host/filter_wnnm.cpp:180-260, process_plane_batched :44-177): the
 * contribution of centre frame `center` (0 <= center < num_frames of the
 * create call). src[t * num_planes + p] (rclip likewise, or NULL): device
 * planes of slot t = 0 .. 2 radius, the frame center + clamp(t - radius,
 * -center, num_frames - 1 - center), strides in floats per entry. Each
 * reference is matched on rclip (or src) by the predictive search over the
 * slots inside the clip, its group shrunk, and its patches added into the
 * num / den sums of their slot. dst[p] (2 (2 radius + 1) plane heights): slot
 * t's num at rows 2t h, den at (2t + 1) h, as nss_hip_bm3d's legacy output;
 * planes with sigma 0 carry the source in the centre slice with weight 1.
 * Rolling output (temporal_mode "rolling") is nss_hip_vaggregate over the
 * contributions of centres n - radius .. n + radius, which the glue keeps.
 * Asynchronous on stream; calls on one handle are serialised. */
int nss_hip_wnnm_process_temporal(nss_hip_wnnm* handle, int center, const float* const* src, const int* src_stride,
                                  const float* const* rclip, const int* rclip_stride, float* const* dst,
                                  const int* dst_stride, void* stream);

/* ---- NCSR ---------------------------------------------------------------
 * nss.NCSR (src/host/filter_ncsr.cpp; the device form src/cuda/ncsr/filter.cpp:16-80,
 * kernels.cu:48-190): the same matcher and eigenbasis as WNNM; each code row
 * of the centred group's PCA is soft-thresholded about its match-weighted
 * mean, over `iters` rounds on an estimate relaxed toward the noisy frame by
 * `delta`. */

/* Resolved arguments (include/nss/params/ncsr.hpp:9-20; defaults
 * include/nss/params.hpp:101-107). sigma on the 0-255 scale. */
typedef struct nss_hip_ncsr_params {
    float sigma[3];
    int block_size;
    int block_step;
    int group_size;
    int bm_range;
    int radius;
    int ps_num;
    int ps_range;
    int iters;
    float delta;
    int has_rclip;
    int memory_limit_mb;
    int device_id;
    int temporal_mode;   /* radius > 0: 0 legacy, 1 rolling (the default) */
} nss_hip_ncsr_params;

typedef struct nss_hip_ncsr nss_hip_ncsr;

/* <ns>.NCSR's argument resolution (src/frontend/ncsr_args.cpp:10-32 parse_ncsr
 * after parse_temporal_mode and the memory limit; rclip and plane checks of
 * src/cuda/ncsr/filter.cpp:26-34). */
int nss_hip_ncsr_parse(const nss_hip_video_info* clip_info, const nss_hip_video_info* rclip_info,
                       const nss_hip_arg* args, int nargs, const char* ns, nss_hip_ncsr_params* out);
int nss_hip_ncsr_create(const nss_hip_ncsr_params* params, const nss_hip_video_info* info, int device,
                        nss_hip_ncsr** out);
/* One frame (ncsrGetFrame at radius 0, src/host/filter_ncsr.cpp:34-190):
 * device planes in and out, rclip (or NULL) guiding round 0's matching. */
int nss_hip_ncsr_process(nss_hip_ncsr* handle, const float* const* src, const int* src_stride,
                         const float* const* rclip, const int* rclip_stride, float* const* dst,
                         const int* dst_stride, void* stream);
int nss_hip_ncsr_destroy(nss_hip_ncsr* handle);

/* nss.NCSR at radius > 0 (ncsrGetFrame with its fat output,
 * src/host/filter_ncsr.cpp:34-190): the contribution of centre frame
 * `center`, with nss_hip_wnnm_process_temporal's arguments and output layout.
 * The iterations run on every slot's estimate: round i > 0 relaxes each slot
 * inside the clip toward its noisy frame by delta, every round matches on the
 * estimates (rclip's slots in round 0) by the predictive search, and each
 * round but the last finishes every slot's estimate (num / den, or the
 * estimate where no patch landed); the last round's sums are the output.
 * Asynchronous on stream; calls on one handle are serialised. */
int nss_hip_ncsr_process_temporal(nss_hip_ncsr* handle, int center, const float* const* src, const int* src_stride,
                                  const float* const* rclip, const int* rclip_stride, float* const* dst,
                                  const int* dst_stride, void* stream);

/* ---- Serving loop ------------------------------------------------------
 * No counterpart in the reference: its filters answer one VapourSynth frame
 * request at a time (src/host/filter_bm3d.cpp bm3dGetFrame). This loop is
 * for callers that feed frames of several clips as they arrive: a scheduler
 * thread admits queued frames into device frame slots under an HBM budget,
 * packs the radius-0 frames admitted so far into one launch sequence per
 * stage round (planes of several clips, sizes and sigmas in the same
 * kernels) while earlier batches still run, and runs a batch whose shape
 * (instances and filtered planes, in order) was seen before as a hipGraph:
 * instantiated once per shape, then updated with each batch's buffers and
 * sigmas (hipGraphExecUpdate) and replayed. Radius > 0 clips keep their
 * rolling state on the device and run on a HIP stream of their own. */
typedef struct nss_hip_server nss_hip_server;

typedef struct nss_hip_server_config {
    int device_id;
    long long hbm_budget_bytes; /* frame slots + rolling state; 0 = a quarter of the device */
    int max_batch_frames;       /* radius-0 frames per launch sequence, 1..16 */
    int max_inflight;           /* batches on the device at once, 1..4 */
    int use_graphs;             /* run repeated batch shapes as hipGraphs */
    int autotune;               /* run nss_hip_bm3d_autotune when a radius-0 stream opens */
} nss_hip_server_config;

typedef struct nss_hip_server_stats {
    /* graph_captures: graphs instantiated; graph_replays: batches run on an
     * existing executable graph after updating its arguments */
    long long submitted, completed, batches, batched_frames, graph_replays, graph_captures;
    long long bytes_in_use, peak_bytes, admissions_deferred;
    long long slots_released; /* idle frame slots released to admit another stream's frame */
} nss_hip_server_stats;

int nss_hip_server_create(const nss_hip_server_config* config, nss_hip_server** out);
/* A clip stream with resolved parameters (nss_hip_bm3d_parse) and format; a
 * radius > 0 stream needs info->num_frames and takes its frames in order.
 * RESOURCE_FAILURE when its rolling state does not fit the budget. */
int nss_hip_server_open_stream(nss_hip_server* server, const nss_hip_bm3d_params* params,
                               const nss_hip_video_info* info, int* stream_id);
/* Queues one frame (device planes, read until the request completes; dst
 * written when it completes: the caller keeps both allocated and untouched
 * until nss_hip_server_query reports the request done). sigma_user > 0
 * overrides the stream's sigma for this frame on every filtered plane and
 * both stages (radius-0 streams); 0 keeps it. ready_stream: the HIP stream
 * on which the caller produced the planes; the request's work waits for what
 * was queued on it before this call (NULL: the planes are complete now). */
int nss_hip_server_submit(nss_hip_server* server, int stream_id, const float* const* src, const int* src_stride,
                          float* const* dst, const int* dst_stride, double sigma_user, void* ready_stream,
                          long long* request_id);
/* NSS_HIP_OK when the request's output is complete on the device (latency_ms:
 * submit to completion seen by the scheduler), NSS_HIP_BYPASS while pending. */
int nss_hip_server_query(nss_hip_server* server, long long request_id, double* latency_ms);
/* Blocks until every submitted request has completed. */
int nss_hip_server_drain(nss_hip_server* server);
int nss_hip_server_get_stats(nss_hip_server* server, nss_hip_server_stats* out);
int nss_hip_server_destroy(nss_hip_server* server);

/* Error-path hooks for tests. No counterpart in the reference's C++ API; they
 * stand in for the fake allocator and launch-error hook its tests put behind
 * their device layer, so that every caller's clean-up and error report can be
 * checked on a healthy device. Process-wide; not for production use.
 * nss_hip_debug_fail_alloc(n): the n-th next device allocation the library
 * makes (n = 0: the next one) fails with hipErrorOutOfMemory; n < 0 disarms.
 * nss_hip_debug_fail_launch(n): the n-th next kernel launch is refused
 * (hipErrorLaunchFailure, nothing launched); n < 0 disarms.
 * nss_hip_debug_live_allocations: the device allocations the library holds
 * now (count, bytes). */
int nss_hip_debug_fail_alloc(int n);
int nss_hip_debug_fail_launch(int n);
int nss_hip_debug_live_allocations(long long* count, long long* bytes);

#ifdef __cplusplus
}
#endif

#endif /* NSS_HIP_H */
