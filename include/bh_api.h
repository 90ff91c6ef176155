/*
 * bh_api.h — C ABI of the MI355X photon-geodesic renderer.
 *
 * This is the drop-in boundary for the per-pixel Schwarzschild march that the
 * reference runs inline in main():
 *   include/blackhole/blackhole_solution_test.cc:161-298   (the march)
 *   include/blackhole/blackhole_solution.h:16-93           (StaticBlackhole)
 *   include/blackhole/camera.h:21-71                       (Camera)
 *   include/blackhole/object/object_manager.h:69-85        (FindCollision)
 *   include/blackhole/object/vector_object.h:159-179,306-356 (Annulus + texel)
 *   include/blackhole/ray_tracer.h:52-96                   (RayTracer::Prograde)
 * The reference has no FFI of its own; its seams are C++ templates.  Every
 * entry point below cites the reference interface it replaces.
 *
 * Conventions: plain C types only, caller-owned buffers, int status codes
 * (BH_OK == 0, negative on error, message via bh_last_error), no exceptions
 * across the boundary.  One context per host thread.
 */
#ifndef BH_API_H_
#define BH_API_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BH_API_VERSION 1
#define BH_MAX_SPHERES 32
#define BH_MAX_TEXTURES 64
#define BH_MAX_RECTS 8
#define BH_MAX_PLANES 4

/* ---- status codes ------------------------------------------------------ */
enum {
  BH_OK = 0,
  BH_ERR_INVALID_ARG = -1,
  BH_ERR_NO_DEVICE = -2,
  BH_ERR_HIP = -3,
  BH_ERR_OUT_OF_MEMORY = -4,
  BH_ERR_EMPTY_TEXTURE = -5, /* Material::SetTexture terminates (material.h:29-35) */
  BH_ERR_UNSUPPORTED = -6,
  BH_ERR_BAD_HANDLE = -7,
  BH_ERR_NOT_BUILT = -8
};

/* ---- hit classes (one byte per pixel) ---------------------------------- */
enum {
  BH_HIT_NONE = 0,    /* nothing hit: black (reference) or background (escape) */
  BH_HIT_HORIZON = 1, /* StaticBlackhole::Collide (blackhole_solution.h:65-88) */
  BH_HIT_DISC = 2,    /* Annulus::Collide (vector_object.h:328-347) */
  BH_HIT_OBJECT = 3   /* any other object: textured sphere, rectangle, plane */
};

/* ---- option enums ------------------------------------------------------ */
enum {
  BH_PREC_FP64 = 0, /* every step in FP64, reference operation order */
  BH_PREC_FP32 = 1  /* FP32 march; FP64 set-up, swept angle and photon-sphere band.
                       Scenes the FP32 march does not cover render as BH_PREC_FP64:
                       bilinear filter, a disc not centred on the hole, nstep > 4096,
                       or the camera at or inside the photon sphere (|C - B| <= 3M) */
};
enum { BH_FILTER_NEAREST = 0, BH_FILTER_BILINEAR = 1 };
enum {
  BH_ROOT_BISECT20 = 0, /* StaticBlackhole::SolveG: 20 bisections, lower end */
  BH_ROOT_EXACT = 1     /* turning point to full FP64 precision (smooth in b, M) */
};

/* ---- images ------------------------------------------------------------ */
/* Row-major rows x cols x channels bytes, channels 3 or 4.  Only the first
 * three channels are used and are passed through in order, exactly like the
 * reference's cv::imread(IMREAD_COLOR) texture (config.h:35-37) whose bytes
 * go straight to the framebuffer (vector_object.h:176-178). */
typedef struct bh_image {
  const uint8_t* data;
  int32_t rows, cols, channels;
} bh_image;

/* ---- scene ------------------------------------------------------------- */
/* Camera<T> (camera.h:30-63): position and the three Object axes
 * (object.h:112-114, camera default (1,0,0),(0,-1,0),(0,0,-1)). */
typedef struct bh_camera {
  double position[3];
  double axis_x[3], axis_y[3], axis_z[3];
  int32_t width, height;
  double fov; /* radians; clamped to [0, nextafter(pi,0)] like Camera::fov */
} bh_camera;

/* StaticBlackhole<T>(position, mass): r_s = 2M (blackhole_solution.h:24,57) */
typedef struct bh_hole {
  double position[3];
  double mass;
} bh_hole;

/* Annulus<T>(p1,p2,p3,p4,r_outer,r_inner) (vector_object.h:320-326).
 * vertex[0] is the centre/pivot and vertex[1..4] are p1..p4 (the Object
 * initializer-list constructor appends to the default vertex, object.h:38,109).
 * normal is normalize((p3-p1) x (p2-p1)) taken at construction. */
typedef struct bh_annulus {
  double vertex[5][3];
  double normal[3];
  double r_outer, r_inner;
  int32_t texture; /* texture slot */
  int32_t pad_;
} bh_annulus;

/* Textured sphere (not in the reference; its Sphere::Collide is unfinished,
 * vector_object.h:281-300).  Hit test = the horizon's segment/sphere test,
 * texture = equirectangular map of the outward normal. */
typedef struct bh_sphere {
  double center[3];
  double radius;
  int32_t texture;
  int32_t pad_;
} bh_sphere;

typedef struct bh_scene {
  bh_camera camera;
  bh_hole hole;
  int32_t has_disc;
  int32_t n_spheres;
  bh_annulus disc;
  bh_sphere spheres[BH_MAX_SPHERES];
  int32_t background; /* texture slot of an equirect background, -1 = none */
  int32_t pad_;
} bh_scene;

/* ---- options ----------------------------------------------------------- */
typedef struct bh_options {
  int32_t precision; /* BH_PREC_*                                           */
  int32_t nstep;     /* reference: 20 (blackhole_solution_test.cc:202); the  */
                     /* FP32 march has an instance compiled for 20 (fastest) */
  int32_t escape;    /* 0: march ends near the camera radius (reference);
                        1: continue outbound until r >= r_escape or step_cap */
  int32_t step_cap;  /* total steps per ray in escape mode (10000)          */
  double r_escape;   /* absolute radius of the escape sphere                */
  int32_t filter;    /* BH_FILTER_*                                         */
  int32_t root;      /* BH_ROOT_*                                           */
  double band;       /* FP32 mode: rays with |b/b_c - 1| < band march in FP64 */
  double edge_width; /* bilinear mode only: soft disc-edge width (0 = hard) */
  double edge_guard; /* FP32 mode: bound on the relative error of an FP32 hit
                        point or escape direction.  A ray whose disc, sphere or
                        background texel coordinate lies within that error of a
                        texel edge is marched again in FP64 in the reference's
                        operation order, so its texel is the FP64 march's.
                        0: the default 2e-6; negative: no re-march.           */
} bh_options;

/* ---- per-pixel outputs (NULL = not wanted) ------------------------------ */
typedef struct bh_outputs {
  uint8_t* rgb;       /* H*W*3, texture channel order (reference: BGR)      */
  uint8_t* hit;       /* H*W hit class                                      */
  double* phi_end;    /* H*W orbit-plane angle at the last step taken       */
  double* hit_point;  /* H*W*3 world position of the hit (NaN when none)    */
  int32_t* steps;     /* H*W integration steps taken                         */
} bh_outputs;

/* ---- differentiable hole + disc scene --------------------------------------- */
#define BH_DIFF_MAX_CROSSINGS 8
/* Parameters of the inverse fit: the hole (r_s = 2M) at the origin, a thin
 * disc in z = 0 centred on it with the texture square (+-r_out, +-r_out) and
 * normal (0,0,-1), and a camera at distance dist, inclination incl from +z and
 * azimuth azim, looking at the hole with +z up.  The render is the reference
 * march (nstep 20) with the exact turning point, bilinear texels and a soft
 * disc edge: coverage alpha = S((rho - r_in)/w) S((r_out - rho)/w),
 * S(x) = smoothstep of clamp(x + 1/2, 0, 1), composited front to back over the
 * disc crossings of a ray until the horizon or full coverage. */
typedef struct bh_diff_params {
  double r_s, incl, dist, r_inner; /* the four fitted parameters */
  double r_outer, azim, fov, edge_width;
  int32_t width, height;
  int32_t texture; /* disc texture slot */
  int32_t pad_;
} bh_diff_params;

/* ---- flat-space scene (ray_tracer_test path, ray_tracer.h:52-96) -------- */
typedef struct bh_rect {
  double vertex[5][3]; /* [0] pivot, [1..4] = p1..p4 (vector_object.h:94-98) */
  int32_t texture;
  int32_t pad_;
} bh_rect;

typedef struct bh_plane {
  double position[3];
  double axis_x[3], axis_y[3], axis_z[3];
  double pattern_size; /* ChessPattern2D(size) (pattern.h:14-50), a whole
                          number in 1..2^29 (the reference's int); 0 = black */
} bh_plane;

typedef struct bh_flat_scene {
  bh_camera camera;
  int32_t n_rects;
  int32_t n_planes;
  bh_rect rects[BH_MAX_RECTS];
  bh_plane planes[BH_MAX_PLANES];
  int32_t steps; /* Prograde step_size (ray_tracer_test.cc:145 uses 10) */
  int32_t pad_;
} bh_flat_scene;

/* ======================================================================== */
/* Context                                                                   */
/* ======================================================================== */
typedef struct bh_ctx bh_ctx;

/* Replaces the reference's process-wide state (ObjectManager singleton,
 * object_manager.h:30-33).  device = HIP ordinal. */
int bh_ctx_create(int device, bh_ctx** out);
void bh_ctx_destroy(bh_ctx* ctx);
/* Message of the last failing call on ctx; with ctx NULL, the message of the
 * last failing call on this thread that had no context (bh_ctx_create, ring
 * waits and releases), or "" if none. */
const char* bh_last_error(const bh_ctx* ctx);
/* Library version and the device name / CU count the context runs on. */
int bh_version(void);
int bh_device_info(bh_ctx* ctx, char* name, int name_len, int* n_cu, int* clock_khz);
/* Run the context's work on a caller stream (hipStream_t), NULL = own stream. */
int bh_ctx_set_stream(bh_ctx* ctx, void* stream);
int bh_synchronize(bh_ctx* ctx);

/* ======================================================================== */
/* Textures — Material::SetTexture (material.h:29-35)                        */
/* ======================================================================== */
/* Upload img into slot (0..BH_MAX_TEXTURES-1); replaces the slot's content.
 * Empty image -> BH_ERR_EMPTY_TEXTURE (the reference calls std::terminate);
 * more than 2^28 texels or a side over 2^24 -> BH_ERR_INVALID_ARG. */
int bh_texture_set(bh_ctx* ctx, int slot, const bh_image* img);
int bh_texture_clear(bh_ctx* ctx, int slot);

/* ======================================================================== */
/* Rendering — the march of blackhole_solution_test.cc:161-298               */
/* ======================================================================== */
/* Validate a scene against the textures bound in ctx (no rendering). */
int bh_scene_check(bh_ctx* ctx, const bh_scene* scene, const bh_options* opt);

/* Render one frame into host buffers (synchronous). */
int bh_render(bh_ctx* ctx, const bh_scene* scene, const bh_options* opt,
              const bh_outputs* out);

/* Render one frame into device buffers on the context stream (asynchronous).
 * Pointers in out are device addresses of the same shapes as bh_render's. */
int bh_render_device(bh_ctx* ctx, const bh_scene* scene, const bh_options* opt,
                     const bh_outputs* out_device);

/* bh_render_device without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate (code under HIP graph capture,
 * tight serving loops).  The scene record and the FP64 queue live in a
 * file-scope device array of the library, 24 MiB on every GPU it runs on,
 * held until the process ends; the caller passes the output planes and binds
 * the textures beforehand (bh_texture_set).  Frames whose FP32 march needs a
 * queue of more than 3 x 2^20 items (larger than about 2560 x 1080) return
 * BH_ERR_INVALID_ARG.  All calls on one GPU share its array: calls that
 * overlap on one GPU (other streams, other contexts) race; calls on one
 * context's stream are ordered.  The kernels of scenes without escape or
 * spheres use no scratch memory; those of escape and sphere scenes spill a
 * few registers, and the HIP runtime reserves scratch memory for a kernel at
 * its first launch on a device, so a caller that cannot allocate renders such
 * a scene once beforehand (with this or any other entry point). */
int bh_render_device_noalloc(bh_ctx* ctx, const bh_scene* scene, const bh_options* opt,
                             const bh_outputs* out_device);

/* ======================================================================== */
/* Camera paths — one frame per camera, replayed as HIP graphs              */
/* ======================================================================== */
typedef struct bh_path bh_path;
/* Capture n_frames renders of scene with cameras[i] into HIP graphs writing
 * rgb/hit planes of frame i at offset i*H*W*{3,1} of device buffers owned by
 * the path.  All cameras must share width/height. */
int bh_path_create(bh_ctx* ctx, const bh_scene* scene, const bh_options* opt,
                   const bh_camera* cameras, int n_frames, bh_path** out);
/* Launch the captured path once; ms (optional) = GPU time from HIP events.
 * When every frame takes the FP32 march, the path is two graphs launched back
 * to back on the context stream: one march launch over all frames, then one
 * FP64 launch over all frames' queues (band and texel-edge rays); otherwise
 * one graph of per-frame launches. */
int bh_path_launch(bh_path* path, float* ms);
/* Average ms of those two launches over iters eager runs (HIP events on the
 * context stream); BH_ERR_UNSUPPORTED for paths with FP64 frames. */
int bh_path_time(bh_path* path, int iters, double* march_ms, double* band_ms);
/* The durations (ms) of the fused march launch and of the FP64 pass in the
 * path's last bh_path_launch, from HIP events on the context stream around
 * each (a path whose frames all take the FP32 march is two graphs, the march
 * over all frames and the FP64 pass over all frames, launched back to back).
 * Other paths: BH_ERR_UNSUPPORTED. */
int bh_path_last_ms(bh_path* path, double* march_ms, double* band_ms);
/* Copy frame i back to host buffers (rgb H*W*3, hit H*W; either may be NULL). */
int bh_path_read(bh_path* path, int frame, uint8_t* rgb, uint8_t* hit);
void bh_path_destroy(bh_path* path);

/* ======================================================================== */
/* Kernel timing for the measurement harness                                 */
/* ======================================================================== */
/* Average duration in ms of the main march kernel over iters eager launches
 * on the context stream, bracketed by HIP events on that stream.  The frame
 * is rendered into device buffers owned by ctx. */
int bh_time_render(bh_ctx* ctx, const bh_scene* scene, const bh_options* opt,
                   int iters, double* avg_ms_main, double* avg_ms_total,
                   int64_t* band_rays);

/* Measurement helper with no counterpart in the reference: the fused
 * multiply-add throughput (TFLOP/s, 2 flops per FMA) that plain, not packed,
 * FP32 and FP64 VALU instruction streams reach on the context's device,
 * eight independent chains per lane at eight waves per SIMD.  This is the
 * ceiling the FP32 march is issued against (the library is built without
 * packed FP32). */
int bh_fma_issue_rate(bh_ctx* ctx, double* fp32_tflops, double* fp64_tflops);

/* Measurement helper with no counterpart in the reference: the sine and
 * cosine of n FP32 angles (|x| < pi, the span of a chord and the angles within
 * it) exactly as the FP32 march's chord tests form them on the context's
 * device; sc_host receives n (sin, cos) pairs.  n in 1..2^24. */
int bh_chord_sincos(bh_ctx* ctx, const float* x_host, float* sc_host, int64_t n);

/* Diagnostic with the same role as the reference's StaticBlackhole::SolveG
 * (blackhole_solution.h:38-55), per pixel: the turning point l (the lower
 * end of the 20-step bisection cell of the root of G) of every pixel of the
 * frame, as the device's per-ray set-up forms it for the FP32 march (closed
 * form, one Newton step and the cell, or the bisection itself near a cell
 * edge and in the photon-sphere band); 0 for a captured or degenerate ray.
 * l_host receives width x height doubles, row-major. */
int bh_turning_points(bh_ctx* ctx, const bh_scene* scene, const bh_options* opt,
                      double* l_host);

/* ======================================================================== */
/* Batches: many independent scenes in one persistent launch (no reference  */
/* counterpart: the reference renders one scene at a time,                  */
/* blackhole_solution_test.cc:161-298)                                      */
/* ======================================================================== */
/* Scheduling of a batch (0 = default).  Modes:
 *  BH_BATCH_GRID   one launch, a wave per 8x8 tile of every scene (default);
 *  BH_BATCH_TILES  persistent waves take ranges of tiles from a device queue;
 *  BH_BATCH_LANES  persistent waves whose lanes each take rays from the queue
 *                  and refill individually once refill_idle of a wave's 64
 *                  lanes are idle (ray compaction; sphere scenes go solo). */
#define BH_BATCH_GRID 1
#define BH_BATCH_TILES 2
#define BH_BATCH_LANES 3
typedef struct bh_batch_tuning {
  int32_t mode;           /* BH_BATCH_* (BH_BATCH_GRID)                         */
  int32_t block;          /* threads per workgroup: 64, 128 or 256 (256)       */
  int32_t waves_per_simd; /* persistent modes: waves per SIMD, 1..8 (4)       */
  int32_t refill_steps;   /* lanes mode: steps between refill checks (4)       */
  int32_t refill_idle;    /* lanes mode: refill once this many lanes idle (16) */
  int32_t chunk;          /* persistent modes: least 8x8 tiles per queue pop
                             (4); pops take max(chunk, remaining / (4 waves))  */
} bh_batch_tuning;

typedef struct bh_batch_stats {
  double total_ms;    /* upload of the batch tables to completion (HIP events) */
  double march_ms;    /* persistent FP32 march                                 */
  double band_ms;     /* FP64 photon-sphere band pass                          */
  int64_t rays;       /* pixels rendered                                       */
  int64_t band_rays;  /* rays marched by the band pass                         */
  int32_t solo_scenes;/* scenes outside the FP32 march, rendered one by one    */
  int32_t pad_;
} bh_batch_stats;

/* Render n scenes; scene i with options[i] into outs[i] (device pointers,
 * planes as for bh_render_device).  Every scene's output equals its own
 * bh_render_device output.  Scenes the FP32 march does not take (FP64,
 * spheres, bilinear) are rendered one by one after the batch.  scene_ms
 * (host, n doubles, optional) receives each scene's completion time from the
 * start of the march; with scene_ms or stats the call waits for completion,
 * otherwise it is asynchronous on the context stream. */
int bh_render_batch_device(bh_ctx* ctx, int n, const bh_scene* scenes,
                           const bh_options* options, const bh_outputs* outs,
                           const bh_batch_tuning* tune, double* scene_ms,
                           bh_batch_stats* stats);
/* Time each of n_cand tunings on this batch (best of reps runs, march plus
 * band pass) and return the fastest in *best (its time in *best_ms;
 * cand_ms, optional, receives every candidate's time). */
int bh_batch_autotune(bh_ctx* ctx, int n, const bh_scene* scenes, const bh_options* options,
                      const bh_outputs* outs, const bh_batch_tuning* cand, int n_cand, int reps,
                      bh_batch_tuning* best, double* best_ms, double* cand_ms);

/* ======================================================================== */
/* Differentiable render and its adjoint (no reference counterpart: the     */
/* reference renders only; these serve the inverse fit of the four          */
/* parameters of bh_diff_params)                                            */
/* ======================================================================== */
/* Forward: rgb_out = H*W*3 float colour in [0, 255], hit_out = H*W class
 * (either may be NULL).  precision BH_PREC_FP32 or BH_PREC_FP64.  Device
 * pointers, asynchronous on the context stream. */
int bh_diff_render_device(bh_ctx* ctx, const bh_diff_params* p, int precision,
                          float* rgb_out, uint8_t* hit_out);
/* Forward and adjoint in one launch: given dL/drgb (H*W*3 float), writes
 * dL/d(r_s, incl, dist, r_inner) to grad_out[4] (4 doubles).  rgb_out and
 * hit_out (optional) receive the forward render.  A ray records at most
 * BH_DIFF_MAX_CROSSINGS disc crossings for its adjoint, and in FP32 only
 * crossings at an orbit angle below 64 rad (ten turns and more; the angle is
 * kept in 32-bit fixed point); overflow_out (optional, one uint32) receives
 * the number of rays that had crossings beyond either limit, which are left
 * out of the gradient.  All pointers are device pointers; asynchronous on
 * the context stream. */
int bh_diff_adjoint_device(bh_ctx* ctx, const bh_diff_params* p, int precision,
                           const float* grad_rgb, float* rgb_out, uint8_t* hit_out,
                           double* grad_out, uint32_t* overflow_out);
/* Average ms of the forward kernel and of the forward+adjoint kernel over
 * iters back-to-back launches (HIP events on the context stream). */
int bh_diff_time(bh_ctx* ctx, const bh_diff_params* p, int precision, int iters,
                 double* fwd_ms, double* fwd_adj_ms);

/* ======================================================================== */
/* Flat-space march — RayTracer::Prograde (ray_tracer.h:52-96)              */
/* ======================================================================== */
int bh_render_flat(bh_ctx* ctx, const bh_flat_scene* scene, const bh_outputs* out);
/* The same into device buffers on the context stream (asynchronous). */
int bh_render_flat_device(bh_ctx* ctx, const bh_flat_scene* scene, const bh_outputs* out_device);
/* Average ms of the flat march kernel over iters back-to-back launches (HIP
 * events on the context stream; rgb and hit into buffers owned by ctx). */
int bh_time_flat(bh_ctx* ctx, const bh_flat_scene* scene, int iters, double* avg_ms);

/* ======================================================================== */
/* Frame ring — the double-buffered screen hand-off of                       */
/* blackhole_solution_test.cc:117-145,327-334 and ray_tracer_test.cc:97-181  */
/* (two cv::Mat screens swapped and cleared by a thread, the finished one   */
/* written out while the next frame renders)                                 */
/* ======================================================================== */
/* depth (2..8) slots, each a device rgb/hit frame and a pinned host copy.
 * A render call queues the march of one frame into the next slot and the
 * copy of its planes to the slot's pinned memory (on a second stream, so it
 * overlaps the next frame's march), and returns at once (waiting first while
 * that slot is still held by the host).  wait gives the
 * host planes once the frame is there; release hands the slot back.  Every
 * pixel is written by the march, so no clearing pass runs between frames.
 * bh_ring_wait and bh_ring_release may run on another thread than the render
 * calls (the writer thread); their error message is then read with
 * bh_last_error(NULL) on that thread. */
typedef struct bh_ring bh_ring;
int bh_ring_create(bh_ctx* ctx, int width, int height, int depth, bh_ring** out);
/* Frame of scene into the next slot; *slot receives its index.  The scene's
 * camera must have the ring's width and height. */
int bh_ring_render(bh_ring* ring, const bh_scene* scene, const bh_options* opt, int* slot);
int bh_ring_render_flat(bh_ring* ring, const bh_flat_scene* scene, int* slot);
/* Block until slot's frame is on the host; rgb (H*W*3) and hit (H*W) point
 * into the slot's pinned memory until it is released; gpu_ms (optional) =
 * time from the start of the frame's march to the end of its copy (HIP
 * events; the copy runs on a second stream and may queue behind the
 * previous frame's). */
int bh_ring_wait(bh_ring* ring, int slot, const uint8_t** rgb, const uint8_t** hit,
                 float* gpu_ms);
int bh_ring_release(bh_ring* ring, int slot);
/* Waits for every queued frame, then frees the ring. */
void bh_ring_destroy(bh_ring* ring);

/* ======================================================================== */
/* Image files (host only, no GPU needed) — texture input in place of       */
/* cv::imread(path, IMREAD_COLOR) (config.h:35-37) and frame output in      */
/* place of cv::VideoWriter (blackhole_solution_test.cc:69-72,330)          */
/* ======================================================================== */
/* Decode a PNG held in memory to rows x cols x 3 bytes in B, G, R order with
 * IMREAD_COLOR semantics: alpha and tRNS dropped (no compositing), grey
 * replicated, grey below 8 bits scaled to 0..255, 16-bit samples keep the
 * high byte, palette looked up, gamma chunks not applied; interlaced files
 * are read.  dst NULL: only *rows / *cols are set.  Errors:
 * BH_ERR_INVALID_ARG (malformed, CRC, truncated, dst too small),
 * BH_ERR_UNSUPPORTED (unknown critical chunk, more than 2^28 pixels). */
int bh_png_decode(const uint8_t* data, size_t size, uint8_t* dst, size_t dst_size, int* rows,
                  int* cols);
/* Decode a baseline (or extended sequential, Huffman, 8-bit) JPEG held in
 * memory to rows x cols x 3 bytes in B, G, R order as libjpeg's default
 * decompression gives it to imread (integer IDCT, triangle upsampling of
 * 2:1 chroma, JFIF YCbCr -> RGB); grey replicated.  dst NULL: only *rows /
 * *cols.  BH_ERR_UNSUPPORTED for progressive, arithmetic, 12-bit and
 * 4-component files; message via bh_jpeg_last_error. */
int bh_jpeg_decode(const uint8_t* data, size_t size, uint8_t* dst, size_t dst_size, int* rows,
                   int* cols);
const char* bh_jpeg_last_error(void);
/* The same from a file holding a PNG, a JPEG or a binary PPM (P6, maxval
 * 255), told apart by their signatures. */
int bh_image_load(const char* path, uint8_t* dst, size_t dst_size, int* rows, int* cols);
/* Write rows x cols x 3 bytes in B, G, R order (the framebuffer's order for
 * textures loaded as above) to path: ".ppm" binary PPM, ".png" 8-bit RGB
 * PNG (both store R, G, B) or ".jpg"/".jpeg" baseline JPEG at quality 95
 * (imwrite's default). */
int bh_image_write(const char* path, const uint8_t* bgr, int rows, int cols);
/* Message of the last failing image-file call on this thread. */
const char* bh_io_last_error(void);
/* Encode rows x cols x 3 bytes in B, G, R order as a baseline JFIF JPEG
 * (4:2:0, quality 1..100 scaling the Annex K tables, optimal Huffman tables)
 * into a malloc'd buffer the caller frees with bh_free.  Frame output in the
 * format of the reference's MJPG video frames (blackhole_solution_test.cc:71-72);
 * message via bh_jpeg_last_error. */
int bh_jpeg_encode(const uint8_t* bgr, int rows, int cols, int quality, uint8_t** out,
                   size_t* out_size);
void bh_free(void* p);

/* Motion-JPEG AVI writer in place of cv::VideoWriter(path,
 * fourcc('M','J','P','G'), 29, Size(W, H)) and `out_capture << frame`
 * (blackhole_solution_test.cc:69-72,330; ray_tracer_test.cc:62-63): a RIFF
 * 'AVI ' file with one 'vids' stream, handler and compression MJPG, one
 * '00dc' JPEG chunk per frame and an idx1 index of key frames.  Files are
 * limited to 1 GiB (RIFF AVI 1.0; no OpenDML extension).  Frames are
 * rows x cols x 3 bytes in B, G, R order.  bh_video_close writes the index,
 * fills in the frame counts and frees the writer, returning the first error
 * of the writer's life.  Messages via bh_io_last_error. */
typedef struct bh_video bh_video;
int bh_video_open(const char* path, int rows, int cols, double fps, int quality,
                  bh_video** out);
int bh_video_write(bh_video* v, const uint8_t* bgr);
int bh_video_close(bh_video* v);

/* ======================================================================== */
/* Host-side helpers shared by every caller (no GPU needed)                  */
/* ======================================================================== */
/* Camera::fov + focal length W/(2 tan(fov/2)) (camera.h:37-41). */
double bh_focal_length(int width, double fov);
/* Object::Rotate{X,Y,Z} with RotationMatrixForAxis (object.h:58-80,
 * matrix.h:342-356).  axes = {x[3], y[3], z[3]}, verts = n x 3 rotated about
 * verts[0] (NULL/0 for a bare frame).  which: 0 = X, 1 = Y, 2 = Z. */
void bh_object_rotate(double axes[9], double* verts, int n_verts, int which, double rad);
/* Camera frame looking from position at target with world up vector: sets
 * axis_x toward target, axis_z = -up (image top), axis_y = axis_z x axis_x,
 * matching the default camera's handedness (camera.h:32). */
void bh_camera_look_at(bh_camera* cam, const double target[3], const double up[3]);
/* Annulus normal and vertex list from p1..p4 around centre (0,0,0)+offset. */
void bh_annulus_init(bh_annulus* a, const double p1[3], const double p2[3],
                     const double p3[3], const double p4[3], double r_outer,
                     double r_inner, int texture);

/* Seeded generators (SplitMix64) shared by every test and benchmark input. */
uint64_t bh_splitmix64(uint64_t* state);
double bh_uniform(uint64_t* state, double lo, double hi); /* [lo, hi) */
/* 2048x1024-style equirect background: uniform u8 noise overlaid with a
 * cells_x x cells_y checkerboard (dark cells scaled to 1/4, light cells
 * lifted by 3/4 toward 255).  channels = 3. */
void bh_gen_background(uint8_t* dst, int rows, int cols, int cells_x, int cells_y, uint64_t seed);
/* Disc texture: radial bands (period band_px) over uniform noise, RGBA. */
void bh_gen_disc_texture(uint8_t* dst, int rows, int cols, int band_px, uint64_t seed);
/* Uniform u8 noise texture, any channel count. */
void bh_gen_noise(uint8_t* dst, size_t n, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif /* BH_API_H_ */
