/*
 * aqsis_hider.h -- C ABI of the MI355X-native REYES hider.
 *
 * The hider takes camera-space surfaces and the hot Ri options, and renders
 * one frame: dice -> bust -> sample -> combine -> filter -> expose ->
 * quantise.  It sits where aqsis's bucket loop sits, so a host renderer that
 * today calls
 *
 *     CqImageBuffer::SetImage()            (libs/core/imagebuffer.cpp:174-228)
 *     CqImageBuffer::PostSurface(surface)  (libs/core/imagebuffer.cpp:353-446)
 *     CqImageBuffer::RenderImage()         (libs/core/imagebuffer.cpp:605-772)
 *     CqDDManager::DisplayBucket(...)      (libs/core/ddmanager/ddmanager.cpp:146-171)
 *
 * calls rh_create / rh_post_surfaces / rh_render / rh_read_* instead.
 *
 * Conventions: every entry point returns an rh_status (0 = success); the
 * message of the last failure on the calling thread is rh_last_error().
 * No exception crosses the ABI.  Host buffers are owned by the caller;
 * device memory and HIP streams are owned by the context.  A context is
 * used by one host thread at a time.
 */
#ifndef AQSIS_HIDER_H
#define AQSIS_HIDER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t rh_status;
enum {
    RH_OK = 0,
    RH_ERR_INVALID = 1,     /* bad argument or option value             */
    RH_ERR_DEVICE = 2,      /* HIP runtime error                        */
    RH_ERR_NOMEM = 3,       /* host or device allocation failed         */
    RH_ERR_STATE = 4,       /* call out of order (e.g. read before render) */
    RH_ERR_CAPACITY = 5     /* request exceeds a fixed capacity         */
};

/* Pixel filters, the RiPixelFilter functions of libs/core/filters.cpp:72-342. */
enum {
    RH_FILTER_BOX = 0,
    RH_FILTER_TRIANGLE = 1,
    RH_FILTER_CATMULLROM = 2,
    RH_FILTER_GAUSSIAN = 3,
    RH_FILTER_SINC = 4,
    RH_FILTER_DISK = 5,
    RH_FILTER_BESSEL = 6,
    RH_FILTER_MITCHELL = 7
};

/* Surface types accepted by rh_post_surfaces. */
enum {
    RH_PRIM_SPHERE = 1,     /* RiSphere, libs/core/geometry/quadrics.cpp         */
    RH_PRIM_BICUBIC = 2,    /* RiPatch "bicubic" (Bezier basis), patch.cpp        */
    RH_PRIM_BILINEAR = 3    /* RiPatch "bilinear", patch.cpp                       */
};

/*
 * The hot options.  Field meanings and defaults follow RiFormat,
 * RiCropWindow, RiProjection, RiClipping, RiPixelSamples, RiPixelFilter,
 * RiExposure, RiQuantize "rgba", RiShutter, RiDepthOfField and Option
 * "limits" (libs/core/api/ri.cpp:730-930, libs/core/options.cpp:273-310,
 * libs/core/optioncache.cpp:48-117).  rh_options_default() fills them in.
 */
typedef struct rh_options {
    int32_t xres, yres;          /* Format                                   */
    float pixel_aspect;          /* Format pixelaspectratio (1)              */
    float crop[4];               /* CropWindow xmin xmax ymin ymax (0 1 0 1) */
    float fov;                   /* Projection "perspective" "fov" (90)      */
    float clip_near, clip_far;   /* Clipping (FLT_EPSILON, FLT_MAX)          */
    int32_t xsamples, ysamples;  /* PixelSamples (2 2)                       */
    int32_t filter;              /* PixelFilter function (gaussian)          */
    float filter_xwidth, filter_ywidth; /* PixelFilter widths (2 2)          */
    float exposure_gain, exposure_gamma;  /* Exposure (1 1)                  */
    float quantize_one, quantize_min, quantize_max, quantize_dither; /* 255 0 255 0.5 */
    float shutter_open, shutter_close;    /* Shutter (0 0)                   */
    int32_t bucket_x, bucket_y;  /* Option "limits" "bucketsize" (16 16)     */
    int32_t eyesplits;           /* Option "limits" "eyesplits" (10)         */
    int32_t jitter;              /* Hider "hidden" "jitter" (1)              */
    int32_t num_displays;        /* display requests that consume dither draws (1) */
    float zthreshold[3];         /* Option "limits" "zthreshold" (1 1 1)     */
    float dof[3];                /* DepthOfField fstop focallength focaldistance
                                    (FLT_MAX x3: pinhole; depth of field is on
                                    when fstop < FLT_MAX, renderer.h:368-377) */
    int32_t depth_filter;        /* Hider "hidden" "depthfilter": RH_DEPTH_MIN (0) */
    int32_t display_z;           /* a display request's mode has z (DMode_Z) (0) */
} rh_options;

/* rh_options.depth_filter (optioncache.cpp:96-105).  The filter sets each
 * sample's depth in CqImagePixel::Combine (imagepixel.cpp:262-305).  With a z
 * display (display_z) it also changes how hits are stored
 * (bucketprocessor.cpp:1076-1081, 1499-1525): max and average keep every hit
 * (nothing is occlusion culled, at the sample or the surface), midpoint keeps
 * the nearest opaque hit and occludes with the second nearest. */
enum { RH_DEPTH_MIN = 0, RH_DEPTH_MIDPOINT = 1, RH_DEPTH_MAX = 2, RH_DEPTH_AVERAGE = 3 };

/*
 * One surface, already in camera space (the state in which
 * CqImageBuffer::PostSurface receives it).  Attributes are the constant
 * shader's inputs (Color, Opacity; shaders/surface/constant.sl) and
 * ShadingRate.
 *
 * Sphere:   sphere[] = radius, zmin, zmax, thetamax (degrees); xform[0] is
 *           the object->camera matrix, element [row][col] at 4*row+col with
 *           RenderMan's row-vector convention (translation in row 3).
 * Bicubic:  P[0..15] are the Bezier control points, row-major in (v,u).
 * Bilinear: P[0..3] are the corners (u,v) = (0,0) (1,0) (0,1) (1,1).
 * nkeys = 2 makes it a deforming surface (geometry given inside RiMotionBegin
 * / RiMotionEnd: CqDeformingSurface, geometry/surface.h:617-720) with keys at
 * key_time[0] < key_time[1]: patches have key 1's camera-space points in P1
 * (same layout as P), spheres key 1's object->camera matrix in xform[1].
 * no_cull_hidden = 1 is Attribute "cull" "hidden" [0]: the surface is never
 * occlusion culled (RenderSurface, bucketprocessor.cpp:945-957); 0, the
 * reference's default, lets a bucket's occlusion tree cull it.
 * focus_factor is GeometricApproximation "focusfactor" (the reference's
 * default is 1; a record built zeroed must set it): with depth of field the
 * shading rate grows by max(1, 0.025 * focus_factor * CoC^2)
 * (CqSurface::AdjustedShadingRate, geometry/surface.cpp:688-724).
 */
typedef struct rh_prim {
    int32_t type;
    int32_t nkeys;
    float shading_rate;
    float Cs[3];
    float Os[3];
    float key_time[2];
    float sphere[4];
    float xform[2][16];
    float P[16][3];
    float P1[16][3];
    int32_t no_cull_hidden;
    float focus_factor;
    int32_t shading_interpolation; /* RH_SHADING_SMOOTH (0, the default) or RH_SHADING_CONSTANT */
    int32_t matte;               /* RiMatte: 0 (the default), 1 matte, 2 matte alpha */
} rh_prim;

/* rh_prim.matte: the attribute "System" "Matte" (CacheGridInfo,
 * micropolygon.cpp:49-55).  A matte surface (1) hides what lies behind it
 * and shows nothing itself: its hits composite to transparent black
 * (imagepixel.cpp:202-213, 301-313); a matte alpha surface (2) is stored as
 * an occluding hit whatever its opacity (bucketprocessor.cpp:1490-1491). */
enum { RH_MATTE_NONE = 0, RH_MATTE = 1, RH_MATTE_ALPHA = 2 };

/* rh_prim.shading_interpolation: RiShadingInterpolation (ri.cpp:1406-1416),
 * the attribute "System" "ShadingInterpolation" (default smooth,
 * attributes.cpp:118).  Smooth interpolates the colour and opacity of a hit
 * bilinearly from the micropolygon's four vertices; constant takes its
 * first vertex's (micropolygon.cpp:1443-1462). */
enum { RH_SHADING_SMOOTH = 0, RH_SHADING_CONSTANT = 1 };

/* Counters of one frame.  spl_* are the reference's own hider counters
 * (libs/core/stats.h:380-382, incremented at bucketprocessor.cpp:1172,
 * 1192, 1482). */
typedef struct rh_stats {
    uint64_t spl_count;        /* (micropolygon, candidate sample) pairs tested */
    uint64_t spl_bound_hits;   /* pairs passing the bound + occlusion tests     */
    uint64_t spl_hits;         /* stored hits                                   */
    uint64_t inside;           /* pairs passing the point-in-micropolygon test  */
    uint64_t inside_nonlinear; /* of those, hits needing a second Newton step   */
    uint64_t grids;            /* diced grids                                   */
    uint64_t micropolygons;    /* busted micropolygons                          */
    uint64_t mp_bucket_refs;   /* (micropolygon, bucket) list entries           */
    uint64_t surfaces;         /* surfaces posted, including split children     */
    uint64_t occlusion_culled; /* occlusion tests that culled a surface         */
    double ms_schedule;        /* host: split/dice scheduling                   */
    double ms_upload;          /* host->device copies of the frame inputs       */
    double ms_dice;            /* device: dice + bust                           */
    double ms_bin;             /* device: per-task micropolygon lists           */
    double ms_sample;          /* device: hit test + depth + composite          */
    double ms_filter;          /* device: filter + expose + quantise            */
    double ms_device;          /* device: whole frame, event to event           */
    double ms_composite;       /* device: transparent hit combine, a part of
                                  ms_sample (0 for an opaque frame)             */
    uint64_t cull_probes;      /* occlusion tests decided by the depths at their
                                  moment (the final depths did not settle them) */
    uint64_t cull_passes;      /* schedules computed for the frame              */
    double ms_cull;            /* host + device: the culling schedule beyond the
                                  first split/dice schedule (a part of ms_schedule) */
    /* parts of ms_schedule */
    double ms_forest;          /* host: split trees of the surfaces              */
    double ms_simulate;        /* host: bucket loops and schedule records, all passes */
    double ms_cull_depth;      /* host + device: final depths, depth pyramids and
                                  the (surface, bucket) tests                   */
    double ms_cull_probes;     /* host + device: rendered probes                 */
} rh_stats;

/* Work and device memory one frame of a scene needs, from the host
 * split/dice schedule alone. */
typedef struct rh_footprint {
    uint64_t surfaces;         /* surfaces posted, including split children     */
    uint64_t grids;            /* grids the schedule dices                      */
    uint64_t micropolygons;
    uint64_t vertices;
    uint64_t list_capacity;    /* task micropolygon-list entries reserved       */
    uint64_t device_bytes;     /* device memory of the frame (shape + schedule) */
} rh_footprint;

/*
 * How a frame is laid out on the device.  None changes a result.
 *   RH_MODE_GRID     one launch per stage over all sampling tasks (default);
 *   RH_MODE_PLAIN    the sampling stage launched once per bucket, in bucket
 *                    order (the reference's one-bucket-at-a-time loop);
 *   RH_MODE_BATCHED  one captured hipGraph per frame; tasks are cut into
 *                    work items by cost (list length x blocks), ordered
 *                    largest first, and pulled from a device queue by
 *                    persistent workgroups.
 */
enum { RH_MODE_GRID = 0, RH_MODE_PLAIN = 1, RH_MODE_BATCHED = 2 };

typedef struct rh_tuning {
    int32_t mode;              /* RH_MODE_*                                     */
    int32_t lds_records;       /* micropolygon records per LDS chunk: 216 or 96 */
    int32_t blocks_per_wave;   /* 64-sample blocks per wave per pass: 4 or 2    */
    uint32_t split_cost;       /* batched: largest work item, in list entries x blocks */
    double frame_ms;           /* rh_autotune: device time of the chosen setting */
} rh_tuning;

typedef struct rh_ctx rh_ctx;

/* Fill *opt with the reference defaults listed above. */
void rh_options_default(rh_options* opt);

/* Create a context for one frame shape.  ~ SqOptionCache::cacheOptions +
 * CqImageBuffer::SetImage + CqOptions::InitialiseCamera. */
rh_status rh_create(const rh_options* opt, rh_ctx** out);

/* Append n camera-space surfaces, in posting order.
 * ~ CqImageBuffer::PostSurface, called once per surface. */
rh_status rh_post_surfaces(rh_ctx* ctx, const rh_prim* prims, int32_t n);

/* Drop all posted surfaces (the frame shape and its sample plan stay). */
rh_status rh_clear_surfaces(rh_ctx* ctx);

/* Render one frame of the posted surfaces.  ~ CqImageBuffer::RenderImage.
 * The host split/dice schedule is recomputed only when surfaces changed. */
rh_status rh_render(rh_ctx* ctx);

/* Run only the device part of the last frame again, on the resident
 * grids and sample plan (dice -> ... -> quantise), for timing. */
rh_status rh_render_device(rh_ctx* ctx);

/* rh_render_device without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops that forbid allocation).  The frame's counters and
 * the sampling workgroups' counter rows live in a fixed device array of
 * RH_NOALLOC_SCRATCH_BYTES that the library holds on each GPU until the
 * process ends (frames of up to 209,702 sampling workgroups: RH_ERR_CAPACITY
 * beyond); every other buffer is the context's, sized by rh_render.  Always
 * the static grid layout, issued eagerly, with the per-stage times of
 * rh_stats.  A frame whose transparent hits outgrow the context's arena
 * fails with RH_ERR_CAPACITY (rh_render_device would grow it).  Calls that
 * overlap on one GPU share its array and race: serialise them. */
#define RH_NOALLOC_SCRATCH_BYTES (4u << 20)
rh_status rh_render_device_noalloc(rh_ctx* ctx);

/* rh_render_device with the kernels issued one by one (also in the batched
 * form, which otherwise replays a frame graph) and a device timestamp
 * between the stages, filling the per-stage times of rh_stats.  Same result
 * and counters.  No counterpart in the reference. */
rh_status rh_render_stages(rh_ctx* ctx);

/* Device layout of the following frames (no counterpart in the reference:
 * it has one CPU bucket loop).  RH_ERR_INVALID for values outside those
 * listed in rh_tuning. */
rh_status rh_set_tuning(rh_ctx* ctx, const rh_tuning* t);
rh_status rh_get_tuning(const rh_ctx* ctx, rh_tuning* t);

/* Time the static grid and every batched setting (LDS chunk, blocks per
 * wave, split cost) on the posted scene, keep the fastest, and return it.
 * Renders frames. */
rh_status rh_autotune(rh_ctx* ctx, rh_tuning* chosen);

/* Width/height of the display region and of the sampled extended region
 * (display region grown by the filter's discrete shift on each side). */
rh_status rh_frame_size(const rh_ctx* ctx, int32_t* width, int32_t* height,
                        int32_t* ext_width, int32_t* ext_height,
                        int32_t* samples_per_pixel);

/* Filtered channels of the display region, 9 floats per pixel in the
 * order Ci.rgb, Oi.rgb, a, z, coverage (the channel buffer built by
 * CqBucketProcessor::FilterBucket, bucketprocessor.cpp:381-744). */
rh_status rh_read_channels(rh_ctx* ctx, float* out);

/* Quantised, dithered 8-bit rgba of the display region (what
 * CqDisplayRequest::FormatBucketForDisplay hands to DspyImageData,
 * ddmanager.cpp:1022-1118). */
rh_status rh_read_rgba8(rh_ctx* ctx, uint8_t* out);

/* Value types of display data, DspyImageData's (include/aqsis/ri/ndspy.h:69-76). */
enum {
    RH_DSPY_FLOAT32 = 1, RH_DSPY_UNSIGNED32 = 2, RH_DSPY_SIGNED32 = 3, RH_DSPY_UNSIGNED16 = 4,
    RH_DSPY_SIGNED16 = 5, RH_DSPY_UNSIGNED8 = 6, RH_DSPY_SIGNED8 = 7
};

/* The value type a display request gets for these options' Quantize
 * settings (selectDataFormat, ddmanager.cpp:249-283): RH_DSPY_FLOAT32 when
 * quantize_one is 0 (no quantisation), else the smallest unsigned
 * (quantize_min >= 0) or signed integer type holding [quantize_min,
 * quantize_max]; *bytes: the bytes of one value.  Host only. */
rh_status rh_display_format(const rh_options* opt, int32_t* type, int32_t* bytes);

/* The rendered frame's r, g, b, a as a display request of value type `type`
 * (RH_DSPY_*) receives them (FormatBucketForDisplay, ddmanager.cpp:
 * 1022-1118): lround(value * one + dither * draw) clamped to [min, max]
 * when quantize_one != 0, then stored as `type` (a narrower type keeps the
 * low bits, as the reference's conversions do on x86-64).  4 values per
 * pixel, width x height pixels row by row.  For RH_DSPY_UNSIGNED8 the bytes
 * equal rh_read_rgba8's. */
rh_status rh_read_display(rh_ctx* ctx, int32_t type, void* out);

/* rh_read_display's conversion on host data, for a caller holding channels
 * (rh_read_channels' layout, 9 floats per pixel) and the per-pixel dither
 * draws of a display (rh_sample_positions' dither): npix pixels, 4 values
 * of `type` each into out.  Host only: needs no device. */
rh_status rh_format_display(const rh_options* opt, const float* channels, const float* dither, int64_t npix,
                            int32_t type, void* out);

/* A display request's mode (RiDisplay's mode string, ri.cpp:947-964):
 * "rgb" RH_MODE_RGB, "a" RH_MODE_A, "z" RH_MODE_Z, combined as "rgbaz". */
enum { RH_MODE_RGB = 1, RH_MODE_A = 2, RH_MODE_Z = 4 };

/* selectDataFormat (ddmanager.cpp:249-283) of a display request's own
 * Quantize settings: *type as rh_display_format's. */
rh_status rh_select_format(float one, float min, float max, int32_t* type);

/* The rendered frame as display request `index` (its place among the
 * num_displays requests, whose dither draws it takes) of mode `mode`
 * receives it (CqDisplayRequest::FormatBucketForDisplay,
 * ddmanager.cpp:1022-1118): per pixel, the request's channels in the order
 * it builds its formats (ddmanager.cpp:456-483): a, then r, g, b, then z
 * (the filtered depth), quantised with quantize = (one, min, max, dither)
 * as rh_read_display does, stored as `type`; width x height pixels row by
 * row.  The settings a request takes (ddmanager.cpp:429-455): its RiDisplay
 * "quantize"/"dither" parameters, else Quantize "z" when its mode has z,
 * else Quantize "rgba" (rh_options' quantize_*).  No counterpart of its own
 * in the reference beyond the data path it restates; for depth images
 * (mode "z") and the z channel of "rgbaz" requests. */
rh_status rh_read_display_request(rh_ctx* ctx, int32_t mode, const float* quantize, int32_t index, int32_t type,
                                  void* out);

/* A display driver's image-data entry point, with DspyImageData's argument
 * meaning (include/aqsis/ri/ndspy.h:158): the rectangle [xmin, xmaxplus1) x
 * [ymin, ymaxplus1), the bytes of one pixel, and the pixels row by row.  A
 * non-zero return stops the delivery. */
typedef int (*rh_dspy_data_fn)(void* image, int xmin, int xmaxplus1, int ymin, int ymaxplus1, int entry_size,
                               const unsigned char* data);

/* Hand the rendered frame's rgba to a display driver the way
 * CqDisplayRequest::DisplayBucket does (ddmanager.cpp:984-1019, 1129-1175):
 * bucket by bucket in the render's bucket order, each bucket's display
 * rectangle (imagebuffer.cpp:745-748, bucketprocessor.cpp:107); or, with
 * scanline_order (PkDspyFlagsWantsScanLineOrder), one row at a time as each
 * row of buckets completes.  Values are of the type rh_display_format
 * selects (entries of 4 values: r, g, b, a; rh_read_display's values).
 * Rectangles are in raster coordinates of the whole image.  With a crop
 * window, buckets wholly outside it are skipped (CqDDManager::DisplayBucket,
 * ddmanager.cpp:146-171) and each rectangle is the bucket's part inside the
 * crop window, rows the crop window's width: the reference also sends a
 * bucket's pixels outside the crop window, whose values it never computes. */
rh_status rh_display(rh_ctx* ctx, rh_dspy_data_fn data, void* image, int32_t scanline_order);

/* rh_display with the values as `type` (RH_DSPY_*), for a driver that asks
 * for another type than the one selected (the file driver turns signed
 * types into unsigned ones, tools/displays/file/display.cpp:455-467). */
rh_status rh_display_as(rh_ctx* ctx, rh_dspy_data_fn data, void* image, int32_t scanline_order, int32_t type);

/* Compression of rh_tiff_write: TIFF's Compression tag values. */
enum { RH_TIFF_NONE = 1, RH_TIFF_LZW = 5 };

/* Write an 8-bit image to a TIFF file as the "file"/"tiff" display driver's
 * integer branch does (tools/displays/file/display.cpp:227-350, WriteTIFF,
 * with the origin and original size set at open, :549-551, 590-593): rgb or rgba
 * (channels 3 or 4; rgba carries ExtraSamples = associated alpha), rows top
 * to bottom, contiguous samples, strips of libtiff's default height,
 * uncompressed or LZW (the driver's default), resolution unit none at 1/1,
 * XPosition/YPosition = origin, PIXAR_IMAGEFULLWIDTH/LENGTH = full_size.
 * origin and full_size may be NULL (0,0 and width,height); description NULL
 * gives a fixed text.  Host only: needs no device.  The pixel data are the
 * driver's; the tags it takes from the scene (world-to-camera and
 * world-to-screen matrices, host name) are not written. */
rh_status rh_tiff_write(const char* path, const uint8_t* pixels, int32_t width, int32_t height,
                        int32_t channels, int32_t compression, const int32_t* origin,
                        const int32_t* full_size, const char* description);

/* rh_tiff_write for the driver's other sample formats (WriteTIFF's second
 * branch, display.cpp:347-411, which adds SampleFormat and StoNits = 1):
 * type RH_DSPY_UNSIGNED8 as rh_tiff_write; RH_DSPY_UNSIGNED16 as 16-bit
 * unsigned samples; RH_DSPY_FLOAT32 as 32-bit IEEE floating point samples;
 * RH_DSPY_UNSIGNED32 as the driver writes it, 32-bit samples tagged as
 * floating point.  pixels: width x height x channels values of that type,
 * little-endian. */
rh_status rh_tiff_write_typed(const char* path, const void* pixels, int32_t width, int32_t height,
                              int32_t channels, int32_t type, int32_t compression, const int32_t* origin,
                              const int32_t* full_size, const char* description);

/* Per-sample visibility of the extended region, sample s of pixel
 * (x, y) at index (y*ext_width + x)*samples_per_pixel + s:
 *   mp_id[]  id of the front-most micropolygon hit (0xffffffff: none);
 *            ids count busted micropolygons in dicing order;
 *   depth[]  camera-space depth of that hit (FLT_MAX: none); for a sample
 *            without transparent hits under the midpoint depth filter,
 *            the combine's depth (the mean of that hit's depth and the
 *            occluding depth, imagepixel.cpp:319-326);
 *   rgbaz[]  7 floats per sample after CqImagePixel::Combine
 *            (imagepixel.cpp:144-332): Ci.rgb, Oi.rgb, z.
 * Any pointer may be NULL. */
rh_status rh_read_samples(rh_ctx* ctx, uint32_t* mp_id, float* depth,
                          float* rgbaz);

rh_status rh_get_stats(const rh_ctx* ctx, rh_stats* out);

/* Host only, no device needed: the split/dice schedule of a scene and the
 * memory one frame of it occupies (used for admission by footprint). */
rh_status rh_estimate(const rh_options* opt, const rh_prim* prims, int32_t n,
                      rh_footprint* out);

/* Host only: the frame's jittered sample positions, 2 floats per sample in
 * the rh_read_samples layout, and display 0's dither value per display
 * pixel (either pointer may be NULL).  This is the sample pattern that
 * CqImagePixel::setSamples builds (imagepixel.cpp:334-359) and that
 * aqsis's mpdump writes with dumpPixelSamples (mpdump.cpp). */
rh_status rh_sample_positions(const rh_options* opt, float* xy, float* dither);

void rh_destroy(rh_ctx* ctx);

const char* rh_last_error(void);

/*
 * Serving (no counterpart in the reference, which renders one frame per
 * process): render jobs of any frame shape are submitted as they arrive and
 * run concurrently on one GPU, each on its own context and HIP stream.
 * Contexts stay resident after a job, keyed by frame shape, so the next job
 * of that shape reuses its sample plan and device buffers.  Jobs are
 * admitted in arrival order when one of `slots` workers is free and the
 * device memory of the resident contexts plus the job's rh_estimate
 * footprint fits `memory_budget` (idle contexts are evicted oldest first).
 * A job's result does not depend on what else runs.
 */
typedef struct rh_server_options {
    int32_t slots;             /* jobs in flight at once (1..64)                */
    int32_t schedule_threads;  /* host threads for job schedules in all (0: 16) */
    uint64_t memory_budget;    /* device bytes for resident contexts            */
} rh_server_options;

typedef struct rh_job_info {
    uint64_t id;
    int32_t status;            /* rh_status of the job                          */
    int32_t state;             /* 0 queued, 1 running, 2 done                   */
    int32_t width, height;     /* display size                                  */
    uint64_t device_bytes;     /* footprint estimate used for admission         */
    uint64_t spl_count;        /* sample tests of the frame                     */
    double ms_queue;           /* submit -> admitted                            */
    double ms_run;             /* admitted -> done (host schedule + device)     */
    double ms_total;           /* submit -> done                                */
    double ms_device;          /* device frame time                             */
    double ms_setup;           /* frame-shape context created for the job (0 when
                                  a resident one was reused)                    */
    double ms_upload;          /* host->device copies of the job's inputs       */
    double ms_render;          /* inputs posted, uploaded and rendered          */
    double ms_readback;        /* channels and 8-bit image copied back          */
} rh_job_info;

typedef struct rh_server rh_server;

rh_status rh_server_create(const rh_server_options* opt, rh_server** out);
/* Queue a job and return at once; inputs are copied.  Its split/dice
 * schedule is computed by the server's scheduler threads; a job that alone
 * exceeds the budget ends with RH_ERR_CAPACITY (reported by wait / poll). */
rh_status rh_server_submit(rh_server* srv, const rh_options* opt, const rh_prim* prims, int32_t n,
                           uint64_t* id);
/* Block until job `id` is done; the job's own status is returned. */
rh_status rh_server_wait(rh_server* srv, uint64_t id, rh_job_info* info);
/* Finished jobs not yet reported by wait or poll (up to max). */
rh_status rh_server_poll(rh_server* srv, rh_job_info* infos, int32_t max, int32_t* count);
/* Copy a finished job's channels (9 floats a pixel, as rh_read_channels)
 * and rgba8; either pointer may be NULL. */
rh_status rh_server_read(rh_server* srv, uint64_t id, float* channels, uint8_t* rgba8);
/* Forget a finished job and its host results. */
rh_status rh_server_release(rh_server* srv, uint64_t id);
/* Device bytes held by resident contexts now, and the most ever held. */
rh_status rh_server_memory(rh_server* srv, uint64_t* resident, uint64_t* high_water);
/* Finish running jobs, fail queued ones, free everything. */
void rh_server_destroy(rh_server* srv);
const char* rh_server_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* AQSIS_HIDER_H */
