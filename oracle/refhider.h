/* ORACLE -- test infrastructure only.  A CPU restatement of aqsis's hider
 * path (libs/core/{imagebuffer,bucketprocessor,micropolygon,imagepixel,...})
 * used by tests/, the smoke check and the CPU-baseline leg of bench.py to
 * check and time the HIP hider.  The product library never links it. */
#ifndef ORACLE_REFHIDER_H
#define ORACLE_REFHIDER_H

#include <stdint.h>
#include "../include/aqsis_hider.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Sizes of the display region and of the sampled extended region. */
int oracle_frame_size(const rh_options* opt, int32_t* width, int32_t* height,
                      int32_t* ext_width, int32_t* ext_height, int32_t* spp);

/* Render one frame on the CPU.  occlusion_cull = 1 follows the reference
 * exactly, including CqOcclusionTree surface culling and reposting
 * (bucketprocessor.cpp:947-960); 0 turns that culling off, which changes
 * only which occluded surfaces get diced and in what order.  Returns -2 for
 * depth of field together with motion keys (not restated).  Output
 * buffers follow the layouts of rh_read_channels / rh_read_rgba8 /
 * rh_read_samples in aqsis_hider.h; any output pointer may be NULL. */
int oracle_render(const rh_options* opt, const rh_prim* prims, int32_t nprims,
                  int32_t occlusion_cull, float* channels, uint8_t* rgba8,
                  uint32_t* mp_id, float* depth, float* rgbaz, rh_stats* stats);

/* Sample positions of the extended region after the frame's preProcess
 * passes (imagepixel.cpp:334-359), 2 floats per sample, and display 0's
 * dither draw per display pixel. */
int oracle_sample_positions(const rh_options* opt, float* xy, float* dither);
/* Shutter time of every sample of the extended image (ext pixels x spp). */
int oracle_sample_times(const rh_options* opt, float* times);

/* Lens offset (depth of field) of every sample of the extended image,
 * 2 floats per sample, and the lens cell it is visited through (may be NULL). */
int oracle_sample_lens(const rh_options* opt, float* xy, int32_t* cell);
/* Depth of field of the frame: circle of confusion at n camera depths
 * (coc[2n]), the lens-cell bounds (cells[4 * spp]) and MinCoCForBound of
 * [z[i], z[i+1]] (minCoc[n-1]).  -1 when the frame has no depth of field. */
int oracle_dof(const rh_options* opt, int32_t n, const float* z, float* coc, float* cells, float* minCoc);

/* Camera-space grid points of one surface diced at u x v
 * (CqSurface::Dice): out[3*((v+1)*(u+1))]. */
int oracle_dice(const rh_prim* prim, int32_t u, int32_t v, float* out);

/* --- Known-answer hooks for the unit tests ----------------------------- */
/* n outputs of the global stream after Reseed(seed) (random.cpp:102-157). */
void oracle_mt_outputs(uint32_t seed, int32_t n, uint32_t* out);
/* n RandomFloat() values after Reseed(seed). */
void oracle_random_floats(uint32_t seed, int32_t n, float* out);
/* lfloor / lceil / lround of math.h:47-70 on floats. */
void oracle_round_f(const float* x, int32_t n, int64_t* fl, int64_t* ce, int64_t* ro);
/* CqInvBilinear(A,B,C,D)(P) of bilinear.h:260-310; uv[2*i], for n points. */
void oracle_inv_bilinear(const float* abcd /* 8 */, const float* p, int32_t n, float* uv);
/* bilerp(A,B,C,D,uv) of bilinear.h:228-236 on 2D values. */
void oracle_bilerp2(const float* abcd /* 8 */, const float* uv, int32_t n, float* out);
/* CqOcclusionTree::treeIndexForPoint of occlusion.cpp:249-281. */
int32_t oracle_tree_index(int32_t numLevels, int32_t splitXFirst, int32_t x, int32_t y);
/* CqOcclusionTree::canCull on the tree of a w x h pixel sample region at
 * (sMinX, sMinY) with xS x yS samples per pixel whose leaves hold
 * cellDepth[y * w * xS + x] (sub-sample cells); bound = min x, min y, max x,
 * max y, min z, max z. */
int32_t oracle_can_cull(int32_t sMinX, int32_t sMinY, int32_t w, int32_t h, int32_t xS, int32_t yS,
                        const float* cellDepth, const float* bound);
/* RiXxxFilter(x, y, xw, yw) of filters.cpp. */
float oracle_filter(int32_t filter, float x, float y, float xw, float yw);
/* The 250-pattern multijitter tables built after Reseed(545)
 * (multijitter.h:90-101): pos2d[2*250*n], pos1d[250*n], shuffled[250*n]. */
void oracle_jitter_tables(int32_t nx, int32_t ny, float* pos2d, float* pos1d,
                          int32_t* shuffled);
/* Point-in-micropolygon (fContains, micropolygon.cpp:1295-1342) of the
 * static micropolygon with raster vertices P[4][3] in grid order
 * (iu,iv),(iu+1,iv),(iu,iv+1),(iu+1,iv+1): returns 1 and depth if hit. */
int32_t oracle_mp_sample(const float* P /* 12 */, float x, float y, float* depth);

/* selectDataFormat's PkDspy* code for the rgba request. */
int32_t oracle_display_type(const rh_options* opt);

/* FormatBucketForDisplay's values (npix x 4 of `type`) from the 9 channels
 * and the dither draw of each pixel. */
void oracle_display_values(const rh_options* opt, const float* channels, const float* dither, int64_t npix,
                           int32_t type, void* out);

#ifdef __cplusplus
}
#endif

#endif
