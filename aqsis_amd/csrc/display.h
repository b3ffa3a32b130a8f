// Display data of a rendered frame: the value type a display request is
// given for the Quantize settings, and the per-value conversion of the
// frame's channels into it.
#pragma once

#include "../../include/aqsis_hider.h"

namespace rh {

// selectDataFormat (ddmanager.cpp:249-283): RH_DSPY_FLOAT32 without
// quantisation (one == 0), else the smallest unsigned (min >= 0) or signed
// integer type whose range holds [min, max].
int display_type(const rh_options& o);
// selectDataFormat of explicit Quantize settings (one, min, max).
int select_data_format(float one, float mn, float mx);

// Bytes of one value of a RH_DSPY_* type; 0 for a type that is not one.
int display_type_bytes(int type);

// FormatBucketForDisplay's values of one pixel (ddmanager.cpp:1055-1118):
// r, g, b from its channels 0..2 and a from channel 6 (ch: the pixel's 9
// channels), quantised as lround(value * one + dither * s) clamped to
// [min, max] when one != 0, and stored as `type` at dst (4 values).
void format_pixel(const rh_options& o, const float* ch, double s, int type, unsigned char* dst);

// A display request's values of one pixel: mode bits RH_MODE_* in the order
// the request builds its formats (ddmanager.cpp:456-483): a (channel 6),
// then r, g, b (0..2), then z (the filtered depth, channel 7); quantised
// with q = (one, min, max, dither) as format_pixel does.  Returns the bytes
// written.
int format_request(const float* q, int mode, const float* ch, double s, int type, unsigned char* dst);

// Values per pixel of a request's mode (0: not a mode).
int mode_channels(int mode);

}  // namespace rh
