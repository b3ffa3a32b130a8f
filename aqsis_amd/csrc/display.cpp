// Display data formats (see display.h).  Host code, compiled like the
// reference's display manager: x86-64 SSE doubles, no FMA contraction.
#include "display.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>

#include "hmath.h"

namespace rh {

int display_type(const rh_options& o) {
    return select_data_format(o.quantize_one, o.quantize_min, o.quantize_max);
}

int select_data_format(float one, float mn, float mx) {
    if (one == 0) return RH_DSPY_FLOAT32;
    if (mn >= 0) {
        if (mx <= 255) return RH_DSPY_UNSIGNED8;
        if (mx <= 65535) return RH_DSPY_UNSIGNED16;
        return RH_DSPY_UNSIGNED32;
    }
    if (mn >= -128 && mx <= 127) return RH_DSPY_SIGNED8;
    if (mn >= -32768 && mx <= 32767) return RH_DSPY_SIGNED16;
    return RH_DSPY_SIGNED32;
}

int display_type_bytes(int type) {
    switch (type) {
    case RH_DSPY_FLOAT32: case RH_DSPY_UNSIGNED32: case RH_DSPY_SIGNED32: return 4;
    case RH_DSPY_UNSIGNED16: case RH_DSPY_SIGNED16: return 2;
    case RH_DSPY_UNSIGNED8: case RH_DSPY_SIGNED8: return 1;
    default: return 0;
    }
}

namespace {
// x86-64's conversion of a double to an integer narrower than 32 bits: the
// 32-bit truncating conversion, then the low bits.  The values reaching it
// are whole numbers in [min, max], or the clamp minimum.
int32_t to_i32(double v) {
    if (!(v >= -2147483648.0 && v < 2147483648.0)) return INT32_MIN;   // cvttsd2si's indefinite value
    return (int32_t)v;
}
}  // namespace

// One value: quantised with q = (one, min, max, dither) and draw s when one
// != 0, stored as `type`; returns the next byte.
static unsigned char* put_value(const float* q, double value, double s, int type, unsigned char* dst) {
    const float zero = 0.0f;
    {
        if (q[0] != 0) {
            value = (double)round_l(zero + value * (double)(q[0] - zero) + ((double)q[3] * s));
            value = clampv<double>(value, q[1], q[2]);
        }
        switch (type) {
        case RH_DSPY_FLOAT32: {
            const float f = (float)value;
            std::memcpy(dst, &f, 4);
            dst += 4;
            break;
        }
        case RH_DSPY_UNSIGNED32: {
            // clamped in double first (ddmanager.cpp:1083-1093)
            value = clampv<double>(value, 0, (double)std::numeric_limits<uint32_t>::max());
            const uint32_t u = (uint32_t)(int64_t)value;
            std::memcpy(dst, &u, 4);
            dst += 4;
            break;
        }
        case RH_DSPY_SIGNED32: {
            value = clampv<double>(value, (double)std::numeric_limits<int32_t>::min(),
                                   (double)std::numeric_limits<int32_t>::max());
            const int32_t i = (int32_t)value;
            std::memcpy(dst, &i, 4);
            dst += 4;
            break;
        }
        case RH_DSPY_UNSIGNED16: case RH_DSPY_SIGNED16: {
            const uint16_t u = (uint16_t)(uint32_t)to_i32(value);
            std::memcpy(dst, &u, 2);
            dst += 2;
            break;
        }
        default: {   // RH_DSPY_UNSIGNED8, RH_DSPY_SIGNED8
            *dst++ = (uint8_t)(uint32_t)to_i32(value);
            break;
        }
        }
    }
    return dst;
}

void format_pixel(const rh_options& o, const float* ch, double s, int type, unsigned char* dst) {
    const float q[4] = {o.quantize_one, o.quantize_min, o.quantize_max, o.quantize_dither};
    const float vals[4] = {ch[0], ch[1], ch[2], ch[6]};
    for (int k = 0; k < 4; ++k) dst = put_value(q, vals[k], s, type, dst);
}

int mode_channels(int mode) {
    if (mode <= 0 || mode > (RH_MODE_RGB | RH_MODE_A | RH_MODE_Z)) return 0;
    return ((mode & RH_MODE_A) ? 1 : 0) + ((mode & RH_MODE_RGB) ? 3 : 0) + ((mode & RH_MODE_Z) ? 1 : 0);
}

int format_request(const float* q, int mode, const float* ch, double s, int type, unsigned char* dst) {
    unsigned char* p = dst;
    if (mode & RH_MODE_A) p = put_value(q, ch[6], s, type, p);
    if (mode & RH_MODE_RGB)
        for (int k = 0; k < 3; ++k) p = put_value(q, ch[k], s, type, p);
    if (mode & RH_MODE_Z) p = put_value(q, ch[7], s, type, p);
    return (int)(p - dst);
}

}  // namespace rh
