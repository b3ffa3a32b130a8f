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
    const float one = o.quantize_one, mn = o.quantize_min, mx = o.quantize_max;
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

void format_pixel(const rh_options& o, const float* ch, double s, int type, unsigned char* dst) {
    const float vals[4] = {ch[0], ch[1], ch[2], ch[6]};
    const float zero = 0.0f;
    for (int k = 0; k < 4; ++k) {
        double value = vals[k];
        if (o.quantize_one != 0) {
            value = (double)round_l(zero + value * (double)(o.quantize_one - zero) + ((double)o.quantize_dither * s));
            value = clampv<double>(value, o.quantize_min, o.quantize_max);
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
}

}  // namespace rh
