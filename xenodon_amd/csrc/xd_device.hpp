// Device-side helpers shared by the traversal kernels (gfx950).
//
// Every floating-point expression follows the GLSL source of the reference
// shaders operation by operation (resources/common.glsl, resources/*.comp):
// IEEE-754 single precision, one rounding per operation, no contraction
// (the library is compiled with -ffp-contract=off), correctly rounded
// division and square root (HIP's default).  GLSL's min/max/normalize/cross/
// mod are spelled out with their specification definitions so the result
// does not depend on how a driver happens to expand them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xd {

struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 splat(float s) { return f3{s, s, s}; }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 mul(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 divv(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ f3 muls(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 smul(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
// GLSL: min(x, y) = y < x ? y : x; max(x, y) = x < y ? y : x.  For finite
// operands the hardware v_min_f32 / v_max_f32 return the same value; they
// can differ only in the sign of a zero result (min(-0, +0)), and no signed
// zero produced here reaches a division or changes a stored pixel (t values
// only feed comparisons, subtractions and products that are summed into a
// non-negative total, and -0 and +0 both store as 0).
__device__ __forceinline__ float gmin(float x, float y) { return __builtin_fminf(x, y); }
__device__ __forceinline__ float gmax(float x, float y) { return __builtin_fmaxf(x, y); }
__device__ __forceinline__ f3 vmin(f3 a, f3 b) { return f3{gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
__device__ __forceinline__ f3 vmax(f3 a, f3 b) { return f3{gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
// common.glsl:58-64
__device__ __forceinline__ float min_elem(f3 v) { return gmin(v.x, gmin(v.y, v.z)); }
__device__ __forceinline__ float max_elem(f3 v) { return gmax(v.x, gmax(v.y, v.z)); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ f3 normalize(f3 v) {
    float len = __builtin_sqrtf(dot(v, v));
    return f3{v.x / len, v.y / len, v.z / len};
}
__device__ __forceinline__ float fabs_(float x) { return x < 0.0f ? -x : x; }
__device__ __forceinline__ float fsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
// GLSL mod(x, y) = x - y * floor(x / y)
__device__ __forceinline__ float gmod(float x, float y) { return x - y * __builtin_floorf(x / y); }

// c / 255 correctly rounded for c in [0, 255] without a division:
// fma(c, h, c * l) with h = fl(1/255) and l = fl(1/255 - h).  The sum
// c*h + fl(c*l) is within about 2^-47 (relative) of c/255, while c/255 lies
// at least about 2^-33 (relative) from every rounding boundary, so the
// fma's single rounding lands on the correctly rounded quotient.  Checked
// exhaustively in exact arithmetic for all 256 inputs.
constexpr float kInv255Hi = 0x1.010102p-8f;
constexpr float kInv255Lo = -0x1.fdfdfep-33f;
__device__ __forceinline__ float unorm8_to_float(uint32_t c) {
    const float x = (float)c;
    return __builtin_fmaf(x, kInv255Hi, x * kInv255Lo);
}
// unpackUnorm4x8(packed).rgb / texelFetch(RGBA8_UNORM).rgb
__device__ __forceinline__ f3 unorm_rgb(uint32_t packed) {
    return f3{unorm8_to_float(packed & 0xFFu), unorm8_to_float((packed >> 8) & 0xFFu),
              unorm8_to_float((packed >> 16) & 0xFFu)};
}

// imageStore to an rgba8 image: clamp to [0,1], scale by 255, round to
// nearest even; NaN stores 0; alpha is 1.0 -> 255.
__device__ __forceinline__ uint32_t unorm8_store(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 1.0f) return 255u;
    return (uint32_t)__builtin_rintf(v * 255.0f);
}
__device__ __forceinline__ uint32_t pack_rgba(f3 c) {
    return unorm8_store(c.x) | (unorm8_store(c.y) << 8) | (unorm8_store(c.z) << 16) | 0xFF000000u;
}

// common.glsl:40-43
__device__ __forceinline__ f3 adjust_ray(f3 rd) {
    const float eps = 1.0f / 8388608.0f;
    return f3{fabs_(rd.x) < eps ? eps : rd.x, fabs_(rd.y) < eps ? eps : rd.y, fabs_(rd.z) < eps ? eps : rd.z};
}

struct Camera {
    f3 forward, up, translation;
};

struct View {
    int32_t out_off[2];
    uint32_t out_ext[2];
    int32_t disp_off[2];
    uint32_t disp_ext[2];
    f3 ratio;
    uint32_t model_dim[3];
    float emission;
};

// common.glsl:45-56, with uv as computed by every main():
// vec2(pixel - display_region.offset) / vec2(display_region.extent)
__device__ __forceinline__ f3 ray_dir(const Camera& cam, const View& v, uint32_t ix, uint32_t iy) {
    int32_t pxl = v.out_off[0] + (int32_t)ix;
    int32_t pyl = v.out_off[1] + (int32_t)iy;
    float uvx = (float)(pxl - v.disp_off[0]) / (float)v.disp_ext[0];
    float uvy = (float)(pyl - v.disp_off[1]) / (float)v.disp_ext[1];
    uvx = uvx - 0.5f;
    uvy = uvy - 0.5f;
    uvy = uvy * ((float)v.disp_ext[1] / (float)v.disp_ext[0]);
    f3 dir = cam.forward;
    f3 up = cam.up;
    f3 right = normalize(cross(up, dir));
    up = normalize(cross(right, dir));
    f3 rd = normalize(add(add(smul(uvx, right), smul(uvy, up)), dir));
    return adjust_ray(normalize(divv(rd, v.ratio)));
}

// common.glsl:68-72
__device__ __forceinline__ float emission_coeff(const View& v, f3 rd) {
    f3 rd2 = mul(rd, rd);
    f3 dim2 = mul(v.ratio, v.ratio);
    return v.emission * __builtin_sqrtf(dot(rd2, dim2) / dot(rd2, splat(1.0f)));
}

__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }

}  // namespace xd
