// sl_device.h -- device-side arithmetic of the comb-scanline pipeline.
//
// Every function reproduces the reference's single-precision expression order exactly
// (the library is compiled with -ffp-contract=off, IEEE division and square root, fp32
// denormals preserved).  Where the reference leans on CUDA/PTX behaviour the rule is
// spelled out: float->int conversions saturate and map NaN to 0, shifts by a count
// outside [0, width) produce 0, round() is half away from zero.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <climits>

namespace sl {
namespace dev {

constexpr int CT_LINEAR = 2, CT_QUAD = 3, CT_CUBIC = 4, CT_RATIONAL = 0x13;

__device__ __forceinline__ uint32_t fb(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bf(uint32_t u) { return __uint_as_float(u); }

// Saturating conversion of an integral float: NaN -> 0, beyond the int range -> INT_MIN /
// INT_MAX.  Written as a clamp into range plus selects, so it compiles without branches.
__device__ __forceinline__ int32_t sat_i32(float f) {
    int32_t r = (int32_t)fminf(fmaxf(f, -2147483648.0f), 2147483520.0f);  // largest float < 2^31
    r = (f >= 2147483648.0f) ? INT_MAX : r;
    return (f != f) ? 0 : r;
}
// __float2int_rd semantics (ras_cut.h:639-657)
__device__ __forceinline__ int32_t floor_i32(float x) { return sat_i32(floorf(x)); }
__device__ __forceinline__ int32_t trunc_i32(float x) { return sat_i32(truncf(x)); }
__device__ __forceinline__ int32_t trunc_i32(double x) {
    double f = trunc(x);
    if (f != f) return 0;
    if (f >= 2147483648.0) return INT_MAX;
    if (f < -2147483648.0) return INT_MIN;
    return (int32_t)f;
}
__device__ __forceinline__ int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
__device__ __forceinline__ uint32_t lsr32(uint32_t v, int32_t s) { return (s < 0 || s >= 32) ? 0u : (v >> s); }
__device__ __forceinline__ uint64_t lsr64(uint64_t v, int32_t s) { return (s < 0 || s >= 64) ? 0ull : (v >> s); }
template <class T> __device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
template <class T> __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }

// Wave-wide inclusive scans with DPP row shifts and row broadcasts (no LDS round trips);
// lane 63 ends with the whole wave's reduction.  Every lane of the wave must be active.
struct OpAdd { __device__ int operator()(int a, int b) const { return a + b; } };
struct OpMin { __device__ int operator()(int a, int b) const { return a < b ? a : b; } };
struct OpMax { __device__ int operator()(int a, int b) const { return a > b ? a : b; } };
struct OpOr { __device__ int operator()(int a, int b) const { return a | b; } };
struct OpMinU { __device__ int operator()(int a, int b) const { return (uint32_t)a < (uint32_t)b ? a : b; } };
struct OpMaxU { __device__ int operator()(int a, int b) const { return (uint32_t)a > (uint32_t)b ? a : b; } };
template <class Op>
__device__ __forceinline__ int dpp_scan(int x, int ident, Op op) {
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x111, 0xf, 0xf, false));  // row_shr:1
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x112, 0xf, 0xf, false));  // row_shr:2
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x114, 0xf, 0xf, false));  // row_shr:4
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x118, 0xf, 0xf, false));  // row_shr:8
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x142, 0xa, 0xf, false));  // row_bcast:15
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return x;
}
// The same scan inside segments of S lanes: S = 16 keeps to the DPP rows, S = 32 adds the
// row_bcast:15 step (rows 1 and 3 take the last lane of rows 0 and 2), S = 64 is dpp_scan.
template <int S, class Op>
__device__ __forceinline__ int dpp_scan_seg(int x, int ident, Op op) {
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x111, 0xf, 0xf, false));  // row_shr:1
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x112, 0xf, 0xf, false));  // row_shr:2
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x114, 0xf, 0xf, false));  // row_shr:4
    x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x118, 0xf, 0xf, false));  // row_shr:8
    if (S >= 32) x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x142, 0xa, 0xf, false));  // row_bcast:15
    if (S >= 64) x = op(x, __builtin_amdgcn_update_dpp(ident, x, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return x;
}
// The reduction of x over the whole wave, in every lane.
template <class Op>
__device__ __forceinline__ int wave_reduce(int x, int ident, Op op) {
    return __builtin_amdgcn_readlane(dpp_scan(x, ident, op), 63);
}

__device__ __forceinline__ float safe_rcp(float a) { return (a != 0.f) ? (1.f / a) : 0.f; }
__device__ __forceinline__ float lerp1(float a, float b, float t) { return a + t * (b - a); }

// One coordinate of a curve at parameter t: de Casteljau lerps; the rational quadratic
// divides by its weight polynomial (interpolateGeneralCurve, ras_cut.h:586-632;
// d_curve_point, bezier_curve.h:57-152).
__device__ __forceinline__ float curve_eval(int type, float w, float v0, float v1, float v2, float v3, float t) {
    if (type == CT_LINEAR) return lerp1(v0, v1, t);
    if (type == CT_QUAD) {
        float l0 = lerp1(v0, v1, t), l1 = lerp1(v1, v2, t);
        return lerp1(l0, l1, t);
    }
    if (type == CT_CUBIC) {
        float q0 = lerp1(v0, v1, t), q1 = lerp1(v1, v2, t), q2 = lerp1(v2, v3, t);
        float l0 = lerp1(q0, q1, t), l1 = lerp1(q1, q2, t);
        return lerp1(l0, l1, t);
    }
    float iv1 = v1 * w;
    float l0 = lerp1(v0, iv1, t), l1 = lerp1(iv1, v2, t);
    float r = lerp1(l0, l1, t);
    float omt = 1.f - t;
    float ww = omt * omt + 2.f * omt * t * w + t * t;
    return r * safe_rcp(ww);
}

// solveQuadEquation (ras_cut.h:33-64)
__device__ __forceinline__ void solve_quad(float a, float b, float c, float &r0, float &r1) {
    if (a == 0.f) {
        float x = -c / b;
        r0 = x;
        r1 = x;
        return;
    }
    float B = b * 0.5f;
    float R = B * B - a * c;
    if (R > 0.0f) {
        float SR = sqrtf(R);
        if (B > 0.0f) {
            float TB = B + SR;
            r0 = -c / TB;
            r1 = -TB / a;
        } else {
            float TB = -B + SR;
            r0 = TB / a;
            r1 = c / TB;
        }
    } else {
        r0 = 0.f;
        r1 = 0.f;
    }
}

// solveQuadBezierQM (ras_cut.h:373-388)
__device__ __forceinline__ float solve_quad_bezier(float iA, float iB, float iC, float iD, float t0, float t1) {
    float a = iA + (-2.f) * iB + iC;
    float b = -2.f * iA + 2.f * iB;
    float c = iA - iD;
    float r0 = 0.f, r1 = 0.f;
    solve_quad(a, b, c, r0, r1);
    if (t0 <= r0 && r0 <= t1) return r0;
    if (t0 <= r1 && r1 <= t1) return r1;
    return t0;
}

// PATH_VISIBLE (ras_define.h:119-126) on the 8 region bits of a path.
__device__ __forceinline__ bool region_visible(uint32_t m) {
    return (m & 0xF8u) && (m & 0x6Bu) && (m & 0x1Fu) && (m & 0xD6u);
}

__device__ __forceinline__ int32_t floor_even(float x) { return floor_i32(x) & (int32_t)0xFFFFFFFE; }

// Texel of a point-sampled, clamped, normalized 256-wide texture.
__device__ __forceinline__ int texel(float u) { return tmin(tmax(floor_i32(u * 256.f), 0), 255); }

}  // namespace dev
}  // namespace sl
