// td2d_arith.hpp -- per-cell arithmetic of the time-domain 2D acoustic
// propagator, written so that every rounding step matches the reference's
// CPU path (torchwi/propagator/td2d/td2d_cpu.cpp, propagator_cpu.cpp):
//
//   * neighbour sums are float, added in the reference's order
//     (order 2: -1, -dimx, +1, +dimx; orders 4/8: -k*dimx, -k, +k, +k*dimx);
//   * the weighted Laplacian is accumulated in double, outermost ring first,
//     then rounded to float (td2d_cpu.cpp:23-25, :52-54, :90-94);
//   * the time update is float: ((2u - u_prev) + m*lap) for order 2
//     (:26) and ((lap*m + 2u) - u_prev) for orders 4/8 (:55, :95),
//     with m = (v*v)*dtoh2;
//   * Keys: uxx double rounded to float, uxt float, the combination double
//     (td2d_cpu.cpp:122-124).
//
// The translation unit must be compiled with -ffp-contract=off so that no
// multiply-add is fused, and with correctly rounded fp32 division.
#pragma once
#include <hip/hip_runtime.h>

namespace tw {

// Central-difference second-derivative weights (td2d_cpu.cpp:4-5, :31-33,
// :62-66): w[0] centre, w[k] the +-k neighbours.  Textbook values
// -2,1; -5/2,4/3,-1/12; -205/72,8/5,-1/5,8/315,-1/560 written as the
// reference's double literals.
template <int NE> struct LapW;
template <> struct LapW<1> {
    static __device__ __host__ constexpr double w(int k) { return k == 0 ? -2.0 : 1.0; }
};
template <> struct LapW<2> {
    static __device__ __host__ constexpr double w(int k)
    {
        return k == 0 ? -2.5 : (k == 1 ? 1.3333333333333333 : -0.08333333333333333);
    }
};
template <> struct LapW<4> {
    static __device__ __host__ constexpr double w(int k)
    {
        return k == 0 ? -2.8472222222222223
             : k == 1 ? 1.6
             : k == 2 ? -0.2
             : k == 3 ? 0.025396825396825397
                      : -0.0017857142857142857;
    }
};

// Keys two-angle ABC constants: cos(pi/6)cos(pi/12), cos(pi/6)+cos(pi/12)
// (td2d_cpu.cpp:107-108).
constexpr double KEYS_CM = 0.8365163037378079;
constexpr double KEYS_CP = 1.831951230073507;

// Interior leapfrog update.  T points at the centre cell of a tile with row
// stride ts; u1 is the previous-step value of the cell; m = (v*v)*dtoh2.
template <int NE>
__device__ __forceinline__ float interior_update(const float* T, int ts, float u1, float m)
{
    const float c = T[0];
    if constexpr (NE == 1) {
        float s = ((T[-1] + T[-ts]) + T[1]) + T[ts];
        float lap = (float)(LapW<1>::w(1) * (double)s + (LapW<1>::w(0) * 2.0) * (double)c);
        return (2.f * c - u1) + m * lap;
    } else {
        float s = ((T[-NE * ts] + T[-NE]) + T[NE]) + T[NE * ts];
        double acc = LapW<NE>::w(NE) * (double)s;
#pragma unroll
        for (int k = NE - 1; k >= 1; --k) {
            s = ((T[-k * ts] + T[-k]) + T[k]) + T[k * ts];
            acc = acc + LapW<NE>::w(k) * (double)s;
        }
        acc = acc + (LapW<NE>::w(0) * 2.0) * (double)c;
        float lap = (float)acc;
        return (lap * m + 2.f * c) - u1;
    }
}

// The same update for four consecutive cells x..x+3 of one row, reading the
// staged field as float4s: rows y-k and y+k at x..x+3 (2*NE reads) and the
// centre row at x-4..x+7 (3 reads).  T must be 16-byte aligned and the row
// stride ts a multiple of 4.  The per-cell arithmetic and its order are
// exactly interior_update's.
__device__ __forceinline__ float f4at(const float4& v, int j)
{
    return j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
}

// The 4-wide update from values already in registers: up[k-1] / dn[k-1] are
// the float4s k rows above / below, L, C, R the centre row at x-4, x, x+4.
template <int NE>
__device__ __forceinline__ float4 lap_update4(const float4 (&up)[NE], const float4 (&dn)[NE],
                                              float4 L, float4 C, float4 R, float4 u1, float4 m)
{
    // centre-row value at offset o in [-4, 7]
    auto row = [&](int o) -> float { return o < 0 ? f4at(L, o + 4) : (o < 4 ? f4at(C, o) : f4at(R, o - 4)); };
    float out[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float c = f4at(C, j);
        const float um = f4at(u1, j), mm = f4at(m, j);
        if constexpr (NE == 1) {
            float s = ((row(j - 1) + f4at(up[0], j)) + row(j + 1)) + f4at(dn[0], j);
            float lap = (float)(LapW<1>::w(1) * (double)s + (LapW<1>::w(0) * 2.0) * (double)c);
            out[j] = (2.f * c - um) + mm * lap;
        } else {
            float s = ((f4at(up[NE - 1], j) + row(j - NE)) + row(j + NE)) + f4at(dn[NE - 1], j);
            double acc = LapW<NE>::w(NE) * (double)s;
#pragma unroll
            for (int k = NE - 1; k >= 1; --k) {
                s = ((f4at(up[k - 1], j) + row(j - k)) + row(j + k)) + f4at(dn[k - 1], j);
                acc = acc + LapW<NE>::w(k) * (double)s;
            }
            acc = acc + (LapW<NE>::w(0) * 2.0) * (double)c;
            float lap = (float)acc;
            out[j] = (lap * mm + 2.f * c) - um;
        }
    }
    return make_float4(out[0], out[1], out[2], out[3]);
}

template <int NE>
__device__ __forceinline__ float4 interior_update4(const float* T, int ts, float4 u1, float4 m)
{
    float4 up[NE], dn[NE];
#pragma unroll
    for (int k = 1; k <= NE; ++k) {
        up[k - 1] = *(const float4*)(T - k * ts);
        dn[k - 1] = *(const float4*)(T + k * ts);
    }
    return lap_update4<NE>(up, dn, *(const float4*)(T - 4), *(const float4*)(T), *(const float4*)(T + 4), u1, m);
}

// Keys update of a boundary cell from its own (um0, uo0) and its first and
// second inward neighbours (um1, uo1, uo2).
__device__ __forceinline__ float keys_update(float um0, float uo0, float um1, float uo1,
                                             float uo2, float v, float h2, float dt2,
                                             float hdt)
{
    float uxx = (float)(((double)uo0 - 2.0 * (double)uo1 + (double)uo2) / (double)h2);
    float uxt = ((uo0 - um0) - (uo1 - um1)) / hdt;
    float a = ((-v) * v) * dt2;
    double r = (double)a / KEYS_CM * ((double)uxx + KEYS_CP / (double)v * (double)uxt)
             + 2.0 * (double)uo0 - (double)um0;
    return (float)r;
}

// Point-source / residual injection term ((w*v)*v)*dt2
// (propagator_cpu.cpp:62, :138).
__device__ __forceinline__ float inject_term(float a, float v, float dt2)
{
    return ((a * v) * v) * dt2;
}

// Virtual source of the imaging condition, (cons/((v*v)*v)) * ((u1-2u2)+u3)
// (propagator_cpu.cpp:70-71).
__device__ __forceinline__ float virt_source(float cons, float v, float u1, float u2, float u3)
{
    float v3 = (v * v) * v;
    return (cons / v3) * ((u1 - 2.f * u2) + u3);
}

// The EXA kernel's own virtual-source expression,
// cons/(v*v*v) * ((u3 - 2u2) + u1) (td2d_cuda.cu:189).
__device__ __forceinline__ float virt_source_exa(float cons, float v, float u1, float u2, float u3)
{
    return (cons / ((v * v) * v)) * ((u3 - 2.f * u2) + u1);
}

}  // namespace tw
