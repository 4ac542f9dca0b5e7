// td2d_vec.hpp -- vectorised batched step kernel (included by td2d_hip.hip).
//
// Same per-cell arithmetic as step_tile (bit-identical results), organised
// for HBM throughput:
//   * a workgroup (256 threads) owns a (4*CX4) x (256/CX4*RYV) tile (CX4 =
//     16 or 32 float4 columns); each thread owns 4 consecutive columns (one
//     float4) of RYV consecutive rows;
//   * every global load the thread needs (u_prev, model, virt, grad) is
//     issued first, for all its rows, before the current wavefield is staged
//     in LDS, so that many 16-byte loads per lane are in flight;
//   * the stencil reads the staged tile (halo 4 columns / max(ne,2) rows);
//   * the model enters as m = (v*v)*dt^2/h^2 and c3 = cons/((v*v)*v), the
//     exact float products the step would form, computed once per model;
//   * tiles are dealt spatial-tile-major, shot-minor and banded per XCD, so
//     the model (shared by the shots of one grid) stays in each XCD's L2.
// Requirements (checked by the host): row stride ld % 4 == 0 and 16-byte
// aligned field bases.
#pragma once

namespace tw {

constexpr int VTHR = 256;   // threads of the one-step batched kernel

// Loads/stores through the global address space: the pointers come from a
// descriptor in memory, so the compiler cannot infer it and would emit flat
// accesses.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) f4v gf4v;
__device__ __forceinline__ float4 gld4(const float* p)
{
    const f4v v = *(const gf4v*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst4(float* p, float4 v)
{
    f4v t = {v.x, v.y, v.z, v.w};
    *(gf4v*)p = t;
}
// Streaming (nontemporal) variants for data touched once per pass: virtual
// sources (written by the forward pass, read once by the adjoint) and
// checkpoints.
__device__ __forceinline__ float4 gld4_nt(const float* p)
{
    const f4v v = __builtin_nontemporal_load((const gf4v*)p);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst4_nt(float* p, float4 v)
{
    f4v t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, (gf4v*)p);
}
// Scalar global-address-space accesses (gathers, residuals, wavelet): a
// plain store through a descriptor pointer compiles to a FLAT store, and a
// pending FLAT operation makes every later wait on the memory counters a wait
// for all outstanding operations.
typedef __attribute__((address_space(1))) float gf1;
__device__ __forceinline__ float gld1(const float* p) { return *(const gf1*)p; }
__device__ __forceinline__ void gst1(float* p, float v) { *(gf1*)p = v; }

__device__ __forceinline__ float4 sel4(bool c, float4 a)
{
    return make_float4(c ? a.x : 0.f, c ? a.y : 0.f, c ? a.z : 0.f, c ? a.w : 0.f);
}

__device__ __forceinline__ float f4get(const float4& v, int j)
{
    return j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
}
__device__ __forceinline__ void f4set(float4& v, int j, float x)
{
    if (j == 0) v.x = x;
    else if (j == 1) v.y = x;
    else if (j == 2) v.z = x;
    else v.w = x;
}

template <int NE, unsigned FLAGS, int RYV, int CX4>
__global__ __launch_bounds__(VTHR) void k_vstep(const ShotDesc* __restrict__ descs,
                                                const int2* __restrict__ tiles, StepArgs a)
{
    constexpr int HALO = NE < 2 ? 2 : NE;   // Keys reads two cells inward
    constexpr int TR = VTHR / CX4;          // thread rows
    constexpr int TYV = TR * RYV;
    constexpr int W4 = CX4 + 2;             // staged row: 4 halo columns each side
    constexpr int W = 4 * W4;               // in floats
    constexpr int H = TYV + 2 * HALO;
    __shared__ float4 T4[H * W4];

    const int L = logical_tile(blockIdx.x, a.per_xcd);
    if (L >= a.ntiles || !resolve_step(a)) return;
    const int2 e = tiles[L];
    const ShotDesc& d = descs[e.x];
    const int x0 = (e.y & 0xffff) * 4 * CX4, y0 = (e.y >> 16) * TYV;
    const int dimx = d.dimx, dimy = d.dimy;
    const long long ld = d.ld;

    const int k0 = a.ring % 3, k1 = (a.ring + 1) % 3, k2 = (a.ring + 2) % 3;
    const float* __restrict__ um = a.adj ? pick3(d.lam, k0) : pick3(d.u, k0);
    const float* __restrict__ uo = a.adj ? pick3(d.lam, k1) : pick3(d.u, k1);
    float* __restrict__ up = a.adj ? pick3(d.lam, k2) : pick3(d.u, k2);

    const int tx = threadIdx.x % CX4, ty = threadIdx.x / CX4;
    const int gx = x0 + 4 * tx;
    const int gy0 = y0 + ty * RYV;
    const bool colok = gx < dimx;

    // 1. global loads of this thread's rows.  Out-of-grid lanes load cell 0
    // (always valid) and discard it: a conditional load would be turned into
    // a select of addresses through scratch memory.
    float4 pum[RYV], pm[RYV], pc3[RYV], pvs[RYV], pg[RYV];
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const float* __restrict__ vslot = d.virt + a.vslot * d.fstride;
#pragma unroll
    for (int r = 0; r < RYV; ++r) {
        const int gy = gy0 + r;
        const bool ok = colok && gy < dimy;
        const long long i = ok ? gy * ld + gx : 0;
        pum[r] = gld4(um + i);
        pm[r] = gld4(d.vm + i);
        if constexpr ((FLAGS & F_VIRT) != 0) pc3[r] = gld4(d.vc3 + i);
        if constexpr ((FLAGS & F_GRAD) != 0) {
            pvs[r] = gld4_nt(vslot + i);
            pg[r] = gld4(d.grad + i);
        }
    }

    // 2. stage u_cur (tile + halo) in LDS
    for (int k = threadIdx.x; k < H * W4; k += VTHR) {
        const int row = k / W4;
        const int c4 = k - row * W4;
        const int gyy = y0 - HALO + row, gxx = x0 - 4 + 4 * c4;
        const bool ok = gyy >= 0 && gyy < dimy && gxx >= 0 && gxx < dimx;
        T4[k] = sel4(ok, gld4(uo + (ok ? gyy * ld + gxx : 0)));
    }
    __syncthreads();
    if (!colok) return;

    const float* Tf = (const float*)T4;
    const bool xin = gx >= NE && gx + 3 < dimx - NE;   // all 4 columns interior
#pragma unroll
    for (int r = 0; r < RYV; ++r) {
        const int gy = gy0 + r;
        if (gy >= dimy) break;
        const long long i = gy * ld + gx;
        const float* Tc = Tf + (ty * RYV + r + HALO) * W + 4 * tx + 4;
        float4 un = z4;
        if (xin && gy >= NE && gy < dimy - NE) {
            un = interior_update4<NE>(Tc, W, pum[r], pm[r]);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int x = gx + j;
                if (x >= dimx) continue;
                const long long ij = i + j;
                const float um0 = f4get(pum[r], j);
                const float uo0 = Tc[j];
                float val;
                if (gy < NE) {
                    val = 0.f;
                } else if (gy >= dimy - NE) {
                    val = keys_update(um0, uo0, um[ij - ld], Tc[j - W], Tc[j - 2 * W], gld1(d.vel + ij), d.h2, d.dt2, d.hdt);
                } else if (x < NE) {
                    val = keys_update(um0, uo0, um[ij + 1], Tc[j + 1], Tc[j + 2], gld1(d.vel + ij), d.h2, d.dt2, d.hdt);
                } else if (x >= dimx - NE) {
                    val = keys_update(um0, uo0, um[ij - 1], Tc[j - 1], Tc[j - 2], gld1(d.vel + ij), d.h2, d.dt2, d.hdt);
                } else {
                    val = interior_update<NE>(Tc + j, W, um0, f4get(pm[r], j));
                }
                f4set(un, j, val);
            }
        }
        if constexpr ((FLAGS & F_SRC) != 0) {
            if (d.isrc >= i && d.isrc < i + 4) {
                const float add = inject_term(gld1(d.w + a.it), gld1(d.vel + d.isrc), d.dt2);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (d.isrc == i + j) f4set(un, j, f4get(un, j) + add);
            }
        }
        const bool onrow = gy == d.rrow;
        if constexpr ((FLAGS & F_RESID) != 0) {
            if (onrow) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int x = gx + j;
                    if (x >= d.rx0 && x < d.rx0 + d.nrec)
                        f4set(un, j, f4get(un, j) + inject_term(gld1(d.resid + (long long)a.it * d.nrec + (x - d.rx0)),
                                                                gld1(d.vel + i + j), d.dt2));
                }
            }
        }
        // columns beyond dimx carry zeros (they are never read as data)
        if (gx + 3 >= dimx) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (gx + j >= dimx) f4set(un, j, 0.f);
        }
        gst4(up + i, un);
        if constexpr ((FLAGS & F_REC) != 0) {
            if (onrow) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int x = gx + j;
                    if (x >= d.rx0 && x < d.rx0 + d.nrec)
                        gst1(d.frd + (long long)a.it * d.nrec + (x - d.rx0), f4get(un, j));
                }
            }
        }
        if constexpr ((FLAGS & F_VIRT) != 0) {
            float4 vs;
            vs.x = pc3[r].x * ((pum[r].x - 2.f * Tc[0]) + un.x);
            vs.y = pc3[r].y * ((pum[r].y - 2.f * Tc[1]) + un.y);
            vs.z = pc3[r].z * ((pum[r].z - 2.f * Tc[2]) + un.z);
            vs.w = pc3[r].w * ((pum[r].w - 2.f * Tc[3]) + un.w);
            gst4_nt(d.virt + a.vslot * d.fstride + i, vs);
        }
        if constexpr ((FLAGS & F_GRAD) != 0) {
            float4 g;
            g.x = pg[r].x + (-pvs[r].x) * un.x;
            g.y = pg[r].y + (-pvs[r].y) * un.y;
            g.z = pg[r].z + (-pvs[r].z) * un.z;
            g.w = pg[r].w + (-pvs[r].w) * un.w;
            gst4(d.grad + i, g);
        }
        if (a.ck0 >= 0) gst4_nt(d.ckpt + (2LL * a.ck0) * d.fstride + i, un);
        if (a.ck1 >= 0) gst4_nt(d.ckpt + (2LL * a.ck1 + 1) * d.fstride + i, un);
    }
}

// Derived model arrays of one padded model: m = (v*v)*dtoh2 and
// c3 = cons/((v*v)*v), exactly the float products the step forms.  Keys
// boundary cells (the bottom ne rows, and the left/right ne columns below the
// top ne rows), which never use m, carry v itself in m, so that a kernel
// reads one model value per cell whatever its region.  Columns at or beyond
// dimx (row padding) get zeros.
__global__ void k_model_prep(const float* __restrict__ v, float* __restrict__ m,
                             float* __restrict__ c3, int dimx, int dimy, int ld, int ne,
                             float dtoh2, float cons)
{
    const long long n = (long long)dimy * ld;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % ld);
        const int y = (int)(i / ld);
        const float vv = v[i];
        const bool in = x < dimx;
        const bool keys = y >= dimy - ne || (y >= ne && (x < ne || x >= dimx - ne));
        m[i] = in ? (keys ? vv : (vv * vv) * dtoh2) : 0.f;
        c3[i] = in ? cons / ((vv * vv) * vv) : 0.f;
    }
}

}  // namespace tw
