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

// The update of the four cells gx..gx+3 of row gy (< dimy, gx < dimx) from
// u_prev (pum, and um for Keys neighbours) and the staged u_cur (Tc: cell
// (gx, gy) in a staged tile with rows of W floats): the interior stencil, the
// Keys strips (bottom rows first, corners included), zero top rows; columns
// at or beyond dimx are left 0.
template <int NE, int W>
__device__ __forceinline__ float4 cell4_update(const ShotDesc& d, const float* __restrict__ um, const float* Tc,
                                               long long i, int gx, int gy, float4 pum, float4 pm)
{
    const int dimx = d.dimx, dimy = d.dimy;
    const long long ld = d.ld;
    const bool xin = gx >= NE && gx + 3 < dimx - NE;   // all 4 columns interior
    float4 un = make_float4(0.f, 0.f, 0.f, 0.f);
    if (xin && gy >= NE && gy < dimy - NE) {
        un = interior_update4<NE>(Tc, W, pum, pm);
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int x = gx + j;
            if (x >= dimx) continue;
            const long long ij = i + j;
            const float um0 = f4get(pum, j);
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
                val = interior_update<NE>(Tc + j, W, um0, f4get(pm, j));
            }
            f4set(un, j, val);
        }
    }
    return un;
}

// Residual injection on the receiver row (adjoint), cells gx..gx+3 of row gy.
__device__ __forceinline__ void inject_resid4(const ShotDesc& d, int it, long long i, int gx, int gy, float4& un)
{
    if (gy != d.rrow) return;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int x = gx + j;
        if (x >= d.rx0 && x < d.rx0 + d.nrec)
            f4set(un, j, f4get(un, j) + inject_term(gld1(d.resid + (long long)it * d.nrec + (x - d.rx0)),
                                                    gld1(d.vel + i + j), d.dt2));
    }
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
#pragma unroll
    for (int r = 0; r < RYV; ++r) {
        const int gy = gy0 + r;
        if (gy >= dimy) break;
        const long long i = gy * ld + gx;
        const float* Tc = Tf + (ty * RYV + r + HALO) * W + 4 * tx + 4;
        float4 un = cell4_update<NE, W>(d, um, Tc, i, gx, gy, pum[r], pm[r]);
        if constexpr ((FLAGS & F_SRC) != 0) {
            if (d.isrc >= i && d.isrc < i + 4) {
                const float add = inject_term(gld1(d.w + a.it), gld1(d.vel + d.isrc), d.dt2);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (d.isrc == i + j) f4set(un, j, f4get(un, j) + add);
            }
        }
        const bool onrow = gy == d.rrow;
        if constexpr ((FLAGS & F_RESID) != 0) inject_resid4(d, a.it, i, gx, gy, un);
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

// Adjoint step of a group of shots that share one model and one gradient
// (imaging condition accumulated over the group).  Not launched at present:
// plans refuse grad_group > 1 because the first MI355X run of this kernel
// ended in a GPU memory fault whose cause has not been found.  A workgroup owns one
// 4*CX4 x (256/CX4) tile and steps every shot of the group through it in
// turn: the model and the group's gradient are read and written once per
// tile instead of once per shot, which removes 12 of the 28 bytes per
// cell-update of the per-shot adjoint from HBM.  The group's contribution of
// this time step, the sum over its shots in shot order of -virt * lambda_new,
// is added to the gradient once; the wavefield update of every shot is the
// per-shot step, bit for bit.
//
// tiles[] hold (leader shot, tile) pairs; the group is shots leader ..
// leader + gcount - 1, all on the leader's grid and model, and the gradient
// is the leader's.  While one shot is computed from the staged tile (LDS
// buffer s & 1), the loads of the next shot are in flight; one barrier per
// shot orders the staging against the stencil reads.
template <int NE, int CX4>
__global__ __launch_bounds__(VTHR) void k_vadj_group(const ShotDesc* __restrict__ descs,
                                                     const int2* __restrict__ tiles, StepArgs a)
{
    constexpr int HALO = NE < 2 ? 2 : NE;
    constexpr int TR = VTHR / CX4;          // tile rows (one per thread)
    constexpr int W4 = CX4 + 2;             // staged row: 4 halo columns each side
    constexpr int W = 4 * W4;
    constexpr int H = TR + 2 * HALO;
    constexpr int NST = (H * W4 + VTHR - 1) / VTHR;   // staged float4 per thread
    __shared__ float4 T4[2][H * W4];

    const int L = logical_tile(blockIdx.x, a.per_xcd);
    if (L >= a.ntiles || !resolve_step(a)) return;
    const int2 e = tiles[L];
    const ShotDesc& d0 = descs[e.x];
    const int G = d0.gcount;
    const int x0 = (e.y & 0xffff) * 4 * CX4, y0 = (e.y >> 16) * TR;
    const int dimx = d0.dimx, dimy = d0.dimy;
    const long long ld = d0.ld;
    const int k0 = a.ring % 3, k1 = (a.ring + 1) % 3, k2 = (a.ring + 2) % 3;

    const int tx = threadIdx.x % CX4, ty = threadIdx.x / CX4;
    const int gx = x0 + 4 * tx, gy = y0 + ty;
    const bool ok = gx < dimx && gy < dimy;
    const long long i = ok ? gy * ld + gx : 0;   // out-of-grid lanes read cell 0 and discard it
    const long long vo = (long long)a.vslot * d0.fstride;

    // staging slots of this thread: tile element k = threadIdx.x + q * VTHR
    long long si[NST];
    bool sok[NST];
#pragma unroll
    for (int q = 0; q < NST; ++q) {
        const int k = threadIdx.x + q * VTHR;
        const int row = k / W4, c4 = k - (k / W4) * W4;
        const int gyy = y0 - HALO + row, gxx = x0 - 4 + 4 * c4;
        sok[q] = k < H * W4 && gyy >= 0 && gyy < dimy && gxx >= 0 && gxx < dimx;
        si[q] = sok[q] ? gyy * ld + gxx : 0;
    }

    const float4 pm = gld4(d0.vm + i);
    const float4 pg = gld4(d0.grad + i);
    // loads of shot 0
    float4 pum = gld4(pick3(d0.lam, k0) + i);
    float4 pvs = gld4_nt(d0.virt + vo + i);
    float4 st[NST];
#pragma unroll
    for (int q = 0; q < NST; ++q) st[q] = gld4(pick3(d0.lam, k1) + si[q]);

    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int s = 0; s < G; ++s) {
        const ShotDesc& d = descs[e.x + s];
        float4* Tb = T4[s & 1];
#pragma unroll
        for (int q = 0; q < NST; ++q) {
            const int k = threadIdx.x + q * VTHR;
            if (k < H * W4) Tb[k] = sel4(sok[q], st[q]);
        }
        const float4 cum = pum, cvs = pvs;
        // issue the next shot's loads (the last shot reloads its own, unused)
        const ShotDesc& dn = descs[e.x + (s + 1 < G ? s + 1 : s)];
        pum = gld4(pick3(dn.lam, k0) + i);
        pvs = gld4_nt(dn.virt + vo + i);
#pragma unroll
        for (int q = 0; q < NST; ++q) st[q] = gld4(pick3(dn.lam, k1) + si[q]);
        __syncthreads();
        if (ok) {
            const float* Tc = (const float*)Tb + (ty + HALO) * W + 4 * tx + 4;
            float4 un = cell4_update<NE, W>(d, pick3(d.lam, k0), Tc, i, gx, gy, cum, pm);
            inject_resid4(d, a.it, i, gx, gy, un);
            if (gx + 3 >= dimx) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (gx + j >= dimx) f4set(un, j, 0.f);
            }
            gst4(pick3(d.lam, k2) + i, un);
            acc.x = acc.x + (-cvs.x) * un.x;
            acc.y = acc.y + (-cvs.y) * un.y;
            acc.z = acc.z + (-cvs.z) * un.z;
            acc.w = acc.w + (-cvs.w) * un.w;
        }
    }
    if (ok) {
        float4 g;
        g.x = pg.x + acc.x;
        g.y = pg.y + acc.y;
        g.z = pg.z + acc.z;
        g.w = pg.w + acc.w;
        gst4(d0.grad + i, g);
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
