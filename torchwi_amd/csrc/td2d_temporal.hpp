// td2d_temporal.hpp -- temporally blocked batched step kernel (included by
// td2d_hip.hip).
//
// One launch advances up to T consecutive time steps ("levels") of every
// shot.  A workgroup owns a 64 x 64 core tile and loads the tile plus a halo
// of R = T*max(ne,2) cells (rounded up to 4) of the two input fields (u_prev,
// u_cur) into LDS.  Level k recomputes the region shrunk by max(ne,2) cells
// per level on every side that is interior to the grid -- the cells the next
// level's stencil reads -- so after T levels the core is exact; sides on a
// grid edge do not shrink, since no cell depends on anything beyond the
// edge.  Redundant halo cells are computed with the same arithmetic as
// everywhere else, so results are bit-identical to one launch per step.
//
// Each level's new values are formed in registers, then (after a barrier)
// written over the level two steps back in LDS: the two buffers alternate
// roles.  Per-step outputs of the core -- gathers, virtual sources, the
// imaging-condition gradient (accumulated in registers across levels in
// the reference's order), checkpoints -- are written as each level is
// computed; the last two levels go to the output pair of the field ring.
// Inputs and outputs are different global buffers (a ring of two pairs), so
// no workgroup overwrites cells another one still reads as halo.
//
// DRAM traffic per cell-step at T = 3 (order 8): about 2*(88/64)^2/3 fields
// read and 2/3 written, against 3 and 1 for one step per launch.
#pragma once

namespace tw {

constexpr int TCX = 64, TCY = 64;   // core tile
constexpr int TTHR = 512;

struct TArgs {
    int t_first;     // time index of level 1
    int dir;         // +1 forward in time, -1 adjoint
    int nsteps;      // levels to compute (1..T)
    int in_pair, out_pair;  // 0: fields (0,1) of the ring, 1: fields (2,3)
    int vslot_first; // virt slot of level 1 (advances with dir)
    int adj;         // 1: adjoint ring
    int K, nseg;     // checkpointing (forward saves)
    int ntiles, per_xcd;
};

template <int NE, int T>
struct TGeom {
    static constexpr int r = NE < 2 ? 2 : NE;              // per-step dependency radius
    static constexpr int R = ((T * r + 3) / 4) * 4;        // halo, multiple of 4
    static constexpr int RW = TCX + 2 * R;                 // region width (cells)
    static constexpr int RW4 = RW / 4;
    static constexpr int RH = TCY + 2 * R;                 // region height
    static constexpr int R4 = R / 4;
    static constexpr int CX4 = TCX / 4;
    static constexpr int NCORE = TCY * CX4;                // core float4 items (1024)
    static constexpr int NHALO = RH * RW4 - NCORE;
    static constexpr int CPT = NCORE / TTHR;               // core items per thread (4)
    static constexpr int HPT = (NHALO + TTHR - 1) / TTHR;  // halo items per thread
};

// Region coordinates (row, float4 column) of halo item h.
template <int NE, int T>
__device__ __forceinline__ void halo_item(int h, int& row, int& c4)
{
    using G = TGeom<NE, T>;
    const int band = G::R * G::RW4;
    if (h < band) {
        row = h / G::RW4;
        c4 = h - row * G::RW4;
    } else if (h < 2 * band) {
        h -= band;
        row = G::R + TCY + h / G::RW4;
        c4 = h % G::RW4;
    } else {
        h -= 2 * band;
        const int per = 2 * G::R4;
        row = G::R + h / per;
        const int c = h % per;
        c4 = c < G::R4 ? c : G::CX4 + c;
    }
}

template <int NE, unsigned FLAGS, int T>
__global__ __launch_bounds__(TTHR) void k_tstep(const ShotDesc* __restrict__ descs,
                                                const int2* __restrict__ tiles, TArgs a)
{
    using G = TGeom<NE, T>;
    constexpr int W = G::RW;           // LDS row stride (floats)
    constexpr int NIT = G::CPT + G::HPT;
    __shared__ float4 LA[G::RH * G::RW4];
    __shared__ float4 LB[G::RH * G::RW4];

    const int L = logical_tile(blockIdx.x, a.per_xcd);
    if (L >= a.ntiles) return;
    const int2 e = tiles[L];
    const ShotDesc& d = descs[e.x];
    const int cx0 = (e.y & 0xffff) * TCX, cy0 = (e.y >> 16) * TCY;
    const int dimx = d.dimx, dimy = d.dimy;
    const long long ld = d.ld;
    const int ox = cx0 - G::R, oy = cy0 - G::R;   // region origin (may be negative)

    float* const* ring = a.adj ? d.lam : d.u;
    const float* __restrict__ in_prev = ring[2 * a.in_pair];
    const float* __restrict__ in_cur = ring[2 * a.in_pair + 1];
    float* __restrict__ out_prev = ring[2 * a.out_pair];
    float* __restrict__ out_cur = ring[2 * a.out_pair + 1];

    // valid-range bounds: sides whose halo reaches past the grid edge never shrink
    const bool ex0 = ox <= 0, ex1 = ox + W >= dimx, ey0 = oy <= 0, ey1 = oy + G::RH >= dimy;

    // this thread's items: CPT core items first, then halo items
    int irow[NIT], ic4[NIT];
#pragma unroll
    for (int q = 0; q < G::CPT; ++q) {
        const int c = threadIdx.x + q * TTHR;
        irow[q] = G::R + c / G::CX4;
        ic4[q] = G::R4 + c % G::CX4;
    }
#pragma unroll
    for (int q = 0; q < G::HPT; ++q) {
        const int h = threadIdx.x + q * TTHR;
        if (h < G::NHALO) halo_item<NE, T>(h, irow[G::CPT + q], ic4[G::CPT + q]);
        else { irow[G::CPT + q] = -1; ic4[G::CPT + q] = 0; }
    }

    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    // 1. loads: model m for every item, c3 / grad / virt slots for core items
    float4 pm[NIT];
#pragma unroll
    for (int q = 0; q < NIT; ++q) {
        const int gy = oy + irow[q], gx = ox + 4 * ic4[q];
        const bool ok = irow[q] >= 0 && gy >= 0 && gy < dimy && gx >= 0 && gx < dimx;
        pm[q] = sel4(ok, gld4(d.vm + (ok ? gy * ld + gx : 0)));
    }
    float4 pc3[G::CPT], pg[G::CPT], pvs[G::CPT][T];
#pragma unroll
    for (int q = 0; q < G::CPT; ++q) {
        const int gy = oy + irow[q], gx = ox + 4 * ic4[q];
        const bool ok = gy < dimy && gx < dimx;   // core cells are never below 0
        const long long i = ok ? gy * ld + gx : 0;
        if constexpr ((FLAGS & F_VIRT) != 0) pc3[q] = gld4(d.vc3 + i);
        if constexpr ((FLAGS & F_GRAD) != 0) {
            pg[q] = gld4(d.grad + i);
#pragma unroll
            for (int k = 0; k < T; ++k) {
                const int slot = a.vslot_first + a.dir * k;
                pvs[q][k] = (k < a.nsteps) ? gld4(d.virt + (long long)slot * d.fstride + i) : z4;
            }
        }
    }
    // 2. stage both input fields of the region
    for (int k = threadIdx.x; k < G::RH * G::RW4; k += TTHR) {
        const int row = k / G::RW4;
        const int c4 = k - row * G::RW4;
        const int gy = oy + row, gx = ox + 4 * c4;
        const bool ok = gy >= 0 && gy < dimy && gx >= 0 && gx < dimx;
        const long long i = ok ? gy * ld + gx : 0;
        LA[k] = sel4(ok, gld4(in_prev + i));
        LB[k] = sel4(ok, gld4(in_cur + i));
    }
    __syncthreads();

    float* Pprev = (float*)LA;  // F_{k-2}
    float* Pcur = (float*)LB;   // F_{k-1}
    float4 nv[NIT];
    for (int k = 1; k <= a.nsteps; ++k) {
        const int t = a.t_first + a.dir * (k - 1);
        const int vx0 = ex0 ? 0 : ox + k * G::r, vx1 = ex1 ? dimx : ox + W - k * G::r;
        const int vy0 = ey0 ? 0 : oy + k * G::r, vy1 = ey1 ? dimy : oy + G::RH - k * G::r;
#pragma unroll
        for (int q = 0; q < NIT; ++q) {
            nv[q] = z4;
            const int row = irow[q];
            const int gy = oy + row, gx = ox + 4 * ic4[q];
            if (row < 0 || gy < vy0 || gy >= vy1 || gx + 3 < vx0 || gx >= vx1) continue;
            const int li = row * W + 4 * ic4[q];
            const float* Tc = Pcur + li;
            const float* Tp = Pprev + li;
            const long long i = gy * ld + gx;
            float4 un = z4;
            if (gx >= vx0 && gx + 3 < vx1 && gx >= NE && gx + 3 < dimx - NE && gy >= NE && gy < dimy - NE) {
                un = interior_update4<NE>(Tc, W, *(const float4*)Tp, pm[q]);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int x = gx + j;
                    if (x < vx0 || x >= vx1 || x >= dimx) continue;
                    float val;
                    if (gy < NE) {
                        val = 0.f;
                    } else if (gy >= dimy - NE) {
                        val = keys_update(Tp[j], Tc[j], Tp[j - W], Tc[j - W], Tc[j - 2 * W], d.vel[i + j], d.h2, d.dt2, d.hdt);
                    } else if (x < NE) {
                        val = keys_update(Tp[j], Tc[j], Tp[j + 1], Tc[j + 1], Tc[j + 2], d.vel[i + j], d.h2, d.dt2, d.hdt);
                    } else if (x >= dimx - NE) {
                        val = keys_update(Tp[j], Tc[j], Tp[j - 1], Tc[j - 1], Tc[j - 2], d.vel[i + j], d.h2, d.dt2, d.hdt);
                    } else {
                        val = interior_update<NE>(Tc + j, W, Tp[j], f4get(pm[q], j));
                    }
                    f4set(un, j, val);
                }
            }
            if constexpr ((FLAGS & F_SRC) != 0) {
                if (d.isrc >= i && d.isrc < i + 4) {
                    const float add = inject_term(d.w[t], d.vel[d.isrc], d.dt2);
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (d.isrc == i + j) f4set(un, j, f4get(un, j) + add);
                }
            }
            if constexpr ((FLAGS & F_RESID) != 0) {
                if (gy == d.rrow) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int x = gx + j;
                        if (x >= d.rx0 && x < d.rx0 + d.nrec && x >= vx0 && x < vx1)
                            f4set(un, j, f4get(un, j) + inject_term(d.resid[(long long)t * d.nrec + (x - d.rx0)],
                                                                    d.vel[i + j], d.dt2));
                    }
                }
            }
            nv[q] = un;
            // per-step outputs of the core
            if (q < G::CPT && gy < dimy && gx < dimx) {
                if constexpr ((FLAGS & F_REC) != 0) {
                    if (gy == d.rrow) {
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int x = gx + j;
                            if (x >= d.rx0 && x < d.rx0 + d.nrec)
                                d.frd[(long long)t * d.nrec + (x - d.rx0)] = f4get(un, j);
                        }
                    }
                }
                if constexpr ((FLAGS & F_VIRT) != 0) {
                    float4 vs;
                    vs.x = pc3[q].x * ((Tp[0] - 2.f * Tc[0]) + un.x);
                    vs.y = pc3[q].y * ((Tp[1] - 2.f * Tc[1]) + un.y);
                    vs.z = pc3[q].z * ((Tp[2] - 2.f * Tc[2]) + un.z);
                    vs.w = pc3[q].w * ((Tp[3] - 2.f * Tc[3]) + un.w);
                    const int slot = a.vslot_first + a.dir * (k - 1);
                    gst4(d.virt + (long long)slot * d.fstride + i, vs);
                }
                if constexpr ((FLAGS & F_GRAD) != 0) {
#pragma unroll
                    for (int kk = 0; kk < T; ++kk) {
                        if (kk == k - 1) {
                            pg[q].x = pg[q].x + (-pvs[q][kk].x) * un.x;
                            pg[q].y = pg[q].y + (-pvs[q][kk].y) * un.y;
                            pg[q].z = pg[q].z + (-pvs[q][kk].z) * un.z;
                            pg[q].w = pg[q].w + (-pvs[q][kk].w) * un.w;
                        }
                    }
                }
                if (a.K > 0 && a.dir > 0 && (FLAGS & F_REC) != 0) {
                    if ((t + 2) % a.K == 0 && (t + 2) / a.K < a.nseg)
                        gst4(d.ckpt + (2LL * ((t + 2) / a.K)) * d.fstride + i, un);
                    if ((t + 1) % a.K == 0 && (t + 1) / a.K < a.nseg)
                        gst4(d.ckpt + (2LL * ((t + 1) / a.K) + 1) * d.fstride + i, un);
                }
            }
        }
        __syncthreads();
        // F_k replaces F_{k-2}: only cells valid at this level are written
#pragma unroll
        for (int q = 0; q < NIT; ++q) {
            const int row = irow[q];
            const int gy = oy + row, gx = ox + 4 * ic4[q];
            if (row < 0 || gy < vy0 || gy >= vy1 || gx + 3 < vx0 || gx >= vx1) continue;
            float* dst = Pprev + row * W + 4 * ic4[q];
            if (gx >= vx0 && gx + 3 < vx1 && gx + 3 < dimx) {
                *(float4*)dst = nv[q];
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int x = gx + j;
                    if (x >= vx0 && x < vx1 && x < dimx) dst[j] = f4get(nv[q], j);
                }
            }
        }
        __syncthreads();
        float* tmp = Pprev;
        Pprev = Pcur;
        Pcur = tmp;
    }

    // 3. core outputs: the last two levels, and the accumulated gradient
#pragma unroll
    for (int q = 0; q < G::CPT; ++q) {
        const int gy = oy + irow[q], gx = ox + 4 * ic4[q];
        if (gy >= dimy || gx >= dimx) continue;
        const long long i = gy * ld + gx;
        const int li = irow[q] * W + 4 * ic4[q];
        float4 p = *(const float4*)(Pprev + li);
        float4 c = *(const float4*)(Pcur + li);
        if (gx + 3 >= dimx) {  // row padding stays zero
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (gx + j >= dimx) { f4set(p, j, 0.f); f4set(c, j, 0.f); }
        }
        gst4(out_prev + i, p);
        gst4(out_cur + i, c);
        if constexpr ((FLAGS & F_GRAD) != 0) gst4(d.grad + i, pg[q]);
    }
}

}  // namespace tw
