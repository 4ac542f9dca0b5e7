// td2d_march.hpp -- temporally blocked batched step kernel (included by
// td2d_hip.hip).
//
// One launch advances T consecutive time steps ("levels" 1..T) of every
// shot.  A workgroup owns a column strip of one shot's padded grid -- a core
// of at most 4*MC4 columns plus a halo of 4T columns on each side -- and a
// chunk of its rows, and marches down the chunk in blocks of RB rows.  At
// iteration i the two input fields (levels -1 and 0) of row block i enter
// LDS, and level k computes the block that lies k*D rows higher
// (D = RB + 4): every row level k reads of level k-1 was computed at an
// earlier iteration, so all T levels of an iteration are independent and one
// barrier per iteration suffices.  The inputs of block i+1 and the model,
// virtual-source and gradient values of the next iteration are loaded while
// the current one computes.
//
// Each field level lives in one of three LDS ring buffers of rows (level k
// in buffer k mod 3): level k overwrites level k-3, which no update reads any
// more, while level k-2 -- read by the Keys update of a boundary cell at its
// neighbours -- stays intact.  The rings carry 4 ghost rows at each end, so
// stencil reads at a fixed offset never wrap.
//
// Levels shrink by 4 cells per level on every side interior to the grid (4 is
// at least the dependency radius max(ne, 2)); sides on a grid edge do not
// shrink.  Redundant halo cells use the same arithmetic as everywhere else,
// so results are bit-identical to one launch per step.
//
// DRAM traffic per cell-step at T = 4: the two input fields and the model
// once per launch (with the column overlap of neighbouring strips), the two
// output fields once per launch, and only the per-step virtual source or
// gather every step -- against three fields and the model per step for one
// step per launch.
#pragma once

namespace tw {

constexpr int MTHR = 320;       // threads per workgroup (5 waves)
constexpr int MC4 = 17;         // widest strip core in float4 columns (68 cells)
constexpr int MWG_TARGET = 512; // workgroups a launch aims for (two per CU)

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

template <int T>
struct MGeom {
    static constexpr int RB = T == 4 ? 4 : (T == 3 ? 5 : 8);  // rows per level per iteration
    static constexpr int D = RB + 4;              // row lag between consecutive levels
    static constexpr int SPAN = T * D + RB + 4;   // rows live in a ring at one iteration
    static constexpr int G = 4;                   // ghost rows at each end of a ring
    static constexpr int PR = SPAN + 2 * G;       // physical rows of a ring
    static constexpr int W4 = MC4 + 2 * T;        // region width in float4
    static constexpr int W = 4 * W4;              // LDS row stride (floats)
    static constexpr int BUF = PR * W;            // floats per ring
    static constexpr int GS = (T - 1) * D + RB;   // rows of the gradient ring
    static_assert(RB * (T * MC4 + T * (T - 1)) <= MTHR, "one update item per thread");
};

// Non-negative remainder.
__device__ __forceinline__ int pmod(int a, int m)
{
    const int r = a % m;
    return r < 0 ? r + m : r;
}

// Store a float4 into ring slot s (and its ghost copy near either end).
// off: ring base plus the float4's column offset, in floats.
template <int T>
__device__ __forceinline__ void ring_put(float* L, int off, int s, float4 v)
{
    using G = MGeom<T>;
    float* r = L + off + (G::G + s) * G::W;
    *(float4*)r = v;
    if (s < G::G) *(float4*)(r + G::SPAN * G::W) = v;
    else if (s >= G::SPAN - G::G) *(float4*)(r - G::SPAN * G::W) = v;
}

template <int NE, unsigned FLAGS, int T>
__global__ __launch_bounds__(MTHR) void k_mstep(const ShotDesc* __restrict__ descs,
                                                const int2* __restrict__ tiles, TArgs a)
{
    using G = MGeom<T>;
    constexpr int W = G::W;
    constexpr int RB = G::RB, D = G::D, SPAN = G::SPAN, GS = G::GS;
    __shared__ float4 LF4[3 * G::BUF / 4];
    __shared__ float4 GR4[(FLAGS & F_GRAD) != 0 ? GS * MC4 : 1];
    float* const LF = (float*)LF4;
    float* const GR = (float*)GR4;

    const int Lt = logical_tile(blockIdx.x, a.per_xcd);
    if (Lt >= a.ntiles) return;
    const int2 e = tiles[Lt];
    const ShotDesc& d = descs[e.x];
    const int dimx = d.dimx, dimy = d.dimy;
    const long long ld = d.ld;
    const int dimx4 = (dimx + 3) >> 2;
    const int strip = e.y & 0xffff, chunk = e.y >> 16;
    const int c4a = strip * dimx4 / d.mstrips, c4b = (strip + 1) * dimx4 / d.mstrips;
    const int Y0 = chunk * dimy / d.mchunks, Y1 = (chunk + 1) * dimy / d.mchunks;
    const int ox4 = c4a - T;                    // region origin, float4 columns
    const bool ex0 = ox4 <= 0, ex1 = c4b + T >= dimx4;
    const int oy = Y0 - 4 * T;
    const bool ey0 = oy <= 0, ey1 = Y1 + 4 * T >= dimy;
    const int O = ey0 ? 0 : oy;                 // region rows [O, yend)
    const int yend = ey1 ? dimy : Y1 + 4 * T;
    const int rl4 = ex0 ? 0 : ox4;              // region float4 columns [rl4, rh4)
    const int rh4 = ex1 ? dimx4 : c4b + T;

    float* const* ring = a.adj ? d.lam : d.u;
    const float* __restrict__ in_prev = ring[2 * a.in_pair];
    const float* __restrict__ in_cur = ring[2 * a.in_pair + 1];
    float* __restrict__ out_prev = ring[2 * a.out_pair];
    float* __restrict__ out_cur = ring[2 * a.out_pair + 1];

    // --- this thread's update item: level k, row j of the block, column g4
    const int tid = threadIdx.x;
    int k = 0, j = 0, g4 = 0;
    {
        int q = tid;
#pragma unroll
        for (int kk = 1; kk <= T; ++kk) {
            const int lo = ex0 ? 0 : ox4 + kk, hi = ex1 ? dimx4 : c4b + T - kk;
            const int n = hi - lo;
            if (k == 0) {
                if (q < RB * n) {
                    k = kk;
                    j = q / n;
                    g4 = lo + (q - j * n);
                } else {
                    q -= RB * n;
                }
            }
        }
    }
    if (k > a.nsteps) k = 0;   // levels past this launch's last step idle
    const int gx = 4 * g4;
    const int lc = 4 * (g4 - ox4);
    const bool core = k > 0 && g4 >= c4a && g4 < c4b;
    const int offC = ((k + 2) % 3) * G::BUF + lc;   // level k-1
    const int offP = ((k + 1) % 3) * G::BUF + lc;   // level k-2
    const int offN = (k % 3) * G::BUF + lc;         // level k
    const int top = ey0 ? 0 : O + 4 * k;
    const int bot = ey1 ? dimy : Y1 + 4 * T - 4 * k;
    const int t = a.t_first + a.dir * (k - 1);
    const int vsl = k > 0 ? a.vslot_first + a.dir * (k - 1) : 0;
    const long long voff = (long long)vsl * d.fstride;
    int p = O - k * D + j;          // row at iteration 0
    int s = pmod(p - O, SPAN);      // its ring slot
    int gs = pmod(p - O, GS);       // its gradient-ring slot
    float* const gcol = GR + 4 * (g4 - c4a);

    // --- this thread's load item: row jl of the block, column rl4 + cl
    const int rw4 = rh4 - rl4;
    const bool lact = tid < RB * rw4;
    const int jl = lact ? tid / rw4 : 0;
    const int cl = lact ? tid - jl * rw4 : 0;
    const int gxl = 4 * (rl4 + cl);
    const int lcl = 4 * (rl4 + cl - ox4);
    const bool lcore = lact && rl4 + cl >= c4a && rl4 + cl < c4b;
    int pl = O + jl;
    int sl = jl;

    // Operands of iteration it+1 are loaded at the start of iteration it, into
    // the other of two register sets (the loop is unrolled by two), before
    // the operands of iteration it are consumed.  Every load is issued by
    // every thread at every iteration, from cell 0 of its array when the
    // operand will not be used, so the in-order counter of outstanding
    // vector-memory operations has the same count on every path and a wait
    // for this iteration's operands never waits for the next iteration's.
    // Values are consumed only under the condition they were loaded for.
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    struct Pref {
        float4 lp, lc;     // input rows: level -1, level 0
        float4 m, vl;      // model m and velocity (boundary / injection cells)
        float4 c3, v, g;   // cons/v^3, virtual source, gradient (first level)
    };
    const bool keysx = gx < NE || gx + 3 >= dimx - NE;   // Keys columns in the item
    auto issue_inputs = [&](Pref& P, int row) {
        const bool ok = lact && row < yend;
        const long long idx = ok ? row * ld + gxl : 0;
        P.lp = gld4(in_prev + idx);
        P.lc = gld4(in_cur + idx);
    };
    auto issue_item = [&](Pref& P, int row) {
        const bool ok = k > 0 && row >= top && row < bot;
        const long long idx = ok ? row * ld + gx : 0;
        P.m = gld4(d.vm + idx);
        const bool kv = ok && (keysx || row >= dimy - NE || row == d.rrow);
        P.vl = gld4(d.vel + (kv ? idx : 0));
        const bool okc = ok && core && row >= Y0 && row < Y1;
        const long long ic = okc ? idx : 0;
        if constexpr ((FLAGS & F_VIRT) != 0) P.c3 = gld4(d.vc3 + ic);
        if constexpr ((FLAGS & F_GRAD) != 0) {
            P.v = gld4_nt(d.virt + voff + ic);
            P.g = gld4(d.grad + (k == 1 ? ic : 0));
        }
    };

    // Per-launch constants, loaded before the loop: the source term of this
    // thread's level, and the residual of every level along the receiver row.
    float srcadd = 0.f;
    if constexpr ((FLAGS & F_SRC) != 0)
        if (k > 0) srcadd = inject_term(gld1(d.w + t), gld1(d.vel + d.isrc), d.dt2);
    __shared__ float RS[(FLAGS & F_RESID) != 0 ? T * W : 1];
    if constexpr ((FLAGS & F_RESID) != 0) {
        const bool rin = d.rrow >= O && d.rrow < yend;
        for (int q = tid; q < T * W; q += MTHR) {
            const int kk = q / W + 1;
            const int x = 4 * ox4 + (q - (kk - 1) * W);
            float r = 0.f;
            if (rin && kk <= a.nsteps && x >= d.rx0 && x < d.rx0 + d.nrec)
                r = gld1(d.resid + (long long)(a.t_first + a.dir * (kk - 1)) * d.nrec + (x - d.rx0));
            RS[q] = r;
        }
        __syncthreads();
    }

    auto step = [&](Pref& P, Pref& N) {
        // 0. operands of the next iteration
        issue_item(N, p + RB);
        issue_inputs(N, pl + RB);

        // 1. input rows of this block: level -1 -> ring 2, level 0 -> ring 0
        if (lact && pl < yend) {
            ring_put<T>(LF, 2 * G::BUF + lcl, sl, P.lp);
            ring_put<T>(LF, lcl, sl, P.lc);
            if (a.nsteps == 1 && lcore && pl >= Y0 && pl < Y1) gst4(out_prev + pl * ld + gxl, P.lc);
        }

        // 2. level k at row p
        if (k > 0 && p >= top && p < bot) {
            const float* Tc = LF + offC + (G::G + s) * W;
            const float* Tp = LF + offP + (G::G + s) * W;
            const long long i = p * ld + gx;
            float4 un = z4;
            if (gx >= NE && gx + 3 < dimx - NE && p >= NE && p < dimy - NE) {
                un = interior_update4<NE>(Tc, W, *(const float4*)Tp, P.m);
            } else {
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    const int x = gx + jj;
                    if (x >= dimx) continue;   // row padding stays zero
                    const float v = f4get(P.vl, jj);
                    float val;
                    if (p < NE) {
                        val = 0.f;
                    } else if (p >= dimy - NE) {
                        val = keys_update(Tp[jj], Tc[jj], Tp[jj - W], Tc[jj - W], Tc[jj - 2 * W], v, d.h2, d.dt2,
                                          d.hdt);
                    } else if (x < NE) {
                        val = keys_update(Tp[jj], Tc[jj], Tp[jj + 1], Tc[jj + 1], Tc[jj + 2], v, d.h2, d.dt2, d.hdt);
                    } else if (x >= dimx - NE) {
                        val = keys_update(Tp[jj], Tc[jj], Tp[jj - 1], Tc[jj - 1], Tc[jj - 2], v, d.h2, d.dt2, d.hdt);
                    } else {
                        val = interior_update<NE>(Tc + jj, W, Tp[jj], f4get(P.m, jj));
                    }
                    f4set(un, jj, val);
                }
            }
            if constexpr ((FLAGS & F_SRC) != 0) {
                if (d.isrc >= i && d.isrc < i + 4) {
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj)
                        if (d.isrc == i + jj) f4set(un, jj, f4get(un, jj) + srcadd);
                }
            }
            if constexpr ((FLAGS & F_RESID) != 0) {
                if (p == d.rrow) {
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) {
                        const int x = gx + jj;
                        if (x >= d.rx0 && x < d.rx0 + d.nrec)
                            f4set(un, jj, f4get(un, jj) + inject_term(RS[(k - 1) * W + lc + jj], f4get(P.vl, jj), d.dt2));
                    }
                }
            }
            // per-step outputs of the core
            if (core && p >= Y0 && p < Y1) {
                if constexpr ((FLAGS & F_REC) != 0) {
                    if (p == d.rrow) {
#pragma unroll
                        for (int jj = 0; jj < 4; ++jj) {
                            const int x = gx + jj;
                            if (x >= d.rx0 && x < d.rx0 + d.nrec)
                                gst1(d.frd + (long long)t * d.nrec + (x - d.rx0), f4get(un, jj));
                        }
                    }
                }
                if constexpr ((FLAGS & F_VIRT) != 0) {
                    float4 vs;
                    vs.x = P.c3.x * ((Tp[0] - 2.f * Tc[0]) + un.x);
                    vs.y = P.c3.y * ((Tp[1] - 2.f * Tc[1]) + un.y);
                    vs.z = P.c3.z * ((Tp[2] - 2.f * Tc[2]) + un.z);
                    vs.w = P.c3.w * ((Tp[3] - 2.f * Tc[3]) + un.w);
                    gst4_nt(d.virt + voff + i, vs);
                }
                if constexpr ((FLAGS & F_GRAD) != 0) {
                    float* gr = gcol + gs * (4 * MC4);
                    const float4 gl = *(const float4*)gr;   // value selects, not an address select
                    const bool first = k == 1;
                    float4 g = make_float4(first ? P.g.x : gl.x, first ? P.g.y : gl.y,
                                           first ? P.g.z : gl.z, first ? P.g.w : gl.w);
                    g.x = g.x + (-P.v.x) * un.x;
                    g.y = g.y + (-P.v.y) * un.y;
                    g.z = g.z + (-P.v.z) * un.z;
                    g.w = g.w + (-P.v.w) * un.w;
                    if (k == a.nsteps) gst4(d.grad + i, g);
                    else *(float4*)gr = g;
                }
                if (a.K > 0 && a.dir > 0 && (FLAGS & F_REC) != 0) {
                    if ((t + 2) % a.K == 0 && (t + 2) / a.K < a.nseg)
                        gst4_nt(d.ckpt + (2LL * ((t + 2) / a.K)) * d.fstride + i, un);
                    if ((t + 1) % a.K == 0 && (t + 1) / a.K < a.nseg)
                        gst4_nt(d.ckpt + (2LL * ((t + 1) / a.K) + 1) * d.fstride + i, un);
                }
                if (k == a.nsteps - 1) gst4(out_prev + i, un);
                if (k == a.nsteps) gst4(out_cur + i, un);
            }
            ring_put<T>(LF, offN, s, un);
        }

        // 3. advance one block
        p += RB;
        s += RB;
        if (s >= SPAN) s -= SPAN;
        gs += RB;
        if (gs >= GS) gs -= GS;
        pl += RB;
        sl += RB;
        if (sl >= SPAN) sl -= SPAN;
        __syncthreads();
    };

    const int nbot = ey1 ? dimy : Y1 + 4 * T - 4 * a.nsteps;
    // iterations, rounded up to even: an extra one finds every row masked
    const int nit = (nbot - O + a.nsteps * D + 2 * RB - 1) / (2 * RB) * 2;
    Pref PA, PB;
    issue_item(PA, p);
    issue_inputs(PA, pl);
    for (int it = 0; it < nit; it += 2) {
        step(PA, PB);
        step(PB, PA);
    }
}

}  // namespace tw
