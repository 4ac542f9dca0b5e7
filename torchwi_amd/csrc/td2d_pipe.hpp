// td2d_pipe.hpp -- level-pipelined, temporally blocked batched step kernel
// (included by td2d_hip.hip).
//
// One launch advances T consecutive time steps ("levels" 1..T) of every
// shot.  A workgroup owns a column strip of one shot's padded grid, one
// wavefront wide (64 float4 columns: a core plus T float4 columns of halo on
// each side that is not a grid edge), and a chunk of its rows plus 4T rows
// of halo above and below (again except at grid edges).  It marches down the
// rows, one row per wave per iteration, with one wave per time level:
//
//   wave 0 (loader) streams the two input fields (levels -1 and 0) from HBM
//          into LDS, PD rows ahead (through register queues);
//   wave k (level k) at iteration i computes level k of row r = O + i - 5k.
//
// Level k lags level k-1 by LAG = 5 rows (the stencil radius, 4, plus one), so
// every row it needs of level k-1 was produced at an earlier iteration and one
// barrier per iteration orders the whole pipeline.  Each level wave keeps the
// nine rows r-4..r+4 of level k-1 in a ten-slot register window (the row loop
// is unrolled ten times, so every window slot is a fixed register); the x
// neighbours of the centre row come from the adjacent lanes through DPP
// wavefront shifts.  LDS carries only what moves between levels: the newest
// row of each level (NCUR = 2 slots), the rows of level k-2 that the next
// level uses as its u_prev (read LAG iterations later: NPRV = 6 slots; the row
// above, which a bottom Keys cell also needs, is the u_prev the wave read one
// iteration earlier and kept), and in the adjoint the gradient sum after each
// level but the last (NGRD = 6 slots; wave 1 loads the stored gradient
// itself), so the gradient is read and written once per launch and
// accumulated in time order.
//
// The halo of a level is valid where its dependency cone stays inside the
// loaded region; it shrinks by one float4 column and four rows per level on
// sides that are not grid edges, and the core stays valid at level T.  Every
// cell uses the same arithmetic as a one-step launch (td2d_arith.hpp), so the
// results are bit-identical.
#pragma once

namespace tw {

struct TArgs {
    int t_first;     // time index of level 1
    int dir;         // +1 forward in time, -1 adjoint
    int nsteps;      // levels to compute (1..T)
    int in_pair, out_pair;  // 0: fields (0,1) of the ring, 1: fields (2,3)
    int ck_in;       // >= 0: the input pair is checkpoint segment ck_in's saved pair, not ring pair in_pair
    int no_out;      // 1: no launch follows that reads the outgoing pair (a segment's last recomputation)
    int vslot_first; // virt slot of level 1 (advances with dir)
    int adj;         // 1: adjoint ring
    int K, nseg;     // checkpointing (forward saves)
    int ntiles, per_xcd;
};

constexpr int PLANES = 64;  // float4 columns per strip (one wavefront)
constexpr int PW = 10;      // register window slots (the row loop's unroll)
constexpr int PQ = 5;       // register prefetch slots (divides PW)
constexpr int PD = 2;      // prefetch distance of the adjoint in iterations (< PQ); the forward steps use PD + 1
// Schedule: a workgroup barrier every BI iterations; level k lags level
// k-1 by LAG = 4 + BI rows, so that each row a level reads was written at
// least one barrier earlier, and the LDS rings are deep enough for the waves
// of a workgroup to be up to BI - 1 iterations apart between barriers.
// BI = 2 (a barrier every two rows, deeper LDS rings and fewer resident
// workgroups) measured 13-16% slower than BI = 1 for the checkpointed gradient
// of 64 shots of 3008 x 1008 (10.9-11.2 vs 9.6 s), and 22% slower on 192
// shots of 1032 x 392.
constexpr int BI = 1;           // iterations per workgroup barrier
constexpr int LAG = 4 + BI;     // rows between consecutive levels: stencil radius + BI
constexpr int NCUR = LAG - 4 + BI;  // LDS slots of the newest rows of a level
constexpr int NPRV = LAG + BI;  // LDS slots of level k-2 rows (read LAG iterations after the write)
constexpr int NGRD = LAG + BI;  // LDS slots of gradient sums (read LAG iterations after the write)
static_assert(NPRV == NGRD, "one slot counter serves both rings");
static_assert(PW % BI == 0, "barriers fall on fixed iterations of the unrolled loop");

template <int T, unsigned FLAGS>
struct PLayout {
    static constexpr int CUR = 0;                          // ring k: NCUR rows of level k
    static constexpr int PRV = CUR + T * NCUR * PLANES;    // ring k: NPRV rows of level k-1
    static constexpr int GRD = PRV + T * NPRV * PLANES;    // ring k-1: NGRD rows of gradient sums after level k
    // ring k-1 holds the model rows of level k, passed to level k+1 (read LAG
    // iterations after the write, like the u_prev ring); the recomputation and
    // the adjoint only
    static constexpr bool MRING = (FLAGS & (F_GRAD | F_VIRT)) != 0;
    static constexpr int MDL = GRD + ((FLAGS & F_GRAD) != 0 ? (T - 1) * NGRD * PLANES : 0);
    static constexpr int KS = MDL + (MRING ? (T - 1) * NPRV * PLANES : 0);  // per level: Keys scratch
    static constexpr int SIZE0 = KS + T * 32;
    // The adjoint's LDS is at least 41.6 KB so that at most 3 workgroups (3
    // waves per SIMD) share a CU: at 4 (its VGPRs would allow them) it
    // measured 4% slower at 64 shots of 3008 x 1008 and 10% slower at 192
    // shots of 1032 x 392.
    static constexpr int SIZE = ((FLAGS & F_GRAD) != 0 && SIZE0 < 2600) ? 2600 : SIZE0;
};

template <int N>
struct IC {
    static constexpr int value = N;
};

// Wavefront shifts: lane l receives the value of lane l-1 (from_left) or of
// lane l+1 (from_right); the lanes at the ends receive 0 (bound_ctrl), so no
// "old" value has to be materialised in the destination first.
__device__ __forceinline__ float dpp_from_left(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float dpp_from_right(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float4 left4(float4 v)
{
    return make_float4(dpp_from_left(v.x), dpp_from_left(v.y), dpp_from_left(v.z), dpp_from_left(v.w));
}
__device__ __forceinline__ float4 right4(float4 v)
{
    return make_float4(dpp_from_right(v.x), dpp_from_right(v.y), dpp_from_right(v.z), dpp_from_right(v.w));
}
template <int NE, unsigned FLAGS, int T>
__global__ __launch_bounds__(64 * (T + 1)) void k_pipe(const ShotDesc* __restrict__ descs,
                                                       const int2* __restrict__ tiles, TArgs a)
{
    using LY = PLayout<T, FLAGS>;
    __shared__ float4 S[LY::SIZE];

    const int Lt = logical_tile(blockIdx.x, a.per_xcd);
    if (Lt >= a.ntiles) return;
    const int2 e = tiles[Lt];
    const ShotDesc& d = descs[e.x];
    const int dimx = d.dimx, dimy = d.dimy;
    const int ld = d.ld;
    const int dimx4 = (dimx + 3) >> 2;
    const int strip = e.y & 0xffff, chunk = e.y >> 16;
    const int c4a = strip * dimx4 / d.mstrips, c4b = (strip + 1) * dimx4 / d.mstrips;
    const int ox4 = c4a == 0 ? 0 : c4a - T;          // float4 column of lane 0
    const int Y0 = chunk * dimy / d.mchunks, Y1 = (chunk + 1) * dimy / d.mchunks;
    const int O = Y0 - 4 * T > 0 ? Y0 - 4 * T : 0;   // loaded rows [O, E)
    const int E = Y1 + 4 * T < dimy ? Y1 + 4 * T : dimy;
    const bool ey0 = O == 0, ey1 = E == dimy;
    const int kx4 = (dimx - NE) >> 2;                // first float4 column holding a right Keys cell
    const bool keys_left = ox4 == 0;                 // lane 0 holds the left Keys columns
    const bool keys_right = ox4 + PLANES - 1 >= kx4; // some lane holds a right Keys column

    const int k = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // 0: loader, else level
    const int lane = threadIdx.x & 63;
    const int g4 = ox4 + lane;
    const int gx = 4 * g4;
    const bool lin = g4 < dimx4;
    const bool lcore = g4 >= c4a && g4 < c4b;

    float* const* ring = a.adj ? d.lam : d.u;
    // the first recomputation launch of a checkpoint segment reads the
    // segment's saved pair (u[t0-2], u[t0-1]) where it lies
    const float* __restrict__ in_prev = a.ck_in >= 0 ? d.ckpt + (2LL * a.ck_in) * d.fstride : ring[2 * a.in_pair];
    const float* __restrict__ in_cur = a.ck_in >= 0 ? d.ckpt + (2LL * a.ck_in + 1) * d.fstride
                                                    : ring[2 * a.in_pair + 1];
    float* __restrict__ out_prev = ring[2 * a.out_pair];
    float* __restrict__ out_cur = ring[2 * a.out_pair + 1];

    // rows of this wave's level: [Ok, Ek) (the loader's are [O, E))
    const int Ok = ey0 ? 0 : O + 4 * k;
    const int Ek = ey1 ? dimy : E - 4 * k;
    const bool active = k <= a.nsteps;
    const int t = a.t_first + a.dir * (k - 1);
    const long long voff = (long long)(a.vslot_first + a.dir * (k - 1)) * d.fstride;

    // ---- prefetch streams: three register queues shared by all roles
    //   A: loader u_prev rows;  levels: m (v at Keys cells)
    //   B: loader u_cur rows;   levels: c3 (virtual-source store) or virt (adjoint)
    //   C: loader gradient rows (adjoint)
    const float* pA;
    const float* pB;
    int aLo, aHi, bLo, bHi;
    bool aLane, bLane;
    bool useB;
    if (k == 0) {
        pA = in_prev;
        pB = in_cur;
        aLo = bLo = O;
        aHi = bHi = E;
        aLane = bLane = lin;
        useB = true;
    } else {
        pA = d.vm;
        aLo = Ok;
        aHi = Ek;
        aLane = lin;
        bLo = Y0 > Ok ? Y0 : Ok;
        bHi = Y1 < Ek ? Y1 : Ek;
        bLane = lcore;
        if constexpr ((FLAGS & F_GRAD) != 0) {
            // a level past this launch's last step has no virtual-source slot
            pB = active ? d.virt + voff : d.vc3;
            useB = active;
        } else if constexpr ((FLAGS & F_VIRT) != 0) {
            pB = d.vc3;
            useB = true;
        } else {
            pB = d.vc3;
            useB = false;
        }
    }
    // Each wave role runs its own copy of the row loop (ROLE 0: the loader;
    // 1: wave 1 of the recomputation and the adjoint, which loads the model
    // rows (and in the adjoint starts from the stored gradient); 2: the other
    // level waves, which there take the model rows from the level before) and
    // issues exactly the loads it uses, the same ones
    // at every iteration (from cell 0 when a row is outside its range), so
    // the in-order count of outstanding vector-memory operations is the same
    // on every path and the wait for one queue slot never waits for the slots
    // behind it.  Loads are not free even when they hit: a placeholder load
    // costs about as much as a real one (1-2% of the forward steps' time).
    // The adjoint prefetches PD iterations ahead, the forward steps PD + 1
    // (their fewer registers leave room: 108-124 VGPRs, still 4 waves per
    // SIMD; the forward 3-6% faster, the gradient at 64 shots of 3008 x 1008
    // 0.6%).
    constexpr int pd = (FLAGS & F_GRAD) != 0 ? PD : PD + 1;
    float4 qA[PQ], qB[PQ], qC[PQ];
    auto issue = [&](auto RC, int slot_i, float4& A, float4& B, float4& C) {
        constexpr int ROLE = decltype(RC)::value;
        const int row = O + slot_i - LAG * k;
        const bool rowA = row >= aLo && row < aHi, rowB = row >= bLo && row < bHi;
        // level waves after the first take the model row from the previous
        // level through LDS
        if constexpr (ROLE != 2 || !LY::MRING) A = gld4(pA + ((aLane && rowA) ? row * ld + gx : 0));
        if constexpr ((FLAGS & F_GRAD) != 0) {
            const int ib = (useB && bLane && rowB) ? row * ld + gx : 0;
            if constexpr (ROLE == 0) B = gld4(pB + ib);
            else B = gld4_nt(pB + ib);
            if constexpr (ROLE == 1) {
                const bool cok = lcore && row >= Y0 && row < Y1;
                C = gld4(d.grad + (cok ? row * ld + gx : 0));
            }
        } else if constexpr (ROLE == 0 || (FLAGS & F_VIRT) != 0) {
            const int ib = (useB && bLane && rowB) ? row * ld + gx : 0;
            B = gld4(pB + ib);
        }
    };

    // ---- per-level constants
    float srcadd = 0.f;
    int srow = -1, sg4 = -1, sj = 0;
    if constexpr ((FLAGS & F_SRC) != 0) {
        if (k > 0 && active) {
            srow = (int)(d.isrc / ld);
            const int sx = (int)(d.isrc - (long long)srow * ld);
            sg4 = sx >> 2;
            sj = sx & 3;
            srcadd = inject_term(gld1(d.w + t), gld1(d.vel + d.isrc), d.dt2);
        }
    }
    float4 inj = make_float4(0.f, 0.f, 0.f, 0.f);
    unsigned injm = 0;
    if constexpr ((FLAGS & F_RESID) != 0) {
        if (k > 0 && active && d.rrow >= Ok && d.rrow < Ek) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int x = gx + j;
                if (lin && x >= d.rx0 && x < d.rx0 + d.nrec) {
                    injm |= 1u << j;
                    f4set(inj, j, inject_term(gld1(d.resid + (long long)t * d.nrec + (x - d.rx0)),
                                              gld1(d.vel + (long long)d.rrow * ld + x), d.dt2));
                }
            }
        }
    }
    int ckA = -1, ckB = -1;
    if ((FLAGS & F_REC) != 0 && a.K > 0 && a.dir > 0 && k > 0) {
        if ((t + 2) % a.K == 0 && (t + 2) / a.K < a.nseg) ckA = 2 * ((t + 2) / a.K);
        if ((t + 1) % a.K == 0 && (t + 1) / a.K < a.nseg) ckB = 2 * ((t + 1) / a.K) + 1;
    }

    // LDS ring bases of this wave: it reads ring k-1 and writes ring k
    const int curR = LY::CUR + (k > 0 ? k - 1 : 0) * NCUR * PLANES + lane;
    const int curW = LY::CUR + k * NCUR * PLANES + lane;
    const int prvR = LY::PRV + (k > 0 ? k - 1 : 0) * NPRV * PLANES + lane;
    const int prvW = LY::PRV + k * NPRV * PLANES + lane;
    const int grdR = LY::GRD + (k > 1 ? k - 2 : 0) * NGRD * PLANES + lane;   // sum after level k-1
    const int grdW = LY::GRD + (k > 0 ? k - 1 : 0) * NGRD * PLANES + lane;   // sum after level k
    const int mdlR = LY::MDL + (k > 1 ? k - 2 : 0) * NPRV * PLANES + lane;   // model rows of level k-1
    const int mdlW = LY::MDL + (k > 0 ? k - 1 : 0) * NPRV * PLANES + lane;   // model rows of level k
    const bool writes_next = k < T && k < a.nsteps;
    const int ksb = LY::KS + (k > 0 ? k - 1 : 0) * 32;   // this level's Keys scratch (float4 index)

    float4 win[PW];
#pragma unroll
    for (int q = 0; q < PW; ++q) win[q] = make_float4(0.f, 0.f, 0.f, 0.f);

    const int ET = ey1 ? dimy : E - 4 * T;
    const int nit = ET - O + LAG * T;
    int cP = 0;   // iteration mod NPRV (= NGRD)
    int cC = 0;   // iteration mod NCUR
    float4 Pk = make_float4(0.f, 0.f, 0.f, 0.f);   // u_prev read this iteration (level waves)

    auto body = [&](auto RC, int i, auto UC) {
        constexpr int ROLE = decltype(RC)::value;
        constexpr int U = decltype(UC)::value;
        issue(RC, i + pd, qA[(U + pd) % PQ], qB[(U + pd) % PQ], qC[(U + pd) % PQ]);
        if constexpr (ROLE == 0) {
            // loader: rows O + i of both input fields (and the gradient)
            S[curW + cC * PLANES] = qB[U % PQ];
            S[prvW + cP * PLANES] = qA[U % PQ];
        } else {
            // append row r+4 of level k-1 (written by wave k-1 LAG-4
            // iterations ago)
            const int cCr = cC + NCUR - (LAG - 4) >= NCUR ? cC - (LAG - 4) : cC + NCUR - (LAG - 4);
            win[U] = S[curR + cCr * PLANES];
            // u_prev of row r (level k-2, written by wave k-1 LAG iterations
            // ago); P1 keeps the one of row r-1
            const float4 P1 = Pk;
            const int cPr = cP + NPRV - LAG >= NPRV ? cP - LAG : cP + NPRV - LAG;   // (i-LAG) mod NPRV
            Pk = S[prvR + cPr * PLANES];
            // gradient sum after level k-1 (wave k-1, LAG iterations ago);
            // read as one 16-byte LDS access on every iteration (wave 1 reads
            // ring 0 and discards it), not inside the selection below, where it
            // would become four conditional 4-byte reads
            float4 Gk = make_float4(0.f, 0.f, 0.f, 0.f);
            if constexpr ((FLAGS & F_GRAD) != 0 && ROLE != 1) Gk = S[grdR + cPr * PLANES];
            // model of row r: wave 1 from its own prefetch queue, the later
            // level waves from wave k-1 (written with its u_prev row)
            float4 Mk;
            if constexpr (ROLE == 2 && LY::MRING) Mk = S[mdlR + cPr * PLANES];
            else Mk = qA[U % PQ];
            const int r = O + i - LAG * k;
            if (active && r >= Ok && r < Ek) {
                const float4 C = win[(U + PW - 4) % PW];
                const float4 P = Pk;
                const float4 M = Mk;
                float4 un;
                if (r < NE) {
                    un = make_float4(0.f, 0.f, 0.f, 0.f);
                } else if (r >= dimy - NE) {
                    // bottom Keys rows: every cell, corners included
                    const float4 C1 = win[(U + PW - 5) % PW], C2 = win[(U + PW - 6) % PW];
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        f4set(un, j, gx + j < dimx ? keys_update(f4at(P, j), f4at(C, j), f4at(P1, j), f4at(C1, j),
                                                                 f4at(C2, j), f4at(M, j), d.h2, d.dt2, d.hdt)
                                                   : 0.f);
                } else {
                    float4 up[NE], dn[NE];
#pragma unroll
                    for (int q = 1; q <= NE; ++q) {
                        up[q - 1] = win[(U + PW - 4 - q) % PW];
                        dn[q - 1] = win[(U + PW - 4 + q) % PW];
                    }
                    const float4 Lf = left4(C), Rf = right4(C);
                    un = lap_update4<NE>(up, dn, Lf, C, Rf, P, M);
                    // Left / right Keys columns (x < NE, x >= dimx-NE) and the
                    // row padding (x >= dimx).  The Keys cells sit in at most
                    // three lanes (lane 0, and the one or two lanes holding
                    // x >= dimx-NE); computing them there would take four
                    // sequential updates per lane.  Instead those owner lanes
                    // put their row segment (x-4 .. x+7 of u_prev and u_cur, and
                    // the model) in a small LDS scratch of this wave, one
                    // worker lane per Keys cell computes it (lanes 0..NE-1 the
                    // left cells, lanes 4..3+NE the right ones), and the owners
                    // read the results back.  LDS operations of one wave are
                    // performed in order, so no workgroup barrier is needed; the
                    // wave barriers only keep the compiler from moving the LDS
                    // accesses of the three phases across each other.
                    if (keys_left || keys_right) {
                        const float4 PL = left4(P), PR = right4(P);
                        float* const ks = (float*)&S[ksb];
                        const bool ownL = keys_left && lane == 0;
                        const bool ownR = keys_right && g4 >= kx4 && g4 < dimx4;
                        const int sR = 1 + (g4 - kx4);
                        if (ownL) {
                            float4* o = (float4*)ks;
                            o[0] = PL; o[1] = P; o[2] = PR; o[3] = Lf; o[4] = C; o[5] = Rf; o[6] = M;
                        }
                        if (ownR) {
                            float4* o = (float4*)(ks + 32 * sR);
                            o[0] = PL; o[1] = P; o[2] = PR; o[3] = Lf; o[4] = C; o[5] = Rf; o[6] = M;
                        }
                        __builtin_amdgcn_wave_barrier();
                        int wx = -1, ws = 0, wd = 0;
                        if (keys_left && lane < NE) {
                            wx = lane;
                            wd = 1;
                        }
                        if (keys_right && lane >= 4 && lane < 4 + NE) {
                            wx = dimx - NE + (lane - 4);
                            ws = 1 + ((wx >> 2) - kx4);
                            wd = -1;
                        }
                        if (wx >= 0) {
                            const float* o = ks + 32 * ws;
                            const int c = wx & 3;
                            // o[0..11]: u_prev at x-4..x+7; o[12..23]: u_cur; o[24..27]: model
                            const float kv = keys_update(o[4 + c], o[16 + c], o[4 + c + wd], o[16 + c + wd],
                                                         o[16 + c + 2 * wd], o[24 + c], d.h2, d.dt2, d.hdt);
                            ks[96 + 4 * ws + c] = kv;
                        }
                        __builtin_amdgcn_wave_barrier();
                        if (ownL) {
                            const float4 kr = *(const float4*)(ks + 96);
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                if (j < NE) f4set(un, j, f4at(kr, j));
                        }
                        if (ownR) {
                            const float4 kr = *(const float4*)(ks + 96 + 4 * sR);
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const int x = gx + j;
                                f4set(un, j, x >= dimx ? 0.f : (x >= dimx - NE ? f4at(kr, j) : f4at(un, j)));
                            }
                        }
                    }
                }
                if constexpr ((FLAGS & F_SRC) != 0) {
                    if (r == srow && g4 == sg4) {
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            if (j == sj) f4set(un, j, f4at(un, j) + srcadd);
                    }
                }
                if constexpr ((FLAGS & F_RESID) != 0) {
                    if (r == d.rrow) {
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            if (injm & (1u << j)) f4set(un, j, f4at(un, j) + f4at(inj, j));
                    }
                }
                const bool orow = r >= Y0 && r < Y1;
                const int i0 = r * ld + gx;
                if (orow && lcore) {
                    if constexpr ((FLAGS & F_REC) != 0) {
                        if (r == d.rrow) {
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                const int x = gx + j;
                                if (x >= d.rx0 && x < d.rx0 + d.nrec)
                                    gst1(d.frd + (long long)t * d.nrec + (x - d.rx0), f4at(un, j));
                            }
                        }
                    }
                    if constexpr ((FLAGS & F_VIRT) != 0) {
                        const float4 c3 = qB[U % PQ];
                        float4 vs;
                        vs.x = c3.x * ((P.x - 2.f * C.x) + un.x);
                        vs.y = c3.y * ((P.y - 2.f * C.y) + un.y);
                        vs.z = c3.z * ((P.z - 2.f * C.z) + un.z);
                        vs.w = c3.w * ((P.w - 2.f * C.w) + un.w);
                        gst4_nt(d.virt + voff + i0, vs);
                    }
                    if (ckA >= 0) gst4_nt(d.ckpt + ckA * d.fstride + i0, un);
                    if (ckB >= 0) gst4_nt(d.ckpt + ckB * d.fstride + i0, un);
                    // The field pair a launch leaves and the gradient are read
                    // back only by the next launch, after every other
                    // workgroup's traffic has passed through the caches:
                    // non-temporal stores (measured on MI355X against ordinary
                    // ones, 64 shots of 3008 x 1008, us per launch: forward
                    // 995 vs 1021, recomputation 1531 vs 1580, adjoint 1663
                    // vs 1761).
                    if (k == a.nsteps && !a.no_out) {
                        gst4_nt(out_prev + i0, C);
                        gst4_nt(out_cur + i0, un);
                    }
                }
                if constexpr ((FLAGS & F_GRAD) != 0) {
                    // wave 1 starts from the stored gradient (its own queue),
                    // wave k > 1 from wave k-1's sum
                    float4 gin;
                    if constexpr (ROLE == 1) gin = qC[U % PQ];
                    else gin = Gk;
                    const float4 v = qB[U % PQ];
                    float4 g;
                    g.x = gin.x + (-v.x) * un.x;
                    g.y = gin.y + (-v.y) * un.y;
                    g.z = gin.z + (-v.z) * un.z;
                    g.w = gin.w + (-v.w) * un.w;
                    if (k == a.nsteps) {
                        if (orow && lcore) gst4_nt(d.grad + i0, g);
                    } else {
                        S[grdW + cP * PLANES] = g;
                    }
                }
                if (writes_next) {
                    S[curW + cC * PLANES] = un;
                    S[prvW + cP * PLANES] = C;
                    if constexpr (LY::MRING) S[mdlW + cP * PLANES] = M;
                }
            }
        }
        cP = cP + 1 == NPRV ? 0 : cP + 1;
        cC = cC + 1 == NCUR ? 0 : cC + 1;
        if constexpr (U % BI == BI - 1) __syncthreads();
    };

    auto march = [&](auto RC) {
#pragma unroll
        for (int q = 0; q < pd; ++q) issue(RC, q, qA[q], qB[q], qC[q]);
        for (int base = 0; base < nit; base += PW) {
            body(RC, base + 0, IC<0>{});
            body(RC, base + 1, IC<1>{});
            body(RC, base + 2, IC<2>{});
            body(RC, base + 3, IC<3>{});
            body(RC, base + 4, IC<4>{});
            body(RC, base + 5, IC<5>{});
            body(RC, base + 6, IC<6>{});
            body(RC, base + 7, IC<7>{});
            body(RC, base + 8, IC<8>{});
            body(RC, base + 9, IC<9>{});
        }
    };
    if (k == 0) march(IC<0>{});
    else if (k == 1 && LY::MRING) march(IC<1>{});
    else march(IC<2>{});
}

}  // namespace tw
