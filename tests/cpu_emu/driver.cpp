// Host-thread run of the library's batched plan (td2d_hip.hip compiled for
// the CPU against the stand-in runtime in hip/hip_runtime.h), for the CPU
// test suite: the kernels' source executes under AddressSanitizer, and the
// gathers and gradients are compared with the oracle.
//
//   driver IN OUT
//
// IN (little-endian): int32 order, nx, ny, nt, S, K, graph, time_block,
// tile_rows; float32 h, dt;
// float32 vel[dimy*ld] (padded model, row stride ld = dimx rounded up to 32);
// float32 w[nt]; float32 pos[S*3] (sx, sy, ry); float32 resid[S*nt*nx].
// OUT: float32 frd[S*nt*nx], float32 grad[S*dimy*ld], int32 launches.
#include "torchwi_hip.h"
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

thread_local uint3_ threadIdx, blockIdx, blockDim, gridDim;
pthread_barrier_t g_bar;
pthread_barrier_t g_wbar[EMU_MAX_WAVES];
int g_dpp[EMU_MAX_WAVES][64];
std::vector<std::function<void()>>* g_capture = nullptr;
long long g_launches = 0;

template <class T>
static void rd(FILE* f, T* p, size_t n)
{
    if (fread(p, sizeof(T), n, f) != n) {
        fprintf(stderr, "short input\n");
        exit(2);
    }
}

int main(int argc, char** argv)
{
    if (argc != 3) return 2;
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[9];
    float hd[2];
    rd(f, hdr, 9);
    rd(f, hd, 2);
    const int order = hdr[0], nx = hdr[1], ny = hdr[2], nt = hdr[3], S = hdr[4], K = hdr[5], graph = hdr[6],
              tb = hdr[7], tile_rows = hdr[8];
    const int dimx = nx + order, dimy = ny + order, ld = (dimx + 31) / 32 * 32;
    const long long nf1 = (long long)dimy * ld, nf = nf1 * S;
    const int RING = 4, slots = K ? K : nt, nseg = K ? (nt + K - 1) / K : 0;
    // every buffer kind its own allocation, so AddressSanitizer sees the
    // bounds of each
    std::vector<float> vel(nf1), vm(nf1), vc3(nf1), w(nt), pos(3 * S), resid((size_t)S * nt * nx);
    rd(f, vel.data(), vel.size());
    rd(f, w.data(), w.size());
    rd(f, pos.data(), pos.size());
    rd(f, resid.data(), resid.size());
    fclose(f);
    std::vector<float> u(RING * nf), lam(K ? RING * nf : 0), frd((size_t)S * nt * nx), virt(slots * nf), grad(nf),
        ckpt(K ? 2LL * nseg * nf : 0);
    TwPlanSpec spec = {};
    spec.order = order;
    spec.nt = nt;
    spec.h = hd[0];
    spec.dt = hd[1];
    spec.ckpt_every = K;
    spec.graph_steps = graph;
    spec.time_block = tb;
    spec.tile_rows = tile_rows;
    // TW_EMU_GRAD_GROUP=G: the group adjoint (a library built with TW_TEST_GRAD_GROUP)
    const char* gg = getenv("TW_EMU_GRAD_GROUP");
    spec.grad_group = gg ? atoi(gg) : 0;
    // TW_EMU_STREAMS=2: the blocked kernel's two shot groups
    const char* sm = getenv("TW_EMU_STREAMS");
    spec.streams = sm ? atoi(sm) : 0;
    std::vector<TwShotSpec> sh(S);
    for (int s = 0; s < S; ++s) {
        TwShotSpec& c = sh[s];
        c = TwShotSpec{};
        c.dimx = dimx;
        c.dimy = dimy;
        c.ld = ld;
        c.sx = pos[3 * s];
        c.sy = pos[3 * s + 1];
        c.ry = pos[3 * s + 2];
        c.rx0 = 0;
        c.nrec = nx;
        c.vel = vel.data();
        c.vm = vm.data();
        c.vc3 = vc3.data();
        c.w = w.data();
        const long long o = s * nf1;
        c.u = u.data() + RING * o;
        c.lam = K ? lam.data() + RING * o : c.u;
        c.frd = frd.data() + (size_t)s * nt * nx;
        c.resid = resid.data() + (size_t)s * nt * nx;
        c.virt = virt.data() + slots * o;
        c.ckpt = K ? ckpt.data() + 2LL * nseg * o : nullptr;
        c.grad = grad.data() + o;
    }
    const size_t ws = tw_plan_workspace_bytes(&spec, sh.data(), S);
    if (!ws) {
        fprintf(stderr, "workspace: %s\n", tw_last_error());
        return 1;
    }
    std::vector<char> work(ws);
    void* plan = nullptr;
    // TW_EMU_STATIC=1: the plan's descriptors in the library's static workspace
    const char* st = getenv("TW_EMU_STATIC");
    const int rc0 = (st && st[0] == '1') ? tw_plan_create_static(&spec, sh.data(), S, nullptr, &plan)
                                         : tw_plan_create(&spec, sh.data(), S, work.data(), ws, nullptr, &plan);
    if (rc0 || tw_plan_forward(plan, nullptr) || tw_plan_backward(plan, 0, nullptr)) {
        fprintf(stderr, "plan: %s\n", tw_last_error());
        return 1;
    }
    // TW_EMU_BACKWARD_TWICE=1: a second backward after the same forward must
    // give the same gradients bit for bit (it recomputes every segment)
    const char* bt = getenv("TW_EMU_BACKWARD_TWICE");
    if (bt && bt[0] == '1') {
        const std::vector<float> first = grad;
        if (tw_plan_backward(plan, 0, nullptr)) {
            fprintf(stderr, "second backward: %s\n", tw_last_error());
            return 1;
        }
        if (std::memcmp(first.data(), grad.data(), sizeof(float) * grad.size()) != 0) {
            fprintf(stderr, "second backward differs from the first\n");
            return 3;
        }
    }
    tw_plan_destroy(plan);
    FILE* g = fopen(argv[2], "wb");
    if (!g) return 2;
    fwrite(frd.data(), sizeof(float), frd.size(), g);
    fwrite(grad.data(), sizeof(float), grad.size(), g);
    const int launches = (int)g_launches;
    fwrite(&launches, sizeof(int), 1, g);
    fclose(g);
    return 0;
}
