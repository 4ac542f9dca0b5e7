/*
 * torchwi_hip.h -- C ABI of the MI355X (gfx950) time-domain 2D acoustic
 * propagator and its adjoint gradient.
 *
 * Conventions shared by every entry point:
 *   - all array arguments are DEVICE pointers (HBM), FP32 unless stated;
 *   - `stream` is a hipStream_t passed as void* (NULL = the legacy default
 *     stream); work is enqueued on it and the call returns without
 *     synchronising;
 *   - the library allocates no device memory: every buffer, including the
 *     batched path's descriptor workspace, is owned by the caller;
 *   - the return value is 0 on success and a negative status otherwise;
 *     tw_last_error() then describes the failure (per host thread).
 *
 * Grid convention (the reference's): padded grid dimx x dimy = (nx+order) x
 * (ny+order), x fastest; halo ne = order/2; interior cells ix in
 * [ne, dimx-ne), iy in [ne, dimy-ne); top rows iy < ne stay zero (free
 * surface); left/right/bottom halo cells carry the Keys absorbing boundary.
 * Source and receiver positions sx, sy, ry are physical coordinates that the
 * library truncates exactly as the reference does: (int)(sx/h) in FP32.
 */
#ifndef TORCHWI_HIP_H
#define TORCHWI_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define TW_OK 0
#define TW_EINVAL -1   /* bad argument (order, sizes, positions, NULL buffer) */
#define TW_EHIP -2     /* a HIP runtime call failed */

/* Text of the last error raised on the calling host thread ("" if none). */
const char *tw_last_error(void);

/* Library version string. */
const char *tw_version(void);

/* ------------------------------------------------------------------------
 * Single-shot entry points.  These are what the reference's extension
 * modules torchwi.propagator.td2d_cuda / td2d_forward_cuda / td2d_exa_cuda
 * bind; the argument lists are theirs with device pointers, the wavelet on
 * the device, and a stream.
 * ---------------------------------------------------------------------- */

/* Forward modelling with full virtual-source storage.
 * Replaces c_fdm_forward (torchwi/propagator/td2d/propagator_cpu.cpp:5-79)
 * and cuda_fdm_forward (propagator_cuda.cu:4-81), bound by
 * fdm_forward (wrapper_cpu.cpp:44-78 / wrapper_cuda.cpp).
 * frd: (nt, nx); virt: (nt, dimx*dimy); u1,u2,u3: dimx*dimy work fields;
 * vpad: dimx*dimy padded model; w: (nt) wavelet. */
int tw_td2d_forward(float *frd, float *virt, float *u1, float *u2, float *u3,
                    const float *vpad, const float *w, int order, int dimx, int dimy,
                    int nt, float h, float dt, float sx, float sy, float ry,
                    void *stream);

/* Forward modelling without storage.
 * Replaces c_fdm_forward_only (propagator_forward_cpu.cpp:5-71) and
 * cuda_fdm_forward_only (propagator_forward_cuda.cu:4-77), bound by
 * fdm_forward (wrapper_forward_cpu.cpp:27-58). */
int tw_td2d_forward_only(float *frd, float *u1, float *u2, float *u3,
                         const float *vpad, const float *w, int order, int dimx,
                         int dimy, int nt, float h, float dt, float sx, float sy,
                         float ry, void *stream);

/* Adjoint propagation of the residual with the imaging condition
 * grad += -virt[it] * lambda[it], accumulated into grad (dimx*dimy).
 * Replaces c_fdm_backward (propagator_cpu.cpp:82-152) and cuda_fdm_backward
 * (propagator_cuda.cu:84-156), bound by fdm_backward (wrapper_cpu.cpp:80-112,
 * which zeroes grad first).  resid: (nt, nx). */
int tw_td2d_backward(float *grad, const float *virt, float *u1, float *u2, float *u3,
                     const float *vpad, const float *resid, int order, int dimx,
                     int dimy, int nt, float h, float dt, float ry, void *stream);

/* Excitation-amplitude forward: exa (dimx*dimy) keeps the largest-magnitude
 * virtual source per cell and iexa (int16) its step.  nt <= 32767.
 * Replaces cuda_fdm_forward of propagator_exa_cuda.cu:5-84, bound by
 * fdm_forward (wrapper_exa_cuda.cpp:9-45). */
int tw_td2d_exa_forward(float *frd, float *exa, short *iexa, float *u1, float *u2,
                        float *u3, const float *vpad, const float *w, int order,
                        int dimx, int dimy, int nt, float h, float dt, float sx,
                        float sy, float ry, void *stream);

/* Excitation-amplitude gradient: grad = -exa * lambda at step iexa
 * (assignment).  Replaces cuda_fdm_backward of propagator_exa_cuda.cu:87-159,
 * bound by fdm_backward (wrapper_exa_cuda.cpp:47-83). */
int tw_td2d_exa_backward(float *grad, const float *exa, const short *iexa, float *u1,
                         float *u2, float *u3, const float *vpad, const float *resid,
                         int order, int dimx, int dimy, int nt, float h, float dt,
                         float ry, void *stream);

/* ------------------------------------------------------------------------
 * Batched shots (in-GPU shot batching).  No counterpart in the reference,
 * which runs one shot per call and spreads shots over processes with Horovod
 * (examples/time/inv_torch.py:123-143, io/DistributedDataloader.py:7-12);
 * this is the replacement for that loop.  All shots of a plan share order,
 * nt, h and dt; each has its own grid, model, wavelet, source, receiver
 * window and buffers.  Shots march in lockstep: one launch advances every
 * shot of the plan (of its shot group, when the plan runs two groups on two
 * streams, TwPlanSpec.streams) by one time step, or by 2..4 with the
 * temporally blocked kernel.
 * ---------------------------------------------------------------------- */

typedef struct TwPlanSpec {
    int order;        /* 2, 4 or 8 */
    int nt;           /* time steps */
    float h, dt;      /* grid spacing, time step */
    int ckpt_every;   /* K > 0: checkpoint every K steps and recompute;
                         0: full virtual-source storage (nt slots) */
    int tile_rows;    /* one-step kernel tile height (64 x 32, 64 x 64, 128 x 16 or
                         128 x 32 with tile_cols), or the blocked kernel's
                         row-chunk height (>= 16); 0 = automatic */
    int graph_steps;  /* > 0: time steps are launched as replays of hipGraphs
                         captured once per plan (this many steps per forward /
                         full-storage adjoint block; one graph per checkpoint
                         segment); 0: one direct launch per step */
    int time_block;   /* 2..4: temporally blocked kernel advancing this many
                         steps per launch (ckpt_every must be a multiple);
                         0 or 1: one step per launch */
    int forward_only; /* 1: modelling only -- no virtual-source storage, no
                         adjoint ring (virt, lam, ckpt may be NULL; ckpt_every
                         must be 0; tw_plan_backward is refused) */
    int tile_cols;    /* one-step kernel tile width: 64 or 128, 0 = automatic */
    int grad_group;   /* 0 or 1: one gradient per shot.  > 1 (an adjoint that
                         accumulates the imaging condition of a group of shots
                         of one model into one gradient) is currently refused
                         with TW_EINVAL: its kernel faulted on MI355X for a
                         cause not yet found */
    int streams;      /* 1 runs all shots on the caller's stream; 2 runs
                         them as two shot groups of about equal work, the
                         second on a stream of the plan's own, each group
                         marching through its steps and checkpoint segments
                         independently (the caller's stream waits for both at
                         the end of each pass); 0 = automatic: 2 whenever the
                         plan has two shots or more (grad_group > 1 keeps one
                         group).  Results are the same either way */
} TwPlanSpec;

typedef struct TwShotSpec {
    int dimx, dimy;   /* padded grid of this shot */
    int ld;           /* row stride in floats (>= dimx, multiple of 4); every
                         field buffer below is 16-byte aligned */
    float sx, sy, ry; /* physical source / receiver-row positions */
    int rx0, nrec;    /* receiver window: interior columns [rx0, rx0+nrec) */
    const float *vel; /* (dimy, ld) padded model */
    float *vm, *vc3;  /* (dimy, ld) derived-model work arrays, filled by the
                         plan; shots sharing vel share them */
    const float *w;   /* (nt) wavelet */
    float *u;         /* forward ring: 4 fields of dimy*ld */
    float *lam;       /* adjoint ring: 4 fields of dimy*ld (may equal u when
                         ckpt_every == 0) */
    float *frd;       /* (nt, nrec) gather */
    const float *resid; /* (nt, nrec) residual, read by tw_plan_backward */
    float *virt;      /* ckpt_every == 0: nt fields; else ckpt_every fields */
    float *ckpt;      /* ckpt_every > 0: 2*ceil(nt/ckpt_every) fields; else NULL */
    float *grad;      /* dimy*ld gradient (padded grid) */
} TwShotSpec;

/* Bytes of device workspace a plan for these shots needs (descriptors and
 * the tile table); 0 if the specs are invalid. */
size_t tw_plan_workspace_bytes(const TwPlanSpec *spec, const TwShotSpec *shots,
                               int nshots);

/* Validate the shots, build the plan and upload its descriptors into the
 * caller's device workspace (stream-ordered).  *plan receives an opaque
 * host handle, released with tw_plan_destroy. */
int tw_plan_create(const TwPlanSpec *spec, const TwShotSpec *shots, int nshots,
                   void *workspace, size_t workspace_bytes, void *stream, void **plan);

/* Forward pass of every shot: fills frd, and virt (full storage) or the
 * checkpoints (with time_block > 1 also the last segment's virtual sources,
 * which the next tw_plan_backward uses instead of recomputing them). */
int tw_plan_forward(void *plan, void *stream);

/* Adjoint pass of every shot (recomputing each checkpoint segment first when
 * ckpt_every > 0): grad = (accumulate ? grad : 0) + sum_t -virt[t]*lambda[t]
 * per shot, in its own grad buffer.  May be repeated after one forward pass
 * (every segment is then recomputed). */
int tw_plan_backward(void *plan, int accumulate, void *stream);

void tw_plan_destroy(void *plan);

/* Shot groups the plan runs on separate streams (1 or 2, see
 * TwPlanSpec.streams), or a negative status. */
int tw_plan_shot_groups(void *plan);

/* Size of the static plan workspace of each device (4 MiB). */
#define TW_STATIC_WORKSPACE_BYTES 4194304

/* tw_plan_create without a caller workspace: the plan's descriptors, tile
 * lists and control words go to a device-resident array the library holds
 * for each device (TW_STATIC_WORKSPACE_BYTES, kept until the process ends),
 * and the call allocates no device memory (no hipMalloc, no allocating
 * library routine; host memory for the plan handle only).  No counterpart in
 * the reference: this is for callers that cannot allocate device memory
 * (code that runs between graph captures, serving loops with a fixed memory
 * footprint, harnesses that forbid allocation).  Limits: the plan must fit
 * the workspace (plan sizes beyond it return TW_EINVAL; tw_plan_workspace_bytes
 * gives the size); graph_steps must be 0 (graph instantiation may allocate)
 * and streams 0 or 1 (a plan stream and events are not created: the plan runs
 * on the caller's stream);
 * at most one such plan per device exists at a time -- a second
 * tw_plan_create_static on a device whose workspace is held returns
 * TW_EINVAL until the holder is destroyed.  The workspace keeps the last
 * plan's contents; every plan writes all of its own before use.  Work of two
 * plans in flight on one device at once cannot share it, which the single
 * holder rule prevents.  The plan runs with tw_plan_forward /
 * tw_plan_backward and is released with tw_plan_destroy, as any plan. */
int tw_plan_create_static(const TwPlanSpec *spec, const TwShotSpec *shots, int nshots,
                          void *stream, void **plan);

/* Kinds of step kernel a plan launches. */
#define TW_KIND_FWD 0      /* forward, gathers (+ checkpoint saves) */
#define TW_KIND_FWD_VIRT 1 /* forward with full virtual-source storage */
#define TW_KIND_RECOMP 2   /* recomputation of a checkpoint segment */
#define TW_KIND_ADJ 3      /* adjoint + imaging condition */
#define TW_NKINDS 4

/* Enable (1) or disable (0) HIP-event timing of this plan's step kernels: an
 * event pair around each direct launch, or around each graph replay (then
 * counted as the replay's number of step kernels), on the launch stream. */
int tw_plan_set_timing(void *plan, int enable);

/* Accumulated event time (ms) and launch count per kind (arrays of
 * TW_NKINDS); waits for the recorded events.  reset != 0 clears the totals. */
int tw_plan_timing(void *plan, int reset, double *ms, long long *count);

/* ------------------------------------------------------------------------
 * Measurement aid.  No counterpart in the reference: a plain streaming copy
 * of `bytes` bytes (a multiple of 16; both pointers 16-byte aligned, device
 * memory, not overlapping), one 16-byte load and store per thread, used by
 * bench.py as the achievable HBM rate of the same device.
 * ---------------------------------------------------------------------- */
int tw_stream_copy(void *dst, const void *src, size_t bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* TORCHWI_HIP_H */
