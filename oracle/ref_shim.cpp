// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// C-linkage wrappers around the reference's own CPU time loops, so the
// reference path (compiled from its sources by oracle/Makefile into
// oracle/_ref/) can be called through ctypes.  The declarations below
// restate the reference prototypes (torchwi/propagator/td2d/propagator_cpu.h);
// the definitions come from the reference's .cpp files at link time.
void c_fdm_forward(float *forward, float *virt, float *u1, float *u2, float *u3,
                   float *vel, float *w, const int order, const int dimx,
                   const int dimy, const int nt, const float h, const float dt,
                   const float sx, const float sy, const float ry);
void c_fdm_backward(float *grad, float *virt, float *u1, float *u2, float *u3,
                    float *vel, float *resid, const int order, const int dimx,
                    const int dimy, const int nt, const float h, const float dt,
                    const float ry);
void c_fdm_forward_only(float *forward, float *u1, float *u2, float *u3, float *vel,
                        float *w, const int order, const int dimx, const int dimy,
                        const int nt, const float h, const float dt, const float sx,
                        const float sy, const float ry);

extern "C" {
void ref_forward(float *frd, float *virt, float *u1, float *u2, float *u3, float *vel,
                 float *w, int order, int dimx, int dimy, int nt, float h, float dt,
                 float sx, float sy, float ry)
{
    c_fdm_forward(frd, virt, u1, u2, u3, vel, w, order, dimx, dimy, nt, h, dt, sx, sy, ry);
}
void ref_backward(float *grad, float *virt, float *u1, float *u2, float *u3, float *vel,
                  float *resid, int order, int dimx, int dimy, int nt, float h, float dt,
                  float ry)
{
    c_fdm_backward(grad, virt, u1, u2, u3, vel, resid, order, dimx, dimy, nt, h, dt, ry);
}
void ref_forward_only(float *frd, float *u1, float *u2, float *u3, float *vel, float *w,
                      int order, int dimx, int dimy, int nt, float h, float dt, float sx,
                      float sy, float ry)
{
    c_fdm_forward_only(frd, u1, u2, u3, vel, w, order, dimx, dimy, nt, h, dt, sx, sy, ry);
}
}
