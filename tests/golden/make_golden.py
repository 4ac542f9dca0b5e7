"""Write the golden fixtures of the time-domain 2D propagator.

The outputs come from the reference's own CPU path (torchwi/propagator/td2d/
td2d_cpu.cpp, propagator_cpu.cpp, propagator_forward_cpu.cpp) compiled from
its sources by ``make -C oracle ref`` into oracle/_ref/libtd2d_ref.so.  Only
inputs and outputs are stored (float32 .npz):

  td2d_o{2,4,8}.npz  seeded random-velocity case on a 37x23 model, nt=160:
                     vel, w, positions, resid -> frd, grad, and two virt slices
  three_layer_100x200_o8.npz     the 100x200 three-layer model behind a 20-cell pad,
                     nt=1000, order 8: forward gather and the L2-misfit
                     gradient against the gather of the +5% blob model
                     (TimeFunction.py crop/transposition applied to grad)

Run:  python tests/golden/make_golden.py   (needs oracle/_ref built)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests.helpers import small_case  # noqa: E402
from torchwi_amd import workloads as W  # noqa: E402


def small(order):
    c = small_case(order, nx=37, ny=23, nt=160, seed=100 + order)
    vpad = O.pad_vel(c["vel"], order)
    geo = (order, c["dimx"], c["dimy"], c["h"], c["dt"])
    frd, virt = O.forward(vpad, c["w"], *geo, c["sx"], c["sy"], c["ry"], impl="ref")
    resid = np.random.default_rng(200 + order).standard_normal(frd.shape).astype(np.float32)
    grad = O.backward(virt, vpad, resid, *geo, c["ry"], impl="ref")
    slices = [40, 120]
    return dict(vel=c["vel"], w=c["w"], order=order, h=c["h"], dt=c["dt"], sx=c["sx"], sy=c["sy"],
                ry=c["ry"], resid=resid, frd=frd, grad=grad, virt_steps=np.array(slices),
                virt_slices=virt[slices])


def three_layer_100x200():
    c = W.three_layer_100x200()
    order, nx, ny = c["order"], c["nx"], c["ny"]
    dimx, dimy = nx + order, ny + order
    vpad = O.pad_vel(c["vel"], order)
    a = (order, dimx, dimy, c["h"], c["dt"], c["sx"][0], c["sy"][0], c["ry"][0])
    d_obs, _ = O.forward(O.pad_vel(c["vel_true"], order), c["w"], *a, store_virt=False, impl="ref")
    frd, virt = O.forward(vpad, c["w"], *a, impl="ref")
    grad = O.backward(virt, vpad, frd - d_obs, order, dimx, dimy, c["h"], c["dt"], c["ry"][0], impl="ref")
    return dict(vel=c["vel"], vel_true=c["vel_true"], w=c["w"], order=order, h=c["h"], dt=c["dt"],
                sx=c["sx"][0], sy=c["sy"][0], ry=c["ry"][0], frd=frd,
                grad=O.crop_grad(grad, order, nx, ny))


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libtd2d_ref.so is not built: run `make -C oracle ref`")
    for order in (2, 4, 8):
        np.savez_compressed(os.path.join(HERE, "td2d_o%d.npz" % order), **small(order))
    np.savez_compressed(os.path.join(HERE, "three_layer_100x200_o8.npz"), **three_layer_100x200())
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
