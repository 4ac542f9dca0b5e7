"""The library's own source on the CPU: td2d_hip.hip (plan code, graph
capture and replay, the one-step and the temporally blocked kernels) is
compiled for the host against a stand-in HIP runtime
(tests/cpu_emu/hip/hip_runtime.h: one host thread per GPU thread, a barrier
for __syncthreads, per-wave barriers for wave barriers and DPP lane shifts)
with AddressSanitizer, and run on small batched plans.  Every access of the
kernels and of the host code is bounds-checked, and the gathers and
per-shot gradients must equal the oracle's bit for bit (the host's IEEE
float and double arithmetic is what the kernels are written to
reproduce)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "torchwi_amd", "csrc")
EMU = os.path.join(ROOT, "tests", "cpu_emu")
CXX = "/opt/rocm/llvm/bin/clang++"


@pytest.fixture(scope="session")
def driver(tmp_path_factory):
    if not os.path.exists(CXX):
        pytest.skip("no host clang++")
    d = str(tmp_path_factory.mktemp("cpu_emu"))
    shutil.copy(os.path.join(CSRC, "td2d_hip.hip"), os.path.join(d, "td2d_hip.cpp"))
    shutil.copy(os.path.join(CSRC, "td2d_arith.hpp"), d)
    vec = open(os.path.join(CSRC, "td2d_vec.hpp")).read().replace("__attribute__((address_space(1)))", "")
    open(os.path.join(d, "td2d_vec.hpp"), "w").write(vec)
    shutil.copy(os.path.join(CSRC, "td2d_pipe.hpp"), d)
    # TW_TEST_GRAD_GROUP lifts the plan's refusal of the group adjoint (switched
    # off on the GPU) in this CPU build only
    flags = [CXX, "-std=c++17", "-O0", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-ffp-contract=off",
             "-pthread", "-w", "-DTW_TEST_GRAD_GROUP", "-I", d, "-I", EMU, "-I", os.path.join(ROOT, "include")]
    subprocess.run(flags + ["-c", "-o", os.path.join(d, "lib.o"), os.path.join(d, "td2d_hip.cpp")], check=True)
    exe = os.path.join(d, "driver")
    subprocess.run(flags + ["-o", exe, os.path.join(EMU, "driver.cpp"), os.path.join(d, "lib.o")], check=True)
    return exe


@pytest.mark.parametrize("order,nx,ny,nt,S,K,graph,tb,rows,static", [
    (8, 20, 12, 10, 2, 0, 4, 1, 0, 0), (4, 37, 15, 9, 3, 4, 0, 1, 0, 0), (2, 70, 18, 8, 2, 3, 3, 1, 0, 0),
    (8, 24, 14, 9, 2, 0, 0, 3, 0, 0), (8, 250, 26, 6, 1, 0, 0, 3, 16, 0), (4, 40, 22, 6, 2, 4, 0, 2, 0, 0),
    (2, 33, 21, 8, 2, 4, 0, 4, 0, 0), (4, 40, 70, 6, 2, 3, 0, 3, 0, 0),
    # the plan descriptors in the static workspace (tw_plan_create_static)
    (8, 24, 14, 9, 2, 3, 0, 3, 0, 1), (4, 30, 16, 8, 3, 0, 0, 1, 0, 1),
    # two shot groups on two streams (static = 2: TW_EMU_STREAMS=2), blocked and one step per launch
    (8, 24, 14, 10, 3, 3, 0, 3, 0, 2), (4, 40, 22, 9, 5, 0, 0, 3, 0, 2), (2, 33, 21, 13, 2, 6, 0, 2, 0, 2),
    (8, 20, 12, 10, 5, 0, 4, 1, 0, 2), (4, 37, 15, 9, 3, 4, 3, 1, 0, 2), (2, 30, 14, 11, 4, 3, 0, 1, 0, 2)])
def test_kernel_source_on_host_threads_matches_oracle(driver, tmp_path, order, nx, ny, nt, S, K, graph, tb, rows,
                                                      static):
    run_case(driver, tmp_path, order, nx, ny, nt, S, K, graph, tb, rows, static)


@pytest.mark.parametrize("order,nx,ny,nt,S,K,graph,tb,rows,static", [
    (8, 24, 14, 10, 3, 3, 0, 3, 0, 2), (4, 40, 22, 14, 2, 4, 0, 2, 0, 0), (8, 20, 12, 9, 2, 30, 0, 3, 0, 0),
    (4, 37, 15, 9, 3, 4, 3, 1, 0, 2)])
def test_second_backward_on_host_threads(driver, tmp_path, order, nx, ny, nt, S, K, graph, tb, rows, static):
    """A second backward after one forward (the blocked forward pass left the
    last segment's virtual sources, which the first backward used and
    overwrote) gives the first backward's gradients bit for bit, and both
    the oracle's; K = 30 > nt: one segment, stored whole by the forward."""
    run_case(driver, tmp_path, order, nx, ny, nt, S, K, graph, tb, rows, static, twice=True)


def run_case(driver, tmp_path, order, nx, ny, nt, S, K, graph, tb, rows, static, twice=False):
    rng = np.random.default_rng(order * 7 + nx)
    h, dt = 10.0, 1e-3
    vel = (1500.0 + 2500.0 * rng.random((nx, ny))).astype(np.float32)
    w = rng.standard_normal(nt).astype(np.float32)
    dimx, dimy = nx + order, ny + order
    ld = (dimx + 31) // 32 * 32
    vpad = O.pad_vel(vel, order).reshape(dimy, dimx)
    vld = np.zeros((dimy, ld), np.float32)
    vld[:, :dimx] = vpad
    pos = np.array([[(rng.integers(0, nx) + 0.5) * h, (rng.integers(0, ny) + 0.5) * h,
                     (rng.integers(0, ny) + 0.5) * h] for _ in range(S)], np.float32)
    resid = rng.standard_normal((S, nt, nx)).astype(np.float32)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        np.array([order, nx, ny, nt, S, K, graph, tb, rows], np.int32).tofile(f)
        np.array([h, dt], np.float32).tofile(f)
        for a in (vld, w, pos, resid):
            a.tofile(f)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", TW_EMU_STATIC=str(int(static == 1)),
               TW_EMU_STREAMS=str(2 if static == 2 else 0), TW_EMU_BACKWARD_TWICE=str(int(twice)))
    r = subprocess.run([driver, fin, fout], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    out = np.fromfile(fout, np.float32)
    frd = out[:S * nt * nx].reshape(S, nt, nx)
    grad = out[S * nt * nx:S * nt * nx + S * dimy * ld].reshape(S, dimy, ld)
    assert out[-1:].view(np.int32)[0] > 0
    for s in range(S):
        sx, sy, ry = (float(v) for v in pos[s])
        frd_o, virt_o = O.forward(vpad, w, order, dimx, dimy, h, dt, sx, sy, ry)
        assert np.array_equal(frd[s], frd_o), f"shot {s} gather"
        g_o = O.backward(virt_o, vpad, resid[s], order, dimx, dimy, h, dt, ry).reshape(dimy, dimx)
        assert np.array_equal(grad[s][:, :dimx], g_o), f"shot {s} gradient"
        assert not np.any(grad[s][:, dimx:] != 0), f"shot {s} row padding"


@pytest.mark.parametrize("order,nx,ny,nt,S,G,K,graph", [
    (8, 40, 20, 120, 5, 2, 0, 100),   # the configuration whose MI355X run faulted
    (4, 30, 18, 20, 3, 4, 0, 0), (2, 33, 16, 16, 6, 3, 7, 0), (8, 26, 14, 18, 4, 4, 9, 6)])
def test_group_adjoint_source_on_host_threads(driver, tmp_path, order, nx, ny, nt, S, G, K, graph):
    """The group adjoint (k_vadj_group, refused by plans on the GPU) on host
    threads under AddressSanitizer: gathers equal the oracle's bit for bit,
    each group leader's gradient equals the sum of its shots' oracle
    gradients to rounding (only the order of the sum over shots and time
    differs), the other shots' gradients stay zero."""
    rng = np.random.default_rng(order * 13 + S)
    h, dt = 10.0, 1e-3
    vel = (1500.0 + 2500.0 * rng.random((nx, ny))).astype(np.float32)
    w = rng.standard_normal(nt).astype(np.float32)
    dimx, dimy = nx + order, ny + order
    ld = (dimx + 31) // 32 * 32
    vpad = O.pad_vel(vel, order).reshape(dimy, dimx)
    vld = np.zeros((dimy, ld), np.float32)
    vld[:, :dimx] = vpad
    pos = np.array([[(rng.integers(0, nx) + 0.5) * h, (rng.integers(0, ny) + 0.5) * h,
                     (rng.integers(0, ny) + 0.5) * h] for _ in range(S)], np.float32)
    resid = rng.standard_normal((S, nt, nx)).astype(np.float32)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        np.array([order, nx, ny, nt, S, K, graph, 1, 0], np.int32).tofile(f)
        np.array([h, dt], np.float32).tofile(f)
        for a in (vld, w, pos, resid):
            a.tofile(f)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", TW_EMU_GRAD_GROUP=str(G))
    r = subprocess.run([driver, fin, fout], capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    out = np.fromfile(fout, np.float32)
    frd = out[:S * nt * nx].reshape(S, nt, nx)
    grad = out[S * nt * nx:S * nt * nx + S * dimy * ld].reshape(S, dimy, ld)
    per = []
    for s in range(S):
        sx, sy, ry = (float(v) for v in pos[s])
        frd_o, virt_o = O.forward(vpad, w, order, dimx, dimy, h, dt, sx, sy, ry)
        assert np.array_equal(frd[s], frd_o), f"shot {s} gather"
        per.append(O.backward(virt_o, vpad, resid[s], order, dimx, dimy, h, dt, ry).reshape(dimy, dimx))
    for lead in range(0, S, G):
        want = np.sum(np.array(per[lead:lead + G], np.float64), axis=0)
        got = grad[lead][:, :dimx].astype(np.float64)
        assert np.linalg.norm(got - want) <= 1e-6 * np.linalg.norm(want), f"group {lead}"
        for s in range(lead + 1, min(lead + G, S)):
            assert not np.any(grad[s] != 0), f"shot {s} is not a group leader"

