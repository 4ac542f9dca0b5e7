"""Host-side checks of the serving loop's admission arithmetic (no GPU)."""
import torch

from torchwi_amd.serving import Request, chunk_plan, chunk_time_block, io_bytes, make_stream
from torchwi_amd.shots import default_ckpt_interval, plan_bytes


def _req(kind, nx=200, ny=80, nt=400, S=10):
    return Request(kind=kind, vel=torch.full((nx, ny), 2000.0), w=torch.zeros(nt), sx=[15.0] * S,
                   sy=[25.0] * S, ry=[25.0] * S, h=10.0, dt=1e-3,
                   d_obs=torch.zeros(1) if kind == "gradient" else None)


def test_gradient_uses_full_storage_when_it_fits_the_cap():
    r = _req("gradient")
    full = plan_bytes([(r.nx, r.ny, 8)], [dict(model=0, nrec=r.nx)] * r.nshots, r.nt, 0)
    per, K, b = chunk_plan(r, full)
    assert (per, K, b) == (r.nshots, 0, full)


def test_gradient_chunks_with_checkpoints_under_a_small_cap():
    r = _req("gradient", nt=900)
    K = default_ckpt_interval(r.nt, 3)
    three = plan_bytes([(r.nx, r.ny, 8)], [dict(model=0, nrec=r.nx)] * 3, r.nt, K)
    per, k2, b = chunk_plan(r, three + 1)
    assert (per, k2, b) == (3, K, three)


def test_forward_requests_hold_no_virtual_sources():
    r = _req("forward", nt=900, S=4)
    per, K, b = chunk_plan(r, 10 ** 12)
    assert (per, K) == (4, 0)
    assert b == plan_bytes([(r.nx, r.ny, 8)], [dict(model=0, nrec=r.nx)] * 4, r.nt, 0, forward_only=True)
    assert b < plan_bytes([(r.nx, r.ny, 8)], [dict(model=0, nrec=r.nx)] * 4, r.nt, 0) / 50


def test_chunking_depends_on_the_request_only():
    r = _req("forward", S=7)
    assert chunk_plan(r, 10 ** 9) == chunk_plan(_req("forward", S=7), 10 ** 9)
    assert io_bytes(r) == 4 * (r.nx * r.ny + r.nt) + 4 * 7 * r.nt * r.nx


def test_stream_generator_is_seeded():
    a = make_stream(5, 2.0, seed=1, grids=((32, 64),), nt_range=(50, 60), shots_range=(1, 3), device="cpu")
    b = make_stream(5, 2.0, seed=1, grids=((32, 64),), nt_range=(50, 60), shots_range=(1, 3), device="cpu")
    assert [(x.kind, x.nt, x.nshots, x.arrival) for x in a] == [(x.kind, x.nt, x.nshots, x.arrival) for x in b]
    assert all(torch.equal(x.vel, y.vel) for x, y in zip(a, b))
    arr = [x.arrival for x in a]
    assert arr == sorted(arr) and arr[0] > 0


def test_chunk_time_block_by_kind_storage_and_size():
    from torchwi_amd.shots import AUTO_BLOCK_CELLS
    big = _req("gradient", nx=3600, ny=1200, nt=6000, S=12)
    cells = (big.nx + 8) * (big.ny + 8)
    n = -(-AUTO_BLOCK_CELLS // cells)
    assert chunk_time_block(big, n, 78) == 3 and chunk_time_block(big, n - 1, 78) == 1
    assert chunk_time_block(big, n, 0) == 3            # full storage
    assert chunk_time_block(big, n, 77) == 1           # not a multiple of the block
    from torchwi_amd.shots import AUTO_BLOCK_CELLS_FORWARD
    fwd = _req("forward", nx=3600, ny=1200, S=12)
    nf = -(-AUTO_BLOCK_CELLS_FORWARD // cells)
    assert chunk_time_block(fwd, nf, 0) == 3 and chunk_time_block(fwd, nf - 1, 0) == 1
    assert nf > n                                       # modelling keeps one step per launch longer
    assert chunk_plan(big, 10 ** 10)[1] % 3 == 0


def test_admission_fcfs_with_bounded_skip_ahead():
    from torchwi_amd.serving import select_admissions
    # the oldest (cost 8) does not fit 6 free bytes: the next two may overtake it
    assert select_admissions([(0, 8, 0), (0, 3, 0), (0, 3, 0)], 4, 6, 1.0, 4) == ([1, 2], [0, 0])
    # ... but not once it has been overtaken max_skips times
    assert select_admissions([(0, 8, 3), (0, 3, 0), (0, 3, 0)], 4, 6, 1.0, 4) == ([1], [0])
    assert select_admissions([(0, 8, 4), (0, 3, 0)], 4, 6, 1.0, 4) == ([], [])
    # slots and arrival times bound the admissions; order is kept
    assert select_admissions([(0, 1, 0), (0, 1, 0), (0, 1, 0)], 2, 10, 1.0, 4) == ([0, 1], [])
    assert select_admissions([(0, 1, 0), (5, 1, 0)], 4, 10, 1.0, 4) == ([0], [])
    assert select_admissions([], 4, 10, 1.0, 4) == ([], [])


def test_batch_key_separates_what_cannot_share_a_plan():
    from torchwi_amd.serving import batch_key
    a, b = _req("gradient"), _req("gradient")
    assert batch_key(a, 0) == batch_key(b, 0)
    assert batch_key(a, 0) != batch_key(a, 30)                       # storage mode
    assert batch_key(a, 0) != batch_key(_req("forward"), 0)          # kind
    assert batch_key(a, 0) != batch_key(_req("gradient", nt=401), 0)  # time axis
    assert batch_key(a, 0) != batch_key(_req("gradient", ny=81), 0)   # grid
