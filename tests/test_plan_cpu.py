"""Host-side plan choices (CPU): kernel selection by size, checkpoint interval."""
from torchwi_amd.batch import _resolve_time_block
from torchwi_amd.shots import AUTO_BLOCK_CELLS, auto_time_block, default_ckpt_interval


def test_auto_time_block_by_plan_size():
    assert auto_time_block(32 * 1032 * 392) == 1            # 13M cells: one step per launch
    assert auto_time_block(64 * 3008 * 1008) == 3           # 194M cells: blocked kernel
    assert auto_time_block(AUTO_BLOCK_CELLS - 1) == 1 and auto_time_block(AUTO_BLOCK_CELLS) == 3
    assert auto_time_block(AUTO_BLOCK_CELLS, forward_only=True) == 1             # modelling: one step longer
    assert auto_time_block(64 * 3008 * 1008, forward_only=True) == 3


def test_resolve_time_block_respects_explicit_choices():
    big = 2 * AUTO_BLOCK_CELLS
    assert _resolve_time_block("auto", big, "auto") == 3
    assert _resolve_time_block("auto", big, 0) == 3          # full storage
    assert _resolve_time_block("auto", big, 90) == 3         # 90 is a multiple of 3
    assert _resolve_time_block("auto", big, 64) == 1         # 64 is not: keep one step per launch
    assert _resolve_time_block(2, big, 64) == 2 and _resolve_time_block(1, big, "auto") == 1


def test_default_ckpt_interval_is_a_multiple_of_the_block():
    for nt in (100, 4000, 8000, 6001):
        for m in (1, 2, 3, 4):
            k = default_ckpt_interval(nt, m)
            assert k % m == 0 and k >= int(nt ** 0.5)
