"""lbm_rules_apply_noalloc, the rule pre-pass that allocates nothing (its rule values go
through the library's own per-GPU device buffer), against lbm_rules_apply with a caller
scratch buffer: the same bits, on consecutive calls of shrinking size, through the C ABI."""

import numpy as np
import pytest

from lbm.backends._native import DTYPE_CODES, library

MAX_RULES = 2**20


def rule_set(rng, n_rules, n_slots, m, dtype):
    """Random rules with distinct targets, term-major (m x n) slots and coefficients."""
    target = rng.choice(n_slots, size=n_rules, replace=False).astype(np.int64)
    slot = rng.integers(0, n_slots, size=(m, n_rules)).astype(np.int64)
    coeff = (rng.standard_normal((m, n_rules)) * (rng.random((m, n_rules)) < 0.8)).astype(dtype)
    constant = rng.standard_normal(n_rules).astype(dtype)
    return target, slot, coeff, constant


def test_sizes_beyond_the_library_scratch_are_refused_before_any_gpu_work():
    lib = library()
    with pytest.raises(ValueError, match="at most 1048576 rules"):
        lib.call("lbm_rules_apply_noalloc", 0, 1, 1, 1, 1, 1, MAX_RULES + 1, 4, None)
    with pytest.raises(ValueError, match="1 <= m <= 128"):
        lib.call("lbm_rules_apply_noalloc", 0, 1, 1, 1, 1, 1, 10, 0, None)
    lib.call("lbm_rules_apply_noalloc", 0, None, None, None, None, None, 0, 4, None)  # empty: nothing to do


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_noalloc_pre_pass_equals_the_scratch_pre_pass_on_consecutive_calls(dtype):
    import torch

    lib = library()
    rng = np.random.default_rng(5)
    n_slots = 9 * 400_000
    base = torch.as_tensor(rng.standard_normal(n_slots).astype(dtype), device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    # two calls in a row, the second smaller (both through the library scratch), then one
    # small enough for the one-workgroup path, then the largest set the scratch holds
    for n_rules, m in ((300_000, 6), (5_000, 3), (1_000, 4), (MAX_RULES, 2)):
        target, slot, coeff, constant = (torch.as_tensor(a, device="cuda") for a in rule_set(
            rng, n_rules, n_slots, m, dtype))
        a, b = base.clone(), base.clone()
        scratch = torch.empty(n_rules, dtype=a.dtype, device="cuda")
        lib.call("lbm_rules_apply", DTYPE_CODES[dtype], a.data_ptr(), slot.data_ptr(), coeff.data_ptr(),
                 constant.data_ptr(), target.data_ptr(), n_rules, m, scratch.data_ptr(), stream)
        lib.call("lbm_rules_apply_noalloc", DTYPE_CODES[dtype], b.data_ptr(), slot.data_ptr(), coeff.data_ptr(),
                 constant.data_ptr(), target.data_ptr(), n_rules, m, stream)
        torch.cuda.synchronize()
        assert torch.equal(a, b), f"{n_rules} rules"
        assert not torch.equal(a, base)
        base = b  # the next call starts from this state
