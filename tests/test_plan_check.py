"""The device launch plan's host side, checked on the CPU (no device).

The plan builder lays out the fronts in the arena and produces every
launch's task list; a bad index there would be a device memory fault. The
library's own checker (sylver_amd_plan_check) rebuilds the plan without a
device and verifies every index each launch will use against the buffer it
addresses. Covered: the configurations of the parity tests and the
benchmark's 80^3 headline case, with the knobs that change the plan
(one-workgroup cutoff, per-front variants, wide-front threshold, the
compute-unit count that decides which contribution blocks ride in the
chained solve launches, the strip TRSM and the per-level assembly on and off).
"""
import numpy as np
import pytest

from sylver_amd import matrices as M


def _akeep(S, A):
    opts = S.default_options()
    akeep, inf, _ = S.spldlt_analyse(A.n, A.ptr, A.row, A.val, options=opts)
    assert inf.flag >= 0
    return akeep


CASES = {
    "lap2d_100": lambda: M.laplacian_2d(100),
    "lap3d_16": lambda: M.laplacian_3d(16),
    "lap3d_24": lambda: M.laplacian_3d(24),
    "lap3d_48": lambda: M.laplacian_3d(48),
    "kkt_8": lambda: M.kkt(8, 100, seed=3),
    "simple": lambda: M.CSCLower(3, [1, 3, 5, 6], [1, 2, 2, 3, 3], [2.0, -1.0, 2.0, -1.0, 2.0]),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("small_max,bf16,cus", [(0, 0, 256), (48, 64, 256), (48, 0, 1), (1024, 0, 8)])
@pytest.mark.parametrize("strip_trsm,level_asm", [(1, 1), (0, 0), (1, 0), (0, 1)])
def test_plan_indices_in_bounds(S, case, small_max, bf16, cus, strip_trsm, level_asm):
    A = CASES[case]()
    akeep = _akeep(S, A)
    S.set_ext_options(strip_trsm=strip_trsm, level_assembly=level_asm)
    try:
        st, why = S.plan_check(akeep, small_front_max=small_max, bf16_min_ncol=bf16, num_cus=cus)
    finally:
        S.set_ext_options(strip_trsm=1, level_assembly=1)
    assert st["violations"] == 0, why
    assert st["factor_launches"] > 0
    if not strip_trsm:
        assert st["strip_tasks"] == 0
    S.spldlt_free_akeep(akeep)


@pytest.mark.parametrize("variant", [0, 1])
def test_plan_indices_with_front_variants(S, variant):
    A = M.laplacian_3d(24)
    akeep = _akeep(S, A)
    sym = S.akeep_symbolic(akeep)
    S.set_front_variants(akeep, np.full(sym["nnodes"], variant, np.int8))
    st, why = S.plan_check(akeep)
    assert st["violations"] == 0, why
    S.spldlt_free_akeep(akeep)


def test_plan_indices_headline_80cubed(S):
    """The benchmark's 3D 7-point 80^3 case (n = 512,000): 7,580-row root,
    arena near 2^31 elements, 60-block chains."""
    A = M.laplacian_3d(80)
    akeep = _akeep(S, A)
    st, why = S.plan_check(akeep)
    assert st["violations"] == 0, why
    assert st["arena_elems"] > 1_000_000_000
    # the top levels' large fronts take the strip TRSM (64-row strips below
    # each diagonal block)
    assert st["strip_tasks"] > 5_000
    S.spldlt_free_akeep(akeep)
