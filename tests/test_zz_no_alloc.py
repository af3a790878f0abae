"""The allocation-free batched front factorization
(sylver_amd_batched_front_potrf_noalloc) against the allocating entry point
(sylver_amd_batched_front_potrf): both run the same per-front kernel, so the
factors and info must be bitwise equal. Two calls in a row on the shared
per-device array, the second with fewer and smaller fronts, so stale
contents of the first call would show."""
import numpy as np
import pytest

from sylver_amd import matrices as M


def _fronts(nf, seed, dtype):
    nrow, ncol = M.dense_front_sizes(nf, seed=seed)
    offs = np.zeros(nf, np.int64)
    tot = 0
    parts = []
    for f in range(nf):
        offs[f] = tot
        parts.append(M.dense_front_values(int(nrow[f]), seed=1000 * seed + f).ravel(order="F"))
        tot += int(nrow[f]) ** 2
    return nrow, ncol, offs, np.concatenate(parts).astype(dtype)


def _run_pair(S, nrow, ncol, offs, host, dtype):
    ref = S.DeviceArray(host)
    rc, info_ref, _ = S.batched_front_potrf(ref, nrow, ncol, offs, dtype)
    assert rc == 0
    got = S.DeviceArray(host)
    # info starts non-zero: the entry point must clear it itself
    info = S.DeviceArray(np.full(len(nrow), 7, np.int32))
    rc = S.batched_front_potrf_noalloc(got, nrow, ncol, offs, dtype, info)
    assert rc == 0
    out_ref, out = ref.numpy(), got.numpy()
    inf = info.numpy()
    for a in (ref, got, info):
        a.free()
    assert np.array_equal(inf, info_ref)
    assert np.array_equal(out.view(np.uint8), out_ref.view(np.uint8))
    return inf


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_noalloc_matches_allocating_entry_point(S, gpu, dtype):
    nrow, ncol, offs, host = _fronts(40, seed=3, dtype=dtype)
    assert int(np.abs(_run_pair(S, nrow, ncol, offs, host, dtype)).sum()) == 0
    # second call: fewer, smaller fronts, one of them not positive definite
    nrow2, ncol2 = np.minimum(nrow[:9], 64).astype(np.int32), np.minimum(ncol[:9], 32).astype(np.int32)
    offs2 = np.concatenate([[0], np.cumsum(nrow2.astype(np.int64) ** 2)[:-1]]).astype(np.int64)
    parts = [M.dense_front_values(int(m), seed=77 + f).ravel(order="F") for f, m in enumerate(nrow2)]
    host2 = np.concatenate(parts).astype(dtype)
    m4 = int(nrow2[4])
    host2[offs2[4]:offs2[4] + m4 * m4] *= -1.0
    inf2 = _run_pair(S, nrow2, ncol2, offs2, host2, dtype)
    assert inf2[4] == 1 and int(np.abs(np.delete(inf2, 4)).sum()) == 0


@pytest.mark.gpu
def test_noalloc_rejects_too_many_fronts(S, gpu):
    n = (1 << 16) + 1
    nrow = np.ones(n, np.int32)
    offs = np.arange(n, dtype=np.int64)
    data = S.DeviceArray(np.ones(n, np.float64))
    info = S.DeviceArray(np.zeros(n, np.int32))
    try:
        rc = S.batched_front_potrf_noalloc(data, nrow, nrow, offs, "float64", info)
    finally:
        data.free()
        info.free()
    assert rc == S.ERROR_CALL_SEQUENCE
