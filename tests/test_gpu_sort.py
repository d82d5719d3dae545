"""GPU radix sort of (key, value) pairs: stable, by the low key_bits bits, against numpy."""
import numpy as np
import pytest
import torch

from torch_kdtree_amd import torch_knn

pytestmark = pytest.mark.gpu


def _sort(keys, vals, key_bits, device):
    k = torch.from_numpy(keys.view(np.int32)).to(device)
    v = torch.from_numpy(vals.view(np.int32)).to(device)
    ko = torch.empty_like(k)
    vo = torch.empty_like(v)
    torch_knn.sort_pairs(k.data_ptr(), v.data_ptr(), ko.data_ptr(), vo.data_ptr(), k.numel(), key_bits,
                         torch.cuda.current_stream(device).cuda_stream)
    torch.cuda.synchronize()
    return ko.cpu().numpy().view(np.uint32), vo.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("n", [1, 100, 4095, 4096, 4097, 100_003, (1 << 20) + 17])
@pytest.mark.parametrize("key_bits", [1, 8, 9, 21, 32])
def test_sort_pairs_is_a_stable_sort(cuda, n, key_bits):
    rng = np.random.default_rng(n * 33 + key_bits)
    # few distinct low-bit values relative to n, so that equal keys are common
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    keys &= np.uint32((1 << min(key_bits, 12)) - 1) | np.uint32(0xF0000000)
    vals = np.arange(n, dtype=np.uint32)
    mask = np.uint32(0xFFFFFFFF) if key_bits == 32 else np.uint32((1 << key_bits) - 1)
    perm = np.argsort(keys & mask, kind="stable")
    ko, vo = _sort(keys, vals, key_bits, cuda)
    np.testing.assert_array_equal(vo, vals[perm])
    np.testing.assert_array_equal(ko, keys[perm])


def test_sort_pairs_repeatable_and_skewed(cuda):
    # one dominant digit value in every pass
    rng = np.random.default_rng(7)
    n = 300_000
    keys = np.where(rng.random(n) < 0.9, 12345, rng.integers(0, 1 << 25, n)).astype(np.uint32)
    vals = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    a = _sort(keys, vals, 25, cuda)
    b = _sort(keys, vals, 25, cuda)
    perm = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(a[1], vals[perm])
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
