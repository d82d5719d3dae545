"""The two calibration kernels the benchmark quotes its rooflines against: the streaming copy
(achievable HBM bandwidth, counter scaling) and the VALU issue probe (the traversal's own bound)."""
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _lib():
    from torch_kdtree_amd import torch_knn
    return torch_knn


def test_stream_copy_copies_exactly():
    torch_knn = _lib()
    dev = torch.device("cuda", 0)
    for n in (4, 1000, (1 << 20) + 12):
        src = torch.randint(-2**31, 2**31 - 1, (n,), dtype=torch.int32, device=dev)
        dst = torch.zeros_like(src)
        torch_knn.copy_gpu(dst.data_ptr(), src.data_ptr(), n * 4, torch.cuda.current_stream(dev).cuda_stream)
        torch.cuda.synchronize()
        assert torch.equal(src, dst)


def test_stream_copy_rejects_unaligned():
    torch_knn = _lib()
    dev = torch.device("cuda", 0)
    src = torch.zeros(64, dtype=torch.int32, device=dev)
    dst = torch.zeros_like(src)
    with pytest.raises(RuntimeError):
        torch_knn.copy_gpu(dst.data_ptr() + 4, src.data_ptr(), 64, 0)
    with pytest.raises(RuntimeError):
        torch_knn.copy_gpu(dst.data_ptr(), src.data_ptr(), 20, 0)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_valu_probe_counts_and_values(mode):
    torch_knn = _lib()
    dev = torch.device("cuda", 0)
    threads, iters = 256 * 8, 64
    out = torch.full((threads,), float("nan"), dtype=torch.float32, device=dev)
    n = torch_knn.valu_probe(out.data_ptr(), threads, iters, mode, torch.cuda.current_stream(dev).cuda_stream)
    torch.cuda.synchronize()
    assert n == threads // 64 * iters * (24 if mode == 2 else 16)
    # the same chains on the host, in float32 (no FMA: the library is built with -ffp-contract=off)
    import numpy as np
    a, b = np.float32(0.999), np.float32(0.001)
    lane = np.arange(threads, dtype=np.int64) % 256
    x = [lane.astype(np.float32) * np.float32(0.001) + np.float32(i) for i in range(8)]
    for _ in range(iters):
        for i in range(8):
            if mode == 0:
                x[i] = (x[i] * a).astype(np.float32)
                x[i] = (x[i] + b).astype(np.float32)
            elif mode == 2:
                x[i] = (x[i] + np.where(x[i] < a, b, -b)).astype(np.float32)
            else:
                x[i] = np.clip(x[i], min(a, b), max(a, b))  # med3(x, a, b) with a, b finite
                x[i] = np.clip(x[i], min(a, b), max(a, b))
    s = np.zeros(threads, dtype=np.float32)
    for i in range(8):
        s = (s + x[i]).astype(np.float32)
    assert np.array_equal(out.cpu().numpy(), s)


def test_valu_probe_rejects_bad_shapes():
    torch_knn = _lib()
    dev = torch.device("cuda", 0)
    out = torch.zeros(512, dtype=torch.float32, device=dev)
    with pytest.raises(RuntimeError):
        torch_knn.valu_probe(out.data_ptr(), 300, 4, 0, 0)
    with pytest.raises(RuntimeError):
        torch_knn.valu_probe(out.data_ptr(), 256, 4, 3, 0)
