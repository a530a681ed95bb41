"""The bench's atomic-throughput probe does the operations it reports."""
import pytest
import torch

from deepcharacters_amd import _lib

pytestmark = pytest.mark.gpu


def test_probe_add_counts_every_op():
    n_addr, n_ops = 1 << 12, 1 << 18
    buf = torch.zeros(n_addr, dtype=torch.float32, device="cuda")
    _lib.call("dcx_probe_atomics", 0, buf.data_ptr(), n_addr, n_ops, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert float(buf.double().sum()) == n_ops  # small integer counts: exact in float
    assert (buf > 0).float().mean() > 0.99      # the hash spreads the ops over the addresses


def test_probe_min_keeps_the_smallest_value():
    n_addr, n_ops = 1 << 12, 1 << 18
    buf = torch.full((n_addr,), 2**62, dtype=torch.int64, device="cuda")
    _lib.call("dcx_probe_atomics", 1, buf.data_ptr(), n_addr, n_ops, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    touched = buf < 2**62
    assert touched.float().mean() > 0.99
    assert int(buf[touched].min()) >= 1 and int(buf[touched].max()) <= n_ops


def test_probe_rejects_sizes_that_are_not_powers_of_two():
    buf = torch.zeros(10, dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.DcxError):
        _lib.call("dcx_probe_atomics", 0, buf.data_ptr(), 10, 100, torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("nontemporal", [0, 1])
def test_probe_copy_copies_every_byte(nontemporal):
    n = (1 << 20) + 4 * 37  # not a multiple of the grid's stride
    src = torch.randint(-2**31, 2**31 - 1, (n,), dtype=torch.int32, device="cuda")
    dst = torch.zeros_like(src)
    _lib.call("dcx_probe_copy", src.data_ptr(), dst.data_ptr(), n * 4, nontemporal, 64,
              torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert torch.equal(src, dst)


def test_probe_copy_rejects_unaligned_sizes():
    buf = torch.zeros(64, dtype=torch.float32, device="cuda")
    with pytest.raises(_lib.DcxError):
        _lib.call("dcx_probe_copy", buf.data_ptr(), buf.data_ptr(), 10, 0, 0, torch.cuda.current_stream().cuda_stream)
