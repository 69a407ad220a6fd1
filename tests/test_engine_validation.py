"""Host-side refusal of mismatched buffers before any pointer reaches a
kernel (shape, dtype, contiguity, device), on CPU tensors."""
import pytest
import torch

from ad_afqmc_amd.engine import check_tensor, walker_shapes


def test_walker_shapes():
    assert walker_shapes(10, (3, 2), "restricted", 7) == ((7, 10, 3), None)
    assert walker_shapes(10, (3, 2), "unrestricted", 7) == ((7, 10, 3), (7, 10, 2))


@pytest.mark.parametrize(
    "t,shape,dtype,msg",
    [
        (torch.zeros(4, 10, 3, dtype=torch.complex64), (4, 10, 3), torch.complex128, "dtype"),
        (torch.zeros(4, 10, 2, dtype=torch.complex128), (4, 10, 3), torch.complex128, "shape"),
        (torch.zeros(4, 3, 10, dtype=torch.complex128).transpose(1, 2), (4, 10, 3), torch.complex128, "contiguous"),
        (torch.zeros(4, 10, 3, dtype=torch.complex128), (4, 10, 3), torch.complex128, "device tensor"),
        (torch.zeros(5, 7, dtype=torch.float64), (5, 8), torch.float64, "shape"),
        (torch.zeros(5, dtype=torch.float32), (5,), torch.float64, "dtype"),
        ([0.0] * 5, (5,), torch.float64, "torch.Tensor"),
    ],
)
def test_check_tensor_refuses(t, shape, dtype, msg):
    with pytest.raises(ValueError, match=msg):
        check_tensor("x", t, shape, dtype)
