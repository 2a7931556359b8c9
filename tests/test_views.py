"""Host-side checks of device views before their raw pointers reach the C
ABI (qg.check_views): a mismatch would make the kernels read or write
outside the caller's tensors, so it must raise.  CPU tensors stand in for
device ones (require_cuda=False); no GPU call is made."""
import pytest
import torch

from pequod_amd import qg


def _v(L=3, g=17, ld=19, dtype=torch.float32):
    return torch.zeros((L, ld, ld), dtype=dtype)


def test_views_accepted():
    qg.check_views([_v(), _v()], 4, 17, 3, "zeta", require_cuda=False)
    # an sdata view of a (G+2)^2 node field, offset to element (1, 1)
    node = torch.zeros((3, 19, 19), dtype=torch.float64)
    qg.check_views([node[:, 1:, 1:]], 8, 17, 3, "psi", require_cuda=False)


@pytest.mark.parametrize("case", ["dtype", "layers", "small", "inner_stride", "batch_strides",
                                  "dims", "empty"])
def test_views_rejected(case):
    views = {
        "dtype": ([_v(dtype=torch.float64)], 4),
        "layers": ([_v(L=2)], 4),
        "small": ([_v(ld=16)], 4),
        "inner_stride": ([torch.zeros((3, 19, 38))[:, :, ::2]], 4),
        "batch_strides": ([_v(), _v(ld=21)], 4),
        "dims": ([torch.zeros((19, 19))], 4),
        "empty": ([], 4),
    }[case]
    with pytest.raises(ValueError):
        qg.check_views(views[0], views[1], 17, 3, "zeta", require_cuda=False)


def test_views_need_device():
    with pytest.raises(ValueError, match="device"):
        qg.check_views([_v()], 4, 17, 3, "zeta")
