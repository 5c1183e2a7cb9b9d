"""The batch wrappers refuse arrays the kernels would misread (host-side checks,
no GPU call): wrong dtype width or kind, wrong row count or row width."""
import pytest
import torch

from mupsi_amd import batch


def test_check_accepts_matching_arrays():
    batch._check(torch.zeros((5, 3), dtype=torch.float32), "f32", 5, 3, None, "p")
    batch._check(torch.zeros(5, dtype=torch.int32), "i32", 5, None, None, "seed")
    batch._check(torch.zeros(5, dtype=torch.uint8), "u8", 5, None, None, "active")


@pytest.mark.parametrize("t,kind,n,cols", [
    (torch.zeros(5, dtype=torch.uint8), "i32", 5, None),      # 1-byte seeds read as 4-byte
    (torch.zeros(5, dtype=torch.int64), "i32", 5, None),
    (torch.zeros((5, 3), dtype=torch.float64), "f32", 5, 3),  # doubles read as floats
    (torch.zeros((5, 3), dtype=torch.float16), "f32", 5, 3),
    (torch.zeros(5, dtype=torch.float32), "u8", 5, None),
])
def test_check_refuses_dtype(t, kind, n, cols):
    with pytest.raises(TypeError):
        batch._check(t, kind, n, cols, None, "x")


@pytest.mark.parametrize("shape,n,cols", [((4,), 5, None), ((5, 2), 5, 3), ((6, 8), 5, 8),
                                          ((5,), 5, 3)])
def test_check_refuses_shape(shape, n, cols):
    with pytest.raises(ValueError):
        batch._check(torch.zeros(shape, dtype=torch.float32), "f32", n, cols, None, "x")


def test_march_refuses_short_seed_before_any_device_call():
    rays = torch.zeros((8, 8), dtype=torch.float32)
    with pytest.raises(ValueError):
        batch.march(None, rays, torch.zeros(7, dtype=torch.int32))
    with pytest.raises(TypeError):
        batch.march(None, rays, torch.zeros(8, dtype=torch.uint8))
