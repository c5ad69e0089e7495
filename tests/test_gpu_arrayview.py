"""The array-level view API, TransposeArrayInPlace (src/py/transpose.py:108-116), on the cases
of the reference's own Python tests (src/py/transpose_unittest.py:38-94, TestArray and
TestFullTranspose): the data is transposed in the tensor's own storage and the result is that
storage viewed with shape (cols, rows); the input tensor keeps its shape; tile factors that do
not divide the dimensions, and arrays that are not 2-D, raise TypeError."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_array_in_place_view(hip, orc):
    import torch
    A, a, B, b = 40, 16, 37, 10  # TestArray, src/py/transpose_unittest.py:82-94
    x = orc.tagged(A * a * B * b, 4, seed=82).view(np.float32).reshape(A * a, B * b)
    d = torch.from_numpy(x.copy()).cuda()
    ptr, storage = d.data_ptr(), d.untyped_storage().data_ptr()
    t = hip.TransposeArrayInPlace(d, a, b)
    torch.cuda.synchronize()
    assert t.shape == (B * b, A * a)
    assert d.shape == (A * a, B * b)  # the input keeps its shape
    assert t.data_ptr() == ptr and t.untyped_storage().data_ptr() == storage
    assert t.is_contiguous()
    want = orc.op_transpose(x.reshape(-1).view(np.uint8), A * a, B * b, es=4)
    assert np.array_equal(t.cpu().numpy().reshape(-1).view(np.uint8), want)
    assert hip.last_route().endswith("3.1-")
    # the view shares storage with the input: writing through one shows in the other
    t[0, 0] = 12345.0
    torch.cuda.synchronize()
    assert d.view(-1)[0].item() == 12345.0


@pytest.mark.parametrize("a", [4, 8, 16, 32, 64])
def test_array_full_transpose_tiles(hip, orc, a):
    """TestFullTranspose (src/py/transpose_unittest.py:38-56): A = 40, B = 37, b = 30."""
    import torch
    A, B, b = 40, 37, 30
    x = orc.tagged(A * a * B * b, 4, seed=a).view(np.float32).reshape(A * a, B * b)
    d = torch.from_numpy(x.copy()).cuda()
    t = hip.TransposeArrayInPlace(d, a, b)
    assert np.array_equal(t.cpu().numpy(), x.T)
    # and back, with the planner's tile factors (a = b = 1)
    u = hip.TransposeArrayInPlace(t)
    assert u.shape == x.shape
    assert np.array_equal(u.cpu().numpy(), x)


def test_array_errors(hip):
    import torch
    d = torch.zeros(60, 35, device="cuda")
    with pytest.raises(TypeError):
        hip.TransposeArrayInPlace(d, 7, 5)  # 7 does not divide 60
    with pytest.raises(TypeError):
        hip.TransposeArrayInPlace(d, 6, 4)  # 4 does not divide 35
    with pytest.raises(TypeError):
        hip.TransposeArrayInPlace(torch.zeros(4, 5, 6, device="cuda"))
    assert d.shape == (60, 35)
