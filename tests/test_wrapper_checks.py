"""The Python wrappers' buffer checks (cola_amd/spmm.py), on host stand-ins
for GPU tensors: every check raises ValueError before anything reaches the
library, so no GPU is needed."""
import pytest

from cola_amd import spmm


class FakeTensor:
    """Shape, strides, element size and device flag of a tensor."""

    def __init__(self, shape, elem=4, cuda=True, strides=None):
        self.shape = tuple(shape)
        self._elem = elem
        self.is_cuda = cuda
        self._strides = strides or ((shape[1], 1) if len(shape) == 2 else (1,))

    def element_size(self):
        return self._elem

    def dim(self):
        return len(self.shape)

    def stride(self, i):
        return self._strides[i]

    def numel(self):
        n = 1
        for s in self.shape:
            n *= s
        return n


def test_matrix_check_accepts_large_enough_row_major_buffers():
    spmm._check_matrix(FakeTensor((10, 8)), "C", 10, 8, 4)
    spmm._check_matrix(FakeTensor((12, 16), strides=(32, 1)), "C", 10, 8, 4)  # padded rows
    spmm._check_matrix(None, "packed", 0, 8, 4)  # nothing to read


@pytest.mark.parametrize("t, rows, dim, elem", [
    (FakeTensor((9, 8)), 10, 8, 4),                   # too few rows
    (FakeTensor((10, 7)), 10, 8, 4),                  # too few columns
    (FakeTensor((10, 8), elem=2), 10, 8, 4),          # element size
    (FakeTensor((8, 10), strides=(1, 8)), 10, 8, 4),  # column-major
    (FakeTensor((80,)), 10, 8, 4),                    # not 2-D
    (FakeTensor((10, 8), cuda=False), 10, 8, 4),      # host memory
    (None, 1, 8, 4),                                  # missing
])
def test_matrix_check_rejects(t, rows, dim, elem):
    with pytest.raises(ValueError):
        spmm._check_matrix(t, "B", rows, dim, elem)


def test_vector_check():
    spmm._check_vector(FakeTensor((5,)), "scales", 5)
    spmm._check_vector(None, "scales", 0)
    for bad in (FakeTensor((4,)), FakeTensor((5,), elem=2), FakeTensor((5,), cuda=False), None):
        with pytest.raises(ValueError):
            spmm._check_vector(bad, "scales", 5)


def _server_stub(shapes):
    """A Server whose residents have these (rows, dim) outputs, without a
    library handle: run_trace's checks run before any library call."""
    s = object.__new__(spmm.Server)
    s._out_shape = dict(shapes)
    s._h = None
    return s


def test_server_run_trace_checks_outputs():
    s = _server_stub({0: (10, 8), 1: (4, 16)})
    arr, res = [0.0, 1.0], [0, 1]
    for outs in ([FakeTensor((9, 8)), None],                  # too few rows
                 [FakeTensor((10, 8)), FakeTensor((4, 8))],   # too few columns
                 [FakeTensor((10, 8), strides=(16, 1)), None],  # row stride != dim
                 [FakeTensor((10, 8), elem=2), None],         # element size
                 [FakeTensor((10, 8))]):                      # one output per request
        with pytest.raises(ValueError):
            s.run_trace(arr, res, outs)
    with pytest.raises(ValueError):
        _server_stub({0: (10, 8)}).run_trace([0.0], [3], [FakeTensor((10, 8))])  # unknown resident
