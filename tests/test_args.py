"""CPU checks of the Python binding's argument validation: every device array of an
assembly or a served request is checked for dtype, length and device before any
pointer reaches the C ABI (a short or narrower array would be read or written past
its end by the kernels)."""
import types

import pytest
import torch

import tianxin_amd as txa

MESH = types.SimpleNamespace(nelem=8, ndofs=27, nnz=343)


def args(**over):
    kw = dict(coef=torch.ones(8, dtype=torch.float64), x=torch.zeros(27, dtype=torch.float64),
              f=torch.zeros(27, dtype=torch.float64), vals=torch.zeros(343, dtype=torch.float64))
    kw.update(over)
    return kw


def reason(**over):
    with pytest.raises(ValueError) as e:
        txa.check_assembly_tensors(MESH, 0, **args(**over))
    return str(e.value)


def test_dtype_is_checked_first():
    assert "float64" in reason(f=torch.zeros(27, dtype=torch.float32))
    assert "float64" in reason(vals=torch.zeros(343, dtype=torch.float32))
    assert "float64" in reason(coef=torch.ones(8, dtype=torch.float32))
    assert "int32" in reason(bc_lids=torch.zeros(3, dtype=torch.int64))
    assert "float64" in reason(bc_lids=torch.zeros(3, dtype=torch.int32), bc_vals=torch.zeros(3, dtype=torch.float32))


def test_lengths_are_checked():
    assert "27 entries" in reason(x=torch.zeros(26, dtype=torch.float64))
    assert "27 entries" in reason(f=torch.zeros(28, dtype=torch.float64))
    assert "27 entries" in reason(xdot=torch.zeros(3, dtype=torch.float64))
    assert "343 entries" in reason(vals=torch.zeros(342, dtype=torch.float64))
    assert "8 entries" in reason(coef=torch.ones(48, dtype=torch.float64))
    assert "3 entries" in reason(bc_lids=torch.zeros(3, dtype=torch.int32), bc_vals=torch.zeros(4, dtype=torch.float64))


def test_required_and_device():
    assert "required" in reason(x=None)
    assert "required" in reason(coef=None)
    # well-formed host tensors are refused as not being device memory
    assert "device tensor" in reason()
    assert "without bc_lids" in reason(bc_vals=torch.zeros(2, dtype=torch.float64))


def test_serve_struct_carries_the_producer_stream():
    names = [n for n, _ in txa.ServeRequest._fields_]
    assert names[-1] == "stream"
