"""GPU: error behaviour of the C ABI (status codes, messages, no partial writes),
the counterpart of the reference's TEUCHOS_TEST_FOR_EXCEPTION checks."""
import ctypes

import numpy as np
import pytest
import torch

import tianxin_amd as txa
from tianxin_amd import synth

pytestmark = pytest.mark.gpu


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=dtype)


def code_of(fn):
    with pytest.raises(txa.TxaError) as e:
        fn()
    assert str(e.value)        # a message comes with every code
    return e.value.code


def brick(engine, topo, n, nf, graph=True):
    m = engine.brick(topo, n, n, n, synth.brick_coords(n, n, n))
    m.build_dofmap(nf)
    if graph:
        m.build_graph()
    return m


def test_bad_arguments_and_state(engine):
    lib = engine.lib
    mid = ctypes.c_int()
    X = synth.brick_coords(2, 2, 2)
    assert lib.txa_mesh_brick(engine.h, txa.HEX8, 0, 2, 2, X.ctypes.data_as(ctypes.c_void_p),
                              ctypes.byref(mid)) == txa.TXA_ERR_INVALID
    assert lib.txa_mesh_brick(engine.h, 99, 2, 2, 2, X.ctypes.data_as(ctypes.c_void_p),
                              ctypes.byref(mid)) == txa.TXA_ERR_INVALID
    assert b"topology" in lib.txa_last_error(engine.h)
    assert lib.txa_mesh_destroy(engine.h, 424242) == txa.TXA_ERR_INVALID
    h = ctypes.c_void_p()
    assert lib.txa_create(1000, ctypes.byref(h)) == txa.TXA_ERR_INVALID

    m = brick(engine, txa.HEX8, 3, 1, graph=False)
    x = dev(np.zeros(m.ndofs))
    k = dev(np.ones(m.nelem))
    f = torch.empty(m.ndofs, device="cuda", dtype=torch.float64)
    assert code_of(lambda: m.assemble(physics=txa.DIFFUSION, coef=k, x=x, f=f)) == txa.TXA_ERR_LOGIC
    m.build_graph()
    assert code_of(lambda: m.assemble(physics=txa.ELASTICITY, coef=k, x=x, f=f)) == txa.TXA_ERR_INVALID
    assert code_of(lambda: m.assemble(physics=txa.DIFFUSION, coef=k, x=x, f=f, transient=True)) == txa.TXA_ERR_INVALID
    assert code_of(lambda: m.assemble(physics=txa.DIFFUSION, coef=k, x=x, f=f,
                                      precision=txa.SPLIT_BF16)) == txa.TXA_ERR_UNSUPPORTED
    m.destroy()


def test_unsupported_precision_on_the_generic_path(engine):
    m = brick(engine, txa.HEX27, 2, 1)
    x = dev(np.zeros(m.ndofs))
    f = torch.empty(m.ndofs, device="cuda", dtype=torch.float64)
    assert code_of(lambda: m.assemble(physics=txa.DIFFUSION, coef=dev(np.ones(m.nelem)), x=x, f=f,
                                      precision=txa.FP32)) == txa.TXA_ERR_UNSUPPORTED
    m.destroy()


def test_dirichlet_lids_outside_the_map_are_reported(engine):
    # a LID outside the map touches no row; the next txa_sync reports it (TXA_ERR_INVALID,
    # std::invalid_argument in the reference) and clears the report
    engine.sync()
    m = brick(engine, txa.HEX8, 4, 1)
    x = dev(synth.uniform(1, m.ndofs, -1, 1))
    k = dev(synth.lognormal(2, m.nelem, 0.5))
    f0 = torch.empty(m.ndofs, device="cuda", dtype=torch.float64)
    v0 = torch.empty(m.nnz, device="cuda", dtype=torch.float64)
    m.assemble(physics=txa.DIFFUSION, coef=k, x=x, f=f0, vals=v0, src=(1.0,))
    engine.sync()
    f1, v1 = torch.empty_like(f0), torch.empty_like(v0)
    bad = dev(np.array([-1, m.ndofs, 1 << 30], np.int32), torch.int32)
    m.assemble(physics=txa.DIFFUSION, coef=k, x=x, f=f1, vals=v1, src=(1.0,), bc_lids=bad)
    assert code_of(engine.sync) == txa.TXA_ERR_INVALID
    assert b"outside the DOF map" in engine.lib.txa_last_error(engine.h)
    engine.sync()                                  # reported once
    assert torch.equal(f0, f1) and torch.equal(v0, v1)
    empty = dev(np.zeros(0, np.int32), torch.int32)
    m.assemble(physics=txa.DIFFUSION, coef=k, x=x, f=f1, vals=v1, src=(1.0,), bc_lids=empty)
    engine.sync()
    assert torch.equal(f0, f1) and torch.equal(v0, v1)
    # the same report from a batch
    b = txa.Batch(engine, [m], precision=txa.FP64)
    b.assemble([dict(physics=txa.DIFFUSION, coef=k, x=x, f=f1, vals=v1, src=(1.0,), bc_lids=bad)],
               use_graph=False)
    assert code_of(engine.sync) == txa.TXA_ERR_INVALID
    b.destroy()
    m.destroy()
