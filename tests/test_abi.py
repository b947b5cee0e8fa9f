"""The C ABI library loads and exports every function include/netflux.h
declares; on a machine without a HIP device nf_create reports NF_E_NOGPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import netflux_amd as nf
from netflux_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "netflux.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nf_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(A.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(n for n, _, _ in A.SIGNATURES) == declared_functions()


def test_struct_layouts():
    assert C.sizeof(A.Params) == 4 * 4 + 8 + 40 + 4 * 4 + 8  # 5th int + tail padding
    assert C.sizeof(A.MeshDesc) == 8 * 8 + 4 * 8  # four mixed-element pointers
    assert C.sizeof(A.RunStats) == 8 + 5 * 8 + 8
    assert C.sizeof(A.RunOpts) == 40
    assert C.sizeof(A.TrainOpts) == 40 and C.sizeof(A.TrainStats) == 24
    assert C.sizeof(A.TuneResult) == 4 + 4 + 8 + 8 + 8 * 8  # n_tried padded to 8


def test_no_gpu_reports_nogpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    mesh = nf.Mesh.box(2)
    ctx = C.c_void_p()
    p = nf.Config().to_params()
    rc = A.lib().nf_create(C.byref(mesh.desc), C.byref(p), C.byref(ctx))
    assert rc == A.NF_E_NOGPU
    descs = (A.MeshDesc * 2)(mesh.desc, mesh.desc)
    params = (A.Params * 2)(p, p)
    assert A.lib().nf_create_batch(2, descs, params, C.byref(ctx)) == A.NF_E_NOGPU


def test_batch_argument_checks():
    # rejected before any device call: bad count, differing shared options
    mesh = nf.Mesh.box(2)
    ctx = C.c_void_p()
    p = nf.Config().to_params()
    q = nf.Config(CFL=0.3).to_params()
    descs = (A.MeshDesc * 2)(mesh.desc, mesh.desc)
    assert A.lib().nf_create_batch(0, descs, (A.Params * 2)(p, p), C.byref(ctx)) == A.NF_E_ARG
    assert A.lib().nf_create_batch(2, descs, (A.Params * 2)(p, q), C.byref(ctx)) == A.NF_E_ARG


def test_mesh_errors_are_reported():
    with pytest.raises(RuntimeError):
        nf.Mesh.box(0)
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    cells = np.array([[0, 1, 2, 7]], np.uint32)  # node 7 does not exist
    m = nf.Mesh.from_arrays(nodes, cells, [[[0, 1, 2]]], ["SlipWall"])
    with pytest.raises(RuntimeError):
        m.fv_grid()
