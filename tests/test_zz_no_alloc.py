"""topt_oc_update_noalloc (static scratch, no allocation) against
topt_oc_update (caller workspace), both through the C ABI: bit-identical
densities and results on consecutive calls, the second on a smaller grid
(the shared scratch keeps the previous call's contents), and against the
oracle's OC update."""

import ctypes as C

import numpy as np
import pytest
import torch

from oracle import simp_np as O

pytestmark = pytest.mark.gpu


def _inputs(nelx, nely, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.001, 1.0, nelx * nely)
    dcf = -rng.uniform(1e-4, 2.0, nelx * nely)
    return rho, dcf


def _oc(lib, fn, nelx, nely, rho_np, dcf_np, volfrac, move, ws=None):
    rho = torch.as_tensor(rho_np, device="cuda").contiguous()
    dcf = torch.as_tensor(dcf_np, device="cuda").contiguous()
    out = torch.full((4,), -1.0, dtype=torch.float64, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = [nelx, nely, C.c_void_p(rho.data_ptr()), C.c_void_p(dcf.data_ptr()), None, volfrac, move,
            C.c_void_p(out.data_ptr())]
    if ws is not None:
        args.append(C.c_void_p(ws.data_ptr()))
    rc = fn(*args, stream)
    assert rc == 0, lib.topt_last_error()
    torch.cuda.synchronize()
    return rho.cpu().numpy(), out.cpu().numpy()


def test_noalloc_oc_equals_workspace_oc_on_consecutive_calls():
    from toptimization_amd import _lib, ops
    lib = _lib.load()
    # 1200x400: a two-level reduction tree (235 blocks); 60x20: one block;
    # then the large grid again after the small one
    for k, (nelx, nely, vf) in enumerate([(1200, 400, 0.3), (60, 20, 0.5), (1200, 400, 0.45)]):
        print(f"noalloc OC call {k + 1}: {nelx}x{nely}", flush=True)
        rho, dcf = _inputs(nelx, nely, 40 + k)
        ws = ops.workspace(nelx, nely)
        r_ws, o_ws = _oc(lib, lib.topt_oc_update, nelx, nely, rho, dcf, vf, 0.2, ws)
        r_na, o_na = _oc(lib, lib.topt_oc_update_noalloc, nelx, nely, rho, dcf, vf, 0.2)
        np.testing.assert_array_equal(r_na, r_ws)
        np.testing.assert_array_equal(o_na, o_ws)
        ref, change, trace = O.oc_update(rho, dcf, np.ones_like(rho), vf, 0.2, return_trace=True)
        assert o_na[3] == len(trace)
        assert o_na[2] == trace[-1][0]
        np.testing.assert_allclose(r_na, ref, rtol=0, atol=1e-15)
        assert o_na[0] == pytest.approx(change, abs=1e-15)
        assert o_na[1] == pytest.approx(np.mean(ref), rel=1e-13)


def test_noalloc_rejects_grids_beyond_its_scratch():
    from toptimization_amd import _lib
    lib = _lib.load()
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    p = C.c_void_p(x.data_ptr())
    rc = lib.topt_oc_update_noalloc(8192, 8192, p, p, None, 0.5, 0.2, p, None)
    assert rc == -1
    assert b"static" in lib.topt_last_error()
