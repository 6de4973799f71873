"""tcscf_compute_F_noalloc through the C ABI: the same contraction as
tcscf_compute_F with its workspace in the library's fixed device array.

Checked on two calls in a row, the second on a smaller input (the array
keeps the first call's contents): the outputs equal tcscf_compute_F's bit for
bit at every precision (same kernels, same workspace contents; the
no-allocation path never splits the BF16 kernel's last round, so the
comparison runs tcscf_compute_F with that split off too), device free
memory does not change across the calls once the first call on the device
has set the array up (the runtime maps it then), and a contraction whose
workspace exceeds the array is rejected before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call_noalloc(tc, g1, g2, kernel, jastrow, same_spin, precision, out):
    from tcscf_amd import _lib
    torch = pytest.importorskip("torch")
    jp = jastrow._struct() if jastrow is not None else None
    rc = tc.lib().tcscf_compute_F_noalloc(
        ctypes.c_void_p(g1.quadratureGrid.points.data_ptr()), ctypes.c_int64(g1.nGrid()),
        ctypes.c_void_p(g2.quadratureGrid.points.data_ptr()), ctypes.c_void_p(g2.quadratureGrid.weights.data_ptr()),
        ctypes.c_void_p(g2.radial_order.data_ptr()), ctypes.c_int64(g2.nGrid()),
        ctypes.c_void_p(g2.basisValues.data_ptr()), g2.nBasis(), kernel,
        ctypes.byref(jp) if jp is not None else None, int(same_spin), _lib.PRECISIONS[precision],
        ctypes.c_double(1.0), ctypes.c_void_p(out.data_ptr()), _lib.F32 if out.dtype == torch.float32 else _lib.F64,
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc


def test_compute_F_noalloc_matches_compute_F(tc, gpu, monkeypatch):
    torch = gpu
    monkeypatch.setenv("TCSCF_BF16_SPLIT", "0")
    from tcscf_amd import _lib, configs
    c = configs.ne_config2()
    big = tc.QMCGrid.radial_lebedev(100, 29, tc.RADIAL_M4, 1.0, 1.0)  # 30,200 points, nb = 30
    big.setBasisFunctions(c["basis"], False)
    small = tc.QMCGrid.radial_lebedev(20, 11, tc.RADIAL_M4, 1.0, 1.0)  # 1,000 points, nb = 9
    basis_s = tc.even_tempered({0: 6, 1: 1}, 0.6, 2.0)
    small.setBasisFunctions(basis_s, False)
    jas = tc.Jastrow.from_b(1.0)
    cases = [("bf16x5", "fp64", _lib.K_COULOMB, None, True), ("fp64", "fp64", _lib.K_COULOMB, None, True),
             ("bf16x3", "fp32", _lib.K_TC_SCALAR, jas, False), ("fp32", "fp64", _lib.K_JASTROW_GRAD, jas, True),
             ("bf16x4", "fp32", _lib.K_COULOMB, None, True)]
    kname = {_lib.K_COULOMB: "coulomb", _lib.K_TC_SCALAR: "tc", _lib.K_JASTROW_GRAD: "grad"}
    # set-up: the first call on the device looks the array up (the runtime maps it)
    warm = torch.empty((small.nGrid(), tc.npair(small.nBasis())), dtype=torch.float64, device="cuda")
    assert _call_noalloc(tc, small, small, _lib.K_COULOMB, None, True, "fp64", warm) == 0, _lib.last_error()
    torch.cuda.synchronize()
    for prec, odt, kernel, j, same in cases:
        refs, outs = [], []
        for g in (big, small):  # two calls in a row, the second smaller
            if kernel == _lib.K_JASTROW_GRAD:
                ref = tc.computeV(g, g, j, precision=prec, out_dtype=odt)
            else:
                ref = tc.computeF(g, g, kernel=kname[kernel], jastrow=j, same_spin=same, precision=prec, out_dtype=odt)
            refs.append(ref)
            outs.append(torch.full_like(ref, float("nan")))
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        for g, out in zip((big, small), outs):
            print(f"step: noalloc {prec} kernel {kernel} n = {g.nGrid()}", flush=True)
            assert _call_noalloc(tc, g, g, kernel, j, same, prec, out) == 0, _lib.last_error()
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free0
        for ref, out in zip(refs, outs):
            assert torch.equal(out, ref), (prec, kernel)
        # the big case again after the small one: the array holds the small
        # call's planes at its start, which this call overwrites
        out = torch.full_like(refs[0], float("nan"))
        assert _call_noalloc(tc, big, big, kernel, j, same, prec, out) == 0
        torch.cuda.synchronize()
        assert torch.equal(out, refs[0])


def test_compute_F_noalloc_rejects_oversize(tc, gpu):
    """A workspace above the fixed array (FP64 pair densities of 60 functions
    on 80,000 points: 80,000 x 1,920 x 8 B = 1.2 GB) is refused with
    TCSCF_EINVAL before any device work; the output is untouched."""
    torch = gpu
    from tcscf_amd import _lib
    ws_bytes = tc.lib().tcscf_F_workspace_bytes
    ws_bytes.restype = ctypes.c_size_t
    assert ws_bytes(ctypes.c_int64(80000), 60, _lib.PRECISIONS["fp64"]) > 128 << 20
    g = tc.QMCGrid.radial_lebedev(10, 7, tc.RADIAL_M4, 1.0, 1.0)
    g.setBasisFunctions(tc.even_tempered({0: 3}, 0.6, 2.0), False)
    out = torch.zeros((g.nGrid(), 6), dtype=torch.float64, device="cuda")
    L = _lib
    rc = tc.lib().tcscf_compute_F_noalloc(
        ctypes.c_void_p(g.quadratureGrid.points.data_ptr()), ctypes.c_int64(g.nGrid()),
        ctypes.c_void_p(g.quadratureGrid.points.data_ptr()), ctypes.c_void_p(g.quadratureGrid.weights.data_ptr()),
        None, ctypes.c_int64(80000), ctypes.c_void_p(g.basisValues.data_ptr()), 60, L.K_COULOMB, None, 1,
        L.PRECISIONS["fp64"], ctypes.c_double(1.0), ctypes.c_void_p(out.data_ptr()), L.F64,
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == L.EINVAL
    assert "exceeds the fixed scratch" in L.last_error()
    torch.cuda.synchronize()
    assert torch.count_nonzero(out) == 0
