"""Outflow faces in the reference's Faces layouts (msw_plan_sweep_faces) on
the GPU: the same vo bits as msw_plan_sweep, the faces of the FP64 sweep
against the oracle's (oracle_sweep_faces, the reference's face arrays with
all eight octants) on perturbed input, and the closed form of the
manufactured problem in every precision."""
import numpy as np
import pytest

import oracle
import minisweep_amd as msw
from parity import assert_close, face_closed_form

pytestmark = pytest.mark.gpu

SHAPES = [
    # ncx, ncy, ncz, ne, nm, na, nu: partial tiles, nm 4/8/16, padded angles
    (9, 7, 5, 2, 16, 64, 4),
    (17, 3, 4, 1, 4, 13, 3),
    (5, 10, 6, 3, 8, 40, 2),
]


def _run(torch, dev, shape, vi0, prec, schedule_kind=0):
    d = msw.dims(*shape)
    a, m = msw.init_matrices(d.nm, d.na)
    ta, tm = torch.from_numpy(a).to(dev), torch.from_numpy(m).to(dev)
    vi = torch.from_numpy(np.ascontiguousarray(vi0)).to(dev)
    ncx, ncy, ncz, ne, nm, na, nu = shape
    fxy = torch.full((8 * ne * ncy * ncx * nu * na,), float("nan"), dtype=torch.float64, device=dev)
    fxz = torch.full((8 * ne * ncz * ncx * nu * na,), float("nan"), dtype=torch.float64, device=dev)
    fyz = torch.full((8 * ne * ncz * ncy * nu * na,), float("nan"), dtype=torch.float64, device=dev)
    plan = msw.Plan(d, prec, engine=msw.ENGINE_MFMA, schedule_kind=schedule_kind)
    vo = torch.full_like(vi, float("nan"))
    plan.sweep_faces(vi, vo, ta, tm, fxy, fxz, fyz)
    vo2 = torch.full_like(vi, float("nan"))
    plan.sweep(vi, vo2, ta, tm)
    torch.cuda.synchronize()
    plan.close()
    return vo.cpu().numpy(), vo2.cpu().numpy(), fxy.cpu().numpy(), fxz.cpu().numpy(), fyz.cpu().numpy()


@pytest.mark.parametrize("shape", SHAPES, ids=[str(s) for s in SHAPES])
def test_faces_fp64_vs_oracle(shape, cuda):
    import torch
    print(f"case faces fp64 {shape}", flush=True)
    od = oracle.dims(*shape)
    rng = np.random.default_rng(3)
    base = oracle.init_state(od)
    vi0 = base * (1.0 + 0.25 * rng.uniform(-1, 1, base.size))
    vo, vo2, fxy, fxz, fyz = _run(torch, cuda, shape, vi0, msw.PREC_FP64)
    np.testing.assert_array_equal(vo, vo2)
    a, m = oracle.init_matrices(od)
    rvo, rxy, rxz, ryz = oracle.sweep_faces(od, a, m, vi0)
    assert_close(vo, rvo, what="vo")
    assert_close(fxy, rxy, what="facexy")
    assert_close(fxz, rxz, what="facexz")
    assert_close(fyz, ryz, what="faceyz")


@pytest.mark.parametrize("prec", [msw.PREC_FP64, msw.PREC_MIXED_FP32, msw.PREC_MIXED_BF16],
                         ids=["fp64", "fp32", "bf16"])
def test_faces_manufactured_closed_form(prec, cuda):
    import torch
    shape = (9, 7, 5, 2, 16, 64, 4)
    print(f"case faces closed form prec {prec}", flush=True)
    base = oracle.init_state(oracle.dims(*shape))
    for kind in (0, 1):
        vo, vo2, fxy, fxz, fyz = _run(torch, cuda, shape, base, prec, schedule_kind=kind)
        np.testing.assert_array_equal(vo, base)
        cxy, cxz, cyz = face_closed_form(shape)
        np.testing.assert_array_equal(fxy, cxy)
        np.testing.assert_array_equal(fxz, cxz)
        np.testing.assert_array_equal(fyz, cyz)


def test_faces_unsupported_plans(cuda):
    import torch
    shape = (8, 8, 8, 1, 16, 32, 4)
    d = msw.dims(*shape)
    a, m = msw.init_matrices(d.nm, d.na)
    ta, tm = torch.from_numpy(a).to(cuda), torch.from_numpy(m).to(cuda)
    vi = torch.zeros(msw.size_state(d), dtype=torch.float64, device=cuda)
    vo = torch.empty_like(vi)
    f = torch.empty(8 * 8 * 8 * 4 * 32, dtype=torch.float64, device=cuda)
    for kw in (dict(engine=msw.ENGINE_VALU), dict(engine=msw.ENGINE_MFMA, z_block=4)):
        plan = msw.Plan(d, msw.PREC_FP64, **kw)
        with pytest.raises(msw.SweepError, match="MFMA engine without z blocks"):
            plan.sweep_faces(vi, vo, ta, tm, f, f, f)
        plan.close()
