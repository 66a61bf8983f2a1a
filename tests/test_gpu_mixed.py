"""GPU mixed-precision sweep (FP32 or BF16 contractions on MFMA, FP32 face
recurrence, FP64 state, vo accumulation and norms) against the FP64 oracle.

Inputs: the deterministic manufactured state, and the same state perturbed
elementwise by (1 + e), e ~ U(-2^-8, 2^-8) with a fixed seed.

Stated error bars, relative to the FP64 result:
  * manufactured state, both forms: vo and normsq exact.  Every operand
    (state, matrices, faces) is a small integer times a power of two that
    BF16/FP32 hold exactly, and all partial sums stay exact in FP32.
  * perturbed state, FP32 form: normsq within 1e-5 of the FP64 oracle (the
    stated target) and vo normwise (max |d| / max |ref|) within 1e-4.  The
    rows of A and M are second differences weighted by 21..55 that cancel
    the smooth part of the input, so roundings show up ~300x larger in the
    result.  Rounding vi itself to FP32 was the largest term (normsq errors
    up to 2.4e-5 on these shapes); the form carries vi's rounding residual in
    BF16 through a second m2a MFMA, which leaves the FP32 accumulation as the
    floor (a CPU emulation over seeds: 0.9e-6 .. 2.4e-6).
  * perturbed state, BF16 form: no accuracy claim.  Rounding the state and
    the solved angular values to BF16 (2^-9) before the same cancellation
    leaves O(1) normwise error (measured 0.5 .. 1.2); the test only checks
    that the result is finite and within that measured band.
"""
import numpy as np
import pytest

import oracle
import minisweep_amd as msw
from parity import closed_form_normsq

pytestmark = pytest.mark.gpu

PRECS = [msw.PREC_MIXED_FP32, msw.PREC_MIXED_BF16]
PREC_IDS = ["fp32", "bf16"]
# perturbed-input bars (the manufactured state must come out exact)
VO_TOL = {msw.PREC_MIXED_FP32: 1e-4, msw.PREC_MIXED_BF16: 4.0}
NORMSQ_TOL = {msw.PREC_MIXED_FP32: 1e-5, msw.PREC_MIXED_BF16: 8.0}

SHAPES = [
    # ncx, ncy, ncz, ne, nm, na, nu
    (8, 8, 8, 4, 16, 32, 4),
    (9, 7, 3, 2, 16, 64, 4),
    (17, 3, 5, 1, 8, 40, 2),
    (3, 20, 9, 3, 4, 13, 3),
    (12, 10, 11, 2, 16, 64, 4),
]


def _sweep(d, vi0, prec, torch, device, nac=0):
    a, m = msw.init_matrices(d.nm, d.na)
    ta = torch.from_numpy(a).to(device)
    tm = torch.from_numpy(m).to(device)
    vi = torch.from_numpy(np.ascontiguousarray(vi0)).to(device)
    vo = torch.full_like(vi, float("nan"))
    plan = msw.Plan(d, prec, engine=msw.ENGINE_MFMA, angles_per_wave=nac)
    plan.sweep(vi, vo, ta, tm)
    torch.cuda.synchronize()
    n, nd = msw.state_norms(d, vi, vo)
    plan.close()
    return vo.cpu().numpy(), n, nd


def _errs(vo, ref):
    normwise = float(np.max(np.abs(vo - ref)) / np.max(np.abs(ref)))
    return normwise


@pytest.mark.parametrize("prec", PRECS, ids=PREC_IDS)
@pytest.mark.parametrize("shape", SHAPES, ids=[str(s) for s in SHAPES])
def test_mixed_manufactured(shape, prec, cuda):
    import torch
    ncx, ncy, ncz, ne, nm, na, nu = shape
    print(f"case {shape} prec {prec}", flush=True)
    d = msw.dims(*shape)
    vi0 = oracle.init_state(oracle.dims(*shape))
    vo, n, nd = _sweep(d, vi0, prec, torch, cuda)
    assert np.isfinite(vo).all()
    err = _errs(vo, vi0)
    exact = float(closed_form_normsq(ncx, ncy, ncz, ne, nm, nu))
    rel_n = abs(n - exact) / exact
    print(f"  vo normwise err {err:.3e}  normsq rel err {rel_n:.3e}", flush=True)
    np.testing.assert_array_equal(vo, vi0)
    assert n == exact and nd == 0.0


@pytest.mark.parametrize("prec", PRECS, ids=PREC_IDS)
@pytest.mark.parametrize("shape", SHAPES, ids=[str(s) for s in SHAPES])
def test_mixed_perturbed_vs_oracle(shape, prec, cuda):
    import torch
    print(f"case {shape} prec {prec}", flush=True)
    d = msw.dims(*shape)
    od = oracle.dims(*shape)
    rng = np.random.default_rng(1234)
    base = oracle.init_state(od)
    vi0 = base * (1.0 + rng.uniform(-2.0 ** -8, 2.0 ** -8, base.size))
    vo, n, nd = _sweep(d, vi0, prec, torch, cuda)
    a, m = oracle.init_matrices(od)
    ref = oracle.sweep(od, a, m, vi0)
    n_ref, _ = oracle.norms(od, vi0, ref)
    err = _errs(vo, ref)
    rel_n = abs(n - n_ref) / n_ref
    print(f"  vo normwise err {err:.3e}  normsq rel err {rel_n:.3e}", flush=True)
    assert np.isfinite(vo).all()
    assert err <= VO_TOL[prec]
    assert rel_n <= NORMSQ_TOL[prec]


@pytest.mark.parametrize("prec", PRECS, ids=PREC_IDS)
def test_mixed_deterministic_and_chunking(prec, cuda):
    """Same bits on a rerun; angle chunking changes only the summation order."""
    import torch
    shape = (9, 7, 5, 2, 16, 64, 4)
    d = msw.dims(*shape)
    vi0 = oracle.init_state(oracle.dims(*shape))
    r1, _, _ = _sweep(d, vi0, prec, torch, cuda, nac=16)
    r2, _, _ = _sweep(d, vi0, prec, torch, cuda, nac=16)
    np.testing.assert_array_equal(r1, r2)
    r3, _, _ = _sweep(d, vi0, prec, torch, cuda, nac=64)
    np.testing.assert_array_equal(r3, vi0)  # exact either way on this input


MODEL_SHAPES = [(6, 5, 7, 2, 16, 64, 2), (9, 7, 3, 2, 16, 64, 4), (7, 3, 4, 1, 8, 40, 3)]
# GPU vs its own rounding model (oracle/mixed_model.py): the model's two
# accumulation variants (one rounding per block sum, or the products aligned
# and truncated as a matrix core may) agree to ~1e-8 normwise in BF16, so
# the bars are loose against that and tight against the O(1) a wrong index
# gives; the BF16 form itself is O(1) from FP64 on this input.
MODEL_VO_TOL = {msw.PREC_MIXED_FP32: 1e-4, msw.PREC_MIXED_BF16: 1e-3}
MODEL_NORMSQ_TOL = {msw.PREC_MIXED_FP32: 1e-5, msw.PREC_MIXED_BF16: 1e-4}


@pytest.mark.parametrize("prec", PRECS, ids=PREC_IDS)
@pytest.mark.parametrize("shape", MODEL_SHAPES, ids=[str(s) for s in MODEL_SHAPES])
def test_mixed_matches_rounding_model(shape, prec, cuda):
    """Each mixed form against a CPU model of its own roundings (BF16 or FP32
    operands, FP32 solve and sums), perturbed input: pins the BF16 form, whose
    error against FP64 is O(1), to the arithmetic it states."""
    import torch
    from oracle import mixed_model
    print(f"case model {shape} prec {prec}", flush=True)
    d = msw.dims(*shape)
    od = oracle.dims(*shape)
    rng = np.random.default_rng(1234)
    base = oracle.init_state(od)
    vi0 = base * (1.0 + rng.uniform(-2.0 ** -8, 2.0 ** -8, base.size))
    vo, n, nd = _sweep(d, vi0, prec, torch, cuda)
    a, m = oracle.init_matrices(od)
    form = "bf16" if prec == msw.PREC_MIXED_BF16 else "fp32"
    ref = mixed_model.sweep(shape, a, m, vi0, form)
    err = _errs(vo, ref)
    n_ref = float(np.sum(ref * ref))
    rel_n = abs(float(np.sum(vo * vo)) - n_ref) / n_ref
    print(f"  vo normwise err vs model {err:.3e}  normsq rel err {rel_n:.3e}", flush=True)
    assert err <= MODEL_VO_TOL[prec]
    assert rel_n <= MODEL_NORMSQ_TOL[prec]


def test_mixed_then_fp64_iterations(cuda):
    """The mixed form with a final FP64 sweep: of the runner's ping-pong
    iterations the first ones in FP32, the last in FP64.  Manufactured input:
    four iterations exact.  Perturbed input: two iterations (FP32, FP64)
    against the oracle's two FP64 iterations within the FP32 form's 1e-5
    normsq bar (a second FP32 iteration of this input leaves FP32's range on
    large grids: each sweep multiplies the perturbation by orders of
    magnitude)."""
    import torch
    shape = (9, 7, 5, 2, 16, 64, 4)
    print(f"case mixed then fp64 {shape}", flush=True)
    d = msw.dims(*shape)
    od = oracle.dims(*shape)
    a, m = msw.init_matrices(d.nm, d.na)
    ta, tm = torch.from_numpy(a).to(cuda), torch.from_numpy(m).to(cuda)
    p32 = msw.Plan(d, msw.PREC_MIXED_FP32, engine=msw.ENGINE_MFMA)
    p64 = msw.Plan(d, msw.PREC_FP64, engine=msw.ENGINE_MFMA)

    def run(plans, x0):
        u = torch.from_numpy(np.ascontiguousarray(x0)).to(cuda)
        w = torch.full_like(u, float("nan"))
        for pl in plans:
            pl.sweep(u, w, ta, tm)
            u, w = w, u
        torch.cuda.synchronize()
        return u.cpu().numpy()

    base = oracle.init_state(od)
    np.testing.assert_array_equal(run([p32, p32, p32, p64], base), base)
    rng = np.random.default_rng(7)
    x0 = base * (1.0 + rng.uniform(-2.0 ** -8, 2.0 ** -8, base.size))
    got = run([p32, p64], x0)
    oa, om = oracle.init_matrices(od)
    ref = oracle.sweep(od, oa, om, oracle.sweep(od, oa, om, x0))
    n_ref = float(np.sum(ref * ref))
    rel_n = abs(float(np.sum(got * got)) - n_ref) / n_ref
    print(f"  two iterations: normsq rel err {rel_n:.3e}", flush=True)
    assert rel_n <= 1e-5
    p32.close()
    p64.close()
