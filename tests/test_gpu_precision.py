"""Pyramid storage precision against the FP64 oracle: 1080p, 16,200 grid points, about 10% of them in near-flat patches
(3 grey of texture). Exact storage (Q6 u8/u16, fp32) must reproduce the oracle to the parity band; fp16 and bf16
storage quantize levels 1 and up and are held to the bounds they meet with margin, on the well-conditioned textured
points. The mixed solve (fp32, fp64 above condition number 1e4) stays within the same bounds as the fp64 solve. The Scharr
gradient variant (no reference counterpart) is held to the same bounds against the oracle's FP64 Scharr tier."""
import pytest

import vtgpu
from gpu_util import precision_eval, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["sobel", "scharr"])
def prec(ctx, oracle, request):
    say("precision eval 1080p", request.param)
    return precision_eval(vtgpu, oracle, ctx, grad=1 if request.param == "scharr" else 0)


def test_workload_shape(prec):
    assert 0.08 <= prec["flat_points"] / prec["points"] <= 0.12
    assert prec["well_conditioned"] > 0.75 * prec["points"]


@pytest.mark.parametrize("combo", ["q6/f64", "q6/mixed", "f32/f64", "f32/mixed"])
def test_exact_storage(prec, combo):
    r = prec[combo]
    assert r["status_equal"] >= 0.999, r
    for part in ("textured", "flat"):
        assert r[part]["p99_px"] <= 1e-4, (part, r)
        assert r[part]["max_px"] <= 0.01, (part, r)


@pytest.mark.parametrize("combo,p99", [("f16/f64", 0.03), ("f16/mixed", 0.03), ("bf16/f64", 0.08),
                                       ("bf16/mixed", 0.08)])
def test_half_storage(prec, combo, p99):
    r = prec[combo]
    assert r["status_equal"] >= 0.99, r
    assert r["textured"]["p50_px"] <= 1e-3, r
    assert r["textured"]["p99_px"] <= p99, r


def test_fallback_exercised(prec):
    # some converged windows are conditioned badly enough for the mixed solve to take its fp64 branch
    assert prec["q6/mixed"]["cond_over_1e4"] > 0.001
