"""Independent SSR instances on per-thread streams: concurrent results agree
with the one-after-another results for instances of different sizes.  The
reductions are deterministic and every stream has its own scratch, but the
tangent assembly accumulates element blocks with FP64 atomics, whose order
varies from run to run; the factor of safety therefore agrees to 1e-9, not
bit for bit."""
import pytest

from slope_stability_octave_amd import problems

pytestmark = pytest.mark.gpu


def test_concurrent_instances_match_sequential():
    probs = [problems.slope_2d_family(3000 + i, h=h)[0] for i, h in enumerate((1.0, 1.5, 1.0, 2.0))]
    p = dict(problems.PARAMS_FAMILY, linear_tol=1e-2)
    seq, t_seq = problems.run_batch(probs, workers=1, params=p)
    par, t_par = problems.run_batch(probs, workers=4, params=p)
    print("sequential", t_seq, "concurrent", t_par, flush=True)
    for a, b in zip(seq, par):
        assert abs(a["FoS"] - b["FoS"]) <= 1e-9
        assert len(a["lambda_hist"]) == len(b["lambda_hist"])
        assert max(abs(x - y) for x, y in zip(a["lambda_hist"], b["lambda_hist"])) <= 1e-9
