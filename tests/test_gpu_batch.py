"""Independent SSR instances on per-thread streams: concurrent results agree
with the one-after-another results for instances of different sizes.  Every
sum on the path has a fixed order (reductions, the gather-form assembly and
internal force) and every stream has its own scratch, so the runs agree bit
for bit."""
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
        assert a["FoS"] == b["FoS"]
        assert len(a["lambda_hist"]) == len(b["lambda_hist"])
        assert all(x == y for x, y in zip(a["lambda_hist"], b["lambda_hist"]))
