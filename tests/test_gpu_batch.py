"""Independent SSR instances on per-thread streams: concurrent results equal
the one-after-another results exactly (every reduction is deterministic and
each stream has its own scratch), for instances of different sizes."""
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
        assert list(a["lambda_hist"]) == list(b["lambda_hist"])
