"""Relaxation budgets of the device contact projection (vx_contact_project):
small round budgets hand passes to the window walker at every point of a
call, and every split gives the oracle's result bit for bit."""
import pytest

from vxsim import contact, scenes

from test_gpu_contact import check_parity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,seed,kw,long_chains", [
    ((16, 12, 10), 5, {"movers": 0.05, "mover_reach": 3.0}, False),
    ((20, 20, 12), 8, {"jitter": 1.5e-3, "sag": 2e-4}, False),
    # every bond broken, large jitter, many fast movers: chains of moves through
    # the whole box, far beyond these budgets (and beyond the evaluation budget)
    ((24, 20, 16), 6, {"broken": 1.0, "jitter": 2e-3, "movers": 0.1, "mover_reach": 3.0}, True),
])
def test_relaxation_handing_passes_to_windows(gpu, dims, seed, kw, long_chains):
    """Small round budgets hand passes to the window walker at every point: from
    the first pass's start, or after passes done by relaxation. Every split
    gives the oracle's result bit for bit."""
    s = scenes.contact_box(dims, seed=seed, **kw)
    targets = s.pop("targets")
    split = []
    try:
        for rounds in (1, 2, 3, 4, 6, 8, 10, 12, 16, 24, 32, 0):
            contact.set_method(contact.AUTO, rounds)
            rep, st = check_parity(s, targets=targets, label="%s seed %d budget %d" % (dims, seed, rounds))
            assert rep["relaxed_passes"] <= rep["passes"]
            split.append((rounds, rep["relaxed_passes"], rep["passes"]))
    finally:
        contact.set_method(contact.AUTO)
    print("budget, relaxed passes, passes:", split)
    if long_chains:
        assert all(r[1] == 0 for r in split)  # the first pass already goes to windows
    else:
        assert any(r[1] == 0 for r in split)  # handed over in the first pass
        assert any(0 < r[1] < r[2] for r in split)  # handed over after relaxed passes
        assert split[-1][1] == split[-1][2]  # the default budget relaxes every pass
