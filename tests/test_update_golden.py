"""The host batch tidy and update (lvgraph.tidy_batch / apply_batch, used by
the benchmarks and the oracle-side tests) against the reference's own
tidyBatchUpdateU + applyBatchUpdateOmpU golden vectors; the rules are in
update_golden.py.  The device update is checked against the same vectors
in test_gpu_update.py."""
import pytest

import update_golden as U


@pytest.mark.parametrize("case", U.load(), ids=lambda c: c[0])
def test_host_update_matches_reference(case):
    import lvgraph
    name, g, (rd, ri), (td, ti), y = case
    got_d, got_i = lvgraph.tidy_batch(g, rd, ri)
    rep = U.check_tidy(got_d, td, rd, f"{name} deletions") + U.check_tidy(got_i, ti, ri, f"{name} insertions")
    U.check_graph(lvgraph.apply_batch(g, rd, ri), y, ri, name)
    print(f"{name}: {len(td['u'])} deletions, {len(ti['u'])} insertions, {rep} repeated pairs with several weights")


def test_golden_cases_hold_repeats_with_different_weights():
    """The vectors exercise the unspecified case: repeated pairs with
    different weights in both lists of every non-tiny case."""
    for name, g, (rd, ri), _, _ in U.load():
        for raw in (rd, ri):
            _, every = U._weights_by_pair(raw)
            n = sum(len(ws) > 1 for ws in every.values())
            assert n >= (1 if name == "tiny1" else 5), (name, n)
