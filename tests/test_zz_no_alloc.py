"""The planner that allocates nothing (``hpx_uneven_prepare_noalloc``): lean drops' tables in the
library's scratch array instead of a caller workspace.  Its launches must give exactly what the
workspace planner's launches give, on two batches in a row (the second smaller), eagerly and
from a captured graph."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _shapes():
    from hermespy_amd.uneven import DropShape

    lean = [DropShape(8, 2048, 23, 1.5, 400.0, 16, 12.0), DropShape(8, 2048, 5, 4.0, 60.0, 4, 6.0),
            DropShape(8, 2048, 14, 2.5, 850.0, 64, 24.0), DropShape(8, 2048, 9, 1.0, 200.0, 16, 30.0)]
    others = [DropShape(8, 1024, 11, 2.0, 300.0, 16, 15.0), DropShape(2, 2048, 7, 3.0, 100.0, 64, 21.0),
              DropShape(1, 256, 6, 5.0, 700.0, 4, 3.0)]
    return lean + others


def _errors(batch):
    import torch

    batch.errors.fill_(-1)
    batch.run()
    torch.cuda.synchronize()
    return batch.errors.cpu().numpy().copy()


def test_noalloc_planner_matches_workspace_planner(cuda):
    import torch

    from hermespy_amd.uneven import UnevenBatch

    pop = _shapes()
    print("no-alloc planner: 7 mixed drops (4 lean) against the workspace planner", flush=True)
    ref = UnevenBatch(pop, seed=77, first_drop=500)
    na = UnevenBatch(pop, seed=77, first_drop=500, no_alloc=True)
    assert na.workspace is None and na.num_lean == 4
    e_ref, e_na = _errors(ref), _errors(na)
    assert (e_ref > 0).all()
    assert np.array_equal(e_na, e_ref)
    # a second, smaller batch through the same scratch array (other drops, other lean tables)
    small = [pop[2], pop[0], pop[5]]
    print("no-alloc planner: second batch of 3 drops", flush=True)
    ref2 = UnevenBatch(small, seed=78, first_drop=900)
    na2 = UnevenBatch(small, seed=78, first_drop=900, no_alloc=True)
    e_ref2, e_na2 = _errors(ref2), _errors(na2)
    assert np.array_equal(e_na2, e_ref2)
    # graph replays of the no-alloc batch, with eager work in between
    na2.capture()
    for rep in range(2):
        na2.errors.fill_(-1)
        torch.cuda.synchronize()
        na2.replay()
        torch.cuda.synchronize()
        assert np.array_equal(na2.errors.cpu().numpy(), e_ref2), rep
        _errors(ref2)
