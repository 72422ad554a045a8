"""c2a_batch_run_graph: the batch's launch sequence captured in a HIP graph
and replayed.  Every replay -- with the program's own eager launches and
synchronisations in between -- gives exactly the results of the direct
c2a_batch_run."""
import numpy as np
import pytest
import torch

import c2a_amd as G
from c2a_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_graph_replay_bit_identical(knot_tris):
    t9, idx = knot_tris
    meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    P = W.knot_batch_queries(512)
    for strict in (0, 1):
        print("graph replay strict=%d" % strict, flush=True)
        b = G.Batch(meshes, G.make_queries(*P),
                    G.default_params(tol_d=1e-3, tol_t=1e-4, with_contacts=1, strict=strict))
        b.run()
        ref = b.fetch()
        x = torch.ones(1 << 20, device="cuda")
        for k in range(4):
            b.run_graph()
            torch.cuda.synchronize()
            got = b.fetch()
            for f in ref.dtype.names:
                assert np.array_equal(ref[f], got[f]), (k, f)
            # eager work between replays: a direct launch and a torch kernel
            if k % 2 == 0:
                b.run()
            x.mul_(1.0001)
            torch.cuda.synchronize()
        assert b.last_ms() > 0


def test_graph_rejects_learned_batches(knot_tris):
    t9, idx = knot_tris
    meshes = [G.Mesh(t9, vidx=idx), G.Mesh(t9, vidx=idx)]
    from c2a_amd import learn as Lm
    s = G.Scorer(path=Lm.DEFAULT_SCORER)
    b = G.Batch(meshes, G.make_queries(*W.knot_batch_queries(8)),
                G.default_params(order=G.ORDER_LEARNED, scorer=s), scorer=s)
    with pytest.raises(G.C2AError) as e:
        b.run_graph()
    assert e.value.code == G.ERR_BAD_ARG
