"""HIP graph replay of the batched solve (needs a GPU).

The captured solve (tables on two streams, binning, exponent, two-point
term, inversion) is replayed several times with eager solves of other
records in between; every replay must agree with an eager solve of the same
records (to the round-off of the FP64 histogram atomics)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _records(zs, seed):
    from sgl_analytic_amd.solver import cosmo_records
    rng = np.random.default_rng(seed)
    p = np.array([[0.674 + 0.02 * rng.standard_normal(), 0.315, 0.811, 0.0493, 0.965, 3402.0,
                   0.0, -1.0 + 0.05 * rng.standard_normal(), 0.0, z] for z in zs])
    return cosmo_records(p, 8)


@pytest.mark.parametrize("clustering", [True, False])
def test_graph_replays_match_eager(clustering):
    from sgl_analytic_amd._lib import PREC_FP32
    from sgl_analytic_amd.solver import Shape, Solver
    sh = Shape(n_z_max=8, n_m=32, n_b=64, n_k=1024, n_out=256, clustering=clustering)
    zs = (0.4, 1.1, 2.5, 6.0)
    recs = [_records(zs, s) for s in range(3)]
    eager = Solver(sh, 4, path="binned", precision=PREC_FP32)
    want = []
    for r in recs:
        p, _ = eager.run(r)
        torch.cuda.synchronize()
        want.append(p.clone())
    g = Solver(sh, 4, path="binned", precision=PREC_FP32)
    g.set_records(recs[0])
    g.capture()
    for rep in range(2):
        for q, r in enumerate(recs):
            print(f"graph replay {rep} records {q} clustering={clustering}", flush=True)
            eager.run(recs[(q + 1) % 3])          # eager work between replays
            torch.cuda.synchronize()
            p, _ = g.run(r)
            torch.cuda.synchronize()
            d = float((p - want[q]).abs().max()) / float(want[q].abs().max())
            assert d < 1e-5, (rep, q, d)
