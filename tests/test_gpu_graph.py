"""HIP graph replay of the batched solve (needs a GPU).

The captured solve (tables on two streams, binning, exponent, two-point
term, inversion) is replayed several times with eager solves of other
records in between; every replay must equal an eager solve of the same
records bit for bit: the histogram's sums are fixed-point integer atomics,
whose result does not depend on the order in which they arrive, and every
other reduction of the batched path runs in a fixed order."""
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
            assert torch.equal(p, want[q]), (rep, q, d)


@pytest.mark.parametrize("prec", [1, 2])
def test_batched_solve_is_reproducible(prec):
    """Two solvers, the same records, two runs each: the same bits, at the
    default 256-mass x 512-column lattice (the histogram's atomics collide
    most there), with and without ellipticity and filaments."""
    from sgl_analytic_amd.solver import Shape, Solver
    zs = (0.4, 1.1, 2.5, 6.0)
    for kw in (dict(), dict(ell=True, filaments=True)):
        sh = Shape(n_z_max=16, n_k=1024, n_out=256, **kw)
        print(f"reproducibility prec={prec} {kw}", flush=True)
        r = _records(zs, 7)
        r[:, 10] = 16
        out = []
        for _ in range(2):
            s = Solver(sh, 4, path="binned", precision=prec)
            for _ in range(2):
                p, d = s.run(r)
                torch.cuda.synchronize()
                out.append((p.clone(), d.clone(), s.hist.clone()))
            del s
        for p, d, h in out[1:]:
            assert torch.equal(p, out[0][0]) and torch.equal(d, out[0][1])
            assert torch.equal(h, out[0][2])
