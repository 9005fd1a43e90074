"""Passes replayed as hipGraphs (use_graph): every replay bit-identical to
the same passes launched directly, with the program's own transfers and
synchronisations between replays, a dumped pass run directly in between and
both the HIO and the ER graph of each buffer state reused."""
import numpy as np
import pytest

from tests.util import blob_phantom

pytestmark = pytest.mark.gpu

# (np = nz, precision, denoise, method, height)
CASES = [(64, 1, True, 1, 32), (64, 1, False, 1, None), (32, 0, True, 0, None), (48, 1, True, 0, 24)]


@pytest.mark.parametrize("n,prec,denoise,method,height", CASES)
def test_graph_replays_match_direct(hip, n, prec, denoise, method, height):
    rng = np.random.default_rng(n + prec)
    dens = blob_phantom((n, n, n), 10, rng, sig=(1.5, 3.0), slab=n // 2)
    proj = hip.project_2d1d(dens, n, -60, 60, 5)
    center = None if height is None else n // 2
    out = {}
    for graph in (False, True):
        p = hip.make_params(n, n, n, -60, 60, 5, padangle=1.0, weight=True, nhio=4, ner=8, denoise=denoise,
                            method=method, height=height, center=center, precision=prec, use_graph=graph)
        vols, rs = [], []
        before, after = np.zeros((n, n, n)), np.zeros((n, n, n))
        with hip.Reconstruction(p) as job:
            job.load_projections(proj)
            job.initial()
            job.set_dump(5, before, after)
            for count in (1, 2, 3, 1, 5):  # 12 passes: HIO 0-3, ER 4-11; pass 5 dumped
                rs.append(job.iterate(count))
                vols.append(job.volume())
        out[graph] = (vols, rs, before, after)
    (v0, r0, b0, a0), (v1, r1, b1, a1) = out[False], out[True]
    for k, (x, y) in enumerate(zip(v0, v1)):
        assert np.array_equal(x, y), "volume after call %d differs" % k
    for x, y in zip(r0, r1):
        assert np.array_equal(x, y)
    if denoise:
        assert np.array_equal(b0, b1) and np.array_equal(a0, a1)
