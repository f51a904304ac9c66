"""The hop kernels' traversal orders (natural, z-block x t tiles) and the
autotuned choice give bit-identical results, equal to the oracle, on a
lattice large enough to be tuned (16^3x64, 131072 sites per parity)."""
import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu


def test_traversals_bit_identical():
    L = [16, 16, 16, 64]
    vh = 16 * 16 * 16 * 64 // 2
    u = oracle.gen_gauge(L, 21, 0.4)
    clov = oracle.clover_build(L, u, 1.0, 0.05)
    inv = oracle.clover_invert(clov[:vh * 72])
    psih = oracle.gen_spinor(L, 22, 1)
    g = q.Geometry(L, q.F32, True)
    ud = q.pack_gauge(g, u)
    invd = g.clover().upload(inv)
    psi = g.spinor().upload(psih, 1)
    D = q.Dslash(g, t_boundary=-1.0)
    outs = []
    try:
        for mode, blk, bz, sl in [(0, 256, 0, 1), (2, 128, 4, 8), (2, 256, 8, 16), (2, 64, 16, 8),
                                  (1, 256, 0, 1)]:
            q.call("qphix_tune_config", mode, blk, bz, sl)
            out = g.spinor()
            D.clov_dslash(out, psi, ud[0], invd, 1, 0)
            outs.append(out.download(0))
        print(q.tune_summary())
    finally:
        q.call("qphix_tune_config", 1, 256, 0, 1)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    ref = oracle.clover_apply(g.clover().upload(inv).download(),
                              oracle.hop(L, u, g.spinor().upload(psih, 1).download(1), 1, 0,
                                         oracle.hop_params(-1.0)))
    assert np.linalg.norm(outs[0] - ref) / np.linalg.norm(ref) < 2e-6
