"""The hop kernels' traversal orders (natural, z-block x t tiles, (y, z) tiles
with LDS reservations) and the autotuned choice give bit-identical results, equal to the oracle, on a
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
        # (y, z) tiles and LDS reservations
        for blk, bz, by, sl, lds in [(256, 4, 8, 1, 0), (128, 2, 16, 4, 40960), (64, 8, 2, 2, 65536),
                                     (256, 16, 4, 1, 54613)]:
            q.call("qphix_tune_traversal", blk, bz, by, sl, lds)
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


def test_multi_rhs_tiles_bit_identical_to_single():
    """Twelve sources through the register tiles under (y, z) tile traversals
    equal the single-source launches bit for bit."""
    L = [16, 16, 16, 64]
    g = q.Geometry(L, q.F32, True)
    g64 = q.Geometry(L, q.F64, False)
    u64 = q.gen_gauge(g64, 23, 0.4)
    inv64 = q.clover_build(g64, u64, 1.0, 0.05, 0, q.CLOV_AINV)
    u = q.convert_gauge(g, g64, u64)
    inv = q.convert_clover(g, g64, inv64)
    psis = [q.gen_spinor(g, 30 + r, 1) for r in range(12)]
    D = q.Dslash(g, t_boundary=-1.0)
    try:
        q.call("qphix_tune_config", 0, 256, 0, 1)
        ref = []
        for p in psis:
            o = g.spinor()
            D.clov_dslash(o, p, u[0], inv, 1, 0)
            ref.append(o.download(0))
        for blk, bz, by, lds, reg, rt in [(128, 4, 8, 40960, 2, 12), (256, 2, 16, 0, 4, 12),
                                          (256, 8, 4, 54613, 2, 4), (128, 4, 8, 65536, 1, 12)]:
            q.call("qphix_tune_traversal", blk, bz, by, 1, lds)
            q.call("qphix_tune_rhs", reg, rt)
            outs = [g.spinor() for _ in psis]
            D.clov_dslash_multi(outs, psis, u[0], inv, 1, 0)
            for r in range(12):
                assert np.array_equal(outs[r].download(0), ref[r]), (blk, bz, by, lds, reg, rt, r)
    finally:
        q.call("qphix_tune_rhs", 0, 0)
        q.call("qphix_tune_config", 1, 256, 0, 1)
