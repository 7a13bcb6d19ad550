"""Host-side generators behind the workloads and the bench's byte model
(no GPU)."""
import math

import numpy as np

import bench
from tbk import structures as S


def test_monkhorst_pack_grid():
    k = S.monkhorst_pack((8, 8, 8))
    assert k.shape == (512, 3)
    # (2r - q - 1) / 2q: symmetric about 0, no point at Gamma for even q
    assert np.allclose(np.sort(np.unique(k[:, 0])), (2 * np.arange(1, 9) - 9) / 16)
    assert not np.any(np.all(k == 0, axis=1))
    assert np.any(np.all(S.monkhorst_pack((5, 5, 5)) == 0, axis=1))


def test_configs_match_their_descriptions():
    c1, c2, c3 = S.config_c1(), S.config_c2(), S.config_c3()
    assert (c1.natoms, c1.nv) == (4, 36)
    assert (c2.natoms, c2.nv) == (32, 288)
    assert (c3.natoms, c3.nv) == (108, 972)
    # 11 electrons per atom
    for st in (c1, c2, c3):
        assert math.isclose(st.zfill, 11.0 * st.natoms)
    # the neighbour search reaches the cutoff in every direction
    for st in (c1, c2, c3):
        rec = np.linalg.inv(st.lat).T
        for i in range(3):
            assert (st.nd[i] - 1) / np.linalg.norm(rec[i]) >= st.params.dnn[0] - 1e-9


def test_uneven_batch_distribution():
    c4 = S.config_c4()
    assert len(c4) == 256
    ns = np.array([st.nv for st, _ in c4])
    qs = np.array([q for _, q in c4])
    assert ns.min() == 9 and ns.max() == 576
    assert qs.min() >= 2 and qs.max() <= 8
    kinds = [st.label[:3] for st, _ in c4]
    for kind in ("fcc", "bcc", "hcp"):
        assert 60 <= kinds.count(kind) <= 110
    # N * q^3 balanced: within a factor 4 of the budget where q is not clipped
    work = ns * qs ** 3
    free = (qs > 2) & (qs < 8)
    assert np.all(work[free] < 4 * 9 * 8 ** 3) and np.all(work[free] > 9 * 8 ** 3 / 4)
    # seeded: the same call gives the same structures
    again = S.config_c4()
    assert all(np.array_equal(a.posn, b.posn) for (a, _), (b, _) in zip(c4[:8], again[:8]))


def test_hetrd_panel_byte_model():
    # the Hermitian products dominate at large n: about n^3/6 complex
    # elements per k, plus the panel terms (O(32 n^2))
    ratios = [bench.hetrd_panel_bytes(n, 16) / 16 / (n ** 3 / 6) for n in (64, 288, 972, 1280)]
    assert all(a > b for a, b in zip(ratios, ratios[1:]))
    assert 1.0 < ratios[-1] < 1.2
    assert bench.hetrd_panel_bytes(288, 8) * 2 == bench.hetrd_panel_bytes(288, 16)


def test_measured_traffic_is_read_from_the_profile():
    t64, t32 = bench.measured_traffic("fp64"), bench.measured_traffic("mixed")
    assert t64 is not None and t32 is not None
    alg64 = bench.hetrd_panel_bytes(288, 16) * 512 / 9
    assert 0.8 < t64 / alg64 < 1.3
