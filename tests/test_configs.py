"""Synthetic workload generators: sizes quoted for the benchmark configurations,
PRNG test vector, orientation of the initial states."""
import numpy as np

from tinyad_amd import configs


def test_splitmix64_test_vector():
    # splitmix64 from state 0: first output 0xE220A8397B1DCDAF (Vigna's reference)
    assert int(configs.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF
    u = configs.uniform(3, 100000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
    z = configs.normal(3, 100000)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1.0) < 0.02


def block_nnz(conns, n_vertices, d):
    pairs = set()
    for conn in conns:
        for a in range(conn.shape[1]):
            for b in range(conn.shape[1]):
                pairs.update(zip(conn[:, a].tolist(), conn[:, b].tolist()))
    pairs.update((v, v) for v in range(n_vertices))
    return d * d * len(pairs)


def test_grid_param_sizes():
    P = configs.grid_param(100, seed=1)
    assert P.terms[0].conn.shape[0] == 19602 and P.n_handles == 10000 and P.n_vars == 20000
    # 4 (V + 2 E) with E = 3 n^2 - 4 n + 1 grid edges
    E = 3 * 100 * 100 - 4 * 100 + 1
    assert 4 * (10000 + 2 * E) == 276808
    assert block_nnz([P.terms[0].conn], P.n_handles, 2) == 276808
    assert configs.min_signed_volume(P) > 0.0


def test_twist_cube_sizes_small():
    P = configs.twist_cube(m=5, seed=7)
    assert P.terms[0].conn.shape[0] == 6 * 125 and P.n_handles == 216
    assert configs.min_signed_volume(P) > 0.0
    assert P.info["n_pins"] == 2 * 36
    # all Kuhn tets have the same volume h^3 / 6 and the weights sum to 1
    assert np.allclose(P.terms[0].params[:, 9], 1.0 / 750)


def test_twist_cube_full_size_counts():
    V, T, h = configs.kuhn_cube(60)
    assert T.shape[0] == 1296000 and V.shape[0] == 226981
    e = np.concatenate([T[:, [a, b]] for a in range(4) for b in range(a + 1, 4)])
    e.sort(axis=1)
    n_edges = np.unique(e[:, 0].astype(np.int64) * V.shape[0] + e[:, 1]).size
    assert n_edges == 1544580
    assert 9 * (V.shape[0] + 2 * n_edges) == 29845269


def test_disk_batch_sizes():
    disks = configs.disk_batch(count=8, seed=3)
    for P in disks:
        assert 20 <= P.info["n"] <= 160
        assert configs.min_signed_volume(P) > 0.0
