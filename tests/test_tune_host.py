"""Host logic of the autotuner: the candidate space, the workload key and the
JSON cache (no device calls)."""
import json
import types

import numpy as np

from felis_amd import tune


def _desc(**kw):
    d = dict(nx=129, ny=129, slice_n=2048, n_harm=3, harmony=[1, 3, 5, 0, 0, 0, 0, 0])
    d.update(kw)
    return types.SimpleNamespace(**d)


def test_candidates_cover_the_space_once():
    c = list(tune.candidates())
    assert len(c) == 2 * 3 * 3 * 2 * 3 * 2
    assert len({tuple(sorted(t.items())) for t in c}) == len(c)
    assert tune.DEFAULT in c


def test_key_separates_shapes_and_rounds_particles():
    k = tune.key(_desc(), 11_670_128, 1)
    assert k == "gfx950:129x129:s2048:h1,3,5:np11700000:pol1"
    assert tune.key(_desc(), 11_671_000, 1) == k
    assert tune.key(_desc(), 11_670_128, 2) != k
    assert tune.key(_desc(nx=101, ny=101), 11_670_128, 1) != k
    assert tune.key(_desc(slice_n=256, n_harm=1), 1_048_576, 1) == "gfx950:129x129:s256:h1:np1050000:pol1"


def test_energy_error_is_relative_to_the_period_total():
    a = np.array([[1.0, 0.0, 2.0, 0.0], [4.0, 1.0, 0.0, 0.0]])
    b = a.copy()
    assert tune._rel_energy_err(a, b) == 0.0
    b[1, 1] += 5e-6 * 5.0
    assert abs(tune._rel_energy_err(a, b) - 5e-6 / (1 + 5e-6)) < 1e-12


def test_cache_round_trip(tmp_path):
    p = tmp_path / "t.json"
    assert tune.read_cache(p) == {}
    tune.write_cache(dict(key="a", best=dict(step_graph=1), ms_per_step=1.0), p)
    tune.write_cache(dict(key="b", best=dict(step_graph=0), ms_per_step=2.0), p)
    tune.write_cache(dict(key="a", best=dict(step_graph=0), ms_per_step=0.5), p)
    c = json.loads(p.read_text())
    assert c == {"a": dict(best=dict(step_graph=0), ms_per_step=0.5), "b": dict(best=dict(step_graph=0), ms_per_step=2.0)}


def test_committed_cache_holds_valid_tunings():
    for k, e in tune.read_cache().items():
        assert k.startswith(tune.ARCH + ":")
        t = e["best"]
        assert set(t) == set(tune.SPACE)
        assert all(t[n] in tune.SPACE[n] for n in t)
