"""Host-side switches of the graph-replayed Davidson passes (no device work): the environment
variables, the SCF setting's default and the graph cache bound."""

from __future__ import annotations


def test_graph_switches_follow_the_environment(monkeypatch):
    from amaterasu_amd import eigen
    from amaterasu_amd.scf import SCFSettings

    monkeypatch.delenv("AMT_DAVIDSON_GRAPHS", raising=False)
    assert not eigen.graphs_enabled()
    assert SCFSettings().davidson_graphs is False
    monkeypatch.setenv("AMT_DAVIDSON_GRAPHS", "1")
    assert eigen.graphs_enabled()
    assert SCFSettings().davidson_graphs is True
    assert SCFSettings(davidson_graphs=False).davidson_graphs is False


def test_graph_cache_bound(monkeypatch):
    from amaterasu_amd import eigen

    monkeypatch.delenv("AMT_DAVIDSON_GRAPHS_MAX", raising=False)
    assert eigen.graphs_per_operator() == 512
    monkeypatch.setenv("AMT_DAVIDSON_GRAPHS_MAX", "3")
    assert eigen.graphs_per_operator() == 3
    monkeypatch.setenv("AMT_DAVIDSON_GRAPHS_MAX", "0")
    assert eigen.graphs_per_operator() == 1


def test_loewdin_bounds_cover_both_block_precisions():
    """The refinement's bound is set per block precision so that its O(E^3) error stays below
    that precision's own rounding: (2e-3)^3 = 8e-9 under FP32's unit roundoff 6e-8, and
    (1e-5)^3 = 1e-15 under the 1e-14 that FP64 inner products over 1e4..1e5 plane waves
    already carry."""
    import torch

    from amaterasu_amd import eigen

    assert set(eigen._LOWDIN_MAX) == {torch.complex64, torch.complex128}
    assert eigen._LOWDIN_MAX[torch.complex64] ** 3 < 2.0 ** -24
    assert eigen._LOWDIN_MAX[torch.complex128] ** 3 < 1e-14
