"""Autotune's parameter file: entries are found again only when the user
parameters match (floats to 1e-6, temperature to 10 %), and a new entry
replaces a matching one.  No GPU: the file logic is host code."""
from __future__ import annotations

import pytest

from rumd_amd.Autotune import Autotune, _same_user

USER = {"deviceName": "gfx950", "N": 1000, "numberEachType": [800, 200], "rho": 1.2, "T": 0.45,
        "potentialTypes": ["potLJ_12_6"], "pot_parameters": [[2.5, 4.0, -4.0]], "integratorType": "IntegratorNVT",
        "timeStep": 0.0025, "mixedPrecision": False}
TECH = {"pb": 32, "tp": 4, "skin": 0.3, "NB_method": "n2"}


def test_matching_rules():
    assert _same_user(USER, dict(USER))
    assert _same_user(USER, dict(USER, T=0.48))          # within 10 %
    assert not _same_user(USER, dict(USER, T=0.6))
    assert not _same_user(USER, dict(USER, N=1001))
    assert _same_user(USER, dict(USER, rho=1.2000002))      # float noise
    assert not _same_user(USER, dict(USER, rho=1.2001))
    assert not _same_user(USER, dict(USER, pot_parameters=[[2.5, 4.0, -4.1]]))
    assert not _same_user(USER, dict(USER, deviceName="other"))


def test_file_round_trip(tmp_path):
    at = Autotune(auto_filename=tmp_path / "autotune.json")
    assert at.CompareSavedParameters(USER) is None
    at.WriteParametersToFile(USER, TECH)
    assert at.CompareSavedParameters(USER) == TECH
    other = dict(USER, N=2000)
    at.WriteParametersToFile(other, dict(TECH, tp=1))
    at.WriteParametersToFile(USER, dict(TECH, skin=0.4))     # replaces the first entry
    assert at.CompareSavedParameters(USER)["skin"] == 0.4
    assert at.CompareSavedParameters(other)["tp"] == 1
    assert len(at._read_file()) == 2


def test_fixed_parameters_validated():
    with pytest.raises(ValueError):
        Autotune(tp=3)
    with pytest.raises(ValueError):
        Autotune(NB_method="cells")
    with pytest.raises(TypeError):
        Autotune(bogus=1)


def test_batch_entries_match_only_batches(tmp_path):
    """Batch user parameters (lists per replica) compare element by element
    and never match a single-sample entry."""
    from rumd_amd.Autotune import _same_user
    batch = {"deviceName": "d", "kind": "ReplicaBatch", "N": [300, 1200], "L": [6.3, 10.0],
             "thermostats": [[0.5, 0.2], [1.5, 0.2]], "timeStep": 0.0025, "skin": 0.3, "cutoffMethod": 2,
             "params": [2.5, 4.0, -4.0]}
    assert _same_user(batch, dict(batch))
    other = dict(batch, N=[300, 1201])
    assert not _same_user(batch, other)
    assert not _same_user(batch, dict(batch, thermostats=[[0.5, 0.2], [1.6, 0.2]]))
    single = {"deviceName": "d", "N": 1500}
    assert not _same_user(batch, single)
