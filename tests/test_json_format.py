"""The on-disk format of the reference (pyANNonGPU/PsiDeep.py:7-44, PsiCNN.py:7-44,
json_numpy.py:5-38) against fixtures that spell it out by hand
(tests/golden/make_json_fixtures.py): from_json reads them into the numbers they hold,
and to_json writes back exactly the same objects (dtypes included), so files move
between the reference and this package in both directions.  The classical network's
format (pyANNonGPU/PsiClassical.py:7-67) is checked by a round trip, with the fully
polarised, a PsiDeep and a PsiCNN reference; its ansatz entries use this package's
PauliExpression encoding.  The GPU test evaluates the loaded
network against the oracle built from the fixture's numbers.
"""
import json
import os

import numpy as np
import pytest

import pyoracle as orc
import annongpu_amd as A

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


def arr(o):
    """the fixture's encoding decoded by hand"""
    if o["complex"]:
        return np.array(o["real"]) + 1j * np.array(o["imag"])
    return np.array(o["data"], dtype=o["dtype"])


def test_psideep_fixture_reads_and_writes_back():
    fix = load("psideep_reference_format.json")
    psi = A.PsiDeep.from_json(fix)
    assert psi.num_sites == fix["num_sites"]
    assert np.array_equal(psi.a, arr(fix["a"]))
    for l in range(2):
        assert np.array_equal(psi.b[l], arr(fix["b"][l]))
        assert np.array_equal(psi.W[l], arr(fix["W"][l]))
        assert np.array_equal(psi.connections[l], arr(fix["connections"][l]))
    assert np.array_equal(psi.final_weights, arr(fix["final_weights"]))
    assert psi.log_prefactor == complex(fix["log_prefactor_re"], fix["log_prefactor_im"])
    assert psi.to_json() == fix


def test_psicnn_fixture_reads_and_writes_back():
    fix = load("psicnn_reference_format.json")
    psi = A.PsiCNN.from_json(fix)
    assert list(psi.extent) == fix["extent"]
    assert np.array_equal(psi.params, arr(fix["params"]))
    assert psi.final_factor == fix["final_factor"]
    out = psi.to_json()
    assert out == fix, {k: (out[k], fix[k]) for k in fix if out.get(k) != fix[k]}


def test_classical_network_round_trip():
    H = A.sigma_z(0) * A.sigma_x(1) * A.sigma_z(2) + A.sigma_y(0) + A.sigma_x(0) * A.sigma_x(1)
    psi = A.new_classical_network(6, 1, H)
    psi.params = np.array([0.1 + 0.2j, -0.3, 0.05j])
    js = psi.to_json()
    assert js["type"] == "PsiClassical" and js["psi_ref"]["type"] == "PsiFullyPolarized"
    back = A.PsiClassical.from_json(json.loads(json.dumps(js)))
    assert np.array_equal(back.params, psi.params) and back.order == 1
    assert [o.expr for o in back.H_local] == [o.expr for o in psi.H_local]
    ref = A.PsiDeep.from_json(load("psideep_reference_format.json"))
    c2 = A.new_classical_network(6, 2, H, psi_ref=ref)
    assert c2.num_params == 3 + ref.num_params
    back2 = A.PsiClassical.from_json(c2.to_json())
    assert np.array_equal(back2.params, c2.params) and back2.order == 2
    # a PsiCNN reference (the reference's PsiClassicalANN)
    cref = A.PsiCNN.from_json(load("psicnn_reference_format.json"))
    c3 = A.new_classical_network(cref.num_sites, 2, A.sigma_x(0) + A.sigma_z(1) * A.sigma_z(2), psi_ref=cref)
    js = c3.to_json()
    assert js["psi_ref"]["type"] == "PsiCNN"
    back3 = A.PsiClassical.from_json(json.loads(json.dumps(js)))
    assert isinstance(back3.psi_ref, A.PsiCNN) and np.array_equal(back3.params, c3.params)
    assert back3.num_params == 2 + cref.num_params


@pytest.mark.gpu
def test_psideep_fixture_evaluates_like_the_oracle():
    fix = load("psideep_reference_format.json")
    psi = A.PsiDeep.from_json(fix)
    o = orc.deep(fix["num_sites"], arr(fix["a"]), [arr(b) for b in fix["b"]],
                 [arr(c) for c in fix["connections"]], [arr(w) for w in fix["W"]], arr(fix["final_weights"]),
                 complex(fix["log_prefactor_re"], fix["log_prefactor_im"]))
    ex = A.ExactSummationSpins(fix["num_sites"])
    ref = orc.exact_log_psi_vector(o, fix["num_sites"])
    assert np.max(np.abs(A.log_psi_vector(psi, ex) - ref)) < 1e-10 * max(1.0, np.abs(ref).max())
