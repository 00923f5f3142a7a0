"""The library's direct evaluation of the singular face-pair integrals (DIRECTFN, singular_*_const.m)
against the reference's own values: tests/golden/galerkin_singular.json holds [ST, EAc, EAo, VAc, VAo]
computed by the reference's DIRECTFN C++ (oracle/_ref) at Np_1D = 20 and 30 for four (ko, dx), written by
tests/golden/make_galerkin_singular.py.  Bar: 1e-12 relative per value (the same rules summed in another
order).  Host computation: no device needed."""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _lib_or_skip():
    from voxscatter_amd import api
    try:
        from voxscatter_amd import capi
        capi.load()
    except Exception as e:
        pytest.skip(f"library not loadable: {e}")
    return api


@pytest.mark.parametrize("case", [0, 1, 2, 3])
@pytest.mark.parametrize("npts", [20, 30])
def test_directfn_matches_reference_values(case, npts):
    api = _lib_or_skip()
    rec = json.load(open(os.path.join(HERE, "golden", "galerkin_singular.json")))[case]
    ref = np.array([complex(*v) for v in rec[f"directfn_{npts}"]])
    got = api.galerkin_singular_directfn(rec["ko"], rec["dx"], npts)
    err = np.abs(got - ref) / np.abs(ref)
    print(f"[directfn] ko={rec['ko']:.3g} dx={rec['dx']:.3g} Np={npts}: max rel err {err.max():.2e}")
    assert np.all(err <= 1e-12), (err, got, ref)


def test_directfn_static_self_term_converges_to_closed_form():
    """k -> 0: the coincident-square integral tends to a^3 (4/3 (1 - sqrt 2) + 4 ln(1 + sqrt 2)) as Np grows."""
    api = _lib_or_skip()
    exact = 4 / 3 * (1 - math.sqrt(2)) + 4 * math.log(1 + math.sqrt(2))
    e20 = abs(api.galerkin_singular_directfn(1e-9, 1.0, 20)[0].real - exact) / exact
    e40 = abs(api.galerkin_singular_directfn(1e-9, 1.0, 40)[0].real - exact) / exact
    assert e20 < 5e-5 and e40 < e20
