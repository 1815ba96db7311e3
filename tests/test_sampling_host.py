"""Flat parameter vectors for sampler callers (no GPU): the vector order and
its mapping onto the pipeline's parameter dict."""
import numpy as np
import pytest

from xpsi_amd import configs
from xpsi_amd._lib import XHError
from xpsi_amd.pipeline import params_from_model
from xpsi_amd.sampling import parameter_names, vector_to_params


def test_config1_order_matches_reference_models():
    # mass, radius, distance, cos_inclination, then the hot region's
    # phase_shift, super_colatitude, super_radius, super_temperature
    assert parameter_names(configs.config1()) == [
        "mass", "radius", "distance", "cos_inclination",
        "phase_shift", "super_colatitude", "super_radius", "super_temperature"]


def test_round_trip_of_the_model_vector():
    m = configs.config1()
    p = params_from_model(m)
    h = p["hot_regions"][0]
    vec = [p["mass"], p["radius"], p["distance"], p["cos_inclination"],
           h["phase_shift"], h["super_colatitude"], h["super_radius"], h["super_temperature"]]
    q = vector_to_params(m, vec)
    assert q["mass"] == p["mass"] and q["cos_inclination"] == p["cos_inclination"]
    assert q["hot_regions"][0] == {k: h[k] for k in q["hot_regions"][0]}
    assert q["hot_regions"][0]["sqrt_num_cells"] == m["hot_regions"][0]["sqrt_num_cells"]


def test_two_regions_with_cede_are_prefixed():
    m = configs.config2("BB")
    names = parameter_names(m)
    assert names[:4] == ["mass", "radius", "distance", "cos_inclination"]
    assert "p__cede_radius" in names and "s__super_temperature" in names
    vec = np.arange(len(names), dtype=float)
    q = vector_to_params(m, vec)
    assert q["hot_regions"][1]["super_temperature"] == names.index("s__super_temperature")
    # a custom order maps by name
    perm = names[::-1]
    q2 = vector_to_params(m, vec[::-1], perm)
    assert q2 == q


def test_wrong_length_or_names_raise():
    m = configs.config1()
    with pytest.raises(XHError):
        vector_to_params(m, np.zeros(7))
    with pytest.raises(XHError):
        vector_to_params(m, np.zeros(8), ["mass"] * 8)
