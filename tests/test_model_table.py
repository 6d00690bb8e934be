"""The model table: the host transform (CPU) and the device transform (GPU)
against the reference's own arrays in the golden fixtures, bit for bit.

The device table is what every kernel reads; it is built by k_model from
the raw deck with the same functions as the host (ysx_model.hpp)."""
import numpy as np
import pytest

from conftest import case_names, load_case
from refdump import MODEL_NAMES
from yspec_amd import api

FIELDS = api.TABLE_FIELDS


def _golden(z):
    return np.array([z["model"][MODEL_NAMES.index(f)] for f in FIELDS])


@pytest.mark.parametrize("name", case_names())
def test_host_table_matches_reference(name):
    deck, p, z = load_case(name)
    got = api.model_table_host(deck)
    np.testing.assert_array_equal(got, _golden(z))


@pytest.mark.gpu
@pytest.mark.parametrize("name", case_names())
def test_device_table_matches_reference(name):
    deck, p, z = load_case(name)
    g = api.YSpec(deck)
    dev = g.model_table()
    g.close()
    np.testing.assert_array_equal(dev, api.model_table_host(deck))
    np.testing.assert_array_equal(dev, _golden(z))
