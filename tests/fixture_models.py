"""The reference's KPM regression models (tests/test_kpm.py:7-22,138-173), built with the model
builder, and the fixture values extracted from tests/baseline_data/kpm/*.pbz
(tests/golden/kpm_ref_pbz.npz, written by tests/golden/extract_pbz.py)."""
import os

import numpy as np

from pybinding_amd import lattice as L
from pybinding_amd.repository import graphene, group6_tmd

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "kpm_ref_pbz.npz")


def golden(name):
    with np.load(GOLDEN) as f:
        return f[name]


def ldos_models():
    return {
        "graphene-pristine": [graphene.monolayer(), L.rectangle(15)],
        "graphene-pristine-oversized": [graphene.monolayer(), L.rectangle(20)],
        "graphene-const_potential": [graphene.monolayer(), L.rectangle(15), L.constant_potential(0.5)],
        "graphene-magnetic_field": [graphene.monolayer(), L.rectangle(15), graphene.constant_magnetic_field(1e3)],
        "mos2": [group6_tmd.monolayer_3band("MoS2"), L.rectangle(6)],
    }


def dos_models():
    return {
        "graphene-const_potential": [graphene.monolayer(), L.rectangle(25), L.constant_potential(0.5)],
        "graphene-magnetic_field": [graphene.monolayer(), L.rectangle(25), graphene.constant_magnetic_field(1e3)],
    }


def cond_models():
    return {
        "graphene-const_potential": [graphene.monolayer(), L.rectangle(20), L.constant_potential(0.5)],
        "graphene-magnetic_field": [graphene.monolayer(), L.rectangle(20), graphene.constant_magnetic_field(1e3)],
    }
