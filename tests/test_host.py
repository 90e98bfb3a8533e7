"""Host-side tests that need no GPU: the product's rod builders against the
oracle's, the workload builders, and the C-ABI library's exported surface."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import oracle
import ropephysics_amd as rp
from ropephysics_amd import _abi, scenes
from helpers import REFERENCE_MATERIAL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ropephysics.h")


def _same(a: dict, b: dict):
    bad = [k for k in a if k in b and not np.array_equal(np.asarray(a[k], np.float64), np.asarray(b[k], np.float64))]
    assert not bad, bad


@pytest.mark.parametrize("tip", [False, True])
def test_straight_builder_bitwise_equals_oracle(tip):
    mat = rp.RodMaterial(3e6, 0.4, 1100.0, 4e-3)
    r = rp.make_straight_rod(37, 0.9, mat, (0.1, -0.3, 0.25), (0.3, -1.0, 0.2))
    rp.clamp_root_exact(r)
    if tip:
        rp.clamp_end_exact(r, True)
    o = oracle.Rod.straight(37, 0.9, mat.as_tuple(), (0.1, -0.3, 0.25), (0.3, -1.0, 0.2))
    o.clamp_end(False)
    if tip:
        o.clamp_end(True)
    _same(o.get(), r.arrays())


def test_centerline_builder_bitwise_equals_oracle():
    pts = [(math.cos(t) * 0.3, math.sin(t) * 0.3, 0.02 * t) for t in np.linspace(0, 7, 57)]
    r = rp.make_rod_from_centerline(pts, REFERENCE_MATERIAL, (0, 1, 0), 0.6)
    o = oracle.Rod.centerline(pts, REFERENCE_MATERIAL.as_tuple(), (0, 1, 0), 0.6)
    _same(o.get(), r.arrays())
    rp.set_rest_from_current(r)
    o.set_rest_from_current()
    _same(o.get(), r.arrays())


def test_uniform_workload_is_seeded_and_shaped():
    a = scenes.uniform_floor_batch(num_rods=64, steps=5)
    b = scenes.uniform_floor_batch(num_rods=64, steps=5)
    assert np.array_equal(a.x0, b.x0) and np.array_equal(a.youngs_scale, b.youngs_scale)
    assert a.x0.shape == (64 * 65, 3)
    assert np.all((a.youngs_scale >= 0.4) & (a.youngs_scale <= 4.0))
    assert np.all((a.friction_scale >= 0.3) & (a.friction_scale <= 0.9))
    # The clamped root particle is not jittered; free ones are, by about 1 mm.
    x = a.x0.reshape(64, 65, 3)
    assert np.all(x[:, 0] == a.prototype.x[0])
    d = (x - a.prototype.x[None])[:, 1:]
    assert 0.7e-3 < d.std() < 1.3e-3
    assert a.segment_substeps == 64 * 64 * 10 * 5


def _declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_abi.LIB_PATH)
    syms = _declared_symbols()
    assert len(syms) >= 14
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # And the Python binding knows every one of them.
    assert set(syms) == set(_abi.SIGNATURES)


def test_tile_configs_listed():
    cfg = rp.tile_configs()
    assert len(cfg) >= 4 and all(isinstance(v, str) and v for v in cfg.values())
    assert _abi.lib().rp_tile_config_name(0) is None


def test_create_without_device_fails_loudly():
    """With no usable GPU the library reports a device error; it never falls
    back to a CPU path."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU device node is present")
    rod = rp.make_straight_rod(8, 1.0, REFERENCE_MATERIAL)
    with pytest.raises(rp.RopePhysicsError) as e:
        rp.Batch(rod, 2)
    assert e.value.code == _abi.RP_ERR_DEVICE
    with pytest.raises(rp.RopePhysicsError):
        rp.device_info()
