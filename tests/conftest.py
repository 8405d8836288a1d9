import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def box_scene_path(tmp_path_factory):
    from lambda_amd import scenes
    return scenes.box_room(str(tmp_path_factory.mktemp("box")), seed=1, width=64)


@pytest.fixture(scope="session")
def box_scene(box_scene_path):
    import lambda_amd as la
    return la.load_nls(box_scene_path)


@pytest.fixture(scope="session")
def field_scene(tmp_path_factory):
    """icosphere field at a test size: 12 instances of a frequency-8 sphere"""
    import lambda_amd as la
    from lambda_amd import scenes
    p = scenes.icosphere_field(str(tmp_path_factory.mktemp("field")), seed=7, n_inst=12, freq=8,
                               width=96, ar=1.5)
    return la.load_nls(p)


@pytest.fixture(scope="session")
def mixed_scene(tmp_path_factory):
    """spheres, planes, a sphere light and a quad light, blinn + thin film +
    lambertian materials, hero 16 on the reference grid (box.nls-like)"""
    import lambda_amd as la
    from tests.scene_text import mixed_scene_text
    d = tmp_path_factory.mktemp("mixed")
    return la.Scene.from_text(mixed_scene_text(str(d)))


def random_rays(rng, n, lo, hi, tmax=3.4e38):
    import lambda_amd as la
    rays = np.zeros(n, dtype=la.RAY_DTYPE)
    rays["o"] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    rays["d"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["tmax"] = tmax
    rays["skip"] = la.LAM_NONE
    return rays


@pytest.fixture(scope="session")
def lights_scene(tmp_path_factory):
    """the mixed scene plus a point, a directional and an ambient light"""
    import lambda_amd as la
    from tests.scene_text import all_lights_scene_text
    d = tmp_path_factory.mktemp("lights")
    return la.Scene.from_text(all_lights_scene_text(str(d)))


@pytest.fixture(scope="session")
def crowd_scene(tmp_path_factory):
    """the mixed scene plus 520 small mesh instances (a large object table,
    small spectra)"""
    import lambda_amd as la
    from tests.scene_text import crowd_scene_text
    d = tmp_path_factory.mktemp("crowd")
    return la.Scene.from_text(crowd_scene_text(str(d)))
