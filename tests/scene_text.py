"""Scene texts used by several tests (written next to their data files)."""
import os

import numpy as np


def mixed_scene_text(d):
    from lambda_amd import scenes
    lam = np.arange(380.0, 781.0, 5.0)
    scenes.write_csv(os.path.join(d, "d65ish.csv"), lam, 1.0 + 0.3 * np.sin(lam / 40.0))
    scenes.write_csv(os.path.join(d, "ior.csv"), lam, 1.5 + 0.02 * (600.0 - lam) / 200.0)
    V, F, N = scenes.geodesic_sphere(4, seed=3)
    scenes.write_obj(os.path.join(d, "ball.obj"), V, F, N)
    return """{
"camera": {"pos": [0, -60, 12], "look_at": [0, 0, 12], "up": [0, 0, 1], "fov": 30, "ar": 1.5, "width": 48},
"lights": [
  {"name": "sl", "type": "sphere", "transform": {"translate": [0, 0, 18]},
   "radiance": {"val": "%(d)s/d65ish.csv", "mult": 2.0}},
  {"name": "ql", "type": "quaddir", "specular": 4,
   "transform": {"scale": [3, 3, 1], "rotate": {"axis": [1, 0, 0], "degrees": 180}, "translate": [8, 0, 23.5]},
   "radiance": {"val": 1.5}}],
"objects": [
  {"name": "s1", "type": "sphere", "transform": {"scale": [6, 6, 6], "translate": [0, 4, 6]}, "material": "reflective"},
  {"name": "s2", "type": "sphere", "transform": {"scale": [5, 5, 5], "translate": [9, -3, 5]}, "material": "refractive"},
  {"name": "ball", "type": "trimesh", "source": "%(d)s/ball.obj",
   "transform": {"scale": [4, 4, 4], "rotate": {"axis": [0, 1, 1], "degrees": 30}, "translate": [-9, -3, 5]},
   "material": "film"},
  {"name": "top", "type": "plane", "transform": {"scale": [32, 32, 1], "translate": [0, 0, 24],
   "rotate": {"axis": [1, 0, 0], "degrees": 180}}, "material": "grey"},
  {"name": "bot", "type": "plane", "transform": {"scale": [32, 32, 1]}, "material": "grey"},
  {"name": "back", "type": "plane", "transform": {"scale": [32, 32, 1], "rotate": {"axis": [1, 0, 0], "degrees": 90},
   "translate": [0, 20, 12]}, "material": "smooth"},
  {"name": "left", "type": "plane", "transform": {"scale": [32, 32, 1], "rotate": {"axis": [0, 1, 0], "degrees": 90},
   "translate": [-15, 0, 12]}, "material": "red"},
  {"name": "right", "type": "plane", "transform": {"scale": [32, 32, 1], "rotate": {"axis": [0, 1, 0], "degrees": -90},
   "translate": [15, 0, 12]}, "material": "green"}],
"materials": [
  {"name": "grey", "type": "lambertian", "albedo": {"val": 0.7}},
  {"name": "red", "type": "lambertian", "albedo": {"val": [0.7, 0.14, 0.14]}},
  {"name": "green", "type": "lambertian", "albedo": {"val": [0.14, 0.7, 0.14]}},
  {"name": "reflective", "type": "blinn", "Ks": {"val": 0.9}, "glossiness": 128},
  {"name": "refractive", "type": "blinn", "glossiness": 1024, "Kt": {"val": 0.9}, "ior": {"val": "%(d)s/ior.csv"}},
  {"name": "smooth", "type": "blinn", "Kd": {"val": 0.7}, "Ks": {"val": 0.3}, "glossiness": 64},
  {"name": "film", "type": "thinfilm", "glossiness": 512, "thickness": 700, "film_ior": {"val": 1.45},
   "base_ior": {"val": 1.6}}]
}""" % {"d": d}


def all_lights_scene_text(d):
    """the mixed scene with every light type of the reference: its sphere and
    quad "dir" lights plus a point, a directional and an ambient light
    (cg.hpp:1560-1622); the directional light enters through the open front"""
    import json
    s = json.loads(mixed_scene_text(d))
    s["lights"] += [
        {"name": "pt", "type": "point", "pos": [-6.0, -8.0, 16.0], "radiant_intensity": {"val": 900.0}},
        {"name": "sun", "type": "direction", "dir": [0.25, 1.0, -0.45], "radiant_intensity": {"val": 1.2}},
        {"name": "sky", "type": "ambient", "radiance": {"val": 0.08}},
    ]
    return json.dumps(s)


def crowd_scene_text(d, n=520):
    """the mixed scene plus n small meshes ("pebbles", a frequency-1
    geodesic sphere each, one shared grey material) scattered over the
    floor: the object table outgrows the shade kernels' LDS budget (n x 112
    bytes) while the spectra stay small, so the plain shade kernel stages
    the spectra alone"""
    import json
    from lambda_amd import scenes
    V, F, N = scenes.geodesic_sphere(1, seed=5)
    scenes.write_obj(os.path.join(d, "pebble.obj"), V, F, N)
    s = json.loads(mixed_scene_text(d))
    rng = np.random.default_rng(11)
    for k in range(n):
        x, y = rng.uniform(-14.0, 14.0), rng.uniform(-12.0, 18.0)
        r = float(rng.uniform(0.3, 0.8))
        s["objects"].append({"name": f"p{k}", "type": "trimesh", "source": os.path.join(d, "pebble.obj"),
                             "transform": {"scale": [r, r, r], "translate": [float(x), float(y), r]},
                             "material": "grey"})
    return json.dumps(s)
