"""Seeded procedural scenes for the benchmark workloads, written as .nls +
OBJ + CSV files in the reference's scene schema (lib/numli/cg.hpp:1991-2474).

  box_room(dir)      closed room of 10 quads (20 tris) + two axis-aligned
                     blocks (24 tris), one 0.5 m x 0.5 m ceiling light with
                     a 6500 K blackbody spectrum, Lambertian surfaces with
                     seeded smooth reflectance spectra (sum of 3 Gaussians,
                     clipped to [0.05, 0.9], 81 samples on 380-780 nm);
                     256x256, hero 4 on 360-830 nm.
  icosphere_field(dir)
                     64 instances of one noise-displaced geodesic sphere
                     (frequency 28: 15,680 tris, ~1.0 M in total) with random
                     rotation, scale U[0.5,2] and position in a 20 m cube,
                     over a ground plane; materials uniform over Lambertian,
                     Blinn (α log-uniform 10-1000) and thin film (d U[200,800]
                     nm, film IOR U[1.33,1.6] over base 1.5); two area
                     lights; 1920x1080.

  box_room_films(dir) the box room with each of its 10 quads and 3 face groups
                     of each block given a Lambertian material of its own (16
                     albedos, 81 samples on 380-780 nm each), plus 64 small
                     thin-film spheres (d U[200,800]
                     nm, film IOR U[1.33,1.6] over base 1.5) on a jittered
                     8 x 8 grid in front of the back wall; 512x512.  The
                     workload of the path-replay adjoint.

  glossy_interior(dir)
                     indirectly lit glossy interior: a 10 m x 10 m x 4 m room
                     shell (Lambertian), 200 instances of a rounded box
                     (subdivision 14: 2,352 tris; ~470 k in total) with Blinn
                     lobes (α uniform 50-500, no transmission), and one small
                     light on a rimmed cabinet top in the far corner, facing
                     the ceiling, hidden from the room, which is lit by the
                     ceiling corner it lights (layout "plate": under the
                     ceiling above a plate instead); 1024x1024.  The workload
                     of the learned proposal.

Each function returns the path of the .nls file.  Same seed, same bytes.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

GRID = {"n": 95, "inf": 360.0, "sup": 830.0}


def blackbody(T, lam_nm):
    """Planck's law, relative spectral radiance."""
    h, c, k = 6.62607015e-34, 2.99792458e8, 1.380649e-23
    lam = lam_nm * 1e-9
    return (2 * h * c * c / lam ** 5) / np.expm1(h * c / (lam * k * T))


def write_csv(path, lam, val):
    with open(path, "w") as f:
        for l, v in zip(lam, val):
            f.write(f"{l:.6g},{v:.9g}\n")


def smooth_reflectance(rng, lam):
    """Sum of 3 Gaussians (μ~U[380,780], σ~U[20,80], a~U[0,1]), clipped."""
    v = np.zeros_like(lam, dtype=np.float64)
    for _ in range(3):
        mu, sig, a = rng.uniform(380, 780), rng.uniform(20, 80), rng.uniform(0, 1)
        v += a * np.exp(-0.5 * ((lam - mu) / sig) ** 2)
    return np.clip(v, 0.05, 0.9)


def write_obj(path, V, F, N=None):
    """OBJ with 1-based indices; vn per vertex when N is given."""
    with open(path, "w") as f:
        for v in V:
            f.write(f"v {v[0]:.7g} {v[1]:.7g} {v[2]:.7g}\n")
        if N is not None:
            for n in N:
                f.write(f"vn {n[0]:.7g} {n[1]:.7g} {n[2]:.7g}\n")
        for t in F:
            a, b, c = (int(x) + 1 for x in t)
            if N is not None:
                f.write(f"f {a}//{a} {b}//{b} {c}//{c}\n")
            else:
                f.write(f"f {a} {b} {c}\n")


def quad(p0, p1, p2, p3):
    """two triangles of the quad p0 p1 p2 p3 (counter-clockwise seen from the
    side the normal points to)"""
    return [p0, p1, p2, p3], [(0, 1, 2), (0, 2, 3)]


def _mesh_from_quads(quads):
    V, F = [], []
    for q in quads:
        vs, fs = quad(*q)
        base = len(V)
        V += vs
        F += [(a + base, b + base, c + base) for a, b, c in fs]
    V = np.array(V, dtype=np.float64)
    # flat per-face normals via duplicated vertices
    N = np.zeros_like(V)
    for a, b, c in F:
        n = np.cross(V[b] - V[a], V[c] - V[a])
        n /= np.linalg.norm(n)
        N[a] = N[b] = N[c] = n
    return V, np.array(F), N


def _block(x0, y0, x1, y1, h):
    z0, z1 = 0.0, h
    P = lambda x, y, z: (x, y, z)
    return [
        (P(x0, y0, z1), P(x1, y0, z1), P(x1, y1, z1), P(x0, y1, z1)),  # top (+z)
        (P(x0, y0, z0), P(x1, y0, z0), P(x1, y0, z1), P(x0, y0, z1)),  # -y
        (P(x1, y1, z0), P(x0, y1, z0), P(x0, y1, z1), P(x1, y1, z1)),  # +y
        (P(x0, y1, z0), P(x0, y0, z0), P(x0, y0, z1), P(x0, y1, z1)),  # -x
        (P(x1, y0, z0), P(x1, y1, z0), P(x1, y1, z1), P(x1, y0, z1)),  # +x
        (P(x0, y1, z0), P(x1, y1, z0), P(x1, y0, z0), P(x0, y0, z0)),  # bottom (-z)
    ]


def box_room(out_dir, seed=1, width=256, hero=4):
    os.makedirs(out_dir, exist_ok=True)
    rng = np.random.default_rng(seed)
    # reflectances on 81 samples, 380-780 nm in 5 nm steps (the grid's bins
    # outside that range receive nothing from the box filter: albedo 0)
    lam_ref = np.arange(380.0, 781.0, 5.0)
    lam_bb = np.arange(355.0, 841.0, 1.0)
    bb = blackbody(6500.0, lam_bb)
    write_csv(os.path.join(out_dir, "blackbody6500.csv"), lam_bb, bb / bb[lam_bb == 560.0][0])
    # room x,y in [-1,1], z in [0,2]; normals point into the room
    P = lambda x, y, z: (x, y, z)
    h = 0.25
    surfaces = {
        "floor": [(P(-1, -1, 0), P(1, -1, 0), P(1, 1, 0), P(-1, 1, 0))],
        "ceiling": [  # four strips around the light opening
            (P(-1, -1, 2), P(-1, 1, 2), P(-h, 1, 2), P(-h, -1, 2)),
            (P(h, -1, 2), P(h, 1, 2), P(1, 1, 2), P(1, -1, 2)),
            (P(-h, -1, 2), P(-h, -h, 2), P(h, -h, 2), P(h, -1, 2)),
            (P(-h, h, 2), P(-h, 1, 2), P(h, 1, 2), P(h, h, 2)),
        ],
        "back": [(P(-1, 1, 0), P(0, 1, 0), P(0, 1, 2), P(-1, 1, 2)),
                 (P(0, 1, 0), P(1, 1, 0), P(1, 1, 2), P(0, 1, 2))],
        "left": [(P(-1, -1, 0), P(-1, 1, 0), P(-1, 1, 2), P(-1, -1, 2))],
        "right": [(P(1, 1, 0), P(1, -1, 0), P(1, -1, 2), P(1, 1, 2))],
        "front": [(P(1, -1, 0), P(-1, -1, 0), P(-1, -1, 2), P(1, -1, 2))],
    }
    # fix winding so normals point inward: flip quads whose normal points out
    def inward(q):
        a, b, c = (np.array(x, dtype=float) for x in q[:3])
        n = np.cross(b - a, c - a)
        center = np.mean(np.array(q, dtype=float), axis=0)
        to_in = np.array([0, 0, 1.0]) - center
        return q if np.dot(n, to_in) > 0 else tuple(reversed(q))
    objects, materials = [], []
    names = list(surfaces) + ["block_tall", "block_short"]
    for nm in names:
        lam_spec = smooth_reflectance(rng, lam_ref)
        write_csv(os.path.join(out_dir, f"albedo_{nm}.csv"), lam_ref, lam_spec)
        materials.append({"name": f"mat_{nm}", "type": "lambertian",
                          "albedo": {"val": os.path.join(out_dir, f"albedo_{nm}.csv")}})
    for nm, qs in surfaces.items():
        V, F, N = _mesh_from_quads([inward(q) for q in qs])
        write_obj(os.path.join(out_dir, f"{nm}.obj"), V, F, N)
    V, F, N = _mesh_from_quads(_block(-0.7, 0.0, -0.1, 0.6, 1.2))
    write_obj(os.path.join(out_dir, "block_tall.obj"), V, F, N)
    V, F, N = _mesh_from_quads(_block(0.1, -0.5, 0.7, 0.1, 0.6))
    write_obj(os.path.join(out_dir, "block_short.obj"), V, F, N)
    for nm in names:
        objects.append({"name": nm, "type": "trimesh", "transform": {},
                        "material": f"mat_{nm}", "source": os.path.join(out_dir, f"{nm}.obj")})
    scene = {
        "spectral_grid": GRID,
        "hero_samples": hero,
        "camera": {"pos": [0, -0.95, 1.0], "look_at": [0, 1, 1.0], "up": [0, 0, 1],
                   "fov": 70, "ar": 1.0, "width": width},
        "lights": [{"name": "ceiling-light", "type": "quaddir", "specular": 1,
                    "transform": {"scale": [h, h, 1], "rotate": {"axis": [1, 0, 0], "degrees": 180},
                                  "translate": [0, 0, 2.0]},
                    "radiance": {"val": os.path.join(out_dir, "blackbody6500.csv"), "mult": 6.0}}],
        "objects": objects,
        "materials": materials,
    }
    path = os.path.join(out_dir, "box_room.nls")
    with open(path, "w") as f:
        json.dump(scene, f, indent=1)
    return path


def sheet_stack(out_dir, seed=5, width=64, n_sheets=8192, hero=4):
    """box_room with n_sheets parallel square sheets (0.8 m, facing the
    camera) stacked along the view axis between y = -0.3 and y = 0.5.  The
    BVH can separate them only along y, so every interior node a camera ray
    descends through has all of its children on the ray: the traversal
    pushes 7 entries per level (a stress case for the traversal stack)."""
    path = box_room(out_dir, seed=seed, width=width, hero=hero)
    with open(path) as f:
        scene = json.load(f)
    quads = []
    for i in range(n_sheets):
        y = -0.3 + 0.8 * i / n_sheets
        quads.append(((0.4, y, 0.6), (-0.4, y, 0.6), (-0.4, y, 1.4), (0.4, y, 1.4)))  # normal -y
    V, F, N = _mesh_from_quads(quads)
    write_obj(os.path.join(out_dir, "sheets.obj"), V, F, N)
    scene["objects"].append({"name": "sheets", "type": "trimesh", "transform": {}, "material": "mat_block_tall",
                             "source": os.path.join(out_dir, "sheets.obj")})
    # the tall block stands behind the sheets; drop it (its material stays)
    scene["objects"] = [o for o in scene["objects"] if o["name"] != "block_tall"]
    path = os.path.join(out_dir, "sheet_stack.nls")
    with open(path, "w") as f:
        json.dump(scene, f, indent=1)
    return path


def box_room_films(out_dir, seed=3, width=512, hero=4, n_films=64, film_freq=3, film_radius=0.06,
                   glossiness=32):
    os.makedirs(out_dir, exist_ok=True)
    rng = np.random.default_rng(seed)
    # 16 albedos on 81 samples, 380-780 nm in 5 nm steps: 16 x 81 CSV values
    lam_ref = np.arange(380.0, 781.0, 5.0)
    lam_bb = np.arange(355.0, 841.0, 1.0)
    bb = blackbody(6500.0, lam_bb)
    write_csv(os.path.join(out_dir, "blackbody6500.csv"), lam_bb, bb / bb[lam_bb == 560.0][0])
    P = lambda x, y, z: (x, y, z)
    h = 0.25
    quads = {  # the 10 quads of the room, normals into the room
        "floor": (P(-1, -1, 0), P(1, -1, 0), P(1, 1, 0), P(-1, 1, 0)),
        "ceiling0": (P(-1, -1, 2), P(-h, -1, 2), P(-h, 1, 2), P(-1, 1, 2)),
        "ceiling1": (P(h, -1, 2), P(1, -1, 2), P(1, 1, 2), P(h, 1, 2)),
        "ceiling2": (P(-h, -1, 2), P(h, -1, 2), P(h, -h, 2), P(-h, -h, 2)),
        "ceiling3": (P(-h, h, 2), P(h, h, 2), P(h, 1, 2), P(-h, 1, 2)),
        "back0": (P(-1, 1, 0), P(-1, 1, 2), P(0, 1, 2), P(0, 1, 0)),
        "back1": (P(0, 1, 0), P(0, 1, 2), P(1, 1, 2), P(1, 1, 0)),
        "left": (P(-1, -1, 0), P(-1, -1, 2), P(-1, 1, 2), P(-1, 1, 0)),
        "right": (P(1, 1, 0), P(1, 1, 2), P(1, -1, 2), P(1, -1, 0)),
        "front": (P(1, -1, 0), P(1, -1, 2), P(-1, -1, 2), P(-1, -1, 0)),
    }

    def inward(q):
        a, b, c = (np.array(x, dtype=float) for x in q[:3])
        n = np.cross(b - a, c - a)
        center = np.mean(np.array(q, dtype=float), axis=0)
        return q if np.dot(n, np.array([0, 0, 1.0]) - center) > 0 else tuple(reversed(q))

    groups = {nm: [inward(q)] for nm, q in quads.items()}
    for nm, (x0, y0, x1, y1, hh) in {"tall": (-0.7, 0.0, -0.1, 0.6, 1.2),
                                     "short": (0.1, -0.5, 0.7, 0.1, 0.6)}.items():
        top, my, py, mx, px, bot = _block(x0, y0, x1, y1, hh)
        groups[f"{nm}_top"] = [top]
        groups[f"{nm}_x"] = [mx, px, bot]
        groups[f"{nm}_y"] = [my, py]
    objects, materials = [], []
    for nm, qs in groups.items():
        write_csv(os.path.join(out_dir, f"albedo_{nm}.csv"), lam_ref, smooth_reflectance(rng, lam_ref))
        materials.append({"name": f"mat_{nm}", "type": "lambertian",
                          "albedo": {"val": os.path.join(out_dir, f"albedo_{nm}.csv")}})
        V, F, N = _mesh_from_quads(qs)
        write_obj(os.path.join(out_dir, f"{nm}.obj"), V, F, N)
        objects.append({"name": nm, "type": "trimesh", "transform": {}, "material": f"mat_{nm}",
                        "source": os.path.join(out_dir, f"{nm}.obj")})
    V, F, N = geodesic_sphere(film_freq, seed, amp=0.0)
    write_obj(os.path.join(out_dir, "film_sphere.obj"), V, F, N)
    side = int(round(math.sqrt(n_films)))
    for i in range(n_films):
        r, c = divmod(i, side)
        x = -0.8 + 1.6 * (c + 0.5 + rng.uniform(-0.25, 0.25)) / side
        z = 0.25 + 1.5 * (r + 0.5 + rng.uniform(-0.25, 0.25)) / max(side, (n_films + side - 1) // side)
        y = float(rng.uniform(0.7, 0.9))
        name = f"film{i}"
        materials.append({"name": name, "type": "thinfilm",
                          "film_ior": {"val": round(float(rng.uniform(1.33, 1.6)), 4)},
                          "base_ior": {"val": 1.5},
                          "thickness": round(float(rng.uniform(200, 800)), 2),
                          "glossiness": glossiness})
        objects.append({"name": f"sphere{i}", "type": "trimesh", "material": name,
                        "source": os.path.join(out_dir, "film_sphere.obj"),
                        "transform": {"scale": [film_radius] * 3, "translate": [float(x), y, float(z)]}})
    scene = {
        "spectral_grid": GRID,
        "hero_samples": hero,
        "camera": {"pos": [0, -0.95, 1.0], "look_at": [0, 1, 1.0], "up": [0, 0, 1],
                   "fov": 70, "ar": 1.0, "width": width},
        "lights": [{"name": "ceiling-light", "type": "quaddir", "specular": 1,
                    "transform": {"scale": [h, h, 1], "rotate": {"axis": [1, 0, 0], "degrees": 180},
                                  "translate": [0, 0, 2.0]},
                    "radiance": {"val": os.path.join(out_dir, "blackbody6500.csv"), "mult": 6.0}}],
        "objects": objects,
        "materials": materials,
    }
    path = os.path.join(out_dir, "box_room_films.nls")
    with open(path, "w") as f:
        json.dump(scene, f, indent=1)
    return path


def geodesic_sphere(freq, seed=2, amp=0.08):
    """Icosahedron faces subdivided freq x freq, projected to the unit sphere,
    displaced radially by a seeded sum of sinusoids; smooth vertex normals."""
    t = (1 + 5 ** 0.5) / 2
    ico_v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                      [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], float)
    ico_f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts, index, faces = [], {}, []

    def vid(p):
        p = p / np.linalg.norm(p)
        key = tuple(np.round(p, 9))
        if key not in index:
            index[key] = len(verts)
            verts.append(p)
        return index[key]

    for a, b, c in ico_f:
        A, B, Cc = ico_v[a], ico_v[b], ico_v[c]
        grid = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                grid[i, j] = vid(A + (B - A) * (i / freq) + (Cc - A) * (j / freq))
        for i in range(freq):
            for j in range(freq - i):
                faces.append((grid[i, j], grid[i + 1, j], grid[i, j + 1]))
                if i + j + 1 < freq:
                    faces.append((grid[i + 1, j], grid[i + 1, j + 1], grid[i, j + 1]))
    V = np.array(verts)
    rng = np.random.default_rng(seed)
    disp = np.zeros(len(V))
    for _ in range(6):
        k = rng.normal(size=3) * rng.uniform(2, 6)
        disp += np.sin(V @ k + rng.uniform(0, 2 * np.pi)) / 6
    V = V * (1 + amp * disp)[:, None]
    F = np.array(faces)
    # orient faces outward
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    fn = np.cross(b - a, c - a)
    flip = np.einsum("ij,ij->i", fn, a + b + c) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    fn = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    N = np.zeros_like(V)
    for k in range(3):
        np.add.at(N, F[:, k], fn)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    return V, F, N


def icosphere_field(out_dir, seed=2, n_inst=64, freq=28, width=1920, ar=1.77778, hero=4, material_p=None):
    """n_inst noise-displaced icospheres (uniform random rotation, scale
    U[0.5, 2], position in a 20 m cube) over a ground plane, two area lights.
    Materials: uniform over Lambertian / Blinn / thin film, or drawn with
    probabilities material_p = (p_lambertian, p_blinn, p_thinfilm)."""
    os.makedirs(out_dir, exist_ok=True)
    rng = np.random.default_rng(seed)
    V, F, N = geodesic_sphere(freq, seed)
    write_obj(os.path.join(out_dir, "icosphere.obj"), V, F, N)
    lam_ref = np.arange(355.0, 836.0, 5.0)
    lam_bb = np.arange(355.0, 841.0, 1.0)
    bb = blackbody(6500.0, lam_bb)
    write_csv(os.path.join(out_dir, "blackbody6500.csv"), lam_bb, bb / bb[lam_bb == 560.0][0])
    materials, objects = [], []
    write_csv(os.path.join(out_dir, "ground.csv"), lam_ref, smooth_reflectance(rng, lam_ref))
    materials.append({"name": "ground", "type": "lambertian",
                      "albedo": {"val": os.path.join(out_dir, "ground.csv")}})
    objects.append({"name": "ground", "type": "plane",
                    "transform": {"scale": [40, 40, 1], "translate": [0, 0, -0.5]}, "material": "ground"})
    for i in range(n_inst):
        kind = rng.integers(0, 3) if material_p is None else rng.choice(3, p=material_p)
        name = f"m{i}"
        if kind == 0:
            p = os.path.join(out_dir, f"albedo_{i}.csv")
            write_csv(p, lam_ref, smooth_reflectance(rng, lam_ref))
            materials.append({"name": name, "type": "lambertian", "albedo": {"val": p}})
        elif kind == 1:
            p = os.path.join(out_dir, f"kd_{i}.csv")
            write_csv(p, lam_ref, 0.7 * smooth_reflectance(rng, lam_ref))
            # even integer exponent: the reference evaluates powf(h·n, α) with
            # h·n < 0 at shading-normal silhouettes (cg.hpp:963), which is NaN
            # (a trap in the reference) for non-integer α
            alpha = 2 * round(math.exp(rng.uniform(math.log(10), math.log(1000))) / 2)
            materials.append({"name": name, "type": "blinn", "Kd": {"val": p}, "Ks": {"val": 0.3},
                              "ior": {"val": 1.5}, "glossiness": alpha})
        else:
            materials.append({"name": name, "type": "thinfilm",
                              "film_ior": {"val": round(float(rng.uniform(1.33, 1.6)), 4)},
                              "base_ior": {"val": 1.5},
                              "thickness": round(float(rng.uniform(200, 800)), 2),
                              "glossiness": 256})
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        objects.append({
            "name": f"sphere{i}", "type": "trimesh", "material": name,
            "source": os.path.join(out_dir, "icosphere.obj"),
            "transform": {"scale": [float(s) for s in [rng.uniform(0.5, 2)] * 3],
                          "rotate": {"axis": [float(x) for x in ax], "degrees": float(rng.uniform(0, 360))},
                          "translate": [float(rng.uniform(-10, 10)), float(rng.uniform(-10, 10)),
                                        float(rng.uniform(1.5, 21.5))]}})
    lights = []
    for j, (x, y) in enumerate([(-6.0, -4.0), (7.0, 5.0)]):
        lights.append({"name": f"area{j}", "type": "quaddir", "specular": 1,
                       "transform": {"scale": [2.5, 2.5, 1], "rotate": {"axis": [1, 0, 0], "degrees": 180},
                                     "translate": [x, y, 30.0]},
                       "radiance": {"val": os.path.join(out_dir, "blackbody6500.csv"), "mult": 2.0}})
    scene = {
        "spectral_grid": GRID,
        "hero_samples": hero,
        "camera": {"pos": [0, -38, 14], "look_at": [0, 0, 10], "up": [0, 0, 1], "fov": 45,
                   "ar": ar, "width": width},
        "lights": lights, "objects": objects, "materials": materials,
    }
    path = os.path.join(out_dir, "icosphere_field.nls")
    with open(path, "w") as f:
        json.dump(scene, f, indent=1)
    return path


def rounded_box(k, p=8.0):
    """cube faces subdivided k x k, pushed onto the superellipsoid
    |x|^p + |y|^p + |z|^p = 1 (rounded edges); smooth vertex normals"""
    verts, index, faces = [], {}, []

    def vid(v):
        key = tuple(np.round(v, 9))
        if key not in index:
            index[key] = len(verts)
            verts.append(v)
        return index[key]

    for axis in range(3):
        for sgn in (-1.0, 1.0):
            u_ax, v_ax = [a for a in range(3) if a != axis]
            grid = {}
            for i in range(k + 1):
                for j in range(k + 1):
                    v = np.zeros(3)
                    v[axis] = sgn
                    v[u_ax] = -1 + 2 * i / k
                    v[v_ax] = -1 + 2 * j / k
                    grid[i, j] = vid(v)
            for i in range(k):
                for j in range(k):
                    a, b, c, d = grid[i, j], grid[i + 1, j], grid[i + 1, j + 1], grid[i, j + 1]
                    faces += [(a, b, c), (a, c, d)]
    V = np.array(verts)
    V = V / (np.sum(np.abs(V) ** p, axis=1) ** (1.0 / p))[:, None]
    F = np.array(faces)
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    fn = np.cross(b - a, c - a)
    flip = np.einsum("ij,ij->i", fn, a + b + c) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    fn = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    N = np.zeros_like(V)
    for t in range(3):
        np.add.at(N, F[:, t], fn)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    return V, F, N


def glossy_interior(out_dir, seed=4, width=1024, hero=4, n_boxes=200, box_subdiv=14, layout="cabinet", plate=1.0,
                    plate_gap=0.35, box_ks=0.35):
    """The learned-proposal workload: a 10 x 10 x 4 m Lambertian room shell,
    200 instanced Blinn rounded boxes (exponent 50-500) and one small quad
    light hidden behind an occluder, so that the room is lit indirectly.

    layout "cabinet": the light sits on top of a 3 m cabinet in the far
    corner, facing the ceiling, inside a 0.4 m rim; no point of the room
    below the rim sees it, and the room is lit by the ceiling corner and the
    upper walls it lights (a bright patch seen by the whole room).
    layout "plate": the light hangs under the ceiling facing down onto a 2 m
    plate `plate_gap` below it; the room is lit through the slit between
    plate and ceiling and then by several diffuse bounces (broad light that
    leaves a learned proposal little to gain)."""
    os.makedirs(out_dir, exist_ok=True)
    rng = np.random.default_rng(seed)
    lam_ref = np.arange(355.0, 836.0, 5.0)
    lam_bb = np.arange(355.0, 841.0, 1.0)
    bb = blackbody(4500.0, lam_bb)
    write_csv(os.path.join(out_dir, "blackbody4500.csv"), lam_bb, bb / bb[lam_bb == 560.0][0])
    P = lambda x, y, z: (x, y, z)
    X, Y, H = 5.0, 5.0, 4.0
    shell = {
        "floor": (P(-X, -Y, 0), P(X, -Y, 0), P(X, Y, 0), P(-X, Y, 0)),
        "ceiling": (P(-X, -Y, H), P(-X, Y, H), P(X, Y, H), P(X, -Y, H)),
        "wall_n": (P(-X, Y, 0), P(X, Y, 0), P(X, Y, H), P(-X, Y, H)),
        "wall_s": (P(X, -Y, 0), P(-X, -Y, 0), P(-X, -Y, H), P(X, -Y, H)),
        "wall_e": (P(X, Y, 0), P(X, -Y, 0), P(X, -Y, H), P(X, Y, H)),
        "wall_w": (P(-X, -Y, 0), P(-X, Y, 0), P(-X, Y, H), P(-X, -Y, H)),
    }

    def inward(q):
        a, b, c = (np.array(x, dtype=float) for x in q[:3])
        n = np.cross(b - a, c - a)
        center = np.mean(np.array(q, dtype=float), axis=0)
        return q if np.dot(n, np.array([0, 0, H / 2]) - center) > 0 else tuple(reversed(q))

    objects, materials = [], []
    for nm, q in shell.items():
        write_csv(os.path.join(out_dir, f"albedo_{nm}.csv"), lam_ref,
                  0.5 + 0.4 * smooth_reflectance(rng, lam_ref))
        materials.append({"name": f"mat_{nm}", "type": "lambertian",
                          "albedo": {"val": os.path.join(out_dir, f"albedo_{nm}.csv")}})
        V, F, N = _mesh_from_quads([inward(q)])
        write_obj(os.path.join(out_dir, f"{nm}.obj"), V, F, N)
        objects.append({"name": nm, "type": "trimesh", "transform": {}, "material": f"mat_{nm}",
                        "source": os.path.join(out_dir, f"{nm}.obj")})
    write_csv(os.path.join(out_dir, "albedo_plate.csv"), lam_ref, 0.8 * np.ones_like(lam_ref))
    materials.append({"name": "mat_plate", "type": "lambertian",
                      "albedo": {"val": os.path.join(out_dir, "albedo_plate.csv")}})

    def block(name, x0, y0, x1, y1, h, z):
        V, F, N = _mesh_from_quads(_block(x0, y0, x1, y1, h))
        write_obj(os.path.join(out_dir, f"{name}.obj"), V, F, N)
        objects.append({"name": name, "type": "trimesh", "material": "mat_plate",
                        "source": os.path.join(out_dir, f"{name}.obj"), "transform": {"translate": [0, 0, z]}})

    # (The light's exponent is even: with an odd one the reference's quad
    # "dir" light gives negative radiance to points behind it,
    # cg.hpp:1625-1668.  Its back emits as its front.)
    if layout == "cabinet":
        # cabinet 1.2 m square, 3 m tall, 1 cm off the corner walls; a 5 cm
        # rim, 0.4 m tall, on its two open sides; the light 0.25 m above the
        # top, facing up
        top, rim, c0 = 3.0, 0.4, 1.2
        x0, y0, x1, y1 = X - c0, Y - c0, X - 0.01, Y - 0.01
        block("cabinet", x0, y0, x1, y1, top, 0.0)
        block("rim_s", x0, y0, x1, y0 + 0.05, rim, top)
        block("rim_w", x0, y0 + 0.05, x0 + 0.05, y1, rim, top)
        light_tf = {"scale": [0.15, 0.15, 1], "translate": [0.5 * (x0 + x1), 0.5 * (y0 + y1), top + 0.25]}
        keep_out = lambda x, y: x > x0 - 0.6 and y > y0 - 0.6
    elif layout == "plate":
        # the light hangs under the ceiling at (lx, ly), facing down; a plate
        # (half-width `plate`) `plate_gap` below the ceiling hides it
        lx, ly = 2.5, 2.5
        block("plate", lx - plate, ly - plate, lx + plate, ly + plate, 0.05, H - plate_gap)
        light_tf = {"scale": [0.15, 0.15, 1], "rotate": {"axis": [1, 0, 0], "degrees": 180},
                    "translate": [lx, ly, H - 0.01]}
        keep_out = lambda x, y: False
    else:
        raise ValueError(f"unknown layout {layout!r}")
    # glossy boxes
    V, F, N = rounded_box(box_subdiv)
    write_obj(os.path.join(out_dir, "rbox.obj"), V, F, N)
    placed = 0
    while placed < n_boxes:
        x, y = rng.uniform(-4.4, 4.4, 2)
        if (x + 4.5) ** 2 + (y + 4.5) ** 2 < 2.0 ** 2 or keep_out(x, y):  # the camera's corner, the cabinet
            continue
        sc = float(rng.uniform(0.15, 0.45))
        z = float(rng.uniform(sc, 3.0))
        kd = os.path.join(out_dir, f"kd_{placed}.csv")
        write_csv(kd, lam_ref, 0.6 * smooth_reflectance(rng, lam_ref))
        alpha = 2 * int(round(rng.uniform(50, 500) / 2))
        materials.append({"name": f"box{placed}", "type": "blinn", "Kd": {"val": kd}, "Ks": {"val": box_ks},
                          "ior": {"val": 1.5}, "glossiness": alpha})
        objects.append({"name": f"box{placed}", "type": "trimesh", "material": f"box{placed}",
                        "source": os.path.join(out_dir, "rbox.obj"),
                        "transform": {"scale": [sc, sc, sc],
                                      "rotate": {"axis": [0, 0, 1], "degrees": float(rng.uniform(0, 90))},
                                      "translate": [float(x), float(y), z]}})
        placed += 1
    scene = {
        "spectral_grid": GRID,
        "hero_samples": hero,
        "camera": {"pos": [-4.5, -4.5, 1.7], "look_at": [1.0, 1.0, 1.2], "up": [0, 0, 1], "fov": 70,
                   "ar": 1.0, "width": width},
        "lights": [{"name": "hidden", "type": "quaddir", "specular": 2, "transform": light_tf,
                    "radiance": {"val": os.path.join(out_dir, "blackbody4500.csv"), "mult": 400.0}}],
        "objects": objects,
        "materials": materials,
    }
    path = os.path.join(out_dir, "glossy_interior.nls")
    with open(path, "w") as f:
        json.dump(scene, f, indent=1)
    return path
