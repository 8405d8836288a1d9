"""Host .nls loader: schema, ordering, defaults and quirks of
nl::cg::load (lib/numli/cg.hpp:1991-2474), and errors instead of crashes."""
import math
import os

import numpy as np
import pytest

import lambda_amd as la
from lambda_amd import scenes


def test_mixed_scene_order_and_lights(mixed_scene):
    s = mixed_scene
    names = [s.object_name(i) for i in range(s.desc.n_objects)]
    # scene objects in file order, then the light objects (loadLights, cg.hpp:2160-2183)
    assert names == ["s1", "s2", "ball", "top", "bot", "back", "left", "right", "sl", "ql"]
    mats = [s.material_name(i) for i in range(s.desc.n_materials)]
    assert mats[:7] == ["grey", "red", "green", "reflective", "refractive", "smooth", "film"]
    assert mats[7].startswith("emitter_") and mats[8].startswith("emitter_")
    m = s.materials()
    assert m[7].type == la.MAT_EMITTER and m[8].type == la.MAT_DIREMITTER
    assert m[8].specular == 4.0
    L = s.lights()
    assert [l.type for l in L] == [la.LIGHT_SPHERE, la.LIGHT_QUADDIR]
    assert L[0].object == 8 and L[1].object == 9
    objs = s.objects()
    assert objs[8].type == la.OBJ_SPHERE and objs[8].material == 7
    assert objs[9].type == la.OBJ_PLANE and objs[9].material == 8
    assert L[0].size == 1.0
    # quaddir inv_area = 1/(|col0| |col1|) of the plane transform: 1/9 here,
    # 4x the 1/area of the [-1,1]^2 quad it bounds (cg.hpp:2098-2099)
    assert L[1].inv_area == np.float32(1.0) / np.float32(9.0)


def test_blinn_init_and_defaults(mixed_scene):
    s = mixed_scene
    m = s.materials()
    spec = s.spectra()
    refl = m[3]  # Ks = 0.9 only
    assert refl.alpha == 128 and refl.p_d == 0.0 and refl.p_spec == 1.0
    d = s.desc
    delta = np.float32((d.grid.sup - d.grid.inf) / np.float32(d.grid.n - 1))
    # w_r = sqrt(Ks|Ks), operator| = sum(v*v) * Δ (cg.hpp:283-288, 931-940)
    ks = spec[refl.spec[1]]
    acc = np.float32(0)
    for v in ks:
        acc = np.float32(acc + np.float32(v * v))
    assert refl.w_r == np.float32(math.sqrt(np.float32(acc * delta)))
    # default ior 1.54 and Le/Kd/Kt 0 (cg.hpp:2366-2371)
    assert np.all(spec[refl.spec[4]] == np.float32(1.54))
    assert np.all(spec[refl.spec[0]] == 0) and np.all(spec[refl.spec[2]] == 0)
    smooth = m[5]
    assert smooth.alpha == 64 and 0 < smooth.p_d < 1 and abs(smooth.p_d + smooth.p_spec - 1) < 1e-6
    film = m[6]
    assert film.type == la.MAT_THINFILM and film.d == 700 and film.alpha == 512
    assert np.all(spec[film.spec[0]] == np.float32(1.45)) and np.all(spec[film.spec[1]] == np.float32(1.6))


def test_thinfilm_defaults(tmp_path):
    text = _minimal(materials='[{"name":"m","type":"thinfilm","film_ior":{"val":1.4}}]')
    s = la.Scene.from_text(text)
    m = s.materials()[0]
    assert m.d == 400 and m.alpha == 1024  # cg.hpp:2379-2381
    assert np.all(s.spectra()[m.spec[1]] == np.float32(1.54))


def test_camera(mixed_scene):
    c = mixed_scene.desc.camera
    assert c.width == 48 and c.height == 32  # ceil(48/1.5)
    z = np.array([0, -60, 0], np.float32) / np.float32(60)
    assert np.allclose(list(c.z), z)
    # x = up ^ z, unnormalised (cg.hpp:2058)
    assert np.allclose(list(c.x), np.cross([0, 0, 1], z))
    delta = 2 * 1.0 * math.tan(0.5 * math.radians(30)) / 32
    assert abs(c.delta - delta) < 1e-7 and abs(c.h - delta * 32) < 1e-6 and abs(c.w - delta * 48) < 1e-6


def test_transform_rotation_is_transposed_rodrigues(mixed_scene):
    """transform::rotate writes the transpose of R(u,θ) (cg.hpp:452-480), so
    'rotate 90° about x' maps e_y to -e_z; the composed matrix is
    T·R·S (cg.hpp:2018, 490-496)."""
    back = mixed_scene.objects()[5]
    M = np.array(back.T.M, dtype=np.float64).reshape(3, 4)
    assert np.allclose(M[:, 1], [0, 0, -32], atol=1e-4)
    assert np.allclose(M[:, 2], [0, 1, 0], atol=1e-6)
    assert np.allclose(M[:, 3], [0, 20, 12])
    Mi = np.array(back.T.Minv, dtype=np.float64).reshape(3, 4)
    A = np.vstack([M, [0, 0, 0, 1]]) @ np.vstack([Mi, [0, 0, 0, 1]])
    assert np.allclose(A, np.eye(4), atol=1e-5)


def test_flattened_triangles_are_world_space(mixed_scene):
    s = mixed_scene
    ball = s.objects()[2]
    tv = s.tri_v()
    assert len(tv) == s.n_tris and np.all(s.tri_obj() == 2)
    M = np.array(ball.T.M, np.float64).reshape(3, 4)
    # every vertex is M * (v, 1) of a unit-ish sphere scaled by 4 around (-9,-3,5)
    c = tv.reshape(-1, 3) - M[:, 3]
    r = np.linalg.norm(c, axis=1)
    assert 3.0 < r.min() and r.max() < 5.0
    assert np.array_equal(s.tri_prim(), np.arange(s.n_tris, dtype=np.uint32))


def test_meshes_shared_by_path(tmp_path):
    """objects naming the same OBJ path share one mesh (loaded once,
    cg.hpp:2295-2303); each instance is flattened with its own transform"""
    V, F, N = scenes.geodesic_sphere(1, seed=1)
    scenes.write_obj(str(tmp_path / "a.obj"), V, F, N)
    scenes.write_obj(str(tmp_path / "b.obj"), V[:3] * 2.0, np.array([[0, 1, 2]]), N[:3])
    objs = ",".join([
        f'{{"name":"a{i}","type":"trimesh","source":"{tmp_path / "a.obj"}","material":"m",'
        f'"transform":{{"translate":[{3 * i},0,0]}}}}' for i in range(3)] +
        [f'{{"name":"b","type":"trimesh","source":"{tmp_path / "b.obj"}","material":"m","transform":{{}}}}'])
    s = la.Scene.from_text(_minimal(objects=f"[{objs}]"))
    meshes = [o.mesh for o in s.objects()]
    assert meshes[0] == meshes[1] == meshes[2] != meshes[3]
    nf = len(F)
    assert s.n_tris == 3 * nf + 1
    tv = s.tri_v().reshape(-1, 3, 3)
    obj = s.tri_obj()
    for i in (1, 2):  # instance i is instance 0 moved by (3i, 0, 0), vertex for vertex
        assert np.allclose(tv[obj == i] - tv[obj == 0], [3.0 * i, 0.0, 0.0], atol=1e-5)


def _minimal(lights='[{"name":"l","type":"point","pos":[0,0,5],"radiant_intensity":{"val":1}}]',
             objects='[{"name":"p","type":"plane","transform":{},"material":"m"}]',
             materials='[{"name":"m","type":"lambertian","albedo":{"val":0.5}}]'):
    return ('{"camera":{"pos":[0,-5,1],"look_at":[0,0,0],"up":[0,0,1],"fov":40,"width":16},'
            f'"lights":{lights},"objects":{objects},"materials":{materials}}}')


@pytest.mark.parametrize("text,code", [
    ("{ not json", la.LAM_E_PARSE),
    (_minimal(lights="[]"), la.LAM_E_PARSE),  # the reference divides by zero (cg.hpp:1554)
    (_minimal(objects='[{"name":"p","type":"plane","transform":{},"material":"nope"}]'), la.LAM_E_PARSE),
    (_minimal(objects='[{"name":"p","type":"cone","transform":{},"material":"m"}]'), la.LAM_E_PARSE),
    (_minimal(objects='[{"name":"p","type":"trimesh","transform":{},"material":"m","source":"/nonexistent.obj"}]'),
     la.LAM_E_PARSE),
    (_minimal(lights='[{"name":"l","type":"sphere","transform":{"scale":[1,2,1]},"radiance":{"val":1}}]'),
     la.LAM_E_PARSE),
    (_minimal(materials='[{"name":"m","type":"blinn"}]'), la.LAM_E_PARSE),  # no lobe: assert in the reference
])
def test_load_errors(text, code):
    with pytest.raises(la.LambdaError) as e:
        la.Scene.from_text(text)
    assert e.value.code == code


def test_missing_file():
    with pytest.raises(la.LambdaError) as e:
        la.load_nls("/nonexistent/scene.nls")
    assert e.value.code == la.LAM_E_IO


def test_unknown_material_type_is_skipped():
    # MaterialType::NONE is not added (cg.hpp:2409): later names still resolve
    s = la.Scene.from_text(_minimal(materials='[{"name":"x","type":"velvet"},{"name":"m","type":"lambertian",'
                                              '"albedo":{"val":0.5}}]'))
    assert s.desc.n_materials == 1 and s.material_name(0) == "m"


def test_group_ignores_own_transform():
    s = la.Scene.from_text(_minimal(objects='[{"name":"g","type":"group","transform":{"translate":[5,5,5]},'
                                            '"children":[{"name":"p","type":"plane","transform":{},"material":"m"}]}]'))
    o = s.objects()[0]
    assert np.allclose(np.array(o.T.M).reshape(3, 4), np.eye(4)[:3])


def test_emitters_share_material_by_name():
    L = ('[{"name":"a","type":"sphere","transform":{"translate":[0,0,3]},"radiance":{"val":2}},'
         '{"name":"b","type":"sphere","transform":{"translate":[1,0,3]},"radiance":{"val":2}}]')
    s = la.Scene.from_text(_minimal(lights=L))
    objs = s.objects()
    assert s.desc.n_materials == 2 and objs[1].material == objs[2].material == 1


def test_spectrum_box_filter_exact_on_linear(tmp_path):
    """coefficientλ(sampledλ) sums trapezoids of the sample segments over
    [l0, l0+Δ] and divides by Δ (cg.hpp:252-272).  With samples on the bin
    edges this is the exact mean of a piecewise-linear input: a + b(l0+Δ/2)
    on a 5 nm grid sampled every 1 nm."""
    lam = np.arange(300.0, 901.0, 1.0)
    a, b = 0.2, 0.001
    p = tmp_path / "lin.csv"
    scenes.write_csv(str(p), lam, a + b * lam)
    text = _minimal(materials='[{"name":"m","type":"lambertian","albedo":{"val":"%s"}}]' % p)
    text = text[:-1] + ',"spectral_grid":{"n":95,"inf":360,"sup":830}}'
    s = la.Scene.from_text(text)
    d = s.desc
    assert (d.grid.sup - d.grid.inf) / (d.grid.n - 1) == 5.0
    l0 = d.grid.inf + np.arange(d.grid.n) * 5.0
    v = s.spectra()[s.materials()[0].spec[0]]
    np.testing.assert_allclose(v, a + b * (l0 + 2.5), rtol=3e-6)
    # reference grid (Δ = 320/79, not aligned): a constant input stays constant
    p2 = tmp_path / "const.csv"
    scenes.write_csv(str(p2), np.arange(300.0, 901.0, 7.0), np.full(86, 0.375))
    s3 = la.Scene.from_text(_minimal(materials='[{"name":"m","type":"lambertian","albedo":{"val":"%s"}}]' % p2))
    np.testing.assert_allclose(s3.spectra()[s3.materials()[0].spec[0]], 0.375, rtol=2e-5)
    # a constant spectrum stays constant; mult scales
    s2 = la.Scene.from_text(_minimal(materials='[{"name":"m","type":"lambertian","albedo":{"val":0.25,"mult":2}}]'))
    assert np.all(s2.spectra()[s2.materials()[0].spec[0]] == np.float32(0.5))


@pytest.mark.parametrize("grid", ['', ',"spectral_grid":{"n":95,"inf":360,"sup":830}'])
def test_spectrum_grad_to_samples_is_the_filter_transpose(tmp_path, grid):
    """The loader's box filter is linear in the CSV values, so its Jacobian
    column k is the filtered unit vector e_k.  The transposed map applied to
    a random coefficient gradient must equal Jᵀ·g built from 81 loads, on the
    95-bin 5 nm grid (bins outside the samples stay 0) and the reference's
    unaligned 80-bin grid; mult scales it."""
    lam = np.arange(380.0, 781.0, 5.0)
    cols = []
    for k in range(len(lam)):
        p = tmp_path / f"e{k}.csv"
        e = np.zeros(len(lam))
        e[k] = 1.0
        scenes.write_csv(str(p), lam, e)
        text = _minimal(materials='[{"name":"m","type":"lambertian","albedo":{"val":"%s"}}]' % p)
        if grid:
            text = text[:-1] + grid + "}"
        s = la.Scene.from_text(text)
        cols.append(s.spectra()[s.materials()[0].spec[0]].astype(np.float64))
    J = np.stack(cols, axis=1)  # (grid.n, 81)
    g = np.random.default_rng(3).normal(size=(2, s.desc.grid.n)).astype(np.float32)
    got = la.spectrum_grad_to_samples(s.desc.grid, lam, g)
    np.testing.assert_allclose(got, g.astype(np.float64) @ J, rtol=1e-5, atol=1e-6 * np.abs(g).max())
    got2 = la.spectrum_grad_to_samples(s.desc.grid, lam, g[0], mult=2.0)
    np.testing.assert_array_equal(got2, 2 * got[0])
    with pytest.raises(la.LambdaError):
        la.spectrum_grad_to_samples(s.desc.grid, lam[::-1], g[0])


def test_rgb_spectrum_is_bt709_basis_sum():
    s = la.Scene.from_text(_minimal(materials='[{"name":"m","type":"lambertian","albedo":{"val":[1,0,0]}}]'))
    s2 = la.Scene.from_text(_minimal(materials='[{"name":"m","type":"lambertian","albedo":{"val":[0,1,0]}}]'))
    s3 = la.Scene.from_text(_minimal(materials='[{"name":"m","type":"lambertian","albedo":{"val":[1,1,0]}}]'))
    r = s.spectra()[s.materials()[0].spec[0]]
    g = s2.spectra()[s2.materials()[0].spec[0]]
    rg = s3.spectra()[s3.materials()[0].spec[0]]
    np.testing.assert_allclose(rg, r + g, rtol=1e-6, atol=1e-7)
    assert r[-1] > r[0] and g.max() > 0.3


def test_reference_grid_default_and_cmf(mixed_scene):
    d = mixed_scene.desc
    assert (d.grid.n, d.grid.inf, d.grid.sup) == (80, 400.0, 720.0) and mixed_scene.hero == 16
    cmf = mixed_scene.cmf()
    # CIE ȳ peaks near 555 nm; the shipped xyz_y table integrates to 113.0
    # (5 nm sum), and its box-filtered coefficients keep that integral
    l0 = 400 + np.arange(80) * (320 / 79)
    assert abs(l0[np.argmax(cmf[1])] - 555) < 8
    assert 100 < cmf[1].sum() * (320 / 79) < 120


def test_obj_without_normals_uses_dummy_normal(tmp_path):
    p = tmp_path / "t.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nf 1 2 3\nf 1 3 4 2\n")
    s = la.Scene.from_text(_minimal(objects='[{"name":"t","type":"trimesh","transform":{},"material":"m",'
                                            '"source":"%s"}]' % p))
    # fast_obj's dummy normal (0,0,1) for faces without vn; quads fan-split
    assert s.n_tris == 3
    n = la._carr(s.desc.tri_n, 9 * 3, np.float32).reshape(3, 3, 3)
    assert np.all(n[..., 2] == 1.0) and np.all(n[..., :2] == 0.0)
