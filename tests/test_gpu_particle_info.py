"""save_particle_info dumps (save_particle_info.cpp:21-128, 271-280): the
files a scene writes hold the scene's fluid state in the reference's names,
separators, header and row order, each number as std::ostream prints a float
by default (%g, six significant digits)."""
import os

import numpy as np
import pytest

from apbf_amd import config as cfg

pytestmark = pytest.mark.gpu

HEADER = "center distance,boundary distance,kernel width,neighbor count,radius,target radius,inverse mass,x,y,z"


def g(x) -> str:
    return format(float(np.float32(x)), "g")


def test_particle_info_files_hold_the_state(device, tmp_path):
    from apbf_amd.scene import Scene
    p, st = cfg.small_water_drop(seed=2)
    s = Scene(p)
    s.set_state(st)
    for _ in range(3):
        s.update(p["dt"])
    folder = str(tmp_path / "particle_data")
    s.save_particle_info(folder)
    a = s.get_state()
    n = len(a["radius"])
    F = np.float32
    pos = a["position"][:, :3].astype(F) / F(262144.0)
    d = pos - np.array([0.0, 10.0, -60.0], F)
    center = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    bdr = a["boundary_distance"].astype(F) / F(262144.0)
    count = np.bincount(a["pairs"][:, 0], minlength=n)
    lists = {"centerDist.txt": center, "radius.txt": a["radius"], "kernelWidth.txt": a["kernel_width"],
             "targetRadius.txt": a["target_radius"], "boundaryDistance.txt": bdr}
    for name, v in lists.items():
        text = open(os.path.join(folder, name)).read()
        assert text == "".join(g(x) + ";" for x in v), name
    assert open(os.path.join(folder, "neighborCount.txt")).read() == "".join(f"{c};" for c in count)
    rows = open(os.path.join(folder, "data.csv")).read().split("\n")
    assert rows[0] == HEADER and rows[-1] == "" and len(rows) == n + 2
    order = np.argsort(center, kind="stable")
    for k in (0, 1, n // 2, n - 1):
        i = order[k]
        want = ",".join([g(center[i]), g(bdr[i]), g(a["kernel_width"][i]), str(count[i]), g(a["radius"][i]),
                         g(a["target_radius"][i]), g(a["inverse_mass"][i]), g(pos[i, 0]), g(pos[i, 1]),
                         g(pos[i, 2])])
        assert rows[1 + k] == want

    s.set_neighborhood_method("binary_search")
    s.save_measurements(str(tmp_path / "particle_svg"), 7, 1.5, 0.25, 0.75)
    m = open(tmp_path / "particle_svg" / "measurements_7.txt").read().split("\n")
    assert m[0] == f"Particle Count      : {n}"
    assert m[1] == f"Neighbor Pair Count : {len(a['pairs'])}"
    assert m[2:6] == ["Simulation Step     : 1.5ms", "Neighborhood Search : 0.25ms", "Constraint Solver   : 0.75ms",
                      "Neighbor Search Type: neighborhood_binary_search.h"]
    s.close()


@pytest.mark.parametrize("color", [0, 1, 2, 3])
def test_planar_svg_drawing(device, tmp_path, color):
    """save_as_svg's drawing of a planar scene (save_particle_info.cpp:130-270)
    equals the restatement from the downloaded state, character for
    character, in every colour mode (transfers in flight for mode 2)."""
    from apbf_amd.scene import Scene
    from svg_text import svg_text
    p, st = cfg.small_water_drop(seed=3, active=True, dims=2)
    s = Scene(p)
    s.set_state(st)
    for _ in range(2):
        s.update(p["dt"])
    a = s.get_state(pairs=False)
    assert color != 2 or len(a["t_source"]) > 0
    vmin, vmax = (-10.0, -8.0), (170.5, 120.25)
    folder = str(tmp_path / "particle_svg")
    s.save_svg(folder, 4, vmin, vmax, render_scale=0.7, max_expected_boundary_distance=30.0, color=color)
    got = open(os.path.join(folder, "particles_4.svg")).read()
    want = svg_text(a, p["boxes"], vmin, vmax, 0.7, 30.0, color, True)
    assert got == want
    s.save_svg(folder, 5, vmin, vmax, color=color, render_boxes=False)
    assert open(os.path.join(folder, "particles_5.svg")).read() == svg_text(a, p["boxes"], vmin, vmax, 0.7, 100.0,
                                                                               color, False)
    s.close()


def test_svg_drawing_skipped_in_3d(device, tmp_path):
    from apbf_amd.scene import Scene
    p, st = cfg.small_water_drop(seed=2)
    s = Scene(p)
    s.set_state(st)
    s.save_svg(str(tmp_path), 1, (0.0, 0.0), (10.0, 10.0))
    assert not os.path.exists(tmp_path / "particles_1.svg")
    s.close()
