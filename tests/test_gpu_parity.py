"""Device (libsamos_hip.so) against the CPU oracle on the same inputs.

Bars:
  pair sets                      bit-exact (sorted half-pair rows)
  forces, torques                max_i |dF_i| / max_i |F_i| <= 1e-12 (FP64)
  energies                       relative 1e-12
  noise-off trajectories         |dx| <= 1e-9 after 1000 steps (FP64)
  noise-on, same Philox stream   |dx| <= 1e-9 after 50 steps (device sin/cos/log
                                 and the host libm differ in the last bit)
  FP32 pair arithmetic           forces <= 1e-6 relative to FP64
"""
import math
import os

import numpy as np
import pytest

import oracle
from tests import scenarios as sc

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def both(gpu, st):
    return sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)


def forces(s):
    return np.stack([s.get("fx"), s.get("fy"), s.get("fz")], 1)


def torques(s):
    return np.stack([s.get("tau_x"), s.get("tau_y"), s.get("tau_z")], 1)


def positions(s):
    return np.stack([s.get("x"), s.get("y"), s.get("z")], 1)


def directors(s):
    return np.stack([s.get("nx"), s.get("ny"), s.get("nz")], 1)


def pos_err(a, b, box, periodic):
    d = a - b
    if periodic:
        for k in range(3):
            d[:, k] = sc.min_image(d[:, k], box[k])
    return float(np.abs(d).max())


# ------------------------------------------------------------ pair sets ---
def test_pairs_c1_plane(gpu):
    st = sc.plane_disks(4096, phi=0.8, seed=11)
    d, o = both(gpu, st)
    d.nlist(2.4, 0.3)
    o.nlist(2.4, 0.3)
    info = d.nlist_info()
    assert info["uses_cells"] == 1 and (info["ncx"], info["ncy"], info["ncz"]) == o.cell_dims()
    assert np.array_equal(d.pairs(), o.pairs())
    assert np.array_equal(d.get("coordination"), o.get("coordination"))


@pytest.mark.parametrize("name", ["plane", "sphere", "lj", "relax"])
def test_pairs_example_inputs(gpu, name):
    import tests.golden.make_golden as mg
    inp = dict(np.load(os.path.join(GOLD, f"inputs_{name}.npz")))
    gold = dict(np.load(os.path.join(GOLD, f"golden_{name}.npz")))
    case = mg.CASES[name]
    d = gpu.System(len(inp["x"]), case["box"], case["periodic"])
    for k, v in inp.items():
        d.set(k, v)
    d.nlist(*case["nlist"])
    assert np.array_equal(d.pairs(), gold["pairs"])


def test_pairs_n_squared_thin_box(gpu):
    rng = np.random.default_rng(5)
    n = 900
    st = {"n": n, "box": (30.0, 30.0, 4.0), "periodic": True,
          "cols": {"x": rng.uniform(-15, 15, n), "y": rng.uniform(-15, 15, n)}}
    d, o = both(gpu, st)
    d.nlist(2.4, 0.3)
    o.nlist(2.4, 0.3)
    assert d.nlist_info()["uses_cells"] == 0
    assert np.array_equal(d.pairs(), o.pairs())


def test_pairs_two_cells_per_side_duplicates(gpu):
    rng = np.random.default_rng(9)
    n, L = 60, 7.0
    st = {"n": n, "box": (L, L, L), "periodic": True, "cols": {k: rng.uniform(-L / 2, L / 2, n) for k in "xyz"}}
    d, o = both(gpu, st)
    d.nlist(2.4, 0.6)
    o.nlist(2.4, 0.6)
    assert np.array_equal(d.pairs(), o.pairs())


def test_pairs_boundary_lattice(gpu):
    L = 27.0
    g = np.arange(-L / 2, L / 2, 1.35)
    x, y = np.meshgrid(g, g)
    st = {"n": x.size, "box": (L, L, 10.0), "periodic": True, "cols": {"x": x.ravel(), "y": y.ravel()}}
    d, o = both(gpu, st)
    d.nlist(2.4, 0.3)
    o.nlist(2.4, 0.3)
    assert np.array_equal(d.pairs(), o.pairs())


def test_empty_and_single_particle(gpu):
    for n in (0, 1):
        d = gpu.System(n, (20.0, 20.0, 20.0))
        d.nlist(2.4, 0.3)
        assert len(d.pairs()) == 0
        d.pair_soft()
        d.constraint("plane")
        d.set_timestep(0.01)
        d.integrator_brownian(v0=0.1, nu=0.0)
        d.run(3)


def test_pairs_after_run_and_rebuild(gpu):
    st = sc.plane_disks(2000, phi=0.7, seed=21)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.constraint("plane")
        s.set_timestep(0.01)
        s.integrator_brownian(v0=1.0, nu=0.0)
        s.run(60)
    assert d.nlist_info()["builds"] > 1
    assert np.array_equal(d.pairs(), o.pairs())


# -------------------------------------------------------- forces/torques ---
@pytest.mark.parametrize("name", ["plane", "sphere", "lj", "relax"])
def test_forces_example_inputs(gpu, name):
    import tests.golden.make_golden as mg
    inp = dict(np.load(os.path.join(GOLD, f"inputs_{name}.npz")))
    gold = dict(np.load(os.path.join(GOLD, f"golden_{name}.npz")))
    case = mg.CASES[name]
    d = gpu.System(len(inp["x"]), case["box"], case["periodic"])
    for k, v in inp.items():
        d.set(k, v)
    d.nlist(*case["nlist"])
    if "soft" in case:
        d.pair_soft(*case["soft"])
    if "lj" in case:
        eps, sig, rc, shifted = case["lj"]
        d.pair_lj(eps, sig, rc, shifted)
    if "polar" in case:
        d.align_polar(case["polar"])
    if "nematic" in case:
        d.align_nematic(case["nematic"])
    d.constraint(*case["constraint"])
    e = d.compute_forces()
    F = forces(d)
    G = np.stack([gold["fx"], gold["fy"], gold["fz"]], 1)
    assert sc.rel_force_err(F, G) <= 1e-12
    if "polar" in case or "nematic" in case:
        T = torques(d)
        H = np.stack([gold["tau_x"], gold["tau_y"], gold["tau_z"]], 1)
        assert sc.rel_force_err(T, H) <= 1e-12
    ep = e["soft"] + e["lj"]
    ea = e["polar"] + e["nematic"]
    assert ep == pytest.approx(float(gold["energy_pair"]), rel=1e-12, abs=1e-12)
    assert ea == pytest.approx(float(gold["energy_align"]), rel=1e-12, abs=1e-12)


def test_forces_c1_soft_polar(gpu):
    st = sc.plane_disks(4096, phi=0.8, seed=12)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.align_polar(0.1, 2.4)
        s.constraint("plane")
    e = d.compute_forces()
    eo = o.compute_forces()
    assert sc.rel_force_err(forces(d), forces(o)) <= 1e-12
    assert sc.rel_force_err(torques(d), torques(o)) <= 1e-12
    assert e["soft"] == pytest.approx(eo["pair"], rel=1e-12)
    assert e["polar"] == pytest.approx(eo["align"], rel=1e-12)


def test_per_type_parameters(gpu):
    rng = np.random.default_rng(4)
    st = sc.plane_disks(1500, phi=0.7, seed=13)
    st["cols"]["type"] = rng.integers(1, 4, 1500).astype(np.int32)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.8, 0.3)
        s.pair_soft(1.0, 2.0)
        s.pair_type_param("soft", 1, 2, 3.0, 2.2)
        s.pair_type_param("soft", 3, 3, 0.5, 1.6)
        s.pair_lj(1.0, 1.0, 2.5, shifted=True)
        s.pair_type_param("lj", 2, 3, 0.7, 1.1, 2.0)
        s.align_nematic(0.4, 2.4)
        s.align_type_param("nematic", 1, 3, 0.9, 2.0)
        s.constraint("plane")
    d.compute_forces()
    o.compute_forces()
    assert sc.rel_force_err(forces(d), forces(o)) <= 1e-12
    assert sc.rel_force_err(torques(d), torques(o)) <= 1e-12


def test_component_calls_match_loop_head(gpu):
    st = sc.plane_disks(1200, phi=0.8, seed=14)
    d = sc.load(gpu.System, st)
    d.nlist(2.4, 0.3)
    d.pair_soft(1.0, 2.0)
    d.align_polar(0.3, 2.4)
    d.constraint("plane")
    e = d.compute_forces()
    F, T = forces(d), torques(d)
    z = np.zeros(st["n"])
    for k in ["fx", "fy", "fz", "tau_x", "tau_y", "tau_z"]:
        d.set(k, z)
    es = d.pair_compute("soft")
    ea = d.align_compute("polar")
    assert np.array_equal(forces(d), F) and np.array_equal(torques(d), T)
    assert es == e["soft"] and ea == e["polar"]


# --------------------------------------------------------- trajectories ---
def test_trajectory_c1_brownian_polar_noise_off(gpu):
    st = sc.plane_disks(4096, phi=0.8, seed=15)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.align_polar(0.1, 2.4)
        s.constraint("plane")
        s.set_timestep(0.01)
        s.integrator_brownian(v0=0.1, nu=0.0, mu=1.0, seed=1)
        s.run(999)
    assert pos_err(positions(d), positions(o), st["box"], True) <= 1e-9
    assert np.abs(directors(d) - directors(o)).max() <= 1e-9
    assert np.array_equal(d.get("ix"), o.get("ix")) and np.array_equal(d.get("iy"), o.get("iy"))


def test_trajectory_sphere_nematic_noise_off(gpu):
    st = sc.sphere_disks(2000, phi=1.0, seed=16)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(10.0, 2.0)
        s.align_nematic(0.5, 2.4)
        s.constraint("sphere", st["R"])
        s.set_timestep(0.005)
        s.integrator_brownian(v0=0.5, nu=0.0, seed=2)
        s.run(999)
    assert pos_err(positions(d), positions(o), st["box"], False) <= 1e-9
    assert np.abs(directors(d) - directors(o)).max() <= 1e-9


@pytest.mark.parametrize("flags", [0, 1, 2])
def test_trajectory_langevin_lj_noise_off(gpu, flags):
    st = sc.lj_lattice(48, rho=0.7, seed=17)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.5, 0.3)
        s.pair_lj(1.0, 1.0, 2.5, shifted=True)
        s.constraint("plane")
        s.set_timestep(0.002)
        s.integrator_langevin(gamma=1.0, seed=3, temp=(0, 0.0, 1.0, 1))
    d.run(999, flags)
    o.run(999)
    assert pos_err(positions(d), positions(o), st["box"], True) <= 1e-9
    v = np.stack([d.get("vx"), d.get("vy")], 1) - np.stack([o.get("vx"), o.get("vy")], 1)
    assert np.abs(v).max() <= 1e-9
    assert sc.rel_force_err(forces(d), forces(o)) <= 1e-9


@pytest.mark.parametrize("flags", [0, 1])
def test_langevin_noise_on_same_stream(gpu, flags):
    st = sc.lj_lattice(40, rho=0.7, seed=18)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.5, 0.3)
        s.pair_lj(1.0, 1.0, 2.5, shifted=True)
        s.constraint("plane")
        s.set_timestep(0.002)
        s.integrator_langevin(gamma=1.0, seed=3, temp=(0, 1.0, 1.0, 1))
    d.run(49, flags)
    o.run(49)
    assert pos_err(positions(d), positions(o), st["box"], True) <= 1e-9
    assert np.abs(d.get("vx") - o.get("vx")).max() <= 1e-9
    assert np.array_equal(d.get("age"), o.get("age"))


@pytest.mark.parametrize("flags", [0, 1])
def test_langevin_sphere_noise_on_same_stream(gpu, flags):
    # off the plane the closing kick has a normal component, which the next
    # loop-head enforce removes: the fused step must do the same
    st = sc.sphere_disks(1500, phi=0.8, seed=29)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.constraint("sphere", st["R"])
        s.set_timestep(0.005)
        s.integrator_langevin(gamma=1.0, seed=4, temp=(0, 0.5, 0.5, 1))
    d.run(99, flags)
    o.run(99)
    assert pos_err(positions(d), positions(o), st["box"], False) <= 1e-9
    assert np.abs(d.get("vz") - o.get("vz")).max() <= 1e-9


@pytest.mark.parametrize("flags", [0, 1])
def test_langevin_linear_temperature_same_stream(gpu, flags):
    # ValueLinear temperature: the fused step reads T at each step's run step
    st = sc.lj_lattice(40, rho=0.7, seed=30)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.5, 0.3)
        s.pair_lj(1.0, 1.0, 2.5, shifted=True)
        s.constraint("plane")
        s.set_timestep(0.002)
        s.integrator_langevin(gamma=1.0, seed=6, temp=(1, 0.2, 1.5, 30))
    d.run(59, flags)
    o.run(59)
    assert pos_err(positions(d), positions(o), st["box"], True) <= 1e-9
    assert np.abs(d.get("vy") - o.get("vy")).max() <= 1e-9


def test_langevin_plane_n_squared_list(gpu):
    # a box too thin for a cell grid: the fused steps keep the 3-D position
    # layout that the N^2 build reads
    rng = np.random.default_rng(6)
    n = 900
    st = {"n": n, "box": (30.0, 30.0, 4.0), "periodic": True,
          "cols": {"x": rng.uniform(-15, 15, n), "y": rng.uniform(-15, 15, n), "nx": np.ones(n)}}
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.constraint("plane")
        s.set_timestep(0.01)
        s.integrator_langevin(gamma=1.0, seed=9, temp=(0, 0.3, 0.3, 1))
    assert d.nlist_info()["uses_cells"] == 0
    d.run(299)
    o.run(299)
    assert pos_err(positions(d), positions(o), st["box"], True) <= 1e-9
    assert np.array_equal(d.pairs(), o.pairs())


def test_fused_langevin_types_and_radii(gpu):
    # the multi-type path of the fused Langevin step (per-type LJ and soft
    # tables, LJ sigma from the radii) in captured chunks, noise on
    st = sc.lj_lattice(48, rho=0.6, seed=35)
    rng = np.random.default_rng(35)
    st["cols"]["type"] = rng.integers(1, 3, st["n"]).astype(np.int32)
    st["cols"]["radius"] = rng.uniform(0.45, 0.55, st["n"])
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.6, 0.3)
        s.pair_lj(1.0, 1.0, 2.5, shifted=True, use_radii=True)
        s.pair_type_param("lj", 1, 2, 0.6, 1.0, 2.2)
        s.pair_soft(2.0, 0.9)
        s.pair_type_param("soft", 2, 2, 1.0, 0.8)
        s.constraint("plane")
        s.set_timestep(0.002)
        s.integrator_langevin(gamma=1.0, seed=15, temp=(0, 0.8, 0.8, 1))
        s.run(299)
    assert pos_err(positions(d), positions(o), st["box"], True) <= 1e-9
    assert np.abs(d.get("vx") - o.get("vx")).max() <= 1e-9
    assert np.array_equal(d.pairs(), o.pairs())


@pytest.mark.parametrize("integrator", ["langevin", "brownian"])
def test_fused_3d_periodic_box_without_constraint(gpu, integrator):
    # constraint none in a periodic 3-D box: the fused steps with the 3-D
    # position layout, periodic wraps in every direction and image flags
    rng = np.random.default_rng(12)
    n, L = 3000, 24.0
    g = np.arange(-L / 2, L / 2, L / 15)[:15]
    x, y, z = np.meshgrid(g, g, g[:14], indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)[:n] + rng.uniform(-0.1, 0.1, (n, 3))
    d3 = rng.normal(size=(n, 3))
    d3 /= np.linalg.norm(d3, axis=1)[:, None]
    st = {"n": n, "box": (L, L, L), "periodic": True,
          "cols": {"x": pts[:, 0], "y": pts[:, 1], "z": pts[:, 2], "nx": d3[:, 0], "ny": d3[:, 1], "nz": d3[:, 2],
                   "vx": rng.normal(0, 0.5, n), "vy": rng.normal(0, 0.5, n), "vz": rng.normal(0, 0.5, n)}}
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.5, 0.4)
        s.constraint("none")
        if integrator == "langevin":
            s.pair_lj(1.0, 1.0, 2.5, shifted=True)
            s.set_timestep(0.002)
            s.integrator_langevin(gamma=1.0, seed=13, temp=(0, 1.0, 1.0, 1))
        else:
            s.pair_soft(1.0, 1.2)
            s.align_polar(0.5, 2.0)
            s.set_timestep(0.01)
            s.integrator_brownian(v0=1.0, nu=0.1, seed=14, temp=(0, 0.1, 0.1, 1))
        s.run(299)
    assert pos_err(positions(d), positions(o), st["box"], True) <= 1e-9
    for k in ("ix", "iy", "iz"):
        assert np.array_equal(d.get(k), o.get(k)), k
    assert np.array_equal(d.pairs(), o.pairs())


@pytest.mark.parametrize("slots", [2, 6, 24])
def test_chunked_fused_steps_equal_direct_launches(gpu, slots):
    # captured chunks stop at the step whose positions need a rebuild and
    # resume after it: same steps, same rebuilds, bit for bit
    st = sc.lj_lattice(48, rho=0.7, seed=31)
    out = []
    for flags in (0, gpu.RUN_NO_GRAPH):
        d = sc.load(gpu.System, st)
        d.nlist(2.5, 0.3)
        d.pair_lj(1.0, 1.0, 2.5, shifted=True)
        d.constraint("plane")
        d.set_timestep(0.004)
        d.integrator_langevin(gamma=1.0, seed=8, temp=(0, 1.5, 1.5, 1))
        d.set_option("chunk_slots", slots)
        b0 = d.nlist_info()["builds"]
        d.run(399, flags)
        out.append(({k: d.get(k) for k in ("x", "y", "vx", "vy", "ix", "iy")}, d.pairs(), d.nlist_info()["builds"] - b0))
    (a, pa, ba), (c, pc, bc) = out
    assert all(np.array_equal(a[k], c[k]) for k in a)
    assert np.array_equal(pa, pc)
    assert ba == bc and ba > 10


def test_fused_langevin_gather_batch_8_equals_4(gpu):
    # the planar single-type step gathers its neighbours in batches of 4 or 8
    # list entries (two 16-byte list groups); the entries and the order of
    # the sums are the same, so the trajectories agree bit for bit
    st = sc.lj_lattice(40, rho=0.7, seed=36)
    out = []
    for batch in (4, 8):
        d = sc.load(gpu.System, st)
        d.nlist(2.5, 0.3)
        d.pair_lj(1.0, 1.0, 2.5, shifted=True)
        d.constraint("plane")
        d.set_timestep(0.004)
        d.integrator_langevin(gamma=1.0, seed=12, temp=(0, 1.0, 1.0, 1))
        d.set_option("gather_batch", batch)
        d.run(299)
        out.append(({k: d.get(k) for k in ("x", "y", "vx", "vy", "fx", "fy")}, d.pairs()))
    (a, pa), (c, pc) = out
    assert all(np.array_equal(a[k], c[k]) for k in a)
    assert np.array_equal(pa, pc)


@pytest.mark.parametrize("lanes", [1, 4, 16])
def test_fused_brownian_lanes_match_oracle(gpu, lanes):
    # each slot's list split over 1, 4 or 16 threads (partial sums added in a
    # fixed butterfly order) against the oracle's sequential sums
    st = sc.plane_disks(2000, phi=0.8, seed=37)
    d, o = both(gpu, st)
    d.set_option("brownian_lanes", lanes)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.align_polar(0.2, 2.4)
        s.constraint("plane")
        s.set_timestep(0.01)
        s.integrator_brownian(v0=0.2, nu=0.1, seed=13, temp=(0, 0.0, 0.0, 1))
        s.run(199)
    assert pos_err(positions(d), positions(o), st["box"], True) <= 1e-9
    assert np.abs(directors(d) - directors(o)).max() <= 1e-9
    assert sc.rel_force_err(torques(d), torques(o)) <= 1e-9


@pytest.mark.parametrize("method", ["spv", "bbk"])
def test_langevin_methods_noise_on(gpu, method):
    st = sc.lj_lattice(30, rho=0.6, seed=19)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.5, 0.3)
        s.pair_lj(1.0, 1.0, 2.5, shifted=True)
        s.constraint("plane")
        s.set_timestep(0.002)
        s.integrator_langevin(gamma=1.0, seed=5, method=method, temp=(0, 0.5, 1.0, 1))
        s.run(49)
    assert pos_err(positions(d), positions(o), st["box"], True) <= 1e-9
    assert np.abs(d.get("vy") - o.get("vy")).max() <= 1e-9


def test_brownian_noise_on_same_stream(gpu):
    st = sc.plane_disks(3000, phi=0.7, seed=20)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.align_polar(0.1, 2.4)
        s.constraint("plane")
        s.set_timestep(0.01)
        s.integrator_brownian(v0=0.1, nu=0.1, seed=7, temp=(0, 0.05, 1.0, 1))
        s.run(49)
    assert pos_err(positions(d), positions(o), st["box"], True) <= 1e-9
    assert np.abs(directors(d) - directors(o)).max() <= 1e-9


@pytest.mark.parametrize("kind", ["brownian", "spv"])
def test_captured_chunks_equal_direct_launches(gpu, kind):
    st = sc.plane_disks(2500, phi=0.7, seed=21)
    runs = []
    for flags in (0, gpu.RUN_NO_GRAPH):
        d = sc.load(gpu.System, st)
        d.nlist(2.4, 0.3)
        d.pair_soft(1.0, 2.0)
        d.constraint("plane")
        d.set_timestep(0.01)
        if kind == "brownian":
            d.align_polar(0.1, 2.4)
            d.integrator_brownian(v0=0.1, nu=0.1, seed=7, temp=(1, 0.0, 0.1, 60))
        else:
            d.integrator_langevin(gamma=1.0, seed=4, method="spv", temp=(1, 0.2, 0.5, 60))
        d.run_range(0, 70, 150, flags)
        d.run_range(70, 151, 150, flags)
        runs.append({k: d.get(k) for k in ("x", "y", "vx", "nx", "fx", "age")})
        d.close()
    for k in runs[0]:
        assert np.array_equal(runs[0][k], runs[1][k]), k


@pytest.mark.parametrize("case", ["flips_velocity_sp", "types_radii_nematic", "lj_polar_noise", "sphere_f32"])
def test_fused_brownian_cases(gpu, case):
    # the fused Brownian step (loop-head forces and torques + integrator in one
    # launch) against the oracle's two-pass loop
    if case == "sphere_f32":
        st = sc.sphere_disks(2000, phi=0.9, seed=32)
    elif case == "lj_polar_noise":
        st = sc.lj_lattice(48, rho=0.7, seed=32)
        th = np.random.default_rng(4).uniform(0.0, 2.0 * np.pi, st["n"])
        st["cols"]["nx"], st["cols"]["ny"] = np.cos(th), np.sin(th)
    else:
        st = sc.plane_disks(2500, phi=0.75, seed=32, polydisperse=(0.8, 1.2) if case == "types_radii_nematic" else None)
    if case == "types_radii_nematic":
        st["cols"]["type"] = np.random.default_rng(3).integers(1, 3, st["n"]).astype(np.int32)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.6, 0.3)
        if case == "flips_velocity_sp":
            s.pair_soft(1.0, 2.0)
            s.align_polar(0.3, 2.4)
            s.self_propulsion(0.5)
            s.constraint("plane")
            s.set_timestep(0.01)
            s.integrator_brownian(v0=0.2, nu=0.2, seed=5, nematic=True, tau=0.3, velocity_align=True,
                                  temp=(0, 0.02, 0.02, 1))
        elif case == "types_radii_nematic":
            s.pair_soft(1.0, 2.0, use_radii=True)
            s.pair_type_param("soft", 1, 2, 2.0, 2.0)
            s.align_nematic(0.4, 2.4)
            s.align_type_param("nematic", 1, 2, 0.7, 2.2)
            s.constraint("plane")
            s.set_timestep(0.01)
            s.integrator_brownian(v0=0.1, nu=0.1, seed=6, temp=(1, 0.0, 0.05, 100))
        elif case == "lj_polar_noise":
            s.pair_lj(1.0, 1.0, 2.5, shifted=True)
            s.pair_soft(1.0, 1.2)
            s.align_polar(0.2, 2.4)
            s.constraint("plane")
            s.set_timestep(0.002)
            s.integrator_brownian(v0=0.5, nu=0.05, seed=7, temp=(0, 0.1, 0.1, 1))
        else:
            s.pair_soft(10.0, 2.0)
            s.align_nematic(0.5, 2.4)
            s.constraint("sphere", st["R"])
            s.set_timestep(0.005)
            s.integrator_brownian(v0=0.5, nu=0.05, seed=8, temp=(0, 0.0, 0.0, 1))
    flags = gpu.RUN_FP32_FORCE if case == "sphere_f32" else 0
    d.run(199, flags)
    o.run(199)
    # FP32 pair terms: positions and directors after 200 steps within 1e-6, the
    # forces of the final positions within 1e-5 (FP32 rounding on top of the drift)
    tol = 1e-6 if case == "sphere_f32" else 1e-9
    assert pos_err(positions(d), positions(o), st["box"], st["periodic"]) <= tol
    assert np.abs(directors(d) - directors(o)).max() <= tol
    assert np.abs(d.get("vx") - o.get("vx")).max() <= tol * 100
    assert sc.rel_force_err(forces(d), forces(o)) <= tol * (10 if case == "sphere_f32" else 1)
    assert np.array_equal(d.get("age"), o.get("age"))


def test_fused_brownian_one_thread_per_slot(gpu):
    # above 4 Ki particles the fused Brownian step uses one thread per slot
    st = sc.plane_disks(40000, phi=0.8, seed=34)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.align_polar(0.2, 2.4)
        s.constraint("plane")
        s.set_timestep(0.01)
        s.integrator_brownian(v0=0.2, nu=0.1, seed=10, temp=(0, 0.05, 0.05, 1))
        s.run(99)
    assert pos_err(positions(d), positions(o), st["box"], True) <= 1e-9
    assert np.abs(directors(d) - directors(o)).max() <= 1e-9
    assert sc.rel_force_err(torques(d), torques(o)) <= 1e-9


@pytest.mark.parametrize("slots", [2, 8])
def test_chunked_brownian_equals_direct_launches(gpu, slots):
    st = sc.plane_disks(3000, phi=0.8, seed=33)
    out = []
    for flags in (0, gpu.RUN_NO_GRAPH):
        d = sc.load(gpu.System, st)
        d.nlist(2.4, 0.2)
        d.pair_soft(1.0, 2.0)
        d.align_polar(0.3, 2.4)
        d.constraint("plane")
        d.set_timestep(0.02)
        d.integrator_brownian(v0=0.5, nu=0.1, seed=9, temp=(0, 0.1, 0.1, 1))
        d.set_option("chunk_slots", slots)
        b0 = d.nlist_info()["builds"]
        d.run(299, flags)
        out.append(({k: d.get(k) for k in ("x", "y", "nx", "ny", "fx", "tau_z", "omega", "ix")}, d.pairs(),
                    d.nlist_info()["builds"] - b0))
    (a, pa, ba), (c, pc, bc) = out
    assert all(np.array_equal(a[k], c[k]) for k in a)
    assert np.array_equal(pa, pc)
    assert ba == bc and ba > 10


def test_brownian_pos_align_self_propulsion(gpu):
    st = sc.sphere_disks(1000, R=12.0, seed=22)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.5)
        s.pair_soft(10.0, 1.0)
        s.align_polar(1.0)
        s.self_propulsion(1.0)
        s.constraint("sphere", 12.0)
        s.set_timestep(0.01)
        s.integrator_brownian(which="brownian_pos", mu=1.0, seed=0, temp=(0, 0.1, 1.0, 1))
        s.integrator_brownian(which="brownian_align", nu=0.01, seed=0)
        s.run(49)
    assert pos_err(positions(d), positions(o), st["box"], False) <= 1e-9
    assert np.abs(directors(d) - directors(o)).max() <= 1e-9


@pytest.mark.parametrize("case", ["sphere_polar", "plane_nematic_flips"])
def test_fused_brownian_align_pos_pair(gpu, case):
    # the production stage of the reference's example confs (brownian_align +
    # brownian_pos with self propulsion), run as one fused launch per step in
    # captured chunks
    if case == "sphere_polar":
        st = sc.sphere_disks(1500, R=14.0, seed=23)
    else:
        st = sc.plane_disks(2500, phi=0.8, seed=23)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.5)
        s.pair_soft(10.0, 1.0 if case == "sphere_polar" else 2.0)
        if case == "sphere_polar":
            s.align_polar(1.0)
            s.constraint("sphere", 14.0)
        else:
            s.align_nematic(0.5, 2.4)
            s.constraint("plane")
        s.self_propulsion(1.0)
        s.set_timestep(0.01)
        s.integrator_brownian(which="brownian_pos", mu=1.0, seed=3, temp=(0, 0.1, 0.1, 1))
        s.integrator_brownian(which="brownian_align", nu=0.01, seed=4, nematic=case != "sphere_polar", tau=0.5)
        s.run(299)
    assert pos_err(positions(d), positions(o), st["box"], st["periodic"]) <= 1e-9
    assert np.abs(directors(d) - directors(o)).max() <= 1e-9
    assert np.abs(d.get("omega") - o.get("omega")).max() <= 1e-9
    assert np.array_equal(d.get("age"), o.get("age"))


@pytest.mark.parametrize("limits", [(0.01, 0.02), (None, None)])
def test_nve_soft_polar_overlaps(gpu, limits):
    # the overlap-relaxation pre-stage of the reference's example confs:
    # NVE with a displacement limit on overlapping soft polar disks
    st = sc.plane_disks(3000, phi=1.0, seed=27)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.align_polar(1.0, 2.4)
        s.constraint("plane")
        s.set_timestep(0.005)
        s.integrator_nve(limit=limits[0], angle_limit=limits[1])
        s.run(199)
    assert pos_err(positions(d), positions(o), st["box"], True) <= 1e-9
    assert np.abs(directors(d) - directors(o)).max() <= 1e-9
    assert np.abs(d.get("omega") - o.get("omega")).max() <= 1e-9
    assert np.abs(d.get("vx") - o.get("vx")).max() <= 1e-9
    assert np.array_equal(d.get("age"), o.get("age"))


def test_nve_soft_nematic_sphere(gpu):
    st = sc.sphere_disks(1500, phi=0.9, seed=28)
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(10.0, 2.0)
        s.align_nematic(0.5, 2.4)
        s.constraint("sphere", st["R"])
        s.set_timestep(0.005)
        s.integrator_nve(limit=0.01, angle_limit=0.05)
        s.run(299)
    assert pos_err(positions(d), positions(o), st["box"], False) <= 1e-9
    assert np.abs(d.get("vy") - o.get("vy")).max() <= 1e-9
    assert np.abs(directors(d) - directors(o)).max() <= 1e-9


def test_fire_parity_polydisperse(gpu):
    st = sc.plane_disks(1024, phi=0.88, seed=23, polydisperse=(0.8, 1.2))
    d, o = both(gpu, st)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0, use_radii=True)
        s.constraint("plane")
        s.set_timestep(0.01)
        s.integrator_fire(dt_max=0.1, alpha=0.1)
        s.run(199)
    fd, fo = d.fire_state(), o.fire_state()
    assert fd["dt"] == pytest.approx(fo["dt"], rel=1e-12)
    assert fd["alpha"] == pytest.approx(fo["alpha"], rel=1e-12)
    assert fd["E"] == pytest.approx(fo["E"], rel=1e-9)
    assert pos_err(positions(d), positions(o), st["box"], True) <= 1e-9


def test_fp32_force_variant_error(gpu):
    st = sc.lj_lattice(64, rho=0.7, seed=24)
    d32 = sc.load(gpu.System, st)
    d64 = sc.load(gpu.System, st)
    for s in (d32, d64):
        s.nlist(2.5, 0.3)
        s.pair_lj(1.0, 1.0, 2.5, shifted=True)
        s.constraint("plane")
        s.set_timestep(0.002)
        s.integrator_langevin(gamma=1.0, seed=3, temp=(0, 0.0, 1.0, 1))
    d32.run(0, gpu.RUN_FP32_FORCE)
    d64.run(0)
    assert sc.rel_force_err(forces(d32), forces(d64)) <= 1e-6


@pytest.mark.parametrize("kind", ["langevin", "brownian"])
def test_run_in_ranges_equals_one_run(gpu, kind):
    """Stopping at observation steps (run_range) leaves the trajectory bit-identical."""
    st = sc.lj_lattice(40, rho=0.7, seed=26) if kind == "langevin" else sc.plane_disks(1500, seed=26)
    a, b = sc.load(gpu.System, st), sc.load(gpu.System, st)
    for s in (a, b):
        s.constraint("plane")
        s.set_timestep(0.002 if kind == "langevin" else 0.01)
        if kind == "langevin":
            s.nlist(2.5, 0.3)
            s.pair_lj(1.0, 1.0, 2.5, shifted=True)
            s.integrator_langevin(gamma=1.0, seed=9, temp=(0, 1.0, 1.0, 1))
        else:
            s.nlist(2.4, 0.3)
            s.pair_soft(1.0, 2.0)
            s.align_polar(0.1, 2.4)
            s.integrator_brownian(v0=0.1, nu=0.1, seed=9)
    steps = 150
    a.run(steps)
    for lo, hi in [(0, 1), (1, 37), (37, 100), (100, 151)]:
        b.run_range(lo, hi, steps)
    for k in ["x", "y", "vx", "vy", "nx", "ny", "fx", "fy"]:
        assert np.array_equal(a.get(k), b.get(k)), k


def test_observables(gpu):
    st = sc.lj_lattice(32, seed=25)
    d = sc.load(gpu.System, st)
    v = np.stack([st["cols"]["vx"], st["cols"]["vy"]], 1)
    assert d.reduce("kinetic_energy") == pytest.approx(0.5 * float(np.sum(v * v)), rel=1e-12)
    assert d.reduce("velocity") == pytest.approx(float(np.mean(np.linalg.norm(v, axis=1))), rel=1e-12)
    d.nlist(2.5, 0.3)
    d.pair_lj(1.0, 1.0, 2.5, shifted=True)
    d.compute_forces()
    F = forces(d)
    assert d.reduce("max_force") == float(np.max(np.linalg.norm(F, axis=1)))
