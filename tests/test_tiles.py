"""Cell-tiled planar engine (tiles.hip) against the CPU oracle.

Planar single-type Langevin BAOAB runs with one common mass and age go
through the tiled engine when the run asks for it (SAMOS_RUN_TILES): state in
cell order, positions of each tile's 3 x (T + 2) surrounding cells staged in
LDS, Verlet lists as 16-bit indices into that stage.  Bars as the list-based path (tests/test_gpu_parity.py):

  noise-off trajectories   |dx|, |dv| <= 1e-9 after the run (FP64)
  forces left by the run   max_i |dF_i| / max_i |F_i| <= 1e-12 on identical
                           positions (the final forces are F(x_end))
  pair set after the run   bit-exact: the slot list is rebuilt at the last
                           tiled build's snapshot
  build count              equal to the oracle's
  FP32 pair arithmetic     forces <= 1e-6 relative to FP64
and, specific to the engine: the halo read from L2 instead of LDS (halo
larger than the LDS stage), runs split into ranges bit-identical to one run,
and agreement with the list-based kernels (the default path) to 1e-9.
"""
import numpy as np
import pytest

import oracle
from tests import scenarios as sc

pytestmark = pytest.mark.gpu

TILES = 8  # SAMOS_RUN_TILES


def setup_lj(s, gamma=1.0, T=0.0, seed=3, pad=0.3):
    s.nlist(2.5, pad)
    s.pair_lj(1.0, 1.0, 2.5, shifted=True)
    s.constraint("plane")
    s.set_timestep(0.002)
    s.integrator_langevin(gamma=gamma, seed=seed, temp=(0, T, 1.0, 1))


def state(s):
    return {k: s.get(k) for k in ["x", "y", "vx", "vy", "fx", "fy", "nx", "ny", "ix", "iy"]}


def pos_err(a, b, box):
    return max(float(np.abs(sc.min_image(a["x"] - b["x"], box[0])).max()),
               float(np.abs(sc.min_image(a["y"] - b["y"], box[1])).max()))


def test_tiled_lj_langevin_matches_oracle(gpu):
    st = sc.lj_lattice(48, rho=0.7, seed=41)
    d, o = sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)
    d.set_option("tile_cells", 4)
    for s in (d, o):
        setup_lj(s)
    print("tiled LJ Langevin, N=%d, 800 steps" % st["n"], flush=True)
    d0, o0 = d.nlist_info()["builds"], o.builds()
    d.run(799, TILES)
    o.run(799)
    info = d.tile_info()
    assert info["runs"] == 1 and info["rebuilds"] >= 1, info
    a, b = state(d), state(o)
    assert pos_err(a, b, st["box"]) <= 1e-9
    assert max(np.abs(a[k] - b[k]).max() for k in ("vx", "vy")) <= 1e-9
    F = np.stack([a["fx"], a["fy"]], 1)
    G = np.stack([b["fx"], b["fy"]], 1)
    assert np.abs(F - G).max() / np.abs(G).max() <= 1e-9
    assert np.array_equal(a["ix"], b["ix"]) and np.array_equal(a["iy"], b["iy"])
    assert np.array_equal(d.pairs(), o.pairs())
    assert d.nlist_info()["builds"] - d0 == o.builds() - o0
    # forces at the identical final positions
    d.compute_forces()
    o.set("x", d.get("x"))
    o.set("y", d.get("y"))
    o.compute_forces()
    F = np.stack([d.get("fx"), d.get("fy")], 1)
    G = np.stack([o.get("fx"), o.get("fy")], 1)
    assert np.abs(F - G).max() / np.abs(G).max() <= 1e-12


def test_tiled_soft_dense_halo_beyond_lds(gpu):
    # rho ~ 1.2 and 50 cells per tile (about 440 particles, two per thread of
    # the step kernel): staged halos exceed the LDS stage, so neighbour
    # positions are read through L2 (the fallback path)
    st = sc.plane_disks(60000, phi=3.8, seed=42)
    d, o = sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)
    d.set_option("tile_cells", 50)
    for s in (d, o):
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.constraint("plane")
        s.set_timestep(0.002)
        s.integrator_langevin(gamma=1.0, seed=5, temp=(0, 0.0, 1.0, 1))
    print("tiled dense soft, N=%d, 40 steps" % st["n"], flush=True)
    d.run(39, TILES)
    o.run(39)
    info = d.tile_info()
    assert info["runs"] == 1 and info["max_halo"] > 1024 and info["tile_cells"] == 50, info
    a, b = state(d), state(o)
    assert pos_err(a, b, st["box"]) <= 1e-9
    assert max(np.abs(a[k] - b[k]).max() for k in ("vx", "vy")) <= 1e-9
    assert np.array_equal(d.pairs(), o.pairs())


@pytest.mark.parametrize("cells", [1, 3, 0])
def test_tiled_noise_on_same_stream(gpu, cells):
    st = sc.lj_lattice(40, rho=0.7, seed=43)
    d, o = sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)
    d.set_option("tile_cells", cells)
    for s in (d, o):
        setup_lj(s, T=1.0, seed=7)
    print("tiled noise-on, tile_cells=%d" % cells, flush=True)
    d.run(199, TILES)
    o.run(199)
    assert d.tile_info()["runs"] == 1
    a, b = state(d), state(o)
    assert pos_err(a, b, st["box"]) <= 1e-9
    assert max(np.abs(a[k] - b[k]).max() for k in ("vx", "vy")) <= 1e-9
    assert np.array_equal(d.pairs(), o.pairs())


def test_tiled_equals_list_kernels(gpu):
    st = sc.lj_lattice(64, rho=0.7, seed=44)
    runs = []
    for flags in (TILES, 0):
        d = sc.load(gpu.System, st)
        setup_lj(d, T=1.0, seed=11)
        d.run(499, flags)
        runs.append((state(d), d.tile_info()["runs"], d.pairs()))
    (a, ta, pa), (b, tb, pb) = runs
    assert ta == 1 and tb == 0
    assert pos_err(a, b, st["box"]) <= 1e-9
    assert max(np.abs(a[k] - b[k]).max() for k in ("vx", "vy")) <= 1e-9
    assert np.array_equal(pa, pb)


def test_tiled_ranges_bit_identical(gpu):
    st = sc.lj_lattice(40, rho=0.7, seed=45)
    a, b = sc.load(gpu.System, st), sc.load(gpu.System, st)
    for s in (a, b):
        setup_lj(s, T=1.0, seed=9)
    steps = 300
    a.run(steps, TILES)
    for lo, hi in [(0, 1), (1, 77), (77, 200), (200, 301)]:
        b.run_range(lo, hi, steps, TILES)
    assert b.tile_info()["runs"] == 4
    sa, sb = state(a), state(b)
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k


def test_tiled_fp32_forces(gpu):
    st = sc.lj_lattice(40, rho=0.7, seed=46)
    d, o = sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)
    for s in (d, o):
        setup_lj(s)
    d.run(20, gpu.RUN_FP32_FORCE | TILES)
    assert d.tile_info()["runs"] == 1
    o.set("x", d.get("x"))
    o.set("y", d.get("y"))
    o.compute_forces()
    F = np.stack([d.get("fx"), d.get("fy")], 1)
    G = np.stack([o.get("fx"), o.get("fy")], 1)
    assert np.abs(F - G).max() / np.abs(G).max() <= 1e-6


def test_tiled_not_used_where_it_does_not_apply(gpu):
    # two particle types: the list-based kernels run
    st = sc.lj_lattice(24, rho=0.7, seed=47)
    d = sc.load(gpu.System, st)
    d.set("type", (np.arange(st["n"]) % 2) + 1)
    setup_lj(d)
    d.run(10, TILES)
    assert d.tile_info()["runs"] == 0
    # not asked for: the list-based kernels run
    d = sc.load(gpu.System, sc.lj_lattice(24, rho=0.7, seed=48))
    setup_lj(d)
    d.run(10)
    assert d.tile_info()["runs"] == 0


# ---- tiled Brownian step: forces and aligner torques from staged positions
# and directors (one "brownian" integrator on a planar single-type system) ----

def brownian_state(s):
    return {k: s.get(k) for k in ["x", "y", "vx", "vy", "nx", "ny", "fx", "fy", "tau_z", "omega", "ix", "iy",
                                  "age"]}


def setup_brownian(s, case):
    if case == "polar_flips_velocity_sp":
        s.nlist(2.6, 0.3)
        s.pair_soft(1.0, 2.0)
        s.align_polar(0.3, 2.4)
        s.self_propulsion(0.5)
        s.constraint("plane")
        s.set_timestep(0.01)
        s.integrator_brownian(v0=0.2, nu=0.2, seed=5, nematic=True, tau=0.3, velocity_align=True,
                              temp=(0, 0.02, 0.02, 1))
    elif case == "nematic_linear_temperature":
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.align_nematic(0.4, 2.4)
        s.constraint("plane")
        s.set_timestep(0.01)
        s.integrator_brownian(v0=0.1, nu=0.1, seed=6, temp=(1, 0.0, 0.05, 100))
    elif case == "lj_polar_noise":
        s.nlist(2.5, 0.3)
        s.pair_lj(1.0, 1.0, 2.5, shifted=True)
        s.pair_soft(1.0, 1.2)
        s.align_polar(0.2, 2.4)
        s.constraint("plane")
        s.set_timestep(0.002)
        s.integrator_brownian(v0=0.5, nu=0.05, seed=7, temp=(0, 0.1, 0.1, 1))
    else:  # noise off, no aligner
        s.nlist(2.4, 0.3)
        s.pair_soft(1.0, 2.0)
        s.constraint("plane")
        s.set_timestep(0.01)
        s.integrator_brownian(v0=0.3, nu=0.0, seed=8)


def brownian_input(case, seed):
    if case == "lj_polar_noise":
        st = sc.lj_lattice(48, rho=0.7, seed=seed)
        th = np.random.default_rng(4).uniform(0.0, 2.0 * np.pi, st["n"])
        st["cols"]["nx"], st["cols"]["ny"] = np.cos(th), np.sin(th)
        return st
    return sc.plane_disks(2500, phi=0.75, seed=seed)


@pytest.mark.parametrize("case", ["polar_flips_velocity_sp", "nematic_linear_temperature", "lj_polar_noise",
                                  "soft_noise_off"])
def test_tiled_brownian_matches_oracle(gpu, case):
    st = brownian_input(case, 51)
    d, o = sc.load(gpu.System, st), sc.load(oracle.OracleSystem, st)
    d.set_option("tile_cells", 4)
    for s in (d, o):
        setup_brownian(s, case)
    print("tiled Brownian %s, N=%d, 200 steps" % (case, st["n"]), flush=True)
    b0, ob0 = d.nlist_info()["builds"], o.builds()
    d.run(199, TILES)
    o.run(199)
    info = d.tile_info()
    assert info["runs"] == 1 and info["rebuilds"] >= 1, info
    a, b = brownian_state(d), brownian_state(o)
    assert pos_err(a, b, st["box"]) <= 1e-9
    for k in ("nx", "ny", "omega"):
        assert np.abs(a[k] - b[k]).max() <= 1e-9, k
    assert max(np.abs(a[k] - b[k]).max() for k in ("vx", "vy")) <= 1e-7
    F = np.stack([a["fx"], a["fy"]], 1)
    G = np.stack([b["fx"], b["fy"]], 1)
    assert np.abs(F - G).max() / np.abs(G).max() <= 1e-9
    if case != "soft_noise_off":
        assert np.abs(a["tau_z"] - b["tau_z"]).max() / np.abs(b["tau_z"]).max() <= 1e-9
    assert np.array_equal(a["ix"], b["ix"]) and np.array_equal(a["iy"], b["iy"])
    assert np.array_equal(a["age"], b["age"])
    assert np.array_equal(d.pairs(), o.pairs())
    assert d.nlist_info()["builds"] - b0 == o.builds() - ob0


def test_tiled_brownian_equals_list_kernels(gpu):
    # a larger active polar system, captured chunks on both engines
    st = sc.plane_disks(30000, phi=0.8, seed=52)
    runs = []
    for flags in (TILES, 0):
        d = sc.load(gpu.System, st)
        setup_brownian(d, "polar_flips_velocity_sp")
        d.run(299, flags)
        runs.append((brownian_state(d), d.tile_info()["runs"], d.pairs(), d.nlist_info()["builds"]))
        d.close()
    (a, ta, pa, ba), (b, tb, pb, bb) = runs
    assert ta == 1 and tb == 0
    assert pos_err(a, b, st["box"]) <= 1e-9
    assert max(np.abs(a[k] - b[k]).max() for k in ("nx", "ny", "omega")) <= 1e-9
    assert np.array_equal(pa, pb) and ba == bb


def test_tiled_brownian_ranges_bit_identical(gpu):
    st = sc.plane_disks(3000, phi=0.8, seed=53)
    a, b = sc.load(gpu.System, st), sc.load(gpu.System, st)
    for s in (a, b):
        setup_brownian(s, "polar_flips_velocity_sp")
    steps = 300
    a.run(steps, TILES)
    for lo, hi in [(0, 1), (1, 90), (90, 220), (220, 301)]:
        b.run_range(lo, hi, steps, TILES)
    assert b.tile_info()["runs"] == 4
    sa, sb = brownian_state(a), brownian_state(b)
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
