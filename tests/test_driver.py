"""The `samos <conf>` driver (samos_amd/bin/samos): conf grammar, error
behaviour and an end-to-end run checked against the oracle.

CPU tests cover errors the reference raises before any particle data exists
(unknown command, unknown box type, bad parameter block, missing input); the
GPU tests run complete confs and compare dumps and logs with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests import scenarios as sc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAMOS = os.path.join(ROOT, "samos_amd", "bin", "samos")


def run_conf(tmp_path, text, timeout=120):
    conf = tmp_path / "run.conf"
    conf.write_text(text)
    return subprocess.run([SAMOS, str(conf)], cwd=tmp_path, capture_output=True, text=True, timeout=timeout)


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(SAMOS):
        pytest.fail("samos driver not built (make -C samos_amd all)")
    return SAMOS


def test_unknown_command(driver, tmp_path):
    r = run_conf(tmp_path, "# comment\nfrobnicate { a = 1 }\n")
    assert r.returncode != 0
    assert "Unknown command : frobnicate" in r.stderr and "line : 2" in r.stderr


def test_unknown_box_type(driver, tmp_path):
    r = run_conf(tmp_path, "box cubic { lx = 10 }\n")
    assert r.returncode != 0 and "Box type cubic is not known." in r.stderr


def test_bad_parameter_block(driver, tmp_path):
    r = run_conf(tmp_path, "box periodic { lx = 10; = 3 }\n")
    assert r.returncode != 0 and "Could not parse parameters" in r.stderr


def test_commands_before_input(driver, tmp_path):
    r = run_conf(tmp_path, "box periodic { lx = 10; ly = 10; lz = 10 }\nnlist { rcut = 2.4 }\n")
    assert r.returncode != 0 and "System has not been defined" in r.stderr


def test_missing_input_file(driver, tmp_path):
    r = run_conf(tmp_path, "box periodic { lx = 10; ly = 10; lz = 10 }\ninput nothere.dat\n")
    assert r.returncode != 0 and "Problem opening file nothere.dat" in r.stderr


def test_out_of_scope_command(driver, tmp_path):
    r = run_conf(tmp_path, "bond harmonic { k = 1 }\n")
    assert r.returncode != 0 and "outside this implementation" in r.stderr


# ------------------------------------------------------------------ GPU ---
def write_input(path, st):
    c = st["cols"]
    n = st["n"]
    with open(path, "w") as fh:
        fh.write("# test input\nkeys: id type radius x y z vx vy vz nx ny nz\n")
        for i in range(n):
            fh.write("%d 1 %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n" % (
                i, c.get("radius", np.ones(n))[i], c["x"][i], c["y"][i], c["z"][i],
                c.get("vx", np.zeros(n))[i], c.get("vy", np.zeros(n))[i], c.get("vz", np.zeros(n))[i],
                c["nx"][i], c["ny"][i], c["nz"][i]))


def read_full(path):
    rows = [ln.split() for ln in open(path) if ln.strip() and not ln.startswith(("#", "keys"))]
    return np.array(rows, dtype=float)


@pytest.mark.gpu
def test_conf_brownian_polar_matches_oracle(gpu, driver, tmp_path):
    st = sc.plane_disks(2000, phi=0.8, seed=31)
    L = st["box"][0]
    write_input(tmp_path / "p.dat", st)
    conf = f"""messages run.msg
box periodic {{ lx = {L!r}; ly = {L!r}; lz = 10.0 }}
input p.dat
nlist {{ rcut = 2.4; pad = 0.3 }}
constraint plane {{ }}
pair_potential soft {{ k = 1.0; a = 2.0 }}
pair_align polar {{ J = 0.1; rcut = 2.4 }}
log run.log {{ freq = 50; velocity; soft_energy; polar_align }}
dump out {{ type = full; start = 0; freq = 200; multi; id; coordinate; velocity; director; keys }}
timestep 0.01
integrator brownian {{ v0 = 0.1; nu = 0.0; mu = 1.0; seed = 1 }}
run 199
"""
    r = run_conf(tmp_path, conf)
    assert r.returncode == 0, r.stderr
    o = sc.load(oracle.OracleSystem, st)
    o.nlist(2.4, 0.3)
    o.pair_soft(1.0, 2.0)
    o.align_polar(0.1, 2.4)
    o.constraint("plane")
    o.set_timestep(0.01)
    o.integrator_brownian(v0=0.1, nu=0.0, mu=1.0, seed=1)
    o.run(199)
    # the dump at step 200 is never due (run 199 ends at time step 199): compare the one at 0
    d0 = read_full(tmp_path / "out_0000000000.dat")
    assert d0.shape == (2000, 10)
    log = np.array([ln.split() for ln in open(tmp_path / "run.log") if not ln.startswith("#")], dtype=float)
    assert list(log[:, 0]) == [0, 50, 100, 150]
    # log columns: step, polar_align, soft_energy, velocity (map order after step)
    o2 = sc.load(oracle.OracleSystem, st)
    o2.nlist(2.4, 0.3)
    o2.pair_soft(1.0, 2.0)
    o2.align_polar(0.1, 2.4)
    o2.constraint("plane")
    e = o2.compute_forces()
    assert log[0, 1] == pytest.approx(e["align"], rel=1e-6)
    assert log[0, 2] == pytest.approx(e["pair"], rel=1e-6)


@pytest.mark.gpu
def test_conf_langevin_lj_final_state(gpu, driver, tmp_path):
    st = sc.lj_lattice(24, rho=0.7, seed=32)
    L = st["box"][0]
    write_input(tmp_path / "p.dat", st)
    conf = f"""box periodic {{ lx = {L!r}; ly = {L!r}; lz = 10.0 }}
input p.dat
nlist {{ rcut = 2.5; pad = 0.3 }}
constraint plane {{ }}
pair_potential lj {{ epsilon = 1.0; sigma = 1.0; rcut = 2.5; shifted }}
dump fin {{ type = full; start = 300; freq = 300; multi; id; coordinate; velocity; keys }}
timestep 0.002
integrator langevin {{ gamma = 1.0; seed = 3; temperature_control = constant; min_val = 0.0 }}
run 400
"""
    r = run_conf(tmp_path, conf)
    assert r.returncode == 0, r.stderr
    o = sc.load(oracle.OracleSystem, st)
    o.nlist(2.5, 0.3)
    o.pair_lj(1.0, 1.0, 2.5, shifted=True)
    o.constraint("plane")
    o.set_timestep(0.002)
    o.integrator_langevin(gamma=1.0, seed=3, temp=(0, 0.0, 1.0, 1))
    o.run(299)  # loop-head state of step 300 = after 300 iterations, then the loop head
    o.compute_forces()
    d = read_full(tmp_path / "fin_0000000300.dat")
    dx = sc.min_image(d[:, 1] - o.get("x"), L)
    dy = sc.min_image(d[:, 2] - o.get("y"), L)
    assert np.abs(dx).max() < 1e-8 and np.abs(dy).max() < 1e-8
    assert np.abs(d[:, 4] - o.get("vx")).max() < 1e-6


@pytest.mark.gpu
def test_conf_nve_limit_then_brownian(gpu, driver, tmp_path):
    # the example confs' pattern: an NVE overlap relaxation with a displacement
    # limit, `disable nve`, then the production integrator
    st = sc.plane_disks(1500, phi=1.0, seed=34)
    L = st["box"][0]
    write_input(tmp_path / "p.dat", st)
    conf = f"""box periodic {{ lx = {L!r}; ly = {L!r}; lz = 10.0 }}
input p.dat
nlist {{ rcut = 2.4; pad = 0.3 }}
constraint plane {{ }}
pair_potential soft {{ k = 1.0; a = 2.0 }}
timestep 0.01
integrator nve {{ limit = 0.0001 }}
run 100
disable nve {{ group = all }}
integrator brownian {{ v0 = 0.1; nu = 0.0; mu = 1.0; seed = 1 }}
dump fin {{ type = full; start = 150; freq = 150; multi; id; coordinate; keys }}
run 200
"""
    r = run_conf(tmp_path, conf)
    assert r.returncode == 0, r.stderr
    o = sc.load(oracle.OracleSystem, st)
    o.nlist(2.4, 0.3)
    o.pair_soft(1.0, 2.0)
    o.constraint("plane")
    o.set_timestep(0.01)
    o.integrator_nve(limit=0.0001)
    o.run(100)
    o2_state = {k: o.get(k) for k in ("x", "y", "vx", "vy", "nx", "ny")}
    b = sc.load(oracle.OracleSystem, {"n": st["n"], "box": st["box"], "periodic": True,
                                       "cols": dict(st["cols"], **o2_state)})
    b.nlist(2.4, 0.3)
    b.pair_soft(1.0, 2.0)
    b.constraint("plane")
    b.set_timestep(0.01)
    b.integrator_brownian(v0=0.1, nu=0.0, mu=1.0, seed=1)
    b.run(48)  # time steps 101..149; the dump at 150 is the loop-head state of step 150
    d = read_full(tmp_path / "fin_0000000150.dat")
    assert d.shape[0] == st["n"]
    dx = sc.min_image(d[:, 1] - b.get("x"), L)
    assert np.abs(dx).max() < 1e-6


@pytest.mark.gpu
def test_conf_unknown_parameter_rejected(gpu, driver, tmp_path):
    st = sc.plane_disks(50, seed=33)
    L = st["box"][0]
    write_input(tmp_path / "p.dat", st)
    r = run_conf(tmp_path, f"""box periodic {{ lx = {L!r}; ly = {L!r}; lz = 10.0 }}
input p.dat
pair_potential soft {{ k = 1.0; kappa = 2 }}
""")
    assert r.returncode != 0
    assert 'Unknown parameter "kappa" in soft potential.' in r.stderr


@pytest.mark.gpu
def test_conf_morse_and_yukawa_defaults(gpu, driver, tmp_path):
    # reference parameter names; morse rcut defaults to 2.5 re, yukawa takes
    # its defaults g = kappa = 1, rcut = 3
    st = sc.plane_disks(1200, phi=0.7, seed=36)
    L = st["box"][0]
    write_input(tmp_path / "p.dat", st)
    conf = f"""box periodic {{ lx = {L!r}; ly = {L!r}; lz = 10.0 }}
input p.dat
nlist {{ rcut = 3.2; pad = 0.3 }}
constraint plane {{ }}
pair_potential morse {{ D = 0.5; a = 2.0; re = 1.1; shifted }}
pair_potential yukawa {{ }}
timestep 0.005
integrator brownian {{ v0 = 0.1; nu = 0.0; mu = 1.0; seed = 2 }}
dump fin {{ type = full; start = 60; freq = 60; multi; id; coordinate; keys }}
run 61
"""
    r = run_conf(tmp_path, conf)
    assert r.returncode == 0, r.stderr
    o = sc.load(oracle.OracleSystem, st)
    o.nlist(3.2, 0.3)
    o.pair_extra("morse", [0.5, 2.0, 1.1, 2.75], shifted=True)
    o.pair_extra("yukawa", [1.0, 1.0, 3.0])
    o.constraint("plane")
    o.set_timestep(0.005)
    o.integrator_brownian(v0=0.1, nu=0.0, mu=1.0, seed=2)
    o.run(59)
    d = read_full(tmp_path / "fin_0000000060.dat")
    dx = sc.min_image(d[:, 1] - o.get("x"), L)
    assert np.abs(dx).max() < 1e-6


@pytest.mark.gpu
def test_conf_gaussian_per_type_rejected(gpu, driver, tmp_path):
    st = sc.plane_disks(50, seed=37)
    L = st["box"][0]
    write_input(tmp_path / "p.dat", st)
    r = run_conf(tmp_path, f"""box periodic {{ lx = {L!r}; ly = {L!r}; lz = 10.0 }}
input p.dat
nlist {{ rcut = 3.2; pad = 0.3 }}
pair_potential gaussian {{ A = 1.0 }}
pair_param gaussian {{ type_1 = 1; type_2 = 1; A = 2.0 }}
""")
    assert r.returncode != 0
    assert "not supported" in r.stderr


@pytest.mark.gpu
def test_conf_torus_brownian(gpu, driver, tmp_path):
    st = sc.surface_disks("torus", 1500, [12.0, 4.0], seed=38, jitter=0.05)
    Lx, Ly, Lz = st["box"]
    write_input(tmp_path / "p.dat", st)
    conf = f"""box fixed {{ lx = {Lx!r}; ly = {Ly!r}; lz = {Lz!r} }}
input p.dat
nlist {{ rcut = 2.4; pad = 0.3 }}
constraint torus {{ c = 12.0; a = 4.0; maxiter = 50 }}
pair_potential soft {{ k = 1.0; a = 2.0 }}
pair_align polar {{ J = 0.5; rcut = 2.4 }}
timestep 0.01
integrator brownian {{ v0 = 0.2; nu = 0.0; mu = 1.0; seed = 2 }}
dump fin {{ type = full; start = 80; freq = 80; multi; id; coordinate; keys }}
run 81
"""
    r = run_conf(tmp_path, conf)
    assert r.returncode == 0, r.stderr
    o = sc.load(oracle.OracleSystem, st)
    o.nlist(2.4, 0.3)
    o.constraint_surface("torus", [12.0, 4.0], max_iter=50)
    o.pair_soft(1.0, 2.0)
    o.align_polar(0.5, 2.4)
    o.set_timestep(0.01)
    o.integrator_brownian(v0=0.2, nu=0.0, mu=1.0, seed=2)
    o.run(79)
    d = read_full(tmp_path / "fin_0000000080.dat")
    assert np.abs(d[:, 1] - o.get("x")).max() < 1e-6
    assert np.abs(d[:, 3] - o.get("z")).max() < 1e-6


@pytest.mark.gpu
def test_conf_constraint_tolerance_is_an_integer(gpu, driver, tmp_path):
    # the reference converts tol with an integer cast: 1e-8 is a conversion error
    st = sc.surface_disks("ellipsoid", 50, [6.0, 5.0, 4.0], seed=39)
    write_input(tmp_path / "p.dat", st)
    L = st["box"][0]
    r = run_conf(tmp_path, f"""box fixed {{ lx = {L!r}; ly = {L!r}; lz = {L!r} }}
input p.dat
constraint ellipsoid {{ a = 6.0; b = 5.0; c = 4.0; tol = 1e-8 }}
""")
    assert r.returncode != 0
    assert 'bad integer value "1e-8" for tol' in r.stderr
