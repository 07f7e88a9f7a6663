"""The `samos` driver on the command and dump mix of the reference's own
example confs, checked against the oracle:

  configurations/two_component_plane/two_component_plane.conf: `group` by
      type, NVE with a displacement limit, `disable`, per-type polar coupling,
      two Brownian integrators on the passive and active groups;
  configurations/soft_on_plane/soft_on_plane.conf: every dump type it uses
      (velocity, xyz, dcd, xyzv of directors and of velocities, full, vtp),
      the NVE stage, self propulsion and the brownian_pos / brownian_align pair.

Particle counts and step counts are scaled down; the bars are those of the
device-vs-oracle tests (noise on, same Philox stream, 1e-9 on positions,
velocities and directors), and each dump format's numbers are checked against
the full dump of the same step.
"""
import os
import struct

import numpy as np
import pytest

import oracle
from tests import scenarios as sc
from tests.test_driver import SAMOS, read_full, run_conf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(SAMOS):
        pytest.fail("samos driver not built (make -C samos_amd all)")
    return SAMOS


def write_typed_input(path, st):
    c = st["cols"]
    n = st["n"]
    tp = c.get("type", np.ones(n, dtype=np.int32))
    with open(path, "w") as fh:
        fh.write("keys: id type radius x y z vx vy vz nx ny nz\n")
        for i in range(n):
            fh.write("%d %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n" % (
                i, tp[i], c.get("radius", np.ones(n))[i], c["x"][i], c["y"][i], c["z"][i],
                c.get("vx", np.zeros(n))[i], c.get("vy", np.zeros(n))[i], c.get("vz", np.zeros(n))[i],
                c["nx"][i], c["ny"][i], c["nz"][i]))


def wrapped_err(a, b, L):
    d = a - b
    d -= L * np.round(d / L)
    return float(np.abs(d).max())


def test_two_component_plane_conf(driver, tmp_path):
    st = sc.plane_disks(2400, phi=0.75, seed=81)
    rng = np.random.default_rng(82)
    st["cols"]["type"] = np.where(rng.uniform(size=st["n"]) < 0.5, 2, 1).astype(np.int32)
    L = st["box"][0]
    write_typed_input(tmp_path / "plane.dat", st)
    conf = f"""messages messages.msg
box periodic {{ lx = {L!r};  ly = {L!r};  lz = 10.0 }}
input plane.dat
nlist {{ rcut = 2.5; pad = 0.5 }}
group passive  {{ type = 1 }}
group active   {{ type = 2 }}
dump two_component {{ type=full; start=0; freq=91; multi; radius; tp; coordinate; velocity; director; header  }}
constraint plane {{  }}
pair_potential soft {{ k = 1.0 }}
pair_align polar {{ J = 1.0 }}
log two_component.log {{ freq = 30; velocity; soft_energy; polar_align }}
integrator nve {{ dt=0.001; limit=0.0001 }}
run 30
disable nve {{ group=all }}
align_param polar {{ type_1 = 1; type_2 = 1; J = 0.0 }}
align_param polar {{ type_1 = 1; type_2 = 2; J = 0.0 }}
align_param polar {{ type_1 = 2; type_2 = 2; J = 1.0 }}
integrator brownian {{ group = active; dt=0.005; seed = 1;  nu = 0.05; mu = 1.0;  v0 = 0.5; }}
integrator brownian {{ group = passive; dt=0.005; seed = 4;  nu = 0.00; mu = 1.0;  v0 = 0.0; temperature_control=constant; min_val=0.3; max_val=0.3; steps = 1000 }}
run 60
"""
    r = run_conf(tmp_path, conf)
    assert r.returncode == 0, r.stderr
    o = sc.load(oracle.OracleSystem, st)
    types = st["cols"]["type"]
    o.group("passive", types == 1)
    o.group("active", types == 2)
    o.nlist(2.5, 0.5)
    o.pair_soft(1.0, 2.0)
    o.align_polar(1.0, 2.5)
    o.constraint("plane")
    o.set_timestep(0.001)
    o.integrator_nve(limit=0.0001)
    o.run(30)
    o.disable("nve", "all")
    o.align_type_param("polar", 1, 1, 0.0, 2.5)
    o.align_type_param("polar", 1, 2, 0.0, 2.5)
    o.align_type_param("polar", 2, 2, 1.0, 2.5)
    o.set_timestep(0.005)
    o.integrator_brownian(v0=0.5, nu=0.05, mu=1.0, seed=1, group="active")
    o.integrator_brownian(v0=0.0, nu=0.0, mu=1.0, seed=4, temp=(0, 0.3, 0.3, 1000), group="passive")
    o.run(59)  # the dump at time step 91 is the loop head of the second run's iteration 60
    d = read_full(tmp_path / "two_component_0000000091.dat")
    # columns: type, radius, x y z, vx vy vz, nx ny nz
    assert np.array_equal(d[:, 0].astype(int), types)
    assert wrapped_err(d[:, 2], o.get("x"), L) <= 1e-9 and wrapped_err(d[:, 3], o.get("y"), L) <= 1e-9
    assert np.abs(d[:, 5] - o.get("vx")).max() <= 1e-6  # %12.7e columns
    assert np.abs(d[:, 8] - o.get("nx")).max() <= 1e-5   # %10.6f columns
    log = [ln.split() for ln in open(tmp_path / "two_component.log") if not ln.startswith("#")]
    assert [int(float(row[0])) for row in log] == [0, 30, 60, 90]


def test_soft_on_plane_dump_types(driver, tmp_path):
    st = sc.plane_disks(1500, phi=0.8, seed=83)
    L = st["box"][0]
    write_typed_input(tmp_path / "plane.dat", st)
    conf = f"""messages messages.msg
box periodic {{ lx = {L!r};  ly = {L!r};  lz = 10.0 }}
input plane.dat
nlist {{ rcut = 2.5; pad = 0.5 }}
dump plane_test {{ type=velocity; start=0; freq=40; multi; header }}
dump plane_test {{ type=xyz; start=0; freq=40; multi; header }}
dump plane_test {{ type=dcd; start=0; freq=20  }}
dump plane_test_dir {{ type=xyzv; start=0; freq=40; scale = 0.25; header  }}
dump plane_test_vel {{ type=xyzv; start=0; freq=40; scale = 0.1; velocity; header  }}
dump plane_test {{ type=full; start=0; freq=40; multi; coordinate; velocity; director; header  }}
dump plane_test {{ type=vtp; start=0; freq=40;  }}
constraint plane {{  }}
pair_potential soft {{ k = 1.0 }}
pair_align polar {{ J = 1.0 }}
timestep 0.001
integrator nve {{ limit=0.0001 }}
run 20
disable nve {{ group=all }}
external self_propulsion {{ alpha = 1.0 }}
timestep 0.01
integrator brownian_pos {{ seed = 1;  mu = 1.0; group = all }}
integrator brownian_align {{ seed = 1;  mur = 1.0; nu = 0.1; group = all }}
run 59
"""
    r = run_conf(tmp_path, conf)
    assert r.returncode == 0, r.stderr
    o = sc.load(oracle.OracleSystem, st)
    o.nlist(2.5, 0.5)
    o.pair_soft(1.0, 2.0)
    o.align_polar(1.0, 2.5)
    o.constraint("plane")
    o.set_timestep(0.001)
    o.integrator_nve(limit=0.0001)
    o.run(20)
    o.disable("nve", "all")
    o.self_propulsion(1.0)
    o.set_timestep(0.01)
    o.integrator_brownian(which="brownian_pos", seed=1, mu=1.0)
    o.integrator_brownian(which="brownian_align", seed=1, mur=1.0, nu=0.1)
    o.run(58)  # time step 80: loop head of the second run's iteration 59
    n = st["n"]
    full = read_full(tmp_path / "plane_test_0000000080.dat")
    assert full.shape == (n, 9)
    x, y, vx, vy, nx, ny = full[:, 0], full[:, 1], full[:, 3], full[:, 4], full[:, 6], full[:, 7]
    assert wrapped_err(x, o.get("x"), L) <= 1e-9 and wrapped_err(y, o.get("y"), L) <= 1e-9
    assert np.abs(vx - o.get("vx")).max() <= 1e-6 and np.abs(nx - o.get("nx")).max() <= 1e-5
    # velocity dump: x - a vx / 2, ..., a vx with a = 1
    vel = np.array([ln.split() for ln in open(tmp_path / "plane_test_0000000080.vel") if not ln.startswith("#")],
                   dtype=float)
    np.testing.assert_allclose(vel[:, 3], vx, atol=2e-6)
    np.testing.assert_allclose(vel[:, 0], x - 0.5 * vx, atol=2e-6)
    # xyz dump
    xyz = open(tmp_path / "plane_test_0000000080.xyz").read().splitlines()
    assert int(xyz[0]) == n and xyz[1] == "Generated by SAMoS code."
    np.testing.assert_allclose(np.array([r.split()[1] for r in xyz[2:]], dtype=float), x, atol=2e-6)
    # xyzv of directors (scale 0.25) and of velocities (scale 0.1), one file each, frames every 40 steps
    dirv = open(tmp_path / "plane_test_dir.xyzv").read().splitlines()
    assert len(dirv) == 3 * (n + 2)  # time steps 0, 40, 80
    assert dirv[1] == "Generated by SAMoS code. Vectors are directors scaled by 0.25"
    last = np.array([r.split()[1:] for r in dirv[-n:]], dtype=float)
    assert all(r.split()[0] == "A" for r in dirv[-n:])
    np.testing.assert_allclose(last[:, 3], 0.25 * nx, atol=2e-6)
    velv = open(tmp_path / "plane_test_vel.xyzv").read().splitlines()
    assert velv[1] == "Generated by SAMoS code. Vectors are velocities scaled by 0.1"
    np.testing.assert_allclose(np.array([r.split()[4] for r in velv[-n:]], dtype=float), 0.1 * vx, atol=2e-7)
    # dcd: 92-byte fixed header + title block + atom count, then 5 frames (steps 0, 20, ..., 80)
    raw = open(tmp_path / "plane_test.dcd", "rb").read()
    head = 4 + 84 + 4 + 4 + 4 + 160 + 4 + 4 + 4 + 4
    assert raw[4:8] == b"CORD" and struct.unpack("<i", raw[head - 8:head - 4])[0] == n
    frame = 4 + 48 + 4 + 3 * (4 + 4 * n + 4)
    assert len(raw) == head + 5 * frame
    f5 = raw[head + 4 * frame:]
    xs = np.frombuffer(f5[4 + 48 + 4 + 4:4 + 48 + 4 + 4 + 4 * n], dtype=np.float32)
    np.testing.assert_allclose(xs, x.astype(np.float32), atol=1e-4)
    # vtp: accepted, nothing written (a build without VTK)
    assert not any(f.endswith(".vtp") for f in os.listdir(tmp_path))
