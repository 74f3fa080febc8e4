"""End to end through the C++ host application on the GPU: TOML configuration ->
FastHLLMUSCLBathyHIPSolver (C ABI, libswe_hip.so) -> NetCDF snapshots, each snapshot bitwise
equal to the oracle's run of the same initial state, plus the sanity-check file."""
import os
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

import oracle
from test_host_app import APP, HOST, config

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def app():
    subprocess.run(["make", "-s", "-C", HOST], check=True)


def read(path):
    f = netcdf_file(str(path), "r", mmap=False)
    # the file is big-endian; values are converted exactly to the native order
    d = {k: np.asarray(f.variables[k][:]).astype(f.variables[k][:].dtype.newbyteorder("="))
         for k in f.variables}
    a = dict(f._attributes)
    f.close()
    return d, a


def test_app_matches_oracle_and_writes_records(tmp_path):
    over = dict(mesh__Nx=50, mesh__Ny=70, mesh__Lx=5.0, mesh__Ly=7.0, initial_condition__x0=3.0,
                time__end_time=0.06,
                time__time_steps=120, time__save_every=40,
                sanity_checks__mass_conservation="true", sanity_checks__debug="true")
    cfg = config(tmp_path, **over)
    r = subprocess.run([APP, cfg], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "HIP solver timing summary" in r.stdout
    d, attrs = read(tmp_path / "out" / "sim.nc")
    assert d["h"].shape == (4, 50, 70)
    assert attrs["backend"] == b"HIP" and attrs["dt"] == 0.06 / 120

    # the oracle from the written initial state (ghosts mirrored as at solver set-up)
    nx, ny, ng = 50, 70, 2
    gr = oracle.grid(nx, ny, ng, 5.0 / nx, 7.0 / ny)
    full = lambda a: np.pad(a, ng)
    h, hu, hv, B = (np.ascontiguousarray(full(a)) for a in
                    (d["h"][0], d["hu"][0], d["hv"][0], d["B"]))
    oracle.apply_bc(gr, B)
    oracle.apply_bc(gr, h, hu, hv)
    dt = 0.06 / 120
    t = 0.0
    for rec in range(1, 4):
        rc, _, _ = oracle.run(gr, h, hu, hv, B, 40, dt=dt)
        assert rc == 0
        for _ in range(40):
            t += dt
        for name, want in (("h", h), ("hu", hu), ("hv", hv)):
            got = d[name][rec]
            assert np.array_equal(got.view(np.uint64),
                                  np.ascontiguousarray(want[ng:-ng, ng:-ng]).view(np.uint64)), (rec, name)
        assert d["time"][rec] == t

    s, sattrs = read(tmp_path / "out" / "sanity.nc")
    assert s["step"].tolist() == [0, 40, 80, 120]
    assert np.all(s["mass_rel_err"] < 1e-12) and np.all(s["h_min"] > 0.5)
    assert sattrs["backend"] == b"HIP"


def test_app_rejects_cfl_violation(tmp_path):
    cfg = config(tmp_path, time__end_time=5.0, time__time_steps=100, time__save_every=50)
    r = subprocess.run([APP, cfg], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 1 and "Time step too large" in r.stderr


@pytest.mark.parametrize("recon,riemann,bathy", [("PiecewiseConst", "Rusanov", "None"),
                                                 ("MUSCL", "ROE", "None"),
                                                 ("PiecewiseConst", "HLL", "GaussHill")])
def test_serial_backend_matches_modular_oracle(tmp_path, recon, riemann, bathy):
    """[backend] type = "HIP_Serial": the reference's modular SerialSolver on the GPU; every
    snapshot bitwise equal to the modular oracle's run of the written initial state."""
    over = dict(mesh__Nx=36, mesh__Ny=64, mesh__Lx=4.5, mesh__Ly=8.0, initial_condition__x0=3.0,
                initial_condition__y0=5.0, time__end_time=0.03, time__time_steps=60,
                time__save_every=30, backend__type='"HIP_Serial"',
                solver__reconstruction=f'"{recon}"', solver__riemann=f'"{riemann}"',
                bathymetry__type=f'"{bathy}"')
    cfg = config(tmp_path, **over)
    r = subprocess.run([APP, cfg], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert f"HIP serial-scheme solver active ({recon} + {riemann})" in r.stdout
    d, attrs = read(tmp_path / "out" / "sim.nc")
    assert attrs["backend"] == b"HIP_Serial"
    assert attrs["riemann_solver"] == riemann.encode()
    assert attrs["reconstruction"] == recon.encode()
    nx, ny, ng = 36, 64, 2
    gr = oracle.grid(nx, ny, ng, 4.5 / nx, 8.0 / ny)
    full = lambda a: np.pad(a, ng)
    h, hu, hv, B = (np.ascontiguousarray(full(a)) for a in
                    (d["h"][0], d["hu"][0], d["hv"][0], d["B"]))
    oracle.mod_apply_bc(gr, B)
    oracle.mod_apply_bc(gr, h, hu, hv)
    bathy_on = bathy != "None"
    rk = {"MUSCL": oracle.RECON_MUSCL, "PiecewiseConst": oracle.RECON_PC}[recon]
    rs = {"HLL": oracle.RIEMANN_HLL, "Rusanov": oracle.RIEMANN_RUSANOV,
          "ROE": oracle.RIEMANN_ROE}[riemann]
    for rec in range(1, 3):
        assert oracle.mod_run(gr, bathy_on, rk, rs, h, hu, hv, B, 30, 0.03 / 60) == 0
        for name, want in (("h", h), ("hu", hu), ("hv", hv)):
            got = d[name][rec]
            assert np.array_equal(got.view(np.uint64),
                                  np.ascontiguousarray(want[ng:-ng, ng:-ng]).view(np.uint64)), (rec, name)


def test_transmissive_boundary_matches_oracle(tmp_path):
    """[boundary] type = "Transmissive" with the optimized HIP scheme: zero-gradient ghosts
    (an extension; the oracle's transmissive fill is pinned by its closed form)."""
    over = dict(mesh__Nx=40, mesh__Ny=66, mesh__Lx=4.0, mesh__Ly=6.6, initial_condition__x0=1.0,
                initial_condition__y0=1.5, time__end_time=0.04, time__time_steps=80,
                time__save_every=40, boundary__type='"Transmissive"')
    cfg = config(tmp_path, **over)
    r = subprocess.run([APP, cfg], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    d, attrs = read(tmp_path / "out" / "sim.nc")
    assert attrs["boundary_condition"] == b"Transmissive"
    nx, ny, ng = 40, 66, 2
    gr = oracle.grid(nx, ny, ng, 4.0 / nx, 6.6 / ny, bc=oracle.BC_TRANSMISSIVE)
    full = lambda a: np.pad(a, ng)
    h, hu, hv, B = (np.ascontiguousarray(full(a)) for a in
                    (d["h"][0], d["hu"][0], d["hv"][0], d["B"]))
    oracle.apply_bc(gr, B)
    oracle.apply_bc(gr, h, hu, hv)
    for rec in range(1, 3):
        rc, _, _ = oracle.run(gr, h, hu, hv, B, 40, dt=0.04 / 80)
        assert rc == 0
        for name, want in (("h", h), ("hu", hu), ("hv", hv)):
            got = d[name][rec]
            assert np.array_equal(got.view(np.uint64),
                                  np.ascontiguousarray(want[ng:-ng, ng:-ng]).view(np.uint64)), (rec, name)
