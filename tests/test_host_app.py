"""The C++ host application (shallow-water_amd/host): configuration loading with the
reference's schema plus the "HIP" backend key, validation, the initial-state generators and
the NetCDF classic output, all without a GPU (--print-config, --initial-state).

Generators are pinned bitwise to the reference's own setup code (apply_initial_condition /
apply_bathymetry compiled from the reference tree into oracle/_ref; golden states in
tests/golden/setup_states.npz) and by their closed forms; the output files are read back with
scipy.io.netcdf_file.
"""
import json
import os
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(REPO, "shallow-water_amd", "host")
APP = os.path.join(HOST, "build", "swe_hip_app")

BASE = {
    "mesh": dict(Lx=10.0, Ly=8.0, nG=2, Nx=40, Ny=32, spatial_unit_x='"m"', spatial_unit_y='"m"',
                 spatial_unit_h='"m"'),
    "bathymetry": dict(type='"GaussHill"', b0=0.1, bathy_x0=4.0, bathy_y0=3.0, bathy_sigma_x=1.5,
                       bathy_sigma_y=0.75, bathy_peak_height=0.4),
    "boundary": dict(type='"ReflectingWalls"'),
    "initial_condition": dict(type='"GaussInitial"', h0=1.0, peak_height=0.5, x0=6.0, y0=4.0,
                              sigma_x=1, sigma_y=0.5, dam_height=2.5, dam_x=5.0, dam_x0=5.0,
                              dam_y0=4.0, dam_radius=2.0),
    "solver": dict(reconstruction='"MUSCL"', riemann='"HLL"', time='"SSPRK3"'),
    "time": dict(end_time=0.1, time_steps=100, cfl=0.45, time_unit='"s"', save_every=50),
    "backend": dict(type='"HIP"', threads=1),
    "sanity_checks": dict(mass_conservation="false", mass_threshold=1e-11, positivity="true",
                          debug="false", output_path='"out/sanity.nc"'),
    "output": dict(compute_eta="false", path='"out/sim.nc"'),
}


def toml_text(cfg):
    lines = []
    for table, kv in cfg.items():
        lines.append(f"[{table}]")
        for k, v in kv.items():
            lines.append(f"{k} = {v}   # comment")
        lines.append("")
    return "\n".join(lines)


def config(tmp_path, **over):
    cfg = {t: dict(kv) for t, kv in BASE.items()}
    for key, v in over.items():
        table, k = key.split("__")
        if v is None:
            cfg[table].pop(k)
        else:
            cfg[table][k] = v
    p = tmp_path / "cfg.toml"
    p.write_text(toml_text(cfg))
    return str(p)


@pytest.fixture(scope="module", autouse=True)
def app():
    subprocess.run(["make", "-s", "-C", HOST], check=True)
    return APP


def run(*args, cwd=None):
    return subprocess.run([APP, *args], capture_output=True, text=True, cwd=cwd)


def test_reference_schema_and_hip_key(tmp_path):
    r = run("--print-config", config(tmp_path))
    assert r.returncode == 0, r.stderr
    vals = dict(line.split(" = ", 1) for line in r.stdout.strip().splitlines())
    assert vals["backend.type"] == "HIP"
    assert vals["mesh.Nx"] == "40" and float(vals["initial_condition.h0"]) == 1.0
    assert float(vals["sanity_checks.mass_threshold"]) == 1e-11
    # integer literals where floats are expected (sigma_x = 1) are accepted, as in toml++
    r = run("--print-config", config(tmp_path, backend__type='"OpenMP"'))
    assert r.returncode == 0 and "backend.type = OpenMP" in r.stdout


def test_other_backends_are_reported(tmp_path):
    r = run(config(tmp_path, backend__type='"CUDA"'))
    assert r.returncode == 1 and "HIP backend only" in r.stderr


def test_hip_serial_key_and_combinations(tmp_path):
    """HIP_Serial accepts exactly the combinations of the reference's run_serial
    (solver_assembly.hpp:315-500); the others fail before any device work."""
    r = run("--print-config", config(tmp_path, backend__type='"HIP_Serial"'))
    assert r.returncode == 0 and "backend.type = HIP_Serial" in r.stdout
    for over in ({"solver__riemann": '"ROE"'},                       # bathymetry: HLL only
                 {"solver__riemann": '"Rusanov"', "bathymetry__type": '"Flat"'},
                 {"boundary__type": '"Transmissive"'}):                # reflecting walls only
        r = run(config(tmp_path, backend__type='"HIP_Serial"', **over))
        assert r.returncode == 1 and "Unsupported serial solver combination" in r.stderr, r.stderr


@pytest.mark.parametrize("over,msg", [
    ({"mesh__Nx": None}, "Missing or invalid TOML value for key 'mesh.Nx'"),
    ({"mesh__Nx": "40.5"}, "Missing or invalid TOML value for key 'mesh.Nx'"),
    ({"time__cfl": '"fast"'}, "Missing or invalid TOML value for key 'time.cfl'"),
    ({"backend__type": '"Vulkan"'}, "Unknown backend type: Vulkan"),
    ({"bathymetry__type": '"Ramp"'}, "Unknown bathymetry type: Ramp"),
    ({"time__save_every": 500}, "save_every must not be larger than time_steps"),
    ({"time__cfl": 1.5}, "cfl should usually be <= 1.0"),
    ({"mesh__nG": 1}, "MUSCL reconstruction requires nG >= 2"),
    ({"initial_condition__sigma_x": 0}, "Gaussian initial-condition sigmas must be positive"),
    ({"bathymetry__bathy_x0": 11.0}, "Gaussian bathymetry center must lie inside the domain"),
])
def test_config_errors(tmp_path, over, msg):
    r = run("--print-config", config(tmp_path, **over))
    assert r.returncode == 1
    assert msg in r.stderr, r.stderr


def test_toml_syntax_errors(tmp_path):
    p = tmp_path / "bad.toml"
    p.write_text(toml_text(BASE) + "\n[mesh]\nLx = 3\n")
    r = run("--print-config", str(p))
    assert r.returncode == 1 and "defined twice" in r.stderr
    p.write_text("[mesh]\nLx = 1\nLx = 2\n")
    r = run("--print-config", str(p))
    assert r.returncode == 1 and "duplicate key" in r.stderr
    p.write_text("[mesh]\nLx = [1, 2]\n")
    r = run("--print-config", str(p))
    assert r.returncode == 1 and "not supported" in r.stderr


def centres(n, L):
    return (np.arange(n) + 0.5) * (L / n)


def read_initial(tmp_path, **over):
    out = tmp_path / "init.nc"
    r = run("--initial-state", str(out), config(tmp_path, **over))
    assert r.returncode == 0, r.stderr
    f = netcdf_file(str(out), "r", mmap=False)
    data = {k: f.variables[k][:].copy() for k in f.variables}
    attrs = dict(f._attributes)
    units = {k: getattr(f.variables[k], "units", None) for k in f.variables}
    f.close()
    return data, attrs, units


def test_gauss_initial_and_gauss_hill(tmp_path):
    d, attrs, units = read_initial(tmp_path)
    x, y = centres(40, 10.0), centres(32, 8.0)
    X, Y = np.meshgrid(x, y, indexing="ij")
    h = 1.0 + 0.5 * np.exp(-((X - 6.0) ** 2) / (2 * 1.0 ** 2) - ((Y - 4.0) ** 2) / (2 * 0.5 ** 2))
    B = 0.1 + 0.4 * np.exp(-((X - 4.0) ** 2) / (2 * 1.5 ** 2) - ((Y - 3.0) ** 2) / (2 * 0.75 ** 2))
    assert d["h"].shape == (1, 40, 32) and d["B"].shape == (40, 32)
    np.testing.assert_allclose(d["h"][0], h, rtol=1e-15, atol=0)
    np.testing.assert_allclose(d["B"], B, rtol=1e-15, atol=0)
    assert not d["hu"].any() and not d["hv"].any()
    np.testing.assert_array_equal(d["x"], x)
    np.testing.assert_array_equal(d["y"], y)
    assert d["time"].tolist() == [0.0]
    assert attrs["save_every"] == 50 and attrs["dt"] == 0.1 / 100
    assert attrs["backend"] == b"HIP" and attrs["bathymetry"] == b"Gaussian Hill"
    assert attrs["riemann_solver"] == b"HLL" and attrs["boundary_condition"] == b"Reflecting Walls"
    assert units["h"] == b"m" and units["time"] == b"s" and units["B"] == b"m"


@pytest.mark.parametrize("ic", ["DamBreak", "DamBreakRadial", "StillWater"])
def test_other_initial_conditions(tmp_path, ic):
    d, _, _ = read_initial(tmp_path, initial_condition__type=f'"{ic}"',
                           bathymetry__type='"Flat"')
    x, y = centres(40, 10.0), centres(32, 8.0)
    X, Y = np.meshgrid(x, y, indexing="ij")
    if ic == "DamBreak":  # h0 below dam_x, dam_height beyond
        want = np.where(X < 5.0, 1.0, 2.5)
    elif ic == "DamBreakRadial":
        want = np.where((X - 5.0) ** 2 + (Y - 4.0) ** 2 <= 2.0 ** 2, 2.5, 1.0)
    else:
        want = np.ones_like(X)
    np.testing.assert_array_equal(d["h"][0], want)
    np.testing.assert_array_equal(d["B"], np.full((40, 32), 0.1))


def test_no_bathymetry(tmp_path):
    d, attrs, _ = read_initial(tmp_path, bathymetry__type='"None"')
    assert not d["B"].any() and attrs["bathymetry"] == b"None"


GOLDEN = os.path.join(REPO, "tests", "golden")
SETUPS = json.load(open(os.path.join(GOLDEN, "setup_states.json")))


def setup_overrides(case):
    """TOML overrides that make the app's configuration the golden case's."""
    m = case["mesh"]
    over = {"mesh__Nx": m["nx"], "mesh__Ny": m["ny"], "mesh__nG": m["ng"],
            "mesh__Lx": repr(float(m["lx"])), "mesh__Ly": repr(float(m["ly"]))}
    for table, key in (("initial_condition", "initial_condition"), ("bathymetry", "bathymetry")):
        for k, v in case[key].items():
            over[f"{table}__{k}"] = f'"{v}"' if k == "type" else repr(float(v))
    if case["bathymetry"]["type"] != "GaussHill":
        # the app validates the hill's centre even when unused; keep it inside the domain
        over["bathymetry__bathy_x0"] = repr(float(m["lx"]) / 2)
        over["bathymetry__bathy_y0"] = repr(float(m["ly"]) / 2)
    if case["initial_condition"]["type"] != "GaussInitial":
        over["initial_condition__sigma_x"] = "1.0"
        over["initial_condition__sigma_y"] = "1.0"
    return over


@pytest.mark.parametrize("name", sorted(SETUPS))
def test_initial_state_bitwise_vs_reference_setup(tmp_path, name):
    """The app's initial arrays (the NetCDF snapshot 0 and B, interior cells as the
    reference writes them) equal, bit for bit, the reference's own generators for every
    initial-condition x bathymetry type and for the TOML forms of configs 1 and 3."""
    case = SETUPS[name]
    d, _, _ = read_initial(tmp_path, **setup_overrides(case))
    z = np.load(os.path.join(GOLDEN, "setup_states.npz"))
    ng = case["mesh"]["ng"]
    inner = (slice(ng, -ng), slice(ng, -ng))
    for k, got in (("h", d["h"][0]), ("hu", d["hu"][0]), ("hv", d["hv"][0]), ("B", d["B"])):
        want = z[f"{name}__{k}"][inner]
        assert got.shape == want.shape, (name, k)
        got = np.ascontiguousarray(got, dtype=np.float64)  # NetCDF data is big-endian
        assert np.array_equal(got.view(np.uint64),
                              np.ascontiguousarray(want).view(np.uint64)), (name, k)


def test_setup_fixtures_are_the_reference_setup_code():
    """The golden initial states are what the reference's compiled setup code produces now
    (pins the fixtures; needs oracle/_ref)."""
    import sys

    sys.path.insert(0, REPO)
    import oracle

    if not oracle.ref_available():
        pytest.skip("reference harness not built (oracle/_ref)")
    z = np.load(os.path.join(GOLDEN, "setup_states.npz"))
    for name, case in SETUPS.items():
        m = case["mesh"]
        ic = dict(case["initial_condition"])
        bt = dict(case["bathymetry"])
        arrays = oracle.ref_initial_state(m["nx"], m["ny"], m["ng"], m["lx"], m["ly"],
                                          ic.pop("type"), ic, bt.pop("type"), bt)
        for k, a in zip(("h", "hu", "hv", "B"), arrays):
            assert np.array_equal(a.view(np.uint64), z[f"{name}__{k}"].view(np.uint64)), (name, k)
