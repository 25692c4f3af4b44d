"""History and restart output of device runs: the Python mirror writing the
device state, and the Fortran driver writing files at its output alerts and
restarting from one of its own restart files."""
import os
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

from conftest import ROOT

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "gamil-dycore_amd", "dycore_test.exe")


def read(path):
    with netcdf_file(str(path), "r", mmap=False) as f:
        gatt = {k: (v.decode() if isinstance(v, bytes) else v) for k, v in f._attributes.items()}
        var = {k: f.variables[k].data.copy() for k in f.variables}
        units = f.variables["time"].units.decode()
    return gatt, var, units


def test_dycore_history_of_device_state(gd, tmp_path):
    N, M = 72, 37
    ic = gd.ic_balanced(N, M, seed=4, nmax=8, n_mountains=1, mountain_seed=2)
    with gd.Dycore(N, M, 900.0, split_scheme="csp2", subcycles=3, use_diffusion=True,
                   diffusion_coef=2e5, zonal_tend_filter_cutoff_wavenumber=[6, 4, 2]) as d:
        d.set_state(*ic)
        d.step(6)
        p = tmp_path / "h.nc"
        d.write_history(str(p), time_value=6 * 900 / 86400.0, split_scheme="csp2", subcycles=3)
        u, v, g = d.get_state()
        div, vor = d.diag_fields()
        tm, te = d.diag()
    gatt, var, units = read(p)
    np.testing.assert_array_equal(var["u"][0], 0.5 * (u + np.roll(u, 1, axis=1)))
    np.testing.assert_array_equal(var["v"][0][1:M - 1], 0.5 * (v[1:] + v[:-1]))
    np.testing.assert_array_equal(var["gh"][0], g + ic[3])
    np.testing.assert_array_equal(var["div"][0], div)
    np.testing.assert_array_equal(var["vor"][0], vor)
    assert var["te"][0] == te[0] and var["tm"][0] == tm[0]
    assert gatt["subcycles"] == 3


NML = """&dycore_params
test_case = 'balanced_sh'
case_name = '{case}'
case_desc = 'io test'
run_days = 1
num_lon = 72
num_lat = 37
time_step_size = 1200
qcon_modified = .true.
split_scheme = 'csp2'
subcycles = 4
use_diffusion = .true.
diffusion_coef = 4.0e5
zonal_tend_filter_lat0 = 75.0
precision = 'fp64'
history_periods = '6 hours'
restart_period = '12 hours'
log_every = 12
{extra}
/
&balanced_sh_params
seed = 42, nmax = 8, n_mountains = 1, mountain_seed = 3
/
"""


def run_driver(tmp_path, case, extra=""):
    nml = tmp_path / f"{case}.nml"
    nml.write_text(NML.format(case=case, extra=extra))
    print("driver run", case, flush=True)
    r = subprocess.run([EXE, str(nml)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_driver_output_and_restart(gd, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("dycore_test.exe not built")
    out = run_driver(tmp_path, "c")
    files = sorted(os.listdir(tmp_path))
    hist = [f for f in files if f.startswith("c.h0.h0.")]
    rest = [f for f in files if f.startswith("c.r.")]
    assert hist == [f"c.h0.h0.0001-01-0{d}T{t}.nc" for d, t in
                    ((1, "06:00:00"), (1, "12:00:00"), (1, "18:00:00"), (2, "00:00:00"))]
    assert rest == ["c.r.0001-01-01T12:00:00.nc", "c.r.0001-01-02T00:00:00.nc"]
    gatt, var, units = read(tmp_path / hist[-1])
    assert units == "days since 0001-01-01T00_00_00" and var["time"].tolist() == [1.0]
    assert gatt["desc"] == "io test" and gatt["split_scheme"] == "csp2"
    final_mass = float([ln for ln in out.splitlines() if "final total_mass" in ln][0].split("=")[1])
    final_energy = float([ln for ln in out.splitlines() if "final total_energy" in ln][0].split("=")[1])
    assert var["tm"][0] == final_mass and var["te"][0] == final_energy

    # restart from the 12 h file: the alerts began at the namelist start, so
    # both ring at once at the restart time (time_mod.F90:181-183)
    out2 = run_driver(tmp_path, "c2", "restart_file = 'c.r.0001-01-01T12:00:00.nc'")
    assert "Reset time to 0001-01-01T12_00_00." in out2
    files = sorted(os.listdir(tmp_path))
    assert [f for f in files if f.startswith("c2.h0.h0.")] == [
        "c2.h0.h0.0001-01-01T12_00_00.nc", "c2.h0.h0.0001-01-01T18:00:00.nc",
        "c2.h0.h0.0001-01-02T00:00:00.nc"]
    assert [f for f in files if f.startswith("c2.r.")] == [
        "c2.r.0001-01-01T12_00_00.nc", "c2.r.0001-01-02T00:00:00.nc"]
    g2, v2, units2 = read(tmp_path / "c2.h0.h0.0001-01-02T00:00:00.nc")
    assert units2 == "days since 0001-01-01T12_00_00" and v2["time"].tolist() == [0.5]
    # with diffusion every step ends with iap_transform of (u, v, gd), so the
    # restarted run continues bit for bit
    _, a, _ = read(tmp_path / "c.r.0001-01-02T00:00:00.nc")
    _, b, _ = read(tmp_path / "c2.r.0001-01-02T00:00:00.nc")
    for k in ("u", "v", "gd", "ghs"):
        np.testing.assert_array_equal(a[k], b[k])
    assert v2["tm"][0] == var["tm"][0] and v2["te"][0] == var["te"][0]


def test_driver_ensemble_history_members(gd, tmp_path):
    """ensemble_members = 2 through the driver: one history file per member
    and output time, each member's history equal to the same run alone."""
    if not os.path.exists(EXE):
        pytest.skip("dycore_test.exe not built")
    nml = NML.format(case="e", extra="").replace("&balanced_sh_params\nseed = 42, nmax = 8, n_mountains = 1, "
                                                  "mountain_seed = 3\n/\n", "")
    nml = nml.replace("log_every = 12\n", "log_every = 12\nensemble_members = 2\nrun_hours = 12\n") \
             .replace("run_days = 1\n", "")
    nml += "&balanced_sh_params\nseed = 42, nmax = 8, n_mountains = 1, mountain_seed = 3, " \
           "pert_seed = 11, pert_nmax = 6, pert_rms = 5.0\n/\n"
    (tmp_path / "e.nml").write_text(nml)
    r = subprocess.run([EXE, str(tmp_path / "e.nml")], capture_output=True, text=True, timeout=120,
                       cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    files = sorted(f for f in os.listdir(tmp_path) if f.endswith(".nc"))
    assert files == ["e.m001.h0.h0.0001-01-01T06:00:00.nc", "e.m001.h0.h0.0001-01-01T12:00:00.nc",
                     "e.m001.r.0001-01-01T12:00:00.nc", "e.m002.h0.h0.0001-01-01T06:00:00.nc",
                     "e.m002.h0.h0.0001-01-01T12:00:00.nc", "e.m002.r.0001-01-01T12:00:00.nc"]
    N, M = 72, 37
    cut = gd.cutoff_from_lat0(N, M, 75.0)
    for m in (1, 2):
        ic = gd.ic_balanced(N, M, seed=42, nmax=8, n_mountains=1, mountain_seed=3, pert_seed=11 + m - 1,
                            pert_nmax=6, pert_rms=5.0)
        with gd.Dycore(N, M, 1200.0, split_scheme="csp2", subcycles=4, use_diffusion=True,
                       diffusion_coef=4.0e5, zonal_tend_filter_cutoff_wavenumber=cut) as d:
            d.set_state(*ic)
            d.step(36)
            u, v, g = d.get_state()
        _, var, _ = read(tmp_path / f"e.m00{m}.h0.h0.0001-01-01T12:00:00.nc")
        np.testing.assert_array_equal(var["gh"][0], g + ic[3])
        np.testing.assert_array_equal(var["u"][0], 0.5 * (u + np.roll(u, 1, axis=1)))
