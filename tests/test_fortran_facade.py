"""The Fortran host side: dycore_mod facade + dycore_test driver over the C ABI
(gamil-dycore_amd/fortran).  CPU tests parse every committed namelist (and
the reference's own run/namelist.* files where that tree is present) with
`dycore_test.exe <nml> --check`; the GPU test runs the driver end to end and
compares it with the same run through the Python mirror of the C ABI."""
import glob
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

EXE = os.path.join(ROOT, "gamil-dycore_amd", "dycore_test.exe")
NAMELISTS = sorted(glob.glob(os.path.join(ROOT, "namelists", "*.nml")))
REF_RUN = "/root/reference/run"


def check(path):
    r = subprocess.run([EXE, path, "--check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {}
    for line in r.stdout.splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            out[k.strip()] = v.strip()
    return out


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        pytest.skip("dycore_test.exe not built")
    return EXE


@pytest.mark.parametrize("path", NAMELISTS, ids=[os.path.basename(p) for p in NAMELISTS])
def test_committed_namelists_parse(exe, gd, path):
    out = check(path)
    nlon, nlat = [int(x) for x in out["grid"].split("x")]
    text = "\n".join(ln.split("!")[0] for ln in open(path).read().splitlines())
    m = re.search(r"zonal_tend_filter_lat0\s*=\s*([0-9.]+)", text)
    if m:
        expect = gd.cutoff_from_lat0(nlon, nlat, float(m.group(1)))
        assert [int(c) for c in out["filter cutoffs"].split()] == expect
    days = int(re.search(r"\b(?:run_)?days\s*=\s*(\d+)", text).group(1))
    dt = float(re.search(r"time_step_size\s*=\s*([0-9.]+)", text).group(1))
    assert int(out["steps"]) == int(np.ceil(days * 86400 / dt))


@pytest.mark.parametrize("name", ["namelist.mz_test", "namelist.rh_test", "namelist.jz_test"])
def test_reference_namelists_parse(exe, name):
    """The superset accepts the reference's own namelists, including the keys
    its params_mod never declared (run/namelist.rh_test:6, jz_test:16-19)."""
    path = os.path.join(REF_RUN, name)
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    out = check(path)
    assert out["case_name"]


def test_missing_namelist_fails_cleanly(exe, tmp_path):
    r = subprocess.run([EXE, str(tmp_path / "nope.nml"), "--check"], capture_output=True, text=True)
    assert r.returncode == 1
    assert "[Error]" in r.stdout


SMALL = """&dycore_params
test_case = '{case}'
case_name = 'small.{case}'
run_days = 1
num_lon = 72
num_lat = 37
time_step_size = 1200
time_scheme = 'predict_correct'
qcon_modified = .true.
split_scheme = 'csp2'
subcycles = 4
use_diffusion = .true.
diffusion_coef = 4.0e5
use_zonal_tend_filter = .true.
zonal_tend_filter_lat0 = 75.0
precision = 'fp64'
log_every = 1
/
&balanced_sh_params
seed = 42, nmax = 8, n_mountains = 1, mountain_seed = 3
/
"""


@pytest.mark.gpu
def test_driver_matches_python_path(exe, gd, tmp_path):
    nml = tmp_path / "small.nml"
    nml.write_text(SMALL.format(case="balanced_sh"))
    r = subprocess.run([exe, str(nml)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith(" => ")]
    assert len(lines) == 72 + 1  # initial diag + one line per step
    mass_f = np.array([float(ln.split()[2]) for ln in lines])
    energy_f = np.array([float(ln.split()[3]) for ln in lines])
    N, M = 72, 37
    cut = gd.cutoff_from_lat0(N, M, 75.0)
    ic = gd.ic_balanced(N, M, seed=42, nmax=8, n_mountains=1, mountain_seed=3)
    with gd.Dycore(N, M, 1200.0, split_scheme="csp2", subcycles=4, use_diffusion=True,
                   diffusion_coef=4.0e5, zonal_tend_filter_cutoff_wavenumber=cut) as d:
        d.set_state(*ic)
        m0, e0 = d.diag()
        mass, energy, beta, _ = d.step(72)
    np.testing.assert_array_equal(mass_f[1:], mass[:, 0])
    np.testing.assert_array_equal(energy_f[1:], energy[:, 0])
    assert mass_f[0] == m0[0] and energy_f[0] == e0[0]


@pytest.mark.gpu
def test_driver_rossby_haurwitz(exe, tmp_path):
    nml = tmp_path / "rh.nml"
    nml.write_text(SMALL.format(case="rossby_haurwitz_wave") + "&rossby_haurwitz_wave_test_params\n/\n")
    r = subprocess.run([exe, str(nml)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "final total_mass" in r.stdout


@pytest.mark.gpu
def test_driver_autotune_gives_the_same_run(exe, tmp_path):
    """autotune_steps in &dycore_params runs gd_autotune on the handed-over
    state before the loop: the per-step log lines (mass, energy, beta printed
    to 21 digits) are those of the untuned run; the matrix-core filter mode is
    reachable from the namelist too."""
    base = SMALL.format(case="balanced_sh")
    runs = {}
    for name, extra in (("plain", ""), ("tuned", "autotune_steps = 5\nautotune_cache = 'tune.txt'\n"),
                        ("mfma", "precision = 'mixed_mfma'\n")):
        text = base.replace("log_every = 1\n", "log_every = 1\n" + extra)
        nml = tmp_path / f"{name}.nml"
        nml.write_text(text)
        print(f"driver run: {name}", flush=True)
        r = subprocess.run([exe, str(nml)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
        assert r.returncode == 0, r.stdout + r.stderr
        runs[name] = r.stdout
    steps = lambda out: [ln for ln in out.splitlines() if ln.startswith(" => ")]
    assert steps(runs["tuned"]) == steps(runs["plain"])
    assert "autotune: fused = " in runs["tuned"]
    assert (tmp_path / "tune.txt").read_text().count("N=72 M=37") == 1
    assert len(steps(runs["mfma"])) == 73
