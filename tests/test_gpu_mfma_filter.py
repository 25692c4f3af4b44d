"""The polar filter of the momentum rows on the matrix cores (precision
'mixed_mfma': FP32 inputs, 'mixed_bf16': BF16 inputs, both with FP32
accumulation) against the oracle's FFTPACK-semantics filter
(filter_mod.F90:105-167) on the filter tables of the 1440x721, 720x361 and
360x181 configurations.  The full rows (dgd) stay on the FP64 FFT.

Measured errors (max over flagged rows of |y - y_ref| / max |y_ref|) are
printed; the bounds are FP32: 2e-5 (measured 4e-7 to 1.5e-6) and BF16 with
split operands (hi.hi + hi.lo + lo.hi): 1e-4.  A single BF16 product
measured 4e-3 to 5.7e-3 here and, through the rescale, up to 0.15 of the
field maximum in the fast-pass dv tendency."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

BOUND = {"mixed_mfma": 2e-5, "mixed_bf16": 1e-4}


@pytest.mark.parametrize("prec", ["mixed_mfma", "mixed_bf16"])
@pytest.mark.parametrize("N,M", [(360, 181), (720, 361), (1440, 721), (90, 46), (75, 38)])
def test_mfma_projector_vs_oracle(gd, oracle_mod, N, M, prec):
    cut = gd.cutoff_from_lat0(N, M, 75.0 if N >= 360 else 40.0)
    o = oracle_mod.Oracle(N, M, 600.0, cutoff=cut)
    rng = np.random.default_rng(N)
    print(f"mfma projector {N}x{M} {prec}", flush=True)
    with gd.Dycore(N, M, 600.0, zonal_tend_filter_cutoff_wavenumber=cut, precision=prec) as d:
        ff, hf, _, hc = o.filter_table()
        xh = rng.standard_normal((M - 1, N))
        yh = d.filter_rows(xh, half=True)
        worst = 0.0
        for j in range(M - 1):
            if hf[j]:
                worst = max(worst, rel_err(yh[j], o.filter_half(j + 1, xh[j])))
            else:
                np.testing.assert_array_equal(yh[j], xh[j])
        print(f"  half rows: {int(hf.sum())} flagged, max rel err {worst:.2e}", flush=True)
        assert worst <= BOUND[prec]
        assert hf.sum() > 0
        # full rows are the mass rows: FP64 FFT in every mode
        x = rng.standard_normal((M, N))
        y = d.filter_rows(x, half=False)
        for j in range(M):
            if ff[j]:
                assert rel_err(y[j], o.filter_full(j + 1, x[j])) < 1e-13


@pytest.mark.parametrize("prec", ["mixed_mfma", "mixed_bf16"])
def test_mfma_space_operators_vs_oracle(gd, oracle_mod, prec):
    N, M = 360, 181
    u, v, h, ghs = gd.ic_balanced(N, M, seed=5, nmax=20, n_mountains=3, mountain_seed=9)
    cut = gd.cutoff_from_lat0(N, M, 75.0)
    o = oracle_mod.Oracle(N, M, 240.0, cutoff=cut)
    o.set_initial_state(u, v, h, ghs)
    o.start()
    print(f"mfma space operators {prec}", flush=True)
    with gd.Dycore(N, M, 240.0, zonal_tend_filter_cutoff_wavenumber=cut, precision=prec) as d:
        d.set_state(u, v, h, ghs)
        for pname, pid in (("all", 0), ("fast", 1), ("slow", 2)):
            o.space_operators(o.old, 0, pid)
            du, dv, dgd = d.space_operators(pname)
            edu = rel_err(du, o.tend(0, oracle_mod.T_DU))
            edv = rel_err(dv, o.tend(0, oracle_mod.T_DV))
            print(f"  {pname}: du {edu:.2e} dv {edv:.2e}", flush=True)
            # the projection's error grows through the rescale s1/s2 of rows
            # whose weighted sum nearly cancels (BF16 split operands: measured
            # 1.4e-4 du, 6.6e-4 dv in the fast pass)
            lim = 5e-4 if prec == "mixed_mfma" else 2e-3
            assert edu < lim and edv < lim
            if pname != "slow":
                assert rel_err(dgd, o.tend(0, oracle_mod.T_DGD)) < 1e-12


@pytest.mark.parametrize("prec", ["mixed_mfma", "mixed_bf16"])
def test_mfma_member_alone_equals_member_in_batch(gd, prec):
    """A member's rows form one row of the batched products: alone (a block
    of one) and at position 17 of 20 (second block) it gets the same bits."""
    N, M = 180, 91
    cut = gd.cutoff_from_lat0(N, M, 60.0)
    kw = dict(split_scheme="csp2", subcycles=3, use_diffusion=True, diffusion_coef=1e5, precision=prec,
              zonal_tend_filter_cutoff_wavenumber=cut)
    ics = [gd.ic_balanced(N, M, seed=8, nmax=10, pert_seed=40 + m, pert_nmax=8, pert_rms=5.0) for m in range(20)]
    print(f"mfma batch invariance {prec}", flush=True)
    with gd.Dycore(N, M, 480.0, members=20, **kw) as d:
        for m, ic in enumerate(ics):
            d.set_state(*ic, member=m)
        mass, energy, beta, _ = d.step(4)
        batch = d.get_state(member=17)
    with gd.Dycore(N, M, 480.0, **kw) as d:
        d.set_state(*ics[17])
        m1, e1, b1, _ = d.step(4)
        alone = d.get_state()
    for x, y in zip(batch, alone):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(mass[:, 17], m1[:, 0])
    np.testing.assert_array_equal(beta[:, 17], b1[:, 0])


def test_mfma_modes_refuse_wide_rows(gd):
    with pytest.raises(gd.DycoreError, match="1440") as e:
        gd.Dycore(2048, 65, 30.0, precision="mixed_mfma",
                  zonal_tend_filter_cutoff_wavenumber=gd.cutoff_from_lat0(2048, 65, 60.0))
    assert e.value.code == gd.GD_ERR_NOT_SUPPORTED
