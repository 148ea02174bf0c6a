"""The input side against the reference's own builders (golden segment tables and maps, tests/golden/), CPU."""

from pathlib import Path

import numpy as np
import pytest
import torch

import blochparty_amd as bp

GOLD = Path(__file__).resolve().parent / "golden"


def builders():
    sys10 = bp.System(rf_raster=10e-6)
    sys20 = bp.System(rf_raster=20e-6)
    return {
        "gre2d_16_ideal": lambda: bp.gre_2d(16, n_dummy=3, te=5e-3),
        "gre2d_12_rfspoil_ramps": lambda: bp.gre_2d(12, n_dummy=4, spoiling="rf", ramps=True),
        "gre2d_8_gradspoil_slice": lambda: bp.gre_2d(8, spoiling="gradient", slice_thickness=5e-3, n_dummy=1),
        "gre3d_6_sinc_slab": lambda: bp.gre_cartesian((6, 6, 6), flip_deg=10.0, tr=8e-3, slice_thickness=0.22,
                                                      rf_duration=2e-3, system=sys10),
        "tse_32_etl4_hard": lambda: bp.tse_2d(32, etl=4, tr=0.5),
        "tse_16_etl4_sinc": lambda: bp.tse_2d(16, etl=4, tr=0.5, slice_thickness=5e-3, rf_duration=2e-3,
                                              system=sys20),
        "tse_16_centric_train": lambda: bp.tse_2d(16, etl=4, tr=0.4, ordering="centric",
                                                  refocus_flip_deg=[150.0, 160.0, 170.0, 180.0]),
        "tse_16_reverse_gradspoil": lambda: bp.tse_2d(16, etl=8, tr=0.4, ordering="reverse_centric",
                                                      spoiling="gradient"),
    }


@pytest.mark.parametrize("name", list(builders()))
def test_segment_tables_match_reference_builders(name):
    z = np.load(GOLD / "sequences.npz")
    seq = builders()[name]()
    flags = (np.asarray(seq.is_rf, np.int32) | (np.asarray(seq.adc, np.int32) << 1)
             | (np.asarray(seq.spoil, np.int32) << 2))
    assert np.array_equal(flags, z[f"{name}/flags"])
    for key, mine in (("dt", seq.dt), ("rf", seq.rf), ("grad", seq.grad), ("adc_phase", seq.adc_phase)):
        ref = z[f"{name}/{key}"]
        np.testing.assert_allclose(mine.detach().numpy(), ref, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(ref).max()))


def test_shepp_logan_maps_and_positions_match_reference():
    z = np.load(GOLD / "phantoms.npz")
    for n, nd in ((32, 2), (12, 3)):
        m = bp.shepp_logan_maps(n, ndim=nd)
        for key in ("label", "pd", "t1", "t2"):
            assert np.array_equal(m[key].numpy(), z[f"sl{nd}d_{n}/{key}"]), (n, nd, key)
    m = bp.shepp_logan_maps(8, ndim=3)
    sub = bp.phantom_from_maps(m["pd"], m["t1"], m["t2"], 0.22, spins_per_voxel=(1, 2, 2))
    jit = bp.phantom_from_maps(m["pd"], m["t1"], m["t2"], 0.22, spins_per_voxel=(2, 2, 2), jitter="stratified", seed=5)
    np.testing.assert_allclose(sub.pos.numpy(), z["sub122/pos"], atol=1e-15)
    np.testing.assert_allclose(sub.pd.numpy(), z["sub122/pd"], atol=1e-15)
    np.testing.assert_allclose(jit.pos.numpy(), z["jit222/pos"], atol=1e-15)
    c = bp.birdcage_sensitivities(sub.pos, 8)
    np.testing.assert_allclose(c.real.numpy(), z["coils8/re"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(c.imag.numpy(), z["coils8/im"], rtol=1e-13, atol=1e-13)
    c = bp.birdcage_sensitivities(sub.pos, 5, radius=0.15, radius_y=0.2, element_radius=0.04, normalize=True)
    np.testing.assert_allclose(c.real.numpy(), z["coils5e/re"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(c.imag.numpy(), z["coils5e/im"], rtol=1e-13, atol=1e-13)


def test_flip_angle_carries_gradient_through_builders():
    flip = torch.tensor(12.0, dtype=torch.float64, requires_grad=True)
    seq = bp.gre_cartesian((4, 4, 4), flip_deg=flip, tr=8e-3, slice_thickness=0.22, rf_duration=2e-3,
                           system=bp.System(rf_raster=10e-6))
    (g,) = torch.autograd.grad(seq.rf.sum(), flip)
    assert torch.isfinite(g) and float(g) != 0.0


def test_workload_sizes():
    from blochparty_amd import workloads as W

    ph, seq, coils = W.gre2d()
    assert (len(ph), len(seq), seq.n_adc, coils.shape) == (2039, 4752, 4096, (2039, 8))
    seq3 = W.gre3d_sequence(8)
    assert sum(seq3.is_rf) == 200 * 64 and seq3.n_adc == 8**3
