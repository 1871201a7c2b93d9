"""The Stokes-Einstein noise model (config noiseModel "stokes_einstein", an
extension of spatial_cell_sim/src/dynamics.cuh:156-215, which scales the
noise by dimensionless constants instead) pinned by the physics it states,
independently of the formula in the code: for a sphere of radius r in a
fluid of viscosity eta at temperature T, each displacement axis has variance
2 D dt with D = kT / (6 pi eta r), and the rotation angle has mean square
4 D_r dt with D_r = kT / (8 pi eta r^3) (small-angle rotational diffusion).
Measured on isolated particles from the oracle's own Philox draws; two
planted mistakes (the displacement scale without its factor 3, the rotation
scale as 1/r instead of 1/r^3) fail the check."""
import numpy as np
import pytest

import oracle as orc
import scenes
import spatial_cell as sc

KB = 1.380649e-23


def _cfg(n, L, **kw):
    c = scenes.config(n, n, L, boundary="periodic", cutoff=10.0, diff=10.0)
    c.update(noiseModel="stokes_einstein", temperature=310.0, viscosity=6.9e-4, timestep=0.1)
    c.update(kw)
    return c


def _lattice(n, spacing):
    side = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    return (g * spacing + spacing / 2).astype(np.float32), side * spacing


def diffusion(r, T=310.0, eta=6.9e-4):
    """D (nm^2/ns) and D_r (rad^2/ns) of a sphere of radius r nm."""
    kT = KB * T
    rm = r * 1e-9
    D = kT / (6 * np.pi * eta * rm) * 1e18 / 1e9
    Dr = kT / (8 * np.pi * eta * rm ** 3) / 1e9
    return D, Dr


def sample(radius, n=27000, plant=None):
    pos, L = _lattice(n, 20.0)
    rows = [{"pos": p, "radius": radius, "type": 1, "id": i} for i, p in enumerate(pos)]
    from test_cpu_oracle import make_particles
    cfg = _cfg(n, L)
    s = sc.Setup(cfg, {"types": scenes.TYPES})
    plain = make_particles(rows, n)
    met = make_particles([], n)
    o = orc.Oracle(s.params, s.tables, sc.Counts(n, 0, n, 0), plain, met, plant=plant)
    o.run_stage(sc.STAGE_MOVE)
    _, p, _ = o.snapshot()
    o.close()
    return p["pos_noise"][:n].astype(np.float64), p["ang_noise"][:n].astype(np.float64), s.params


def check_moments(plant=None):
    dt = 0.1
    for radius in (1.0, 2.5, 4.0):
        x, q, P = sample(radius, plant=plant)
        D, Dr = diffusion(radius)
        var = (x ** 2).mean(axis=0)
        assert np.all(np.abs(x.mean(axis=0)) < 0.05 * np.sqrt(2 * D * dt)), radius
        np.testing.assert_allclose(var, 2 * D * dt, rtol=0.05, err_msg=f"displacement r={radius}")
        theta = 2 * np.arccos(np.clip(np.abs(q[:, 3]), 0, 1))
        np.testing.assert_allclose((theta ** 2).mean(), 4 * Dr * dt, rtol=0.05, err_msg=f"rotation r={radius}")


def test_stokes_einstein_moments():
    check_moments()


def test_stokes_einstein_coefficients_at_310k():
    """The coefficients the config yields: D(2.5 nm) = 0.1317 nm^2/ns in water
    near body temperature (0.69 mPa s), a per-axis step of 0.162 nm at
    0.1 ns; the reference noise model leaves them zero."""
    s = sc.Setup(_cfg(8, 60.0), {"types": scenes.TYPES})
    assert s.params.noise_model == sc.NOISE_STOKES_EINSTEIN
    D, _ = diffusion(2.5)
    assert abs(D - 0.1317) < 5e-4
    assert abs(np.sqrt(s.params.se_translation / 2.5 / 3) - np.sqrt(2 * D * 0.1)) < 1e-6
    s0 = sc.Setup(scenes.config(8, 8, 60.0), {"types": scenes.TYPES})
    assert s0.params.noise_model == sc.NOISE_REFERENCE and s0.params.se_translation == 0.0


@pytest.mark.parametrize("plant", [41, 42])
def test_planted_noise_mistakes_fail(plant):
    with pytest.raises(AssertionError):
        check_moments(plant=plant)
