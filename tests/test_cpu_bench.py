"""The benchmark's host arithmetic: the per-stage algorithmic bytes add up to
the closed form 4292 N_P + 1216 N_L + 2544 N_M bytes per step at
C/R/D = 16/32/8 with 3 radix passes (the sum of the per-stage budgets of
dynamics.cuh, metabolites.cuh and interactions.cuh kernels: key 24, sort
4 + 16 d, permute 180 / 228, ranges 12, move 48, coordinate 8 C, applyNoise
84, relax 108 R, metabolic relax 44 + 36 per lipid and back-reaction 8 + 40
per lipid + 32 per metabolic on the R/2 even passes, applyVelocities 112,
complexify 52, transition 16, diffusion 96 D), the radix pass count of the
reference's power-of-two grid (config.cuh:66-74), and the refusal of more
than one GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


class _P:
    def __init__(self, L, cell=10.0, C=16, R=32, D=8):
        self.sim_size, self.grid_cell_size = L, cell
        self.noise_coordination_steps, self.relaxation_steps, self.metabolite_diffusion_steps = C, R, D


@pytest.mark.parametrize("n_p,n_l,n_m", [(471858, 209715, 576718), (1, 0, 0), (0, 0, 1), (5, 5, 0)])
def test_stage_bytes_closed_form(n_p, n_l, n_m):
    P = _P(1000.0)
    assert bench.radix_passes(P) == 3
    total = sum(bench.stage_bytes(n_p, n_l, n_m, 3, P).values())
    assert total == 4292 * n_p + 1216 * n_l + 2544 * n_m


@pytest.mark.parametrize("L,cell,passes", [(400.0, 10.0, 3), (1000.0, 10.0, 3), (2000.0, 10.0, 4),
                                           (10.0, 10.0, 1), (5000.0, 10.0, 4), (40.0, 10.0, 1)])
def test_radix_passes(L, cell, passes):
    # bits = log2 of the power-of-two cells per axis; 3 bits per key axis plus the sentinel bit, 8 per pass
    assert bench.radix_passes(_P(L, cell)) == passes


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout
