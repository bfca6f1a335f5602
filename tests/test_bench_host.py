"""Host-side pieces of bench.py (no GPU): the refusal of more than one GPU,
and its per-launch byte model of the march kernels against the hardware
counters committed under profiles/ (rocprofv3 FETCH_SIZE + WRITE_SIZE,
calibrated on a plain device copy), which the bench line's roofline reads."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("args,env", [(["--gpus", "2"], {}), ([], {"WORLD_SIZE": "8"})])
def test_more_than_one_gpu_is_refused(args, env):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, capture_output=True, text=True,
                       env=dict(os.environ, **env), timeout=120)
    assert r.returncode == 2
    assert "is not supported" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("precision", [64, 32])
def test_byte_model_of_the_march_kernels_matches_the_hbm_counters(precision):
    """512^2 x 100, two propagations of one wavelength, N_w = 2: every march
    kernel moves what the byte model says, give or take the few re-fetches
    that the counters see (at most 15 %); a dropped or doubled plane in the
    model would fall outside."""
    c = 16 if precision == 64 else 8
    path = os.path.join(ROOT, "profiles", "hbm_traffic_random512_fp%d.txt" % precision)
    kernels = {"solve_x": "k_xsolve", "solve_y": "k_ysolve", "boundary": "k_boundary_point", "coefs": "k_slab_coefs"}
    for fam, kernel in kernels.items():
        measured, src = bench.measured_traffic(kernel, path)
        assert measured is not None and src.startswith("profiles"), kernel
        model = bench.algorithmic_bytes(fam, c, 2, 1, 511 * 511, 512 * 512, 3, 2)
        assert 0.98 <= measured / model <= 1.15, (kernel, measured / model)


def test_missing_traffic_summary_gives_none():
    assert bench.measured_traffic("k_ysolve", os.path.join(ROOT, "profiles", "no_such_file.txt")) == (None, None)
