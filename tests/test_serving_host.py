"""Host-side parts of the serving loop: the seeded trace generator, the
latency percentiles, and a server that refuses to start without a device."""
import numpy as np
import pytest

from nemaktis_amd import serving


def test_poisson_trace_is_seeded_and_bounded():
    a = serving.poisson_trace(40, rate_hz=2.0, seed=3)
    b = serving.poisson_trace(40, rate_hz=2.0, seed=3)
    c = serving.poisson_trace(40, rate_hz=2.0, seed=4)
    assert a == b
    assert a != c
    t = [s["arrival_s"] for s in a]
    assert all(t1 > t0 for t0, t1 in zip(t, t[1:]))
    # mean inter-arrival close to 1/rate for 40 draws (loose bound)
    assert 0.25 < np.mean(np.diff([0.0] + t)) < 1.0
    for s in a:
        assert tuple(s["mesh"]) in [tuple(m) for m in serving.MESHES]
        assert s["generator"] in serving.GENERATORS
        assert 1 <= len(s["wavelengths"]) <= 21
        assert all(0.40 <= w <= 0.80 for w in s["wavelengths"])
        assert 1 <= len(s["wavevectors"]) <= 7
        assert s["wavevectors"][0] == [0.0, 0.0]  # on-axis first


def test_percentiles():
    p = serving.percentiles(list(range(1, 101)))
    assert p[50] == pytest.approx(50.5)
    assert p[99] == pytest.approx(99.01)


def test_job_settings_round_trip():
    s = serving.poisson_trace(1, rate_hz=1.0, seed=1)[0]
    lc, mask, formula, sp = serving.director((16, 12, 5), s["generator"])
    assert lc.shape == (5, 12, 16, 3)
    st = serving.job_settings(dict(s, mesh=(16, 12, 5)), sp)
    co = st["Physics settings"]["Coefficients"]
    assert co["Wavelengths"] == [float(w) for w in s["wavelengths"]]
    assert len(co["Wavevectors"]) == 2 * len(s["wavevectors"])


def test_server_refuses_to_start_without_device():
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(serving.ServingError, match="HIP|device"):
        serving.Server(max_active=2)
