"""The bandwidth probes behind the bench's roofline figures.

copy_probe (one array in, one out) calibrates the HBM byte counters;
stream_probe copies in the sweeps' access shape (four arrays in, four out)
and gives the rate that pattern reaches with no arithmetic.  Both must return
a positive rate below the MI355X's published 8 TB/s, and reject bad sizes
before touching the GPU.
"""
import pytest

import hydro_amd

PEAK_GBS = 8000.0


def test_stream_probe_rejects_bad_sizes_before_any_gpu_call():
    # argument checks come first, so this runs without a GPU too
    with pytest.raises(hydro_amd.HydroError, match="stream probe"):
        hydro_amd.stream_probe(0, 1)
    with pytest.raises(hydro_amd.HydroError, match="stream probe"):
        hydro_amd.stream_probe(1000, 0)
    with pytest.raises(hydro_amd.HydroError, match="copy probe"):
        hydro_amd.copy_probe(4, 1)


@pytest.mark.gpu
def test_probe_rates_are_physical():
    for cells in (1, 1000, 8196 * 8196):
        r = hydro_amd.stream_probe(cells, 3)
        print(f"stream probe, {cells} cells: {r:.0f} GB/s")
        assert 0 < r < PEAK_GBS * 1.02
    r = hydro_amd.copy_probe(1 << 28, 3)
    print(f"copy probe, 256 MiB: {r:.0f} GB/s")
    assert 0 < r < PEAK_GBS * 1.02
