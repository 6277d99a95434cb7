"""bench.py reads the per-kernel HBM bytes of the committed traffic profile
and keeps, for the tunable kernels, only the launch-shape variant the run
uses (host logic only)."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _profile(tmp_path, line):
    p = tmp_path / "traffic.md"
    p.write_text("# summary\n\n" + line + "\n")
    return str(p)


def test_variants_are_matched_per_kernel(bench, tmp_path, monkeypatch):
    line = ('traffic_bytes_per_launch: {"tuning": {"fft_cols": 0, "green": 3, "bodyforce": 5}, "bytes": '
            '{"green_z": {"3": 816.0e6}, "fft_y": {"0": 805.3e6}, "bodyforce": {"4": 1.5e9, "5": 1.4e9}, '
            '"stress_corr": 4803.0e6}}')
    monkeypatch.setattr(bench, "TRAFFIC_PROFILE", _profile(tmp_path, line))
    out, why = bench.measured_traffic({"fft_cols": 0, "green": 3, "bodyforce": 4})
    assert why is None
    assert out == {"green_z": 816.0e6, "fft_y": 805.3e6, "bodyforce": 1.5e9, "stress_corr": 4803.0e6}
    out, why = bench.measured_traffic({"fft_cols": 1, "green": 3, "bodyforce": 2})
    assert "fft_y" not in out and "bodyforce" not in out and out["stress_corr"] == 4803.0e6
    assert "fft_y variant 1" in why and "bodyforce variant 2" in why


def test_plain_numbers_belong_to_the_profiled_tuning(bench, tmp_path, monkeypatch):
    line = ('traffic_bytes_per_launch: {"tuning": {"fft_cols": 0, "green": 3, "bodyforce": 3}, "bytes": '
            '{"green_z": 816.0e6, "bodyforce": 1779.6e6, "stress_corr": 4803.0e6}}')
    monkeypatch.setattr(bench, "TRAFFIC_PROFILE", _profile(tmp_path, line))
    out, why = bench.measured_traffic({"fft_cols": 0, "green": 3, "bodyforce": 3})
    assert why is None and out["bodyforce"] == 1779.6e6
    out, why = bench.measured_traffic({"fft_cols": 0, "green": 3, "bodyforce": 5})
    assert "bodyforce" not in out and "bodyforce variant 5" in why


def test_the_committed_profile_parses(bench):
    out, why = bench.measured_traffic({"fft_cols": 0, "green": 3, "bodyforce": 5})
    assert out is not None and out.get("stress_corr", 0) > 0, why
