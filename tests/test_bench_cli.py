"""bench.py's host-side behaviour without a GPU: one GPU only, and the HBM
traffic figure taken only from a profile line measured at the run's own
tuning."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and r.stdout == ""
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr


def test_traffic_only_from_a_profile_at_the_same_tuning(tmp_path, monkeypatch):
    b = _bench()
    prof = tmp_path / "profiles"
    prof.mkdir()
    (prof / "k.md").write_text("text\ntraffic_bytes_per_launch 16,32,2048,4096,8,1048576: 633207194\n"
                               "earlier build, traffic per launch 16,32,1024,4096,0,1048576: 646043750\n")
    monkeypatch.setattr(b, "ROOT", str(tmp_path))
    tun = dict(stack_depth=16, refill_min=32, grid_blocks=2048, shade_blocks=4096, waves_per_simd=8,
               tail_paths=1048576, claim_rays=128)
    assert b.profiled_traffic("k.md", tun) == (633207194, "16,32,2048,4096,8,1048576")
    assert b.profiled_traffic("k.md", dict(tun, waves_per_simd=4)) == (None, None)        # no line at this tuning
    assert b.profiled_traffic("k.md", dict(tun, grid_blocks=1024, waves_per_simd=0)) == (None, None)   # renamed
    assert b.profiled_traffic("k.md", dict(tun, claim_rays=256)) == (None, None)          # profiles use 128
    assert b.profiled_traffic("missing.md", tun) == (None, None)
    assert b.profiled_tunings("k.md") == [dict(stack_depth=16, refill_min=32, grid_blocks=2048, shade_blocks=4096,
                                               waves_per_simd=8, tail_paths=1048576)]
    assert b.profiled_tunings("missing.md") == []


def test_parse_tuning():
    b = _bench()
    assert b.parse_tuning("16,32,2048,4096,4,1048576", "--tuning") == dict(
        stack_depth=16, refill_min=32, grid_blocks=2048, shade_blocks=4096, waves_per_simd=4, tail_paths=1048576)
