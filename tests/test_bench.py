"""CPU tests of bench.py's contract and its measurement helpers: the one-GPU
refusal, the HBM roofline record, the VALU issue fraction read from the
committed profile, and the parity block's classification."""
import importlib.util
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU only" in p.stderr and p.stdout == ""


def test_roofline_record(bench):
    n, scene, t_ms = 1_000_000, 5_000_000, 2.0
    r = bench.roofline("k", n, t_ms, scene, cache_b=700 * n, per_ray={"nodes": 7.0}, traffic=120_000_000)
    compulsory = 48 * n + scene
    assert r["algorithmic_bytes_per_launch"] == compulsory
    assert r["achieved"] == pytest.approx(compulsory / 2e-3 / 1e9, abs=0.05)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0, abs=1e-4)
    assert r["bound"] == "hbm" and r["peak"] == 8000.0 and r["unit"] == "GB/s"
    assert r["traffic"] == 120_000_000
    assert r["hbm_measured"]["gbs"] == pytest.approx(120e6 / 2e-3 / 1e9, abs=0.05)
    assert r["hbm_measured"]["traffic_over_algorithmic"] == pytest.approx(120e6 / compulsory, abs=1e-3)
    assert r["cache_served"]["gbs"] == pytest.approx(700 * n / 2e-3 / 1e9, abs=0.05)
    assert "traffic" in r and bench.roofline("k", n, t_ms, scene)["traffic"] is None


def test_valu_issue_uses_the_committed_ceiling(bench):
    rec = json.loads((ROOT / "profiles" / "valu_issue.json").read_text())
    n = rec["rays"]
    k = rec["kernels"]["cwbvh_closest"]
    v = bench.valu_issue(rec["workload"], "cwbvh_closest", n, 2.0)
    assert v["valu_insts_per_launch"] == k["valu_insts_per_launch"]
    assert v["frac"] == pytest.approx(k["valu_insts_per_launch"] / 2e-3 / rec["ceiling_insts_per_s"], abs=1e-4)
    # the ceiling is the fastest calibrated stream, and the traversal stays below it
    assert rec["ceiling_insts_per_s"] == pytest.approx(max(p["insts_per_s"] for p in rec["probes"].values()), rel=1e-9)
    assert 0 < k["frac_of_ceiling_at_traced_ms"] <= 1.0
    # another workload or ray count has no record
    assert bench.valu_issue(rec["workload"], "cwbvh_closest", n + 1, 2.0) is None
    assert bench.valu_issue("C9", "cwbvh_closest", n, 2.0) is None


def test_parity_block_flags_a_planted_mismatch(bench, orc):
    from cwbvh_amd import workloads as W

    import cwbvh_amd as cw

    wl = W.c0(num_points=2048, rays_per_point=4)
    _, lay = cw.build_layouts(wl.mesh)
    ref = orc.trace_cwbvh(lay.nodes, lay.node_stride, lay.woop, wl.rays)
    ok = bench.parity_block(orc, ref.copy(), ref, wl.rays, lay.woop, False, "same")
    assert ok["ok"] and ok["bitwise_mismatches"] == 0 and ok["rays"] == len(ref)
    # plant a wrong triangle on a ray that hits something: a triangle that the
    # oracle's Woop test rejects for that ray
    hit = np.nonzero(orc.hit_ids(ref) >= 0)[0]
    i = int(hit[0])
    wrong = next(a for a in range(0, len(lay.woop) - 2, 3) if not orc.woop_test(lay.woop[a:a + 3], wl.rays[i])[0])
    bad = ref.copy()
    bad[i, 1] = np.int32(wrong).view(np.float32)
    out = bench.parity_block(orc, bad, ref, wl.rays, lay.woop, False, "planted")
    assert not out["ok"] and out["bitwise_mismatches"] == 1 and out["classified_sample"]["mismatches"] == 1
    # a distance one ulp off on the right triangle is within the 1e-5 tolerance:
    # counted as a bitwise difference, accepted by the contract
    near = ref.copy()
    near[i, 0] = np.nextafter(near[i, 0], np.float32(np.inf))
    out = bench.parity_block(orc, near, ref, wl.rays, lay.woop, False, "one ulp")
    assert out["ok"] and out["bitwise_mismatches"] == 1
    # any hit: a ray reported as a miss is an occlusion mismatch
    anyref = orc.trace_cwbvh(lay.nodes, lay.node_stride, lay.woop, wl.rays, any_hit=True)
    miss = anyref.copy()
    miss[i, 0] = wl.rays[i, 7]
    miss[i, 1] = np.int32(-1).view(np.float32)
    out = bench.parity_block(orc, miss, anyref, wl.rays, lay.woop, True, "planted any")
    assert out["occlusion_mismatches"] == 1 and not out["ok"]
