"""The compiled drop-in: tests/dropin/main.cpp is a C++ caller written like the
reference's driver (src/main.cpp:88-245).  It includes
include/TraversalKernelBVH2.h and include/TraversalKernelCWBVH.h with its own
float4 / Ray / Hit and links libcwbvh_amd.so.  The CPU test compiles and links
it and checks that the four entry points bind by their C++ (mangled) names,
as a caller built against src/TraversalKernelCWBVH.h:3-11 and
src/TraversalKernelBVH2.h:3-11 would; the GPU test runs it and compares both
kernels' hits with the oracle on the layouts and rays it wrote."""
import importlib.util
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "tests" / "dropin" / "dropin_main"

# the reference signatures, mangled for HIP's float4 and global ::Ray / ::Hit
MANGLED = {
    "_Z15rtBindCWBVHDataPK15HIP_vector_typeIfLj4EES2_PKi",
    "_Z12rtTraceCWBVHP3RayP3Hiti",
    "_Z14rtBindBVH2DataPK15HIP_vector_typeIfLj4EES2_PKi",
    "_Z11rtTraceBVH2P3RayP3Hiti",
}


def _builder():
    spec = importlib.util.spec_from_file_location("cwbvh_amd_build_t", ROOT / "cwbvh_amd" / "build.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _symbols(path, *flags):
    out = subprocess.run(["nm", *flags, str(path)], capture_output=True, text=True, check=True).stdout
    return out.splitlines()


def test_dropin_compiles_and_links_by_cxx_names(cw, tmp_path):
    exe = _builder().build_dropin(tmp_path / "dropin_main")
    undefined = {l.split()[-1] for l in _symbols(exe) if l.strip().startswith("U ")}
    assert MANGLED <= undefined, sorted(MANGLED - undefined)
    # no C-linkage rt* reference: the C++ declarations were the ones used
    assert not {"rtTraceCWBVH", "rtTraceBVH2", "rtBindCWBVHData", "rtBindBVH2Data"} & undefined
    exported = {l.split()[-1] for l in _symbols(cw.LIB_PATH, "-D", "--defined-only") if " T " in l}
    assert MANGLED <= exported
    ldd = subprocess.run(["ldd", str(exe)], capture_output=True, text=True, check=True).stdout
    assert "libcwbvh_amd.so" in ldd and "not found" not in ldd.split("libcwbvh_amd.so", 1)[1].splitlines()[0]


@pytest.mark.gpu
def test_dropin_driver_matches_oracle(orc, tmp_path):
    assert BIN.exists(), "tests/dropin/dropin_main is built by __graft_entry__.build()"
    print(f"[gpu-dropin] running {BIN.name} on the C0 heightfield, 131072 x 8 rays", flush=True)
    sys.stdout.flush()
    proc = subprocess.run([str(BIN), str(tmp_path), "131072", "8"], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "rtTraceBVH2 No Dynamic Fetch" in proc.stdout and "rtTraceCWBVH Dynamic Fetch" in proc.stdout
    meta = dict(l.split() for l in (tmp_path / "meta.txt").read_text().splitlines())
    n = int(meta["rays"])

    def arr(name, cols):
        return np.fromfile(tmp_path / name, np.float32).reshape(-1, cols)

    rays = arr("rays.f32", 8)
    assert len(rays) == n == 1048576
    for kind, ref in (("bvh2", orc.trace_bvh2(arr("bvh2_nodes.f32", 4), arr("bvh2_woop.f32", 4), rays)),
                      ("cwbvh", orc.trace_cwbvh(arr("cwbvh_nodes.f32", 4), int(meta["cwbvh_node_stride"]),
                                                arr("cwbvh_woop.f32", 4), rays))):
        got = arr(f"hits_{kind}.f32", 4)
        diff = (got.view(np.uint32) != ref.view(np.uint32)).any(axis=1)
        assert not diff.any(), f"{kind}: {int(diff.sum())} of {n} rays differ from the oracle"
        assert 0.05 < (orc.hit_ids(ref) >= 0).mean() < 0.95
