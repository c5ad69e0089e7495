"""C-ABI library: loads without a GPU, exports every symbol the headers declare, host logic."""
import os
import re
import subprocess

import pytest

from tests.conftest import ROOT


def header_symbols():
    syms = set()
    for h in ("hip_marshal.h", "marshal.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^(?:const\s+)?\w+\s*\*?\s*((?:hip|gpu)_\w+)\s*\(", text,
                             flags=re.M):
            syms.add(m.group(1))
    return syms


def test_library_loads_and_exports_all_declared_symbols():
    import libmarshal_amd as lm
    L = lm.lib()
    declared = header_symbols()
    assert len(declared) >= 30
    for s in declared:
        assert hasattr(L, s), s
    assert set(lm.EXPORTED) == declared
    nm = subprocess.run(["nm", "-D", "--defined-only", lm.LIB_PATH], capture_output=True,
                        text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in nm.splitlines() if " T " in ln}
    assert declared <= exported
    assert b"gfx950" in L.hip_marshal_version()


def test_library_contains_gfx950_code_object():
    import libmarshal_amd as lm
    data = open(lm.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_fastdiv_host_check(tmp_path):
    src = os.path.join(ROOT, "tests", "host", "fastdiv_check.cpp")
    exe = str(tmp_path / "fdc")
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-o", exe, src], check=True,
                   capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert "fails=0" in r.stdout


def test_flag_bytes_host_logic():
    import libmarshal_amd as lm
    L = lm.lib()
    # 100(16, 2^18, 64) f32: one bit per 256-byte super-element
    assert L.hip_marshal_flag_bytes_0100(1, 16, 1 << 18, 64, 4) == 16 * (1 << 18) // 8
    # element-granular fallback on 250007 x 199999 bytes: one bit per element
    n = 250007 * 199999
    assert L.hip_marshal_flag_bytes_0100(1, 250007, 199999, 1, 1) == (n + 31) // 32 * 4
    # invalid input gives 0
    assert L.hip_marshal_flag_bytes_0100(1, 0, 5, 1, 4) == 0
    assert L.hip_marshal_flag_bytes_0100(1, 2, 5, 1, 3) == 0


def test_invalid_arguments_rejected_without_gpu_work():
    """Argument validation happens before any device call (error convention: True)."""
    import ctypes
    import libmarshal_amd as lm
    L = lm.lib()
    nul = ctypes.c_void_p(0)
    assert L.hip_transpose_es(nul, nul, 4, 2, 3, 4, 5, 1, 3, None) is True
    assert L.hip_transpose_es(nul, ctypes.c_void_p(256), 3, 2, 3, 4, 5, 1, 3, None) is True
    assert L.hip_transpose_es(nul, ctypes.c_void_p(256), 4, 0, 3, 4, 5, 1, 3, None) is True
    assert L.hip_aos_asta_es(nul, ctypes.c_void_p(256), 4, 10, 3, 4, 0, None) is True  # 10 % 4
    assert L.hip_padding_es(nul, ctypes.c_void_p(256), 4, 9, 3, 8, None) is True       # x > pad
    assert L.hip_transpose_es(nul, ctypes.c_void_p(258), 4, 2, 3, 4, 5, 1, 3, None) is True


def test_lds_budget_knob():
    import libmarshal_amd as lm
    default = lm.lds_budget()
    assert default >= 64 * 1024
    lm.set_lds_budget(4096)
    assert lm.lds_budget() == 4096
    lm.set_lds_budget(0)
    assert lm.lds_budget() == default


def test_planner_route_choice_host_only():
    """The cost-model planner runs on the host: square matrices take the tile-pair swaps,
    non-square ones within the scratch cap the through-scratch route, ones beyond it the
    stripe route; a scratch cap of 0 leaves only in-place routes."""
    import libmarshal_amd as lm
    keep = lm.get_tuning("st_scratch_mb")
    try:
        assert lm.plan_transpose(4096, 4096, 8)[2] == 8
        assert lm.plan_transpose(3001, 7001, 4)[2] == 9
        assert lm.plan_transpose(98304, 131071, 8)[2] in (6, 7)
        assert lm.plan_transpose(250007, 199999, 1)[2] in (6, 7)
        # 2.4 GB: within the default 8 GiB cap through scratch, beyond a 2 GiB cap in stripes
        assert lm.plan_transpose(20011, 15013, 8)[2] == 9
        lm.set_tuning("st_scratch_mb", 2048)
        assert lm.plan_transpose(20011, 15013, 8)[2] in (6, 7)
        lm.set_tuning("st_scratch_mb", 0)
        for (M, N, es) in ((3001, 7001, 4), (1031, 997, 1), (1800, 7200, 4)):
            assert lm.plan_transpose(M, N, es)[2] not in (6, 7, 9), (M, N, es)
    finally:
        lm.set_tuning("st_scratch_mb", keep)
    assert lm.plan_transpose(3001, 7001, 4)[2] == 9  # knob changes re-plan


def test_tuning_keys_defaults_and_limits():
    import libmarshal_amd as lm
    defaults = {"st_scratch_mb": 8192, "batch_solo_kb": 64, "batch_solo_lanes": 4, "st_skew": 1, "copy_nt": 1,
                "oop_nt": 2, "pad_nt": 1, "gather_tile_kb": 24, "gather_row_bytes": 256,
                "bs_upt": 16, "bs_blocks_per_cu": 192, "oop_elem_tile_kb": 8,
                "copy_blocks_per_cu": 64, "copy_nt_store_mb": 256, "oop_pitched": 1, "oop_min_rows": 256, "oop_min_tiles": 512, "st_super_bytes": 8192,
                "st_max_height": 4096,
                "autotune": 0, "oop_order": 2, "cl_max_wgs": 16, "cl_min_bytes": 256 << 10,
                "cl_min_inst": 64, "cl_fuse": 1, "cl_spin": 1 << 22,
                "batch_arena_mb": 1024, "batch_arena_max_mb": 16,
                "batch_level_lanes": 4, "pt010_lead_cached": 2}
    for k, v in defaults.items():
        assert lm.get_tuning(k) == v, k
    assert lm.get_tuning("no_such_key") == 0
    with pytest.raises(ValueError):
        lm.set_tuning("no_such_key", 1)
    with pytest.raises(ValueError):
        lm.set_tuning("bs_upt", 0)  # below its minimum of 1
    lm.set_tuning("st_skew", 0)  # 0 allowed
    assert lm.get_tuning("st_skew") == 0
    lm.set_tuning("st_skew", 1)
    lm.set_tuning("oop_elem_tile_kb", 0)  # 0 allowed: element-store tiles use st_tile_kb
    assert lm.get_tuning("oop_elem_tile_kb") == 0
    lm.set_tuning("oop_elem_tile_kb", 8)
    with pytest.raises(ValueError):
        lm.set_tuning("copy_blocks_per_cu", 0)
    for k in ("cl_max_wgs", "cl_fuse", "oop_order", "batch_arena_mb"):  # 0 switches these off
        keep = lm.get_tuning(k)
        lm.set_tuning(k, 0)
        assert lm.get_tuning(k) == 0
        lm.set_tuning(k, keep)
    with pytest.raises(ValueError):
        lm.set_tuning("cl_spin", 0)  # the barrier wait needs at least one poll


def test_planner_survives_random_shapes_host_only():
    """Planning (cost model over every candidate route, including the cluster-staged and fused
    forms) for random shapes and element sizes returns a valid stage code and never crashes."""
    import random
    import libmarshal_amd as lm
    rng = random.Random(11)
    valid = {2, 22, 3, 32, 4, -1, 5, 6, 7, 8, 9}
    for _ in range(300):
        M = int(2 ** rng.uniform(1, 17))
        N = int(2 ** rng.uniform(1, 17))
        es = rng.choice((1, 2, 4, 8, 16))
        a, b, stages = lm.plan_transpose(M, N, es)
        assert stages in valid, (M, N, es, stages)
        assert a >= 1 and b >= 1 and M % a == 0 and N % b == 0, (M, N, es, a, b, stages)


def test_cycle_count_matches_the_oracle():
    """The host's gcd / Moebius count of the transpose map's cycles equals the oracle's walk
    (include/plan.hpp:57-73 restated) on every small grid, and gives the 16 x 2^18 grid of the 2^24 x 16 SoA conversion."""
    import libmarshal_amd as lm
    import oracle
    for R in range(2, 34):
        for C in range(2, 34):
            n, k = lm.cycle_count(R, C, max_order=1 << 20)
            assert n == oracle.num_cycles(R, C), (R, C)
            assert k >= 1
            # every cycle length divides the order: R^k = 1 (mod RC - 1)
            assert pow(R, k, R * C - 1) == 1 % (R * C - 1), (R, C, k)
    assert lm.cycle_count(16, 1 << 18) == (381300, 11)   # 100(16, 2^18, 64): 2^24 structs
    assert lm.cycle_count(1 << 18, 16) == (381300, 11)   # and its inverse
    assert lm.cycle_count(1800, 7200, max_order=64) == (0, 0)  # order far above 64


def test_c_caller_host_checks(tmp_path):
    """A C++ program built against include/*.h and linked to the library (no Python in
    between) calls the entry points that need no device: version, planner, knobs, argument
    validation."""
    from tests.capi import build as capi
    exe = capi.build(str(tmp_path / "capi_check"))
    env = dict(os.environ)  # built outside the tree: its relative rpath does not reach the .so
    env["LD_LIBRARY_PATH"] = os.path.join(ROOT, "libmarshal_amd") + ":" + env.get(
        "LD_LIBRARY_PATH", "")
    r = subprocess.run([exe, "host"], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "capi host ok" in r.stdout


def test_leader_ordering_covers_every_cycle_once(tmp_path):
    """The host-built leader lists of the cycle-following 010 (one start per cycle, ordered
    for few sectors per wave instruction, padded to groups of 64): over 1610 tile shapes and
    element sizes, the walks from the listed starts cover every moved slot exactly once, and
    only tiles with a cycle longer than 64 slots are left unordered."""
    src = os.path.join(ROOT, "tests", "capi", "order_check.cpp")
    exe = str(tmp_path / "order_check")
    subprocess.run(["g++", "-O2", "-std=c++17", src, "-L" + os.path.join(ROOT, "libmarshal_amd"),
                    "-lmarshal_hip", "-L/opt/rocm/lib", "-lamdhip64", "-o", exe], check=True,
                   capture_output=True)
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = os.path.join(ROOT, "libmarshal_amd") + ":/opt/rocm/lib:" + env.get(
        "LD_LIBRARY_PATH", "")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "fails=0" in r.stdout


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asked for more it exits with a message (before touching a
    device)."""
    import json
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
