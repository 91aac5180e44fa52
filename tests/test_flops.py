"""Per-unit work figures that bench.py's roofline uses.

* The reference formulation: the op-counting FP64 oracle on the benchmark workload, per (receiver, source) pair and
  per (receiver, triangle side) of the free-surface correction (AngSetupStrainFSC).
* The engine's forward-mode formulation of the dominant kernel's unit (receiver, mesh edge): the device header
  compiled for the host with a counting scalar (tests/flops/edge_count.cpp).  The count of the three-component
  evaluation with each shared quantity formed once is the unit's algorithmic work; the kernel evaluates the shared
  part once per component (three times), which the identity below ties to the product's own functions.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import bench
from cutanddisplacejulia_amd import api, meshes
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROCM_INCLUDE = "/opt/rocm/include"


def test_flops_per_pair_half_space():
    c = meshes.config2()
    P1, P2, P3 = c["P1"], c["P2"], c["P3"]
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    rng = np.random.default_rng(0)
    ri = rng.integers(0, n, 4000)
    si = rng.integers(0, n, 4000)
    cnt = O.count_traction_flops(nrm, mid, P1, P2, P3, 0.25, 1.0, 0.4, 1, ri, si)
    per_pair = (cnt["add"] + cnt["mul"] + cnt["div"]) / cnt["pairs"]
    assert abs(per_pair - bench.FLOPS_PER_PAIR_HS) / bench.FLOPS_PER_PAIR_HS < 0.01


def test_reference_flops_per_side_half_space():
    """One triangle side's free-surface strain (three slip modes) at one receiver, in the reference formulation."""
    c = meshes.config2()
    P1, P2, P3 = c["P1"], c["P2"], c["P3"]
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    rng = np.random.default_rng(0)
    ri = rng.integers(0, n, 6000)
    si = rng.integers(0, 3 * n, 6000)
    cnt = O.count_fsc_side_flops(nrm, mid, P1, P2, P3, 0.25, ri, si)
    per_unit = (cnt["add"] + cnt["mul"] + cnt["div"]) / cnt["pairs"]
    assert abs(per_unit - bench.REF_FLOPS_PER_EDGE_HS) / bench.REF_FLOPS_PER_EDGE_HS < 0.01
    # the three sides are most of the pair's work (AngSetupStrainFSC dominates the half-space strain)
    assert 0.8 < 3 * per_unit / bench.FLOPS_PER_PAIR_HS < 0.9


@pytest.fixture(scope="module")
def edge_counts(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None or not os.path.isdir(os.path.join(ROCM_INCLUDE, "hip")):
        pytest.skip("g++ or the HIP headers are not available")
    exe = str(tmp_path_factory.mktemp("flops") / "edge_count")
    subprocess.run([gxx, "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I" + ROCM_INCLUDE, "-w",
                    os.path.join(HERE, "flops", "edge_count.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def _flops(c):
    return c["add"] + c["mul"] + c["div"]


def test_edge_unit_count_matches_the_kernels_functions(edge_counts):
    """fsc_side_wk<0> + <1> + <2> = the single pass + twice the shared part, operation by operation: the counter's
    single pass evaluates exactly the product's statements, with the shared part once."""
    c = edge_counts
    for op in ("add", "mul", "div", "sqrt", "log", "atan"):
        assert c["wk0"][op] + c["wk1"][op] + c["wk2"][op] == c["single"][op] + 2 * c["shared"][op], op
    # every component evaluates three outputs: none is empty
    assert min(_flops(c["wk%d" % k]) - _flops(c["shared"]) for k in range(3)) > 500


def test_edge_unit_live_counts(edge_counts):
    """Only operations that reach the outputs count: the displacement values' logarithms and arctangents are dead
    (the unit needs gradients), the square roots and the traction part are all live, and liveness never adds."""
    c = edge_counts
    for name in ("wk0", "wk1", "wk2", "single", "traction"):
        live = c["live_" + name]
        assert live["log"] == 0 and live["atan"] == 0, name
        assert live["sqrt"] == c[name]["sqrt"], name
        for op in ("add", "mul", "div"):
            assert live[op] <= c[name][op], (name, op)
    assert _flops(c["live_traction"]) >= _flops(c["traction"]) - 2


def test_edge_unit_flops_frozen_in_bench(edge_counts):
    c = edge_counts
    assert _flops(c["live_single"]) + _flops(c["live_traction"]) == bench.FLOPS_PER_EDGE_HS
    assert (sum(_flops(c["live_wk%d" % k]) for k in range(3)) + _flops(c["live_traction"])
            == bench.KERNEL_FLOPS_PER_EDGE_HS)
    # the forward-mode formulation is far cheaper than the reference's strain closed form for the same unit
    assert bench.FLOPS_PER_EDGE_HS < 0.4 * bench.REF_FLOPS_PER_EDGE_HS
