"""The command-line drivers keep the reference's argv and output file
format; the File layer writes the reference's file set."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import pyoracle as O
from helpers import load_golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "lpa_amd", "bin")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def copy_case(case, tmp_path):
    dst = tmp_path / "edges.txt"
    shutil.copy(os.path.join(GOLDEN, case, "edges.txt"), dst)
    return str(dst)


@pytest.mark.parametrize("prog", ["nova", "novaplus", "cglp"])
def test_usage_on_wrong_argc(prog):
    r = subprocess.run([os.path.join(BIN, prog), "only_one_arg"], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr


@pytest.mark.parametrize("case", ["mixed400_t32", "rmat10_t16", "lattice16_t0"])
def test_file_layer_writes_reference_file_set(case, tmp_path):
    """File::load_graph on a cache miss writes the same files as the
    reference (the engine step may fail here without a GPU)"""
    g = load_golden(case)
    src = copy_case(case, tmp_path)
    t = g["params"]["small_thresh"]
    subprocess.run([os.path.join(BIN, "nova"), src, "1", str(g["params"]["one_way"]), str(t),
                    str(tmp_path / "out.txt"), "--iters", "1"], capture_output=True, timeout=120)
    ref = os.path.join(GOLDEN, case, "edges.txt")
    for suf in (".graph", ".offset", ".vertices.bin", ".row", ".clm", ".graph.small.bin", ".offset.small.bin"):
        if os.path.exists(ref + suf):
            assert open(src + suf, "rb").read() == open(ref + suf, "rb").read(), suf
    assert open(src + ".lpa_params").read().split() == [str(t), str(g["params"]["one_way"]), "1"]


@pytest.mark.gpu
@pytest.mark.parametrize("prog", ["nova", "novaplus", "cglp"])
def test_driver_labels_match_oracle(prog, tmp_path):
    case = "mixed400_t32"
    g = load_golden(case)
    src = copy_case(case, tmp_path)
    out = tmp_path / "out.txt"
    lab_path = tmp_path / "labels.bin"
    print(f"[gpu-case] driver {prog} on {case}", flush=True)
    r = subprocess.run([os.path.join(BIN, prog), src, "1", "0", "32", str(out), "--iters", "6",
                        "--labels-out", str(lab_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(lab_path, np.uint32)
    want, _ = O.lp_run_csr(g["row"], g["clm"], np.arange(g["n"], dtype=np.uint32), 6)
    assert np.array_equal(got, want)
    lines = open(out).read().splitlines()
    assert lines[0] == f"{prog} file: {src}" and lines[1].startswith(" running time: ") \
        and lines[1].endswith("thre: 32") and lines[2] == "~~~~~"


@pytest.mark.parametrize("case", ["mixed400_t32", "multiseg6000_t0"])
def test_reference_file_set_loads_without_edge_list(case, tmp_path):
    """a file set the reference wrote (no .lpa_params, no edge list) is
    loaded as the reference loads it (src/file.cpp:29-60), not rebuilt"""
    g = load_golden(case)
    ref = os.path.join(GOLDEN, case, "edges.txt")
    src = str(tmp_path / "edges.txt")
    for suf in (".graph", ".offset", ".vertices.bin", ".row", ".clm", ".graph.small.bin", ".offset.small.bin"):
        if os.path.exists(ref + suf):
            shutil.copy(ref + suf, src + suf)
    t = g["params"]["small_thresh"]
    r = subprocess.run([os.path.join(BIN, "nova"), src, "1", str(g["params"]["one_way"]), str(t),
                        str(tmp_path / "out.txt"), "--iters", "1"], capture_output=True, text=True, timeout=120)
    assert f"clm_size: {len(g['clm'])} row_size: {g['n'] + 1}" in r.stdout, (r.stdout, r.stderr)
    assert f"vertices: {g['n']}" in r.stdout
    assert not os.path.exists(src + ".lpa_params")
    assert open(src + ".graph", "rb").read() == open(ref + ".graph", "rb").read()


def test_reference_file_set_of_another_threshold_is_rebuilt(tmp_path):
    """asked for threshold 32 on a set written with threshold 0 and no edge
    list: the set is inconsistent, the rebuild needs the edge list, and the
    driver stops instead of running on the wrong split"""
    ref = os.path.join(GOLDEN, "mixed400_t0", "edges.txt")
    src = str(tmp_path / "edges.txt")
    for suf in (".graph", ".offset", ".vertices.bin", ".row", ".clm"):
        shutil.copy(ref + suf, src + suf)
    r = subprocess.run([os.path.join(BIN, "nova"), src, "1", "1", "32", str(tmp_path / "out.txt"), "--iters", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "clm_size" not in r.stdout
    assert "cannot open" in r.stderr
