"""The command-line front ends on the CPU: usage and argument errors of tensorsvm-train
(tensorsvm-train.cu:1061-1156 parses the same options) and bench.py's one-GPU guard.
Every case here ends before a GPU context is created."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tensorsvm_amd", "tensorsvm-train")


def _run(args):
    if not os.path.exists(EXE):
        pytest.fail("tensorsvm-train is not built (make -C tensorsvm_amd)")
    return subprocess.run([EXE] + args, capture_output=True, text=True, timeout=60)


def test_usage_without_training_file():
    r = _run([])
    assert r.returncode != 0
    assert "Usage: tensorsvm-train" in r.stdout + r.stderr


@pytest.mark.parametrize("args,msg", [
    (["-precision", "fp4", "x"], "unknown precision fp4"),
    (["-stop", "gapp", "x"], "unknown stop rule gapp"),
    (["-init", "zero", "x"], "unknown initial point zero"),
    (["-t", "1", "x"], "unimplemented -t 1"),
    (["-k"], "missing value for -k"),
])
def test_rejects_bad_options(args, msg):
    r = _run(args)
    assert r.returncode != 0
    assert msg in r.stdout + r.stderr


def test_missing_training_file(tmp_path):
    r = _run(["-q", str(tmp_path / "absent.libsvm")])
    assert r.returncode != 0
    assert "Can't open" in r.stderr


def test_malformed_training_file(tmp_path):
    p = tmp_path / "bad.libsvm"
    p.write_text("+1 1:0.5 2:x\n")
    r = _run(["-q", str(p)])
    assert r.returncode != 0
    assert "Wrong input format" in r.stderr


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_rank_limit_checked_before_reading():
    """-k above the solver's limit is refused before the training file is read (the factor
    would otherwise run for minutes before the interior-point solver rejects it)."""
    r = _run(["-t", "2", "-k", "2048", "absent.libsvm"])
    assert r.returncode != 0
    assert "-k 2048 outside 1..1024" in r.stderr


def test_three_class_labels_reported(tmp_path):
    """Labels outside the first two seen are reported (the reference prints an error per
    such label and continues, tensorsvm-train.cu:890-895)."""
    p = tmp_path / "three.libsvm"
    p.write_text("3 1:0.5\n7 1:0.1\n9 1:0.2\n3 1:0.4\n")
    r = _run(["-t", "2", "-k", "2", str(p)])
    assert "Error: 1 labels outside {3, 7}" in r.stdout
