"""libtsvm's C-ABI on the CPU: the library loads, exports every declared entry point, and
its host-only functions (Omega stream, LIBSVM reader, label map, model writers) match the
oracle and the golden files.  No GPU is touched."""
import os
import re

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = os.path.join(ROOT, "tests", "golden")


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            names |= set(re.findall(r"\b(tsvm_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import tensorsvm_amd as T
    L = T.load()
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_gaussrand_bit_exact_vs_glibc():
    import tensorsvm_amd as T
    for count in (1, 2, 7, 8192, 1_000_003):
        a = T.gaussrand_fill(count, nthreads=3)
        b = O.gaussrand(count)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), count
    g = np.load(os.path.join(G, "glibc_rand.npz"))
    a = T.gaussrand_fill(g["gauss_bits"].size)
    assert np.array_equal(a.view(np.uint32), g["gauss_bits"])


def test_reader_matches_oracle_and_golden(tmp_path):
    import tensorsvm_amd as T
    g = np.load(os.path.join(G, "tiny_libsvm.npz"))
    lab, x = T.capi.read_libsvm(os.path.join(G, "tiny.libsvm"))
    assert np.array_equal(x, g["x"]) and np.array_equal(lab, g["labels"])
    rng = np.random.default_rng(0)
    p = tmp_path / "r.libsvm"
    with open(p, "w") as f:
        for i in range(500):
            idx = np.sort(rng.choice(40, rng.integers(1, 12), replace=False)) + 1
            vals = rng.normal(size=idx.size)
            f.write(("+1" if i % 3 else "-1") + " " +
                    " ".join(f"{a}:{b:.6g}" for a, b in zip(idx, vals)) + "\n")
    l1, x1 = T.capi.read_libsvm(str(p))
    l2, x2 = O.libsvm_read(str(p))
    assert np.array_equal(l1, l2) and np.array_equal(x1, x2)
    l3, x3 = T.capi.read_libsvm(str(p), scale256=True)
    l4, x4 = O.libsvm_read(str(p), scale256=True)
    assert np.array_equal(x3, x4)


def test_reader_many_chunks_and_writer_round_trip(tmp_path):
    """A file of several MiB is parsed by several threads (chunks cut at line ends); the
    rows, labels and index range match the oracle's single-pass reader.  The writer's text
    reads back to the same floats."""
    import tensorsvm_amd as T
    rng = np.random.default_rng(5)
    p = tmp_path / "big.libsvm"
    seps = [" ", "\t", "  ", " \t"]
    with open(p, "w") as f:
        for i in range(60000):
            idx = np.sort(rng.choice(60, rng.integers(1, 16), replace=False)) + 1
            vals = rng.normal(size=idx.size) * 10.0 ** rng.integers(-6, 6, idx.size)
            sp = seps[i % 4]
            end = "\r\n" if i % 7 == 0 else (" \n" if i % 5 == 0 else "\n")
            f.write(("+1" if i % 3 else "-1") + sp +
                     sp.join(f"{a}:{b:.8g}" for a, b in zip(idx, vals)) + end)
    assert p.stat().st_size > 4 << 20
    l1, x1 = T.capi.read_libsvm(str(p))
    l2, x2 = O.libsvm_read(str(p))
    assert x1.shape == x2.shape == (60000, 60)
    assert np.array_equal(l1, l2) and np.array_equal(x1.view(np.uint32), x2.view(np.uint32))
    # writer -> reader: exact floats, zeros dropped, the width set by the last non-zero column
    x = (rng.standard_normal((3001, 17)) * 10.0 ** rng.integers(-30, 30, (3001, 17))).astype(np.float32)
    x[rng.random(x.shape) < 0.3] = 0.0
    x[:, -1] = 1.5  # keeps d = 17
    lab = np.where(rng.random(3001) < 0.5, 1.0, -1.0).astype(np.float32)
    q = tmp_path / "w.libsvm"
    T.capi.write_libsvm(q, lab, x)
    l3, x3 = T.capi.read_libsvm(str(q))
    assert np.array_equal(l3, lab) and np.array_equal(x3.view(np.uint32), x.view(np.uint32))
    l4, x4 = O.libsvm_read(str(q))
    assert np.array_equal(x4, x3)


def test_reader_rejects_bad_input(tmp_path):
    import tensorsvm_amd as T
    for bad in ("+1 2:1 1:3\n", "+1 a:1\n", "x 1:2\n", "+1 1:2\n\n"):
        p = tmp_path / "b.libsvm"
        p.write_text(bad)
        with pytest.raises(T.TsvmError):
            T.capi.read_libsvm(str(p))
    with pytest.raises(T.TsvmError):
        T.capi.read_libsvm(str(tmp_path / "missing"))


def test_label_map():
    import tensorsvm_amd as T
    lab = np.array([3, 3, 7, 3, 7, 9], dtype=np.float32)
    m1, p1, n1, b1 = T.capi.map_labels(lab)
    m2, p2, n2, b2 = O.label_map(lab)
    assert (p1, n1, b1) == (p2, n2, b2) == (3, 7, 1)
    assert np.array_equal(m1, m2)


def test_model_writers_golden(tmp_path):
    import tensorsvm_amd as T
    m = np.load(os.path.join(G, "model_inputs.npz"))
    p = str(tmp_path / "m.txt")
    T.capi.write_model_rbf(p, m["alpha"], m["labels"], m["inst"], float(m["C"]), float(m["gamma"]),
                           float(m["b"]), int(m["pos"]), int(m["neg"]))
    assert open(p).read() == open(os.path.join(G, "model_rbf.txt")).read()
    p2 = str(tmp_path / "l.txt")
    T.capi.write_model_linear(p2, m["w"], float(m["b"]), int(m["pos"]), int(m["neg"]))
    assert open(p2).read() == open(os.path.join(G, "model_linear.txt")).read()


def test_context_without_device_is_an_error():
    import torch

    import tensorsvm_amd as T
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(T.TsvmError):
        T.Context(0)

def test_model_reader_on_golden_models(tmp_path):
    """The golden model files parse back to the values that were written (no GPU)."""
    import tensorsvm_amd as T
    m = np.load(os.path.join(G, "model_inputs.npz"))
    md = T.Model(os.path.join(G, "model_rbf.txt"))
    assert md.kernel_type == 2
    assert md.gamma == float("%.7f" % float(m["gamma"]))
    assert md.rho == float("%f" % -float(m["b"]))
    assert md.labels == (int(m["pos"]), int(m["neg"]))
    assert md.nsv == int((m["alpha"] > 1e-6).sum())
    assert md.d <= m["inst"].shape[1]
    ml = T.Model(os.path.join(G, "model_linear.txt"))
    assert ml.kernel_type == 0 and ml.d == m["w"].shape[0]
    assert abs(ml.rho + float(m["b"])) <= 1e-15 * (1 + abs(float(m["b"])))
    bad = tmp_path / "bad.model"
    bad.write_text("svm_type c_svc\nkernel_type poly\nSV\n")
    with pytest.raises(T.TsvmError):
        T.Model(str(bad))
    with pytest.raises(T.TsvmError):
        T.Model(str(tmp_path / "missing.model"))

