"""CPU-side checks of the C ABI: the library builds, loads, exports every
symbol include/caps_amd.h declares, and the descriptor helpers behave like
verifyDescriptor / numEntriesPerProc (matrix.cpp:3-12).  No compute calls:
there is no GPU here."""
import ctypes
import os
import re

import pytest

import caps_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "caps_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(caps_[a-z_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(caps_amd.LIB_PATH):
        from caps_amd import build
        build.build()
    return caps_amd.load()


def test_every_declared_symbol_is_exported(lib):
    names = _declared()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    # and the Python mirror binds exactly those
    assert sorted(caps_amd.SIGNATURES) == names


def test_descriptor_layout_matches_reference():
    # struct MatDescriptor is 7 ints in this order (matrix.h:8-18)
    assert ctypes.sizeof(caps_amd.MatDescriptor) == 7 * 4
    names = [f[0] for f in caps_amd.MatDescriptor._fields_]
    assert names == ["lda", "nrec", "nproc", "nprocr", "nprocc", "nproc_summa", "bs"]


def test_verify_descriptor(lib):
    D = caps_amd.MatDescriptor.one_device
    caps_amd.verify_descriptor(D(1024, 3, 1))
    caps_amd.verify_descriptor(D(392, 3, 7))
    with pytest.raises(caps_amd.CapsError) as e:
        caps_amd.verify_descriptor(D(100, 3, 1))       # 100 % 8 != 0
    assert e.value.code == 1
    with pytest.raises(caps_amd.CapsError):
        caps_amd.verify_descriptor(D(28, 2, 2))        # 28 % (4*2) != 0
    assert "multiple" in lib.caps_last_error().decode()
    with pytest.raises(caps_amd.CapsError):
        caps_amd.verify_descriptor(caps_amd.MatDescriptor(28, 1, 7, 7, 1, 1, 1))  # nproc > 1
    assert "nproc" in lib.caps_last_error().decode()


def test_num_entries(lib):
    assert caps_amd.num_entries_per_proc(caps_amd.MatDescriptor.one_device(7168, 3)) == 7168 ** 2
    # numEntriesPerProc = lda^2 / nproc / nproc_summa (matrix.cpp:9-12)
    assert caps_amd.num_entries_per_proc(caps_amd.MatDescriptor(7168, 3, 7, 7, 1, 1, 1)) == \
        7168 ** 2 // 7


def test_host_array_checks_before_device(lib):
    import numpy as np
    d = caps_amd.MatDescriptor.one_device(8, 1)
    a = np.zeros(64)
    with pytest.raises(caps_amd.CapsError):
        caps_amd.multiply(a, a, np.zeros(63), d)
    with pytest.raises(caps_amd.CapsError):
        caps_amd.multiply(a.astype(np.float32), a, a, d)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(caps_amd, "_lib", None)
    with pytest.raises(caps_amd.CapsError) as e:
        caps_amd.load(str(tmp_path / "nope.so"))
    assert "no CPU fallback" in str(e.value)


def test_candidate_enumeration(lib):
    D = caps_amd.MatDescriptor.one_device
    # n = 8192, nrec = 2 (leaves of 2048): cutoffs 0..2, every b/d pattern of
    # the recursed levels, both FP64 block tiles
    fp = caps_amd.candidates(D(8192, 2, 1), caps_amd.MASK_FP64)
    assert sorted((c.cutoff, c.pattern.decode(), c.tile) for c in fp) == sorted(
        (cut, p, t) for cut, p in [(0, "bb"), (0, "bd"), (0, "db"), (0, "dd"), (1, "bb"),
                                   (1, "db"), (2, "bb")] for t in (0, 1))
    # prefix patterns come first (the learned planner keeps the first of equal features)
    assert [c.pattern.decode() for c in fp if c.cutoff == 0 and c.tile == 0] == [
        "bb", "bd", "dd", "db"]
    crt = caps_amd.candidates(D(8192, 2, 1), caps_amd.MASK_CRT_INT8)
    # 256-aligned leaves: INT8 tiles 256 and 128
    assert {c.moduli for c in crt} == {12, 13, 14, 15, 16} and len(crt) == 7 * 5 * 2
    assert {c.tile for c in crt} == {0, 1}
    # leaves not a multiple of 128: no cutoff
    small = caps_amd.candidates(D(392, 3, 7), caps_amd.MASK_FP64 | caps_amd.MASK_CRT_INT8)
    assert all(c.cutoff == 0 for c in small)
    # 8 patterns; FP64 two tiles, residue one (leaves of 49 take the 128 tiles)
    assert len(small) == 8 * 2 + 8 * 5
    deep = caps_amd.candidates(D(8192, 6, 1), caps_amd.MASK_FP64)
    # beyond 4 recursed levels only the prefix patterns
    assert len({c.pattern.decode() for c in deep if c.cutoff == 0}) == 7
    f = caps_amd.candidate_features(D(8192, 2, 1), crt[0], 10.0)
    # mode ordinal 1 (residue INT8); leaves of 2048: alignment log2(256) / 8 = 1
    assert list(f[:1]) == [13.0] and f[1] == 10.0 and f[6] == 1.0 and f[7] == 1.0
    f = caps_amd.candidate_features(D(392, 3, 7), small[0], 0.0)
    assert f[6] == 0.0 and f[7] == 0.0  # FP64; leaves of 49 are odd
