"""CPU checks of host-side rules that need no device: the memory budget that
drives the automatic BFS/DFS choice (memory.cpp:22-33, multiply.cpp:43-66),
the validation of tuning files, and the persistent-grid tile schedule."""
import os
import subprocess

import pytest

import caps_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
D = caps_amd.MatDescriptor.one_device


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(caps_amd.LIB_PATH):
        from caps_amd import build
        build.build()
    return caps_amd.load()


@pytest.fixture
def budget(lib):
    yield caps_amd.set_available_memory
    caps_amd.set_available_memory(0)  # back to the default for later tests


def _pattern(desc, mode=caps_amd.MODE_FP64, moduli=0, flags=0):
    return caps_amd.plan_choice(desc, mode, moduli, flags)["pattern"]


def test_tiny_budget_forces_dfs(budget):
    d = D(8192, 2, 1)
    budget(1)
    assert _pattern(d) == "dd"
    assert _pattern(d, caps_amd.MODE_CRT_INT8, 14) == "dd"
    budget(1 << 40)  # no device here, so no free-memory cap: every level batched
    assert _pattern(d) == "bb"
    assert _pattern(d, caps_amd.MODE_CRT_INT8, 14) == "bb"


def test_budget_is_greedy_from_the_top_and_monotone(budget):
    # levels are decided top-down, as multiplyInternal decides them while it
    # recurses: a level is 'b' if the plan with it batched still fits.  A
    # batched level costs more the higher it sits, so as the budget grows the
    # number of 'b' levels never drops, and at equal counts the 'b's move up.
    d = D(8192, 3, 1)
    seen = []
    for k in range(0, 41):
        budget(1 << k)
        p = _pattern(d)
        if not seen or seen[-1] != p:
            seen.append(p)
    assert seen[0] == "ddd" and seen[-1] == "bbb", seen
    for a, b in zip(seen, seen[1:]):
        ca, cb = a.count("b"), b.count("b")
        assert cb >= ca, seen
        if ca == cb:  # 'b' as 0, 'd' as 1: a smaller string has its 'b's higher
            assert b.replace("b", "0").replace("d", "1") < a.replace("b", "0").replace("d", "1")
    assert "bdb" in seen and "ddb" in seen, seen


def test_default_budget_without_a_device(budget):
    d = D(8192, 2, 1)
    budget(0)
    default = _pattern(d)
    budget(1 << 28)  # memory.cpp:4: 2^28 doubles when no device answers
    assert _pattern(d) == default


def test_reference_rule_ignores_the_budget(budget):
    budget(1 << 40)
    assert _pattern(D(8192, 2, 1), flags=caps_amd.FLAG_REFERENCE_AUTO) == "dd"


def test_negative_budget_is_rejected(lib):
    with pytest.raises(caps_amd.CapsError) as e:
        caps_amd.set_available_memory(-5)
    assert e.value.code == 1


HEADER = "# lda nrec bs mode_mask max_err -> mode moduli cutoff pattern ms err\n"


@pytest.mark.parametrize("line", [
    "1024 2 1 4 1e-14 7 14 0 bb 1.0 1e-15",       # unknown mode
    "1024 2 1 4 1e-14 2 40 0 bb 1.0 1e-15",       # moduli outside 2..16
    "1024 2 1 4 1e-14 2 1 0 bb 1.0 1e-15",        # moduli outside 2..16
    "1024 2 1 4 1e-14 0 14 0 bb 1.0 1e-15",       # FP64 with moduli
    "1024 2 1 4 1e-14 2 14 3 bb 1.0 1e-15",       # cutoff > nrec
    "1024 2 1 4 1e-14 2 14 0 bx 1.0 1e-15",       # pattern char
    "1024 2 1 4 1e-14 2 14 0 bbb 1.0 1e-15",      # pattern length != nrec
    "1024 2 1 4 0 2 14 0 bb 1.0 1e-15",           # error target
    "1024 2 1 4 1e-14 2 14 0 bb",                 # truncated
])
def test_bad_tuning_file_is_rejected(lib, tmp_path, line):
    f = tmp_path / "bad.txt"
    f.write_text(HEADER + "1024 2 1 4 1e-14 2 14 0 bb 1.0 1e-15\n" + line + "\n")
    with pytest.raises(caps_amd.CapsError) as e:
        caps_amd.tune_load(str(f))
    assert e.value.code == 1
    assert "line 3" in lib.caps_last_error().decode()


def test_good_tuning_file_loads(lib, tmp_path):
    f = tmp_path / "good.txt"
    f.write_text(HEADER + "\n1024 2 1 4 1e-14 2 14 0 bb 1.0 1e-15\n"
                 "2048 2 1 1 1e-13 0 0 1 bd 2.0 1e-14\n1024 1 1 4 1e-14 2 0 0 - 1.0 -1\n")
    caps_amd.tune_load(str(f))
    out = tmp_path / "saved.txt"
    caps_amd.tune_save(str(out))
    rows = [r.split() for r in out.read_text().splitlines() if not r.startswith("#")]
    assert ["1024", "2", "1", "4"] in [r[:4] for r in rows]
    assert ["2048", "2", "1", "1"] in [r[:4] for r in rows]


def test_tile_schedule_is_a_permutation(tmp_path):
    exe = tmp_path / "tilemap_check"
    src = os.path.join(ROOT, "tests", "native", "tilemap_check.cpp")
    r = subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "-x", "hip",
                        "--offload-arch=gfx950", src, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
