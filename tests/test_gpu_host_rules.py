"""GPU checks of two host rules.

* The memory budget of setAvailableMemory (memory.cpp:31-33) drives the
  automatic BFS/DFS choice (multiply.cpp:43-66 with enoughMemory,
  memory.cpp:22-29): a tiny budget gives all-DFS, whose counters read
  "DFS,DFS," (setExecutionType, counters.cpp:49-59), a large one all-BFS;
  in FP64 both give the same bits (BFS and DFS evaluate the same Winograd
  expression trees).
* A cached plan owns its device workspace per stream: the same buffers
  multiplied on two streams at once give the single-stream bits (the pattern
  has only batched levels, so C itself is written once, by the final
  combination; with recursive levels C's quadrants hold intermediate values,
  multiply.cpp:411-416, and calls sharing C must not overlap)."""
import numpy as np
import pytest

import caps_amd

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _setup():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    caps_amd.load()
    yield
    caps_amd.set_available_memory(0)
    caps_amd.release()


def _operands(n, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    A = torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    B = torch.rand(n * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    return A, B


@pytest.mark.parametrize("mode", [caps_amd.MODE_FP64, caps_amd.MODE_CRT_INT8])
def test_budget_flips_bfs_to_dfs(mode):
    n, r = 1024, 2
    d = caps_amd.MatDescriptor.one_device(n, r, 1)
    A, B = _operands(n, 3)
    out = {}
    for budget, want in ((1, "DFS,DFS,"), (1 << 40, "BFS,BFS,")):
        print(f"[gpu] budget {budget} doubles mode {mode}", flush=True)
        caps_amd.set_available_memory(budget)
        assert caps_amd.plan_choice(d, mode)["pattern"] == want.replace("DFS,", "d").replace(
            "BFS,", "b")
        C = torch.full_like(A, float("nan"))
        caps_amd.reset_counters()
        caps_amd.multiply_device(A, B, C, d, None, mode=mode)
        torch.cuda.synchronize()
        assert caps_amd.counters()["pattern"] == want
        out[budget] = C
    caps_amd.set_available_memory(0)
    # FP64: BFS and DFS are the same expression trees; residue mode: exact
    # modular arithmetic either way
    assert torch.equal(out[1], out[1 << 40])


def test_same_buffers_on_two_streams():
    n, r = 2048, 2
    d = caps_amd.MatDescriptor.one_device(n, r, 1)
    A, B = _operands(n, 4)
    ref = torch.empty_like(A)
    caps_amd.multiply_device(A, B, ref, d, "bb")
    torch.cuda.synchronize()
    C = torch.full_like(A, float("nan"))
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    print("[gpu] same buffers on two streams", flush=True)
    for _ in range(4):
        caps_amd.multiply_device(A, B, C, d, "bb", stream=s1)
        caps_amd.multiply_device(A, B, C, d, "bb", stream=s2)
    torch.cuda.synchronize()
    assert torch.equal(C, ref)


@pytest.mark.parametrize("mode,n,r,pat,cutoff", [
    (caps_amd.MODE_FP64, 1024, 2, "bb", 0),
    (caps_amd.MODE_FP64, 1344, 2, "bd", 0),     # leaves of 336: edge-checked tiles
    (caps_amd.MODE_FP64, 2048, 3, "bbb", 1),    # quadtree operands under the cutoff
    (caps_amd.MODE_CRT_INT8, 2048, 2, "bb", 0),  # leaves of 512: 256 vs 128 tiles
    (caps_amd.MODE_CRT_INT8, 4096, 3, "bbb", 1),
    (caps_amd.MODE_CRT_INT8, 2560, 2, "bd", 0),  # leaves of 640: partial 256-tile edges
])
def test_small_tile_gives_the_same_bits(mode, n, r, pat, cutoff):
    """The autotuner's tile dimension changes no arithmetic: every output
    element sums its products in the same K order with either block tile."""
    d = caps_amd.MatDescriptor.one_device(n, r, 1)
    A, B = _operands(n, 5)
    out = []
    for flags in (0, caps_amd.FLAG_SMALL_TILE):
        print(f"[gpu] tile flags={flags} mode={mode} n={n} {pat} cutoff={cutoff}", flush=True)
        C = torch.full_like(A, float("nan"))
        caps_amd.multiply_device(A, B, C, d, pat, mode=mode, flags=flags, cutoff=cutoff)
        torch.cuda.synchronize()
        out.append(C)
    assert torch.equal(out[0], out[1])
    caps_amd.release()


def test_destroyed_planner_model_is_uninstalled():
    """The library does not own an installed planner model: destroying it
    uninstalls it, and later CAPS_MODE_AUTO choices fall back to the rule
    without a model instead of reading freed memory."""
    d = caps_amd.MatDescriptor.one_device(2048, 2, 1)
    m = caps_amd.MLP(seed=1)
    norm = np.concatenate([np.zeros(8), np.ones(8), np.zeros(2), np.ones(2)]).astype(np.float32)
    m.use_as_planner(norm)
    print("[gpu] planner model installed, then destroyed", flush=True)
    caps_amd.plan_choice(d, caps_amd.MODE_AUTO)  # consults the installed model
    m.close()
    caps_amd.MLP._installed = None
    c = caps_amd.plan_choice(d, caps_amd.MODE_AUTO)
    assert c["predicted_ms"] < 0
