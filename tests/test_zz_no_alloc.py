"""hf_step_noalloc (no device allocation, scratch in a static device array)
against the context entry point hf_run, one step at a time: every field of
every block (ghosts included) and every TimeStep result bitwise, over calls in
a row whose second case is smaller than the first (the scratch keeps the
previous call's contents)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")   # before the library: one HIP runtime in the process

import goldens  # noqa: E402
from gpu_util import say  # noqa: E402
from hyperflow_amd import case as C  # noqa: E402
from hyperflow_amd import native  # noqa: E402

pytestmark = pytest.mark.gpu

TORCH_DTYPE = {"fp64": torch.float64, "fp32": torch.float32}
NP_DTYPE = {"fp64": np.float64, "fp32": np.float32}


def dense(fields, precision):
    return np.concatenate([np.asarray(U, NP_DTYPE[precision]).ravel() for U in fields])


def split(flat, case):
    out, o = [], 0
    g = case.ghost
    for b in case.blocks:
        n = 4 * (b.ny + 2 * g) * (b.nx + 2 * g)
        out.append(flat[o:o + n].reshape(4, b.ny + 2 * g, b.nx + 2 * g))
        o += n
    return out


def context_steps(case, fields, nsteps, precision):
    """The existing entry point: one hf_run step at a time, downloads after each."""
    outs = []
    with native.Solver(case, precision=precision) as s:
        s.upload_all(fields)
        for _ in range(nsteps):
            s.run(1)
            s.synchronize()
            outs.append(s.download_all())
        hist = s.dt_history()
    return outs, hist


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_steps_match_hf_run(precision):
    big = C.step_channel(40, "MinBee", 0.5)             # 8 blocks: compat dt, fix-up launches
    small = C.quad_case(37, 23, 4, scheme="Godunov", cfl=0.8)
    for case, nsteps in ((big, 3), (small, 2)):
        say(f"no-alloc {precision}: {len(case.blocks)} block(s) {case.scheme}, {nsteps} steps")
        f = case.initial_fields()
        ref, ref_hist = context_steps(case, f, nsteps, precision)
        a = torch.from_numpy(dense(f, precision)).cuda()
        b = torch.empty_like(a)
        cps = 1 if case.scheme == "Godunov" else 1 + len(case.blocks)
        dts = torch.zeros(cps, dtype=torch.float64, device="cuda")
        calls = np.zeros(1, np.int64)
        hist = []
        for k in range(nsteps):
            native.step_noalloc(case, a, b, calls, dts, precision=precision)
            out = split(b.double().cpu().numpy(), case)
            for blk, (x, y) in enumerate(zip(out, ref[k])):
                assert goldens.same(x, y), f"block {blk} after step {k + 1}"
            hist += dts.cpu().tolist()
            a, b = b, a
        assert calls[0] == nsteps * cps
        assert goldens.same(np.array(hist), ref_hist)


def test_noalloc_rejects_what_it_cannot_do():
    case = C.quad_case(20, 20, 1, scheme="Godunov")
    f = torch.from_numpy(dense(case.initial_fields(), "fp64")).cuda()
    with pytest.raises(native.HyperFlowError, match="different arrays"):
        native.step_noalloc(case, f, f, np.zeros(1, np.int64))
    # 20 000 MUSCL problems of 16^2 cells: 144 ghost cells each, a 46 MB ghost
    # table, beyond the 32 MiB scratch (rejected on the host before any device
    # work; the buffers are real anyway)
    many = [C.quad_case(16, 16, k % 50, scheme="MUSCLHancock", limiter="MinBee", cfl=0.5) for k in range(20000)]
    n = 20000 * 4 * 20 * 20
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    with pytest.raises(native.HyperFlowError, match="bytes of device scratch"):
        native.step_noalloc(many, x, y, np.zeros(20000, np.int64))
