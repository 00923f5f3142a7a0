"""Handle teardown in any order: a fresh interpreter builds and frees three small Galerkin systems (operator,
2-level preconditioner, device buffers, a GMRES-DR checkpoint) on its own context, closes the context
while some of its handles are still alive, releases the rest afterwards, leaves more for the exit hook,
and must exit with status 0 and no C++ runtime abort."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import numpy as np
from voxscatter_amd import api, capi
from voxscatter_amd.galerkin import GalerkinSystem

ctx = api.Context(0)
keep = []
for i, n in enumerate((16, 24, 32)):
    rng = np.random.default_rng(i)
    eps = (1.5 + rng.uniform(size=(n, n, n))) ** 2 + 0.01j
    ko = 2 * np.pi
    s = GalerkinSystem.build(eps, 1.0 / 35, ko, precisions=capi.VS_FP64 | capi.VS_FP32, ctx=ctx)
    pc = s.preconditioner(2, precisions=capi.VS_FP64 | capi.VS_FP32)
    x, flag, relres, it, res = api.gmres(s.op, s.b, 20, 1e-3, 2, pc)
    buf = api.DeviceArray.from_numpy(s.b, np.complex128, ctx)
    rec = api.Recycle(ctx)
    print(f"[teardown] system {i}: n={n} unknowns={s.op.n} gmres flag={flag} relres={relres:.2e}", flush=True)
    if i == 0:
        pc.close(); s.op.close(); buf.close(); del rec
    elif i == 1:
        keep.append((s, pc, buf, rec))   # left to the exit hook
    else:
        ctx.close()          # the context goes first; its handles still work and are freed afterwards
        y = pc.apply(s.b)
        s.op.close(); pc.close(); buf.close(); rec.close()
print("[teardown] done", flush=True)
"""


def test_handles_released_in_any_order_exit_cleanly():
    env = dict(os.environ)
    env["PYTHONPATH"] = ROOT + os.pathsep + env.get("PYTHONPATH", "")
    r = subprocess.run([sys.executable, "-c", SCRIPT], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    print(r.stderr[-4000:])
    assert r.returncode == 0
    assert "terminate called" not in r.stderr
    assert "[teardown] done" in r.stdout
