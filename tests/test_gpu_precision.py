"""FP32 storage against the FP64 run (the precision variant of the plain case).

With TPLS_FP32, phi, u, v, w, their histories, the WENO5 fluxes and the
Helmholtz sweeps are FP32.  The pressure CG (its inner products and residual),
the divergence and the volume sums stay FP64.  Both runs start from the same
seeded state.  The tests report, per checkpoint, the max and L2 relative
errors of phi, u, v, w and p (p modulo its mean) and the relative difference
of the liquid volume V = sum H_eps(phi) dV.  The bounds are the errors FP32
storage is expected to produce, not parity tolerances.
"""
import dataclasses

import numpy as np
import pytest

from tests.compare import demean, rows
from tpls_amd.config import CaseParams

pytestmark = pytest.mark.gpu


def errors(a, b):
    d = a - b
    return float(np.abs(d).max() / np.abs(b).max()), float(np.linalg.norm(d) / np.linalg.norm(b))


def test_volume_matches_numpy():
    from tpls_amd.case import Case
    from tpls_amd.initial import make_initial
    p = CaseParams(nx=32, ny=16, nz=24, seed=4)
    with Case(p) as c:
        c.load(*make_initial(p))
        c.step(2)
        phi = c.field("phi2")[1:p.maxl, 1:p.maxm, 1:p.maxn]
        v_gpu = c.volume()
    sw = p.smooth_width
    pi = float(np.float32(3.1415926535))
    H = np.where(phi < -sw, 0.0, np.where(phi > sw, 1.0,
                 0.5 + phi / (2 * sw) + np.sin(pi * phi / sw) / (2 * pi)))
    v_np = H.sum() * p.dx ** 3
    print(f"[volume] gpu {v_gpu:.15e} numpy {v_np:.15e}", flush=True)
    assert v_gpu == pytest.approx(v_np, rel=1e-12)
    assert 0.4 < v_gpu / ((p.maxl - 1) * (p.maxm - 1) * (p.maxn - 1) * p.dx ** 3) < 0.6


@pytest.mark.parametrize("pc", ["mg"])
def test_fp32_against_fp64_128x32x64(pc):
    from tpls_amd.case import Case
    from tpls_amd.config import CHANNEL_128x32x64
    from tpls_amd.initial import make_initial
    p = dataclasses.replace(CHANNEL_128x32x64, ksp_pc=pc)
    init = make_initial(p)
    c64, c32 = Case(p, precision="fp64"), Case(p, precision="fp32")
    c64.load(*init)
    c32.load(*init)
    P = p.maxm - 1
    worst = {}
    for s in range(1, 101):
        c64.step(1)
        c32.step(1)
        if s in (1, 10, 50, 100):
            line = []
            for f in ("phi2", "u2", "v2", "w2"):
                mx, l2 = errors(rows(c32.field(f), P), rows(c64.field(f), P))
                worst[f] = max(worst.get(f, 0.0), mx)
                line.append(f"{f} {mx:.1e}/{l2:.1e}")
            mx, l2 = errors(rows(demean(c32.field("pres"), p.maxl, P, p.maxn), P),
                            rows(demean(c64.field("pres"), p.maxl, P, p.maxn), P))
            worst["pres"] = max(worst.get("pres", 0.0), mx)
            v64, v32 = c64.volume(), c32.volume()
            dv = abs(v32 - v64) / v64
            worst["volume"] = max(worst.get("volume", 0.0), dv)
            print(f"[fp32] step {s} its {c64.diag()['ksp_its']}/{c32.diag()['ksp_its']} "
                  + " ".join(line) + f" pres {mx:.1e}/{l2:.1e} volume {dv:.1e}", flush=True)
    c64.close()
    c32.close()
    # measured: velocity <= 6e-6, phi <= 6e-7, p <= 4e-6, volume <= 1.4e-8
    # after 100 steps
    for f in ("phi2", "u2", "v2", "w2"):
        assert worst[f] < 3e-5, (f, worst[f])
    assert worst["pres"] < 3e-5
    assert worst["volume"] < 2e-7
