"""The MRT cavity on the oracle: the reference's own acceptance test, and where the
reference's scheme (collide_mrt_2d! after stream_2d! and the Zou–He lid,
src/kernels/collide_mrt_2d.jl:43-125, src/kernels/boundary_2d.jl:13-44, loop of
test/test_mrt.jl:19-25) is stable in FP64.

The stability cases decide whether a cavity that diverges in the batched FP32-storage runs
diverges because of its storage or because the scheme itself has no stable solution there:
the FP64 oracle runs the reference's operation order, so a divergence here is the
reference's."""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _mrt_cavity(kmi, N, nu, u_lid):
    return kmi.cavity_desc(N, nu, u_lid, storage="f64", collision="mrt")


def test_mrt_cavity_re100_matches_ghia(po, kmi):
    """test/test_mrt.jl:6-44: 64², Re = 100, 15000 steps, max error on the six Ghia points
    below 0.15 and no NaN."""
    g = json.load(open(os.path.join(GOLDEN, "ghia_re100_mrt.json")))
    N, Re, ul = 64, 100.0, 0.1
    f = po.rest((N, N))
    po.run(_mrt_cavity(kmi, N, ul * N / Re, ul), f, None, 15000)
    ux = po.moments(f)["ux"]
    assert not np.any(np.isnan(ux))
    u = ux[N // 2 - 1, :] / ul  # i_center = N ÷ 2 (1-based)
    err = max(abs(u[min(max(int(round(y * (N - 1))), 0), N - 1)] - ug) for y, ug in zip(g["y"], g["u"]))
    assert err < g["tolerance"], err


def _first_nonfinite(po, desc, N, steps, chunk=250):
    f = po.rest((N, N))
    done = 0
    while done < steps:
        po.run(desc, f, None, chunk)
        done += chunk
        if not np.all(np.isfinite(f)):
            return done
    return None


@pytest.mark.parametrize("N,Re,u_lid", [
    (64, 2818.6, 0.1),    # the lowest-viscosity cavity of the seed-2024 ensemble (tau = 0.5068)
    (128, 2400.0, 0.12),  # tau = 0.5192
    (128, 1200.0, 0.06),  # tau = 0.5192 at half the lid speed
])
def test_reference_scheme_diverges_below_tau_0522(po, kmi, N, Re, u_lid):
    """In FP64 the reference's MRT cavity blows up within 2000 steps from rest once
    tau = 3 nu + 1/2 falls to ~0.52, whatever the lid speed: these cavities have no stable
    solution in the reference's scheme, so no storage precision can converge them."""
    nu = u_lid * N / Re
    at = _first_nonfinite(po, _mrt_cavity(kmi, N, nu, u_lid), N, 2000)
    assert at is not None, f"tau {3 * nu + 0.5:.4f} stayed finite for 2000 steps"


@pytest.mark.parametrize("N,Re,u_lid", [(128, 2000.0, 0.12), (128, 1000.0, 0.06)])
def test_reference_scheme_stable_at_tau_0523(po, kmi, N, Re, u_lid):
    """tau = 0.523 (the lowest of the learned warm start's parameter range) stays finite over
    6000 steps from rest, its speed within twice the lid's (the corner overshoot of an
    under-resolved Re = 2000 cavity reaches ~1.3 u_lid)."""
    nu = u_lid * N / Re
    desc = _mrt_cavity(kmi, N, nu, u_lid)
    f = po.rest((N, N))
    po.run(desc, f, None, 6000)
    assert np.all(np.isfinite(f))
    m = po.moments(f)
    assert float(np.max(np.hypot(m["ux"], m["uy"]))) < 2.0 * u_lid
