"""GPU parity of nrg_broaden against the oracle (SpectrumRealFreq::continuous with BR_NEW,
c++/spectrum.hpp:180-209, c++/spectral.hpp:39-71). Tolerance: every mesh value within 1e-13 of the
same sum taken over |w| (the device exp/log differ from the host's by a few ulp, and the sums are
taken in another order)."""
import math

import numpy as np
import pytest

import oracle
from nrgljubljana_amd import NrgError

pytestmark = pytest.mark.gpu


def make_mesh(broaden_max=10.0, broaden_min=1e-8, ratio=1.05):
    """make_mesh (c++/spectrum.hpp:173-178): E = broaden_max, /ratio while E > broaden_min."""
    out, E = [], broaden_max
    while E > broaden_min:
        out.append(E)
        E /= ratio
    return np.array(out)


def binned(rng, log10emin=-12.0, log10emax=2.0, bpd=1000, trim=1e-16):
    """A smooth spectrum on the standard grid with noise, trimmed as Bins::trim does (c++/bins.hpp:164-185)."""
    n = int((log10emax - log10emin) * bpd + 1)
    e = 10.0 ** (log10emin + np.arange(n) / bpd)
    w = np.exp(-((np.log10(e) + 3.0) / 2.0) ** 2) * e * rng.uniform(0.5, 1.5, n)
    w[rng.uniform(size=n) < 0.3] = 0.0
    keep = np.abs(w[:-1]) >= trim * np.diff(e)
    keep = np.append(keep, True)   # the last bin is always kept
    return e[keep], w[keep]


def run(ctx, pe, pw, ne, nw, mesh, alpha, omega0):
    import torch
    t = [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda() for a in (pe, pw, ne, nw, mesh)]
    op = torch.zeros(len(mesh), dtype=torch.float64, device="cuda")
    on = torch.zeros_like(op)
    ctx.broaden(*t, alpha, omega0, op, on)
    torch.cuda.synchronize()
    return op.cpu().numpy(), on.cpu().numpy()


@pytest.mark.parametrize("seed,alpha,omega0", [(1, 0.3, 1e-9), (2, 0.6, 1e-4), (3, 0.15, 1e-6)])
def test_matches_oracle(ctx, seed, alpha, omega0):
    rng = np.random.default_rng(seed)
    pe, pw = binned(rng)
    ne, nw = binned(rng)
    mesh = make_mesh()
    print(f"broaden seed {seed}: {len(pe)} + {len(ne)} peaks, {len(mesh)} mesh points", flush=True)
    gp, gn = run(ctx, pe, pw, ne, nw, mesh, alpha, omega0)
    rp, rn = oracle.broaden(pe, pw, ne, nw, mesh, alpha, omega0)
    sp, sn = oracle.broaden(pe, np.abs(pw), ne, np.abs(nw), mesh, alpha, omega0)
    assert np.all(np.abs(gp - rp) <= 1e-13 * sp + 1e-300)
    assert np.all(np.abs(gn - rn) <= 1e-13 * sn + 1e-300)


def test_total_weight(ctx):
    # a single peak far above omega0 integrates to its weight on a fine mesh
    mesh = np.exp(np.linspace(math.log(1e-4), math.log(1e2), 6000))
    gp, gn = run(ctx, [0.5], [2.5], [], [], mesh, 0.3, 1e-8)
    assert np.trapezoid(gp, mesh) == pytest.approx(2.5, rel=1e-6)
    assert not gn.any()


def test_empty_and_bad_arguments(ctx):
    import torch
    mesh = make_mesh()
    gp, gn = run(ctx, [], [], [], [], mesh, 0.3, 1e-8)
    assert not gp.any() and not gn.any()
    z = torch.zeros(3, dtype=torch.float64, device="cuda")
    with pytest.raises(NrgError):
        ctx.broaden(z, z, z, z, z, -0.3, 1e-8, z, z)
    with pytest.raises(NrgError):
        ctx.broaden(z, z, z, z, z, 0.3, 0.0, z, z)
