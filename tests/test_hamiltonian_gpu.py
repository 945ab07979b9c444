"""GPU parity of nrg_hamiltonian against the oracle (hamiltonian(), c++/core.hpp:72-88, and the
building steps of c++/matrix.hpp:32-111).

Tolerance: none. The kernel applies every element's terms in the caller's order with separately
rounded products and sums, as the oracle does, so the arenas must be bit-identical."""
import numpy as np
import pytest

import oracle
from nrgljubljana_amd import NRG_SYEV_FP64, NrgError
from tests.fdm_cases import np_ptr
from tests.ham_cases import arena, block_term, diag_term, ident_term, random_step

pytestmark = pytest.mark.gpu


def device_terms(terms):
    import torch
    keep = {}

    def ptr(a):
        if id(a) not in keep:   # the array is held too, so its id cannot be reused meanwhile
            keep[id(a)] = (a, torch.from_numpy(a).cuda())
        return keep[id(a)][1].data_ptr()

    return [t.to_struct(ptr) for t in terms], keep


def gpu_assemble(ctx, sizes, terms, fill=np.nan):
    import torch
    off, tot = arena(sizes)
    H = torch.full((max(tot, 1),), fill, dtype=torch.float64, device="cuda")
    structs, keep = device_terms(terms)
    ctx.hamiltonian(sizes, off, H, structs)
    torch.cuda.synchronize()
    return H.cpu().numpy()[:tot], H


def cpu_assemble(sizes, terms):
    off, tot = arena(sizes)
    H = np.full(max(tot, 1), np.nan)
    oracle.hamiltonian(sizes, off, H, [t.to_struct(np_ptr) for t in terms])
    return H[:tot]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_steps_bit_identical(ctx, seed):
    sizes, terms = random_step(seed, nblk=16, nmax=400)
    print(f"hamiltonian parity: seed {seed}, {len(sizes)} blocks (max n {max(sizes)}), {len(terms)} terms", flush=True)
    got, _ = gpu_assemble(ctx, sizes, terms)
    ref = cpu_assemble(sizes, terms)
    bad = np.nonzero(got != ref)[0]
    assert bad.size == 0, (f"{bad.size} elements differ, first at {bad[:5]}, max |diff| "
                           f"{np.max(np.abs(got - ref)):.3e}, max |ref| {np.max(np.abs(ref)):.3e}")


def test_tile_edges_and_large_block(ctx):
    # sub-blocks straddling 64-element tile edges, an identity coupling crossing tiles, a 1100 block
    rng = np.random.default_rng(7)
    n = 1100
    terms = [diag_term(0, 0, rng.uniform(0, 5, n), factor=1.3),
             block_term(0, 63, 127, rng.standard_normal((65, 600)), 0.5, pad=3),
             block_term(0, 1, 700, rng.standard_normal((399, 60)), -0.25, trans=True),
             ident_term(0, 30, 600, 500, 0.75),
             diag_term(0, 500, None, shift=-0.5, count=600)]
    got, _ = gpu_assemble(ctx, [n, 1, 64, 65], terms + [diag_term(1, 0, [2.0]), diag_term(3, 0, None, shift=1.0,
                                                                                       count=65)])
    assert np.array_equal(got, cpu_assemble([n, 1, 64, 65], terms + [diag_term(1, 0, [2.0]),
                                                                      diag_term(3, 0, None, shift=1.0, count=65)]))


def test_feeds_the_eigensolver(ctx):
    # identity coupling of two m-blocks: eigenvalues (a+b)/2 -+ sqrt(((a-b)/2)^2 + t^2), m times each
    import torch
    m, a, b, t = 200, 0.25, -0.5, 0.8
    sizes = [2 * m]
    off, tot = arena(sizes)
    H = torch.empty(tot, dtype=torch.float64, device="cuda")
    structs, keep = device_terms([diag_term(0, 0, None, shift=a, count=m), diag_term(0, m, None, shift=b, count=m),
                                  ident_term(0, 0, m, m, t)])
    ctx.hamiltonian(sizes, off, H, structs)
    w = torch.zeros(2 * m, dtype=torch.float64, device="cuda")
    ctx.syev_batched(H, np.array(sizes, dtype=np.int64), off, w, np.zeros(1, dtype=np.int64), NRG_SYEV_FP64)
    torch.cuda.synchronize()
    r = np.sqrt(((a - b) / 2) ** 2 + t * t)
    want = np.array([(a + b) / 2 - r] * m + [(a + b) / 2 + r] * m)
    assert np.max(np.abs(w.cpu().numpy() - want)) <= 1e-14


def test_empty_and_bad_arguments(ctx):
    import torch
    got, H = gpu_assemble(ctx, [5, 3], [])
    assert np.all(got == 0)    # every element written
    structs, keep = device_terms([block_term(0, 0, 3, np.ones((2, 2)), 1.0)])
    off, _ = arena([5, 3])
    bad = [type(structs[0]).from_buffer_copy(structs[0]) for _ in range(4)]
    bad[0].block = 2
    bad[1].col = 4          # 3 + 2 > 5
    bad[2].src = None
    bad[3].kind = 7
    for s in bad:
        with pytest.raises(NrgError):
            ctx.hamiltonian([5, 3], off, H, [s])
    ctx.hamiltonian([5, 3], off, H, structs)
    torch.cuda.synchronize()
