"""CPU: the host-built transform tables, executed by the oracle's C loop, agree
bit-for-bit with the oracle's independent dense permute (U1) and Koszul-sign
permute (fZ2⊠U1).  Also pins the oracle itself with the reference's
permutation invariants (test/tensors/indexmanipulations.jl:55-98)."""
import numpy as np
import pytest

from tkmi.sectors import U1Irrep, FermionParity, ProductSector
from tkmi.spaces import GradedSpace, ProductSpace, HomSpace, U1Space
from tkmi.treetransformers import abelian_table

from oracle import abelian as O
from oracle import kernels as OK
from helpers import cfg1_spaces, cfg2_spaces, oracle_legs, kind_of, random_u1_space, random_perm


def _check_table(W, p1, p2, rng, alpha=1.0, beta=0.0):
    Wd = W.permute(p1, p2)
    tab = abelian_table(Wd, W, p1, p2)
    src = rng.standard_normal(W.dim())
    dst0 = rng.standard_normal(Wd.dim())
    dst = dst0.copy()
    OK.abelian_permute(tab, tuple(p1) + tuple(p2), src, dst, alpha, beta)
    cod, dom = oracle_legs(W)
    ref, ncod, ndom = O.permute(src, cod, dom, p1, p2, kind_of(W), alpha)
    assert (ncod, ndom) == oracle_legs(Wd)
    if beta != 0.0:
        # numpy rounds β·dst and (α·coeff)·src separately, as Julia's
        # scale(C, β) + scale(A, α) does in the reference's tensoradd!
        ref = beta * dst0 + ref
    np.testing.assert_array_equal(dst, ref)
    return dst0, ref


def _fused(beta, c, o):
    """round(β·c + o) with one rounding (a fused multiply-add), exactly."""
    from fractions import Fraction
    return np.array([float(Fraction(beta) * Fraction(float(ci)) + Fraction(float(oi))) for ci, oi in zip(c, o)])


@pytest.mark.parametrize("alpha,beta", [(0.3, 0.7), (-1.7, 1.0 / 3.0)])
def test_nondyadic_scalars_are_not_fused(alpha, beta):
    """β·dst + (α·coeff)·src with β, α not powers of two: the oracle loop (built with
    -ffp-contract=off) rounds both products, and these draws tell that apart from a
    fused multiply-add, so a kernel that fuses fails the same check on the GPU."""
    V, P = cfg1_spaces()
    W = HomSpace(ProductSpace(V, P), ProductSpace(V))
    rng = np.random.default_rng(11)
    Wd = W.permute((2, 0), (1,))
    dst0, ref = _check_table(W, (2, 0), (1,), rng, alpha, beta)
    cod, dom = oracle_legs(W)
    src = np.random.default_rng(11).standard_normal(W.dim())
    o, _, _ = O.permute(src, cod, dom, (2, 0), (1,), kind_of(W), alpha)
    assert Wd.dim() == len(o)
    fused = _fused(beta, dst0, o)
    assert np.count_nonzero(fused != ref) > 0


def test_cfg1_permute_table():
    V, P = cfg1_spaces()
    W = HomSpace(ProductSpace(V, P), ProductSpace(V))
    _check_table(W, (2, 0), (1,), np.random.default_rng(0))


@pytest.mark.parametrize("seed", range(16))
def test_random_u1_tables(seed):
    rng = np.random.default_rng(100 + seed)
    n1, n2 = int(rng.integers(1, 4)), int(rng.integers(0, 3))
    W = HomSpace(ProductSpace(*[random_u1_space(rng) for _ in range(n1)]),
                 ProductSpace(*[random_u1_space(rng) for _ in range(n2)]))
    if W.dim() == 0:
        pytest.skip("empty space")
    p1, p2 = random_perm(rng, n1, n2)
    alphas, betas = [1.0, -0.5, 0.3, -1.7], [0.0, 2.0, 0.7, 1.0 / 3.0]
    _check_table(W, p1, p2, rng, alpha=alphas[seed % 4], beta=betas[(seed // 4) % 4])


def _fz2u1_space(rng, dual):
    return GradedSpace({ProductSector(FermionParity(p), U1Irrep(q)): int(rng.integers(1, 3))
                        for p in (0, 1) for q in (-1, 0, 1) if rng.random() < 0.8}, dual=dual)


@pytest.mark.parametrize("seed", range(10))
def test_fermionic_tables_koszul(seed):
    rng = np.random.default_rng(200 + seed)
    n1, n2 = int(rng.integers(1, 3)), int(rng.integers(1, 3))
    W = HomSpace(ProductSpace(*[_fz2u1_space(rng, bool(rng.integers(0, 2))) for _ in range(n1)]),
                 ProductSpace(*[_fz2u1_space(rng, bool(rng.integers(0, 2))) for _ in range(n2)]))
    if W.dim() == 0:
        pytest.skip("empty space")
    p1, p2 = random_perm(rng, n1, n2)
    _check_table(W, p1, p2, rng, *((1.0, 0.0), (0.3, 0.7))[seed % 2])


def test_oracle_invariants():
    """The oracle's permute is norm- and inner-product-preserving and
    permute∘inverse = id (test/tensors/indexmanipulations.jl:55-75)."""
    rng = np.random.default_rng(3)
    V, P = cfg1_spaces()
    W = HomSpace(ProductSpace(V, P), ProductSpace(V, V.dual()))
    cod, dom = oracle_legs(W)
    for kind, legs in ((O.U1Kind, (cod, dom)),):
        x = rng.standard_normal(W.dim())
        y = rng.standard_normal(W.dim())
        for p1, p2 in [((3, 0), (1, 2)), ((), (2, 1, 0, 3)), ((1, 2, 3, 0), ())]:
            px, c2, d2 = O.permute(x, cod, dom, p1, p2, kind)
            py, _, _ = O.permute(y, cod, dom, p1, p2, kind)
            assert np.isclose(np.linalg.norm(px), np.linalg.norm(x), rtol=1e-14)
            assert np.isclose(px @ py, x @ y, rtol=1e-12)
            p = list(p1) + list(p2)
            ip = [p.index(i) for i in range(len(p))]
            n1 = len(p1)
            back, c3, d3 = O.permute(px, c2, d2, tuple(ip[:len(cod)]), tuple(ip[len(cod):]), kind)
            assert (c3, d3) == (cod, dom)
            np.testing.assert_array_equal(back, x)


def test_oracle_fermion_sign_closed_forms():
    """Koszul sign: swapping two odd legs gives -1, even-odd +1, and a
    permutation applied twice as an involution returns the identity."""
    K = O.FZ2U1Kind
    odd, even = (1, 0), (0, 0)
    assert O.koszul_sign((odd, odd), (1, 0), (), 2, 0, K) == -1
    assert O.koszul_sign((odd, even), (1, 0), (), 2, 0, K) == 1
    # moving a codomain leg to the domain alone never crosses (linearised order kept)
    assert O.koszul_sign((odd, odd, odd), (0, 1), (2,), 3, 0, K) == 1
    # cyclic shift of three odd legs = two transpositions
    assert O.koszul_sign((odd, odd, odd), (1, 2, 0), (), 3, 0, K) == 1
