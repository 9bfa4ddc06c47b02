"""CPU: the plan caches (LRU, `src/auxiliary/caches.jl:2-19`) and plan pinning for
captured graphs (no device: plain objects stand in for plans)."""
import gc
import weakref

from tkmi import treetransformers as TT


class _Plan:
    pass


def test_lru_evicts_least_recently_used():
    c = TT._LRU(2)
    a, b, d = _Plan(), _Plan(), _Plan()
    c.put("a", a)
    c.put("b", b)
    assert c.get("a") is a          # a becomes the most recent
    c.put("d", d)                   # evicts b
    assert c.get("b") is None and c.get("a") is a and c.get("d") is d
    assert c.hits == 3 and c.misses == 1


def test_pinning_keeps_evicted_plans_alive():
    c = TT._LRU(1)
    p1 = _Plan()
    c.put("k1", p1)
    ref = weakref.ref(p1)
    del p1
    with TT.pinning() as outer:
        with TT.pinning() as inner:
            assert c.get("k1") is ref()
        c.put("k2", _Plan())        # evicts k1 while pinned
    gc.collect()
    assert ref() is not None        # still held by the pins
    assert ref() in inner and ref() in outer and len(outer) == 2
    del outer, inner
    gc.collect()
    assert ref() is None            # released with the pins
    assert c.get("k2") is not None  # outside a scope nothing is pinned
    assert TT._PIN_STACK == []


def test_empty_globalcaches_clears_every_plan_cache():
    from tkmi import fusedcontract as FC
    from tkmi import tensors as T
    for cache in (TT._TRANSFORMER_CACHE, T._GEMM_CACHE, FC._FUSED_CACHE):
        cache.put(("probe",), _Plan())
    TT.empty_globalcaches()
    for cache in (TT._TRANSFORMER_CACHE, T._GEMM_CACHE, FC._FUSED_CACHE):
        assert cache.get(("probe",)) is None


def test_space_structure_is_shared_and_cleared():
    """Equal spaces share their block structure (the reference's global caches keyed by
    the space); scalar spaces of different sector types do not; empty_globalcaches drops
    it, and the structure recomputed afterwards is the same."""
    from tkmi import treetransformers as TT
    from tkmi.sectors import U1Irrep, SU2Irrep
    from tkmi.spaces import HomSpace, ProductSpace, U1Space
    V = U1Space((0, 2), (1, 1), (-1, 3))
    W1 = HomSpace(ProductSpace(V, V), ProductSpace(V))
    ds = W1.degeneracystructure()
    W2 = HomSpace(ProductSpace(U1Space((0, 2), (1, 1), (-1, 3)), V), ProductSpace(V))
    assert W2 == W1 and W2._cache is W1._cache and "ds" in W2._cache
    s1 = HomSpace(ProductSpace(), ProductSpace(), unit=U1Irrep(0))
    s2 = HomSpace(ProductSpace(), ProductSpace(), unit=SU2Irrep(0))
    assert s1._cache is not s2._cache
    TT.empty_globalcaches()
    W3 = HomSpace(ProductSpace(V, V), ProductSpace(V))
    assert W3._cache is not W1._cache and not W3._cache
    assert W3.degeneracystructure() == ds


def test_factorization_caches_are_bounded():
    """svd / eigh plans own streams, handles and workspace: their caches are LRUs that
    evict beyond a fixed size and that empty_globalcaches() empties."""
    from tkmi import factorizations as F
    for cache in (F._SVD_CACHE, F._EIGH_CACHE):
        assert isinstance(cache, TT._LRU) and cache in TT._ALL_CACHES
        assert cache.maxsize == F.FACTORIZATION_CACHE_SIZE
        for i in range(cache.maxsize + 5):
            cache.put(("space", i), _Plan())
        assert len(cache.d) == cache.maxsize
        TT.empty_globalcaches()
        assert len(cache.d) == 0
