"""The oracle's NTT / iNTT / coset LDE / transpose_rev, pinned by the
mathematics: the O(N^2) DFT, direct polynomial evaluation (Horner, pure
Python), round trips, delta and constant inputs, and the brute-force
transpose definition (wrappers/rust/tests/utils.rs:15-27)."""
import numpy as np
import pytest

import oracle as O

P = O.P


def horner(coeffs, x):
    acc = 0
    for c in reversed([int(v) for v in coeffs]):
        acc = (acc * x + c) % P
    return acc


@pytest.mark.parametrize("lg", [1, 2, 3, 5, 8])
def test_ntt_forward_is_evaluation_at_powers_of_omega(lg):
    n = 1 << lg
    x = O.random(n, 30 + lg)
    out = O.ntt(x, lg, 1)
    w = pow(7, (P - 1) >> lg, P)
    assert [int(v) for v in out] == [horner(x, pow(w, i, P)) for i in range(n)]


@pytest.mark.parametrize("lg", [1, 4, 7, 10])
@pytest.mark.parametrize("inverse", [False, True])
def test_ntt_matches_naive_dft(lg, inverse):
    x = O.random(3 << lg, 40 + lg)
    out = O.ntt(x, lg, 3, inverse=inverse)
    n = 1 << lg
    ninv = pow(n, P - 2, P)
    for b in range(3):
        ref = O.dft_naive(x[b * n:(b + 1) * n], lg, inverse=inverse)
        if inverse:
            ref = np.array([(int(v) * ninv) % P for v in ref], dtype=np.uint64)
        assert (out[b * n:(b + 1) * n] == ref).all()


@pytest.mark.parametrize("lg", [1, 6, 12, 16])
def test_round_trip(lg):
    x = O.random(2 << lg, 50 + lg)
    assert (O.ntt(O.ntt(x, lg, 2), lg, 2, inverse=True) == x).all()


def test_delta_and_constant():
    lg, n = 9, 512
    delta = np.zeros(n, dtype=np.uint64)
    delta[0] = 5
    assert (O.ntt(delta, lg, 1) == 5).all()
    const = np.full(n, 3, dtype=np.uint64)
    out = O.ntt(const, lg, 1)
    assert out[0] == (3 * n) % P and (out[1:] == 0).all()


def test_coset_forward_and_inverse_quirk():
    lg, n, g = 6, 64, 7
    x = O.random(n, 60)
    scaled = np.array([(int(v) * pow(g, i, P)) % P for i, v in enumerate(x)], dtype=np.uint64)
    assert (O.ntt(x, lg, 1, with_coset=True, coset_gen=g) == O.ntt(scaled, lg, 1)).all()
    # inverse with coset multiplies the iNTT output by g^i (native/ntt/ntt.cu:137-139)
    plain = O.ntt(x, lg, 1, inverse=True)
    quirk = O.ntt(x, lg, 1, inverse=True, with_coset=True, coset_gen=g)
    assert [int(v) for v in quirk] == [(int(v) * pow(g, i, P)) % P for i, v in enumerate(plain)]


@pytest.mark.parametrize("lg_n,rate_bits", [(1, 1), (3, 3), (5, 2), (7, 1)])
def test_lde_is_evaluation_on_the_coset(lg_n, rate_bits):
    n, N = 1 << lg_n, 1 << (lg_n + rate_bits)
    batches = 2
    c = O.random(batches * n, 70 + lg_n)
    out = O.lde(c, lg_n, rate_bits, batches)
    w = pow(7, (P - 1) >> (lg_n + rate_bits), P)
    for b in range(batches):
        col = c[b * n:(b + 1) * n]
        expect = [horner(col, 7 * pow(w, i, P) % P) for i in range(N)]
        assert [int(v) for v in out[b * N:(b + 1) * N]] == expect


def test_lde_matches_dft_of_padded_shifted_input():
    lg_n, r = 10, 3
    c = O.random(1 << lg_n, 80)
    pad = np.zeros(1 << (lg_n + r), dtype=np.uint64)
    pad[: 1 << lg_n] = c
    assert (O.lde(c, lg_n, r, 1) == O.dft_naive(pad, lg_n + r, shift=7)).all()


def test_inverse_lde_definition():
    lg_n, r = 4, 2
    c = O.random(1 << lg_n, 81)
    out = O.lde(c, lg_n, r, 1, inverse=True, with_coset=True)
    pad = np.zeros(1 << (lg_n + r), dtype=np.uint64)
    pad[: 1 << lg_n] = c
    assert (out == O.ntt(pad, lg_n + r, 1, inverse=True, with_coset=True)).all()


@pytest.mark.parametrize("lg_n,batches", [(0, 3), (1, 1), (5, 7), (8, 33)])
def test_transpose_rev_definition(lg_n, batches):
    n = 1 << lg_n
    x = O.random(n * batches, 90 + lg_n)
    out = O.transpose_rev(x, lg_n, batches).reshape(n, batches)
    m = x.reshape(batches, n)
    for i in range(n):
        r = int(format(i, "0%db" % lg_n)[::-1], 2) if lg_n else 0
        assert (out[r] == m[:, i]).all()


def test_poly_eval_matches_python_horner():
    rng = np.random.default_rng(3)
    for n in (0, 1, 2, 17, 300):
        c = O.random(n, 40 + n) if n else np.zeros(0, dtype=np.uint64)
        for x in [0, 1, P - 1, 7] + [int(v) for v in rng.integers(0, 2**63, 3)]:
            assert O.poly_eval(c, x) == horner(c, x), (n, x)
