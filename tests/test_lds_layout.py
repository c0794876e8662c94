"""CPU model of the axis kernels' LDS exchange layout (axis_kernel.hpp):
slot of element i = i + (i >> K), K from lds_pad_shift (5 for rows whose
transform spans at least 32 threads, else 4; 4 for column tiles).

Checks, for every shape the library compiles: (1) the pad map is linear
over each access group, so padded(base + d) == padded(base) + d + (d >> K)
-- the kernels rely on it to fold d into the instruction's offset field;
(2) every slot stays inside the per-transform region of N + N/16 slots;
(3) in the bank model (64 four-byte banks, eight-byte accesses served 32
lanes a cycle) the chosen K is no worse than the other one."""
import pytest


def sched(N, E):
    rx = min(E, 16)
    ln, lr = N.bit_length() - 1, rx.bit_length() - 1
    nfull = ln // lr if lr else 0
    rem = 1 << (ln % lr if lr else ln)
    nst = nfull + (1 if rem > 1 else 0)
    return [rem if (rem > 1 and s == 0) else rx for s in range(nst)]


def pad_shift(N, E, cfast):
    return 5 if (not cfast and N // E >= 32) else 4


def accesses(N, E, TB, cfast, K):
    """Yield (instruction id, lane slots, linear ok) for every exchange."""
    pad = lambda i: i + (i >> K)
    mul = TB if cfast else 1
    off = lambda d: (d + (d >> K)) * mul
    P = N // E
    NT = TB * P
    stride = N + N // 16

    def index(t, i):
        return pad(i) * TB + t if cfast else t * stride + pad(i)

    def lane(tid):
        return (tid % TB, tid // TB) if cfast else (tid // P, tid % P)

    radix = sched(N, E)
    span = 1
    for S in range(len(radix) - 1):
        r, Ns, Q = radix[S], span, E // radix[S]
        for q in range(Q):
            for m in range(r):
                slots, ok = [], True
                for tid in range(NT):
                    t, p = lane(tid)
                    j = p + q * P
                    dst0 = (j // Ns) * Ns * r + (j % Ns)
                    w = index(t, dst0) + off(m * Ns)
                    ok &= w == index(t, dst0 + m * Ns)
                    slots.append(w)
                yield slots, ok
        for e in range(E):
            slots, ok = [], True
            for tid in range(NT):
                t, p = lane(tid)
                w = index(t, p) + off(P * e)
                ok &= w == index(t, p + P * e)
                slots.append(w)
            yield slots, ok
        span *= r


def conflict_cost(N, E, TB, cfast, K):
    cycles = ideal = 0
    for slots, _ in accesses(N, E, TB, cfast, K):
        for w0 in range(0, len(slots), 64):
            for h in range(2):
                grp = slots[w0 + 32 * h: w0 + 32 * h + 32]
                if not grp:
                    continue
                cnt = {}
                for x in grp:
                    cnt[x % 32] = cnt.get(x % 32, 0) + 1
                cycles += max(cnt.values())
                ideal += 1
    return cycles / max(ideal, 1)


# (N, E, TB, cfast): the axis variants (kernels_axis.hip), the batch kernel's
# classes (many_kernel.hpp: 256- and 1024-thread blocks) and the fused real rows
ROWS = [(n, 16, tb, False) for n, tb in [(2, 256), (4, 256), (8, 256), (16, 256), (32, 128), (64, 64), (128, 32),
                                         (256, 16), (512, 8), (1024, 4), (2048, 2), (4096, 1), (4096, 2), (256, 4),
                                         (256, 1), (64, 16), (128, 16)]]
MANY = [(1 << L, 16, max(1, nt // max(1, (1 << L) // 16)), False) for nt in (256, 1024) for L in range(4, 15)
        if max(1, (1 << L) // 16) <= nt] + \
       [(1 << L, 8, nt // ((1 << L) // 8), False) for nt in (256, 1024) for L in range(3, 14) if (1 << L) // 8 <= nt]
COLS = [(n, e, tb, True) for n, e, tb in [(2, 2, 64), (4, 4, 64), (8, 8, 64), (16, 16, 64), (32, 16, 32), (64, 16, 16),
                                          (128, 16, 16), (256, 16, 16), (512, 16, 8), (1024, 16, 4), (1024, 16, 8),
                                          (2048, 16, 2), (2048, 16, 4), (4096, 16, 1), (4096, 16, 2), (4096, 16, 4),
                                          (4096, 32, 2), (4096, 32, 4), (2048, 32, 4), (2048, 32, 8), (512, 16, 16),
                                          (512, 16, 32), (256, 16, 32), (256, 16, 64), (256, 16, 4), (256, 4, 1), (512, 8, 1),
                                          (1024, 16, 1), (2048, 16, 1)]]
SHAPES = sorted(set(ROWS + MANY + COLS))


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "n%d_e%d_tb%d_%s" % (s[0], s[1], s[2], "col" if s[3] else "row"))
def test_pad_map_linear_and_in_region(shape):
    N, E, TB, cfast = shape
    if N // E > 1024 or TB * (N // E) > 1024:
        pytest.skip("not a compiled shape")
    K = pad_shift(N, E, cfast)
    limit = TB * (N + N // 16)
    for slots, ok in accesses(N, E, TB, cfast, K):
        assert ok, (shape, K)
        assert 0 <= min(slots) and max(slots) < limit, (shape, K)


@pytest.mark.parametrize("shape", [s for s in SHAPES if not s[3] and s[0] >= 64 and s[0] // s[1] <= 1024],
                         ids=lambda s: "n%d_e%d_tb%d" % s[:3])
def test_row_pad_shift_choice(shape):
    N, E, TB, cfast = shape
    chosen = pad_shift(N, E, cfast)
    other = 9 - chosen  # 4 <-> 5
    assert conflict_cost(N, E, TB, cfast, chosen) <= conflict_cost(N, E, TB, cfast, other) + 1e-9, shape


def test_model_sees_the_conflicts_it_should():
    # a planted mistake: no padding at all makes the stride-16 scatter of a
    # 4096-point row hit 2 banks per half-wave
    assert conflict_cost(4096, 16, 1, False, 30) > 4
    assert conflict_cost(4096, 16, 1, False, 5) < conflict_cost(4096, 16, 1, False, 4)
