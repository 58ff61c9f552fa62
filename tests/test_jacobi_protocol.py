"""Model check (CPU) of the synchronisation protocol of the blocked Jacobi
kernel (jacobi_body in oqupy_amd/csrc/pte_kernels.hip): nb column blocks,
nb/2 pair slots of R workgroups (row slices) each, a round-robin tournament
of block pairs, per-block versions that release a block to its next pair,
a per-pair arrival counter (pair_done) whose last arrival of a round releases
the pair's blocks, a per-pair barrier, a grid barrier closing each sweep, and
the skipping of pairs found converged whose blocks have not been rotated
since. Every workgroup is a generator that yields at each wait; a seeded
random scheduler interleaves them under sequentially consistent memory. The
test asserts that no interleaving deadlocks and that the slices of a pair
always agree on skipping; the planted variant without the pair barrier in
skipped rounds (a fast slice then finishes the next round before a slow one
finishes this round, and the release count of this round is never met)
deadlocks."""

import random

import pytest


def _tournament_pair(nplayers, rnd, slot):
    if slot == 0:
        return nplayers - 1, rnd % (nplayers - 1)
    return ((rnd + slot) % (nplayers - 1),
            (rnd - slot + nplayers - 1) % (nplayers - 1))


def _simulate(nb, slices, sweeps, seed, need_prob, barrier_when_skipped=True):
    rng = random.Random(seed)
    npairs = nb // 2
    pair_bar, pair_done = [0] * npairs, [0] * npairs
    blockver, blk_mod = [0] * nb, [0] * nb
    pair_ok, need_of, decisions = {}, {}, {}
    grid = [0]

    def workgroup(pair, part):
        gen = pair_epoch = epoch = 0
        for sweep in range(sweeps):
            for rnd in range(nb - 1):
                bi, bj = _tournament_pair(nb, rnd, pair)
                key = (min(bi, bj), max(bi, bj))
                while blockver[bi] < gen or blockver[bj] < gen:
                    yield "wait"
                ok = pair_ok.get(key, 0)
                skip = ok != 0 and blk_mod[bi] < ok and blk_mod[bj] < ok
                assert decisions.setdefault((pair, gen), skip) == skip, \
                    "slices of a pair disagree on skipping"
                yield "step"
                if slices > 1 and (barrier_when_skipped or not skip):
                    pair_epoch += 1
                    pair_bar[pair] += 1
                    while pair_bar[pair] < pair_epoch * slices:
                        yield "wait"
                if not skip:
                    need = need_of.setdefault((pair, gen),
                                              rng.random() < need_prob[sweep])
                    if part == 0:
                        if need:
                            blk_mod[bi] = blk_mod[bj] = gen + 1
                            pair_ok[key] = 0
                        else:
                            pair_ok[key] = gen + 1
                    yield "step"
                old = pair_done[pair]
                pair_done[pair] += 1
                if old == (gen + 1) * slices - 1:
                    blockver[bi] = blockver[bj] = gen + 1
                gen += 1
                yield "step"
            epoch += 1
            grid[0] += 1
            while grid[0] < epoch * npairs * slices:
                yield "wait"

    alive = {(p, s): workgroup(p, s) for p in range(npairs)
             for s in range(slices)}
    idle = 0
    while alive:
        key = rng.choice(list(alive))
        try:
            state = next(alive[key])
        except StopIteration:
            del alive[key]
            continue
        idle = idle + 1 if state == "wait" else 0
        if idle > 50000:
            return False
    return True


NEED = [0.9, 0.6, 0.3, 0.1, 0.0, 0.0]


@pytest.mark.parametrize("nb,slices", [(4, 1), (6, 2), (18, 5), (22, 6)])
def test_protocol_never_deadlocks(nb, slices):
    for seed in range(12):
        assert _simulate(nb, slices, 6, seed, NEED), (nb, slices, seed)


def test_protocol_without_barrier_in_skipped_rounds_deadlocks():
    assert not all(_simulate(6, 2, 6, seed, NEED, barrier_when_skipped=False)
                   for seed in range(12))
