"""CPU model of the LDS hash tables of the count and support kernels: every
insert and every probe chain ends, membership is exact, and a key's slot word
stays with it.

The kernel (trianglecount_amd/csrc/count_engine.hip: table_insert, table_mark,
probe_round, drain_queue) keeps a row's keys in 2^tl slots grouped into
4-slot buckets.  A key's home bucket is hquad(key) = (key * 0x9E3779B1 mod 2^32)
>> (34 - tl); an insert tries the slots of a bucket in order with
compare-and-swap and moves to the next bucket (wrapping) when all four are
taken, marking its home as overflowed.  table_mark then orders slots 0 and 1 of
every bucket by that mark (overflowed: slot0 > slot1).  A probe reads the home
bucket; only when it misses there and the home is marked overflowed does it
walk on, from the next bucket, until it hits or reads a bucket whose last slot
is empty.  The device loops are bounded by one lap of the buckets; the model
below fails if any insert or chain would need more than that, i.e. if the loop
bound could ever change a result.

Why chains end: a slot is written only after the slots before it in its bucket
were seen taken, and slots never empty again, so a bucket with its last slot
taken is full.  Tables are filled to at most half of their slots, so at most
half of the buckets are full and every lap passes a bucket with an empty last
slot.  The model checks that by simulation over random interleavings of the
individual compare-and-swap steps (any schedule of the wave's atomics is one of
them), at every table size the kernels use, loads up to 1/2, and adversarial
keys (all keys on one home at the end of the table, so chains wrap around).

The support kernels (trianglecount_amd/csrc/support.hip: sup_build, sup_probe)
use the same buckets, marks and probe walk with tables from 16 slots
(sup_tlog2) up to the workgroup's 2^12, and give every slot a word holding the
key's tile index; their marking swaps the words of slots 0 and 1 with the keys.
"""
import random

import numpy as np
import pytest

PHI = 0x9E3779B1
PHI_INV = pow(PHI, -1, 1 << 32)
EMPTY = 0xFFFFFFFF
NO_ID = 0xFFFFFFFE  # the id that past-the-end positions probe
SLICE_LOG2 = 10
BIG_LOG2 = {2: 11, 4: 12, 8: 13}  # workgroup table of 1024 slots per wave, by waves per workgroup
SUP_GROUP_LOG2 = 12  # support.hip kSupGroupLog2: the workgroup table; waves take 1/4 or 1/8 of it


def hquad(x, tl):
    return ((x * PHI) & 0xFFFFFFFF) >> (34 - tl)


def sup_tlog2(t):
    """support.hip sup_tlog2: smallest l >= 4 with 2^l >= 2 * t."""
    l = 4
    while (1 << l) < 2 * t:
        l += 1
    return l


def table_log2_for(keys, scale, max_log2):
    """count_engine.hip table_log2_for: smallest l >= 6 with 2^l >= scale * keys, at most max_log2."""
    l = 6
    while l < max_log2 and (1 << l) < scale * keys:
        l += 1
    return l


class Table:
    def __init__(self, tl):
        self.tl = tl
        self.nb = 1 << (tl - 2)
        self.qmask = self.nb - 1
        self.slot = [EMPTY] * (4 * self.nb)
        self.word = [None] * (4 * self.nb)  # the support kernels' slot word: the key's tile index
        self.ovf = [False] * self.nb

    def insert_all(self, keys, threads, rng):
        """Inserts with `threads` concurrent inserters whose single CAS steps interleave at random.

        Each inserter follows table_insert: for step in 0..qmask, try slots 0..3 of bucket h in
        order, then h+1.  Returns the largest number of buckets any insert visited."""
        pending = list(enumerate(keys))
        live = []  # [key, home, h, slot, buckets_visited, tile index]
        worst = 0
        while pending or live:
            while pending and len(live) < threads:
                idx, k = pending.pop()
                home = hquad(k, self.tl)
                live.append([k, home, home, 0, 1, idx])
            i = rng.randrange(len(live))
            st = live[i]
            k, home, h, s, visited, idx = st
            if self.slot[4 * h + s] == EMPTY:  # the CAS succeeds
                self.slot[4 * h + s] = k
                self.word[4 * h + s] = idx
                if h != home:
                    self.ovf[home] = True
                worst = max(worst, visited)
                live[i] = live[-1]
                live.pop()
                continue
            if s < 3:
                st[3] = s + 1
            else:
                assert visited <= self.qmask, "an insert found no free slot within one lap"
                st[2] = (h + 1) & self.qmask
                st[3] = 0
                st[4] = visited + 1
        return worst

    def mark(self, words=True):
        for b in range(self.nb):
            x, y = self.slot[4 * b], self.slot[4 * b + 1]
            if (x < y) if self.ovf[b] else (x > y):
                self.slot[4 * b], self.slot[4 * b + 1] = y, x
                if words:
                    self.word[4 * b], self.word[4 * b + 1] = self.word[4 * b + 1], self.word[4 * b]

    def probe(self, x):
        """(found, buckets read, slot) as probe_round + drain_queue (sup_probe) compute them."""
        home = hquad(x, self.tl)
        v = self.slot[4 * home: 4 * home + 4]
        if x in v:
            return True, 1, 4 * home + v.index(x)
        if not v[0] > v[1]:  # not marked overflowed: the chain ends at home
            return False, 1, None
        q = (home + 1) & self.qmask
        for step in range(self.qmask):  # drain_queue's lap bound
            w = self.slot[4 * q: 4 * q + 4]
            if x in w:
                return True, step + 2, 4 * q + w.index(x)
            if w[3] == EMPTY:
                return False, step + 2, None
            q = (q + 1) & self.qmask
        raise AssertionError(f"probe chain of {x:#x} did not end within one lap")


def keys_with_home(home, tl, count, rng):
    """`count` distinct ids whose home bucket is `home` (inverse of the Fibonacci hash)."""
    lo, width = home << (34 - tl), 1 << (34 - tl)
    ys = rng.sample(range(width), count)
    out = [((lo + y) * PHI_INV) & 0xFFFFFFFF for y in ys]
    return [k for k in out if k < NO_ID]


def key_sets(tl, n, rng):
    nb = 1 << (tl - 2)
    yield "random", rng.sample(range(NO_ID), n)
    yield "consecutive", list(range(1000, 1000 + n))
    yield "one home at the end", keys_with_home(nb - 1, tl, n, rng)
    hot = [keys_with_home(h, tl, 8, rng) for h in rng.sample(range(nb), max(1, nb // 8))]
    mix = [k for grp in hot for k in grp][: n // 2]
    mix += [k for k in rng.sample(range(NO_ID), n) if k not in set(mix)][: n - len(mix)]
    yield "clustered homes", mix


def check_table(tl, keys, threads, rng):
    t = Table(tl)
    worst_insert = t.insert_all(keys, threads, rng)
    assert worst_insert <= t.nb
    # a bucket with its last slot taken is full, and at most half of them are
    full = [all(t.slot[4 * b + i] != EMPTY for i in range(4)) for b in range(t.nb)]
    for b in range(t.nb):
        if t.slot[4 * b + 3] != EMPTY:
            assert full[b]
    assert sum(full) <= t.nb // 2
    t.mark()
    keyset = set(keys)
    for i, k in enumerate(keys):
        found, _, slot = t.probe(k)
        assert found and t.word[slot] == i
    absent = [NO_ID] + [x for x in rng.sample(range(NO_ID), 200) if x not in keyset]
    absent += [x for x in keys_with_home(hquad(keys[0], tl), tl, 16, rng) if x not in keyset]
    for x in absent:
        assert not t.probe(x)[0]
    return worst_insert


@pytest.mark.parametrize("tl", [4, 5, 6, 7, 8, 9, 10, 11, 12, 13])
def test_chains_end_and_membership_exact(tl):
    rng = random.Random(tl)
    slots = 1 << tl
    for load in (1 / 16, 1 / 6, 1 / 4, 1 / 2):
        n = max(2, int(slots * load))
        for name, keys in key_sets(tl, n, rng):
            for threads in (1, 64, 256):
                check_table(tl, keys, threads, rng)


def test_full_home_chain_wraps_around():
    """Half the slots, every key on the last bucket: the chain runs over the end of the table."""
    rng = random.Random(7)
    for tl in (4, 6, 10, 13):
        nb = 1 << (tl - 2)
        keys = keys_with_home(nb - 1, tl, 2 * nb, rng)
        worst = check_table(tl, keys, 64, rng)
        assert worst >= nb // 2  # the chain really wrapped and ran long


def test_kernel_loads_stay_at_or_below_half():
    """Every table the kernel builds is at most half full, for every row length and table_scale
    the tuning allows (resolve_tune clamps table_scale to 2..16 and small_max to 64..1024/scale;
    small rows take a 1024-slot slice, big rows tiles of half the workgroup's table of 1024
    slots per wave)."""
    for scale in range(2, 17):
        small_max = max(64, 1024 // scale)
        for k in range(2, small_max + 1):
            tl = table_log2_for(k, scale, SLICE_LOG2)
            assert 2 * k <= (1 << tl), (scale, k, tl)
        for waves, big in BIG_LOG2.items():
            for tk in range(1, (1 << big) // 2 + 1):  # big rows are tiled by half the table
                tl = table_log2_for(tk, scale, big)
                assert 2 * tk <= (1 << tl), (scale, waves, tk, tl)


def test_support_loads_stay_at_or_below_half():
    """Every support table is at most half full: tiles of up to half the table the kernel gives
    them (the workgroup's 2^12 slots, or a wave's quarter or eighth of it), at least 16 slots."""
    for table_log2 in (SUP_GROUP_LOG2, SUP_GROUP_LOG2 - 2, SUP_GROUP_LOG2 - 3):
        tile = (1 << table_log2) // 2
        for t in range(1, tile + 1):
            tl = sup_tlog2(t)
            assert 4 <= tl <= table_log2 and 2 * t <= (1 << tl), (table_log2, t, tl)


def test_hash_matches_kernel_arithmetic():
    """hquad is the high bits of the 32-bit product (numpy uint32 wrap-around as on the device)."""
    x = np.array([0, 1, 2, 12345, 0xFFFFFFFD, NO_ID], np.uint32)
    for tl in (4, 6, 10, 12):
        dev = ((x * np.uint32(PHI)).astype(np.uint32) >> np.uint32(34 - tl)).tolist()
        assert dev == [hquad(int(v), tl) for v in x]
        assert max(dev) < (1 << (tl - 2))


def test_model_catches_a_planted_mistake():
    """With the overflow marks left out (table_mark skipped), lookups that must walk past a full
    home stop there: the model has to report missed keys, so its membership check has teeth."""
    rng = random.Random(3)
    tl = 8
    nb = 1 << (tl - 2)
    keys = keys_with_home(nb - 1, tl, 24, rng)  # 24 keys on one home: 20 of them live beyond it
    t = Table(tl)
    t.insert_all(keys, 64, rng)
    t.slot[4 * (nb - 1)], t.slot[4 * (nb - 1) + 1] = sorted(t.slot[4 * (nb - 1): 4 * (nb - 1) + 2])
    missed = [k for k in keys if not t.probe(k)[0]]
    assert len(missed) == 20
    t.mark()
    assert all(t.probe(k)[0] for k in keys)


def test_model_catches_unswapped_words():
    """With the marks swapping keys but not their words, every bucket whose first two keys the
    marking reorders reports each other's tile index for them: the word check has to see it."""
    rng = random.Random(5)
    tl = 6
    keys = keys_with_home(1, tl, 12, rng)  # 12 keys on one home: it overflows into the next ones
    t = Table(tl)
    t.insert_all(keys, 64, rng)
    before = list(t.slot)
    t.mark(words=False)
    swaps = sum(t.slot[4 * b] != before[4 * b] for b in range(t.nb))
    wrong = [k for i, k in enumerate(keys) if t.word[t.probe(k)[2]] != i]
    assert swaps >= 1 and len(wrong) == 2 * swaps
