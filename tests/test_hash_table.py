"""Hash table: the oracle pinned on CPU; the GPU kernels against the oracle (bulk brick table,
bit-exact slots) and against the reference test's semantics (tests/hash_table_test.cpp)."""
from __future__ import annotations

import numpy as np
import pytest

HASH_NONE = 0xFFFFFFFF
K_HIT, K_INSERTED, K_PENDING, K_FULL, K_NO_PAYLOAD, K_EMPTY, K_MISSING = range(7)


def random_words(rng, count, word_count):
    return rng.integers(0, 2**32, size=(count, word_count), dtype=np.uint64).astype(np.uint32)


def brick_keys(n, seed, extent=512):
    rng = np.random.default_rng(seed)
    flat = rng.choice(extent**3, size=n, replace=False)
    return np.stack([flat % extent, (flat // extent) % extent, flat // extent**2], axis=1).astype(np.int32) - extent // 2


# ---------------------------------------------------------------------------------------- CPU


def test_splitmix64_published_vector(oracle):
    # Vigna's splitmix64 from state 0: the first output is 0xE220A8397B1DCDAF.
    assert oracle.splitmix64(0) == 0xE220A8397B1DCDAF


def test_fingerprint_forced_words_and_pairs(oracle):
    # All-zero words: h = splitmix64(0); key_lo/hi are its halves (both non-zero here).
    for count in (1, 2):
        fp = oracle.hash_fingerprint(np.zeros((1, count), np.uint32))[0]
        assert fp[0] == 0x7B1DCDAF and fp[1] == 0xE220A839
    # Three words = two rounds: splitmix64(splitmix64(w0 | w1 << 32) ^ w2)
    w = np.array([[1, 2, 3]], np.uint32)
    h = oracle.splitmix64((2 << 32) | 1)
    h = oracle.splitmix64(h ^ 3)
    fp = oracle.hash_fingerprint(w)[0]
    assert fp[0] == max(h & 0xFFFFFFFF, 1) and fp[1] == ((h >> 32) or 1)
    rng = np.random.default_rng(11)
    for count in range(1, 9):
        fps = oracle.hash_fingerprint(random_words(rng, 500, count))
        assert np.all(fps[:, 0] >= 1) and np.all(fps[:, 1] != 0)


@pytest.mark.parametrize("n,load", [(1, 0.5), (100, 0.7), (5000, 0.7), (5000, 0.95), (4096, None)])
def test_oracle_bulk_layout_is_the_ordered_linear_probing_table(oracle, n, load):
    """The layout of a sequential linear-probing insert in (home, key index) order, characterised
    without the oracle's code: (1) no empty slot between a key's home and its slot; (2) keys that did
    not wrap (slot >= home) appear in slot order with (home, index) increasing; (3) the wrapped keys
    (slot < home), in (home, index) order, occupy exactly the first free slots that the non-wrapped
    keys leave from slot 0."""
    keys = brick_keys(n, 7 + n)
    slot_count = 16384 if load is None else int(np.ceil(n / load)) + (1 if load >= 0.95 else 0)
    slots, longest = oracle.hash_build_bricks(keys, slot_count)
    fps = oracle.hash_fingerprint(keys.view(np.uint32))
    pow2 = slot_count & (slot_count - 1) == 0
    homes = (fps[:, 0] & (slot_count - 1)) if pow2 else (fps[:, 0] % slot_count)
    occupied = slots[:, 0] != 0
    assert occupied.sum() == n
    position = np.full(n, -1, np.int64)
    position[slots[occupied, 2]] = np.nonzero(occupied)[0]
    assert np.all(position >= 0)
    assert np.array_equal(slots[position, 0], fps[:, 0]) and np.array_equal(slots[position, 1], fps[:, 1])
    disp = (position - homes.astype(np.int64)) % slot_count
    assert disp.max(initial=0) == longest
    for k in range(n):  # (1)
        idx = (homes[k] + np.arange(disp[k] + 1)) % slot_count
        assert occupied[idx].all()
    wrapped = position < homes
    straight = np.nonzero(~wrapped)[0]
    by_slot = straight[np.argsort(position[straight])]
    pairs = list(zip(homes[by_slot].tolist(), by_slot.tolist()))
    assert pairs == sorted(pairs)  # (2)
    wrapped_keys = np.nonzero(wrapped)[0]
    wrapped_keys = wrapped_keys[np.lexsort((wrapped_keys, homes[wrapped_keys]))]
    free = np.ones(slot_count, bool)
    free[position[straight]] = False
    first_free = np.nonzero(free)[0][: len(wrapped_keys)]
    assert np.array_equal(position[wrapped_keys], first_free)  # (3)
    # lookups: all hit within longest + 1 probes, absent keys miss
    for k in range(0, n, max(1, n // 50)):
        payload, probes = oracle.hash_get_brick(slots, longest + 1, *map(int, keys[k]))
        assert payload == k and probes == disp[k] + 1
    payload, _ = oracle.hash_get_brick(slots, longest + 1, 10**6, 10**6, 10**6)
    assert payload == HASH_NONE


# ---------------------------------------------------------------------------------------- GPU


@pytest.mark.gpu
def test_gpu_fingerprint_matches_oracle(cx, oracle):
    import torch

    rng = np.random.default_rng(11)
    for count in range(1, 9):
        words = random_words(rng, 4096, count)
        words[0] = 0
        out = torch.empty(4096 * 2, dtype=torch.int32, device="cuda")
        w = torch.from_numpy(words.view(np.int32)).cuda()
        cx.check(cx.load().cx_hash_fingerprint(w.data_ptr(), count, 4096, out.data_ptr(), cx.stream_handle()), "fp")
        got = out.cpu().numpy().view(np.uint32).reshape(-1, 2)
        assert np.array_equal(got, oracle.hash_fingerprint(words)), count


@pytest.mark.gpu
@pytest.mark.parametrize("n,load", [(1, 0.5), (1000, 0.7), (100000, 0.7), (70000, 0.97), (65536, None)])
def test_gpu_bulk_table_matches_oracle(cx, oracle, n, load):
    import torch

    print(f"bulk brick table n={n} load={load}", flush=True)
    keys = brick_keys(n, 100 + n)
    slot_count = 1 << 18 if load is None else int(np.ceil(n / load)) + 1
    table = cx.BrickHashTable(torch.from_numpy(keys).cuda(), slot_count=slot_count)
    torch.cuda.synchronize()
    got = table.slots.cpu().numpy().view(np.uint32).reshape(-1, 4)
    expected, longest = oracle.hash_build_bricks(keys, slot_count)
    bad = np.nonzero(np.any(got[:, :3] != expected[:, :3], axis=1))[0]
    assert bad.size == 0, f"{bad.size} slots differ, first {bad[:5]}"
    info = table.info.cpu().numpy()
    assert info[0] == longest + 1 and info[1] == longest
    payloads, probes = table.get(torch.from_numpy(keys).cuda(), with_probes=True)
    assert np.array_equal(payloads.cpu().numpy(), np.arange(n))
    absent = brick_keys(2000, 5, extent=600) + 1000
    p2, _ = table.get(torch.from_numpy(absent).cuda())
    assert np.all(p2.cpu().numpy().view(np.uint32) == HASH_NONE)


@pytest.mark.gpu
def test_gpu_bulk_table_config_size_load_07(cx, oracle):
    """1,048,576 brick keys at load 0.7 (1,497,966 slots)."""
    import torch

    n = 1 << 20
    keys = brick_keys(n, 2, extent=512)
    slot_count = int(np.ceil(n / 0.7))
    assert slot_count == 1497966
    table = cx.BrickHashTable(torch.from_numpy(keys).cuda(), slot_count=slot_count)
    torch.cuda.synchronize()
    got = table.slots.cpu().numpy().view(np.uint32).reshape(-1, 4)
    expected, longest = oracle.hash_build_bricks(keys, slot_count)
    assert np.array_equal(got[:, :3], expected[:, :3])
    payloads, probes = table.get(torch.from_numpy(keys).cuda(), with_probes=True)
    assert np.array_equal(payloads.cpu().numpy(), np.arange(n))
    mean_probes = probes.cpu().numpy().mean()
    # Knuth: expected successful probes for linear probing at load a: (1 + 1 / (1 - a)) / 2 = 2.17
    assert 1.9 < mean_probes < 2.5, mean_probes


@pytest.mark.gpu
def test_gpu_one_key_from_65536_threads_inserts_once(cx):
    import torch

    t = cx.DynamicHashTable(1024, 3)
    first = t.maintain(2)
    assert first["live"] == 0 and t.top() == 1024
    words = torch.from_numpy(np.tile(np.array([[7, 8, 9]], np.int32), (65536, 1))).cuda()
    results, payloads = t.lookup(words, 2, insert=True)
    torch.cuda.synchronize()
    r = results.cpu().numpy()
    assert np.all(np.isin(r, [K_INSERTED, K_PENDING, K_HIT]))
    assert (r == K_INSERTED).sum() == 1
    c = t.counters_host()
    assert c["inserted"] == 1 and c["pending"] == int((r == K_PENDING).sum())
    assert t.top() == 1023
    payload = int(payloads.cpu().numpy()[r == K_INSERTED][0])
    again, p2 = t.lookup(words, 2, insert=False)
    assert np.all(again.cpu().numpy() == K_HIT) and np.all(p2.cpu().numpy() == payload)
    m = t.maintain(3)
    assert m["live"] == 1 and m["active"] == 1
    r3, p3 = t.lookup(words[:1], 3, insert=False)
    assert int(r3.cpu()[0]) == K_HIT and int(p3.cpu()[0]) == payload


@pytest.mark.gpu
def test_gpu_forty_same_home_keys_fill_the_window(cx, oracle):
    import torch

    t = cx.DynamicHashTable(1024, 2)
    t.maintain(2)
    mask = t.slot_count - 1
    rng = np.random.default_rng(31)
    keys = []
    home = None
    while len(keys) < 40:
        w = random_words(rng, 1, 2)
        h = int(oracle.hash_fingerprint(w)[0, 0]) & mask
        home = h if home is None else home
        if h == home:
            keys.append(w[0])
    keys = np.array(keys, np.uint32)
    inputs = np.tile(keys, (64, 1))
    results, _ = t.lookup(torch.from_numpy(inputs.view(np.int32)).cuda(), 2, insert=True)
    torch.cuda.synchronize()
    c = t.counters_host()
    assert c["inserted"] == 8
    r = results.cpu().numpy()
    assert (r == K_INSERTED).sum() == 8
    assert np.all(np.isin(r, [K_INSERTED, K_PENDING, K_HIT, K_FULL]))


@pytest.mark.gpu
def test_gpu_free_stack_exhaustion_and_eviction(cx):
    import torch

    t = cx.DynamicHashTable(256, 3)
    t.maintain(2)
    rng = np.random.default_rng(41)
    words = random_words(rng, 600, 3)
    results, payloads = t.lookup(torch.from_numpy(words.view(np.int32)).cuda(), 2, insert=True)
    torch.cuda.synchronize()
    c = t.counters_host()
    r = results.cpu().numpy()
    assert c["inserted"] == 256
    assert c["failed_payload"] == int((r == K_NO_PAYLOAD).sum()) > 0
    # keys without a payload are empty entries on a second lookup in the same frame
    again, _ = t.lookup(torch.from_numpy(words.view(np.int32)).cuda(), 2, insert=False)
    a = again.cpu().numpy()
    assert np.all(a[r == K_NO_PAYLOAD] == K_EMPTY)
    # Not touched for max_age frames: evicted and the stack is full again.
    m = t.maintain(2 + 10, max_age=5)
    assert m["evicted"] == 256 and m["live"] == 0
    assert t.top() == 256


# --------------------------------------- the reference test's table checks (hash_table_test.cpp:318-470)


def table_keys(t) -> dict:
    """(key_lo, key_hi) -> the slots holding it."""
    slots = t.keys.cpu().numpy().view(np.uint32).reshape(-1, 4)
    out: dict = {}
    for i in np.nonzero(slots[:, 0])[0]:
        out.setdefault((int(slots[i, 0]), int(slots[i, 1])), []).append(int(i))
    return out


def payload_headers(t) -> np.ndarray:
    return t.payloads.cpu().numpy().view(np.uint32).reshape(t.capacity, -1)


def check_inserts(oracle, t, keys, key_of, results, payloads) -> dict:
    """CheckInserts: per key at most one insert, every hit and insert on one payload, every key once in
    its window, no payload twice, every payload header holding its key. Returns key index -> payload."""
    r = results.cpu().numpy().view(np.uint32)
    p = payloads.cpu().numpy().view(np.uint32)
    assert np.all(r <= K_EMPTY)
    payload_of: dict = {}
    inserted = np.zeros(len(keys), np.int64)
    for i in np.nonzero((r == K_INSERTED) | (r == K_HIT))[0]:
        assert p[i] < t.capacity
        assert payload_of.setdefault(int(key_of[i]), int(p[i])) == int(p[i])
    np.add.at(inserted, key_of[r == K_INSERTED], 1)
    assert np.all(inserted <= 1)
    slots = t.keys.cpu().numpy().view(np.uint32).reshape(-1, 4)
    headers = payload_headers(t)
    fps = oracle.hash_fingerprint(keys)
    key_index = {(int(a), int(b)): k for k, (a, b) in enumerate(fps)}
    seen_payloads = set()
    for key, where in table_keys(t).items():
        assert len(where) == 1 and key[1] != 0
        assert ((where[0] - key[0]) & (t.slot_count - 1)) < 8
        payload = int(slots[where[0], 2])
        if payload == HASH_NONE:
            continue
        assert payload not in seen_payloads
        seen_payloads.add(payload)
        assert headers[payload, 0] in (1, 2)
        fp = oracle.hash_fingerprint(headers[payload, 2:2 + t.key_words][None, :])[0]
        assert (int(fp[0]), int(fp[1])) == key
        k = key_index.get(key)
        if k is not None and k in payload_of:
            assert payload_of[k] == payload
    table = table_keys(t)
    for k, payload in payload_of.items():
        where = table[(int(fps[k, 0]), int(fps[k, 1]))]
        assert int(slots[where[0], 2]) == payload
    return payload_of


def check_partition(t) -> None:
    """CheckPartition: the free stack's entries and the payloads that are not free partition the pool."""
    stack = t.stack.cpu().numpy().view(np.int32)
    top = int(stack[0])
    assert top >= 0
    headers = payload_headers(t)
    seen = np.zeros(t.capacity, np.int64)
    entries = stack[1:1 + top].view(np.uint32)
    assert np.all(entries < t.capacity)
    assert np.all(headers[entries, 0] == 0)
    np.add.at(seen, entries, 1)
    seen += headers[:, 0] != 0
    assert np.all(seen == 1)


def check_rebuilt(oracle, t, frame, counters) -> None:
    """CheckRebuilt: after a maintenance every live payload's key is in the table once, on that payload;
    the active list holds exactly the payloads touched at frame - 1; live and active count them."""
    table = table_keys(t)
    slots = t.keys.cpu().numpy().view(np.uint32).reshape(-1, 4)
    headers = payload_headers(t)
    active = t.active.cpu().numpy().view(np.uint32)[: counters["active"]]
    assert len(set(active.tolist())) == len(active)
    live_ids = np.nonzero(headers[:, 0] != 0)[0]
    if len(live_ids):
        fps = oracle.hash_fingerprint(np.ascontiguousarray(headers[live_ids, 2:2 + t.key_words]))
        for p, fp in zip(live_ids, fps):
            where = table[(int(fp[0]), int(fp[1]))]
            assert len(where) == 1 and int(slots[where[0], 2]) == p
    is_active = headers[live_ids, 1] == frame - 1
    assert set(active.tolist()) == set(live_ids[is_active].tolist())
    assert np.all(headers[active, 0] != 0)
    assert len(table) == len(live_ids) == counters["live"]
    assert counters["active"] == int(is_active.sum())


@pytest.mark.gpu
def test_gpu_equal_key_lo_pair(cx, oracle):
    """Two keys with the same key_lo (so the same home) and different key_hi, found by a birthday search:
    one dispatch inserts at least one; a second dispatch of the same frame finds it published and inserts
    the other in the next slot; both survive the maintenance (hash_table_test.cpp:655-704)."""
    import torch

    rng = np.random.default_rng(31)
    words = random_words(rng, 400000, 2)
    fps = oracle.hash_fingerprint(words)
    order = np.argsort(fps[:, 0], kind="stable")
    lo = fps[order, 0]
    dup = np.nonzero((lo[1:] == lo[:-1]) & (fps[order[1:], 1] != fps[order[:-1], 1]))[0]
    assert len(dup) > 0
    a, b = order[dup[0]], order[dup[0] + 1]
    keys = np.stack([words[a], words[b]])
    t = cx.DynamicHashTable(1024, 2)
    t.maintain(2)
    key_of = np.arange(1024) & 1
    inputs = torch.from_numpy(keys[key_of].view(np.int32)).cuda()
    first = check_inserts(oracle, t, keys, key_of, *t.lookup(inputs, 2, insert=True))
    assert len(first) >= 1
    second = check_inserts(oracle, t, keys, key_of, *t.lookup(inputs, 2, insert=True))
    assert len(second) == 2 and second[0] != second[1]
    m = t.maintain(3)
    check_rebuilt(oracle, t, 3, m)
    assert m["live"] == 2
    look, payloads = t.lookup(torch.from_numpy(keys.view(np.int32)).cuda(), 3, insert=False)
    assert np.all(look.cpu().numpy() == K_HIT)
    assert payloads.cpu().numpy().tolist() == [second[0], second[1]]
    check_partition(t)


@pytest.mark.gpu
def test_gpu_random_frames_keep_the_pool_partitioned(cx, oracle):
    """40 frames of random key subsets from a universe of 6000 five-word keys, up to 12 threads per key
    (some frames overfill the 2048-payload pool), each followed by a maintenance (max age 6): every
    dispatch passes CheckInserts, every maintenance CheckRebuilt and CheckPartition
    (hash_table_test.cpp:806-855)."""
    import torch

    rng = np.random.default_rng(51)
    t = cx.DynamicHashTable(2048, 5)
    universe = random_words(rng, 6000, 5)
    frame = 2
    m = t.maintain(frame, max_age=6)
    for round_ in range(40):
        distinct = 100 + int(rng.integers(0, 3000))
        chosen = rng.choice(len(universe), size=distinct, replace=False)
        keys = universe[chosen]
        threads = distinct * (1 + int(rng.integers(0, 12)))
        key_of = rng.integers(0, distinct, size=threads)
        inputs = torch.from_numpy(np.ascontiguousarray(keys[key_of]).view(np.int32)).cuda()
        check_inserts(oracle, t, keys, key_of, *t.lookup(inputs, frame, insert=True))
        frame += 1
        m = t.maintain(frame, max_age=6)
        check_rebuilt(oracle, t, frame, m)
        check_partition(t)
    assert m["live"] > 0
