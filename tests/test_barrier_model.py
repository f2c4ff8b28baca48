"""Host model of the grid barrier's release protocol (bfs_device.hpp,
grid_barrier): workgroup b arrives on leaf b % 16; each waiter reads its
leaf's release word before arriving and spins until the word changes; the
last arriver publishes the next generation.

Launches with fewer than 16 workgroups advance only the low leaves' words,
so a later full-grid launch starts with words that disagree.  Publishing
"the last arriver's leaf value + 1" to every leaf can then hand a leaf that
is exactly one generation ahead its current value, and its waiters never
leave.  The kernels read the generation from one global word instead and
wait for their leaf word to equal generation + 1, which releases everyone
whatever the leaf words held before.
"""
import itertools

FAN_IN = 16


def run_barrier(state, nblocks, last_block, protocol):
    """One barrier over nblocks workgroups whose last arrival is last_block.
    state = {"release": [leaf words], "gen": global word}.  Returns the
    leaves whose waiters would never be released."""
    release = state["release"]
    leaves = min(nblocks, FAN_IN)
    if protocol == "global":
        gen = state["gen"]                                  # read before arriving
        state["gen"] = gen + 1
        for l in range(leaves):
            release[l] = gen + 1
        return [l for l in range(leaves) if release[l] != gen + 1]
    seen = {l: release[l] for l in range(leaves)}          # read before arriving
    gen_last = release[last_block % FAN_IN]
    for l in range(leaves):
        release[l] = release[l] + 1 if protocol == "per_leaf" else gen_last + 1
    return [l for l in range(leaves) if release[l] == seen[l]]


def launches(release, grids, last_of, protocol):
    state = {"release": release, "gen": 0}
    stuck = []
    for k, (nblocks, barriers) in enumerate(grids):
        for i in range(barriers):
            stuck += run_barrier(state, nblocks, last_of(k, i, nblocks), protocol)
    return stuck


def test_old_protocol_can_strand_a_leaf():
    # grid 3 with one barrier, grid 1 with one barrier, then a full grid
    # whose first barrier completes on leaf 1 (one generation behind leaf 0)
    release = [0] * FAN_IN
    grids = [(3, 1), (1, 1), (512, 1)]
    stuck = launches(release, grids, lambda k, i, n: 1 if n == 512 else 0, protocol="last_leaf")
    assert stuck == [0]


def test_fixed_protocols_always_release():
    # every order of small and large launches, every last-arriving leaf
    for protocol in ("per_leaf", "global"):
        for sizes in itertools.product([1, 3, 7, 16, 512], repeat=3):
            for last in range(FAN_IN):
                release = [0] * FAN_IN
                grids = [(n, b) for n, b in zip(sizes, (1, 2, 3))]
                stuck = launches(release, grids, lambda k, i, n: last % n, protocol=protocol)
                assert stuck == [], (protocol, sizes, last)


def others_arrived(root, leaf_counts, nblocks):
    """The test wait_other_arrivals (bfs_device.hpp) spins on."""
    leaves = min(nblocks, FAN_IN)
    leaf0 = (nblocks + FAN_IN - 1) // FAN_IN
    return root == leaves - 1 and leaf_counts[0] == leaf0 - 1


def test_wait_for_other_arrivals_predicate():
    # Arrivals as grid_barrier performs them: a workgroup increments its
    # leaf's counter; the leaf's last arriver resets it and increments the
    # root.  Block 0 has not arrived.  The predicate must hold exactly once
    # every other workgroup has arrived, in any arrival order.
    import random
    rng = random.Random(7)
    for nblocks in list(range(2, 40)) + [255, 256, 512, 768]:
        leaves = min(nblocks, FAN_IN)
        for trial in range(20):
            order = list(range(1, nblocks))
            rng.shuffle(order)
            leaf_counts = [0] * FAN_IN
            root = 0
            for i, b in enumerate(order):
                assert not others_arrived(root, leaf_counts, nblocks), (nblocks, i)
                leaf = b % FAN_IN
                size = (nblocks - leaf + FAN_IN - 1) // FAN_IN
                leaf_counts[leaf] += 1
                if leaf_counts[leaf] == size:
                    leaf_counts[leaf] = 0
                    root += 1
            assert others_arrived(root, leaf_counts, nblocks), nblocks
            assert root <= leaves - 1
