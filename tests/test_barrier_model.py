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


# ---------------------------------------------------------------------------
# Weak-order model.  On gfx950 relaxed stores of one thread to different
# addresses (different L2 channels) may become visible in any order, and a
# relaxed atomic RMW is not ordered after the thread's earlier relaxed
# stores; only a release fence drains them.  The model gives every thread a
# buffer of pending stores that reach memory one at a time in any order
# (same-address order is kept; a thread reads its own pending stores), and
# explores every interleaving of grid_barrier's thread-0 protocol over a few
# consecutive barriers.  It checks that no workgroup leaves barrier k before
# every workgroup has arrived at it, and that the grid never hangs.
# Acquire fences are no-ops here (loads read memory directly), which only
# allows more behaviours.
# ---------------------------------------------------------------------------
REGS = ("gen", "arr", "r")


def barrier_program(b, nblocks, fan_in, nbar, fence_after_reset, fence_before_release):
    """Thread 0 of workgroup b through nbar consecutive grid_barrier calls."""
    leaf = b % fan_in
    leaves = min(nblocks, fan_in)
    leaf_size = (nblocks - leaf + fan_in - 1) // fan_in
    prog, labels = [], {}
    for k in range(nbar):
        wait, done = ("wait", k), ("done", k)
        prog += [("load", "gen", ("gen",)),
                 ("fence",),                                    # release before arriving
                 ("rmw", "arr", ("lc", leaf), k),               # arrival at barrier k
                 ("jne", "arr", leaf_size - 1, wait),
                 ("fence",),                                    # acq_rel of the leaf's last arriver
                 ("store", ("lc", leaf), None, 0)]
        if fence_after_reset:
            prog.append(("fence",))
        prog += [("rmw", "r", ("root",), None),
                 ("jne", "r", leaves - 1, wait),
                 ("fence",),                                    # acq_rel of the last arriver
                 ("store", ("root",), None, 0),
                 ("store", ("gen",), "gen", 1)]
        if fence_before_release:
            prog.append(("fence",))
        prog += [("store", ("rel", l), "gen", 1) for l in range(leaves)]
        prog.append(("jmp", done))
        labels[wait] = len(prog)
        prog.append(("spin", ("rel", leaf), "gen", 1))
        labels[done] = len(prog)
        prog.append(("pass", k))
    prog.append(("end",))
    out = []
    for ins in prog:
        if ins[0] == "jne":
            ins = ins[:3] + (labels[ins[3]],)
        elif ins[0] == "jmp":
            ins = ("jmp", labels[ins[1]])
        out.append(ins)
    return out


def check_weak_barrier(nblocks, fan_in, nbar, fence_after_reset=True, fence_before_release=True):
    """('ok', states) or the first violation found: ('early pass', ...) / ('hang', ...)."""
    addrs = [("gen",), ("root",)] + [("lc", l) for l in range(fan_in)] + [("rel", l) for l in range(fan_in)]
    ai = {a: i for i, a in enumerate(addrs)}
    progs = [barrier_program(b, nblocks, fan_in, nbar, fence_after_reset, fence_before_release)
             for b in range(nblocks)]
    init = (tuple(0 for _ in addrs), tuple((0, (0, 0, 0), ()) for _ in range(nblocks)), (0,) * nblocks)
    seen, stack = {init}, [init]
    while stack:
        mem, ths, arrived = stack.pop()
        succ = []
        for t, (pc, regs, pend) in enumerate(ths):
            pending = dict(pend)
            r = dict(zip(REGS, regs))

            def with_thread(pc2, regs2=regs, pend2=pend, t=t):
                nt = list(ths)
                nt[t] = (pc2, regs2, pend2)
                return tuple(nt)

            def setreg(name, v, r=r):
                r2 = dict(r)
                r2[name] = v
                return tuple(r2[k] for k in REGS)

            for a, v in pend:                                  # one pending store lands
                m = list(mem)
                m[a] = v
                succ.append((tuple(m), with_thread(pc, regs, tuple(x for x in pend if x[0] != a)), arrived))
            ins = progs[t][pc]
            op = ins[0]
            if op == "load":
                a = ai[ins[2]]
                succ.append((mem, with_thread(pc + 1, setreg(ins[1], pending.get(a, mem[a]))), arrived))
            elif op == "fence":
                if not pend:
                    succ.append((mem, with_thread(pc + 1), arrived))
            elif op == "rmw":
                a = ai[ins[2]]
                m, p2 = list(mem), pend
                if a in pending:                               # same-address order
                    m[a] = pending[a]
                    p2 = tuple(x for x in pend if x[0] != a)
                old = m[a]
                m[a] = old + 1
                arr2 = arrived
                if ins[3] is not None:
                    arr2 = tuple(x + 1 if i == t else x for i, x in enumerate(arrived))
                succ.append((tuple(m), with_thread(pc + 1, setreg(ins[1], old), p2), arr2))
            elif op == "jne":
                succ.append((mem, with_thread(ins[3] if r[ins[1]] != ins[2] else pc + 1), arrived))
            elif op == "jmp":
                succ.append((mem, with_thread(ins[1]), arrived))
            elif op == "store":
                p2 = dict(pending)
                p2[ai[ins[1]]] = ins[3] + (r[ins[2]] if ins[2] else 0)
                succ.append((mem, with_thread(pc + 1, regs, tuple(sorted(p2.items()))), arrived))
            elif op == "spin":
                a = ai[ins[1]]
                if pending.get(a, mem[a]) == r[ins[2]] + ins[3]:
                    succ.append((mem, with_thread(pc + 1), arrived))
            elif op == "pass":
                if any(x <= ins[1] for x in arrived):
                    return ("early pass", t, ins[1], len(seen))
                succ.append((mem, with_thread(pc + 1), arrived))
        if not succ and any(progs[t][pc][0] != "end" for t, (pc, _, _) in enumerate(ths)):
            return ("hang", [pc for pc, _, _ in ths], len(seen))
        for s in succ:
            if s not in seen:
                seen.add(s)
                stack.append(s)
    return ("ok", len(seen))


def test_weak_order_model_finds_unfenced_hazards():
    # No fence between the last arriver's generation store and the release
    # words: a released workgroup can read the old generation at the next
    # barrier and pass it at once.
    assert check_weak_barrier(2, 2, 2, True, False)[0] == "early pass"
    assert check_weak_barrier(3, 2, 2, False, False)[0] == "early pass"
    # No fence between a leaf's counter reset and its root arrival: the reset
    # can land after a released workgroup's next arrival and erase it.
    assert check_weak_barrier(3, 2, 2, False, True)[0] == "hang"


def test_fenced_barrier_is_safe_under_weak_order():
    for nblocks, fan_in, nbar in [(2, 2, 2), (3, 2, 2), (3, 2, 3), (2, 1, 3), (3, 3, 2)]:
        verdict = check_weak_barrier(nblocks, fan_in, nbar)
        assert verdict[0] == "ok", (nblocks, fan_in, nbar, verdict)
