"""The LDS stack of the group walk (k_walk32g, galic_tree.hip) never overflows.

The kernel takes up to 64 pending entries at a time while the stack has room for
64 x 7 more plus a reserve, else one at a time; an opened node pushes its (at most
8) children.  This model of that discipline runs on trees the kernel can meet --
random ones, and the worst case where every node opens into 8 children down to the
deepest level -- and checks the peak against the stack's size.  The constants are
read from the kernel's source, so the check follows any change to them.
"""
import pathlib
import random
import re

SRC = pathlib.Path(__file__).resolve().parents[1] / "galic_amd" / "csrc" / "galic_tree.hip"


def constants():
    text = SRC.read_text()
    depth = int(re.search(r"constexpr int kMaxDepth = (\d+);", text).group(1))
    m = re.search(r"constexpr int kGStack = (\d+), kGList = (\d+), kGDeep = 7 \* \(kMaxDepth \+ (\d+)\);", text)
    return int(m.group(1)), int(m.group(2)), 7 * (depth + int(m.group(3))), depth


def walk(opens, stack_cap, list_cap, deep, max_steps=4000):
    """Peak stack and list use of one group's walk; opens(depth) -> children of an opened node (0: accepted)."""
    stk, nl, peak, peak_list = [0], 0, 1, 0
    for _ in range(max_steps):
        if not stk:
            break
        top = len(stk)
        k = min(64, top) if stack_cap - top >= 7 * 64 + deep else 1
        if nl + k > list_cap:
            nl = 0
        # the k entries are taken first (one per lane), then the lanes' children pushed
        taken = [stk.pop() for _ in range(k)]
        for d in taken:
            nch = opens(d)
            if nch == 0:
                nl += 1
            else:
                stk.extend([d + 1] * nch)
        peak = max(peak, len(stk))
        peak_list = max(peak_list, nl)
        assert len(stk) <= stack_cap and nl <= list_cap
    return peak, peak_list


def test_group_walk_stack_bound():
    stack_cap, list_cap, deep, depth = constants()
    peaks = []
    # every node opens into 8 children down to the deepest level (the leaves are accepted)
    peaks.append(walk(lambda d: 8 if d < depth else 0, stack_cap, list_cap, deep)[0])
    for seed in range(12):
        rnd = random.Random(seed)
        p_deep = rnd.uniform(0.2, 0.6)

        def opens(d, rnd=rnd, p_deep=p_deep):
            if d >= depth:
                return 0
            p = 0.95 if d < 10 else p_deep
            return rnd.randint(1, 8) if rnd.random() < p else 0

        peaks.append(walk(opens, stack_cap, list_cap, deep)[0])
    assert max(peaks) <= stack_cap
