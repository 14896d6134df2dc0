"""The two-pair walk of the angular loops (PairWalk2 in fortnet_amd/csrc/fnet_acsf.hip), compiled
for the host with g++ from the kernel source itself and run for many list lengths.

The walk must visit every pair the reference's double loop over a group's lists visits
(acsf.F90:1377-1492: ordered pairs j != k of one list, counted here once with weight 2, or every
(j, k) of two lists) exactly once, with the two pairs of a step sharing j, and never name an
entry outside the lists, valid pair or not.  A wrong wrap, offset limit or half-offset rule drops
or repeats pairs and fails here.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "fortnet_amd", "csrc", "fnet_acsf.hip")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")

HARNESS = r"""
#include <cstdio>
#include <cstdlib>
#define __device__
#define __forceinline__ inline
%s
template <bool SAME> void run(int A0, int LA, int B0, int LB) {
  for (int lane = 0; lane < 64; ++lane) {
    PairWalk2<SAME> w;
    w.init(lane, A0, LA, B0, LB);
    int ej, ek1, ek2, s = 0;
    bool v1, v2;
    while (w.next(ej, ek1, ek2, v1, v2))
      printf("%%d %%d %%d %%d %%d %%d %%d\n", lane, s++, ej, ek1, ek2, (int)v1, (int)v2);
  }
}
int main(int argc, char** argv) {
  const int same = atoi(argv[1]), A0 = atoi(argv[2]), LA = atoi(argv[3]), B0 = atoi(argv[4]),
            LB = atoi(argv[5]);
  if (same) run<true>(A0, LA, B0, LB); else run<false>(A0, LA, B0, LB);
  return 0;
}
"""


@pytest.fixture(scope="module")
def walker(tmp_path_factory):
    src = open(SRC).read()
    m = re.search(r"template <bool SAME>\nstruct PairWalk2 \{.*?\n\};\n", src, flags=re.S)
    assert m, "PairWalk2 not found"
    d = tmp_path_factory.mktemp("walk")
    cpp = d / "walk.cpp"
    cpp.write_text(HARNESS % m.group(0))
    exe = d / "walk"
    subprocess.run(["g++", "-O1", "-o", str(exe), str(cpp)], check=True)

    def walk(same, a0, la, b0, lb):
        out = subprocess.run([str(exe), str(int(same)), str(a0), str(la), str(b0), str(lb)],
                             check=True, capture_output=True, text=True).stdout
        rows = np.array([[int(x) for x in l.split()] for l in out.splitlines()], dtype=np.int64)
        return rows.reshape(-1, 7)
    return walk


SAME_LENGTHS = [0, 1, 2, 3, 4, 5, 7, 8, 31, 32, 33, 63, 64, 65, 77, 96, 127, 128, 129, 150]
DIFF_LENGTHS = [(0, 5), (1, 1), (1, 64), (3, 2), (13, 38), (38, 77), (64, 64), (65, 1),
                (77, 38), (100, 7), (130, 129)]


def _check(rows, pairs_expected, a_range, b_range):
    seen = {}
    for lane, step, ej, ek1, ek2, v1, v2 in rows:
        assert a_range[0] <= ej < a_range[1]
        for ek in (ek1, ek2):
            assert b_range[0] <= ek < b_range[1]
        for ek, v in ((ek1, v1), (ek2, v2)):
            if v:
                seen[(ej, ek)] = seen.get((ej, ek), 0) + 1
    assert all(c == 1 for c in seen.values()), "a pair is visited twice"
    assert set(seen) == pairs_expected


@pytest.mark.parametrize("n", SAME_LENGTHS)
def test_same_list_visits_every_unordered_pair_once(walker, n):
    a0 = 5
    rows = walker(True, a0, n, a0, n)
    got = rows if len(rows) else np.zeros((0, 7), np.int64)
    # unordered pairs {j, k}, j != k, each as one (j, k) orientation
    seen = set()
    for lane, step, ej, ek1, ek2, v1, v2 in got:
        for ek, v in ((ek1, v1), (ek2, v2)):
            if v:
                key = (min(ej, ek), max(ej, ek))
                assert ej != ek
                assert key not in seen, "pair visited twice"
                seen.add(key)
        assert a0 <= ej < a0 + n and a0 <= ek1 < a0 + n and a0 <= ek2 < a0 + n
    want = {(a0 + j, a0 + k) for j in range(n) for k in range(j + 1, n)}
    assert seen == want


@pytest.mark.parametrize("la,lb", DIFF_LENGTHS)
def test_two_lists_visit_every_pair_once(walker, la, lb):
    a0, b0 = 3, 3 + la + 4
    rows = walker(False, a0, la, b0, lb)
    if la == 0 or lb == 0:
        assert len(rows) == 0
        return
    # the longer list is walked by j
    if lb > la:
        want = {(b0 + k, a0 + j) for j in range(la) for k in range(lb)}
        _check(rows, want, (b0, b0 + lb), (a0, a0 + la))
    else:
        want = {(a0 + j, b0 + k) for j in range(la) for k in range(lb)}
        _check(rows, want, (a0, a0 + la), (b0, b0 + lb))


def test_lanes_of_a_step_touch_consecutive_entries(walker):
    """64 lanes of one step take 64 consecutive j (mod the list length): the loads and the LDS
    accumulator updates of a step stay free of bank conflicts."""
    rows = walker(True, 0, 100, 0, 100)
    for step in range(int(rows[:, 1].max()) + 1):
        r = rows[rows[:, 1] == step]
        j = r[np.argsort(r[:, 0]), 2]
        assert np.all((np.diff(j) % 100) == 1)
