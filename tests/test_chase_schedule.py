"""Host check of the bulge chase's dependency rule (csrc/bulge.hip).

Every step of the chase -- the right and the left reflector step of task
(s, t) -- touches a block of the band; the GPU kernel runs the sweeps on
different waves and orders them only by progress keys.  This
enumerates every pair of steps of different sweeps whose blocks meet and
checks that the waits order them as in the sequential chase (transitively
through the intermediate sweeps).  The footprints restate the kernel's task
geometry: task t of sweep s covers r = s (t = 0) or s + 1 + (t-1) NB,
c = s + 1 or r + NB, L = min(NB, C - c); right step rows r..c+L-1 x columns
c..c+L-1, left step rows c..c+L-1 x columns c..min(C, c+L+NB)-1."""
import pytest

NB = 32


def _tasks(C, s):
    if s > C - 2:
        return []
    rest = C - 3 - s
    n = 1 + (rest // NB if rest >= 0 else 0)
    out = []
    for t in range(n):
        r = s if t == 0 else s + 1 + (t - 1) * NB
        c = s + 1 if t == 0 else r + NB
        out.append((r, c, min(NB, C - c)))
    return out


def _foot(tk, t, step):
    r, c, L = tk[t]
    if step == 0:
        return (r, c + L - 1, c, c + L - 1)
    return (c, c + L - 1, c, min(C_CUR[0], c + L + NB) - 1)


C_CUR = [0]


def _meet(a, b):
    return not (a[1] < b[0] or b[1] < a[0] or a[3] < b[2] or b[3] < a[2])


def violations(C, need):
    """need(step, t, ntp): half-steps of sweep s-1 (task t' right = 2t'+1,
    left = 2t'+2) done before step (s, t, step) starts."""
    C_CUR[0] = C
    T = [_tasks(C, s) for s in range(C - 1)]
    F = [[[_foot(T[s], t, st) for st in (0, 1)] for t in range(len(T[s]))] for s in range(C - 1)]
    bad = 0
    for s in range(1, C - 1):
        for t in range(len(T[s])):
            for st in (0, 1):
                B = F[s][t][st]
                pref = need(st, t, len(T[s - 1]))
                sp = s - 1
                while sp >= 0:
                    for tp in range(len(T[sp])):
                        for stp in (0, 1):
                            if 2 * tp + stp >= pref and _meet(F[sp][tp][stp], B):
                                bad += 1
                    if pref <= 0 or sp == 0:
                        # nothing of the sweeps below is guaranteed either
                        for q in range(sp - 1, -1, -1):
                            for tp in range(len(T[q])):
                                for stp in (0, 1):
                                    if _meet(F[q][tp][stp], B):
                                        bad += 1
                        break
                    h = pref - 1  # last guaranteed half-step of sweep sp waited on sp-1
                    pref = need(h % 2, h // 2, len(T[sp - 1]))
                    sp -= 1
    return bad


def kernel_rule(step, t, ntp):
    # k_bulge: task (s, t) -- both its steps -- after tasks 0..t+2 of s-1
    return min(2 * t + 6, 2 * ntp)


def least_rule(step, t, ntp):
    # the least each step needs: a right step tasks 0..t+1, a left step also
    # the right step of task t+2
    return min(2 * t + 4 if step == 0 else 2 * t + 5, 2 * ntp)


@pytest.mark.parametrize("C", [40, 67, 100, 161])
def test_rules_order_every_conflict(C):
    assert violations(C, kernel_rule) == 0
    assert violations(C, least_rule) == 0


@pytest.mark.parametrize("C", [67, 161])
def test_half_step_less_than_least_fails(C):
    """The check is not vacuous: one half-step less than the least rule, for
    either step, misses a conflict."""
    weaker_right = lambda st, t, n: min(2 * t + 3 if st == 0 else 2 * t + 5, 2 * n)
    weaker_left = lambda st, t, n: min(2 * t + 4, 2 * n)
    assert violations(C, weaker_right) > 0
    assert violations(C, weaker_left) > 0
