"""Lockstep model of the team kernel's work hand-out (the wave-local queue
refilled from per-XCD cursors, csrc/pfchem_team_impl.hpp, the block under
"work stealing: hand out cells"), run on the CPU for every team width.

The hand-out is the one piece of the team kernel shared by several teams of
a wavefront: a ballot of the team leaders that need a cell, ranks among
them, a wave-uniform queue [qnext, qend) refilled by one atomic on the
cursor of this XCD's range and then of the others.  Each wavefront here is
64 lanes advanced in lockstep, statement for statement as the kernel
executes them, and wavefronts interleave their atomics in random order.
Invariants: every index of [0, nwork) is handed out exactly once, a team
holds at most one cell, every team ends exhausted, and no team is told the
work is gone while a range still holds cells.  The shapes with 64 and 8
teams per wavefront (1xR and 8x1) are covered as well as the shipped ones.
"""
import numpy as np
import pytest

CLAIM_ALIGN = 16  # kClaimAlign


class Launch:
    def __init__(self, nwork, nxcd, blocks, block, L, listed, rng, slip=0):
        self.nwork, self.nxcd, self.L, self.slip = nwork, nxcd, L, slip
        self.cursors = [0] * nxcd
        self.chunk = (nwork + nxcd - 1) // nxcd
        waves = blocks * (block // 64)
        batch = max(64 // L, CLAIM_ALIGN)
        if listed and nwork < 4 * batch * waves:
            batch = 64 // L
        self.batch = batch
        self.handed = np.zeros(nwork, dtype=np.int64)
        self.waves = [Wave(self, b % nxcd, rng) for b in range(blocks) for _ in range(block // 64)]

    def atomic_add(self, xr, v):
        old = self.cursors[xr]
        self.cursors[xr] += v
        return old


class Wave:
    def __init__(self, launch, xcd, rng):
        self.k, self.xcd, self.rng = launch, xcd, rng
        nteam = 64 // launch.L
        self.cell = [-1] * nteam          # team-uniform state
        self.busy = [0] * nteam           # rounds left on the current cell
        self.exhausted = [False] * nteam
        self.qnext = self.qend = 0
        self.steal = 0

    def done(self):
        return all(self.exhausted)

    def round(self):
        """one pass of the kernel's loop body: hand-out, then one unit of
        work on each team's cell"""
        k, L = self.k, self.k.L
        nteam = 64 // L
        # needy leaders (lane t == 0 of each team) and the ballot
        needy = [self.cell[tm] < 0 and not self.exhausted[tm] for tm in range(nteam)]
        needmask = [needy[lane // L] if lane % L == 0 else False for lane in range(64)]
        got = [-1] * 64
        if any(needmask):
            cnt = sum(needmask)
            rk = [sum(needmask[:lane]) if needmask[lane] else 0 for lane in range(64)]
            avail = self.qend - self.qnext
            for lane in range(64):
                if needmask[lane] and rk[lane] < avail:
                    got[lane] = self.qnext + rk[lane]
            taken = min(avail, cnt)
            self.qnext += taken
            while taken < cnt and self.steal < k.nxcd:
                xr = (self.xcd + self.steal) % k.nxcd
                lo = xr * k.chunk
                hi = min(k.nwork, lo + k.chunk)
                base = k.atomic_add(xr, k.batch)  # lane 0, then broadcast
                if lo + base >= hi:
                    self.steal += 1
                    continue
                self.qnext = lo + base
                self.qend = min(hi, self.qnext + k.batch - k.slip)
                avail = self.qend - self.qnext
                need = cnt - taken
                for lane in range(64):
                    if needmask[lane] and taken <= rk[lane] < taken + avail:
                        got[lane] = self.qnext + (rk[lane] - taken)
                take = min(avail, need)
                self.qnext += take
                taken += take
            for lane in range(64):
                if needmask[lane] and got[lane] < 0:
                    got[lane] = k.nwork
        # got = __shfl(got, team_in_wave * L): every lane takes its leader's
        for tm in range(nteam):
            g = got[tm * L]
            if self.cell[tm] < 0 and not self.exhausted[tm] and g >= 0:
                if g >= k.nwork:
                    # only when every range is used up
                    assert all(k.cursors[x] >= min(k.nwork, x * k.chunk + k.chunk) - x * k.chunk
                               for x in range(k.nxcd)), "told exhausted while cells remain"
                    self.exhausted[tm] = True
                else:
                    k.handed[g] += 1
                    self.cell[tm] = g
                    self.busy[tm] = int(self.rng.integers(1, 6))
        # one unit of work; a team whose cell is done takes a new one next round
        for tm in range(nteam):
            if self.cell[tm] >= 0:
                self.busy[tm] -= 1
                if self.busy[tm] == 0:
                    self.cell[tm] = -1


@pytest.mark.parametrize("L", [1, 2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("case", [
    # (nwork, blocks, block, listed)
    (1000, 16, 256, False),     # fewer cells than teams in a few waves
    (5000, 8, 64, False),       # more cells than teams: several refills
    (37, 32, 256, True),        # a short hand-over list: one cell per team
    (3001, 16, 128, True),
    (0, 8, 64, False),          # no work
])
def test_every_cell_handed_out_once(L, case):
    nwork, blocks, block, listed = case
    rng = np.random.default_rng(1000 * L + nwork)
    k = Launch(nwork, 8, blocks, block, L, listed, rng)
    _run(k, rng)
    assert all(w.done() for w in k.waves), "a team never ended"
    assert np.all(k.handed == 1), (np.flatnonzero(k.handed != 1)[:10], k.handed.max())


def _run(k, rng):
    for _ in range(200000):
        live = [w for w in k.waves if not w.done()]
        if not live:
            break
        # wavefronts run in any order; one round of one of them at a time
        live[int(rng.integers(len(live)))].round()


def test_the_model_catches_a_planted_queue_slip():
    """the same hand-out with the refilled queue one cell shorter than the
    batch its atomic claimed (a plausible off-by-one) skips cells, which the
    invariants catch"""
    rng = np.random.default_rng(3)
    k = Launch(4096, 8, 8, 64, 16, False, rng, slip=1)
    _run(k, rng)
    assert not np.all(k.handed == 1)
