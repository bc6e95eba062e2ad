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

The one-lane kernel's hand-out (csrc/pfchem_lane.hpp) is modelled the same
way, with its cell-list mode and the rescue filter: every index claimed once,
every listed handed-over cell stepped once, the others skipped.
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


# ---------------------------------------------------------------------------
# the one-lane kernel's hand-out (csrc/pfchem_lane.hpp, "work stealing: hand
# out cells"): one cell per lane, 64-index claims, an optional cell list and
# the rescue filter (a pass over handed-over cells skips the others)
# ---------------------------------------------------------------------------
LANE_CLAIM = 64  # kClaim


class LaneLaunch:
    def __init__(self, nwork, nxcd, waves, listed, rescue, rng):
        self.nwork, self.nxcd = nwork, nxcd
        self.cursors = [0] * nxcd
        self.chunk = (nwork + nxcd - 1) // nxcd
        self.list = rng.permutation(4 * max(nwork, 1))[:nwork] if listed else None
        ncell = 4 * max(nwork, 1) if listed else nwork
        self.escalated = rng.random(ncell) < 0.5 if rescue else np.ones(ncell, dtype=bool)
        self.rescue = rescue
        self.claimed = np.zeros(nwork, dtype=np.int64)  # per list / range index
        self.stepped = np.zeros(ncell, dtype=np.int64)  # per cell
        self.waves = [LaneWave(self, w % nxcd, rng) for w in range(waves)]


class LaneWave:
    def __init__(self, launch, xcd, rng):
        self.k, self.xcd, self.rng = launch, xcd, rng
        self.cell = [-1] * 64
        self.busy = [0] * 64
        self.exhausted = [False] * 64
        self.qn = self.qe = 0
        self.steal = 0

    def done(self):
        return all(c < 0 for c in self.cell) and all(self.exhausted)

    def round(self):
        k = self.k
        needy = [self.cell[l] < 0 and not self.exhausted[l] for l in range(64)]
        mask = list(needy)
        while any(mask):
            if self.qn >= self.qe:
                if self.steal >= k.nxcd:
                    for l in range(64):
                        if needy[l]:
                            self.exhausted[l] = True
                        needy[l] = False
                    break
                xr = self.xcd + self.steal if self.xcd + self.steal < k.nxcd else \
                    self.xcd + self.steal - k.nxcd
                lo = xr * k.chunk
                hi = min(lo + k.chunk, k.nwork)
                base = k.cursors[xr]          # atomicAdd by the first active lane
                k.cursors[xr] += LANE_CLAIM
                if lo + base >= hi:
                    self.steal += 1
                    continue
                self.qn = lo + base
                self.qe = min(self.qn + LANE_CLAIM, hi)
            avail = self.qe - self.qn
            cnt = sum(mask)
            for l in range(64):
                rk = sum(mask[:l])
                if needy[l] and rk < avail:
                    idx = self.qn + rk
                    k.claimed[idx] += 1
                    c = int(k.list[idx]) if k.list is not None else idx
                    self.cell[l] = c
                    needy[l] = False
                    if k.rescue and not k.escalated[c]:
                        self.cell[l] = -1
                        needy[l] = True
                    else:
                        k.stepped[c] += 1
                        self.busy[l] = int(self.rng.integers(1, 8))
            self.qn += min(cnt, avail)
            mask = list(needy)
        for l in range(64):
            if self.cell[l] >= 0:
                self.busy[l] -= 1
                if self.busy[l] == 0:
                    self.cell[l] = -1


@pytest.mark.parametrize("listed,rescue", [(False, False), (False, True), (True, True), (True, False)])
@pytest.mark.parametrize("nwork,waves", [(1, 8), (1000, 8), (20000, 16), (0, 8), (4099, 24)])
def test_lane_hand_out(listed, rescue, nwork, waves):
    rng = np.random.default_rng(nwork * 7 + waves + 2 * listed + rescue)
    k = LaneLaunch(nwork, 8, waves, listed, rescue, rng)
    for _ in range(500000):
        live = [w for w in k.waves if not w.done()]
        if not live:
            break
        live[int(rng.integers(len(live)))].round()
    assert all(w.done() for w in k.waves)
    assert np.all(k.claimed == 1)
    want = np.zeros_like(k.stepped)
    cells = k.list if listed else np.arange(nwork)
    want[cells] = k.escalated[cells]
    np.testing.assert_array_equal(k.stepped, want)
