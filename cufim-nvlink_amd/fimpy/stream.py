"""Continuous batching: a stream of solves through the instances of one
converge-mode solver.  A converged instance is handed to the caller
(on_map), cleared (fim_recycle) and seeded with the next solve while the
other instances keep running, so the end of one map's front no longer
leaves the rest of the batch idle.  Every map comes out bit-identical to a
solve started on its own (each instance's time window restarts when it is
seeded again)."""
from __future__ import annotations


def stream_solve(solver, sets, scales=None, on_map=None, check_iters=6, max_iters=10_000_000):
    """Solves len(sets) maps (pacing vertex arrays; scales: optional per-solve
    4-vectors of regional speed factors) through solver.n_instances slots.
    on_map(solve_index, inst) is called for every converged map before its
    instance is reused.  Returns the number of global iterations run, as the
    device counts them (iterations that processed work; fim_advance runs
    whole captured graphs, so one call may run more than check_iters);
    max_iters caps that count."""
    n = solver.n_instances
    slot = [-1] * n
    solver.reset()

    def start(insts, first):
        for k, i in enumerate(insts):
            slot[i] = first + k
            if scales is not None:
                solver.set_scaling(scales[first + k], inst=i)
        solver.seed_batch([sets[first + k] for k in range(len(insts))], 0.0, insts=insts)

    first = list(range(min(n, len(sets))))
    if first:
        start(first, 0)
    nxt, done, iters, reset = len(first), 0, 0, True
    while done < len(sets):
        if iters >= max_iters:
            raise RuntimeError("iteration cap reached before the stream was solved")
        solver.advance(check_iters, reset_stats=reset)
        reset = False
        iters = int(solver.stats()["global_iters"])  # accumulated since the first advance
        act = solver.active()
        fin = [i for i in range(n) if slot[i] >= 0 and not act[i]]
        for i in fin:
            if on_map is not None:
                on_map(slot[i], i)
            slot[i] = -1
            done += 1
        refill = fin[:max(0, len(sets) - nxt)]
        if refill:
            solver.recycle(refill)
            start(refill, nxt)
            nxt += len(refill)
    return iters
