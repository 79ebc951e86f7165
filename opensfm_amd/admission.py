"""Many references on one GPU, in groups admitted by device memory.

The reference's dense driver cuts the views into clusters of at most
``max_cluster_size`` views (opensfm/dense/clustering.py:164-216) and hands
them to per-device workers that run each cluster's references one after the
other (opensfm/dense/depthmaps.py:70-151; cluster_estimator.cc:82-156).
Here the unit of admission is device memory: references are built one at a
time, each reports the device bytes its run allocates
(``DepthmapEstimator.device_bytes``), and a group takes references in order
while their sum fits the budget (by default 90 % of the device memory free
when the run starts).  A reference larger than the budget runs alone.

A group runs as one ``DepthmapBatch`` (one launch per kernel for all its
references, captured in a HIP graph), level by level when the references
have several pyramid levels, is refined by Gauss-Newton when asked,
and hands each estimator to ``on_result`` while its results are on the
device; then the group's device memory is released and its estimators are
dropped.  While a group runs on the device, a worker thread builds the next
group's estimators (host work: views and parameters) and, with
``overlap_staging``, stages it (host conversion, uploads and graph capture
on the batch's own stream): each group then gets half the budget, so that
two groups fit at once.  With ``overlap_staging`` the worker also takes a
finished group's results and release, while the next group runs; it stages
the group after that only once they are done, so at most two groups hold
device memory at any time.  The host holds at most two groups' images.  During
the run the library keeps released device buffers for reuse
(``DepthmapEstimator.set_buffer_cache``, up to the budget): groups of
same-shaped references then reuse the blocks of an earlier group instead of
freeing and re-allocating them.
"""
from __future__ import annotations

import dataclasses
import time
from concurrent.futures import ThreadPoolExecutor
from typing import Any, Callable, Dict, List, Optional, Sequence


@dataclasses.dataclass
class GroupStats:
    refs: List[int]            # reference indices of the group
    device_bytes: int          # sum of the references' device_bytes()
    build_s: float = 0.0       # host: building the estimators (views, parameters), on the worker
    wait_s: float = 0.0        # time the run waited for the worker (the part not overlapped)
    prepare_s: float = 0.0     # staging, first-use tuning and graph capture (on the worker when overlapped)
    run_ms: float = 0.0        # device time of the batched PatchMatch run
    refine_ms: float = 0.0     # device time of the Gauss-Newton steps (sum over the group)
    results_s: float = 0.0     # on_result calls (downloads; on the worker when overlapped)
    release_s: float = 0.0     # releasing the device memory (on the worker when overlapped)
    wall_s: float = 0.0        # the whole group, end to end
    stage_ms: Dict[str, List[float]] = dataclasses.field(default_factory=dict)  # {stage: [ms, launches]}, eager runs


def admit(sizes: Sequence[int], budget: int) -> List[List[int]]:
    """Split indices 0..len(sizes)-1, in order, into groups whose sizes sum
    to at most ``budget``; an item larger than the budget forms a group of
    its own."""
    if budget <= 0:
        raise ValueError("admit: the budget must be positive")
    groups: List[List[int]] = []
    cur: List[int] = []
    used = 0
    for i, s in enumerate(sizes):
        if s < 0:
            raise ValueError("admit: negative size")
        if cur and used + s > budget:
            groups.append(cur)
            cur, used = [], 0
        cur.append(i)
        used += s
    if cur:
        groups.append(cur)
    return groups


def default_budget(pydense, device: int = 0, fraction: float = 0.9) -> int:
    """``fraction`` of the device memory free now (release cached torch
    blocks first if torch shares the device)."""
    return int(pydense.DepthmapEstimator.device_free_memory(device) * fraction)


def run_admitted(pydense, count: int, build: Callable[[int], Any], budget: Optional[int] = None,
                 refine_steps: int = 0, damping: float = 1e-3,
                 on_result: Optional[Callable[[int, Any], Any]] = None, device: int = 0,
                 batch_config: Optional[Dict[str, Any]] = None,
                 on_group: Optional[Callable[[GroupStats], None]] = None, reuse_buffers: bool = True,
                 overlap_staging: bool = False):
    """Run references 0..count-1 in memory-admitted groups.

    ``build(i)`` returns a configured ``DepthmapEstimator`` of reference i
    (views added, parameters set, on ``device``); it is called in index
    order on a worker thread.  ``on_result(i, est)`` is called once per
    reference after its group's run and refinement; its return value is
    collected (default: ``est.download(["depth"])``).  ``batch_config`` goes
    to ``DepthmapBatch.set_config``.  ``on_group`` is called with each
    finished group's timing.  With ``overlap_staging`` both callbacks run on
    the worker thread, one group at a time and in group order.  ``reuse_buffers``: keep released device
    buffers for the next group (the cache is emptied at the end).
    ``overlap_staging``: stage the next group while the current one runs
    (groups of half the budget), and collect a group's results and release
    it on the worker while the next one runs.  Returns
    ``(results: {i: value}, groups: [GroupStats])``.
    """
    if count < 0:
        raise ValueError("run_admitted: negative count")
    if budget is None:
        budget = default_budget(pydense, device)
    if budget <= 0:
        raise ValueError("run_admitted: no device memory to admit references into")
    if on_result is None:
        on_result = lambda i, est: est.download(["depth"])  # noqa: E731
    results: Dict[int, Any] = {}
    groups: List[GroupStats] = []
    group_budget = budget // 2 if overlap_staging else budget

    def make_batch(ests):
        batch = pydense.DepthmapBatch(device)
        for e in ests:
            batch.add(e)
        if batch_config:
            batch.set_config(**batch_config)
        return batch

    def next_group(pending, i):
        """Estimators from index i on (after ``pending``, built earlier but
        not admitted) while they fit the group budget; staged when
        overlapping."""
        t0 = time.perf_counter()
        ests, refs, used = [], [], 0
        while True:
            if pending is not None:
                idx, est, need = pending
                pending = None
            elif i < count:
                idx, est = i, build(i)
                need = int(est.device_bytes())
                i += 1
            else:
                break
            if refs and used + need > group_budget:
                pending = (idx, est, need)
                break
            ests.append(est)
            refs.append(idx)
            used += need
        build_s = time.perf_counter() - t0
        batch, prep_s = None, 0.0
        if overlap_staging and ests:
            batch = make_batch(ests)
            t0 = time.perf_counter()
            batch.prepare()
            prep_s = time.perf_counter() - t0
        return ests, refs, used, pending, i, build_s, batch, prep_s

    def finish(ests, refs, st, t_group):
        """Results, then release, of a group whose device work is done."""
        t0 = time.perf_counter()
        for idx, e in zip(refs, ests):
            results[idx] = on_result(idx, e)
        st.results_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        for e in ests:
            e.release_device()
        ests.clear()
        st.release_s = time.perf_counter() - t0
        st.wall_s = time.perf_counter() - t_group
        groups.append(st)
        if on_group is not None:
            on_group(st)

    est_cls = pydense.DepthmapEstimator
    if reuse_buffers:
        est_cls.set_buffer_cache(device, budget)
    try:
        with ThreadPoolExecutor(max_workers=1) as pool:
            finishing = []
            fut = pool.submit(next_group, None, 0)
            while fut is not None:
                t_group = time.perf_counter()
                ests, refs, used, pending, i, build_s, batch, prep_s = fut.result()
                wait_s = time.perf_counter() - t_group
                # with overlap_staging the worker runs this after the previous
                # group's finish(): that group's memory is free before the
                # next one is staged
                fut = pool.submit(next_group, pending, i) if (pending is not None or i < count) else None
                if not ests:
                    break
                st = GroupStats(refs=refs, device_bytes=used, build_s=build_s, wait_s=wait_s, prepare_s=prep_s)
                if batch is None:
                    batch = make_batch(ests)
                    t0 = time.perf_counter()
                    batch.prepare()
                    st.prepare_s = time.perf_counter() - t0
                # coarse to fine: a batch starts at level 0 of every reference's
                # pyramid; each finer level starts from the level below it
                # (the references of a batch share their number of levels)
                def run_level():
                    ms = batch.run()
                    for k, (sm, n) in batch.stage_timing().items():   # zero after a graph run
                        acc = st.stage_ms.setdefault(k, [0.0, 0])
                        acc[0] += sm
                        acc[1] += n
                    return ms

                st.run_ms = run_level()
                for level in range(1, int(ests[0].params()["hierarchy_levels"])):
                    batch.set_level(level)
                    st.run_ms += run_level()
                for e in ests:
                    if refine_steps > 0:
                        e.gauss_newton(refine_steps, damping)
                        st.refine_ms += e.last_refine_ms()
                del batch
                if overlap_staging:
                    finishing.append(pool.submit(finish, list(ests), refs, st, t_group))
                    ests.clear()
                    # a failed finish ends the run before the next group starts
                    while finishing and finishing[0].done():
                        finishing.pop(0).result()
                else:
                    finish(ests, refs, st, t_group)
            for f in finishing:
                f.result()
    finally:
        if reuse_buffers:
            est_cls.set_buffer_cache(device, 0)
    return results, groups
