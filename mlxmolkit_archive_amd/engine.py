"""Device-side entry points: thin wrappers over the C ABI operating on torch tensors.

Everything here keeps data on the GPU. The drop-in APIs (minimize.py, optimizer.py,
kernel.py, pipeline.py) are built on these functions.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib
from .packing import DeviceBatch, HostBatch, lpt_order, pack

TERM_NAMES = ("bond", "angle", "strbend", "oop", "torsion", "vdw", "ele")


def _p(t) -> int:
    return 0 if t is None else int(t.data_ptr())


def device_batch(params_list, conf_counts, device="cuda") -> DeviceBatch:
    torch = _lib.require_gpu()
    return DeviceBatch(pack(params_list, conf_counts), device=device)


def energy_grad(db: DeviceBatch, x, family: int = _lib.FAMILY_TG, diel: int = 1, mask: int = 0x7F,
                grad: bool = True, terms: bool = False, stream=None):
    """(E float64 [C], G float32 like x or None, E_terms float64 [C,7] or None) on the GPU."""
    torch = _lib.require_gpu()
    L = _lib.load()
    assert x.dtype == torch.float32 and x.is_cuda and x.numel() == db.host.total_coords
    C_ = db.n_confs
    e = torch.empty(C_, dtype=torch.float64, device=x.device)
    g = torch.empty_like(x) if grad else None
    et = torch.empty((C_, 7), dtype=torch.float64, device=x.device) if terms else None
    _lib.check(L.mmffhip_energy_grad(C.byref(db.desc), family, diel, _p(x), _p(e), _p(et), _p(g), mask,
                                     _lib.stream_ptr(stream)))
    return e, g, et


@dataclass
class NkResult:
    x: "object"          # float32 [total] on device
    energy: "object"     # float64 [C]
    status: "object"     # int32 [C]
    counters: "object"   # int32 [C,3] iterations, gradient evals, energy evals


def nk_options(max_iters=500, grad_tol=1e-4, use_lbfgs=False, lbfgs_m=8, diel=1, hess_place=_lib.HESS_AUTO,
               block=0, ls_grad=False, precond=None, table_lds_bytes=0) -> _lib.NkOpts:
    o = _lib.NkOpts()
    o.algo = _lib.LBFGS if use_lbfgs else _lib.BFGS
    o.max_iters = int(max_iters)
    o.grad_tol = float(grad_tol)
    o.diel = int(diel)
    o.lbfgs_m = int(lbfgs_m)
    o.hess_place = int(hess_place)
    o.block = int(block)
    o.ls_grad = int(bool(ls_grad))
    o.precond = _p(precond)
    o.table_lds_bytes = int(table_lds_bytes)
    return o


def xcd_split(sel: np.ndarray, conf_mol: np.ndarray, n3: np.ndarray, n_xcd: int = 8):
    """Split a launch's conformers into per-XCD work queues, whole molecules per queue, so the
    conformers that share a molecule's term tables run on the XCD whose L2 holds them.
    Molecules go to the least-loaded queue, costliest first (cost: sum of n^2 over their
    conformers); each queue keeps the longest-first order. Returns (order, offsets[n_xcd + 1])."""
    import heapq
    sel = np.asarray(sel, np.int32)
    cm = conf_mol[sel]
    mols, inv = np.unique(cm, return_inverse=True)
    cost = np.bincount(inv, weights=n3[sel].astype(np.float64) ** 2, minlength=len(mols))
    bins = np.zeros(len(mols), np.int32)
    heap = [(0.0, x) for x in range(n_xcd)]
    for mi in np.argsort(-cost, kind="stable"):
        load, x = heapq.heappop(heap)
        bins[mi] = x
        heapq.heappush(heap, (load + cost[mi], x))
    b = bins[inv]
    perm = np.lexsort((-n3[sel], b))  # by queue, then longest first (stable)
    order = np.ascontiguousarray(sel[perm])
    offsets = np.zeros(n_xcd + 1, np.int32)
    offsets[1:] = np.cumsum(np.bincount(b, minlength=n_xcd))
    return order, offsets


def resident_workgroups(desc, o) -> int:
    """Workgroups the persistent minimize_nk launch keeps resident on the device for this
    batch descriptor's molecules and these options (the grid of an unbounded batch)."""
    L = _lib.load()
    d = _lib.Batch.from_buffer_copy(desc)
    d.n_confs = 1 << 30
    g = C.c_int32()
    _lib.check(L.mmffhip_minimize_nk_plan(C.byref(d), C.byref(o), C.byref(g), None, None, None))
    return int(g.value)


def planned_placement(desc, o) -> int:
    """The Hessian placement the minimize_nk launch will use (HESS_LDS / _GLOBAL /
    _GLOBAL_PACKED / _GLOBAL_SYM): AUTO resolved, and a symmetric or packed request with a
    learned start or L-BFGS mapped to the full matrix, as the planner does."""
    L = _lib.load()
    g, hp = C.c_int32(), C.c_int32()
    _lib.check(L.mmffhip_minimize_nk_plan(C.byref(desc), C.byref(o), C.byref(g), None, None, C.byref(hp)))
    return int(hp.value)


def table_lds_bytes(hb: HostBatch, pairs: bool = True) -> np.ndarray:
    """Per-molecule bytes of the LDS staging layout of its term tables (csrc table_lds_bytes);
    without ``pairs``, of its bonded tables only (what a global-Hessian launch stages)."""
    d = hb.mol_desc.reshape(-1, 16).astype(np.int64)
    nb, na, nsb, noop, nt, npair = d[:, 2], d[:, 4], d[:, 6], d[:, 8], d[:, 10], d[:, 12]
    npair = npair if pairs else 0 * npair
    return 16 * (nt + nsb + npair) + 8 * (2 * na + nsb + noop + nt + nb) + 4 * (nb + nsb + noop + npair)


# Size buckets for the nk minimizer: (max 3N, threads per workgroup, Hessian placement).
# Each bucket is its own persistent launch sized for its largest molecule. Every default
# bucket keeps the inverse Hessian in a global-memory slot per resident workgroup (L2 /
# Infinity Cache resident) as symmetric 4x4 blocks, one per pair of row quads: half the bytes
# of the full matrix, each block read once per iteration for both halves of H v. Those
# workgroups need little LDS, so registers set the residency (three waves per SIMD up to 256
# threads) and one conformer's memory waits overlap the others' term evaluation. Small
# molecules run one conformer per wave (64 threads, 12 conformers per CU). Measured on MI355X
# by autotune_buckets on 8,192-conformer samples of each class of the ragged workload, ms
# (sym = symmetric blocks, global = full matrix, lds = packed triangle in LDS):
#   n <= 96:        64 sym 10.6, 64 global 11.7, 128 global 12.3, 128 lds 16.1, 64 lds 21.5
#   96 < n <= 192:  128 sym 39.3, 128 global 41.9, 192 global 43.7, 256 global 47.8, 64 sym 47.9
#   192 < n <= 264: 256 sym 122.5, 128 sym 125.8, 256 global 136.8, 192 global 159.9
#   264 < n <= 360: 256 sym 256.4, 512 sym 405.6, 256 global 429.3, 192 global 452.0, 512 global 474.8
#   n > 360:        256 sym 514.9, 512 sym 698.7, 512 global 846.1, 640 global 942.9, 256 global 1006.2
# With the full matrix, one 512-thread workgroup per CU did best above n = 360 (three large
# Hessians per CU outgrew the Infinity Cache); at half the bytes, 256 threads do.
_SYM = _lib.HESS_GLOBAL_SYM
DEFAULT_BUCKETS = ((96, 64, _SYM), (192, 128, _SYM), (360, 256, _SYM), (1 << 20, 256, _SYM))
# L-BFGS keeps its (s, y) history in global memory: 128 threads up to n = 600, 256 above.
# Whole classes, m = 10, maxIters 2000, ms per class: ragged n <= 192 (119,560 conformers)
# 704 with global history against 1233 with the history in LDS; ragged 192 < n <= 600
# (80,440): 128 threads 2313, 192: 2690, 256: 2676; chain n > 600 (10,928): 128: 5144,
# 192: 3899, 256: 3149, 512: 4939. The chain workload's 5,072 conformers of n <= 600 fill
# the 128-thread launch only three times over, so they fold into the 256-thread one.
LBFGS_BUCKETS = ((600, 128, _lib.HESS_GLOBAL), (1 << 20, 256, _lib.HESS_GLOBAL))


class NkRunner:
    """Pre-planned minimize_nk launches for one batch (one launch per size bucket).

    ``buckets``: sequence of (max 3N, block, hess_place) or (max 3N, block, hess_place, algo);
    conformers go to the first bucket whose bound covers them (``autotune_buckets`` picks them
    by measurement). ``algo`` ("bfgs" or "lbfgs") chooses the method per bucket, so dense BFGS
    or L-BFGS is chosen per conformer by its size; without it every bucket runs the method of
    ``use_lbfgs``. ``order``: optional conformer order (default: longest first); conformers left
    out of it are not minimized.

    ``min_fill``: a bucket whose conformers number fewer than ``min_fill`` times the workgroups
    its persistent launch keeps resident (``mmffhip_minimize_nk_plan``) is folded into the next
    bucket of the same method. Launches run one after another, and each ends with a tail in
    which its last conformers leave most of the device idle; a bucket that fills the device
    only a few times over loses more to its tail than its better shape gains (medium workload:
    163.8 ms per step with a 128-thread launch for its 4,510 conformers of 96 < n <= 192,
    against 149.7 ms with them in the 256-thread launch). 0 keeps every bucket.

    ``xcd_queues``: give each launch eight work queues, one per XCD, with whole molecules per
    queue (``xcd_split``), so conformers sharing a molecule's term tables run under one L2;
    a workgroup drains its own XCD's queue, then the others. Results are the same as with one
    queue (each conformer's workgroup computes alone); ragged workload 4,048 against 4,109 ms
    per step, medium 142.3 against 145.4 ms.
    """

    def __init__(self, db: DeviceBatch, order: Optional[np.ndarray] = None, buckets=None, min_fill: int = 8,
                 xcd_queues: bool = True, **opts):
        torch = _lib.require_gpu()
        self.db = db
        self._precond = opts.get("precond")
        use_lbfgs = bool(opts.get("use_lbfgs", False))
        user_block = int(opts.pop("block", 0) or 0)
        user_place = int(opts.pop("hess_place", _lib.HESS_AUTO) or 0)
        stage = bool(opts.pop("stage_tables", True))
        tbytes_mol = table_lds_bytes(db.host)
        bbytes_mol = table_lds_bytes(db.host, pairs=False)
        if buckets is None:
            buckets = LBFGS_BUCKETS if use_lbfgs else DEFAULT_BUCKETS
        hb = db.host
        n3 = 3 * hb.n_atoms[hb.conf_mol]
        if order is None:
            order = lpt_order(hb)
        order = np.asarray(order, np.int32)
        dev = db.t["conf_mol"].device
        L = _lib.load()
        self.launches = []
        assigned = np.zeros(hb.n_confs, bool)
        algos = [bk[3] if len(bk) > 3 else ("lbfgs" if use_lbfgs else "bfgs") for bk in buckets]
        for k, bk in enumerate(buckets):
            bound, block, place = bk[:3]
            algo = algos[k]
            if algo not in ("bfgs", "lbfgs"):
                raise ValueError(f"bucket algo must be 'bfgs' or 'lbfgs', not {algo!r}")
            sel = order[(n3[order] <= bound) & ~assigned[order]]
            if len(sel) == 0:
                continue
            desc = _lib.Batch.from_buffer_copy(db.desc)
            desc.n_confs = len(sel)
            desc.max_atoms = int(hb.n_atoms[hb.conf_mol[sel]].max())
            hp = user_place or place
            per_mol = bbytes_mol if hp in (_lib.HESS_GLOBAL, _lib.HESS_GLOBAL_SYM) else tbytes_mol
            tb = int(per_mol[hb.conf_mol[sel]].max()) if stage else 0
            bopts = dict(opts, use_lbfgs=algo == "lbfgs")
            o = nk_options(block=user_block or block, hess_place=hp, table_lds_bytes=tb, **bopts)
            # too few to fill the device many times over: the next bucket (same method, larger
            # bound) takes them
            fold = min_fill > 0 and k + 1 < len(buckets) and algos[k + 1] == algo and buckets[k + 1][0] >= bound
            if fold and len(sel) < min_fill * resident_workgroups(desc, o):
                continue
            assigned[sel] = True
            xoff = None
            if xcd_queues:
                sel, xoff = xcd_split(sel, hb.conf_mol, n3)
                o.xcd_offsets = xoff.ctypes.data
            nbytes = C.c_size_t()
            _lib.check(L.mmffhip_minimize_nk_workspace(C.byref(desc), C.byref(o), C.byref(nbytes)))
            ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=dev)
            ordt = torch.from_numpy(np.ascontiguousarray(sel)).to(dev)
            self.launches.append(dict(desc=desc, opts=o, ws=ws, ws_bytes=int(nbytes.value), order=ordt,
                                      n=len(sel), max_n=3 * desc.max_atoms, tbytes=tb, algo=algo, xcd_off=xoff,
                                      placement=planned_placement(desc, o)))
        assert assigned[order].all(), "a conformer fits no bucket"
        C_ = db.n_confs
        self.e = torch.empty(C_, dtype=torch.float64, device=dev)
        self.st = torch.empty(C_, dtype=torch.int32, device=dev)
        self.cnt = torch.empty((C_, 3), dtype=torch.int32, device=dev)

    def run(self, x_in, x_out=None, stream=None, overlap: bool = False, events=None) -> NkResult:
        """Minimize every conformer of the plan. The size-bucket launches run one after another
        on ``stream`` (ragged workload on MI355X: 2.66 s per step, against 2.68-2.74 s with the
        launches overlapped; 4.29 against 4.49 s with the full-matrix Hessians, which competed
        for the Infinity Cache). With
        ``overlap`` they go to up to three side streams instead (forked from and joined back to
        ``stream``); the launches write disjoint conformers, so results are the same.
        ``events``: optional list of (start, end) torch.cuda.Event pairs, one per launch,
        recorded around each launch on the stream it runs on (per-launch kernel time)."""
        torch = _lib.require_gpu()
        assert x_in.dtype == torch.float32 and x_in.numel() == self.db.host.total_coords
        if x_out is None:
            x_out = torch.empty_like(x_in)
        L = _lib.load()
        main = stream if stream is not None else torch.cuda.current_stream()
        streams = [main]
        if overlap and len(self.launches) > 1:
            if not hasattr(self, "_side"):
                self._side = [torch.cuda.Stream(device=x_in.device) for _ in range(3)]
            streams = self._side[:min(3, len(self.launches))]
            for s in streams:
                s.wait_stream(main)
        for k, ln in enumerate(self.launches):
            s = streams[k % len(streams)]
            if events is not None:
                events[k][0].record(s)
            _lib.check(L.mmffhip_minimize_nk(C.byref(ln["desc"]), C.byref(ln["opts"]), _p(x_in), _p(x_out),
                                             _p(self.e), _p(self.st), _p(self.cnt), _p(ln["order"]), _p(ln["ws"]),
                                             ln["ws_bytes"], int(s.cuda_stream)))
            if events is not None:
                events[k][1].record(s)
        if streams[0] is not main:
            for s in streams:
                main.wait_stream(s)
            # the tensors are used on the side streams: keep them alive for the caching allocator
            # (not needed, and not allowed, while a graph is being captured)
            if not torch.cuda.is_current_stream_capturing():
                for t in (x_in, x_out, self.e, self.st, self.cnt):
                    for s in streams:
                        t.record_stream(s)
        return NkResult(x_out, self.e, self.st, self.cnt)


    def capture(self, x_in, x_out):
        """Record this runner's launches (queue resets and every bucket's minimize kernel) into
        a HIP graph via torch.cuda.CUDAGraph. ``graph.replay()`` then re-runs the whole
        minimization of x_in into x_out with one host call; results land in the same tensors
        as run()'s. The launches clear their work queues with a fill kernel, never a memset,
        so every replay starts from the same state."""
        torch = _lib.require_gpu()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self.run(x_in, x_out)  # plans, LDS attributes and code objects before capture
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run(x_in, x_out)
        return g


# candidate (block, placement) pairs per size class for the autotuner; LDS candidates
# whose plan does not fit are skipped
TUNE_CLASSES = (
    (96, ((64, _lib.HESS_GLOBAL), (128, _lib.HESS_GLOBAL), (64, _lib.HESS_LDS), (128, _lib.HESS_LDS), (64, _SYM))),
    (192, ((64, _lib.HESS_GLOBAL), (128, _lib.HESS_GLOBAL), (192, _lib.HESS_GLOBAL), (256, _lib.HESS_GLOBAL),
           (192, _lib.HESS_LDS), (64, _SYM), (128, _SYM))),
    (264, ((128, _lib.HESS_GLOBAL), (192, _lib.HESS_GLOBAL), (256, _lib.HESS_GLOBAL), (640, _lib.HESS_LDS),
           (128, _SYM), (256, _SYM))),
    (360, ((192, _lib.HESS_GLOBAL), (256, _lib.HESS_GLOBAL), (512, _lib.HESS_GLOBAL), (256, _SYM), (512, _SYM))),
    (1 << 20, ((256, _lib.HESS_GLOBAL), (512, _lib.HESS_GLOBAL), (640, _lib.HESS_GLOBAL), (256, _SYM), (512, _SYM))),
)


def autotune_buckets(db: DeviceBatch, x, classes=TUNE_CLASSES, sample: int = 8192, reps: int = 1, seed: int = 0,
                     **opts):
    """Pick the block size (and hence waves per workgroup and workgroups per CU) and the
    Hessian placement of each size class by timing candidate launches on a seeded sample of
    that class's conformers. A persistent launch holds up to 3,072 conformers at once on
    MI355X (256 CUs x 12 waves), so a sample needs several times that many conformers to time
    the steady state rather than the launch's tail; the default takes up to 8,192 per class.

    Returns a bucket tuple for NkRunner and the timings {bound: {(block, placement): ms}}.
    """
    torch = _lib.require_gpu()
    hb = db.host
    n3 = 3 * hb.n_atoms[hb.conf_mol]
    rng = np.random.default_rng(seed)
    lo = 0
    buckets, timings = [], {}
    for bound, cands in classes:
        idx = np.nonzero((n3 > lo) & (n3 <= bound))[0]
        lo = bound
        if len(idx) == 0:
            continue
        pick = np.sort(rng.choice(idx, size=min(sample, len(idx)), replace=False)).astype(np.int32)
        best, tb = None, {}
        for blk, place in cands:
            try:
                r = NkRunner(db, order=pick[np.argsort(-n3[pick], kind="stable")], buckets=((1 << 20, blk, place),),
                             **opts)
                r.run(x)
            except RuntimeError:
                continue  # the plan does not fit (LDS) or has no kernel: not a candidate
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                r.run(x)
            e1.record()
            torch.cuda.synchronize()
            tb[(blk, place)] = e0.elapsed_time(e1) / reps
            if best is None or tb[(blk, place)] < tb[best]:
                best = (blk, place)
        timings[bound] = tb
        if best is not None:
            buckets.append((bound, *best))
    return tuple(buckets), timings


def minimize_nk(db: DeviceBatch, x, **opts) -> NkResult:
    r = NkRunner(db, **opts)
    res = r.run(x)
    return NkResult(res.x, res.energy.clone(), res.status.clone(), res.counters.clone())


def bb_fused(db: DeviceBatch, x, max_iters=500, max_step=0.3, grad_lr=0.1, stream=None):
    torch = _lib.require_gpu()
    L = _lib.load()
    x_out = torch.empty_like(x)
    e = torch.empty(db.n_confs, dtype=torch.float64, device=x.device)
    _lib.check(L.mmffhip_bb_fused(C.byref(db.desc), _p(x), _p(x_out), _p(e), int(max_iters), float(max_step),
                                  float(grad_lr), _lib.stream_ptr(stream)))
    return x_out, e


def multimol(db: DeviceBatch, x, max_iters=200, eval_freq=10, max_step_per_atom=0.3, lbfgs_m=10, grad_scale=0.1,
             grad_tol=1e-4, stream=None, batch_mode=False):
    """Lock-step L-BFGS of mmff_optimize_metal_multi_mol, fully on the device. ``batch_mode``
    runs the mmff_optimize_metal_batch form instead: every iteration up to max_iters, no polish
    pass, grad_norm = ||g_sc||_2 of the last iterate."""
    torch = _lib.require_gpu()
    L = _lib.load()
    nbytes = C.c_size_t()
    _lib.check(L.mmffhip_multimol_workspace(C.byref(db.desc), int(lbfgs_m), C.byref(nbytes)))
    dev = x.device
    ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=dev)
    C_ = db.n_confs
    xb = torch.empty_like(x)
    e = torch.empty(C_, dtype=torch.float64, device=dev)
    gn = torch.empty(C_, dtype=torch.float32, device=dev)
    pci = torch.empty(C_, dtype=torch.int32, device=dev)
    conv = torch.empty(C_, dtype=torch.int32, device=dev)
    used = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.check(L.mmffhip_multimol_run_ex(C.byref(db.desc), _p(x), _p(xb), _p(e), _p(gn), _p(pci), _p(conv),
                                         _p(used), int(max_iters), int(eval_freq), float(max_step_per_atom),
                                         int(lbfgs_m), float(grad_scale), float(grad_tol),
                                         _lib.MM_BATCH if batch_mode else 0, _p(ws), int(nbytes.value),
                                         _lib.stream_ptr(stream)))
    return dict(x=xb, energy=e, grad_norm=gn, per_conf_iters=pci, converged=conv, iters_used=used)
